"""GPU polynomial AMG smoother (poly_degree > 0) against the oracle's
restatement: one smoother application, the V-cycle and the whole solve, with
the solver's default Chebyshev step sizes and with step sizes set through the
C-ABI."""
import numpy as np
import pytest

import jxpamg_all_amd as jx
from oracle.cpu_ref import Oracle
from tests._common import relerr, system

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

OP_TOL = 1e-11
CYCLE_TOL = 1e-9


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("name,deg,prec", [("c1", 4, 0), ("lvl3", 4, 0), ("lvl3", 3, 0), ("layered", 2, 0),
                                           ("lvl3", 4, 1)])
def test_poly_smoother_vs_oracle(name, deg, prec):
    d, A, rhs, Ad, bd, H, meta = system(name)
    p = jx.default_params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                          stage2_maxit=int(meta["stage2_maxit"]), poly_degree=deg, precond_precision=prec)
    print(f"[gpu] case={name} poly degree {deg} precision {prec}", flush=True)
    s = jx.Solver(Ad, H, p)
    L = s.num_levels
    smooth_levels = range(max(L - 1, 1))
    omegas = [s.get_poly(l, deg)[0] for l in smooth_levels]
    lams = [s.get_poly(l, deg)[1] for l in smooth_levels]
    assert all(0.5 < lam < 3.0 for lam in lams), lams  # D^{-1}A of these M-matrices: lambda_max in (1, 2]
    o = Oracle(Ad, H)
    o.set_poly([list(w) for w in omegas] + [list(omegas[0])] * (H.num_levels - len(omegas)))
    tol_op, tol_cyc = (OP_TOL, CYCLE_TOL) if prec == 0 else (2e-5, 1e-4)
    rng = np.random.default_rng(deg)
    for l in smooth_levels:
        n = len(H.A[l][0]) - 1
        f, u0 = rng.standard_normal(n), rng.standard_normal(n)
        u = dev(u0)
        s.hsgs(l, dev(f).data_ptr(), u.data_ptr())
        assert relerr(host(u), o.poly_smooth(l, f, u0, omegas[l])) < tol_op
    n0 = len(H.A[0][0]) - 1
    f = rng.standard_normal(n0)
    uv = dev(np.zeros(n0))
    s.vcycle(dev(f).data_ptr(), uv.data_ptr())
    assert relerr(host(uv), o.vcycle(f)) < tol_cyc
    # step sizes replaced through the C-ABI are the ones used
    w2 = np.linspace(0.5, 1.2, deg)
    s.set_poly(0, w2)
    assert np.array_equal(s.get_poly(0, deg)[0], w2)
    u = dev(np.zeros(n0))
    s.hsgs(0, dev(f).data_ptr(), u.data_ptr())
    assert relerr(host(u), o.poly_smooth(0, f, np.zeros(n0), w2)) < tol_op
    s.set_poly(0, omegas[0])
    x, st, hist = s.solve(bd)
    xo, sto, _, _ = o.gmres(bd, o.params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                                         stage2_maxit=int(meta["stage2_maxit"])))
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    print(f"[gpu] {name} poly{deg} prec {prec}: iters {st.iterations} (oracle {sto.iterations}, reference relax-6 "
          f"{int(d['iterations'][0])}) true_rel {true:.3e}", flush=True)
    assert st.converged == 1 and true <= meta["rtol"]
    assert abs(st.iterations - sto.iterations) <= (1 if prec == 0 else 2)
    assert s.error_flag() == 0
    with pytest.raises(RuntimeError):
        s.set_poly(0, np.ones(deg + 1))  # degree must match
    s.close()
