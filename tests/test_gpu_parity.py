"""GPU parity: the HIP path through the C-ABI against the reference's golden
vectors and the oracle.  FP64 throughout; the GPU sums with fused multiply-add
and tree reductions, so operator results agree with the reference to a few
ulps (tolerances below), and whole solves are held to the parity bar:
iteration count within +-2 of the reference and a true FP64 relative residual
at or below rtol."""
import numpy as np
import pytest

import jxpamg_all_amd as jx
from oracle.cpu_ref import Oracle
from tests._common import ranks_system, relerr, system

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

OP_TOL = 1e-11      # single operators, relative to max |expected|
X_TOL = 1e-6        # solution vectors after a full solve


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _solver(name, **kw):
    d, A, rhs, Ad, bd, H, meta = system(name)
    p = jx.default_params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                          stage2_maxit=int(meta["stage2_maxit"]), **kw)
    print(f"[gpu] case={name} nb={Ad.nb} b={Ad.b} levels={H.num_levels if H else 0}", flush=True)
    return d, Ad, bd, H, meta, jx.Solver(Ad, H, p)


@pytest.mark.parametrize("name", ["c1", "tiny", "b4", "b2", "lvl3", "nodec"])
def test_bsr_matvec(name):
    d, Ad, bd, H, meta, s = _solver(name)
    o = Oracle(Ad)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(Ad.n), rng.standard_normal(Ad.n)
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (2.5, -0.5)]:
        yd = dev(y)
        s.bsr_matvec(alpha, dev(x).data_ptr(), beta, yd.data_ptr())
        assert relerr(host(yd), o.matvec(alpha, x, beta, y)) < OP_TOL
    s.close()


@pytest.mark.parametrize("name", ["c1", "tiny", "b4", "b2", "lvl3", "nodec", "timpes"])
def test_smoothers_and_cycle_vs_reference(name):
    d, Ad, bd, H, meta, s = _solver(name)
    u = dev(d["hsgs_u_in"])
    s.hsgs(0, dev(d["hsgs_f"]).data_ptr(), u.data_ptr())
    assert relerr(host(u), d["hsgs_u_out"]) < OP_TOL
    uc = dev(np.zeros(len(d["coarse_u"])))
    s.coarse_solve(dev(d["coarse_f"]).data_ptr(), uc.data_ptr())
    assert relerr(host(uc), d["coarse_u"]) < 1e-9  # explicit inverse vs unpivoted elimination
    uv = dev(np.zeros(len(d["vcycle_u"])))
    s.vcycle(dev(d["vcycle_f"]).data_ptr(), uv.data_ptr())
    assert relerr(host(uv), d["vcycle_u"]) < 1e-9
    xb = dev(d["bhgs_x_in"])
    s.bhgs(xb.data_ptr(), dev(d["bhgs_b"]).data_ptr(), 0.7, 0.9)
    assert relerr(host(xb), d["bhgs_x_out"]) < OP_TOL
    xc = dev(np.zeros(Ad.n))
    s.precond(dev(d["cpr_r"]).data_ptr(), xc.data_ptr())
    assert relerr(host(xc), d["cpr_x"]) < 1e-9
    assert s.error_flag() == 0
    s.close()


@pytest.mark.parametrize("name", ["c1", "tiny", "b4", "b2", "lvl3", "nodec", "restart", "layered", "timpes"])
def test_full_solve_vs_reference(name):
    d, Ad, bd, H, meta, s = _solver(name)
    x, st, hist = s.solve(bd)
    ref_it = int(d["iterations"][0])
    S = Ad.to_scipy()
    true = np.linalg.norm(bd - S @ x) / np.linalg.norm(bd)
    print(f"[gpu] {name}: iters {st.iterations} (reference {ref_it}) true_rel {true:.3e} "
          f"solve {st.t_solve_ms:.2f} ms", flush=True)
    assert abs(st.iterations - ref_it) <= 2
    assert st.converged == 1
    assert true <= meta["rtol"]
    assert abs(st.true_rel_res - true) <= 1e-3 * true + 1e-15
    if "x_final" in d:
        assert relerr(x, d["x_final"]) < X_TOL
    assert relerr(hist[: min(len(hist), len(d["res_history"]))][:4], d["res_history"][:4]) < 1e-6
    assert s.error_flag() == 0
    s.close()


def _threads_per_level(H, relax_threads=0, relax_chunk=0):
    out = []
    for ia, _, _, _ in H.A:
        n = len(ia) - 1
        T = relax_threads if relax_threads > 1 else (-(-n // relax_chunk) if relax_chunk > 0 else 1)
        out.append(min(T, n))
    return out


@pytest.mark.parametrize("name,T", [("lvl3_t4", 4), ("tiny_t16", 16), ("tiny_t64", 64)])
def test_threaded_smoother_vs_reference(name, T):
    """relax_threads = T against the OMP_NUM_THREADS=T reference run; on
    tiny_t64 also relax_chunk = 4 (the benchmark's chunk), which gives the
    fine level the same 64 chunks."""
    if name == "tiny_t64":
        d, Ad, bd, H, meta, s = _solver(name, relax_chunk=4)
        u = dev(d["hsgs_u_in"])
        s.hsgs(0, dev(d["hsgs_f"]).data_ptr(), u.data_ptr())
        assert relerr(host(u), d["hsgs_u_out"]) < OP_TOL
        s.close()
    d, Ad, bd, H, meta, s = _solver(name, relax_threads=T)
    u = dev(d["hsgs_u_in"])
    s.hsgs(0, dev(d["hsgs_f"]).data_ptr(), u.data_ptr())
    assert relerr(host(u), d["hsgs_u_out"]) < OP_TOL
    uv = dev(np.zeros(len(d["vcycle_u"])))
    s.vcycle(dev(d["vcycle_f"]).data_ptr(), uv.data_ptr())
    assert relerr(host(uv), d["vcycle_u"]) < 1e-9
    xc = dev(np.zeros(Ad.n))
    s.precond(dev(d["cpr_r"]).data_ptr(), xc.data_ptr())
    assert relerr(host(xc), d["cpr_x"]) < 1e-9
    x, st, hist = s.solve(bd)
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    assert abs(st.iterations - int(d["iterations"][0])) <= 2 and st.converged == 1 and true <= meta["rtol"]
    assert relerr(x, d["x_final"]) < X_TOL
    assert s.error_flag() == 0
    s.close()


@pytest.mark.parametrize("name,chunk", [("c1", 8), ("tiny", 3), ("b4", 5), ("b2", 1), ("lvl3", 16), ("nodec", 64),
                                        ("layered", 32)])
def test_chunked_smoother_vs_oracle(name, chunk):
    """relax_chunk = c: ceil(n_l / c) chunks per level, against the oracle's
    threaded relax 6 with the same chunk counts; the solve stays within two
    iterations of the one-thread reference."""
    d, Ad, bd, H, meta, s = _solver(name, relax_chunk=chunk)
    o = Oracle(Ad, H)
    T = _threads_per_level(H, relax_chunk=chunk)
    o.set_relax_threads(T)
    rng = np.random.default_rng(chunk)
    for l in range(max(H.num_levels - 1, 1)):
        n = len(H.A[l][0]) - 1
        f, u0 = rng.standard_normal(n), rng.standard_normal(n)
        u = dev(u0)
        s.hsgs(l, dev(f).data_ptr(), u.data_ptr())
        assert relerr(host(u), o.relax6_threads(l, f, u0, T[l])) < OP_TOL
    f = rng.standard_normal(len(H.A[0][0]) - 1)
    uv = dev(np.zeros_like(f))
    s.vcycle(dev(f).data_ptr(), uv.data_ptr())
    assert relerr(host(uv), o.vcycle(f)) < 1e-9
    r = rng.standard_normal(Ad.n)
    p = o.params(stage2_maxit=int(meta["stage2_maxit"]))
    xc = dev(np.zeros(Ad.n))
    s.precond(dev(r).data_ptr(), xc.data_ptr())
    assert relerr(host(xc), o.cpr(r, p)) < 1e-9
    x, st, hist = s.solve(bd)
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    print(f"[gpu] {name} chunk {chunk}: iters {st.iterations} (reference one-thread {int(d['iterations'][0])}) "
          f"true_rel {true:.3e}", flush=True)
    assert abs(st.iterations - int(d["iterations"][0])) <= 2
    assert st.converged == 1 and true <= meta["rtol"]
    assert s.error_flag() == 0
    s.close()


def test_solve_is_deterministic_and_device_entry_agrees():
    d, Ad, bd, H, meta, s = _solver("lvl3")
    x1, st1, _ = s.solve(bd)
    x2, st2, _ = s.solve(bd)
    assert np.array_equal(x1, x2) and st1.iterations == st2.iterations
    bdev, xdev = dev(bd), dev(np.zeros(Ad.n))
    st3 = s.solve_device(bdev.data_ptr(), xdev.data_ptr())
    assert np.array_equal(host(xdev), x1) and st3.iterations == st1.iterations
    s.close()


@pytest.mark.parametrize("name", ["lvl3", "restart", "b4"])
def test_arnoldi_kernels_agree(name):
    """The one-launch MGS step (cooperative, new vector in registers) and the
    launch-per-projection MGS give the same Arnoldi process: same iteration
    count, residual histories and solutions to rounding (the reductions group
    partial sums differently); the one-launch step by an ordinary launch
    gives the cooperative launch's results bit for bit."""
    d, Ad, bd, H, meta, s0 = _solver(name, arnoldi_kernel=0)
    x0, st0, h0 = s0.solve(bd)
    s0.close()
    d, Ad, bd, H, meta, s1 = _solver(name, arnoldi_kernel=1)
    x1, st1, h1 = s1.solve(bd)
    s1.close()
    d, Ad, bd, H, meta, s2 = _solver(name, arnoldi_kernel=2)
    x2, st2, h2 = s2.solve(bd)
    assert s2.error_flag() == 0
    s2.close()
    print(f"[gpu] {name}: arnoldi kernels iterations {st0.iterations} / {st1.iterations} / {st2.iterations}",
          flush=True)
    assert st0.iterations == st1.iterations and st0.converged == st1.converged == 1
    n = st0.iterations + 1
    assert relerr(h0[:n], h1[:n]) < 1e-9
    assert relerr(x0, x1) < 1e-9
    # the same kernel and grid launched without the cooperative launch: bit for bit
    assert st2.iterations == st0.iterations and np.array_equal(h2[:n], h0[:n]) and np.array_equal(x2, x0)


@pytest.mark.parametrize("name", ["lvl3", "restart", "b4"])
def test_speculative_cycle_start_bit_exact(name):
    """An unprofiled solve scales v_0 on the device and enqueues its first
    Arnoldi step before the norms reach the host, and looks a step ahead; a
    profiled solve does neither (host-side 1 / ||r_0||, one step at a time).
    Same iterations, residual histories and x, bit for bit."""
    d, Ad, bd, H, meta, s = _solver(name)
    x0, st0, h0 = s.solve(bd)
    s.profile(True)
    x1, st1, h1 = s.solve(bd)
    s.profile(False)
    x2, st2, h2 = s.solve(bd)
    assert s.error_flag() == 0
    s.close()
    n = st0.iterations + 1
    print(f"[gpu] {name}: iterations {st0.iterations} / {st1.iterations} / {st2.iterations}", flush=True)
    assert st0.converged == st1.converged == 1
    assert st0.iterations == st1.iterations == st2.iterations
    assert np.array_equal(h0[:n], h1[:n]) and np.array_equal(x0, x1)
    assert np.array_equal(h0[:n], h2[:n]) and np.array_equal(x0, x2)


def test_solve_starts_that_stop_at_once():
    """The speculative first step is left unused when the solve stops before
    its first cycle: a zero right-hand side (zero residual: x stays zero, no
    iteration) and an initial guess that already meets the tolerance (the
    converged x of a first solve: no iteration, x unchanged); a fresh solve
    afterwards gives the first solve's x bit for bit."""
    d, Ad, bd, H, meta, s = _solver("lvl3")
    x0, st0, _ = s.solve(bd)
    assert st0.converged == 1 and st0.iterations > 0
    xz, stz, _ = s.solve(np.zeros(Ad.n))
    assert stz.iterations == 0 and not np.any(xz)
    xc, stc, _ = s.solve(bd, x0=x0)
    assert stc.iterations == 0 and stc.converged == 1 and np.array_equal(xc, x0)
    x3, st3, _ = s.solve(bd)
    assert s.error_flag() == 0
    s.close()
    assert st3.iterations == st0.iterations and np.array_equal(x3, x0)


@pytest.mark.parametrize("name", ["lvl3", "restart", "b4"])
def test_solution_updates_agree(name):
    """x += Z y with the Arnoldi steps' z_j = M^{-1} v_j kept, against the
    reference's x += M^{-1}(V y): for the FP64 preconditioner (linear) the
    same Arnoldi process and the same x to rounding, with one preconditioner
    application fewer per cycle."""
    d, Ad, bd, H, meta, s0 = _solver(name, solution_update=0)
    x0, st0, h0 = s0.solve(bd)
    s0.close()
    d, Ad, bd, H, meta, s1 = _solver(name, solution_update=1)
    x1, st1, h1 = s1.solve(bd)
    s1.close()
    print(f"[gpu] {name}: solution updates iterations {st0.iterations} / {st1.iterations}, applications "
          f"{st0.num_precond} / {st1.num_precond}", flush=True)
    assert st0.iterations == st1.iterations and st0.converged == st1.converged == 1
    n = st0.iterations + 1
    assert relerr(h0[:n], h1[:n]) < 1e-9
    assert relerr(x0, x1) < 1e-9
    assert st0.true_rel_res <= meta["rtol"] and st1.true_rel_res <= meta["rtol"]
    assert st0.num_precond == st0.iterations
    assert st1.num_precond == st1.iterations + st1.num_restarts + 1


def test_solution_updates_mixed_precision():
    """With an FP16 stage 1 (a preconditioner linear only to FP16 rounding)
    both updates reach the FP64 true residual in the same number of
    iterations to one; the kept-basis update is the flexible-GMRES one, for
    which the Arnoldi relation holds exactly."""
    its = []
    for update in (0, 1):
        d, Ad, bd, H, meta, s = _solver("lvl3", solution_update=update, precond_precision=2, poly_degree=2,
                                        relax_chunk=4, stage2_chunk=4, stage2_precision=2)
        x, st, h = s.solve(bd)
        s.close()
        print(f"[gpu] lvl3 FP16 stage 1, update {update}: {st.iterations} iterations, true "
              f"{st.true_rel_res:.3e}", flush=True)
        assert st.converged == 1 and st.true_rel_res <= meta["rtol"]
        its.append(st.iterations)
    assert abs(its[0] - its[1]) <= 1


def test_nan_rhs_is_reported_as_error():
    d, Ad, bd, H, meta, s = _solver("tiny")
    b = bd.copy()
    b[3] = np.nan
    with pytest.raises(RuntimeError):
        s.solve(b)
    # the solver stays usable afterwards
    x, st, _ = s.solve(bd)
    assert st.converged == 1
    s.close()


def test_c2_host_setup_and_gpu_solve_vs_reference():
    """Full 60x220x85 system: host-built hierarchy (bit-identical to the
    reference's) and the GPU solve against the reference's own run."""
    from tests.conftest import fixture_gen_params, load_fixture
    d = load_fixture("c2")
    A, rhs = jx.generate(**fixture_gen_params(d))
    Ad, bd = jx.decouple_abf(A, rhs)
    ia, ja, a = jx.extract_pressure(Ad)
    H = jx.amg_setup(ia, ja, a)
    assert [len(m[0]) - 1 for m in H.A] == list(d["level_sizes"])
    print(f"[gpu] case=c2 nb={Ad.nb} levels={H.num_levels}", flush=True)
    S = Ad.to_scipy()
    ref_it = int(d["iterations"][0])
    runs = [(0, {}), (jx.BENCH_RELAX_CHUNK, {}),
            (jx.BENCH_RELAX_CHUNK, dict(stage2_chunk=jx.BENCH_STAGE2_CHUNK)),
            # the mixed configuration: FP16 stage 1, degree-2 polynomial, one
            # FP16-block stage-2 sweep; FP64 GMRES must still reach rtol
            (0, dict(precond_precision=2, poly_degree=2, stage2_maxit=1, stage2_weight=1.0, stage2_omega=0.8,
                     stage2_chunk=jx.BENCH_STAGE2_CHUNK, stage2_precision=2))]
    for chunk, kw in runs:
        s = jx.Solver(Ad, H, jx.default_params(tol=1e-6, max_iter=100, relax_chunk=chunk, **kw))
        x, st, hist = s.solve(bd)
        true = np.linalg.norm(bd - S @ x) / np.linalg.norm(bd)
        print(f"[gpu] c2 chunk {chunk} {kw}: iters {st.iterations} (reference {ref_it}) true_rel {true:.3e} "
              f"solve {st.t_solve_ms:.1f} ms", flush=True)
        assert abs(st.iterations - ref_it) <= 2
        assert st.converged == 1 and true <= 1e-6
        if chunk == 0 and not kw:
            n = min(len(hist), len(d["res_history"]))
            assert relerr(hist[:n], d["res_history"][:n]) < 1e-6
        assert s.error_flag() == 0
        s.close()


def _node_major_csr(A, one_based=True, seed=0):
    """BSR -> CSR with b consecutive rows per cell, entries shuffled inside rows."""
    S = A.to_scipy().tocsr()
    S.sort_indices()
    rng = np.random.default_rng(seed)
    ia, ja, a = [0], [], []
    for i in range(S.shape[0]):
        cols, vals = S.indices[S.indptr[i]:S.indptr[i + 1]], S.data[S.indptr[i]:S.indptr[i + 1]]
        order = rng.permutation(len(cols))
        ja += list(cols[order])
        a += list(vals[order])
        ia.append(len(ja))
    ia, ja = np.array(ia, np.int32), np.array(ja, np.int32)
    if one_based:
        ia += 1
        ja += 1
    return ia, ja, np.array(a)


def test_solver_interface_case_220():
    """JXFPAMG_Solver_Interface arguments, solver 220: node-major 1-based CSR
    in, BSR CPR-GMRES (no decoupling, one stage-2 sweep) on the GPU."""
    from tests.conftest import fixture_gen_params, load_fixture
    d = load_fixture("c1")
    A, rhs = jx.generate(**fixture_gen_params(d))
    ia, ja, a = _node_major_csr(A)
    x = np.zeros(A.n)
    print("[gpu] case=interface-220 c1", flush=True)
    iters = jx.solver_interface(A.n, ia, ja, a, x, rhs.copy(), 1e-6, 220, A.b)
    # in-place effects of the reference entry: 0-based, diagonal first
    assert ia[0] == 0 and all(ja[ia[i]] == i for i in range(A.n))
    true = np.linalg.norm(rhs - A.to_scipy() @ x) / np.linalg.norm(rhs)
    assert true <= 1e-6
    # same system through the BSR entry and through the oracle
    Ap = jx.extract_pressure(A)
    H = jx.amg_setup(*Ap)
    s = jx.Solver(A, H, jx.default_params(tol=1e-6, max_iter=1000, stage2_maxit=1))
    x2, st, _ = s.solve(rhs)
    s.close()
    assert st.iterations == iters and relerr(x, x2) < 1e-12
    o = Oracle(A, H)
    xo, sto, _, _ = o.gmres(rhs, o.params(tol=1e-6, max_iter=1000, stage2_maxit=1))
    print(f"[gpu] interface-220: iters {iters} oracle {sto.iterations} true_rel {true:.3e}", flush=True)
    assert abs(iters - sto.iterations) <= 2
    with pytest.raises(RuntimeError):
        jx.solver_interface(A.n, ia, ja, a, x, rhs.copy(), 1e-6, 22, A.b)


def _pow2_at_least(mx):
    """the smallest power of two at or above mx > 0"""
    m, ex = np.frexp(mx)
    return float(np.ldexp(1.0, int(ex) - 1 if m == 0.5 else int(ex)))


def _bilu_solve_f16(Ad, lu, dinv, r):
    """(LU)^{-1} r in FP64 with the factors as the FP16 storage holds them:
    the blocks of block row i in one triangle divided by the smallest power
    of two at or above their largest magnitude, rounded to FP16 (through
    FP32, as the upload does), multiplied back; U's inverted diagonal blocks
    rounded to FP32."""
    b, nb = Ad.b, Ad.nb
    blk = lu.reshape(-1, b, b)
    dv = dinv.reshape(nb, b, b).astype(np.float32).astype(np.float64)
    ia, ja = Ad.ia, Ad.ja

    def stored(ks):
        if not ks:
            return []
        sc = _pow2_at_least(max(np.abs(blk[k]).max() for k in ks))
        return [(ja[k], sc * (blk[k] / sc).astype(np.float32).astype(np.float16).astype(np.float64)) for k in ks]

    y = r.reshape(nb, b).copy()
    for i in range(nb):
        for j, m in stored([k for k in range(ia[i], ia[i + 1]) if ja[k] < i]):
            y[i] -= m @ y[j]
    x = np.zeros_like(y)
    for i in range(nb - 1, -1, -1):
        acc = y[i].copy()
        for j, m in stored([k for k in range(ia[i], ia[i + 1]) if ja[k] > i]):
            acc -= m @ x[j]
        x[i] = dv[i] @ acc
    return x.ravel()


@pytest.mark.parametrize("name,prec,s2p", [("c1", 0, 0), ("b4", 0, 0), ("b2", 0, 0), ("lvl3", 0, 0), ("lvl3", 1, 0),
                                           ("c1", 0, 1), ("c1", 1, 1), ("lvl3", 0, 2), ("b4", 0, 2), ("b2", 0, 2)])
def test_bilu0_stage2_vs_oracle(name, prec, s2p):
    """Stage-2 solver type 5: x += (LU)^{-1}(r - A x) with exact BILU(0)
    factors stored in FP64, in FP32 (prec = 1 or stage2_precision s2p = 1),
    or in FP16 with a power-of-two scale per block row and FP32 inverted
    diagonals (s2p = 2).  ("c1", 1, 1) is the 16x16x8 mixed configuration:
    AMG hierarchy and BILU(0) factors in FP32, FP64 GMRES and true residual,
    reported against the FP64 oracle's iterations.  The FP16 solve is checked against its restatement
    on the rounded factors (1e-10: only the summation order differs) and,
    within 1e-2, against the FP64 factors; its solve reaches rtol within
    three iterations of the oracle's FP64 BILU(0) solve."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    p = jx.default_params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                          stage2_maxit=1, stage2_solver_type=5, precond_precision=prec, stage2_precision=s2p)
    print(f"[gpu] case={name} bilu0 precision={prec} stage2_precision={s2p}", flush=True)
    s = jx.Solver(Ad, H, p)
    o = Oracle(Ad, H)
    r = np.random.default_rng(4).standard_normal(Ad.n)
    z = dev(np.zeros(Ad.n))
    s.bilu_solve(dev(r).data_ptr(), z.data_ptr())
    if s2p == 2:
        lu, dinv = o.bilu0()
        e16, e64 = relerr(host(z), _bilu_solve_f16(Ad, lu, dinv, r)), relerr(host(z), o.bilu0_solve(r))
        print(f"[gpu] {name} FP16 BILU(0) solve: vs rounded factors {e16:.2e}, vs FP64 factors {e64:.2e}", flush=True)
        assert e16 < 1e-10 and e64 < 1e-2
    else:
        assert relerr(host(z), o.bilu0_solve(r)) < (OP_TOL if prec == 0 and s2p == 0 else 1e-5)
    x, st, hist = s.solve(bd)
    xo, sto, _, _ = o.gmres(bd, o.params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                                         stage2_maxit=1, stage2_solver_type=5))
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    print(f"[gpu] {name} bilu0: iters {st.iterations} (oracle {sto.iterations}) true_rel {true:.3e}", flush=True)
    assert st.converged == 1 and true <= meta["rtol"]
    exact = prec == 0 and s2p == 0
    assert abs(st.iterations - sto.iterations) <= (0 if exact else 2 if s2p < 2 else 3)
    if exact:
        assert relerr(x, xo) < X_TOL
    assert s.error_flag() == 0
    s.close()


@pytest.mark.parametrize("name,chunk", [("c1", 4), ("b4", 3), ("b2", 1), ("lvl3", 8), ("layered", 4), ("nodec", 4)])
def test_stage2_rank_partitions_vs_oracle(name, chunk):
    """stage2_chunk = c: stage 2 as run on ceil(nb / c) ranks, against the
    oracle's restatement; the solve stays within two iterations of the
    one-rank reference."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    p = jx.default_params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                          stage2_maxit=int(meta["stage2_maxit"]), stage2_chunk=chunk, relax_chunk=4)
    print(f"[gpu] case={name} stage2 chunk {chunk}", flush=True)
    s = jx.Solver(Ad, H, p)
    o = Oracle(Ad, H)
    parts = -(-Ad.nb // chunk)
    rng = np.random.default_rng(chunk)
    x0, r = rng.standard_normal(Ad.n), rng.standard_normal(Ad.n)
    xb = dev(x0)
    s.bhgs(xb.data_ptr(), dev(r).data_ptr(), 0.7, 0.9)
    assert relerr(host(xb), o.bhgs_parts(x0, r, parts)) < OP_TOL
    o.set_relax_threads([min(len(m[0]) - 1, -(-(len(m[0]) - 1) // 4)) for m in H.A])
    op = o.params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                  stage2_maxit=int(meta["stage2_maxit"]), stage2_parts=parts)
    xc = dev(np.zeros(Ad.n))
    s.precond(dev(r).data_ptr(), xc.data_ptr())
    assert relerr(host(xc), o.cpr(r, op)) < 1e-9
    x, st, hist = s.solve(bd)
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    print(f"[gpu] {name} stage2 chunk {chunk}: iters {st.iterations} (reference one-rank {int(d['iterations'][0])}) "
          f"true_rel {true:.3e}", flush=True)
    assert abs(st.iterations - int(d["iterations"][0])) <= 2
    assert st.converged == 1 and true <= meta["rtol"]
    assert s.error_flag() == 0
    s.close()


def test_exact_abf_4x4_solve_vs_oracle():
    """4-unknown blocks with the exact decoupling (the reference's 4x4 block
    inverse is wrong and its CPR-GMRES stalls on such systems): the GPU solve
    against the oracle's on the same decoupled system, and the solution
    against the raw, undecoupled system."""
    A, rhs = jx.generate(nx=14, ny=12, nz=10, block_size=4, field=1, sigma_lnk=2.0, corr_len=5.0, seed=7)
    Ad, bd = jx.decouple_abf(A, rhs, exact=True)
    H = jx.amg_setup(*jx.extract_pressure(Ad))
    print(f"[gpu] exact ABF b=4 nb={Ad.nb} levels={H.num_levels}", flush=True)
    o = Oracle(Ad, H)
    xo, sto, _, _ = o.gmres(bd, o.params(tol=1e-8, max_iter=100))
    for kw in ({}, {"relax_chunk": 4, "stage2_chunk": 4}):
        s = jx.Solver(Ad, H, jx.default_params(tol=1e-8, max_iter=100, **kw))
        x, st, _ = s.solve(bd)
        true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
        raw = np.linalg.norm(rhs - A.to_scipy() @ x) / np.linalg.norm(rhs)
        print(f"[gpu] exact ABF {kw}: iters {st.iterations} (oracle {sto.iterations}) true_rel {true:.3e} "
              f"raw {raw:.3e}", flush=True)
        assert st.converged == 1 and true <= 1e-8
        if not kw:
            assert abs(st.iterations - sto.iterations) <= 2
            assert relerr(x, xo) < X_TOL
        assert raw < 1e-5  # D^-1 scaling: the raw relative residual is within cond(D) of the decoupled one
        s.close()


@pytest.mark.parametrize("name,rows", [("lvl3", 2000), ("layered", 1500)])
def test_truncated_vcycle_vs_oracle(name, rows):
    """coarse_rows: the V-cycle ends at the first level with at most `rows`
    rows (here level 1) and solves it with an explicit inverse; the oracle
    runs the same V-cycle on the hierarchy cut there, its coarsest level
    solved by Gaussian elimination (relax 9)."""
    d, Ad, bd, H, meta, s = _solver(name, coarse_rows=rows)
    s.close()
    Ht = jx.Hierarchy(H.A[:2], H.P[:1])
    assert len(H.A[1][0]) - 1 <= rows < len(H.A[0][0]) - 1
    p = jx.default_params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                          stage2_maxit=int(meta["stage2_maxit"]), coarse_rows=rows)
    s = jx.Solver(Ad, H, p)
    assert s.num_levels == 2
    o = Oracle(Ad, Ht)
    f = np.random.default_rng(4).standard_normal(len(H.A[0][0]) - 1)
    uv = dev(np.zeros_like(f))
    s.vcycle(dev(f).data_ptr(), uv.data_ptr())
    assert relerr(host(uv), o.vcycle(f)) < 1e-9
    x, st, hist = s.solve(bd)
    xo, sto, histo, rc = o.gmres(bd, o.params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]),
                                             tol=meta["rtol"], stage2_maxit=int(meta["stage2_maxit"])))
    print(f"[gpu] truncated {name} at {rows} rows: iters {st.iterations} oracle {sto.iterations}", flush=True)
    assert st.converged == 1 and abs(st.iterations - sto.iterations) <= 2
    assert np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd) <= meta["rtol"]
    s.close()


@pytest.mark.parametrize("name", ["lvl3", "layered"])
def test_exact_sweep_every_level_vs_oracle(name):
    """The exact one-thread relax 6 (forward then backward Gauss-Seidel,
    par_relax_6.c:224-273) on every smoothed level, from a nonzero guess:
    level 0 (short rows) runs one lane per row, level 1 (19-21 entries per
    row) row groups of 4 lanes over the plain CSR rows; both against the
    oracle's sequential sweep."""
    d, Ad, bd, H, meta, s = _solver(name)
    o = Oracle(Ad, H)
    rng = np.random.default_rng(11)
    for l in range(H.num_levels - 1):
        n = len(H.A[l][0]) - 1
        f, u0 = rng.standard_normal(n), rng.standard_normal(n)
        u = dev(u0)
        s.hsgs(l, dev(f).data_ptr(), u.data_ptr())
        err = relerr(host(u), o.relax6(l, f, u0))
        print(f"[gpu] exact relax 6 {name} level {l} ({n} rows): relerr {err:.2e}", flush=True)
        assert err < OP_TOL
    s.close()


@pytest.mark.parametrize("b", [2, 3, 4])
def test_stage2_partitions_vs_reference_mpi_ranks(b):
    """stage2_chunk = c on 24 block rows is stage 2 as the reference runs it
    on P = ceil(24 / c) MPI ranks: one sweep against the reference's own
    multi-rank sweep (ranks.npz), for c = 8 down to 1 (uneven at c = 5)."""
    from tests.conftest import load_fixture
    d = load_fixture("ranks")
    A = ranks_system(d, b)
    H = jx.amg_setup(*jx.extract_pressure(A))
    for c in (8, 6, 5, 4, 3, 2, 1):
        P = -(-A.nb // c)
        s = jx.Solver(A, H, jx.default_params(stage2_chunk=c))
        x = dev(d[f"b{b}_x0"])
        s.bhgs(x.data_ptr(), dev(d[f"b{b}_rhs"]).data_ptr(), 0.7, 0.9)
        err = relerr(host(x), d[f"b{b}_x_p{P}"])
        print(f"[gpu] b={b} stage2_chunk {c} = {P} ranks: relerr {err:.2e}", flush=True)
        assert err < OP_TOL
        s.close()
