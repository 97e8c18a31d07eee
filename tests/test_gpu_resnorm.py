"""GPU: the standalone true-residual norm on the reference BSR layout
(pgmres.c:1505-1511) against a float64 scipy evaluation of ||b - A x|| and
||b||, through the C-ABI."""
import numpy as np
import pytest

import jxpamg_all_amd as jx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-12  # relative; the GPU sums in a different (fixed) order


def _dev(a, dt):
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


def _norms(A, x, b):
    ia, ja, a = _dev(A.ia, torch.int32), _dev(A.ja, torch.int32), _dev(A.a, torch.float64)
    xd, bd = _dev(x, torch.float64), _dev(b, torch.float64)
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
    jx.bsr_residual_norm(A.nb, A.b, ia.data_ptr(), ja.data_ptr(), a.data_ptr(), xd.data_ptr(), bd.data_ptr(),
                         out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _random_bsr(nb, b, rng, max_off=4):
    """random block pattern, diagonal first, some rows with the diagonal only"""
    ia, ja = [0], []
    for i in range(nb):
        k = int(rng.integers(0, max_off + 1))
        offs = sorted(set(int(c) for c in rng.integers(0, nb, size=k)) - {i})
        ja += [i] + offs
        ia.append(len(ja))
    a = rng.standard_normal(len(ja) * b * b)
    return jx.BSRMatrix(nb, b, np.array(ia), np.array(ja), a)


@pytest.mark.parametrize("b", [1, 2, 3, 4])
def test_residual_norm_random_patterns(b):
    rng = np.random.default_rng(10 + b)
    for nb in (1, 37, 5000):
        A = _random_bsr(nb, b, rng)
        x, rhs = rng.standard_normal(nb * b), rng.standard_normal(nb * b)
        print(f"[gpu] resnorm b={b} nb={nb}", flush=True)
        got = _norms(A, x, rhs)
        ax = A.to_scipy() @ x
        want = np.array([np.linalg.norm(rhs - ax), np.linalg.norm(rhs)])
        assert np.all(np.abs(got - want) <= TOL * want), (got, want)
        assert np.array_equal(got, _norms(A, x, rhs))  # deterministic


def test_residual_norm_reservoir_system_and_exact_solution():
    A, rhs = jx.generate(**dict(jx.CONFIGS["C1"], nx=30, ny=30, nz=10))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(A.n)
    b = A.to_scipy() @ x
    got = _norms(A, x, b)
    assert got[0] <= 1e-13 * got[1]  # x solves A x = b up to rounding
    got = _norms(A, x, rhs)
    want = np.linalg.norm(rhs - b)
    assert abs(got[0] - want) <= TOL * want


def test_residual_norm_empty_system_and_bad_arguments():
    z = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
    jx.bsr_residual_norm(0, 3, z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    assert out.cpu().tolist() == [0.0, 0.0]
    L = jx.lib()
    for nb, b in ((-1, 3), (4, 0), (4, 5)):
        assert L.jxamd_bsr_residual_norm(nb, b, z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                                         z.data_ptr(), out.data_ptr(), None) == 4
    assert L.jxamd_bsr_residual_norm(4, 3, None, None, None, None, None, out.data_ptr(), None) == 4
