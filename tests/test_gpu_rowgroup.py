"""Row-group kernels (G = 4..32 lanes per CSR row, chosen by mean row length)
on one-level hierarchies with long rows: the chunked relax-6 smoother and the
polynomial smoother against the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import jxpamg_all_amd as jx
from oracle.cpu_ref import Oracle
from tests._common import relerr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(x):
    return torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def wide_level(n, rowlen, seed):
    """Random diagonally dominant CSR with about rowlen entries per row, the
    diagonal stored first; couplings mostly near the diagonal (in and out of
    the chunks) plus a few far ones."""
    rng = np.random.default_rng(seed)
    ia, ja, a = [0], [], []
    for i in range(n):
        near = i + rng.integers(-3 * rowlen, 3 * rowlen, rowlen // 2)
        far = rng.integers(0, n, rowlen - rowlen // 2)
        cols = np.unique(np.concatenate([near, far]))
        cols = cols[(cols >= 0) & (cols < n) & (cols != i)]
        vals = -rng.uniform(0.1, 1.0, len(cols))
        ja += [i] + list(cols)
        a += [1.1 * np.abs(vals).sum() + 0.5] + list(vals)
        ia.append(len(ja))
    return np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)


def trivial_bsr(n):
    return jx.BSRMatrix(n, 2, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
                        np.tile(np.eye(2).ravel(), n))


@pytest.mark.parametrize("rowlen", [16, 32, 64, 100, 140])
@pytest.mark.parametrize("prec", [0, 1])
def test_rowgroup_smoothers_vs_oracle(rowlen, prec):
    n = 2000
    ia, ja, a = wide_level(n, rowlen, rowlen)
    H = jx.Hierarchy([(ia, ja, a, n)], [])
    A = trivial_bsr(n)
    print(f"[gpu] row groups: mean row {len(ja) / n:.1f} precision {prec}", flush=True)
    rng = np.random.default_rng(rowlen + prec)
    f, u0 = rng.standard_normal(n), rng.standard_normal(n)
    tol = 1e-11 if prec == 0 else 2e-5
    for chunk in (1, 4, 13):
        s = jx.Solver(A, H, jx.default_params(relax_chunk=chunk, precond_precision=prec))
        T = min(n, -(-n // chunk))
        u = dev(u0)
        s.hsgs(0, dev(f).data_ptr(), u.data_ptr())
        assert relerr(host(u), Oracle(A, H).relax6_threads(0, f, u0, T)) < tol, chunk
        assert s.error_flag() == 0
        s.close()
    s = jx.Solver(A, H, jx.default_params(poly_degree=4, precond_precision=prec))
    w, lam = s.get_poly(0, 4)
    u = dev(u0)
    s.hsgs(0, dev(f).data_ptr(), u.data_ptr())
    assert relerr(host(u), Oracle(A, H).poly_smooth(0, f, u0, w)) < tol
    s.close()
