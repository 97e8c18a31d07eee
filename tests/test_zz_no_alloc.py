"""GPU: the allocation-free true-residual norm (static device scratch)
against the allocating entry point, through the C-ABI: two calls in a row,
the second on a smaller system, each compared with the allocating entry
point's result (same kernels and reduction order, so bit for bit) and with a
float64 scipy evaluation."""
import numpy as np
import pytest

import jxpamg_all_amd as jx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dt):
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


def _call(fn, A, x, b):
    ia, ja, a = _dev(A.ia, torch.int32), _dev(A.ja, torch.int32), _dev(A.a, torch.float64)
    xd, bd = _dev(x, torch.float64), _dev(b, torch.float64)
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(jx.lib(), fn)(A.nb, A.b, ia.data_ptr(), ja.data_ptr(), a.data_ptr(), xd.data_ptr(), bd.data_ptr(),
                               out.data_ptr(), None)
    assert rc == 0, (fn, rc)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_on_two_calls():
    big, _ = jx.generate(**dict(jx.CONFIGS["C1"], nx=40, ny=40, nz=20, seed=5))   # 32,000 cells, 96,000 rows
    small, _ = jx.generate(**dict(jx.CONFIGS["C1"], nx=6, ny=5, nz=3, block_size=2, seed=6))
    rng = np.random.default_rng(1)
    for A in (big, small):  # the second call is smaller and must not see the first call's partials
        x, b = rng.standard_normal(A.n), rng.standard_normal(A.n)
        print(f"[gpu] no-alloc residual norm nb={A.nb} b={A.b}", flush=True)
        ref = _call("jxamd_bsr_residual_norm", A, x, b)
        got = _call("jxamd_bsr_residual_norm_noalloc", A, x, b)
        assert np.array_equal(got, ref), (got, ref)
        want = np.array([np.linalg.norm(b - A.to_scipy() @ x), np.linalg.norm(b)])
        assert np.all(np.abs(got - want) <= 1e-12 * want)


def test_noalloc_rejects_oversized_and_invalid_input():
    L = jx.lib()
    z = torch.zeros(4, dtype=torch.int32, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    p = z.data_ptr()
    assert L.jxamd_bsr_residual_norm_noalloc(1 << 30, 4, p, p, p, p, p, out.data_ptr(), None) == 4
    assert L.jxamd_bsr_residual_norm_noalloc(4, 5, p, p, p, p, p, out.data_ptr(), None) == 4
    assert L.jxamd_bsr_residual_norm_noalloc(4, 3, p, p, p, p, p, None, None) == 4
