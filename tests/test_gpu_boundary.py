"""GPU: the drop-in entries of the boundary through their C-ABI.  The
reference-signature JXFPAMG_Solver_Interface (libjxamd_jxfpamg.so) gives the
same solve as jxamd_solver_interface; the JXF_PtrToSolverFcn-shaped CPR
callback on host vectors equals the device CPR application and the oracle's
CPR (jxf_cpr.c:424-610)."""
import ctypes
import os

import numpy as np
import pytest

import jxpamg_all_amd as jx
from oracle.cpu_ref import Oracle
from tests._common import relerr, system
from tests.conftest import ROOT, fixture_gen_params, load_fixture

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHIM = os.path.join(ROOT, "jxpamg_all_amd", "libjxamd_jxfpamg.so")


class PrecondCtx(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_void_p), ("vector_data", ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p))]


@pytest.mark.parametrize("name", ["c1", "b4"])
def test_precond_callback_matches_device_apply_and_oracle(name):
    d, A, rhs, Ad, bd, H, meta = system(name)
    s = jx.Solver(Ad, H, jx.default_params(stage2_maxit=int(meta["stage2_maxit"])))
    # "vectors" are the numpy arrays themselves: their host array is their address
    vec_data = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p)(lambda v: v)
    ctx = PrecondCtx(s._s.value, vec_data)
    r = np.ascontiguousarray(d["cpr_r"])
    x = np.full(Ad.n, np.nan)
    print(f"[gpu] precond callback {name}", flush=True)
    rc = jx.lib().jxamd_cpr_precond_callback(ctypes.addressof(ctx), None, r.ctypes.data, x.ctypes.data)
    assert rc == 0
    xd = torch.zeros(Ad.n, dtype=torch.float64, device="cuda")
    s.precond(torch.tensor(r, device="cuda").data_ptr(), xd.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(x, xd.cpu().numpy())
    assert relerr(x, d["cpr_x"]) < 1e-9  # the reference's own CPR application
    o = Oracle(Ad, H)
    assert relerr(x, o.cpr(r, o.params(stage2_maxit=int(meta["stage2_maxit"])))) < 1e-9
    s.close()


@pytest.mark.skipif(not os.path.exists(SHIM), reason="libjxamd_jxfpamg.so not built (no mpi.h)")
def test_reference_signature_entry_solves_like_the_csr_entry():
    from tests.test_gpu_parity import _node_major_csr
    d = load_fixture("c1")
    A, rhs = jx.generate(**fixture_gen_params(d))
    L = ctypes.CDLL(SHIM)
    f = L.JXFPAMG_Solver_Interface
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5 + \
        [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    comm_world = 0x44000000  # MPI_COMM_WORLD of the image's MPICH (MPI not initialised: one rank)
    ia, ja, a = _node_major_csr(A)
    x, b = np.zeros(A.n), rhs.copy()
    it = ctypes.c_int(0)
    print("[gpu] JXFPAMG_Solver_Interface c1", flush=True)
    rc = f(0, None, comm_world, A.n, jx._p(ia), jx._p(ja), jx._p(a), jx._p(x), jx._p(b), 1e-6, 220, A.b,
           ctypes.byref(it))
    assert rc == 0
    ia2, ja2, a2 = _node_major_csr(A)
    x2 = np.zeros(A.n)
    it2 = jx.solver_interface(A.n, ia2, ja2, a2, x2, rhs.copy(), 1e-6, 220, A.b)
    assert it.value == it2 and np.array_equal(x, x2)
    assert np.linalg.norm(rhs - A.to_scipy() @ x) / np.linalg.norm(rhs) <= 1e-6
