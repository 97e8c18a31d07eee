"""The drop-in boundary on the host (no GPU calls): the reference-signature
JXFPAMG_Solver_Interface in libjxamd_jxfpamg.so (built against the image's
mpi.h) has exactly the reference's prototype and rejects what the reference
entry would not run here before touching the device; the preconditioner
callback has the shape of the reference's JXF_PtrToSolverFcn."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import jxpamg_all_amd as jx
from tests.conftest import ROOT

MPI_INC = "/opt/conda/include"
SHIM = os.path.join(ROOT, "jxpamg_all_amd", "libjxamd_jxfpamg.so")

needs_mpi = pytest.mark.skipif(not os.path.exists(os.path.join(MPI_INC, "mpi.h")) or not os.path.exists(SHIM),
                               reason="no mpi.h in this image (the MPI-signature entry is not built)")

# the reference's declaration (src/jxpamg_interface.c:118-119), with its FP64 types
REFERENCE_DECL = r"""
typedef int JXF_Int;
typedef double JXF_Real;
JXF_Int JXFPAMG_Solver_Interface(JXF_Int argc, char *argv[],MPI_Comm comm,JXF_Int row, JXF_Int *ia, JXF_Int  *ja,
                                 JXF_Real *a, JXF_Real *ser_x, JXF_Real *ser_b, JXF_Real tol, JXF_Int solver_id,
                                 JXF_Int num_functions, JXF_Int *iter);
typedef JXF_Int (*JXF_PtrToSolverFcn)(void *, void *, void *, void *);
"""


@needs_mpi
def test_reference_prototype_is_the_exported_one_and_links():
    """A translation unit holding both our declaration and the reference's
    compiles only if the two are the same type; it then links against the
    library and registers the CPR callback where the reference takes a
    JXF_PtrToSolverFcn."""
    src = ('#include "jxamd_jxfpamg.h"\n#include "jxamd.h"\n' + REFERENCE_DECL + r"""
int main(void) {
    JXF_PtrToSolverFcn precond = (JXF_PtrToSolverFcn)jxamd_cpr_precond_callback;
    JXF_Int (*entry)(JXF_Int, char **, MPI_Comm, JXF_Int, JXF_Int *, JXF_Int *, JXF_Real *, JXF_Real *, JXF_Real *,
                     JXF_Real, JXF_Int, JXF_Int, JXF_Int *) = JXFPAMG_Solver_Interface;
    return (precond != 0 && entry != 0) ? 0 : 1;
}
""")
    with tempfile.TemporaryDirectory() as tmp:
        c, exe = os.path.join(tmp, "t.c"), os.path.join(tmp, "t")
        open(c, "w").write(src)
        libdir = os.path.join(ROOT, "jxpamg_all_amd")
        # the executable references no MPI symbol itself (MPI_Comm is a type);
        # the libraries bring their own dependencies
        r = subprocess.run(["gcc", "-std=gnu11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I", MPI_INC,
                            c, "-o", exe, "-L", libdir, "-ljxamd_jxfpamg", "-ljxamd", f"-Wl,-rpath,{libdir}",
                            "-Wl,--allow-shlib-undefined"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert subprocess.run([exe]).returncode == 0


@needs_mpi
def test_reference_entry_rejects_other_solvers_before_the_device():
    L = ctypes.CDLL(SHIM)
    f = L.JXFPAMG_Solver_Interface
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    ia = np.array([0, 1], np.int32)
    ja = np.array([0], np.int32)
    a = np.array([1.0])
    x, b = np.zeros(1), np.ones(1)
    it = ctypes.c_int(-1)
    comm_world = 0x44000000  # MPI_COMM_WORLD of the image's MPICH (mpi.h)
    rc = f(0, None, comm_world, 1, jx._p(ia), jx._p(ja), jx._p(a), jx._p(x), jx._p(b), 1e-6, 221, 1, ctypes.byref(it))
    assert rc == 4 and it.value == -1


def test_precond_callback_rejects_missing_context():
    L = jx.lib()
    f = L.jxamd_cpr_precond_callback
    f.argtypes = [ctypes.c_void_p] * 4
    f.restype = ctypes.c_int
    assert f(None, None, None, None) == 4
