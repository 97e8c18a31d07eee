"""MI355X-native CPR-GMRES solve phase (host bindings).

The product is the C-ABI shared library ``libjxamd.so`` (HIP kernels for
gfx950 + host C), declared in ``include/jxamd.h``.  This module loads it with
ctypes and offers thin Python wrappers used by the tests and the benchmark.
Device vectors are passed as raw pointers (``torch.Tensor.data_ptr()``);
torch only provides device memory.

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjxamd.so")
MAX_LEVELS = 25

SUCCESS, ERROR_GENERIC, ERROR_MEMORY, ERROR_ARG, ERROR_CONV, ERROR_DEVICE = 0, 1, 2, 4, 256, 512
FIELD_IID, FIELD_SMOOTH, FIELD_LAYERED = 0, 1, 2


class GenParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("block_size", ctypes.c_int),
                ("field", ctypes.c_int), ("sigma_lnk", ctypes.c_double), ("corr_len", ctypes.c_double),
                ("n_smooth_layers", ctypes.c_int), ("channel_ratio", ctypes.c_double),
                ("kz_ratio", ctypes.c_double), ("n_wells", ctypes.c_int), ("well_boost", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class BSR(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int), ("ncols", ctypes.c_int), ("nnzb", ctypes.c_int), ("b", ctypes.c_int),
                ("ia", ctypes.c_void_p), ("ja", ctypes.c_void_p), ("a", ctypes.c_void_p)]


class CSR(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ncols", ctypes.c_int), ("nnz", ctypes.c_int),
                ("ia", ctypes.c_void_p), ("ja", ctypes.c_void_p), ("a", ctypes.c_void_p)]


class HierarchyC(ctypes.Structure):
    _fields_ = [("num_levels", ctypes.c_int), ("A", CSR * MAX_LEVELS), ("P", CSR * MAX_LEVELS)]


class AMGParams(ctypes.Structure):
    _fields_ = [("max_levels", ctypes.c_int), ("coarse_threshold", ctypes.c_int),
                ("strong_threshold", ctypes.c_double), ("max_row_sum", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("k_dim", ctypes.c_int), ("max_iter", ctypes.c_int), ("is_check_restarted", ctypes.c_int),
                ("tol", ctypes.c_double), ("a_tol", ctypes.c_double),
                ("pressure_index", ctypes.c_int), ("stage1_maxit", ctypes.c_int), ("stage2_maxit", ctypes.c_int),
                ("stage1_solver_type", ctypes.c_int), ("stage2_solver_type", ctypes.c_int),
                ("print_level", ctypes.c_int),
                ("threshold", ctypes.c_double), ("stage2_weight", ctypes.c_double),
                ("stage2_omega", ctypes.c_double), ("relax_threads", ctypes.c_int), ("relax_chunk", ctypes.c_int),
                ("precond_precision", ctypes.c_int), ("poly_degree", ctypes.c_int), ("stage2_chunk", ctypes.c_int),
                ("stage2_precision", ctypes.c_int), ("arnoldi_kernel", ctypes.c_int),
                ("solution_update", ctypes.c_int), ("convfac_threshold", ctypes.c_double),
                ("coarse_rows", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("num_precond", ctypes.c_int),
                ("num_restarts", ctypes.c_int), ("final_rel_res", ctypes.c_double),
                ("true_rel_res", ctypes.c_double), ("b_norm", ctypes.c_double),
                ("t_solve_ms", ctypes.c_double), ("t_precond_ms", ctypes.c_double),
                ("num_false_conv", ctypes.c_int), ("diverged", ctypes.c_int)]


_lib = None


def lib():
    """The loaded libjxamd.so (raises if it has not been built).

    torch, where installed, is imported first: its wheel carries its own HIP
    runtime, and libjxamd.so must bind to that one (the dynamic linker then
    finds libamdhip64 already loaded).  Loaded before torch, the library would
    bring up the system HIP runtime beside torch's, two runtimes with their
    own queues in one process: device memory from torch and the library's
    synchronisations would no longer refer to the same device state."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no CPU fallback)")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sigs = {
        "jxamd_gen_sizes": [vp, vp, vp],
        "jxamd_gen_reservoir": [vp, vp, vp, vp, vp],
        "jxamd_bsr_decouple_abf": [i32, i32, vp, vp, vp, vp],
        "jxamd_bsr_decouple_abf_exact": [i32, i32, vp, vp, vp, vp],
        "jxamd_bsr_decouple": [i32, i32, vp, vp, vp, vp, i32, i32],
        "jxamd_bsr_decouple_device": [i32, i32, vp, vp, vp, vp, i32, i32, i32, vp, vp],
        "jxamd_bsr_extract_pressure_device": [i32, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp],
        "jxamd_bsr_extract_pressure": [i32, i32, vp, vp, vp, i32, vp, vp, vp],
        "jxamd_bsr_write_binary": [ctypes.c_char_p, i32, i32, i32, i32, vp, vp, vp],
        "jxamd_bsr_read_binary_sizes": [ctypes.c_char_p, vp, vp, vp, vp],
        "jxamd_bsr_read_binary": [ctypes.c_char_p, vp, vp, vp],
        "jxamd_rhs_write_text": [ctypes.c_char_p, i32, vp],
        "jxamd_rhs_read_text": [ctypes.c_char_p, i32, vp],
        "jxamd_params_default": [vp],
        "jxamd_solver_create": [vp, vp, vp, vp],
        "jxamd_solver_destroy": [vp],
        "jxamd_solver_solve": [vp, vp, vp, vp, vp],
        "jxamd_solver_solve_device": [vp, vp, vp, vp, vp],
        "jxamd_cpr_precond_device": [vp, vp, vp],
        "jxamd_cpr_precond_callback": [vp, vp, vp, vp],
        "jxamd_op_bsr_matvec": [vp, f64, vp, f64, vp],
        "jxamd_op_hsgs": [vp, i32, vp, vp],
        "jxamd_op_coarse_solve": [vp, vp, vp],
        "jxamd_op_vcycle": [vp, vp, vp],
        "jxamd_op_bhgs": [vp, vp, vp, f64, f64],
        "jxamd_solver_error_flag": [vp, vp],
        "jxamd_solver_set_spin_limit": [vp, i32],
        "jxamd_op_bsr_matvec_timed": [vp, vp, vp, i32, vp],
        "jxamd_amg_params_default": [vp],
        "jxamd_solver_profile": [vp, i32],
        "jxamd_solver_profile_read": [vp, i32, vp, vp, vp],
        "jxamd_solver_profile_read_level": [vp, i32, vp, vp, vp],
        "jxamd_amg_setup": [vp, vp, vp],
        "jxamd_hierarchy_free": [vp],
        "jxamd_cpr_precond_timed": [vp, vp, vp, i32, vp],
        "jxamd_solver_interface": [i32, vp, i32, vp, vp, vp, vp, vp, f64, i32, i32, vp],
        "jxamd_op_bilu_solve": [vp, vp, vp],
        "jxamd_solver_set_poly": [vp, i32, i32, vp],
        "jxamd_solver_get_poly": [vp, i32, i32, vp, vp],
        "jxamd_solver_num_levels": [vp, vp],
        "jxamd_mlp_forward": [vp, vp, i32, vp, vp],
        "jxamd_mlp_train": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, ctypes.c_float, i32, vp, vp],
        "jxamd_solver_poly_features": [vp, i32, vp],
        "jxamd_solver_vcycle_factor": [vp, i32, vp],
        "jxamd_solver_apply_mlp": [vp, vp, vp, vp, vp],
        "jxamd_solver_judge_poly": [vp, vp, vp, vp],
        "jxamd_solver_set_level_smoother": [vp, i32, i32],
        "jxamd_solver_get_level_smoother": [vp, i32, vp],
        "jxamd_bilu0_factor": [i32, i32, vp, vp, vp, vp, vp],
        "jxamd_bsr_residual_norm": [i32, i32, vp, vp, vp, vp, vp, vp, vp],
        "jxamd_bsr_residual_norm_noalloc": [i32, i32, vp, vp, vp, vp, vp, vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = None if name in ("jxamd_params_default", "jxamd_amg_params_default",
                                     "jxamd_hierarchy_free") else ctypes.c_int


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with error flag {rc}")


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
CONFIGS = {
    # 16x16x8, 3x3 FP64, ln k ~ N(0, 1.5^2) iid
    "C1": dict(nx=16, ny=16, nz=8, block_size=3, field=FIELD_IID, sigma_lnk=1.5, corr_len=0.0,
               n_smooth_layers=0, channel_ratio=1.0, kz_ratio=0.1, n_wells=0, well_boost=0.0, seed=1),
    # 60x220x85: top 35 layers smooth lognormal (sigma 1, corr 8), bottom 50 sinuous channels (k ratio 1e3),
    # 5 well columns boosting the pressure diagonal by 1e4 x mean T
    "C2": dict(nx=60, ny=220, nz=85, block_size=3, field=FIELD_LAYERED, sigma_lnk=1.0, corr_len=8.0,
               n_smooth_layers=35, channel_ratio=1e3, kz_ratio=0.1, n_wells=5, well_boost=1e4, seed=2),
    # 100^3, 4x4 blocks, smooth lognormal sigma 2, corr 5
    "C4": dict(nx=100, ny=100, nz=100, block_size=4, field=FIELD_SMOOTH, sigma_lnk=2.0, corr_len=5.0,
               n_smooth_layers=0, channel_ratio=1.0, kz_ratio=0.1, n_wells=0, well_boost=0.0, seed=4),
    # 120x440x85 layered (C2 style, 4x larger)
    "C5": dict(nx=120, ny=440, nz=85, block_size=3, field=FIELD_LAYERED, sigma_lnk=1.0, corr_len=8.0,
               n_smooth_layers=35, channel_ratio=1e3, kz_ratio=0.1, n_wells=5, well_boost=1e4, seed=2001),
}


# rows per relax-6 chunk on the GPU for the large workloads: the reference's
# threaded hybrid SSOR (par_relax_6.c) with ceil(n_l / 4) chunks on level l
# (on the 60x220x85 system: 9 iterations against 8 for one thread, and the
# fastest of chunk lengths 1..16)
BENCH_RELAX_CHUNK = 4
# block rows per stage-2 partition on the GPU for the large workloads: the
# stage-2 sweeps as run on ceil(nb / 4) ranks (par_relax_bsr.c), 9 iterations
# on the 60x220x85 system for partitions of 4..64 rows
BENCH_STAGE2_CHUNK = 4


def gen_params(**kw):
    d = dict(CONFIGS["C1"])
    d.update(kw)
    return GenParams(**d)


class BSRMatrix:
    """Host BSR system (0-based int32 ia/ja, row-major b x b float64 blocks)."""

    def __init__(self, nb, b, ia, ja, a):
        self.nb, self.b = int(nb), int(b)
        self.ia = np.ascontiguousarray(ia, dtype=np.int32)
        self.ja = np.ascontiguousarray(ja, dtype=np.int32)
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.nnzb = int(self.ia[-1])

    @property
    def n(self):
        return self.nb * self.b

    def c_struct(self):
        return BSR(self.nb, self.nb, self.nnzb, self.b, _p(self.ia).value, _p(self.ja).value, _p(self.a).value)

    def to_scipy(self):
        import scipy.sparse as sp
        b = self.b
        return sp.bsr_matrix((self.a.reshape(-1, b, b), self.ja, self.ia), shape=(self.n, self.n)).tocsr()


def generate(**kw):
    """Seeded synthetic reservoir system -> (BSRMatrix, rhs)."""
    L = lib()
    p = gen_params(**kw)
    nb, nnzb = ctypes.c_int(), ctypes.c_int()
    check(L.jxamd_gen_sizes(ctypes.byref(p), ctypes.byref(nb), ctypes.byref(nnzb)), "jxamd_gen_sizes")
    b = p.block_size
    ia = np.zeros(nb.value + 1, np.int32)
    ja = np.zeros(nnzb.value, np.int32)
    a = np.zeros(nnzb.value * b * b)
    rhs = np.zeros(nb.value * b)
    check(L.jxamd_gen_reservoir(ctypes.byref(p), _p(ia), _p(ja), _p(a), _p(rhs)), "jxamd_gen_reservoir")
    return BSRMatrix(nb.value, b, ia, ja, a), rhs


def solver_interface(row, ia, ja, a, ser_x, ser_b, tol, solver_id, num_functions, argv=()):
    """JXFPAMG_Solver_Interface (src/jxpamg_interface.c:118-257), solver_id 220.
    ia/ja/a are numpy arrays modified in place as by the reference (1-based
    indices decremented, diagonal swapped to the front of each row); the
    solution is written into ser_x.  Returns the iteration count."""
    for arr, dt in ((ia, np.int32), (ja, np.int32), (a, np.float64), (ser_x, np.float64), (ser_b, np.float64)):
        if arr.dtype != dt or not arr.flags.c_contiguous:
            raise TypeError("solver_interface works in place on contiguous int32 / float64 arrays")
    args = [b"jxamd"] + [str(v).encode() for v in argv]
    cargv = (ctypes.c_char_p * (len(args) + 1))(*args, None)
    it = ctypes.c_int(0)
    check(lib().jxamd_solver_interface(len(args), ctypes.cast(cargv, ctypes.c_void_p), int(row), _p(ia), _p(ja),
                                       _p(a), _p(ser_x), _p(ser_b), float(tol), int(solver_id), int(num_functions),
                                       ctypes.byref(it)), "jxamd_solver_interface")
    return it.value


def bsr_residual_norm(nb, b, d_ia, d_ja, d_a, d_x, d_b, d_norms, stream=None):
    """d_norms[0:2] = (||b - A x||, ||b||) for device BSR arrays (raw device
    addresses, reference layout); pgmres.c:1505-1511.  Synchronous."""
    check(lib().jxamd_bsr_residual_norm(int(nb), int(b), d_ia, d_ja, d_a, d_x, d_b, d_norms, stream),
          "jxamd_bsr_residual_norm")


def bilu0_factor(A):
    """Exact block ILU(0) factors of a BSR matrix: (lu on A's pattern, dinv)."""
    bb = A.b * A.b
    lu = np.zeros(A.nnzb * bb)
    dinv = np.zeros(A.nb * bb)
    check(lib().jxamd_bilu0_factor(A.nb, A.b, _p(A.ia), _p(A.ja), _p(A.a), _p(lu), _p(dinv)), "jxamd_bilu0_factor")
    return lu, dinv


def decouple_abf(A, rhs=None, exact=False):
    """ABF decoupling (copy); returns (A_dec, rhs_dec).  exact=False is the
    reference's arithmetic bit for bit; exact=True uses true block inverses
    and an out-of-place rhs product (the reference's 4x4 inverse is wrong)."""
    a = A.a.copy()
    r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).copy()
    fn = "jxamd_bsr_decouple_abf_exact" if exact else "jxamd_bsr_decouple_abf"
    check(getattr(lib(), fn)(A.nb, A.b, _p(A.ia), _p(A.ja), _p(a), None if r is None else _p(r)), fn)
    return BSRMatrix(A.nb, A.b, A.ia, A.ja, a), r


DECOUP_NONE, DECOUP_ABF, DECOUP_ANL, DECOUP_SEM, DECOUP_QI, DECOUP_TIMPES, DECOUP_TIMPES2 = range(7)


def decouple(A, rhs=None, dtype=DECOUP_ABF, is_thermal=0):
    """jxf_BSRMatrixDecouple(A, rhs, dtype, is_thermal) on a copy, bit for bit
    with the reference (jxf_bsr_decoup.c:1273-1348); returns (A_dec, rhs_dec)."""
    a = A.a.copy()
    r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).copy()
    check(lib().jxamd_bsr_decouple(A.nb, A.b, _p(A.ia), _p(A.ja), _p(a), None if r is None else _p(r), int(dtype),
                                   int(is_thermal)), "jxamd_bsr_decouple")
    return BSRMatrix(A.nb, A.b, A.ia, A.ja, a), r


def decouple_device(nb, b, d_ia, d_ja, d_a, d_rhs, dtype=DECOUP_ABF, is_thermal=0, pidx=0, d_p_a=None, stream=None):
    """In-place decoupling of device BSR arrays (raw device addresses); with
    d_p_a, also the decoupled pressure values.  Returns the error flag (0 ok,
    4 for a pattern True-IMPES cannot transpose)."""
    return lib().jxamd_bsr_decouple_device(int(nb), int(b), d_ia, d_ja, d_a, d_rhs, int(dtype), int(is_thermal),
                                           int(pidx), d_p_a, stream)


def extract_pressure(A, pidx=0):
    """(ia, ja, a) of the pressure matrix A_p."""
    ia = np.zeros(A.nb + 1, np.int32)
    ja = np.zeros(A.nnzb, np.int32)
    a = np.zeros(A.nnzb)
    check(lib().jxamd_bsr_extract_pressure(A.nb, A.b, _p(A.ia), _p(A.ja), _p(A.a), pidx, _p(ia), _p(ja), _p(a)),
          "jxamd_bsr_extract_pressure")
    return ia, ja, a


def write_bsr_binary(path, A):
    check(lib().jxamd_bsr_write_binary(path.encode(), A.nb, A.nb, A.nnzb, A.b, _p(A.ia), _p(A.ja), _p(A.a)),
          "jxamd_bsr_write_binary")


def read_bsr_binary(path):
    L = lib()
    nb, nc, nnzb, b = (ctypes.c_int() for _ in range(4))
    check(L.jxamd_bsr_read_binary_sizes(path.encode(), ctypes.byref(nb), ctypes.byref(nc), ctypes.byref(nnzb),
                                        ctypes.byref(b)), "jxamd_bsr_read_binary_sizes")
    ia = np.zeros(nb.value + 1, np.int32)
    ja = np.zeros(nnzb.value, np.int32)
    a = np.zeros(nnzb.value * b.value * b.value)
    check(L.jxamd_bsr_read_binary(path.encode(), _p(ia), _p(ja), _p(a)), "jxamd_bsr_read_binary")
    return BSRMatrix(nb.value, b.value, ia, ja, a)


def write_rhs_text(path, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    check(lib().jxamd_rhs_write_text(path.encode(), len(v), _p(v)), "jxamd_rhs_write_text")


def read_rhs_text(path, n):
    v = np.zeros(n)
    check(lib().jxamd_rhs_read_text(path.encode(), n, _p(v)), "jxamd_rhs_read_text")
    return v


class Hierarchy:
    """Pressure-AMG hierarchy: lists of (ia, ja, a, ncols) for A_l and P_l."""

    def __init__(self, A_levels, P_levels):
        self.A = [tuple(np.ascontiguousarray(x, dt) for x, dt in zip(m[:3], (np.int32, np.int32, np.float64)))
                  + (int(m[3]),) for m in A_levels]
        self.P = [tuple(np.ascontiguousarray(x, dt) for x, dt in zip(m[:3], (np.int32, np.int32, np.float64)))
                  + (int(m[3]),) for m in P_levels]

    @property
    def num_levels(self):
        return len(self.A)

    def c_struct(self):
        h = HierarchyC()
        h.num_levels = len(self.A)
        for l, (ia, ja, a, nc) in enumerate(self.A):
            h.A[l] = CSR(len(ia) - 1, nc, int(ia[-1]), _p(ia).value, _p(ja).value, _p(a).value)
        for l, (ia, ja, a, nc) in enumerate(self.P):
            h.P[l] = CSR(len(ia) - 1, nc, int(ia[-1]), _p(ia).value, _p(ja).value, _p(a).value)
        return h

    @classmethod
    def from_fixture(cls, d):
        L = int(d["num_levels"][0])
        As, Ps = [], []
        for l in range(L):
            n, m, nnz = d[f"A{l}_shape"]
            As.append((d[f"A{l}_ia"], d[f"A{l}_ja"], d[f"A{l}_a"], m))
            if l < L - 1:
                n, m, nnz = d[f"P{l}_shape"]
                Ps.append((d[f"P{l}_ia"], d[f"P{l}_ja"], d[f"P{l}_a"], m))
        return cls(As, Ps)


def amg_setup(ia, ja, a, **kw):
    """Pressure-AMG hierarchy built on the host (jxamd_amg_setup) -> Hierarchy."""
    L = lib()
    p = AMGParams()
    L.jxamd_amg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    n = len(ia) - 1
    A0 = CSR(n, n, int(ia[-1]), _p(ia).value, _p(ja).value, _p(a).value)
    h = HierarchyC()
    check(L.jxamd_amg_setup(ctypes.byref(A0), ctypes.byref(p), ctypes.byref(h)), "jxamd_amg_setup")
    try:
        def grab(c):
            m, nnz, nc = c.n, c.nnz, c.ncols
            gi = np.ctypeslib.as_array(ctypes.cast(c.ia, ctypes.POINTER(ctypes.c_int)), (m + 1,)).copy()
            gj = np.ctypeslib.as_array(ctypes.cast(c.ja, ctypes.POINTER(ctypes.c_int)), (max(nnz, 1),))[:nnz].copy()
            ga = np.ctypeslib.as_array(ctypes.cast(c.a, ctypes.POINTER(ctypes.c_double)), (max(nnz, 1),))[:nnz].copy()
            return (gi, gj, ga, nc)
        As = [grab(h.A[l]) for l in range(h.num_levels)]
        Ps = [grab(h.P[l]) for l in range(h.num_levels - 1)]
    finally:
        L.jxamd_hierarchy_free(ctypes.byref(h))
    return Hierarchy(As, Ps)


def default_params(**kw):
    p = Params()
    lib().jxamd_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Solver:
    """Device-resident CPR-GMRES solver (jxamd_solver_*)."""

    def __init__(self, A, H, params=None):
        self._L = lib()
        self.A, self.H = A, H
        self.params = params if params is not None else default_params()
        self._a = A.c_struct()
        self._h = H.c_struct()
        self._s = ctypes.c_void_p()
        check(self._L.jxamd_solver_create(ctypes.byref(self._s), ctypes.byref(self._a), ctypes.byref(self._h),
                                          ctypes.byref(self.params)), "jxamd_solver_create")

    def close(self):
        if self._s:
            self._L.jxamd_solver_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, x0=None, history=True):
        x = np.zeros(self.A.n) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        hist = np.zeros(self.params.max_iter + 1)
        st = Stats()
        rc = self._L.jxamd_solver_solve(self._s, _p(b), _p(x), _p(hist) if history else None, ctypes.byref(st))
        check(rc, "jxamd_solver_solve")
        return x, st, hist[: st.iterations + 1]

    def solve_device(self, d_b, d_x, hist=None):
        st = Stats()
        rc = self._L.jxamd_solver_solve_device(self._s, ctypes.c_void_p(d_b), ctypes.c_void_p(d_x),
                                               None if hist is None else _p(hist), ctypes.byref(st))
        check(rc, "jxamd_solver_solve_device")
        return st

    def precond(self, d_r, d_x):
        check(self._L.jxamd_cpr_precond_device(self._s, ctypes.c_void_p(d_r), ctypes.c_void_p(d_x)),
              "jxamd_cpr_precond_device")

    def bsr_matvec(self, alpha, d_x, beta, d_y):
        check(self._L.jxamd_op_bsr_matvec(self._s, alpha, ctypes.c_void_p(d_x), beta, ctypes.c_void_p(d_y)),
              "jxamd_op_bsr_matvec")

    def hsgs(self, level, d_f, d_u):
        check(self._L.jxamd_op_hsgs(self._s, level, ctypes.c_void_p(d_f), ctypes.c_void_p(d_u)), "jxamd_op_hsgs")

    def coarse_solve(self, d_f, d_u):
        check(self._L.jxamd_op_coarse_solve(self._s, ctypes.c_void_p(d_f), ctypes.c_void_p(d_u)),
              "jxamd_op_coarse_solve")

    def vcycle(self, d_f, d_u):
        check(self._L.jxamd_op_vcycle(self._s, ctypes.c_void_p(d_f), ctypes.c_void_p(d_u)), "jxamd_op_vcycle")

    @property
    def num_levels(self):
        n = ctypes.c_int()
        check(self._L.jxamd_solver_num_levels(self._s, ctypes.byref(n)), "jxamd_solver_num_levels")
        return n.value

    def set_poly(self, level, omega):
        w = np.ascontiguousarray(omega, dtype=np.float64)
        check(self._L.jxamd_solver_set_poly(self._s, level, len(w), _p(w)), "jxamd_solver_set_poly")

    def get_poly(self, level, degree):
        w = np.zeros(degree)
        lam = ctypes.c_double()
        check(self._L.jxamd_solver_get_poly(self._s, level, degree, _p(w), ctypes.byref(lam)), "jxamd_solver_get_poly")
        return w, lam.value

    def poly_features(self, level):
        f = np.zeros(16, np.float32)
        check(self._L.jxamd_solver_poly_features(self._s, level, _p(f)), "jxamd_solver_poly_features")
        return f

    def vcycle_factor(self, ncycles=4):
        v = ctypes.c_double()
        check(self._L.jxamd_solver_vcycle_factor(self._s, ncycles, ctypes.byref(v)), "jxamd_solver_vcycle_factor")
        return v.value

    def apply_mlp(self, d_params):
        """Learned polynomial roots, judged against relax 6: each level keeps them
        only where they lower the V-cycle contraction, else falls back to relax 6.
        Returns (levels on learned steps, final factor, factor of relax 6 everywhere)."""
        used, fl, fp = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(self._L.jxamd_solver_apply_mlp(self._s, ctypes.c_void_p(d_params), ctypes.byref(used), ctypes.byref(fl),
                                             ctypes.byref(fp)), "jxamd_solver_apply_mlp")
        return used.value, fl.value, fp.value

    def judge_poly(self):
        """The polynomial steps in use (default Chebyshev roots unless set),
        judged against relax 6 level by level as apply_mlp does:
        (levels on the polynomial, final factor, factor of relax 6 everywhere)."""
        used, f, fp = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(self._L.jxamd_solver_judge_poly(self._s, ctypes.byref(used), ctypes.byref(f), ctypes.byref(fp)),
              "jxamd_solver_judge_poly")
        return used.value, f.value, fp.value

    def set_level_smoother(self, level, plain):
        check(self._L.jxamd_solver_set_level_smoother(self._s, int(level), int(plain)),
              "jxamd_solver_set_level_smoother")

    def level_smoothers(self):
        """per level: "R" relax 6, "D" a polynomial on the default (or set)
        roots, "L" a polynomial on the network's roots, "C" the coarsest
        level (a direct solve)"""
        out = []
        for l in range(self.num_levels):
            v = ctypes.c_int()
            check(self._L.jxamd_solver_get_level_smoother(self._s, l, ctypes.byref(v)), "get_level_smoother")
            out.append("DRLC"[v.value])
        return "".join(out)

    def bilu_solve(self, d_r, d_z):
        check(self._L.jxamd_op_bilu_solve(self._s, ctypes.c_void_p(d_r), ctypes.c_void_p(d_z)), "jxamd_op_bilu_solve")

    def bhgs(self, d_x, d_b, w, omega):
        check(self._L.jxamd_op_bhgs(self._s, ctypes.c_void_p(d_x), ctypes.c_void_p(d_b), w, omega),
              "jxamd_op_bhgs")

    def bsr_matvec_timed(self, d_x, d_y, reps):
        ms = ctypes.c_double()
        check(self._L.jxamd_op_bsr_matvec_timed(self._s, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), reps,
                                                ctypes.byref(ms)), "jxamd_op_bsr_matvec_timed")
        return ms.value

    def precond_timed(self, d_r, d_x, reps):
        ms = ctypes.c_double()
        check(self._L.jxamd_cpr_precond_timed(self._s, ctypes.c_void_p(d_r), ctypes.c_void_p(d_x), reps,
                                              ctypes.byref(ms)), "jxamd_cpr_precond_timed")
        return ms.value

    PROF_KINDS = ["gs_sweep", "bhgs_sweep", "bsr_spmv", "csr_spmv", "coarse_solve", "mgs", "vector_ops", "bilu_solve"]

    def profile(self, enable=True):
        check(self._L.jxamd_solver_profile(self._s, 1 if enable else 0), "jxamd_solver_profile")

    def profile_read(self):
        """{kind: (launches, total_ms, total_algorithmic_bytes)}"""
        out = {}
        for k, name in enumerate(self.PROF_KINDS):
            n, ms, by = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
            check(self._L.jxamd_solver_profile_read(self._s, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)),
                  "jxamd_solver_profile_read")
            out[name] = (n.value, ms.value, by.value)
        return out

    def profile_read_levels(self):
        """{level: (launches, total_ms, total_algorithmic_bytes)} of the
        smoother launches (one kernel per level), levels with launches only"""
        out = {}
        for l in range(self.num_levels):
            n, ms, by = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
            check(self._L.jxamd_solver_profile_read_level(self._s, l, ctypes.byref(n), ctypes.byref(ms),
                                                          ctypes.byref(by)), "jxamd_solver_profile_read_level")
            if n.value:
                out[l] = (n.value, ms.value, by.value)
        return out

    def set_spin_limit(self, limit):
        check(self._L.jxamd_solver_set_spin_limit(self._s, int(limit)), "jxamd_solver_set_spin_limit")

    def error_flag(self):
        f = ctypes.c_int()
        check(self._L.jxamd_solver_error_flag(self._s, ctypes.byref(f)), "jxamd_solver_error_flag")
        return f.value
