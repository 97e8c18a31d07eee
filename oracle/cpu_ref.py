"""TEST INFRASTRUCTURE: ctypes binding of the oracle (oracle/_ref/libcpuref.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  Built by `make -C oracle cpuref`.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libcpuref.so")
MAX_LEVELS = 25


class CSR(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ncols", ctypes.c_int), ("ia", ctypes.c_void_p), ("ja", ctypes.c_void_p),
                ("a", ctypes.c_void_p)]


class Hier(ctypes.Structure):
    _fields_ = [("num_levels", ctypes.c_int), ("A", CSR * MAX_LEVELS), ("P", CSR * MAX_LEVELS),
                ("relax_threads", ctypes.c_int * MAX_LEVELS), ("poly_degree", ctypes.c_int),
                ("poly_omega", (ctypes.c_double * 8) * MAX_LEVELS)]


class BSR(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int), ("b", ctypes.c_int), ("ia", ctypes.c_void_p), ("ja", ctypes.c_void_p),
                ("a", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("k_dim", ctypes.c_int), ("max_iter", ctypes.c_int), ("is_check_restarted", ctypes.c_int),
                ("tol", ctypes.c_double), ("a_tol", ctypes.c_double), ("pressure_index", ctypes.c_int),
                ("stage1_maxit", ctypes.c_int), ("stage2_maxit", ctypes.c_int), ("stage2_weight", ctypes.c_double),
                ("stage2_omega", ctypes.c_double), ("stage2_solver_type", ctypes.c_int), ("bilu_lu", ctypes.c_void_p),
                ("bilu_dinv", ctypes.c_void_p), ("stage2_parts", ctypes.c_int), ("convfac_threshold", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("num_precond", ctypes.c_int),
                ("final_rel_res", ctypes.c_double), ("b_norm", ctypes.c_double), ("num_false_conv", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle cpuref`")
        _lib = ctypes.CDLL(LIB_PATH)
        vp, f64, i32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        for name, args, res in [
            ("cref_bsr_matvec", [vp, f64, vp, f64, vp], None),
            ("cref_csr_matvec", [vp, f64, vp, f64, vp], None),
            ("cref_csr_matvecT", [vp, f64, vp, f64, vp], None),
            ("cref_dot", [i32, vp, vp], f64),
            ("cref_relax6", [vp, vp, vp], None),
            ("cref_relax6_threads", [vp, vp, vp, i32], None),
            ("cref_relax9", [vp, vp, vp], i32),
            ("cref_poly_smooth", [vp, vp, vp, i32, vp], None),
            ("cref_gselim", [vp, vp, i32], i32),
            ("cref_bsr_block_matvec_stable", [f64, vp, vp, f64, vp, i32], None),
            ("cref_bsr_block_invmatvec_stable", [vp, vp, vp, i32], None),
            ("cref_bsr_hgs", [vp, vp, vp, vp, f64, f64, i32], None),
            ("cref_bsr_hgs_parts", [vp, vp, vp, vp, f64, f64, i32], None),
            ("cref_pamg_solve", [vp, vp, vp], i32),
            ("cref_bilu0_factor", [vp, vp, vp], i32),
            ("cref_bilu0_solve", [vp, vp, vp, vp, vp], None),
            ("cref_cpr_precond", [vp, vp, vp, vp, vp], i32),
            ("cref_gmres_solve", [vp, vp, vp, vp, vp, vp, vp], i32),
        ]:
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = res
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class Oracle:
    """Holds ctypes views of one system (BSR matrix + hierarchy) for the oracle."""

    def __init__(self, A, H=None):
        self.L = lib()
        self.A = A
        self.keep = [A.ia, A.ja, A.a]
        self.bsr = BSR(A.nb, A.b, _p(A.ia).value, _p(A.ja).value, _p(A.a).value)
        self.hier = Hier()
        if H is not None:
            self.hier.num_levels = H.num_levels
            for l, (ia, ja, a, nc) in enumerate(H.A):
                self.hier.A[l] = CSR(len(ia) - 1, nc, _p(ia).value, _p(ja).value, _p(a).value)
                self.keep += [ia, ja, a]
            for l, (ia, ja, a, nc) in enumerate(H.P):
                self.hier.P[l] = CSR(len(ia) - 1, nc, _p(ia).value, _p(ja).value, _p(a).value)
                self.keep += [ia, ja, a]

    def params(self, k_dim=30, max_iter=100, is_check_restarted=1, tol=1e-6, a_tol=0.0, pressure_index=0,
               stage1_maxit=1, stage2_maxit=3, stage2_weight=0.7, stage2_omega=0.9, stage2_solver_type=2,
               stage2_parts=0, convfac_threshold=1.0e4):
        lu = dinv = None
        if stage2_solver_type == 5:
            lu, dinv = self.bilu0()
            lu, dinv = _p(lu).value, _p(dinv).value
        return Params(k_dim, max_iter, is_check_restarted, tol, a_tol, pressure_index, stage1_maxit, stage2_maxit,
                      stage2_weight, stage2_omega, stage2_solver_type, lu, dinv, stage2_parts, convfac_threshold)

    def bilu0(self):
        """Exact block ILU(0) factors (lu on A's pattern, dinv = U_ii^{-1}), cached."""
        if getattr(self, "_bilu", None) is None:
            bb = self.A.b * self.A.b
            lu = np.zeros(self.A.nnzb * bb)
            dinv = np.zeros(self.A.nb * bb)
            rc = self.L.cref_bilu0_factor(ctypes.byref(self.bsr), _p(lu), _p(dinv))
            assert rc == 0, "singular pivot block in BILU(0)"
            self._bilu = (lu, dinv)
        return self._bilu

    def bilu0_solve(self, r):
        lu, dinv = self.bilu0()
        z = np.zeros(self.A.n)
        self.L.cref_bilu0_solve(ctypes.byref(self.bsr), _p(lu), _p(dinv), _p(np.ascontiguousarray(r)), _p(z))
        return z

    def matvec(self, alpha, x, beta, y):
        y = np.array(y, dtype=np.float64, copy=True)
        self.L.cref_bsr_matvec(ctypes.byref(self.bsr), alpha, _p(np.ascontiguousarray(x)), beta, _p(y))
        return y

    def relax6(self, level, f, u):
        u = np.array(u, dtype=np.float64, copy=True)
        self.L.cref_relax6(ctypes.byref(self.hier.A[level]), _p(np.ascontiguousarray(f)), _p(u))
        return u

    def set_relax_threads(self, per_level):
        for l, t in enumerate(per_level):
            self.hier.relax_threads[l] = int(t)

    def set_poly(self, omegas):
        """Polynomial smoother on every level: omegas[l] = step sizes w_k."""
        deg = len(omegas[0])
        self.hier.poly_degree = deg
        for l, w in enumerate(omegas):
            assert len(w) == deg
            for k, v in enumerate(w):
                self.hier.poly_omega[l][k] = float(v)

    def poly_smooth(self, level, f, u, omegas):
        u = np.array(u, dtype=np.float64, copy=True)
        w = np.ascontiguousarray(omegas, dtype=np.float64)
        self.L.cref_poly_smooth(ctypes.byref(self.hier.A[level]), _p(np.ascontiguousarray(f)), _p(u), len(w), _p(w))
        return u

    def relax6_threads(self, level, f, u, threads):
        u = np.array(u, dtype=np.float64, copy=True)
        self.L.cref_relax6_threads(ctypes.byref(self.hier.A[level]), _p(np.ascontiguousarray(f)), _p(u), threads)
        return u

    def relax9(self, f):
        L = self.hier.num_levels
        u = np.zeros(self.hier.A[L - 1].n)
        self.L.cref_relax9(ctypes.byref(self.hier.A[L - 1]), _p(np.ascontiguousarray(f)), _p(u))
        return u

    def vcycle(self, f, u=None):
        u = np.zeros(self.hier.A[0].n) if u is None else np.array(u, dtype=np.float64, copy=True)
        self.L.cref_pamg_solve(ctypes.byref(self.hier), _p(np.ascontiguousarray(f)), _p(u))
        return u

    def bhgs(self, x, rhs, w=0.7, omega=0.9, fwd=1):
        x = np.array(x, dtype=np.float64, copy=True)
        work = np.zeros_like(x)
        self.L.cref_bsr_hgs(ctypes.byref(self.bsr), _p(x), _p(np.ascontiguousarray(rhs)), _p(work), w, omega, fwd)
        return x

    def bhgs_parts(self, x, rhs, nparts, w=0.7, omega=0.9):
        x = np.array(x, dtype=np.float64, copy=True)
        work = np.zeros_like(x)
        self.L.cref_bsr_hgs_parts(ctypes.byref(self.bsr), _p(x), _p(np.ascontiguousarray(rhs)), _p(work), w, omega,
                                  int(nparts))
        return x

    def cpr(self, r, params=None):
        p = params if params is not None else self.params()
        x = np.zeros(self.A.n)
        rc = self.L.cref_cpr_precond(ctypes.byref(self.bsr), ctypes.byref(self.hier), ctypes.byref(p),
                                     _p(np.ascontiguousarray(r)), _p(x))
        assert rc == 0
        return x

    def gmres(self, b, params=None, x0=None):
        p = params if params is not None else self.params()
        x = np.zeros(self.A.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        hist = np.zeros(p.max_iter + 1)
        st = Stats()
        rc = self.L.cref_gmres_solve(ctypes.byref(self.bsr), ctypes.byref(self.hier), ctypes.byref(p),
                                     _p(np.ascontiguousarray(b)), _p(x), _p(hist), ctypes.byref(st))
        return x, st, hist[: st.iterations + 1], rc
