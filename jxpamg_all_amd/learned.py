"""Learned polynomial AMG smoother: data from the plain solver, GPU training.

The network (csrc/hip/jxamd_mlp.hip) maps 16 features of an AMG level to the
m roots (m <= 4, its first m outputs; one network per degree) of the level's
degree-m polynomial smoother in units of lambda_max (steps
w_k = 1 / (tau_k lambda_max)).  Training data come from the solver itself: on
small seeded systems (iid lognormal permeability as in the 16x16x8 case, and
the layered smooth-over-channel fields with wells of the large cases, a few
grid sizes so that coarse levels appear too) the roots of each level are
chosen, level by level, from Chebyshev candidates on [lo, hi] by the V-cycle
contraction factor the solver measures on the GPU; the network learns
features -> the best roots.  No reference counterpart; the solver judges
the predicted roots level by level against the plain smoother (relax 6),
which is also the fallback (jxamd_solver_apply_mlp).
"""
import ctypes
import math

import numpy as np

from . import CONFIGS, FIELD_IID, FIELD_LAYERED, Solver, amg_setup, check, decouple_abf, default_params, extract_pressure, \
    generate, lib

NPARAM, NFEAT, NOUT = 5508, 16, 4
LO = (0.03, 0.06, 0.1, 0.15, 0.2, 0.3, 0.45)
HI = (0.9, 1.0, 1.1, 1.25, 1.4)


def chebyshev_steps(lo, hi, m=NOUT):
    """Step sizes (in units of 1/lambda_max) with roots of the degree-m
    Chebyshev polynomial on [lo, hi]."""
    c, d = 0.5 * (hi + lo), 0.5 * (hi - lo)
    return np.array([1.0 / (c + d * math.cos(math.pi * (2 * k + 1) / (2 * m))) for k in range(m)])


def training_system(seed, nx=16, ny=16, nz=8, layered=False):
    """A small system: iid lognormal permeability (sigma 1.5), or the layered
    field of the large cases (smooth lognormal over sinuous channels, wells)
    scaled to the grid."""
    if layered:
        kw = dict(CONFIGS["C2"], nx=nx, ny=ny, nz=nz, seed=seed, n_smooth_layers=max(1, nz * 35 // 85),
                  corr_len=4.0, n_wells=2)
    else:
        kw = dict(CONFIGS["C1"], nx=nx, ny=ny, nz=nz, field=FIELD_IID, seed=seed)
    A, rhs = generate(**kw)
    Ad, bd = decouple_abf(A, rhs)
    H = amg_setup(*extract_pressure(Ad))
    return Ad, bd, H


def level_targets(solver, ncycles=4, m=NOUT):
    """Greedy per-level search of the best Chebyshev roots of the degree-m
    smoother by the measured V-cycle contraction; returns (features [L][16],
    roots [L][m], best factor, default factor)."""
    Ls = max(solver.num_levels - 1, 1)
    f_def = solver.vcycle_factor(ncycles)
    feats, targets = [], []
    for l in range(Ls):
        w0, lam = solver.get_poly(l, m)
        best = (solver.vcycle_factor(ncycles), w0 * lam)
        for lo in LO:
            for hi in HI:
                z = chebyshev_steps(lo, hi, m)
                solver.set_poly(l, z / lam)
                f = solver.vcycle_factor(ncycles)
                if f < best[0]:
                    best = (f, z)
        solver.set_poly(l, best[1] / lam)
        feats.append(solver.poly_features(l))
        targets.append(1.0 / best[1])
    return np.array(feats, np.float32), np.array(targets, np.float32), best[0], f_def


# (nx, ny, nz, layered): iid grids as in the 16x16x8 case, layered grids with
# the 60 : 220 : 85 aspect of the large cases at 1/3 to 3/5 of their size
GRIDS = ((16, 16, 8, False), (30, 110, 42, True), (24, 24, 12, False), (20, 74, 28, True), (20, 20, 10, False),
         (36, 132, 51, True), (32, 16, 8, False), (24, 88, 34, True))


def make_dataset(seeds, precision=1, grids=GRIDS, m=NOUT):
    """Labelled levels (features, the best degree-m roots) of small systems,
    grid sizes and fields cycling over grids."""
    X, T, info = [], [], []
    for q, sd in enumerate(seeds):
        nx, ny, nz, layered = grids[q % len(grids)]
        Ad, bd, H = training_system(sd, nx=nx, ny=ny, nz=nz, layered=layered)
        s = Solver(Ad, H, default_params(poly_degree=m, precond_precision=precision))
        f, t, fb, fd = level_targets(s, m=m)
        s.close()
        X.append(f)
        T.append(t)
        info.append((sd, fb, fd))
    return np.concatenate(X), np.concatenate(T), info


class MLP:
    """FP32 master weights, Adam state and gradient in HBM (torch tensors as
    storage); every forward, backward and update runs in the HIP kernels."""

    def __init__(self, seed=0, device="cuda"):
        import torch
        self.torch = torch
        rng = np.random.default_rng(seed)
        P = np.zeros(NPARAM, np.float32)
        o = 0
        for shp in ((64, 16), (64,), (64, 64), (64,), (4, 64), (4,)):
            n = int(np.prod(shp))
            if len(shp) == 2:
                P[o:o + n] = rng.normal(0.0, 1.0 / math.sqrt(shp[1]), n)
            o += n
        self.P = torch.tensor(P, device=device)
        self.M = torch.zeros_like(self.P)
        self.V = torch.zeros_like(self.P)
        self.G = torch.zeros_like(self.P)
        self.t = 0

    def _dev(self, a, cols):
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1, cols)
        return self.torch.tensor(a, device=self.P.device)

    def forward(self, X):
        x = self._dev(X, NFEAT)
        y = self.torch.zeros((len(x), NOUT), dtype=self.torch.float32, device=self.P.device)
        check(lib().jxamd_mlp_forward(ctypes.c_void_p(self.P.data_ptr()), ctypes.c_void_p(x.data_ptr()), len(x),
                                      ctypes.c_void_p(y.data_ptr()), None), "jxamd_mlp_forward")
        self.torch.cuda.synchronize()
        return y.cpu().numpy()

    def _train(self, X, T, steps, lr, update):
        nout = np.asarray(T).reshape(len(np.asarray(X).reshape(-1, NFEAT)), -1).shape[1]
        x, t = self._dev(X, NFEAT), self._dev(T, nout)
        loss = self.torch.zeros(steps, dtype=self.torch.float32, device=self.P.device)
        check(lib().jxamd_mlp_train(ctypes.c_void_p(self.P.data_ptr()), ctypes.c_void_p(self.M.data_ptr()),
                                    ctypes.c_void_p(self.V.data_ptr()), ctypes.c_void_p(self.G.data_ptr()),
                                    ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(t.data_ptr()), len(x), nout,
                                    steps, self.t, ctypes.c_float(lr), int(update),
                                    ctypes.c_void_p(loss.data_ptr()), None),
              "jxamd_mlp_train")
        self.torch.cuda.synchronize()
        if update:
            self.t += steps
        return loss.cpu().numpy()

    def gradient(self, X, T):
        """(loss, gradient) of one batch (first 32 samples), no update; T
        holds the targets of the first T.shape[1] outputs."""
        loss = self._train(X, T, 1, 0.0, False)
        return float(loss[0]), self.G.cpu().numpy()

    def fit(self, X, T, steps=3000, lr=3e-3):
        return self._train(X, T, steps, lr, True)

    @property
    def params(self):
        return self.P.cpu().numpy()

    def load(self, P):
        self.P.copy_(self.torch.tensor(np.asarray(P, np.float32)))
