"""GPU learned polynomial smoother: the BF16 MFMA network against the NumPy
checker (forward values and gradients), training on the GPU, and what the
trained network gains over the default Chebyshev steps on unseen systems
(with the fallback exercised)."""
import numpy as np
import pytest

import jxpamg_all_amd as jx
from oracle import mlp_ref
from tests._common import relerr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
L = pytest.importorskip("jxpamg_all_amd.learned")

FWD_TOL = 1e-2   # BF16 operands: a tanh value rounded across a BF16 tie flips one 2^-8 step
GRAD_TOL = 3e-2  # relative to the largest gradient entry


def test_mlp_forward_and_gradient_vs_checker():
    net = L.MLP(seed=3)
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (29, 16)).astype(np.float32)
    T = rng.normal(0, 1, (29, 4)).astype(np.float32)
    P = net.params.astype(np.float64)
    print("[gpu] mlp forward/gradient n=29", flush=True)
    y = net.forward(X)
    y_ref, _ = mlp_ref.forward(P, X)
    assert relerr(y, y_ref) < FWD_TOL
    loss, g = net.gradient(X, T)
    loss_ref, g_ref = mlp_ref.loss_and_grad(P, X, T)
    assert abs(loss - loss_ref) < 1e-2 * loss_ref
    assert np.max(np.abs(g - g_ref)) < GRAD_TOL * np.max(np.abs(g_ref))
    assert np.corrcoef(g, g_ref)[0, 1] > 0.999
    # targets for the first two outputs only: the others get no gradient
    loss2, g2 = net.gradient(X, T[:, :2])
    loss2_ref, g2_ref = mlp_ref.loss_and_grad(P, X, T[:, :2], 2)
    assert abs(loss2 - loss2_ref) < 1e-2 * loss2_ref
    assert np.max(np.abs(g2 - g2_ref)) < GRAD_TOL * np.max(np.abs(g2_ref))
    p2 = mlp_ref.unpack(g2)
    assert not np.any(p2["W3"][2:]) and not np.any(p2["b3"][2:])
    # more samples than one batch: every batch of 32 runs the same network
    X2 = rng.normal(0, 1, (77, 16)).astype(np.float32)
    assert relerr(net.forward(X2), mlp_ref.forward(P, X2)[0]) < FWD_TOL


def test_mlp_training_on_gpu_fits_a_smooth_map():
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (96, 16)).astype(np.float32)
    W = rng.normal(0, 0.5, (16, 4))
    T = np.tanh(X @ W).astype(np.float32)
    net = L.MLP(seed=1)
    print("[gpu] mlp training 96 samples", flush=True)
    losses = net.fit(X, T, steps=1500, lr=3e-3)
    final = mlp_ref.loss_and_grad(net.params.astype(np.float64), X, T)[0]
    first = float(np.mean(losses[:3]))
    print(f"[gpu] mlp loss {first:.4f} -> {final:.5f}", flush=True)
    assert final < 0.05 * first


@pytest.mark.parametrize("deg,prec", [(4, 1), (2, 2)])
def test_learned_smoother_gain_on_unseen_systems(deg, prec):
    """Train on systems the plain solver labels itself, then compare the
    V-cycle contraction and the GMRES iterations of learned and default
    steps on unseen seeds; learned steps are kept only where they win."""
    print(f"[gpu] learned smoother degree {deg} precision {prec}: dataset", flush=True)
    X, T, info = L.make_dataset(range(1001, 1021), precision=prec, m=deg)
    print(f"[gpu] dataset {X.shape} best/default factors {[(round(b, 3), round(d, 3)) for _, b, d in info]}",
          flush=True)
    net = L.MLP(seed=0)
    losses = net.fit(X, T, steps=3000, lr=3e-3)
    print(f"[gpu] training loss {losses[0]:.4f} -> {losses[-1]:.5f}", flush=True)
    assert losses[-1] < 0.2 * losses[0]
    gains = []
    for seed, grid, layered in ((2001, (20, 20, 10), False), (2002, (24, 24, 12), False), (2003, (30, 70, 26), True)):
        Ad, bd, H = L.training_system(seed, *grid, layered=layered)
        p = jx.default_params(poly_degree=deg, precond_precision=prec, tol=1e-6)
        s_def = jx.Solver(Ad, H, p)
        _, st_def, _ = s_def.solve(bd)
        f_def = s_def.vcycle_factor()
        s_def.close()
        s = jx.Solver(Ad, H, p)
        used, f_l, f_d = s.apply_mlp(net.P.data_ptr())
        x, st, _ = s.solve(bd)
        true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
        print(f"[gpu] seed {seed}: learned used={used} factor {f_l:.4f} vs default {f_d:.4f}; iterations "
              f"{st.iterations} vs {st_def.iterations}", flush=True)
        assert abs(f_d - f_def) < 1e-6
        assert st.converged == 1 and true <= 1e-6
        assert abs(s.vcycle_factor() - f_l) < 1e-6  # the steps in place give the reported factor
        if used:
            assert f_l < f_d
        else:
            # fallback: the default steps are back in place
            assert abs(f_l - f_d) < 1e-6
            assert st.iterations == st_def.iterations
        gains.append(f_d - f_l)
        s.close()
    # the network must win somewhere (otherwise it adds nothing over the defaults)
    assert max(gains) > 0.0
