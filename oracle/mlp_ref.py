"""TEST INFRASTRUCTURE: NumPy restatement of the learned-smoother MLP
(jxpamg_all_amd/csrc/hip/jxamd_mlp.hip) for checking the GPU kernels.

Same network and the same BF16 rounding (round to nearest even) of every
matrix-product operand, in float64 arithmetic.  Only tests/ import it."""
import numpy as np

SIZES = [("W1", (64, 16)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (4, 64)), ("b3", (4,))]
NPARAM = 5508


def bf16(x):
    """float32 -> nearest-even bfloat16, returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) >> 16
    return (a.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)


def unpack(P):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(P[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def forward(P, X):
    p = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    H1 = np.tanh(bf16(X) @ bf16(p["W1"]).T + p["b1"])
    H2 = np.tanh(bf16(H1) @ bf16(p["W2"]).T + p["b2"])
    Y = bf16(H2) @ bf16(p["W3"]).T + p["b3"]
    return Y, (X, H1, H2)


def loss_and_grad(P, X, T, nout=4):
    """Mean over samples of the summed squared error of the first nout
    outputs (T: [n][nout]), and its gradient."""
    p = unpack(P)
    Y, (X, H1, H2) = forward(P, X)
    nb = len(X)
    R = np.zeros_like(Y)
    R[:, :nout] = Y[:, :nout] - np.asarray(T, dtype=np.float64).reshape(nb, nout)
    loss = float((R * R).sum() / nb)
    dY = 2.0 * R / nb
    g = {"W3": bf16(dY).T @ bf16(H2), "b3": dY.sum(0)}
    dZ2 = (bf16(dY) @ bf16(p["W3"])) * (1.0 - H2 * H2)
    g["W2"] = bf16(dZ2).T @ bf16(H1)
    g["b2"] = dZ2.sum(0)
    dZ1 = (bf16(dZ2) @ bf16(p["W2"])) * (1.0 - H1 * H1)
    g["W1"] = bf16(dZ1).T @ bf16(X)
    g["b1"] = dZ1.sum(0)
    return loss, pack(g)


def loss_exact(P, X, T, nout=4):
    """The same loss without BF16 rounding (float64 throughout)."""
    p = unpack(P)
    H1 = np.tanh(X @ p["W1"].T + p["b1"])
    H2 = np.tanh(H1 @ p["W2"].T + p["b2"])
    Y = H2 @ p["W3"].T + p["b3"]
    R = Y[:, :nout] - np.asarray(T).reshape(len(X), nout)
    return float((R * R).sum() / len(X))
