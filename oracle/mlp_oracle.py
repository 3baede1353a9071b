"""CPU oracle of the learned con2prim guess (test infrastructure only: the
tests import it as the checker, nothing in the product path does).

The MLP 5 -> 64 -> 64 -> 2 (ReLU) of the GPU path, restated in numpy:
  parameter vector  W1[64][5] | b1[64] | W2[64][64] | b2[64] | W3[2][64] | b3[2]
  features of q = Q/tau  ln(q_t)/10, |q|/q_t, q_eta/q_t, n_b/q_t, sqrt(1-(|q|/q_t)^2)
  targets            |v|, e/q_t  (from the reference con2prim, transformPV,
                     src/rmn.cpp:28-139)
  loss               (1/n) sum_i sum_o (out_o - y_o)^2
The GPU inference rounds the inputs, W1 | b1, the first hidden layer and W2 to
BF16 (round to nearest even) and accumulates in FP32; `rounding=True`
reproduces those rounding points (accumulating in FP64).  The backward pass
treats the BF16 roundings as the identity (straight-through), as the GPU
training kernels do.

Pinning: `backward` with rounding=False is checked against central finite
differences of `forward` in FP64 (tests/test_mlp_oracle.py); the model has no
counterpart in the reference, so nothing else can pin it.
"""
import numpy as np

NIN, NH, NOUT = 5, 64, 2
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, NPARAM = 0, 320, 384, 4480, 4544, 4672, 4674


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NH, NIN), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(NH, NH), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(NOUT, NH), b3=p[OFF_B3:NPARAM])


def features(Q):
    """Features (float32, n x 5) of states Q (n x 7, already divided by tau)."""
    Q = np.asarray(Q, dtype=np.float64)
    Qt = Q[:, 0]
    ok = Qt > 0
    M = np.sqrt(Q[:, 1] ** 2 + Q[:, 2] ** 2 + Q[:, 3] ** 2)
    M = np.where(M > Qt, 0.9999 * Qt, M)
    safe = np.where(ok, Qt, 1.0)
    x = (M / safe).astype(np.float32)
    f = np.zeros((len(Q), 5), dtype=np.float32)
    f[:, 0] = np.float32(0.1) * np.log(safe.astype(np.float32))
    f[:, 1] = x
    f[:, 2] = (Q[:, 3] / safe).astype(np.float32)
    f[:, 3] = (Q[:, 4] / safe).astype(np.float32)
    g = np.float32(1) - x * x
    f[:, 4] = np.sqrt(np.maximum(g, np.float32(0)))
    f[~ok] = 0
    return f


def targets(Q, prim):
    Q = np.asarray(Q, dtype=np.float64)
    v = np.sqrt(prim[:, 5] ** 2 + prim[:, 6] ** 2 + prim[:, 7] ** 2)
    eq = np.where(Q[:, 0] > 0, prim[:, 0] / np.where(Q[:, 0] > 0, Q[:, 0], 1.0), 0.0)
    return np.stack([v, eq], axis=1).astype(np.float32)


def forward(params, X, rounding=True, keep=False):
    """Outputs (n x 2) of the MLP for features X (n x 5)."""
    w = unpack(params)
    r = (lambda a: bf16(a).astype(np.float64)) if rounding else (lambda a: np.asarray(a, np.float64))
    x = r(X)
    W1, b1, W2 = r(w["W1"]), r(w["b1"]), r(w["W2"])
    z1 = x @ W1.T + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ W2.T + w["b2"]
    h2 = np.maximum(z2, 0.0)
    out = h2 @ w["W3"].T + w["b3"]
    if keep:
        return out, dict(x=x, W1=W1, W2=W2, z1=z1, h1=h1, z2=z2, h2=h2)
    return out


def loss(params, X, Y, rounding=True):
    out = forward(params, X, rounding)
    return float(np.mean(np.sum((out - Y) ** 2, axis=1)))


def backward(params, X, Y, rounding=True):
    """(gradient vector of the loss, loss)."""
    w = unpack(params)
    out, a = forward(params, X, rounding, keep=True)
    n = len(X)
    d = out - np.asarray(Y, np.float64)
    L = float(np.mean(np.sum(d * d, axis=1)))
    dout = 2.0 * d / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (dout.T @ a["h2"]).ravel()
    g[OFF_B3:NPARAM] = dout.sum(0)
    dz2 = (dout @ w["W3"]) * (a["z2"] > 0)
    g[OFF_W2:OFF_B2] = (dz2.T @ a["h1"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ a["W2"]) * (a["z1"] > 0)
    g[OFF_W1:OFF_B1] = (dz1.T @ a["x"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return g, L


def c2p_conformal(Q):
    """Exact con2prim root of p = e/3 for states Q (n x 7): (e, |v|), the
    root v = 3M / (2 Q_t + sqrt(4 Q_t^2 - 3 M^2)) of M v^2 - 4 Q_t v + 3 M = 0."""
    Q = np.asarray(Q, dtype=np.float64)
    Qt = Q[:, 0]
    M = np.sqrt(Q[:, 1] ** 2 + Q[:, 2] ** 2 + Q[:, 3] ** 2)
    M = np.where(M > Qt, 0.9999 * Qt, M)
    v = 3 * M / (2 * Qt + np.sqrt(4 * Qt * Qt - 3 * M * M))
    return Qt - M * v, v
