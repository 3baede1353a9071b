"""CPU checks of the learned-con2prim oracle (oracle/mlp_oracle.py): its
backward pass against central finite differences of its forward pass, its
BF16 rounding against known bit patterns, and its features against the
closed-form conformal root."""
import numpy as np

from oracle import mlp_oracle as mo


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.NPARAM)
    p[mo.OFF_W1:mo.OFF_B1] = rng.uniform(-1, 1, 320) * np.sqrt(6 / 5) * scale
    p[mo.OFF_B1:mo.OFF_W2] = rng.uniform(-0.1, 0.1, 64)
    p[mo.OFF_W2:mo.OFF_B2] = rng.uniform(-1, 1, 4096) * np.sqrt(6 / 64) * scale
    p[mo.OFF_B2:mo.OFF_W3] = rng.uniform(-0.1, 0.1, 64)
    p[mo.OFF_W3:mo.OFF_B3] = rng.uniform(-1, 1, 128) * np.sqrt(6 / 64)
    p[mo.OFF_B3:] = rng.uniform(-0.1, 0.1, 2)
    return p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e-3], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 1.5*2^-8 rounds up
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) / 3.0e-3 < 2 ** -8
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rand_params(5)
    X = rng.uniform(-1, 1, (40, 5))
    Y = rng.uniform(0, 1, (40, 2))
    g, L = mo.backward(p, X, Y, rounding=False)
    assert abs(L - mo.loss(p, X, Y, rounding=False)) < 1e-12 * max(1.0, L)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 10),
                          np.arange(mo.OFF_B1, mo.OFF_B1 + 5),
                          rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 20),
                          np.arange(mo.OFF_B2, mo.OFF_B2 + 5),
                          rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 10),
                          np.arange(mo.OFF_B3, mo.NPARAM)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss(pp, X, Y, rounding=False) - mo.loss(pm, X, Y, rounding=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_features_and_conformal_root():
    rng = np.random.default_rng(7)
    e = 10 ** rng.uniform(-4, 3, 500)
    v = rng.uniform(0, 0.99, 500)
    n = rng.normal(size=(500, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    g2 = 1 / (1 - v * v)
    p = e / 3
    Q = np.zeros((500, 7))
    Q[:, 0] = (e + p) * g2 - p
    Q[:, 1:4] = ((e + p) * g2 * v)[:, None] * n
    ee, vv = mo.c2p_conformal(Q)
    assert np.max(np.abs(ee - e) / e) < 1e-12 and np.max(np.abs(vv - v)) < 1e-12
    f = mo.features(Q)
    assert np.allclose(f[:, 1], np.linalg.norm(Q[:, 1:4], axis=1) / Q[:, 0], rtol=1e-6)
    assert np.allclose(f[:, 4], np.sqrt(1 - f[:, 1].astype(np.float64) ** 2), atol=1e-6)
