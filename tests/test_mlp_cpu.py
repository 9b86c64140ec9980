"""CPU pin of the prescreen network's oracle (tests/mlp_ref.py): its float64 gradients equal
central finite differences of its loss, its BF16 rounding is round-to-nearest-even, and the
Adam step matches the textbook update (Kingma & Ba 2015, Algorithm 1)."""
import numpy as np

import mlp_ref


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-39, 65504.0], dtype=np.float32)
    got = mlp_ref.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0); 1 + 3*2^-8 ties to 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5 and got[6] == 65536.0


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, size=mlp_ref.N_PARAMS)
    X = rng.normal(0, 1, size=(16, 64))
    t = (rng.random(16) < 0.5).astype(np.float64)
    g = mlp_ref.grads(p, X, t, emulate=False)
    idx = np.concatenate([rng.choice(64 * 256, 40, replace=False),                  # W1
                          64 * 256 + rng.choice(256, 10, replace=False),             # b1
                          64 * 256 + 256 + rng.choice(256 * 256, 40, replace=False),  # W2
                          mlp_ref.N_PARAMS - 1 - 256 - 256 + rng.choice(256, 10, replace=False),  # b2
                          mlp_ref.N_PARAMS - 1 - 256 + rng.choice(256, 10, replace=False),        # w3
                          [mlp_ref.N_PARAMS - 1]])                                   # b3
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_ref.loss(p + e, X, t, emulate=False) - mlp_ref.loss(p - e, X, t, emulate=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_adam_matches_textbook():
    p, g = np.array([1.0, -2.0]), np.array([0.5, -0.25])
    m = v = np.zeros(2)
    p1, m1, v1 = mlp_ref.adam(p, g, m, v, 1, 0.1)
    # first step: m_hat = g, v_hat = g^2 -> p - lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
