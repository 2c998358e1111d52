"""The FP64 network oracle (oracle/mlp_ref.py) against central finite
differences and closed forms (CPU only): pins the backpropagation formulas
the GPU kernels are checked against."""
import numpy as np

from oracle import mlp_ref


def _net(rng, shapes):
    return np.concatenate([np.concatenate([rng.standard_normal(a * b) / np.sqrt(a), rng.standard_normal(b) * 0.1])
                           for a, b in shapes])


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    shapes = [(5, 7), (7, 7), (7, 6), (6, 4)]
    p = _net(rng, shapes)
    x = rng.standard_normal((3, 5))
    t = rng.standard_normal((3, 4))
    loss, g = mlp_ref.gradients(p, shapes, x, t)
    h = 1e-6
    for i in rng.choice(p.size, size=60, replace=False):
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_ref.gradients(p + e, shapes, x, t)[0] - mlp_ref.gradients(p - e, shapes, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_linear_network_closed_form():
    # two layers with non-negative hidden pre-activations: the network is affine
    rng = np.random.default_rng(1)
    shapes = [(4, 3), (3, 2)]
    W1 = np.abs(rng.standard_normal((4, 3)))
    b1 = np.ones(3)
    W2 = rng.standard_normal((3, 2))
    b2 = rng.standard_normal(2)
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2])
    x = np.abs(rng.standard_normal((5, 4)))
    y, _ = mlp_ref.forward(p, shapes, x)
    assert np.allclose(y, (x @ W1 + b1) @ W2 + b2)
    # loss gradient of the output bias = 2 mean over rows of (y - t), per output / width
    t = rng.standard_normal((5, 2))
    _, g = mlp_ref.gradients(p, shapes, x, t)
    gb2 = g[-2:]
    assert np.allclose(gb2, 2.0 * (y - t).sum(axis=0) / y.size)


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 lies halfway between BF16 neighbours 1 and 1 + 2^-7: ties to even (1)
    assert mlp_ref.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mlp_ref.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert mlp_ref.bf16_round(np.float32(-2.5)) == -2.5


def test_adam_oracle_hand_worked_steps():
    """Adam pinned by hand: at step 1 the bias-corrected moments are g and
    g^2, so the move is lr g / (|g| + eps) (lr sign(g) up to eps); at step 2
    with gradients g1, g2: m = 0.09 g1 + 0.1 g2, v = 0.000999 g1^2 + 0.001 g2^2,
    m_hat = m / 0.19, v_hat = v / 0.001999."""
    from oracle import mlp_ref
    w0 = np.array([1.0, -2.0, 0.5, 0.0])
    g1 = np.array([0.3, -4.0, 1e-3, 0.0])
    g2 = np.array([-0.6, 2.0, 1e-3, 5.0])
    lr = 0.01
    s1, s2 = mlp_ref.adam_steps(w0, [g1, g2], lr)
    exp1 = w0 - lr * g1 / (np.abs(g1) + 1e-8)
    assert np.allclose(s1, exp1, rtol=0, atol=1e-15)
    m = 0.09 * g1 + 0.1 * g2
    v = 0.000999 * g1 ** 2 + 0.001 * g2 ** 2
    exp2 = exp1 - lr * (m / 0.19) / (np.sqrt(v / 0.001999) + 1e-8)
    assert np.allclose(s2, exp2, rtol=0, atol=1e-14)
    # a planted mistake (no bias correction) must differ
    wrong = w0 - lr * (0.1 * g1) / (np.sqrt(0.001 * g1 ** 2) + 1e-8)
    assert not np.allclose(s1, wrong, rtol=0, atol=1e-6)
