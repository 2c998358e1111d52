"""GPU checks of the warm-start network's kernels (bo_mlp_*, bo_gemm_bf16).

* The MFMA product kernel against the exact product of the same BF16
  operands (FP64 on the CPU), for every transpose combination and ragged
  sizes (tile edges in M, N and K).
* Gradients of the whole network (forward, backward: weight, bias and input
  gradients through ReLU) against the FP64 oracle oracle/mlp_ref.py, which
  quantises to BF16 where the kernels do; the oracle itself is pinned by
  finite differences in tests/test_mlp_oracle.py.
* Training with Adam reduces the loss of a small regression.
"""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _bf(pkg, a):
    return pkg.f32_to_bf16_bits(a)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("M,N,K", [(70, 33, 45), (64, 64, 32), (5, 130, 300), (129, 7, 1),
                                   (64, 100, 5000), (33, 70, 4001)])  # the last two split K
def test_gemm_matches_exact_product(pkg, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    b = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    ab, bb = _bf(pkg, a), _bf(pkg, b)
    c = pkg.gemm_bf16(ab, bb, ta, tb)
    A = pkg.bf16_bits_to_f32(ab).astype(np.float64)
    B = pkg.bf16_bits_to_f32(bb).astype(np.float64)
    A = A.T if ta else A
    B = B.T if tb else B
    ref = A @ B
    scale = np.abs(A) @ np.abs(B)
    print("case: gemm", M, N, K, ta, tb, "max err", np.max(np.abs(c - ref) / (scale + 1e-30)), flush=True)
    assert c.shape == (M, N)
    assert np.all(np.abs(c - ref) <= 2e-6 * scale + 1e-30)


@pytest.mark.parametrize("n_in,width,n_out", [(37, 64, 29), (700, 96, 600)])  # the second splits K
def test_network_gradients_match_oracle(pkg, n_in, width, n_out):
    rng = np.random.default_rng(5)
    B = 8
    net = pkg.Mlp(n_in, width, n_out, layers=4, max_batch=16, seed=3)
    shapes = net.shapes()
    p = rng.standard_normal(net.nparam).astype(np.float32) * 0.3
    net.set_params(p)
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    t = rng.standard_normal((B, n_out)).astype(np.float32)
    loss, g = net.gradients(x, t)
    y = net.forward(x)
    yq, _ = mlp_ref.forward(p, shapes, x, quant=True)
    assert np.max(np.abs(y - yq)) <= 1e-3 * (1 + np.max(np.abs(yq)))
    loss_ref, g_ref = mlp_ref.gradients(p, shapes, x, t, quant=True)
    print("loss gpu %.8f oracle %.8f" % (loss, loss_ref), flush=True)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    o = 0
    for din, dout in shapes:
        for size, what in ((din * dout, "W"), (dout, "b")):
            gg, gr = g[o:o + size], g_ref[o:o + size]
            err = np.linalg.norm(gg - gr) / (np.linalg.norm(gr) + 1e-30)
            print("layer", din, dout, what, "rel err %.2e" % err, flush=True)
            assert err <= 1e-3, (din, dout, what, err)
            o += size


def test_training_reduces_loss(pkg):
    rng = np.random.default_rng(9)
    n_in, n_out = 48, 40
    M = rng.standard_normal((n_in, n_out)) / np.sqrt(n_in)
    x = rng.standard_normal((256, n_in)).astype(np.float32)
    t = np.tanh(x @ M).astype(np.float32)
    net = pkg.Mlp(n_in, 128, n_out, layers=4, max_batch=64, seed=2)
    first = None
    for step in range(300):
        idx = rng.integers(0, 256, size=64)
        loss = net.train_step(x[idx], t[idx], 2e-3)
        first = loss if first is None else first
    final, _ = net.gradients(x[:64], t[:64])
    print("loss %.5f -> %.5f" % (first, final), flush=True)
    assert final < 0.1 * first


def test_adam_steps_match_oracle(pkg):
    """k_adam (FP32 master weights) against the FP64 Adam of oracle/mlp_ref.py
    (pinned by hand-worked steps in tests/test_mlp_oracle.py): three
    training steps from set parameters, each step's gradient taken with
    bo_mlp_gradients at the parameters the previous step left; the
    parameters after each step must match the oracle's to FP32 rounding of
    the update."""
    rng = np.random.default_rng(21)
    n_in, width, n_out, B = 30, 64, 20, 8
    net = pkg.Mlp(n_in, width, n_out, layers=3, max_batch=16, seed=4)
    p0 = (rng.standard_normal(net.nparam) * 0.2).astype(np.float32)
    net.set_params(p0)  # also resets Adam's moments and step count
    x = rng.standard_normal((B, n_in)).astype(np.float32)
    t = rng.standard_normal((B, n_out)).astype(np.float32)
    lr = 1e-3
    grads, got = [], []
    for _ in range(3):
        _, g = net.gradients(x, t)  # gradient at the current parameters (the one train_step uses)
        grads.append(g.astype(np.float64))
        net.train_step(x, t, lr)
        got.append(net.params().astype(np.float64))
    want = mlp_ref.adam_steps(p0.astype(np.float64), grads, lr)
    for step, (a, b) in enumerate(zip(got, want), start=1):
        move = np.abs(b - p0)
        err = np.max(np.abs(a - b) / (move + 1e-7 * np.abs(b) + 1e-12))
        print("adam step", step, "max relative error of the move %.2e" % err, flush=True)
        assert np.all(np.abs(a - b) <= 1e-3 * move + 2e-7 * np.abs(b) + 1e-9), step
