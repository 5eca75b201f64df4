"""Neural radiance cache kernels against the CPU oracle (oracle/nrc_oracle.py).

  * forward: the oracle mirrors the device arithmetic (bf16 weights, features and
    activations, FP64 accumulation in place of FP32): outputs agree within 1e-2 relative
    (+1e-3 absolute) -- bf16 rounding of an activation may differ where FP32 and FP64
    accumulation straddle a rounding boundary;
  * gradients: per-layer relative L2 error of the device gradient (bf16 backprop, FP32
    accumulation) against the oracle's analytic gradient below 3e-2;
  * training: Adam on a smooth target reduces the relative loss at least tenfold.
"""
import numpy as np
import pytest

import nrc_oracle as N
from crender_amd import Context, Nrc

pytestmark = pytest.mark.gpu


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 13), np.float32)
    q[:, :3] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    q[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    q[:, 6:9] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    q[:, 9:12] = rng.random((n, 3))
    q[:, 12] = rng.random(n)
    return q


@pytest.fixture(scope="module")
def nrc():
    ctx = Context(0)
    m = Nrc(ctx, max_batch=8192, seed=5)
    yield m
    m.close()
    ctx.close()


def test_param_layout(nrc):
    p = nrc.params()
    assert p.size == N.N_PARAMS == 24576
    W = N.split(p.astype(np.float64))
    assert np.all(W[-1][:, 3:] == 0.0) and np.all(W[0][83:, :] == 0.0)


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4096])
def test_forward_matches_oracle(nrc, n):
    q = _queries(n, n)
    got = nrc.infer(q)
    ref, _ = N.forward(nrc.params(), N.encode(q, True), True)
    ref = ref[:, :3]
    close = np.isclose(got, ref, rtol=1e-2, atol=1e-3)
    assert close.mean() >= 0.995, (close.mean(), np.abs(got - ref).max())
    assert np.abs(got - ref).max() < 5e-2 * max(1.0, np.abs(ref).max())


def test_gradient_matches_oracle(nrc):
    q = _queries(512, 7)
    t = (np.random.default_rng(8).random((512, 3)) * 2).astype(np.float32)
    loss, g = nrc.gradient(q, t)
    lref, gref = N.loss_and_grad(nrc.params(), q, t, round_bf16=True)
    assert abs(loss - lref) <= 2e-2 * abs(lref)
    for (o, (k, n)) in zip(N.offsets(), N.LAYERS):
        a, b = g[o:o + k * n].astype(np.float64), gref[o:o + k * n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err < 3e-2, (o, err)


def test_training_reduces_loss():
    ctx = Context(0)
    m = Nrc(ctx, max_batch=4096, seed=2, lr=3e-3)
    rng = np.random.default_rng(0)
    losses = []
    for step in range(300):
        q = _queries(4096, 1000 + step)
        t = np.stack([0.5 + 0.4 * np.sin(3 * q[:, 0]) * np.cos(2 * q[:, 1]),
                      0.3 + 0.2 * q[:, 2] + 0.2 * q[:, 9],
                      0.6 + 0.3 * np.cos(4 * q[:, 0] + q[:, 1])], 1).astype(np.float32)
        losses.append(m.train_step(q, t))
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-5:])
    m.close()
    ctx.close()
