"""GPU network kernels (csrc/pv_cnn.hip) against the NumPy reference.

The reference rounds to BF16 where the GPU stores BF16 (activations, the
32 -> 32 weights, the pre-activation gradients), so the remaining difference
is FP32 accumulation order and the odd BF16 rounding flip it causes:
bounds below are relative L2 norms.  These tests also pin the operand
layout of the MFMA builtin the 32 -> 32 layers use.
"""
import numpy as np
import pytest

import cnn_ref as R
import pvof

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(gpu):
    p = pvof.config1_params(NXM=32, NYM=64)
    s = pvof.Solver(p)
    return pvof.Network(s, seed=7)


def l2rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("px", [0, 1])
def test_forward_matches_reference(net, px):
    rng = np.random.default_rng(11 + px)
    prm = R.init_params(21).astype(np.float32)
    net.set_params(prm)
    x = rng.standard_normal((2, 32, 48, 3)).astype(np.float32)
    out = net.eval(x, px=px)
    ref, _ = R.forward(prm.astype(np.float64), x.astype(np.float64), px, use_bf16=True)
    assert l2rel(out, ref) < 5e-3, l2rel(out, ref)


@pytest.mark.parametrize("px", [0, 1])
def test_gradient_matches_reference(net, px):
    rng = np.random.default_rng(31 + px)
    prm = R.init_params(22).astype(np.float32)
    net.set_params(prm)
    x = rng.standard_normal((2, 32, 32, 3)).astype(np.float32)
    tgt = rng.standard_normal((2, 32, 32)).astype(np.float32)
    out, loss, grad = net.eval(x, tgt, px=px, border=3, want_grad=True)
    p64 = prm.astype(np.float64)
    ref, acts = R.forward(p64, x.astype(np.float64), px, use_bf16=True)
    rloss, gout = R.loss_and_gout(ref, tgt.astype(np.float64), 3)
    rgrad = R.backward(p64, acts, gout, px, use_bf16=True)
    assert abs(loss - rloss) <= 2e-3 * rloss
    spans = [(0, 896)] + [(896 + k * 9248, 896 + (k + 1) * 9248) for k in range(4)] + [(37888, 38177)]
    for a, b in spans:
        e = l2rel(grad[a:b], rgrad[a:b])
        assert e < 3e-2, (a, e)


def test_training_reduces_loss(net):
    """Regress a fixed 3x3 smoothing of input channel 0 plus channel 1."""
    rng = np.random.default_rng(5)
    net.set_params(R.init_params(23).astype(np.float32))
    x = rng.standard_normal((4, 32, 32, 3)).astype(np.float32)
    xp = np.pad(x[..., 0], ((0, 0), (1, 1), (1, 1)))
    tgt = sum(xp[:, ky:ky + 32, kx:kx + 32] for ky in range(3) for kx in range(3)) / 9.0 + 0.5 * x[..., 1]
    losses = []
    for t in range(1, 301):
        _, loss, _ = net.eval(x, tgt, border=2)
        losses.append(loss)
        net.adam_step(2e-3, t)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
