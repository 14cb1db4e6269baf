"""Pins the NumPy network reference (tests/cnn_ref.py) on the CPU: its
backward pass against central finite differences in float64, for both
padding modes and entries of every layer (weights and biases)."""
import numpy as np
import pytest

import cnn_ref as R


@pytest.mark.parametrize("px", [0, 1])
def test_backward_matches_finite_differences(px):
    rng = np.random.default_rng(3 + px)
    p = R.init_params(5 + px)
    x = rng.standard_normal((2, 16, 16, 3))
    tgt = rng.standard_normal((2, 16, 16))

    def loss(q):
        out, _ = R.forward(q, x, px)
        return R.loss_and_gout(out, tgt, 2)[0]

    out, acts = R.forward(p, x, px)
    _, gout = R.loss_and_gout(out, tgt, 2)
    g = R.backward(p, acts, gout, px)
    idx = [0, 500, 863, 870, 896 + 17, 896 + 9216 + 3, 896 + 9248 + 4000, 896 + 2 * 9248 + 9000,
           896 + 3 * 9248 + 123, 896 + 3 * 9248 + 9230, 37888, 38000, 38176]
    for k in idx:
        e = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        fd = (loss(pp) - loss(pm)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-10, (k, fd, g[k])


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0])
    r = R.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
