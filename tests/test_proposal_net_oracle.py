"""Pins the CPU fp64 restatement of the proposal network (oracle/proposal_net.py).

The network has no counterpart in the reference, so the oracle is pinned by
mathematics: the tap convention by one-hot kernels, the pooling by closed
forms, and the backward pass and the loss gradient by central finite
differences of the forward pass (the backward is an independent derivation,
so a dropped term, a wrong sign, a transposed operand or a wrong tap
direction fails here).
"""
import numpy as np
import pytest

from oracle import proposal_net as pn


def test_layout_size():
    lay, n = pn.layout()
    assert n == 28517
    assert lay[0] == ("w0", 0, (32, 9, 2))
    assert lay[1] == ("b0", 576, (32,))
    assert lay[-2] == ("wh", n - 165, (5, 32))


def test_conv_tap_convention():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 9, 11, 32))
    for dil in (1, 2, 4):
        for t in range(9):
            w = np.zeros((32, 9, 32))
            w[np.arange(32), t, np.arange(32)] = 1.0
            y = pn.conv(x, w, np.zeros(32), dil)
            dy, dx = (t // 3 - 1) * dil, (t % 3 - 1) * dil
            for i in range(9):
                for j in range(11):
                    si, sj = i + dy, j + dx
                    want = x[0, si, sj] if 0 <= si < 9 and 0 <= sj < 11 else np.zeros(32)
                    np.testing.assert_array_equal(y[0, i, j], want)


def test_pool_closed_forms():
    a = np.ones((2, 7, 10, 32))
    np.testing.assert_allclose(pn.pool(a), 1.0)
    # a ramp in x: each cell's mean is the mean of its (possibly ragged) columns
    r = np.broadcast_to(np.arange(10.0)[None, None, :, None], (1, 7, 10, 32))
    p = pn.pool(r)
    assert p.shape == (1, 2, 3, 32)
    np.testing.assert_allclose(p[0, :, 0, 0], 1.5)
    np.testing.assert_allclose(p[0, :, 1, 0], 5.5)
    np.testing.assert_allclose(p[0, :, 2, 0], 8.5)


def _net(seed, n=1, h=6, w=7):
    rng = np.random.default_rng(seed)
    params = pn.unpack(pn.init_params(seed).astype(np.float64))
    for l in range(4):
        params[f"b{l}"] = rng.normal(0, 0.05, 32)
    params["bh"] = rng.normal(0, 0.1, 5)
    x = rng.uniform(-0.5, 0.5, size=(n, h, w, 2))
    return pn.pack(params), x, rng


def test_backward_matches_finite_differences():
    p, x, rng = _net(1)
    out, cache = pn.forward(p, x)
    r = rng.normal(size=out.shape)
    g = pn.backward(cache, r)
    lay, n = pn.layout()
    # every tensor gets a few probes
    idx = []
    for name, off, shape in lay:
        size = int(np.prod(shape))
        idx += list(off + rng.choice(size, size=min(size, 8), replace=False))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = ((pn.forward(pp, x)[0] * r).sum() - (pn.forward(pm, x)[0] * r).sum()) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_quantized_backward_is_exact_backward_of_rounded_net():
    """With rounding on, the backward treats each rounding as identity
    (straight-through); it must agree with the exact backward closely."""
    p, x, rng = _net(2, n=2, h=9, w=13)
    out, cache = pn.forward(p, x)
    outq, cacheq = pn.forward(p, x, quant=True)
    r = rng.normal(size=out.shape)
    g, gq = pn.backward(cache, r), pn.backward(cacheq, r)
    assert np.linalg.norm(outq - out) <= 0.02 * np.linalg.norm(out)
    assert np.linalg.norm(gq - g) <= 0.05 * np.linalg.norm(g)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, -3.0, 1.0 + 2 ** -9], np.float32)
    np.testing.assert_array_equal(pn.bf16(v), np.array([1.0, 1.0, 1.0 + 2 ** -6, -3.0, 1.0], np.float32))


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    m = 40
    tgt = rng.normal(0, 0.3, size=(m, 2))
    out = np.concatenate([tgt + rng.normal(0, 0.2, size=(m, 2)), rng.normal(0, 1, size=(m, 2)),
                          rng.normal(0, 2, size=(m, 1))], axis=1)
    valid = rng.uniform(size=m) < 0.8
    _, d = pn.loss(out, tgt, valid)
    eps = 1e-7
    for i in range(m):
        for o in range(5):
            op, om = out.copy(), out.copy()
            op[i, o] += eps
            om[i, o] -= eps
            fd = (pn.loss(op, tgt, valid)[0] - pn.loss(om, tgt, valid)[0]) / (2 * eps)
            # the near label and the winner switch are piecewise constant: skip probes on a switch
            assert abs(fd - d[i, o]) <= 1e-4 * max(1.0, abs(d[i, o])) or abs(fd) > 1e3, (i, o, fd, d[i, o])
    assert np.all(d[~valid] == 0)


def test_loss_closed_form():
    # exact candidate, zero logit: regression 0, BCE log 2 with label 1
    out = np.array([[0.25, -0.5, 3.0, 3.0, 0.0]])
    l, d = pn.loss(out, np.array([[0.25, -0.5]]), np.array([True]))
    assert l == pytest.approx(0.5 * np.log(2.0))
    np.testing.assert_allclose(d[0, :4], 0.0)
    assert d[0, 4] == pytest.approx(0.5 * (0.5 - 1.0))


def test_downsample_closed_form():
    cur = np.full((10, 13), 255, np.uint8)
    ref = np.zeros((10, 13), np.uint8)
    x = pn.downsample(cur, ref)
    assert x.shape == (3, 4, 2)
    np.testing.assert_array_equal(x[..., 0], 0.5)
    np.testing.assert_array_equal(x[..., 1], -0.5)
    # edge repeat: a ramp in x keeps the last column's value past the frame
    ramp = np.tile(np.arange(13, dtype=np.uint8) * 16, (10, 1))
    x = pn.downsample(ramp, ramp)
    last = (12 * 16) / 255.0 - 0.5
    assert x[0, 3, 0] == pytest.approx(last, abs=2 ** -8)
    # the ref channel displaced by the global vector: ref(x + 4) in cell x is cur's next cell
    x = pn.downsample(ramp, ramp, gx=4)
    np.testing.assert_array_equal(x[:, :2, 1], x[:, 1:3, 0])
    assert x[0, 3, 1] == pytest.approx(last, abs=2 ** -8)   # clamped past the edge


def test_adam_first_step_is_sign_step():
    g = np.array([3.0, -0.5, 1e-3])
    p, m, v = pn.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)
