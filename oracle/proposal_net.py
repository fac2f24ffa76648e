"""CPU fp64 restatement of the learned-seeding proposal network.

TEST INFRASTRUCTURE ONLY: the checker of the network's HIP forward and
backward kernels (libvc1_amd/csrc/vc1me_nn.hip).  The product never imports
it.  The network has no counterpart in the reference.

Parameter layout (as vc1me_nn_set_params takes it): for each of the four
3x3 convolutions, weights [32][9][cin] (tap t = 3 * (dy + 1) + (dx + 1) at
offset (dy, dx) * dilation) then 32 biases; then the head's weights [5][32]
and its 5 biases.  Activations are NHWC.
"""
from __future__ import annotations

import numpy as np

CIN = (2, 32, 32, 32)
DIL = (1, 2, 4, 8)


def layout():
    """(name, offset, shape) of every parameter tensor."""
    out, off = [], 0
    for l, cin in enumerate(CIN):
        out.append((f"w{l}", off, (32, 9, cin)))
        off += 32 * 9 * cin
        out.append((f"b{l}", off, (32,)))
        off += 32
    out.append(("wh", off, (5, 32)))
    off += 160
    out.append(("bh", off, (5,)))
    off += 5
    return out, off


def unpack(params: np.ndarray) -> dict:
    lay, n = layout()
    assert params.size == n, (params.size, n)
    return {name: params[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for name, o, s in lay}


def pack(d: dict) -> np.ndarray:
    lay, n = layout()
    out = np.zeros(n, np.float64)
    for name, o, s in lay:
        out[o:o + int(np.prod(s))] = d[name].reshape(-1)
    return out


def init_params(seed: int) -> np.ndarray:
    """He-normal convolution weights, zero biases, small head."""
    rng = np.random.default_rng(seed)
    d = {}
    for l, cin in enumerate(CIN):
        d[f"w{l}"] = rng.normal(0.0, np.sqrt(2.0 / (9 * cin)), size=(32, 9, cin))
        d[f"b{l}"] = np.zeros(32)
    d["wh"] = rng.normal(0.0, 0.05, size=(5, 32))
    d["bh"] = np.zeros(5)
    return pack(d).astype(np.float32)


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """y[n, i, j] = x[n, i + dy, j + dx], zero outside."""
    n, h, w, c = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(h, h - dy)
    xs, xe = max(0, -dx), min(w, w - dx)
    if ys < ye and xs < xe:
        y[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def _taps(dil: int):
    return [((t // 3 - 1) * dil, (t % 3 - 1) * dil) for t in range(9)]


def conv(x, w, b, dil):
    out = np.zeros(x.shape[:3] + (32,))
    for t, (dy, dx) in enumerate(_taps(dil)):
        out += _shift(x, dy, dx) @ w[:, t, :].T
    return out + b


def cells(h4: int, w4: int):
    mbh, mbw = (h4 + 3) // 4, (w4 + 3) // 4
    return mbh, mbw


def pool(a: np.ndarray) -> np.ndarray:
    n, h, w, c = a.shape
    mbh, mbw = cells(h, w)
    p = np.zeros((n, mbh, mbw, c))
    for my in range(mbh):
        for mx in range(mbw):
            blk = a[:, my * 4:min(h, my * 4 + 4), mx * 4:min(w, mx * 4 + 4)]
            p[:, my, mx] = blk.mean(axis=(1, 2))
    return p


def forward(params: np.ndarray, x: np.ndarray, quant: bool = False):
    """x: (n, h4, w4, 2).  Returns (out (n, mbh, mbw, 5), cache).

    quant=True rounds where the HIP kernels round: the input, the convolution
    weights and every activation to bf16 (sums stay wide)."""
    d = unpack(params)
    q = bf16 if quant else (lambda a: a)
    acts, pre = [q(np.asarray(x, np.float64)).astype(np.float64)], []
    for l in range(4):
        w = q(d[f"w{l}"]).astype(np.float64)
        z = conv(acts[-1], w, d[f"b{l}"], DIL[l])
        pre.append(z)
        acts.append(q(np.maximum(z, 0.0)).astype(np.float64))
    p = pool(acts[-1])
    out = p @ d["wh"].T + d["bh"]
    return out, (d, acts, pre, p, quant)


def backward(cache, dout: np.ndarray) -> np.ndarray:
    """Gradient of sum(out * dout) with respect to every parameter (packed)."""
    d, acts, pre, p, quant = cache
    q = bf16 if quant else (lambda a: a)
    g = {}
    g["wh"] = np.einsum("nyxo,nyxc->oc", dout, p)
    g["bh"] = dout.sum(axis=(0, 1, 2))
    dp = dout @ d["wh"]
    n, h, w, c = acts[-1].shape
    mbh, mbw = cells(h, w)
    da = np.zeros_like(acts[-1])
    for my in range(mbh):
        for mx in range(mbw):
            ys, ye, xs, xe = my * 4, min(h, my * 4 + 4), mx * 4, min(w, mx * 4 + 4)
            cnt = (ye - ys) * (xe - xs)
            da[:, ys:ye, xs:xe] += dp[:, my, mx][:, None, None, :] / cnt
    # ReLU masks: the kernels test the stored (rounded) activation, the exact
    # form the pre-activation; they agree but for underflow
    mask = [(acts[l + 1] > 0) if quant else (pre[l] > 0) for l in range(4)]
    dz = q(da).astype(np.float64) * mask[3]
    for l in range(3, -1, -1):
        x = acts[l]
        taps = _taps(DIL[l])
        g[f"w{l}"] = np.stack([np.einsum("nyxo,nyxc->oc", dz, _shift(x, dy, dx)) for dy, dx in taps], axis=1)
        g[f"b{l}"] = dz.sum(axis=(0, 1, 2))
        if l:
            wl = q(d[f"w{l}"]).astype(np.float64)
            dx_ = np.zeros_like(x)
            for t, (dy, dx) in enumerate(taps):
                dx_ += _shift(dz, -dy, -dx) @ wl[:, t, :]
            dz = q(dx_).astype(np.float64) * mask[l - 1]
    return pack(g)


def downsample(cur: np.ndarray, ref: np.ndarray, gx: int = 0, gy: int = 0) -> np.ndarray:
    """Network input of one frame pair: 4x4 box means of the two luma planes
    (edge samples repeated past the frame), the ref block displaced by the
    global vector (gx, gy) px, scaled to [-0.5, 0.5] in float32 and rounded
    to bf16.  (h4, w4, 2), h4 = ceil(H / 4), w4 = ceil(W / 4)."""
    H, W = cur.shape
    h4, w4 = (H + 3) // 4, (W + 3) // 4
    out = np.zeros((h4, w4, 2), np.float32)
    for k, (plane, ox, oy) in enumerate(((cur, 0, 0), (ref, gx, gy))):
        ys = np.clip(np.arange(h4 * 4) + oy, 0, H - 1)
        xs = np.clip(np.arange(w4 * 4) + ox, 0, W - 1)
        e = plane.astype(np.int64)[ys][:, xs]
        s = e.reshape(h4, 4, w4, 4).sum(axis=(1, 3))
        out[..., k] = s.astype(np.float32) / np.float32(4080.0) - np.float32(0.5)
    return bf16(out)


def _sl1(e):
    a = np.abs(e)
    return np.where(a < 1, 0.5 * e * e, a - 0.5)


def _dsl1(e):
    return np.where(np.abs(e) < 1, e, np.sign(e))


def loss(out: np.ndarray, tgt: np.ndarray, valid: np.ndarray):
    """Training loss and its gradient with respect to out.

    out (m, 5): two candidate vectors (64-px units) and a confidence logit;
    tgt (m, 2) in 64-px units; valid (m,) bool.  Per valid macroblock: the
    nearer candidate (smooth-L1 on errors scaled to 8-px units) is regressed,
    and the logit learns whether that candidate lies within 8 px of the
    target in both components (weight 0.5, binary cross-entropy).  The loss is
    the mean over valid macroblocks."""
    out = np.asarray(out, np.float64).reshape(-1, 5)
    tgt = np.asarray(tgt, np.float64).reshape(-1, 2)
    valid = np.asarray(valid, bool).reshape(-1)
    cnt = int(valid.sum())
    inv = 1.0 / cnt if cnt else 0.0
    s = 8.0
    e = np.stack([_sl1((out[:, 2 * k] - tgt[:, 0]) * s) + _sl1((out[:, 2 * k + 1] - tgt[:, 1]) * s) for k in range(2)], 1)
    kb = (e[:, 1] < e[:, 0]).astype(int)
    rows = np.arange(out.shape[0])
    cx, cy = out[rows, 2 * kb], out[rows, 2 * kb + 1]
    near = (np.maximum(np.abs(cx - tgt[:, 0]), np.abs(cy - tgt[:, 1])) * 64.0 <= 8.0).astype(np.float64)
    z = out[:, 4]
    bce = np.maximum(z, 0) - z * near + np.log1p(np.exp(-np.abs(z)))
    per = e[rows, kb] + 0.5 * bce
    total = float((per * valid).sum() * inv)
    dout = np.zeros_like(out)
    dout[rows, 2 * kb] = _dsl1((cx - tgt[:, 0]) * s) * s * inv
    dout[rows, 2 * kb + 1] = _dsl1((cy - tgt[:, 1]) * s) * s * inv
    dout[:, 4] = 0.5 * (1.0 / (1.0 + np.exp(-z)) - near) * inv
    dout[~valid] = 0.0
    return total, dout


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps), m, v


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bf16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)
