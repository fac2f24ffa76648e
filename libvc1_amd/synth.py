"""Deterministic synthetic YUV 4:2:0 sequences for motion-estimation runs.

The sequences follow the workload definitions the backend is measured on:

* ``cif_translation``: 352x288, every frame a global quarter-pel translation
  of the previous one (each component uniform in [-8, 8] px), resampled with
  the VC-1 luma law (libvc1 ``luma_mc_sample``,
  src/encoder_motion_compensation.cpp:367-396, rnd=0, clamped fetch) and the
  VC-1 chroma law (``chroma_mc_sample``, :417-447), plus Gaussian noise
  sigma 2, rounded and clipped.  Frame 0 is uniform noise blurred with a
  Gaussian of sigma 3 px.
* ``pan_rectangles``: a camera pan of up to +/-16 px/frame over a
  blurred-noise background (blur sigma 2-6 px) with eight textured
  rectangles (64-512 px per side) moving at independent quarter-pel
  velocities of up to +/-``max_obj_px`` px/frame, painted in fixed z-order,
  plus Gaussian noise sigma 2.  Optional hard cuts to fresh textures and
  exact duplicate frames model the hard-motion 4K workload.

Textures are periodic (blurred through the FFT), so a pan never runs off the
texture and no edge smear accumulates.  Every draw comes from
``numpy.random.Generator(PCG64(seed))``; a sequence is a pure function of its
arguments.
"""
from __future__ import annotations

import dataclasses

import numpy as np

_TAPS = {1: (-4, 53, 18, -3), 2: (-1, 9, 9, -1), 3: (-3, 18, 53, -4)}
_SHIFT_VALUE = (0, 5, 1, 5)


def _floor_div4(v: int) -> int:
    return v // 4  # python floor division matches floor_div4 for all ints


def _arshift(v: np.ndarray, n: int) -> np.ndarray:
    return np.right_shift(v, n)  # numpy >> on signed ints is an arithmetic (floor) shift


def _filter4(a, b, c, d, mode: int):
    t = _TAPS[mode]
    return t[0] * a + t[1] * b + t[2] * c + t[3] * d


def vc1_luma_translate(plane: np.ndarray, xq: int, yq: int, *, wrap: bool = False,
                       rnd: int = 0, out_shape: tuple[int, int] | None = None) -> np.ndarray:
    """Whole-plane VC-1 bicubic prediction at a quarter-pel offset.

    out[y, x] = luma_mc_sample(plane, x, y, xq, yq, rnd) with clamped fetch
    (or periodic fetch when ``wrap``), src/encoder_motion_compensation.cpp:367-396.
    ``out_shape`` (wrap only) asks for a window of another size.
    """
    ph, pw = plane.shape
    h, w = out_shape if out_shape is not None else (ph, pw)
    p = plane.astype(np.int32)
    ix, iy = _floor_div4(xq), _floor_div4(yq)
    hm, vm = xq - 4 * ix, yq - 4 * iy
    xs = np.arange(w) + ix
    ys = np.arange(h) + iy

    def cols(off):
        v = xs + off
        return np.mod(v, pw) if wrap else np.clip(v, 0, pw - 1)

    def rows(off):
        v = ys + off
        return np.mod(v, ph) if wrap else np.clip(v, 0, ph - 1)

    def S(dx, dy):
        return p[np.ix_(rows(dy), cols(dx))]

    if not hm and not vm:
        out = S(0, 0)
    elif hm and vm:
        shift = (_SHIFT_VALUE[hm] + _SHIFT_VALUE[vm]) >> 1
        r1 = (1 << (shift - 1)) + rnd - 1
        t = [_arshift(_filter4(S(k, -1), S(k, 0), S(k, 1), S(k, 2), vm) + r1, shift)
             for k in (-1, 0, 1, 2)]
        out = _arshift(_filter4(t[0], t[1], t[2], t[3], hm) + 64 - rnd, 7)
    elif vm:
        bias = (7 + rnd) if vm == 2 else (31 + rnd)
        out = _arshift(_filter4(S(0, -1), S(0, 0), S(0, 1), S(0, 2), vm) + bias, 4 if vm == 2 else 6)
    else:
        bias = (8 - rnd) if hm == 2 else (32 - rnd)
        out = _arshift(_filter4(S(-1, 0), S(0, 0), S(1, 0), S(2, 0), hm) + bias, 4 if hm == 2 else 6)
    return np.clip(out, 0, 255).astype(np.uint8)


def _chroma_mv(q: int) -> int:
    """P-picture chroma MV derivation (chroma_mc_sample, src/encoder_motion_compensation.cpp:426-435)."""
    uv = (q + (1 if (q & 3) == 3 else 0)) >> 1
    if uv < 0:
        uv += uv & 1
    else:
        uv -= uv & 1
    return uv


def vc1_chroma_translate(plane: np.ndarray, xq: int, yq: int, *, wrap: bool = False,
                         rnd: int = 0, out_shape: tuple[int, int] | None = None) -> np.ndarray:
    """Whole-plane VC-1 chroma prediction for a luma quarter-pel vector
    (chroma_mc_sample, src/encoder_motion_compensation.cpp:417-447)."""
    ph, pw = plane.shape
    h, w = out_shape if out_shape is not None else (ph, pw)
    p = plane.astype(np.int32)
    uvx, uvy = _chroma_mv(xq), _chroma_mv(yq)
    ix, iy = _floor_div4(uvx), _floor_div4(uvy)
    fx, fy = (uvx - 4 * ix) * 2, (uvy - 4 * iy) * 2
    xs = np.arange(w) + ix
    ys = np.arange(h) + iy

    def S(dx, dy):
        c = xs + dx
        r = ys + dy
        c = np.mod(c, pw) if wrap else np.clip(c, 0, pw - 1)
        r = np.mod(r, ph) if wrap else np.clip(r, 0, ph - 1)
        return p[np.ix_(r, c)]

    a = S(0, 0)
    if not fx and not fy:
        return a.astype(np.uint8)
    v = (a * (8 - fx) * (8 - fy) + S(1, 0) * fx * (8 - fy) + S(0, 1) * (8 - fx) * fy
         + S(1, 1) * fx * fy)
    return ((v + (28 if rnd else 32)) >> 6).astype(np.uint8)


def _gauss_kernel(sigma: float) -> np.ndarray:
    r = max(1, int(np.ceil(3.0 * sigma)))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def blur_reflect(img: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian blur with reflected edges, in double precision."""
    k = _gauss_kernel(sigma)
    r = len(k) // 2
    a = np.pad(img.astype(np.float64), ((0, 0), (r, r)), mode="reflect")
    a = sum(k[i] * a[:, i:i + img.shape[1]] for i in range(len(k)))
    a = np.pad(a, ((r, r), (0, 0)), mode="reflect")
    return sum(k[i] * a[i:i + img.shape[0], :] for i in range(len(k)))


def periodic_texture(rng: np.random.Generator, h: int, w: int, sigma: float, make: bool = True):
    """Uniform noise blurred by a periodic Gaussian (FFT), stretched back to
    roughly the full 8-bit range, rounded.  Tiles seamlessly.  make=False
    only consumes the random draws (returns None)."""
    noise = rng.uniform(0.0, 255.0, size=(h, w))
    if not make:
        return None
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.fftfreq(w)[None, :]
    gain = np.exp(-2.0 * (np.pi * sigma) ** 2 * (fx * fx + fy * fy))
    b = np.real(np.fft.ifft2(np.fft.fft2(noise) * gain))
    b = (b - b.mean()) / max(b.std(), 1e-9) * 40.0 + 128.0
    return np.clip(np.rint(b), 0, 255).astype(np.uint8)


def _add_noise(rng: np.random.Generator, plane: np.ndarray, sigma: float) -> np.ndarray:
    if sigma <= 0:
        return plane
    n = rng.normal(0.0, sigma, size=plane.shape)
    return np.clip(np.rint(plane.astype(np.float64) + n), 0, 255).astype(np.uint8)


@dataclasses.dataclass
class Picture:
    """One 8-bit 4:2:0 picture with tight planes (stride = width)."""
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray

    @property
    def width(self) -> int:
        return int(self.y.shape[1])

    @property
    def height(self) -> int:
        return int(self.y.shape[0])


def cif_translation(seed: int, frames: int = 16, width: int = 352, height: int = 288,
                    max_px: int = 8, noise: float = 2.0) -> list[Picture]:
    """Blurred noise translated frame to frame at quarter-pel (see module doc)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    y = np.clip(np.rint(blur_reflect(rng.uniform(0, 255, size=(height, width)), 3.0)), 0, 255).astype(np.uint8)
    u = np.clip(np.rint(blur_reflect(rng.uniform(0, 255, size=(height // 2, width // 2)), 3.0)), 0, 255).astype(np.uint8)
    v = np.clip(np.rint(blur_reflect(rng.uniform(0, 255, size=(height // 2, width // 2)), 3.0)), 0, 255).astype(np.uint8)
    out = [Picture(y, u, v)]
    for _ in range(frames - 1):
        xq = int(rng.integers(-4 * max_px, 4 * max_px + 1))
        yq = int(rng.integers(-4 * max_px, 4 * max_px + 1))
        y = _add_noise(rng, vc1_luma_translate(y, xq, yq), noise)
        u = _add_noise(rng, vc1_chroma_translate(u, xq, yq), noise)
        v = _add_noise(rng, vc1_chroma_translate(v, xq, yq), noise)
        out.append(Picture(y, u, v))
    return out


@dataclasses.dataclass
class _Layer:
    ty: np.ndarray
    tu: np.ndarray
    tv: np.ndarray
    x0q: int  # top-left position at frame 0, luma quarter-pel
    y0q: int
    vxq: int  # velocity, quarter-pel per frame
    vyq: int
    w: int
    h: int


def _new_background(rng, width, height, make=True):
    sigma = float(rng.uniform(2.0, 6.0))
    th = 1 << int(np.ceil(np.log2(height + 64)))
    tw = 1 << int(np.ceil(np.log2(width + 64)))
    return (periodic_texture(rng, th, tw, sigma, make),
            periodic_texture(rng, th // 2, tw // 2, sigma / 2, make),
            periodic_texture(rng, th // 2, tw // 2, sigma / 2, make))


def _new_rects(rng, width, height, count, max_obj_px, make=True):
    rects = []
    for _ in range(count):
        w = int(rng.integers(32, 257)) * 2
        h = int(rng.integers(32, 257)) * 2
        sigma = float(rng.uniform(1.5, 4.0))
        tw = 1 << int(np.ceil(np.log2(w)))
        th = 1 << int(np.ceil(np.log2(h)))
        rects.append(_Layer(
            periodic_texture(rng, th, tw, sigma, make),
            periodic_texture(rng, th // 2, tw // 2, sigma / 2, make),
            periodic_texture(rng, th // 2, tw // 2, sigma / 2, make),
            int(rng.integers(-w // 2, width)) * 4, int(rng.integers(-h // 2, height)) * 4,
            int(rng.integers(-4 * max_obj_px, 4 * max_obj_px + 1)),
            int(rng.integers(-4 * max_obj_px, 4 * max_obj_px + 1)),
            w, h))
    return rects


def _sample_periodic(tex: np.ndarray, oxq: int, oyq: int, h: int, w: int, chroma: bool) -> np.ndarray:
    """An h x w window of a periodic texture whose origin sits at quarter-pel
    (oxq, oyq), interpolated with the VC-1 luma or chroma law."""
    fn = vc1_chroma_translate if chroma else vc1_luma_translate
    return fn(tex, oxq, oyq, wrap=True, out_shape=(h, w))


def pan_rectangles(seed: int, frames: int, width: int = 1920, height: int = 1080,
                   max_pan_px: int = 16, max_obj_px: int = 48, rects: int = 8,
                   noise: float = 2.0, cut_every: int = 0, dup_fraction: float = 0.0,
                   frame_indices: list[int] | None = None) -> dict[int, Picture]:
    """Camera pan over a blurred-noise background with moving textured
    rectangles (see module doc).  Returns {frame index: Picture} for the
    requested indices (all frames by default)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    wanted = set(range(frames)) if frame_indices is None else set(frame_indices)
    bg = _new_background(rng, width, height)
    pan = (int(rng.integers(-4 * max_pan_px, 4 * max_pan_px + 1)),
           int(rng.integers(-4 * max_pan_px, 4 * max_pan_px + 1)))
    layers = _new_rects(rng, width, height, rects, max_obj_px)
    dup = rng.uniform(0.0, 1.0, size=frames) < dup_fraction
    dup[0] = False
    noise_seeds = rng.integers(0, 2**63 - 1, size=frames)
    out: dict[int, Picture] = {}
    scene_start = 0
    last: Picture | None = None
    last_needed = max(wanted) if wanted else -1
    for k in range(last_needed + 1):
        if cut_every and k and k % cut_every == 0:
            bg = _new_background(rng, width, height)
            layers = _new_rects(rng, width, height, rects, max_obj_px)
            scene_start = k
        if dup[k] and last is not None:
            if k in wanted:
                out[k] = Picture(last.y.copy(), last.u.copy(), last.v.copy())
            continue
        if k not in wanted and not (k + 1 < frames and dup[k + 1]):
            continue  # no duplicate depends on this frame: skip it
        t = k - scene_start
        y = _sample_periodic(bg[0], t * pan[0], t * pan[1], height, width, False)
        u = _sample_periodic(bg[1], t * pan[0], t * pan[1], height // 2, width // 2, True)
        v = _sample_periodic(bg[2], t * pan[0], t * pan[1], height // 2, width // 2, True)
        for L in layers:
            span_x = (width + L.w) * 4
            span_y = (height + L.h) * 4
            pxq = (L.x0q + t * L.vxq + L.w * 4) % span_x - L.w * 4
            pyq = (L.y0q + t * L.vyq + L.h * 4) % span_y - L.h * 4
            x0, y0 = pxq // 4, pyq // 4
            fxq, fyq = pxq - 4 * x0, pyq - 4 * y0
            xa, xb = max(0, x0), min(width, x0 + L.w)
            ya, yb = max(0, y0), min(height, y0 + L.h)
            if xa >= xb or ya >= yb:
                continue
            patch = _sample_periodic(L.ty, (xa - x0) * 4 - fxq, (ya - y0) * 4 - fyq, yb - ya, xb - xa, False)
            y[ya:yb, xa:xb] = patch
            cxa, cxb, cya, cyb = xa // 2, (xb + 1) // 2, ya // 2, (yb + 1) // 2
            cx0, cy0 = x0 // 2, y0 // 2
            if cxa < cxb and cya < cyb:
                u[cya:cyb, cxa:cxb] = _sample_periodic(L.tu, (cxa - cx0) * 4 - fxq, (cya - cy0) * 4 - fyq,
                                                       cyb - cya, cxb - cxa, True)
                v[cya:cyb, cxa:cxb] = _sample_periodic(L.tv, (cxa - cx0) * 4 - fxq, (cya - cy0) * 4 - fyq,
                                                       cyb - cya, cxb - cxa, True)
        nrng = np.random.Generator(np.random.PCG64(int(noise_seeds[k])))
        pic = Picture(_add_noise(nrng, y, noise), _add_noise(nrng, u, noise), _add_noise(nrng, v, noise))
        last = pic
        if k in wanted:
            out[k] = pic
    return out


def scene_events(seed: int, frames: int, width: int = 1920, height: int = 1080, max_pan_px: int = 16,
                 max_obj_px: int = 48, rects: int = 8, cut_every: int = 0,
                 dup_fraction: float = 0.0) -> dict[str, list[int]]:
    """The frames of pan_rectangles(seed, frames, ...) that start a new scene
    ("cuts") and that repeat the previous frame exactly ("duplicates"),
    found by replaying the generator's random draws without building any
    texture."""
    rng = np.random.Generator(np.random.PCG64(seed))
    _new_background(rng, width, height, make=False)
    rng.integers(-4 * max_pan_px, 4 * max_pan_px + 1)
    rng.integers(-4 * max_pan_px, 4 * max_pan_px + 1)
    _new_rects(rng, width, height, rects, max_obj_px, make=False)
    dup = rng.uniform(0.0, 1.0, size=frames) < dup_fraction
    dup[0] = False
    cuts = [k for k in range(1, frames) if cut_every and k % cut_every == 0]
    return {"cuts": cuts, "duplicates": [int(k) for k in np.nonzero(dup)[0]]}


def ibbp_p_pairs(frames: int, gop: int = 15, spacing: int = 3) -> list[tuple[int, int]]:
    """(current, reference) frame indices of the P pictures of an IBBP GOP
    structure: P at offsets spacing, 2*spacing, ... of each GOP, each
    predicted from the anchor `spacing` frames back."""
    pairs = []
    for g0 in range(0, frames, gop):
        off = spacing
        while off < gop and g0 + off < frames:
            pairs.append((g0 + off, g0 + off - spacing))
            off += spacing
    return pairs


def p_only_pairs(frames: int) -> list[tuple[int, int]]:
    return [(k, k - 1) for k in range(1, frames)]


def fade(pic: Picture, gain: float, offset: float) -> Picture:
    """A brightness fade of `pic`: luma gain * Y + offset, chroma pulled
    toward 128 by the same gain, rounded and clipped to 8 bits."""
    y = np.clip(np.rint(gain * pic.y.astype(np.float64) + offset), 0, 255).astype(np.uint8)
    uv = lambda p: np.clip(np.rint(128.0 + gain * (p.astype(np.float64) - 128.0)), 0, 255).astype(np.uint8)
    return Picture(y, uv(pic.u), uv(pic.v))


def zoom(pic: Picture, scale: float) -> Picture:
    """`pic` magnified by `scale` about its centre (bilinear resampling,
    clamped edges): a camera zoom that no single translation models."""
    def resample(p: np.ndarray) -> np.ndarray:
        h, w = p.shape
        ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
        sy = (ys - (h - 1) / 2.0) / scale + (h - 1) / 2.0
        sx = (xs - (w - 1) / 2.0) / scale + (w - 1) / 2.0
        y0 = np.clip(np.floor(sy).astype(int), 0, h - 1)
        x0 = np.clip(np.floor(sx).astype(int), 0, w - 1)
        y1, x1 = np.clip(y0 + 1, 0, h - 1), np.clip(x0 + 1, 0, w - 1)
        fy, fx = np.clip(sy - y0, 0, 1), np.clip(sx - x0, 0, 1)
        a = p.astype(np.float64)
        v = (a[y0, x0] * (1 - fx) + a[y0, x1] * fx) * (1 - fy) + (a[y1, x0] * (1 - fx) + a[y1, x1] * fx) * fy
        return np.clip(np.rint(v), 0, 255).astype(np.uint8)
    return Picture(resample(pic.y), resample(pic.u), resample(pic.v))
