"""The 2160p hard-motion workload's cases (extended MV range 1024): the
3840x2160 geometry, motion far beyond the local range, a hard scene cut and
an exact duplicate frame, bit-exact against the CPU oracle through the C ABI;
and the learned step on two 2160p streams against the plain one.

The 2160p pictures are built from a quarter-resolution smoothed-noise texture
(pixel-repeated) and integer translations, which exercises the same geometry
as the content generator at a fraction of its cost; the hard-motion cases use
the content generator at 640x368.
"""
import numpy as np
import pytest

from libvc1_amd import abi, synth
from _oracle import describe_diff, oracle_analyze

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _texture(rng, h, w):
    # quarter-resolution noise, 3x3 box-smoothed, pixel-repeated 4x
    small = rng.integers(0, 256, size=(h // 4 + 4, w // 4 + 4)).astype(np.int32)
    box = sum(small[dy:dy + h // 4 + 2, dx:dx + w // 4 + 2] for dy in range(3) for dx in range(3))
    return np.repeat(np.repeat((box // 9).astype(np.uint8), 4, axis=0), 4, axis=1)


def _uhd_pair(seed, pan=(37, -11), obj=(-150, 60)):
    """3840x2160 (cur, ref): the background moved by `pan` pixels, a
    768x512 object by `obj` pixels, noise sigma 2."""
    rng = np.random.default_rng(seed)
    H, W = 2160, 3840
    big = _texture(rng, H + 400, W + 400)
    ref_y = big[200:200 + H, 200:200 + W].copy()
    cur_y = big[200 - pan[1]:200 - pan[1] + H, 200 - pan[0]:200 - pan[0] + W].copy()
    ox, oy = 1400, 900
    tex = _texture(rng, 512 + 400, 768 + 400)
    ref_y[oy:oy + 512, ox:ox + 768] = tex[200:712, 200:968]
    cur_y[oy + obj[1]:oy + obj[1] + 512, ox + obj[0]:ox + obj[0] + 768] = tex[200:712, 200:968]
    noise = lambda a: np.clip(a.astype(np.int32) + np.rint(rng.normal(0, 2, a.shape)).astype(np.int32), 0, 255).astype(np.uint8)
    chroma = lambda: np.full((H // 2, W // 2), 128, np.uint8)
    return synth.Picture(noise(cur_y), chroma(), chroma()), synth.Picture(noise(ref_y), chroma(), chroma())


def _params(w, h):
    return abi.default_params(w, h, pqindex=8, motion_search_range=1024, motion_local_search_range=32,
                              mode=abi.MODE_MIXED_MV)


def _check(backend, cur, ref, p, label):
    _say(f"[case] {label} {p.width}x{p.height} R={p.motion_search_range}")
    want = oracle_analyze(p, cur, ref)
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        got = backend.analyze(fc, fr, p)
    finally:
        fc.release()
        fr.release()
    assert not got.diff(want), f"{label}\n" + describe_diff(got, want, p.width)
    return got


def test_uhd_pair_extended_range(backend):
    cur, ref = _uhd_pair(501)
    got = _check(backend, cur, ref, _params(3840, 2160), "2160p pan + far object")
    assert got.mbs == 240 * 135


def test_hard_motion_cut_and_duplicate(backend):
    """Motion up to 200 px per frame, a hard cut to a fresh scene every third
    frame, and an exact duplicate."""
    d = synth.pan_rectangles(502, 5, width=640, height=368, max_pan_px=200, max_obj_px=200, cut_every=3,
                             frame_indices=[0, 1, 2, 3, 4])
    p = _params(640, 368)
    _check(backend, d[1], d[0], p, "fast motion")
    _check(backend, d[3], d[2], p, "scene cut")
    got = _check(backend, d[4], synth.Picture(d[4].y.copy(), d[4].u.copy(), d[4].v.copy()), p, "duplicate")
    assert not got.mvs.any() and not got.screen_sads.any()


def test_uhd_learned_streams_keep_cost(backend):
    """The learned step on two 2160p streams: every macroblock gets a
    result, and its SAD + lambda*R of the final 1-MV vectors stays within 1%
    of the plain step's."""
    from libvc1_amd import learned
    from libvc1_amd.backend import ProposalNet
    pairs = [_uhd_pair(503, pan=(12, 5), obj=(-40, 20)), _uhd_pair(504, pan=(-20, 9), obj=(30, -25))]
    p = _params(3840, 2160)
    frames = [(backend.upload(c), backend.upload(r)) for c, r in pairs]
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    _say("[case] learned step, two 2160p streams")
    try:
        r = learned.cost_compare(backend, net, frames, [p, p], 0.5)
    finally:
        net.close()
        for a, b in frames:
            a.release()
            b.release()
    _say(f"   {r}")
    assert r["macroblocks"] == 2 * 240 * 135
    assert r["cost_rise"] <= 0.01
