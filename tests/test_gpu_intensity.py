"""Intensity compensation on the GPU (vc1me_estimate_intensity_comp,
vc1me_frame_intensity_compensated; src/encoder_motion_compensation.cpp:175-348)
through the C ABI: bit-exact against the reference-written fixtures and the
oracle, and the compensated reference feeding the core as the reference's
weighted pass does (src/encoder_motion.cpp:573-596)."""
import ctypes as C

import numpy as np
import pytest

from libvc1_amd import abi, synth
from libvc1_amd.backend import BackendError
from _oracle import describe_diff, oracle_analyze, oracle_estimate_intensity, oracle_intensity_compensate
from test_oracle_intensity import FIXTURES, intensity_inputs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _estimate(backend, cur, ref, mvs):
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        return backend.estimate_intensity(fc, fr, mvs)
    finally:
        fc.release()
        fr.release()


@pytest.mark.parametrize("k", range(len(FIXTURES["estimate"])))
def test_estimate_matches_fixture(backend, k):
    case = FIXTURES["estimate"][k]
    _say(f"[case] intensity estimate fixture {k} {case['kind']} {case['size']}")
    cur, ref, mvs = intensity_inputs(case)
    assert list(_estimate(backend, cur, ref, mvs)) == case["expected"]


def test_estimate_matches_oracle_odd_rasters_and_far_vectors(backend):
    rng = np.random.default_rng(12)
    for w, h in ((1920, 1080), (210, 94), (18, 10), (8, 8), (34, 6)):
        base = synth.Picture(rng.integers(10, 245, size=(h, w)).astype(np.uint8),
                             rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8),
                             rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8))
        cur = synth.fade(base, 0.82, 17.0)
        mbw, mbh = abi.mb_grid(w, h)
        for mvs in (np.zeros((mbw * mbh, 2), np.int32), rng.integers(-5000, 5000, size=(mbw * mbh, 2)).astype(np.int32)):
            _say(f"[case] intensity estimate {w}x{h}")
            assert _estimate(backend, cur, base, mvs) == oracle_estimate_intensity(cur, base, mvs)


def test_compensated_frames_match_oracle(backend):
    """Every luma and chroma map value of many codes, including LUMSCALE 0
    (inverting) and disabled; rasters whose plane bytes end mid-word."""
    rng = np.random.default_rng(13)
    codes = [(0, 0, 0), (0, 19, 10)] + [(1, 0, s) for s in range(0, 64, 7)] + \
            [(1, l, s) for l in range(1, 64, 5) for s in (0, 1, 31, 32, 63)]
    for w, h in ((32, 16), (18, 10), (1920, 1080), (66, 50)):
        pic = synth.Picture(rng.integers(0, 256, size=(h, w)).astype(np.uint8),
                            rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8),
                            rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8))
        f = backend.upload(pic)
        try:
            for ic in (codes if w * h <= 4096 else codes[:6]):
                _say(f"[case] compensate {w}x{h} {ic}")
                g = backend.intensity_compensated(f, ic)
                try:
                    got = backend.download(g)
                finally:
                    g.release()
                want = oracle_intensity_compensate(pic, ic)
                assert all(np.array_equal(a, b) for a, b in zip(got, want)), (w, h, ic)
        finally:
            f.release()


def test_weighted_pass_on_compensated_reference(backend):
    """The reference's fade path: estimate on the chosen pass's vectors, build
    the compensated reference, and run the core on it; the result equals the
    oracle's core on the oracle's compensated planes."""
    d = synth.pan_rectangles(404, 4, width=320, height=176, frame_indices=[0, 3])
    cur = synth.fade(d[3], 0.75, 20.0)
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=128, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    fc, fr = backend.upload(cur), backend.upload(d[0])
    try:
        first = backend.analyze(fc, fr, p)
        ic = backend.estimate_intensity(fc, fr, first.mvs)
        assert ic == oracle_estimate_intensity(cur, d[0], first.mvs) and ic[0] == 1
        _say(f"[case] weighted pass with {ic}")
        comp = backend.intensity_compensated(fr, ic)
        try:
            got = backend.analyze(fc, comp, p)
        finally:
            comp.release()
    finally:
        fc.release()
        fr.release()
    y, u, v = oracle_intensity_compensate(d[0], ic)
    want = oracle_analyze(p, cur, synth.Picture(y, u, v))
    assert not got.diff(want), describe_diff(got, want, 320)
    assert got.struct.motion_sad < first.struct.motion_sad


def test_intensity_argument_errors(backend):
    z = lambda h, w: np.zeros((h, w), np.uint8)
    a = backend.upload(synth.Picture(z(16, 32), z(8, 16), z(8, 16)))
    b = backend.upload(synth.Picture(z(16, 16), z(8, 8), z(8, 8)))
    try:
        out = abi.Intensity()
        mvs = np.zeros(4, np.int32)
        rc = backend.lib.vc1me_estimate_intensity_comp(backend.ctx, a.handle, b.handle,
                                                        mvs.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(out))
        assert rc == abi.VC1ME_EINVAL                      # rasters differ
        with pytest.raises(BackendError) as e:
            backend.intensity_compensated(a, (1, 64, 0))  # not a 6-bit code
        assert e.value.code == abi.VC1ME_EINVAL
        with pytest.raises(ValueError):
            backend.estimate_intensity(a, a, np.zeros((3, 2), np.int32))
    finally:
        a.release()
        b.release()
