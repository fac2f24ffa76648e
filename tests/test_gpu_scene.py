"""scene_change_score on the GPU (vc1me_scene_change_score,
src/encoder_motion.cpp:599-676) through the C ABI, bit-exact against the
reference-written scores and the oracle."""
import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi, synth
from libvc1_amd.backend import BackendError
from _scene_cases import scene_inputs
from test_oracle_scene import FIX, PARAMS, bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _score(backend, cur, ref, p, fade, thr):
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        return backend.scene_change_score(fc, fr, p, fade, thr)
    finally:
        fc.release()
        fr.release()


@pytest.mark.parametrize("k", range(len(FIX)), ids=lambda k: f"{FIX[k]['case']}-f{FIX[k]['fade']}-t{FIX[k]['threshold']:g}")
def test_scene_matches_fixture(backend, k):
    f = FIX[k]
    _say(f"[case] scene fixture {f['case']} fade={f['fade']} thr={f['threshold']}")
    cur, ref = scene_inputs(f["case"])
    v = _score(backend, cur, ref, abi.default_params(320, 176, **PARAMS), f["fade"], f["threshold"])
    assert bits(v) == f["score_bits"], (v, f["score"])


def test_scene_ranges_and_rasters_match_oracle(backend):
    d = synth.pan_rectangles(95, 3, width=640, height=368, max_pan_px=24, frame_indices=[0, 2])
    for cur, ref, label in ((d[2], d[0], "pan"), (synth.fade(d[2], 0.7, 20.0), d[0], "fade"),
                            (synth.zoom(d[0], 1.03), d[0], "zoom")):
        for rng, thr in ((0, 28.0), (16, 5.0), (512, 2.0)):
            p = abi.default_params(640, 368, pqindex=8, motion_search_range=rng, motion_local_search_range=32)
            _say(f"[case] scene 640x368 {label} range {rng} thr {thr}")
            v = _score(backend, cur, ref, p, True, thr)
            w = O.oracle_scene_change_score(p, cur, ref, True, thr)
            assert bits(v) == bits(w), (label, rng, thr, v, w)


def test_scene_then_analysis_on_the_same_context(backend):
    """The score's global-only search leaves the context usable: a P
    analysis right after it equals the oracle's."""
    cur, ref = scene_inputs("consecutive")
    p = abi.default_params(320, 176, **PARAMS)
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        backend.scene_change_score(fc, fr, p, True, 3.0)
        got = backend.analyze(fc, fr, p)
        with pytest.raises(BackendError) as e:
            backend.scene_change_score(fc, fr, abi.default_params(64, 48), True, 28.0)
        assert e.value.code == abi.VC1ME_EINVAL
    finally:
        fc.release()
        fr.release()
    want = O.oracle_analyze(p, cur, ref)
    assert not got.diff(want), O.describe_diff(got, want, 320)
