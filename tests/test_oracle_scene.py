"""The oracle's scene_change_score (src/encoder_motion.cpp:599-676) against
scores written by the compiled reference (tests/golden/scene.json: a pan,
a cut, a fade with and without compensation, and two zooms that the
local-motion rescue accepts below the threshold) and, where it is built,
against the reference on more thresholds."""
import json
import os

import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi
from _scene_cases import scene_inputs

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene.json")))
needs_ref = pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref not built (needs the reference sources)")
PARAMS = dict(pqindex=8, motion_search_range=128, motion_local_search_range=32)


def bits(v):
    return int(np.float64(v).view(np.uint64))


@pytest.mark.parametrize("k", range(len(FIX)), ids=lambda k: f"{FIX[k]['case']}-f{FIX[k]['fade']}-t{FIX[k]['threshold']:g}")
def test_oracle_scene_matches_fixture(k):
    f = FIX[k]
    cur, ref = scene_inputs(f["case"])
    v = O.oracle_scene_change_score(abi.default_params(320, 176, **PARAMS), cur, ref, f["fade"], f["threshold"])
    assert bits(v) == f["score_bits"], (v, f["score"])


def test_scene_fixtures_cover_the_paths():
    by = {(f["case"], f["fade"], f["threshold"]): f["score"] for f in FIX}
    # fade compensation lowers a fade's score; the zoom rescue replaces the
    # score with the local MAD only below a low threshold
    assert by[("fade", 1, 28.0)] < by[("fade", 0, 28.0)]
    assert by[("zoom_sharp", 1, 6.0)] < 6.0 * 0.86 < by[("zoom_sharp", 1, 28.0)]
    assert by[("cut", 1, 6.0)] == by[("cut", 1, 28.0)]


@needs_ref
@pytest.mark.parametrize("case", ["consecutive", "cut", "fade", "zoom_sharp", "zoom_pan"])
def test_oracle_scene_matches_reference(case):
    cur, ref = scene_inputs(case)
    for fade, thr in ((1, 3.0), (0, 10.0), (1, 40.0), (0, 0.0)):
        for rng in (0, 24, 128):
            p = abi.default_params(320, 176, **dict(PARAMS, motion_search_range=rng))
            a = O.oracle_scene_change_score(p, cur, ref, fade, thr)
            b = O.reference_scene_change_score(p, cur, ref, fade, thr)
            assert bits(a) == bits(b), (case, fade, thr, rng, a, b)
