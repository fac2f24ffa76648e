"""The oracle's motion-failure check (has_good_motion_match,
src/encoder_motion.cpp:678-751) and its SAD-only refinement
(refine_qpel_block, :3596-3685) against fixtures written by the compiled
reference and, where it is built, against the reference itself."""
import json
import os

import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi
from _motion_cases import H, W, motion_inputs

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "motion_check.json")))
needs_ref = pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref not built (needs the reference sources)")


def good_params(g):
    return abi.default_params(W, H, pqindex=8, motion_search_range=g["range"],
                              long_range_search_mode=g["long_range_search_mode"])


def refine_params(r):
    return abi.default_params(W, H, pqindex=8, fast_subpixel_refinement=r["fast"], subpixel_refinement=r["subpel"],
                              subpixel_refinement_range=r["subpel_range"], display_width=r["display_width"])


@pytest.mark.parametrize("k", range(len(FIX["good"])),
                         ids=lambda k: "{case}-f{fade}-lr{long_range_search_mode}-r{range}".format(**FIX["good"][k]))
def test_oracle_good_match_matches_fixture(k):
    g = FIX["good"][k]
    cur, ref = motion_inputs(g["case"])
    assert O.oracle_has_good_motion_match(good_params(g), cur, ref, g["fade"]) == g["good"]


def test_good_match_fixtures_have_both_answers():
    answers = {g["good"] for g in FIX["good"]}
    assert answers == {True, False}


def test_oracle_refine_qpel_matches_fixture():
    cur, ref = motion_inputs("pan")
    bad = []
    for r in FIX["refine"]:
        got = O.oracle_refine_qpel_block(refine_params(r), cur, ref, r["x0"], r["y0"], r["bw"], r["bw"],
                                         tuple(r["seed"]), r["range"], r["known"])
        if list(got) != r["expected"]:
            bad.append((r, got))
    assert not bad, bad[:3]


@needs_ref
def test_oracle_refine_qpel_matches_reference_fuzz():
    cur, ref = motion_inputs("far")
    rs = np.random.default_rng(5)
    for t in range(300):
        bw = 8 if t % 3 == 0 else 16
        p = abi.default_params(W, H, pqindex=8, fast_subpixel_refinement=t % 2, display_width=W - 10 if t % 5 == 0 else 0)
        args = (int(rs.integers(0, W // bw)) * bw, int(rs.integers(0, H // bw)) * bw, bw, bw,
                (int(rs.integers(-200, 200)), int(rs.integers(-120, 120))), int(rs.choice([1, 16, 128, 1024])),
                2**64 - 1 if t % 3 else int(rs.integers(0, 9000)))
        assert O.oracle_refine_qpel_block(p, cur, ref, *args) == O.reference_refine_qpel_block(p, cur, ref, *args), args


@needs_ref
@pytest.mark.parametrize("case", ["pan", "late_match", "noise_vs_flat"])
def test_oracle_good_match_matches_reference(case):
    cur, ref = motion_inputs(case)
    for fast, algo, rng in ((1, 0, 256), (0, 1, 48), (1, 1, 0)):
        p = abi.default_params(W, H, pqindex=8, motion_search_range=rng, fast_subpixel_refinement=fast,
                               me_algorithm=algo)
        for fade in (0, 1):
            assert O.oracle_has_good_motion_match(p, cur, ref, fade) == O.reference_has_good_motion_match(p, cur, ref, fade)
