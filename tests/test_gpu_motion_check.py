"""The motion-failure check on the GPU (vc1me_has_good_motion_match,
src/encoder_motion.cpp:678-751) through the C ABI: the answer against the
reference-written fixtures, and every macroblock's refined vector against
the oracle's scan run to the end, under both macroblock schedules."""
import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi, synth
from _motion_cases import H, W, motion_inputs
from test_oracle_motion_check import FIX, good_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _check(backend, cur, ref, p, fade):
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        return backend.has_good_motion_match(fc, fr, p, fade, want_field=True)
    finally:
        fc.release()
        fr.release()


@pytest.mark.parametrize("k", range(len(FIX["good"])),
                         ids=lambda k: "{case}-f{fade}-lr{long_range_search_mode}-r{range}".format(**FIX["good"][k]))
def test_good_match_matches_fixture(backend, k):
    g = FIX["good"][k]
    _say("[case] motion check {case} fade={fade} lr={long_range_search_mode} range={range}".format(**g))
    cur, ref = motion_inputs(g["case"])
    got, _ = _check(backend, cur, ref, good_params(g), g["fade"])
    assert got == g["good"]


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_RELAX, abi.SCHEDULE_DIAGONAL], ids=["relax", "diagonal"])
def test_motion_check_field_matches_oracle(backend, schedule):
    backend.set_schedule(schedule)
    try:
        for name, fast, algo, rng, fade in (("pan", 0, 0, 256, 1), ("far", 1, 0, 512, 0), ("fade", 0, 1, 48, 1),
                                            ("late_match", 1, 0, 128, 1)):
            p = abi.default_params(W, H, pqindex=8, motion_search_range=rng, fast_subpixel_refinement=fast,
                                   me_algorithm=algo)
            cur, ref = motion_inputs(name)
            _say(f"[case] motion-check field {name} fast={fast} algo={algo} range={rng} schedule={schedule}")
            got, field = _check(backend, cur, ref, p, fade)
            want, wfield = O.oracle_motion_check_field(p, cur, ref, fade)
            assert got == want
            bad = np.argwhere((field != wfield).any(axis=1))[:, 0]
            assert not bad.size, f"{bad.size} MBs differ, first {bad[:5].tolist()}: {field[bad[:5]].tolist()} vs {wfield[bad[:5]].tolist()}"
    finally:
        backend.set_schedule(abi.SCHEDULE_RELAX)


def test_motion_check_odd_display_raster(backend):
    d = synth.pan_rectangles(98, 3, width=200, height=120, frame_indices=[0, 2])
    p = abi.default_params(200, 120, pqindex=8, motion_search_range=96, display_width=194, display_height=114)
    _say("[case] motion check 200x120 display 194x114")
    got, field = _check(backend, d[2], d[0], p, True)
    want, wfield = O.oracle_motion_check_field(p, d[2], d[0], True)
    assert got == want and np.array_equal(field, wfield)
