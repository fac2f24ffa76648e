"""B-picture analysis on the GPU (vc1me_analyze_b_frame,
src/encoder_motion.cpp:753-1004) through the C ABI: bit-exact against the
reference-written fixtures and the oracle, under both macroblock schedules."""
import numpy as np
import pytest

from libvc1_amd import abi, synth
from libvc1_amd.backend import BackendError
from _oracle import oracle_analyze_b
from test_oracle_bframes import B_FILES, describe_b, load_b

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _b(backend, cur, past, fut, p, binp):
    fs = [backend.upload(x) for x in (cur, past, fut)]
    try:
        return backend.analyze_b(fs[0], fs[1], fs[2], p, binp)
    finally:
        for f in fs:
            f.release()


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_RELAX, abi.SCHEDULE_DIAGONAL], ids=["relax", "diagonal"])
@pytest.mark.parametrize("path", B_FILES, ids=lambda p: p.rsplit("/", 1)[-1][:-4])
def test_b_matches_reference_fixture(backend, path, schedule):
    cur, past, fut, p, binp, want = load_b(path)
    _say(f"[case] B fixture {path.rsplit('/', 1)[-1]} schedule={schedule}")
    backend.set_schedule(schedule)
    try:
        got = _b(backend, cur, past, fut, p, binp)
    finally:
        backend.set_schedule(abi.SCHEDULE_RELAX)
    assert not got.diff(want), describe_b(got, want)


def test_b_ibbp_group_matches_oracle(backend):
    """An IBBP group at 480x272: the P anchor analysed on the GPU, then both B
    pictures (1/3 qpel and 2/3 bilinear half-pel) in one batched step with a
    fixture's picture of another raster, each against the oracle (or the
    reference-written fixture)."""
    d = synth.pan_rectangles(70, 4, width=480, height=272, max_pan_px=12, max_obj_px=30, frame_indices=[0, 1, 2, 3])
    base = dict(pqindex=9, motion_search_range=256, motion_local_search_range=32)
    fs = {k: backend.upload(v) for k, v in d.items()}
    fc, fp, ff, fp_, fbin, fwant = load_b(B_FILES[0])
    extra = [backend.upload(x) for x in (fc, fp, ff)]
    try:
        anchor = backend.analyze(fs[3], fs[0], abi.default_params(480, 272, mode=abi.MODE_MIXED_MV, **base))
        items, wants = [], []
        for k, frac, mode in ((1, (1, 3), abi.MODE_ONE_MV_QPEL), (2, (2, 3), abi.MODE_ONE_MV_HPEL_BILINEAR)):
            p = abi.default_params(480, 272, mode=mode, **base)
            binp = abi.BInputs(anchor.mvs, anchor.use_4mv, fraction=frac)
            items.append((fs[k], fs[0], fs[3], p, binp))
            wants.append(oracle_analyze_b(p, d[k], d[0], d[3], binp))
        items.append((extra[0], extra[1], extra[2], fp_, fbin))
        wants.append(fwant)
        _say("[case] B step: two 480x272 pictures and one 176x144 fixture")
        got = backend.analyze_b_step(items)
        for g, w in zip(got, wants):
            assert not g.diff(w), describe_b(g, w)
        _say("[case] the same pictures one by one")
        for (c, pa, fu, p, binp), w in zip(items, wants):
            g = backend.analyze_b(c, pa, fu, p, binp)
            assert not g.diff(w), describe_b(g, w)
    finally:
        for f in list(fs.values()) + extra:
            f.release()


def test_b_shortcut_same_frames_and_errors(backend):
    d = synth.pan_rectangles(71, 3, width=64, height=48, frame_indices=[0, 1, 2])
    z = np.zeros((12, 2), np.int32)
    p = abi.default_params(64, 48, motion_search_range=0)
    _say("[case] B shortcut")
    r = _b(backend, d[1], d[0], d[2], p, abi.BInputs(z, np.zeros(12, np.uint8), zero_motion_shortcut=1))
    assert (r.modes == abi.B_FORWARD).all() and r.struct.forward_macroblocks == 12
    # past and future the same resident frame, intensity on the forward side
    q = abi.default_params(64, 48, motion_search_range=32)
    binp = abi.BInputs(z, np.zeros(12, np.uint8), intensity=(1, 36, 2), fraction=(1, 2))
    fc, fr = backend.upload(d[1]), backend.upload(d[0])
    try:
        _say("[case] B with one reference frame for both directions")
        got = backend.analyze_b(fc, fr, fr, q, binp)
        with pytest.raises(BackendError) as e:
            backend.analyze_b(fc, fr, fr, abi.default_params(64, 48, mode=abi.MODE_MIXED_MV), binp)
        assert e.value.code == abi.VC1ME_EINVAL
        with pytest.raises(BackendError):
            backend.analyze_b(fc, fr, fr, q, abi.BInputs(z, np.zeros(12, np.uint8), fraction=(3, 4)))
    finally:
        fc.release()
        fr.release()
    want = oracle_analyze_b(q, d[1], d[0], d[0], binp)
    assert not got.diff(want), describe_b(got, want)
