"""The oracle's P prediction picture (motion_compensate_mb /
motion_compensate_4mv, src/encoder_motion_compensation.cpp:474-571) against
digests of the compiled reference's pictures (scalar and SIMD paths) and,
where it is built, against the reference itself."""
import hashlib
import json
import os

import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi, synth
from _predict_cases import bi_inputs, result_from_fixture

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = json.load(open(os.path.join(GOLDEN, "predict.json")))
BI_FIX = json.load(open(os.path.join(GOLDEN, "predict_bi.json")))
needs_ref = pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref not built (needs the reference sources)")


def fixture_ref(f):
    return synth.pan_rectangles(sum(map(ord, f["case"])) % 1000, 4, width=f["width"], height=f["height"],
                                frame_indices=[0])[0]


def digest(planes):
    return hashlib.sha256(b"".join(p.tobytes() for p in planes)).hexdigest()


@pytest.mark.parametrize("k", range(len(FIX)), ids=lambda k: FIX[k]["case"])
def test_oracle_prediction_matches_fixture(k):
    f = FIX[k]
    got = O.oracle_predict_p_frame(fixture_ref(f), f["mode"], result_from_fixture(f))
    assert digest(got) == f["sha256"]


def test_prediction_closed_forms():
    """Zero vectors copy the reference; a whole-pel vector of (+8, -4) px is
    a shift with clamped edges; chroma of a 4-MV macroblock follows the
    median-4 vector."""
    rng = np.random.default_rng(8)
    ref = synth.Picture(rng.integers(0, 256, size=(48, 64)).astype(np.uint8),
                        rng.integers(0, 256, size=(24, 32)).astype(np.uint8),
                        rng.integers(0, 256, size=(24, 32)).astype(np.uint8))
    res = abi.ResultArrays(64, 48)
    got = O.oracle_predict_p_frame(ref, abi.MODE_ONE_MV_QPEL, res)
    assert all(np.array_equal(a, b) for a, b in zip(got, (ref.y, ref.u, ref.v)))
    res.mvs[:] = (32, -16)
    y, u, _ = O.oracle_predict_p_frame(ref, abi.MODE_ONE_MV_QPEL, res)
    ys, xs = np.clip(np.arange(48) - 4, 0, 47), np.clip(np.arange(64) + 8, 0, 63)
    assert np.array_equal(y, ref.y[np.ix_(ys, xs)])
    cys, cxs = np.clip(np.arange(24) - 2, 0, 23), np.clip(np.arange(32) + 4, 0, 31)
    assert np.array_equal(u, ref.u[np.ix_(cys, cxs)])
    res.mvs[:] = 0
    res.use_4mv[:] = 1
    res.block_mvs[:] = [(32, 0), (0, -16), (32, -16), (64, 16)]   # medians: x (32+32)/2, y (-16+0)/2
    _, u4, _ = O.oracle_predict_p_frame(ref, abi.MODE_MIXED_MV, res)
    cys, cxs = np.clip(np.arange(24) - 1, 0, 23), np.clip(np.arange(32) + 4, 0, 31)
    assert np.array_equal(u4, ref.u[np.ix_(cys, cxs)])


@needs_ref
def test_oracle_prediction_matches_reference_wild_fields():
    rng = np.random.default_rng(9)
    d = synth.pan_rectangles(31, 2, width=176, height=144, frame_indices=[0])
    for mode in (0, 1, 2, 3):
        res = abi.ResultArrays(176, 144)
        res.mvs[:] = rng.integers(-400, 400, size=res.mvs.shape)
        if mode in (1, 2):
            res.mvs &= ~1
        res.block_mvs[:] = rng.integers(-400, 400, size=res.block_mvs.shape)
        res.use_4mv[:] = rng.integers(0, 2, size=res.use_4mv.shape) if mode == 3 else 0
        p = abi.default_params(176, 144, mode=mode)
        for simd in (0, 1):
            want = O.reference_predict_p_frame(p, d[0], res, simd)
            assert digest(O.oracle_predict_p_frame(d[0], mode, res)) == digest(want), (mode, simd)


@pytest.mark.parametrize("k", range(len(BI_FIX)), ids=lambda k: BI_FIX[k]["case"])
def test_oracle_bi_prediction_matches_fixture(k):
    f = BI_FIX[k]
    past, fut, fmv, bmv = bi_inputs(f["case"], f["width"], f["height"], f["mode"])
    assert digest(O.oracle_predict_bi_frame(past, fut, f["mode"], fmv, bmv)) == f["sha256"]


def test_bi_prediction_closed_form():
    """Equal pictures and equal whole-pel vectors average to the plain
    prediction; zero vectors average the two pictures."""
    rng = np.random.default_rng(11)
    a = synth.Picture(*(rng.integers(0, 256, size=s).astype(np.uint8) for s in ((32, 48), (16, 24), (16, 24))))
    b = synth.Picture(*(rng.integers(0, 256, size=s).astype(np.uint8) for s in ((32, 48), (16, 24), (16, 24))))
    z = np.zeros((6, 2), np.int32)
    got = O.oracle_predict_bi_frame(a, b, abi.MODE_ONE_MV_QPEL, z, z)
    for g, x, y in zip(got, (a.y, a.u, a.v), (b.y, b.u, b.v)):
        assert np.array_equal(g, ((x.astype(int) + y + 1) >> 1).astype(np.uint8))
    w = np.tile(np.array([[8, -4]], np.int32), (6, 1))
    res = abi.ResultArrays(48, 32)
    res.mvs[:] = (8, -4)
    assert digest(O.oracle_predict_bi_frame(a, a, abi.MODE_ONE_MV_QPEL, w, w)) == \
        digest(O.oracle_predict_p_frame(a, abi.MODE_ONE_MV_QPEL, res))


# ---------------------------------------------------------------- padded pictures

from _predict_cases import BI_PADDED_CASES, PADDED_CASES, bi_padded_inputs, padded_inputs  # noqa: E402

PAD_FIX = json.load(open(os.path.join(GOLDEN, "predict_padded.json")))
P_CASES = {c[0]: c for c in PADDED_CASES}
B_CASES = {c[0]: c for c in BI_PADDED_CASES}


@pytest.mark.parametrize("k", range(len(PAD_FIX)), ids=lambda k: PAD_FIX[k]["case"])
def test_oracle_padded_prediction_matches_fixture(k):
    f = PAD_FIX[k]
    if f["kind"] == "p":
        name, w, h, mode, padded = P_CASES[f["case"]]
        ref, res = padded_inputs(name, w, h, mode, padded)
        got = O.oracle_predict_p_frame_padded(ref, w, h, mode, res)
    else:
        name, w, h, mode, padded = B_CASES[f["case"]]
        past, fut, fm, bm = bi_padded_inputs(name, w, h, mode, padded)
        got = O.oracle_predict_bi_frame_padded(past, fut, w, h, mode, fm, bm)
    assert digest(got) == f["sha256"]


def test_padded_prediction_closed_forms():
    """On the coded raster the padded picture is the plain one (both clamp to
    the reference's raster); zero vectors copy a reference that has the
    padded raster and replicate the edges of one that has the coded raster."""
    rng = np.random.default_rng(12)
    w, h = 58, 38                                           # 4 x 3 macroblocks: padded 64 x 48
    ref = synth.Picture(rng.integers(0, 256, size=(h, w)).astype(np.uint8),
                        rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8),
                        rng.integers(0, 256, size=(h // 2, w // 2)).astype(np.uint8))
    res = abi.ResultArrays(w, h)
    res.mvs[:] = rng.integers(-90, 90, size=res.mvs.shape)
    res.block_mvs[:] = rng.integers(-90, 90, size=res.block_mvs.shape)
    res.use_4mv[::2] = 1
    py, pu, pv = O.oracle_predict_p_frame_padded(ref, w, h, abi.MODE_MIXED_MV, res)
    assert py.shape == (48, 64) and pu.shape == (24, 32)
    y, u, v = O.oracle_predict_p_frame(ref, abi.MODE_MIXED_MV, res)
    assert np.array_equal(py[:h, :w], y) and np.array_equal(pu[:h // 2, :w // 2], u)
    assert np.array_equal(pv[:h // 2, :w // 2], v)
    zero = abi.ResultArrays(w, h)
    zy, zu, _ = O.oracle_predict_p_frame_padded(ref, w, h, abi.MODE_ONE_MV_QPEL, zero)
    assert np.array_equal(zy, ref.y[np.ix_(np.clip(np.arange(48), 0, h - 1), np.clip(np.arange(64), 0, w - 1))])
    assert np.array_equal(zu, ref.u[np.ix_(np.clip(np.arange(24), 0, h // 2 - 1), np.clip(np.arange(32), 0, w // 2 - 1))])
    big = synth.Picture(zy, zu, zu.copy())                   # a reference at the padded raster
    again = O.oracle_predict_p_frame_padded(big, w, h, abi.MODE_ONE_MV_QPEL, zero)
    assert np.array_equal(again[0], zy) and np.array_equal(again[1], zu)
    fm = rng.integers(-90, 90, size=(12, 2)).astype(np.int32)
    bm = rng.integers(-90, 90, size=(12, 2)).astype(np.int32)
    other = synth.Picture(*(rng.integers(0, 256, size=x.shape).astype(np.uint8) for x in (ref.y, ref.u, ref.v)))
    by, bu, _ = O.oracle_predict_bi_frame_padded(ref, other, w, h, abi.MODE_ONE_MV_QPEL, fm, bm)
    ty, tu, _ = O.oracle_predict_bi_frame(ref, other, abi.MODE_ONE_MV_QPEL, fm, bm)
    assert np.array_equal(by[:h, :w], ty) and np.array_equal(bu[:h // 2, :w // 2], tu)


@needs_ref
def test_oracle_padded_prediction_matches_reference_wild_fields():
    """Vectors far outside the picture, references at the coded and at the
    padded raster, every mode."""
    rng = np.random.default_rng(13)
    w, h = 170, 98
    mbw, mbh = abi.mb_grid(w, h)
    coded = synth.pan_rectangles(41, 1, width=w, height=h, frame_indices=[0])[0]
    padded = synth.pan_rectangles(42, 1, width=mbw * 16, height=mbh * 16, frame_indices=[0])[0]
    for mode in (0, 1, 2, 3):
        res = abi.ResultArrays(w, h)
        res.mvs[:] = rng.integers(-900, 900, size=res.mvs.shape)
        res.block_mvs[:] = rng.integers(-900, 900, size=res.block_mvs.shape)
        res.use_4mv[:] = rng.integers(0, 2, size=res.use_4mv.shape) if mode == 3 else 0
        p = abi.default_params(w, h, mode=mode)
        for ref in (coded, padded):
            want = O.reference_predict_p_frame_padded(p, ref, res, 1)
            assert digest(O.oracle_predict_p_frame_padded(ref, w, h, mode, res)) == digest(want), (mode, ref.width)
        if mode < 3:
            fm = rng.integers(-900, 900, size=(mbw * mbh, 2)).astype(np.int32)
            bm = rng.integers(-900, 900, size=(mbw * mbh, 2)).astype(np.int32)
            want = O.reference_predict_bi_frame_padded(p, padded, coded, fm, bm, 1)
            assert digest(O.oracle_predict_bi_frame_padded(padded, coded, w, h, mode, fm, bm)) == digest(want), mode
