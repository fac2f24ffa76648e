"""The P prediction picture on the GPU (vc1me_predict_p_frame,
src/encoder_motion_compensation.cpp:474-571) through the C ABI, bit-exact
against the reference-written digests and the oracle."""
import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi, synth
from _predict_cases import bi_inputs, result_from_fixture
from test_oracle_predict import BI_FIX, FIX, digest, fixture_ref

pytestmark = pytest.mark.gpu


def _predict(backend, ref, p, res):
    fr = backend.upload(ref)
    try:
        g = backend.predict_p_frame(fr, p, res)
        try:
            return backend.download(g)
        finally:
            g.release()
    finally:
        fr.release()


@pytest.mark.parametrize("k", range(len(FIX)), ids=lambda k: FIX[k]["case"])
def test_prediction_matches_fixture(backend, k):
    f = FIX[k]
    print(f"[case] prediction {f['case']}", flush=True)
    p = abi.default_params(f["width"], f["height"], mode=f["mode"])
    assert digest(_predict(backend, fixture_ref(f), p, result_from_fixture(f))) == f["sha256"]


def test_prediction_of_gpu_analysis_matches_oracle(backend):
    """The prediction of the GPU's own 1080p Mixed-MV analysis, and wild
    vectors far outside the picture, against the oracle."""
    d = synth.pan_rectangles(33, 4, width=1920, height=1080, frame_indices=[0, 3])
    p = abi.default_params(1920, 1080, pqindex=8, motion_search_range=512, mode=abi.MODE_MIXED_MV)
    fc, fr = backend.upload(d[3]), backend.upload(d[0])
    try:
        res = backend.analyze(fc, fr, p)
    finally:
        fc.release()
        fr.release()
    print(f"[case] prediction 1080p, {int(res.use_4mv.sum())} 4-MV macroblocks", flush=True)
    got = _predict(backend, d[0], p, res)
    assert digest(got) == digest(O.oracle_predict_p_frame(d[0], p.mode, res))
    rng = np.random.default_rng(10)
    res.mvs[:] = rng.integers(-9000, 9000, size=res.mvs.shape)
    res.block_mvs[:] = rng.integers(-9000, 9000, size=res.block_mvs.shape)
    print("[case] prediction 1080p, wild vectors", flush=True)
    assert digest(_predict(backend, d[0], p, res)) == digest(O.oracle_predict_p_frame(d[0], p.mode, res))


@pytest.mark.parametrize("k", range(len(BI_FIX)), ids=lambda k: BI_FIX[k]["case"])
def test_bi_prediction_matches_fixture_and_oracle(backend, k):
    f = BI_FIX[k]
    print(f"[case] bi prediction {f['case']}", flush=True)
    past, fut, fmv, bmv = bi_inputs(f["case"], f["width"], f["height"], f["mode"])
    fp, ff = backend.upload(past), backend.upload(fut)
    try:
        g = backend.predict_bi_frame(fp, ff, abi.default_params(f["width"], f["height"], mode=f["mode"]), fmv, bmv)
        try:
            got = backend.download(g)
        finally:
            g.release()
    finally:
        fp.release()
        ff.release()
    assert digest(got) == f["sha256"]
    assert digest(got) == digest(O.oracle_predict_bi_frame(past, fut, f["mode"], fmv, bmv))


# ---------------------------------------------------------------- padded pictures

from _predict_cases import BI_PADDED_CASES, PADDED_CASES, bi_padded_inputs, padded_inputs  # noqa: E402
from test_oracle_predict import PAD_FIX  # noqa: E402

P_CASES = {c[0]: c for c in PADDED_CASES}
B_CASES = {c[0]: c for c in BI_PADDED_CASES}


def _fetch(backend, g):
    try:
        return backend.download(g)
    finally:
        g.release()


@pytest.mark.parametrize("k", range(len(PAD_FIX)), ids=lambda k: PAD_FIX[k]["case"])
def test_padded_prediction_matches_fixture_and_oracle(backend, k):
    f = PAD_FIX[k]
    print(f"[case] padded prediction {f['case']}", flush=True)
    w, h, mode = f["width"], f["height"], f["mode"]
    p = abi.default_params(w, h, mode=mode)
    if f["kind"] == "p":
        ref, res = padded_inputs(*P_CASES[f["case"]])
        fr = backend.upload(ref)
        try:
            got = _fetch(backend, backend.predict_p_frame_padded(fr, p, res))
        finally:
            fr.release()
        want = O.oracle_predict_p_frame_padded(ref, w, h, mode, res)
    else:
        past, fut, fm, bm = bi_padded_inputs(*B_CASES[f["case"]])
        fp, ff = backend.upload(past), backend.upload(fut)
        try:
            got = _fetch(backend, backend.predict_bi_frame_padded(fp, ff, p, fm, bm))
        finally:
            fp.release()
            ff.release()
        want = O.oracle_predict_bi_frame_padded(past, fut, w, h, mode, fm, bm)
    mbw, mbh = abi.mb_grid(w, h)
    assert got[0].shape == (mbh * 16, mbw * 16)
    assert digest(got) == f["sha256"]
    assert digest(got) == digest(want)


def test_padded_prediction_1080p(backend):
    """The loop-reconstruction picture of a 1080p analysis (1088 coded rows
    in the padded raster) from the coded reference and from a padded one,
    and the padded bi-directional picture, against the oracle."""
    d = synth.pan_rectangles(34, 4, width=1920, height=1080, frame_indices=[0, 3])
    p = abi.default_params(1920, 1080, pqindex=8, motion_search_range=512, mode=abi.MODE_MIXED_MV)
    fc, fr = backend.upload(d[3]), backend.upload(d[0])
    try:
        res = backend.analyze(fc, fr, p)
        print(f"[case] padded prediction 1080p, {int(res.use_4mv.sum())} 4-MV macroblocks", flush=True)
        got = _fetch(backend, backend.predict_p_frame_padded(fr, p, res))
        assert got[0].shape == (1088, 1920)
        assert digest(got) == digest(O.oracle_predict_p_frame_padded(d[0], 1920, 1080, p.mode, res))
        padded_ref = synth.Picture(*got)
        fpad = backend.upload(padded_ref)
        try:
            print("[case] padded prediction 1080p from a padded reference", flush=True)
            again = _fetch(backend, backend.predict_p_frame_padded(fpad, p, res))
            assert digest(again) == digest(O.oracle_predict_p_frame_padded(padded_ref, 1920, 1080, p.mode, res))
            fm = np.ascontiguousarray(res.mvs.reshape(-1, 2))
            bm = np.ascontiguousarray(-res.mvs.reshape(-1, 2))
            q = abi.default_params(1920, 1080, mode=abi.MODE_ONE_MV_QPEL)
            print("[case] padded bi prediction 1080p, padded past and coded future", flush=True)
            bi = _fetch(backend, backend.predict_bi_frame_padded(fpad, fc, q, fm, bm))
            assert digest(bi) == digest(O.oracle_predict_bi_frame_padded(padded_ref, d[3], 1920, 1080, q.mode, fm, bm))
        finally:
            fpad.release()
    finally:
        fc.release()
        fr.release()
