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
