"""Inputs of the P prediction-picture cases (tests and fixture writer)."""
import numpy as np

from libvc1_amd import abi, synth

PREDICT_CASES = [("pan_qpel", 240, 136, abi.MODE_ONE_MV_QPEL), ("pan_hpel", 240, 136, abi.MODE_ONE_MV_HPEL),
                 ("pan_bilinear", 240, 136, abi.MODE_ONE_MV_HPEL_BILINEAR), ("pan_mixed", 240, 136, abi.MODE_MIXED_MV),
                 ("odd_mixed", 202, 118, abi.MODE_MIXED_MV)]


def predict_inputs(name, w, h, mode, analyze):
    """(ref, params, analysis result) of a case; `analyze(params, cur, ref)`
    supplies the vector field (the reference's or the oracle's analysis)."""
    d = synth.pan_rectangles(sum(map(ord, name)) % 1000, 4, width=w, height=h, frame_indices=[0, 3])
    p = abi.default_params(w, h, pqindex=8, motion_search_range=256, mode=mode)
    res = analyze(p, d[3], d[0])
    if mode == abi.MODE_MIXED_MV:
        rng = np.random.default_rng(len(name))
        res.block_mvs[::3] = rng.integers(-300, 300, size=res.block_mvs[::3].shape)
        res.use_4mv[::3] = 1
    return d[0], p, res


def result_from_fixture(f):
    """An analysis result carrying a fixture's vector field."""
    res = abi.ResultArrays(f["width"], f["height"])
    res.mvs[...] = np.asarray(f["mvs"], np.int32).reshape(res.mvs.shape)
    res.block_mvs[...] = np.asarray(f["block_mvs"], np.int32).reshape(res.block_mvs.shape)
    res.use_4mv[...] = np.asarray(f["use_4mv"], np.uint8)
    return res


BI_CASES = [("bi_qpel", 240, 136, abi.MODE_ONE_MV_QPEL), ("bi_hpel", 240, 136, abi.MODE_ONE_MV_HPEL),
            ("bi_bilinear_odd", 202, 118, abi.MODE_ONE_MV_HPEL_BILINEAR)]


def bi_inputs(name, w, h, mode):
    """(past, future, forward field, backward field) of a bi-directional case:
    odd quarter-pel components included, so both chroma roundings occur."""
    d = synth.pan_rectangles(sum(map(ord, name)) % 1000, 4, width=w, height=h, frame_indices=[0, 3])
    mbw, mbh = abi.mb_grid(w, h)
    rng = np.random.default_rng(len(name) + mode)
    f = rng.integers(-200, 200, size=(mbw * mbh, 2)).astype(np.int32)
    b = rng.integers(-200, 200, size=(mbw * mbh, 2)).astype(np.int32)
    if mode != abi.MODE_ONE_MV_QPEL:
        f &= ~1
        b &= ~1
    return d[0], d[3], f, b


# padded (loop-reconstruction) pictures: (name, width, height, mode, rasters
# of the references: True = the padded macroblock raster, False = the coded one)
PADDED_CASES = [("pad_qpel_coded", 202, 118, abi.MODE_ONE_MV_QPEL, False),
                ("pad_hpel_padded", 240, 136, abi.MODE_ONE_MV_HPEL, True),
                ("pad_bilinear_padded", 202, 118, abi.MODE_ONE_MV_HPEL_BILINEAR, True),
                ("pad_mixed_coded", 240, 136, abi.MODE_MIXED_MV, False),
                ("pad_mixed_padded", 202, 118, abi.MODE_MIXED_MV, True)]
BI_PADDED_CASES = [("bipad_qpel_padded", 202, 118, abi.MODE_ONE_MV_QPEL, (True, True)),
                   ("bipad_hpel_coded", 240, 136, abi.MODE_ONE_MV_HPEL, (False, False)),
                   ("bipad_bilinear_mixed_rasters", 202, 118, abi.MODE_ONE_MV_HPEL_BILINEAR, (True, False))]


def _ref_picture(seed, w, h, padded):
    mbw, mbh = abi.mb_grid(w, h)
    rw, rh = (mbw * 16, mbh * 16) if padded else (w, h)
    return synth.pan_rectangles(seed, 1, width=rw, height=rh, frame_indices=[0])[0]


def padded_inputs(name, w, h, mode, padded):
    """(ref, analysis result with a random field) of a padded P case: vectors
    up to +-75 px (even for the half-pel modes), a third of the macroblocks
    4-MV in Mixed-MV mode."""
    seed = sum(map(ord, name)) % 1000
    ref = _ref_picture(seed, w, h, padded)
    rng = np.random.default_rng(seed)
    res = abi.ResultArrays(w, h)
    res.mvs[...] = rng.integers(-300, 301, size=res.mvs.shape)
    res.block_mvs[...] = rng.integers(-300, 301, size=res.block_mvs.shape)
    if mode in (abi.MODE_ONE_MV_HPEL, abi.MODE_ONE_MV_HPEL_BILINEAR):
        res.mvs &= ~1
    if mode == abi.MODE_MIXED_MV:
        res.use_4mv[::3] = 1
    return ref, res


def bi_padded_inputs(name, w, h, mode, padded):
    """(past, future, forward field, backward field) of a padded bi case."""
    seed = sum(map(ord, name)) % 1000
    past = _ref_picture(seed, w, h, padded[0])
    fut = _ref_picture(seed + 1, w, h, padded[1])
    mbw, mbh = abi.mb_grid(w, h)
    rng = np.random.default_rng(seed + mode)
    f = rng.integers(-200, 200, size=(mbw * mbh, 2)).astype(np.int32)
    b = rng.integers(-200, 200, size=(mbw * mbh, 2)).astype(np.int32)
    if mode != abi.MODE_ONE_MV_QPEL:
        f &= ~1
        b &= ~1
    return past, fut, f, b
