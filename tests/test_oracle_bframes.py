"""The oracle's B-picture analysis (analyze_b_picture_mode,
src/encoder_motion.cpp:753-1004) against fixtures written by the compiled
reference and, where it is built, against the reference itself."""
import glob
import json
import os

import numpy as np
import pytest

import _oracle as O
from libvc1_amd import abi, synth
from _oracle import oracle_analyze_b

BDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bframes")
B_FILES = sorted(glob.glob(os.path.join(BDIR, "*.npz")))
needs_ref = pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref not built (needs the reference sources)")


def load_b(path):
    """(cur, past, future, params, B inputs, expected result) of a fixture."""
    z = np.load(path, allow_pickle=False)
    pic = lambda k: synth.Picture(z[f"{k}_y"], z[f"{k}_u"], z[f"{k}_v"])
    cur, past, fut = pic("cur"), pic("past"), pic("fut")
    h, w = cur.y.shape
    p = abi.default_params(w, h, **json.loads(str(z["params"])))
    bkw = json.loads(str(z["bparams"]))
    bkw = {k: tuple(v) if isinstance(v, list) else v for k, v in bkw.items()}
    binp = abi.BInputs(z["anchor_mvs"], z["anchor_4mv"], **bkw)
    want = abi.BResultArrays(w, h)
    for k in abi.BResultArrays.ARRAYS:
        getattr(want, k)[...] = z[f"out_{k}"]
    for k, v in json.loads(str(z["scalars"])).items():
        setattr(want.struct, k, v)
    return cur, past, fut, p, binp, want


def describe_b(got, want):
    bad = got.diff(want)
    lines = [f"fields differ: {bad}"]
    for k in bad:
        if k in abi.BResultArrays.SCALARS:
            lines.append(f"  {k}: got {getattr(got.struct, k)} want {getattr(want.struct, k)}")
        else:
            a, b = getattr(got, k), getattr(want, k)
            idx = np.argwhere(a.reshape(a.shape[0], -1) != b.reshape(b.shape[0], -1))[:5, 0]
            lines.append(f"  {k}: first MBs {idx.tolist()} got {a[idx].tolist()} want {b[idx].tolist()}")
    return "\n".join(lines)


def test_b_fixtures_present():
    assert len(B_FILES) >= 6


@pytest.mark.parametrize("path", B_FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_b_matches_fixture(path):
    cur, past, fut, p, binp, want = load_b(path)
    got = oracle_analyze_b(p, cur, past, fut, binp)
    assert not got.diff(want), describe_b(got, want)


def test_b_fixtures_cover_every_mode():
    seen = set()
    for path in B_FILES:
        seen |= set(np.unique(np.load(path)["out_modes"]).tolist())
    assert seen == {abi.B_FORWARD, abi.B_BACKWARD, abi.B_INTERPOLATED, abi.B_DIRECT, abi.B_INTRA}


def test_b_zero_motion_shortcut_and_errors():
    d = synth.pan_rectangles(50, 3, width=64, height=48, frame_indices=[0, 1, 2])
    p = abi.default_params(64, 48, motion_search_range=0)
    binp = abi.BInputs(np.zeros((12, 2), np.int32), np.zeros(12, np.uint8), zero_motion_shortcut=1)
    r = oracle_analyze_b(p, d[1], d[0], d[2], binp)
    assert (r.modes == abi.B_FORWARD).all() and not r.forward_mvs.any() and not r.backward_mvs.any()
    assert r.struct.forward_macroblocks == 12 and r.struct.motion_bits == 0
    assert (r.inter_intra_cost_ratio == -1.0).all() and (r.forward_distant_local_mae == -1.0).all()
    c, pa, f = (abi.PlaneSet.of(x) for x in (d[1], d[0], d[2]))
    out = abi.BResultArrays(64, 48)
    for frac, mode in (((2, 2), abi.MODE_ONE_MV_QPEL), ((1, 4), abi.MODE_ONE_MV_QPEL), ((1, 2), abi.MODE_MIXED_MV)):
        q = abi.default_params(64, 48, mode=mode)
        bi = abi.BInputs(np.zeros((12, 2), np.int32), np.zeros(12, np.uint8), fraction=frac)
        import ctypes as C
        assert O.oracle().oracle_analyze_b_frame(C.byref(q), C.byref(c.struct), C.byref(pa.struct), C.byref(f.struct),
                                                 C.byref(bi.struct), C.byref(out.struct)) == -1


@needs_ref
@pytest.mark.parametrize("case", ["pan_qpel_third", "pan_bilinear_half_ic", "pan_hpel_twothirds_nointra",
                                  "shortcut", "wild_anchor"])
def test_oracle_b_matches_reference(case):
    d = synth.pan_rectangles(60, 7, width=240, height=136, frame_indices=[0, 1, 2, 3, 6])
    mbw, mbh = abi.mb_grid(240, 136)
    base = dict(pqindex=10, motion_search_range=256, motion_local_search_range=32)
    anchor = O.oracle_analyze(abi.default_params(240, 136, mode=abi.MODE_MIXED_MV, **base), d[3], d[0])
    mode, frac, kw, cur = abi.MODE_ONE_MV_QPEL, (1, 3), {}, d[1]
    amv, a4 = anchor.mvs, anchor.use_4mv
    if case == "pan_bilinear_half_ic":
        mode, frac, kw, cur = abi.MODE_ONE_MV_HPEL_BILINEAR, (1, 2), dict(intensity=(1, 34, 62)), d[2]
    elif case == "pan_hpel_twothirds_nointra":
        mode, frac, kw, cur = abi.MODE_ONE_MV_HPEL, (2, 3), dict(allow_intra=0), d[2]
    elif case == "shortcut":
        base["motion_search_range"] = 0
        kw = dict(zero_motion_shortcut=1)
    elif case == "wild_anchor":
        rng = np.random.default_rng(61)
        amv = rng.integers(-3000, 3000, size=(mbw * mbh, 2)).astype(np.int32)
        a4 = rng.integers(0, 2, size=mbw * mbh).astype(np.uint8)
    p = abi.default_params(240, 136, mode=mode, **base)
    binp = abi.BInputs(amv, a4, fraction=frac, **kw)
    want = O.reference_analyze_b(p, cur, d[0], d[3], binp)
    got = oracle_analyze_b(p, cur, d[0], d[3], binp)
    assert not got.diff(want), describe_b(got, want)
