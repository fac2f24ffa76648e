"""Test-side loaders of the CPU checkers under oracle/ (never used by the product)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from libvc1_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libvc1me_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvc1ref.so")
P = C.POINTER


class OracleOps(C.Structure):
    _fields_ = [("sad_px", C.c_uint64), ("interp_taps", C.c_uint64), ("satd_ops", C.c_uint64),
                ("sig_ops", C.c_uint64), ("frame_px", C.c_uint64)]


_oracle = None
_ref = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_analyze_p_frame.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.PResult), P(OracleOps)]
        lib.oracle_score_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(abi.Candidate),
                                           C.c_size_t, P(C.c_uint64)]
        lib.oracle_score_batch.restype = None
        lib.oracle_luma_mc_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_int, C.c_int]
        lib.oracle_mb_sad.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int]
        lib.oracle_mb_sad.restype = C.c_uint64
        lib.oracle_block_cost.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.oracle_block_cost.restype = C.c_uint64
        lib.oracle_frame_sad.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int]
        lib.oracle_frame_sad.restype = C.c_uint64
        lib.oracle_global_motion.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, P(C.c_int32),
                                             P(C.c_int32), P(C.c_uint64)]
        lib.oracle_global_motion.restype = None
        lib.oracle_mv_rate_bits.argtypes = [P(abi.Params), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.oracle_score_cmp.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        lib.oracle_hidden_padding.argtypes = [P(abi.Params)] + [C.c_int] * 8
        lib.oracle_distant_decision.argtypes = [P(abi.Params), C.c_uint64, C.c_uint64, C.c_int, C.c_int]
        lib.oracle_integer_search.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes)] + [C.c_int] * 7 + [P(C.c_int32), P(C.c_int32), P(C.c_uint64)]
        lib.oracle_integer_search.restype = None
        lib.oracle_mb_costs20.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(C.c_int32), P(C.c_int32),
                                          P(C.c_uint8), P(C.c_uint64)]
        lib.oracle_mb_costs20.restype = None
        lib.oracle_estimate_intensity_comp.argtypes = [P(abi.Planes), P(abi.Planes), P(C.c_int32), P(abi.Intensity)]
        lib.oracle_intensity_compensate.argtypes = [P(abi.Planes), P(abi.Intensity), C.c_void_p, C.c_void_p,
                                                    C.c_void_p]
        lib.oracle_intensity_compensate.restype = None
        lib.oracle_intensity_map_y.argtypes = [C.c_uint8, P(abi.Intensity)]
        lib.oracle_intensity_map_y.restype = C.c_uint8
        lib.oracle_intensity_map_uv.argtypes = [C.c_uint8, P(abi.Intensity)]
        lib.oracle_intensity_map_uv.restype = C.c_uint8
        lib.oracle_analyze_b_frame.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.Planes),
                                               P(abi.BParams), P(abi.BResult)]
        lib.oracle_scene_change_score.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_double,
                                                  P(C.c_double)]
        lib.oracle_has_good_motion_match.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, P(C.c_int),
                                                     C.c_void_p]
        lib.oracle_refine_qpel_block.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes)] + [C.c_int] * 7 + \
            [C.c_uint64, P(C.c_int32), P(C.c_int32), P(C.c_uint64)]
        lib.oracle_predict_p_frame.argtypes = [P(abi.Planes), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]
        lib.oracle_predict_p_frame.restype = None
        lib.oracle_predict_bi_frame.argtypes = [P(abi.Planes), P(abi.Planes), C.c_int] + [C.c_void_p] * 5
        lib.oracle_predict_bi_frame.restype = None
        lib.oracle_predict_p_frame_padded.argtypes = [P(abi.Planes), C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
        lib.oracle_predict_p_frame_padded.restype = None
        lib.oracle_predict_bi_frame_padded.argtypes = [P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int] + \
            [C.c_void_p] * 5
        lib.oracle_predict_bi_frame_padded.restype = None
        lib.oracle_analyze_p_frame_seeded.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.PResult),
                                                      P(abi.PResult), P(OracleOps)]
        _oracle = lib
    return _oracle


def reference_available() -> bool:
    return os.path.exists(REF_SO)


def reference() -> C.CDLL:
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        lib.ref_analyze_p_core.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.PResult), C.c_int]
        lib.ref_block_costs.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, P(abi.Candidate), C.c_size_t, P(C.c_uint64)]
        lib.ref_mb_sad.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int]
        lib.ref_mb_sad.restype = C.c_uint64
        lib.ref_frame_sad.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_int, C.c_int]
        lib.ref_frame_sad.restype = C.c_uint64
        lib.ref_global_motion.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, P(C.c_int32),
                                          P(C.c_int32), P(C.c_uint64)]
        lib.ref_luma_mc_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int]
        lib.ref_time_p_frames.argtypes = [P(abi.Params), C.c_int, P(abi.Planes), P(abi.Planes), C.c_int, C.c_int]
        lib.ref_time_p_frames.restype = C.c_double
        lib.ref_last_error.restype = C.c_char_p
        lib.ref_estimate_intensity_comp.argtypes = [P(abi.Planes), P(abi.Planes), P(C.c_int32), P(abi.Intensity)]
        lib.ref_intensity_compensate.argtypes = [P(abi.Planes), P(abi.Intensity), C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_int]
        lib.ref_analyze_b.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.Planes), P(abi.BParams),
                                      P(abi.BResult)]
        lib.ref_scene_change_score.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, C.c_double,
                                               P(C.c_double)]
        lib.ref_has_good_motion_match.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), C.c_int, P(C.c_int)]
        lib.ref_refine_qpel_block.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes)] + [C.c_int] * 7 + \
            [C.c_uint64, P(C.c_int32), P(C.c_int32), P(C.c_uint64)]
        lib.ref_predict_p_frame.argtypes = [P(abi.Params), P(abi.Planes)] + [C.c_void_p] * 6 + [C.c_int]
        lib.ref_predict_bi_frame.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes)] + [C.c_void_p] * 5 + [C.c_int]
        lib.ref_predict_p_frame_padded.argtypes = [P(abi.Params), P(abi.Planes)] + [C.c_void_p] * 6 + [C.c_int]
        lib.ref_predict_bi_frame_padded.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes)] + \
            [C.c_void_p] * 5 + [C.c_int]
        lib.ref_analyze_p_core_seeded.argtypes = [P(abi.Params), P(abi.Planes), P(abi.Planes), P(abi.PResult),
                                                  P(abi.PResult), C.c_int]
        _ref = lib
    return _ref


def oracle_analyze(params: abi.Params, cur, ref, ops: OracleOps | None = None) -> abi.ResultArrays:
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = abi.ResultArrays(params.width, params.height)
    rc = oracle().oracle_analyze_p_frame(C.byref(params), C.byref(c.struct), C.byref(r.struct), C.byref(out.struct),
                                         C.byref(ops) if ops is not None else None)
    assert rc == 0
    return out


def reference_analyze(params: abi.Params, cur, ref) -> abi.ResultArrays:
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = abi.ResultArrays(params.width, params.height)
    rc = reference().ref_analyze_p_core(C.byref(params), C.byref(c.struct), C.byref(r.struct), C.byref(out.struct), 0)
    assert rc == 0, reference().ref_last_error()
    return out


def oracle_analyze_seeded(params: abi.Params, cur, ref, seed: abi.ResultArrays,
                          ops: OracleOps | None = None) -> abi.ResultArrays:
    """The seeded core (analyze_p_picture_core with a shared seed): `seed`'s
    1-MV vectors start the staged refinement of each macroblock."""
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = abi.ResultArrays(params.width, params.height)
    rc = oracle().oracle_analyze_p_frame_seeded(C.byref(params), C.byref(c.struct), C.byref(r.struct),
                                                C.byref(seed.struct), C.byref(out.struct),
                                                C.byref(ops) if ops is not None else None)
    assert rc == 0
    return out


def reference_analyze_seeded(params: abi.Params, cur, ref, seed: abi.ResultArrays) -> abi.ResultArrays:
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = abi.ResultArrays(params.width, params.height)
    rc = reference().ref_analyze_p_core_seeded(C.byref(params), C.byref(c.struct), C.byref(r.struct),
                                               C.byref(seed.struct), C.byref(out.struct), 0)
    assert rc == 0, reference().ref_last_error()
    return out


def _b(fn, params, cur, past, future, binp, *err):
    c, p, f = abi.PlaneSet.of(cur), abi.PlaneSet.of(past), abi.PlaneSet.of(future)
    out = abi.BResultArrays(params.width, params.height)
    rc = fn(C.byref(params), C.byref(c.struct), C.byref(p.struct), C.byref(f.struct), C.byref(binp.struct),
            C.byref(out.struct))
    assert rc == 0, err[0]() if err else rc
    return out


def oracle_analyze_b(params: abi.Params, cur, past, future, binp: abi.BInputs) -> abi.BResultArrays:
    """analyze_b_picture_mode in the oracle (params.mode = the B mode)."""
    return _b(oracle().oracle_analyze_b_frame, params, cur, past, future, binp)


def reference_analyze_b(params: abi.Params, cur, past, future, binp: abi.BInputs) -> abi.BResultArrays:
    return _b(reference().ref_analyze_b, params, cur, past, future, binp, lambda: reference().ref_last_error())


def _scene(fn, params, cur, ref, fade, threshold, *err):
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = C.c_double()
    rc = fn(C.byref(params), C.byref(c.struct), C.byref(r.struct), int(fade), float(threshold), C.byref(out))
    assert rc == 0, err[0]() if err else rc
    return out.value


def oracle_scene_change_score(params: abi.Params, cur, ref, fade: bool = True, threshold: float = 28.0) -> float:
    """scene_change_score in the oracle under the ME settings `params`."""
    return _scene(oracle().oracle_scene_change_score, params, cur, ref, fade, threshold)


def reference_scene_change_score(params: abi.Params, cur, ref, fade: bool = True, threshold: float = 28.0) -> float:
    return _scene(reference().ref_scene_change_score, params, cur, ref, fade, threshold,
                  lambda: reference().ref_last_error())


def _good(fn, params, cur, ref, fade, *extra, err=None):
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    out = C.c_int()
    rc = fn(C.byref(params), C.byref(c.struct), C.byref(r.struct), int(fade), C.byref(out), *extra)
    assert rc == 0, err() if err else rc
    return bool(out.value)


def oracle_has_good_motion_match(params: abi.Params, cur, ref, fade: bool = True) -> bool:
    """has_good_motion_match in the oracle under the ME settings `params`."""
    return _good(oracle().oracle_has_good_motion_match, params, cur, ref, fade, None)


def oracle_motion_check_field(params: abi.Params, cur, ref, fade: bool = True):
    """(answer, every macroblock's refined vector) of the motion-failure scan
    run to the end."""
    mbw, mbh = abi.mb_grid(params.width, params.height)
    f = np.zeros((mbw * mbh, 2), np.int32)
    g = _good(oracle().oracle_has_good_motion_match, params, cur, ref, fade, f.ctypes.data)
    return g, f


def reference_has_good_motion_match(params: abi.Params, cur, ref, fade: bool = True) -> bool:
    return _good(reference().ref_has_good_motion_match, params, cur, ref, fade, err=lambda: reference().ref_last_error())


def _refine_qpel(fn, params, cur, ref, x0, y0, bw, bh, seed, rng, known):
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    x, y, s = C.c_int32(), C.c_int32(), C.c_uint64()
    assert fn(C.byref(params), C.byref(c.struct), C.byref(r.struct), x0, y0, bw, bh, seed[0], seed[1], rng,
              known, C.byref(x), C.byref(y), C.byref(s)) == 0
    return (x.value, y.value, s.value)


def oracle_refine_qpel_block(params, cur, ref, x0, y0, bw, bh, seed, rng, known=2**64 - 1):
    """refine_qpel_block: (xq, yq, sad)."""
    return _refine_qpel(oracle().oracle_refine_qpel_block, params, cur, ref, x0, y0, bw, bh, seed, rng, known)


def reference_refine_qpel_block(params, cur, ref, x0, y0, bw, bh, seed, rng, known=2**64 - 1):
    return _refine_qpel(reference().ref_refine_qpel_block, params, cur, ref, x0, y0, bw, bh, seed, rng, known)


def _pred_arrays(ref, result):
    r = abi.PlaneSet.of(ref)
    y, u, v = np.empty_like(r.y), np.empty_like(r.u), np.empty_like(r.v)
    mvs = np.ascontiguousarray(result.mvs, np.int32)
    bmv = np.ascontiguousarray(result.block_mvs, np.int32)
    four = np.ascontiguousarray(result.use_4mv, np.uint8)
    return r, y, u, v, mvs, bmv, four


def oracle_predict_p_frame(ref, mode: int, result: abi.ResultArrays):
    """The P prediction picture (y, u, v) of an analysis result."""
    r, y, u, v, mvs, bmv, four = _pred_arrays(ref, result)
    oracle().oracle_predict_p_frame(C.byref(r.struct), mode, mvs.ctypes.data, bmv.ctypes.data, four.ctypes.data,
                                    y.ctypes.data, u.ctypes.data, v.ctypes.data)
    return y, u, v


def reference_predict_p_frame(params: abi.Params, ref, result: abi.ResultArrays, simd_auto: int = 0):
    r, y, u, v, mvs, bmv, four = _pred_arrays(ref, result)
    rc = reference().ref_predict_p_frame(C.byref(params), C.byref(r.struct), mvs.ctypes.data, bmv.ctypes.data,
                                         four.ctypes.data, y.ctypes.data, u.ctypes.data, v.ctypes.data, simd_auto)
    assert rc == 0, reference().ref_last_error()
    return y, u, v


def _bi(fn, head, past, future, fmvs, bmvs, *tail):
    pa, fu = abi.PlaneSet.of(past), abi.PlaneSet.of(future)
    y, u, v = np.empty_like(pa.y), np.empty_like(pa.u), np.empty_like(pa.v)
    f = np.ascontiguousarray(fmvs, np.int32)
    b = np.ascontiguousarray(bmvs, np.int32)
    rc = fn(*head, C.byref(pa.struct), C.byref(fu.struct), *tail[:1], f.ctypes.data, b.ctypes.data, y.ctypes.data,
            u.ctypes.data, v.ctypes.data, *tail[1:])
    assert rc in (None, 0)
    return y, u, v


def oracle_predict_bi_frame(past, future, mode: int, fmvs, bmvs):
    """The bi-directional prediction picture (y, u, v)."""
    return _bi(oracle().oracle_predict_bi_frame, (), past, future, fmvs, bmvs, mode)


def reference_predict_bi_frame(params: abi.Params, past, future, fmvs, bmvs, simd_auto: int = 0):
    pa, fu = abi.PlaneSet.of(past), abi.PlaneSet.of(future)
    y, u, v = np.empty_like(pa.y), np.empty_like(pa.u), np.empty_like(pa.v)
    f = np.ascontiguousarray(fmvs, np.int32)
    b = np.ascontiguousarray(bmvs, np.int32)
    rc = reference().ref_predict_bi_frame(C.byref(params), C.byref(pa.struct), C.byref(fu.struct), f.ctypes.data,
                                          b.ctypes.data, y.ctypes.data, u.ctypes.data, v.ctypes.data, simd_auto)
    assert rc == 0, reference().ref_last_error()
    return y, u, v


def _padded_planes(width: int, height: int):
    """Empty y, u, v planes of the macroblock raster of a width x height
    picture (mbw*16 x mbh*16)."""
    mbw, mbh = abi.mb_grid(width, height)
    return (np.empty((mbh * 16, mbw * 16), np.uint8), np.empty((mbh * 8, mbw * 8), np.uint8),
            np.empty((mbh * 8, mbw * 8), np.uint8))


def oracle_predict_p_frame_padded(ref, width: int, height: int, mode: int, result: abi.ResultArrays):
    """The padded P prediction picture (y, u, v) of a width x height
    analysis, the reference at its own raster."""
    r = abi.PlaneSet.of(ref)
    y, u, v = _padded_planes(width, height)
    mbw, mbh = abi.mb_grid(width, height)
    mvs = np.ascontiguousarray(result.mvs, np.int32)
    bmv = np.ascontiguousarray(result.block_mvs, np.int32)
    four = np.ascontiguousarray(result.use_4mv, np.uint8)
    oracle().oracle_predict_p_frame_padded(C.byref(r.struct), mbw, mbh, mode, mvs.ctypes.data, bmv.ctypes.data,
                                           four.ctypes.data, y.ctypes.data, u.ctypes.data, v.ctypes.data)
    return y, u, v


def reference_predict_p_frame_padded(params: abi.Params, ref, result: abi.ResultArrays, simd_auto: int = 0):
    r = abi.PlaneSet.of(ref)
    y, u, v = _padded_planes(params.width, params.height)
    mvs = np.ascontiguousarray(result.mvs, np.int32)
    bmv = np.ascontiguousarray(result.block_mvs, np.int32)
    four = np.ascontiguousarray(result.use_4mv, np.uint8)
    rc = reference().ref_predict_p_frame_padded(C.byref(params), C.byref(r.struct), mvs.ctypes.data, bmv.ctypes.data,
                                                four.ctypes.data, y.ctypes.data, u.ctypes.data, v.ctypes.data,
                                                simd_auto)
    assert rc == 0, reference().ref_last_error()
    return y, u, v


def oracle_predict_bi_frame_padded(past, future, width: int, height: int, mode: int, fmvs, bmvs):
    """The padded bi-directional prediction picture (y, u, v)."""
    pa, fu = abi.PlaneSet.of(past), abi.PlaneSet.of(future)
    y, u, v = _padded_planes(width, height)
    mbw, mbh = abi.mb_grid(width, height)
    f = np.ascontiguousarray(fmvs, np.int32)
    b = np.ascontiguousarray(bmvs, np.int32)
    oracle().oracle_predict_bi_frame_padded(C.byref(pa.struct), C.byref(fu.struct), mbw, mbh, mode, f.ctypes.data,
                                            b.ctypes.data, y.ctypes.data, u.ctypes.data, v.ctypes.data)
    return y, u, v


def reference_predict_bi_frame_padded(params: abi.Params, past, future, fmvs, bmvs, simd_auto: int = 0):
    pa, fu = abi.PlaneSet.of(past), abi.PlaneSet.of(future)
    y, u, v = _padded_planes(params.width, params.height)
    f = np.ascontiguousarray(fmvs, np.int32)
    b = np.ascontiguousarray(bmvs, np.int32)
    rc = reference().ref_predict_bi_frame_padded(C.byref(params), C.byref(pa.struct), C.byref(fu.struct),
                                                 f.ctypes.data, b.ctypes.data, y.ctypes.data, u.ctypes.data,
                                                 v.ctypes.data, simd_auto)
    assert rc == 0, reference().ref_last_error()
    return y, u, v


def _mv_pairs(mvs: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(mvs, np.int32).reshape(-1, 2))


def oracle_estimate_intensity(cur, ref, mvs: np.ndarray) -> tuple:
    """estimate_intensity_comp: (enabled, lumscale, lumshift)."""
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    m = _mv_pairs(mvs)
    out = abi.Intensity()
    assert oracle().oracle_estimate_intensity_comp(C.byref(c.struct), C.byref(r.struct),
                                                   m.ctypes.data_as(P(C.c_int32)), C.byref(out)) == 0
    return out.as_tuple()


def reference_estimate_intensity(cur, ref, mvs: np.ndarray) -> tuple:
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    m = _mv_pairs(mvs)
    out = abi.Intensity()
    rc = reference().ref_estimate_intensity_comp(C.byref(c.struct), C.byref(r.struct),
                                                 m.ctypes.data_as(P(C.c_int32)), C.byref(out))
    assert rc == 0, reference().ref_last_error()
    return out.as_tuple()


def _compensate(fn, ref, ic: tuple, *extra):
    r = abi.PlaneSet.of(ref)
    y, u, v = np.empty_like(r.y), np.empty_like(r.u), np.empty_like(r.v)
    rc = fn(C.byref(r.struct), C.byref(abi.Intensity(*ic)), y.ctypes.data, u.ctypes.data, v.ctypes.data, *extra)
    assert rc in (None, 0)
    return y, u, v


def oracle_intensity_compensate(ref, ic: tuple):
    """intensity_compensated_reference: the (y, u, v) planes."""
    return _compensate(oracle().oracle_intensity_compensate, ref, ic)


def reference_intensity_compensate(ref, ic: tuple, simd_auto: int = 0):
    return _compensate(reference().ref_intensity_compensate, ref, ic, simd_auto)


def oracle_score_batch(cur: np.ndarray, ref: np.ndarray, x0, y0, bw, bh, cands, stride=1, bilinear=0, rnd=0,
                       metric=0) -> np.ndarray:
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    h, w = cur.shape
    cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
    out = np.zeros(len(cands), np.uint64)
    oracle().oracle_score_batch(cur.ctypes.data, ref.ctypes.data, w, h, x0, y0, bw, bh, stride, bilinear, rnd, metric,
                                cands.ctypes.data_as(P(abi.Candidate)), len(cands), out.ctypes.data_as(P(C.c_uint64)))
    return out


def oracle_mb_costs20(params: abi.Params, cur, ref, one_mvs: np.ndarray, result: abi.ResultArrays) -> np.ndarray:
    """20 * (SAD + lambda * R) per macroblock at the given 1-MV vectors, with
    the result's block vectors and intra flags as the predictor field."""
    c, r = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    one = np.ascontiguousarray(one_mvs, np.int32)
    bmv = np.ascontiguousarray(result.block_mvs, np.int32)
    intra = np.ascontiguousarray(result.use_intra, np.uint8)
    out = np.zeros(result.mbs, np.uint64)
    oracle().oracle_mb_costs20(C.byref(params), C.byref(c.struct), C.byref(r.struct), one.ctypes.data_as(P(C.c_int32)),
                               bmv.ctypes.data_as(P(C.c_int32)), intra.ctypes.data_as(P(C.c_uint8)),
                               out.ctypes.data_as(P(C.c_uint64)))
    return out


def describe_diff(a: abi.ResultArrays, b: abi.ResultArrays, width: int) -> str:
    bad = a.diff(b)
    lines = [f"fields differ: {bad}"]
    mbw = (width + 15) // 16
    for k in bad[:4]:
        x, y = getattr(a, k, None), getattr(b, k, None)
        if isinstance(x, np.ndarray):
            xx = x.reshape(x.shape[0], -1)
            yy = y.reshape(y.shape[0], -1)
            if x.dtype == np.float64:
                idx = np.nonzero((xx.view(np.uint64) != yy.view(np.uint64)).any(1))[0]
            else:
                idx = np.nonzero((xx != yy).any(1))[0]
            for i in idx[:3]:
                lines.append(f"  {k}[mb {i} = ({i % mbw},{i // mbw})]: {xx[i].tolist()} vs {yy[i].tolist()}")
            lines.append(f"  {k}: {len(idx)} macroblocks differ")
        else:
            lines.append(f"  {k}: {a.scalars()[k]} vs {b.scalars()[k]}")
    return "\n".join(lines)
