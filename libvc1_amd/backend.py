"""Python binding of the C ABI in include/vc1me.h (libvc1_amd/libvc1me.so).

This is plumbing for tests and benchmarks: it loads the in-tree HIP library
and forwards every call through ctypes.  There is no fallback: if the library
or a HIP device is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvc1me.so")

# every symbol declared in include/vc1me.h
EXPORTS = (
    "vc1me_params_default", "vc1me_create", "vc1me_destroy", "vc1me_last_error", "vc1me_abi_version",
    "vc1me_device_name", "vc1me_score_batch", "vc1me_frame_upload", "vc1me_frame_release",
    "vc1me_score_batch_frames", "vc1me_analyze_p_frame", "vc1me_analyze_step", "vc1me_last_timing",
    "vc1me_set_profiling", "vc1me_last_work", "vc1me_counters_get", "vc1me_counters_reset",
    "vc1me_block_costs", "vc1me_nn_param_count", "vc1me_nn_create", "vc1me_nn_destroy", "vc1me_nn_set_params",
    "vc1me_nn_get_params", "vc1me_nn_forward_frames", "vc1me_nn_forward_backward", "vc1me_nn_train_step",
    "vc1me_analyze_step_learned", "vc1me_analyze_step_proposals", "vc1me_set_cost_report", "vc1me_last_mb_costs",
    "vc1me_set_graphs", "vc1me_set_schedule", "vc1me_last_relax_passes", "vc1me_analyze_p_frame_seeded",
    "vc1me_analyze_step_seeded", "vc1me_estimate_intensity_comp", "vc1me_frame_intensity_compensated",
    "vc1me_frame_download", "vc1me_analyze_b_frame", "vc1me_analyze_b_step", "vc1me_scene_change_score",
    "vc1me_has_good_motion_match", "vc1me_predict_p_frame", "vc1me_predict_bi_frame",
    "vc1me_set_learned_candidates", "vc1me_score_batch_frames_noalloc", "vc1me_predict_p_frame_padded",
    "vc1me_predict_bi_frame_padded", "vc1me_score_batch_gen",
)


class BackendError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vc1me error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libvc1me.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: run `python -m libvc1_amd.build` first")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.vc1me_params_default.argtypes = [P(abi.Params), C.c_int32, C.c_int32]
    lib.vc1me_params_default.restype = None
    lib.vc1me_create.argtypes = [C.c_int, P(C.c_void_p)]
    lib.vc1me_destroy.argtypes = [C.c_void_p]
    lib.vc1me_destroy.restype = None
    lib.vc1me_last_error.restype = C.c_char_p
    lib.vc1me_device_name.argtypes = [C.c_void_p]
    lib.vc1me_device_name.restype = C.c_char_p
    lib.vc1me_score_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      P(abi.Candidate), C.c_size_t, P(C.c_uint64)]
    lib.vc1me_score_batch_gen.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, P(abi.Candidate), C.c_size_t, P(C.c_uint64)]
    lib.vc1me_frame_upload.argtypes = [C.c_void_p, P(abi.Planes), P(C.c_void_p)]
    lib.vc1me_frame_release.argtypes = [C.c_void_p, C.c_void_p]
    lib.vc1me_frame_release.restype = None
    lib.vc1me_score_batch_frames.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             P(abi.Candidate), C.c_size_t, P(C.c_uint64)]
    lib.vc1me_score_batch_frames_noalloc.argtypes = lib.vc1me_score_batch_frames.argtypes
    lib.vc1me_analyze_p_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), P(abi.PResult)]
    lib.vc1me_analyze_step.argtypes = [C.c_void_p, P(abi.Job), C.c_int]
    lib.vc1me_analyze_p_frame_seeded.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), P(abi.PResult),
                                                 P(abi.PResult)]
    lib.vc1me_estimate_intensity_comp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(C.c_int32), P(abi.Intensity)]
    lib.vc1me_frame_intensity_compensated.argtypes = [C.c_void_p, C.c_void_p, P(abi.Intensity), P(C.c_void_p)]
    lib.vc1me_analyze_b_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params),
                                          P(abi.BParams), P(abi.BResult)]
    lib.vc1me_analyze_b_step.argtypes = [C.c_void_p, P(abi.BJob), C.c_int]
    lib.vc1me_scene_change_score.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), C.c_int, C.c_double,
                                             P(C.c_double)]
    lib.vc1me_has_good_motion_match.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), C.c_int,
                                                P(C.c_int), C.c_void_p]
    lib.vc1me_predict_p_frame.argtypes = [C.c_void_p, C.c_void_p, P(abi.Params), P(abi.PResult), P(C.c_void_p)]
    lib.vc1me_predict_bi_frame.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), C.c_void_p, C.c_void_p,
                                           P(C.c_void_p)]
    lib.vc1me_predict_p_frame_padded.argtypes = lib.vc1me_predict_p_frame.argtypes
    lib.vc1me_predict_bi_frame_padded.argtypes = lib.vc1me_predict_bi_frame.argtypes
    lib.vc1me_set_learned_candidates.argtypes = [C.c_void_p, C.c_int]
    lib.vc1me_frame_download.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.vc1me_analyze_step_seeded.argtypes = [C.c_void_p, P(abi.Job), C.POINTER(P(abi.PResult)), C.c_int]
    lib.vc1me_last_timing.argtypes = [C.c_void_p, P(abi.Timing)]
    lib.vc1me_set_profiling.argtypes = [C.c_void_p, C.c_int]
    lib.vc1me_last_work.argtypes = [C.c_void_p, P(abi.Work)]
    lib.vc1me_block_costs.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(abi.Params), C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, P(abi.Candidate), C.c_size_t, P(C.c_uint64)]
    lib.vc1me_counters_get.argtypes = [C.c_void_p, P(abi.Counters)]
    lib.vc1me_counters_reset.argtypes = [C.c_void_p]
    lib.vc1me_counters_reset.restype = None
    lib.vc1me_nn_param_count.argtypes = []
    lib.vc1me_nn_create.argtypes = [C.c_void_p, P(C.c_void_p)]
    lib.vc1me_nn_destroy.argtypes = [C.c_void_p]
    lib.vc1me_nn_destroy.restype = None
    lib.vc1me_nn_set_params.argtypes = [C.c_void_p, P(C.c_float)]
    lib.vc1me_nn_get_params.argtypes = [C.c_void_p, P(C.c_float)]
    lib.vc1me_nn_forward_frames.argtypes = [C.c_void_p, P(C.c_void_p), P(C.c_void_p), P(C.c_int32), C.c_int,
                                            P(C.c_float)]
    lib.vc1me_nn_forward_backward.argtypes = [C.c_void_p, P(C.c_float), C.c_int, C.c_int, C.c_int, P(C.c_float),
                                              P(C.c_float), P(C.c_float)]
    lib.vc1me_nn_train_step.argtypes = [C.c_void_p, P(C.c_void_p), P(C.c_void_p), P(C.c_int32), P(C.c_int32),
                                        P(C.c_uint8), C.c_int, C.c_float, P(C.c_float)]
    lib.vc1me_analyze_step_learned.argtypes = [C.c_void_p, C.c_void_p, P(abi.Job), C.c_int, C.c_float]
    lib.vc1me_analyze_step_proposals.argtypes = [C.c_void_p, P(abi.Job), C.c_int, P(P(C.c_float)), C.c_float]
    lib.vc1me_set_cost_report.argtypes = [C.c_void_p, C.c_int]
    lib.vc1me_set_graphs.argtypes = [C.c_void_p, C.c_int]
    lib.vc1me_set_schedule.argtypes = [C.c_void_p, C.c_int]
    lib.vc1me_last_relax_passes.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_double), C.c_int]
    lib.vc1me_last_mb_costs.argtypes = [C.c_void_p, C.c_int, P(C.c_uint64), P(C.c_int32)]
    _lib = lib
    return lib


class Frame:
    """A picture resident in device memory (vc1me_frame handle)."""

    def __init__(self, backend: "Backend", handle: int, width: int, height: int):
        self.backend, self.handle, self.width, self.height = backend, handle, width, height

    def release(self) -> None:
        if self.handle:
            self.backend.lib.vc1me_frame_release(self.backend.ctx, self.handle)
            self.handle = None


class Backend:
    """One vc1me context on a HIP device."""

    def __init__(self, device: int = 0):
        self.device = device
        self.lib = load_library()
        ctx = C.c_void_p()
        self._check(self.lib.vc1me_create(device, C.byref(ctx)))
        self.ctx = ctx

    def _check(self, rc: int) -> None:
        if rc != abi.VC1ME_OK:
            raise BackendError(rc, self.lib.vc1me_last_error().decode())

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.vc1me_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_name(self) -> str:
        return self.lib.vc1me_device_name(self.ctx).decode()

    def upload(self, pic) -> Frame:
        ps = abi.PlaneSet.of(pic)
        h = C.c_void_p()
        self._check(self.lib.vc1me_frame_upload(self.ctx, C.byref(ps.struct), C.byref(h)))
        return Frame(self, h.value, ps.struct.width, ps.struct.height)

    def download(self, frame: Frame):
        """(y, u, v) planes of a resident frame (vc1me_frame_download)."""
        y = np.empty((frame.height, frame.width), np.uint8)
        u = np.empty((frame.height // 2, frame.width // 2), np.uint8)
        v = np.empty_like(u)
        self._check(self.lib.vc1me_frame_download(self.ctx, frame.handle, y.ctypes.data, u.ctypes.data,
                                                  v.ctypes.data))
        return y, u, v

    def estimate_intensity(self, cur: Frame, ref: Frame, mvs) -> tuple:
        """vc1me_estimate_intensity_comp: (enabled, lumscale, lumshift) from
        the quarter-pel 1-MV field `mvs` (mbw * mbh pairs)."""
        m = np.ascontiguousarray(np.asarray(mvs, np.int32).reshape(-1, 2))
        mbw, mbh = abi.mb_grid(cur.width, cur.height)
        if m.shape[0] != mbw * mbh:
            raise ValueError(f"{m.shape[0]} vectors for a {mbw}x{mbh} macroblock grid")
        out = abi.Intensity()
        self._check(self.lib.vc1me_estimate_intensity_comp(self.ctx, cur.handle, ref.handle,
                                                           m.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(out)))
        return out.as_tuple()

    def predict_p_frame(self, ref: Frame, params: abi.Params, result: abi.ResultArrays) -> Frame:
        """vc1me_predict_p_frame: the prediction picture as a new resident frame."""
        h = C.c_void_p()
        self._check(self.lib.vc1me_predict_p_frame(self.ctx, ref.handle, C.byref(params), C.byref(result.struct),
                                                   C.byref(h)))
        return Frame(self, h.value, ref.width, ref.height)

    def predict_bi_frame(self, past: Frame, future: Frame, params: abi.Params, fmvs, bmvs) -> Frame:
        """vc1me_predict_bi_frame: a new resident frame."""
        f = np.ascontiguousarray(np.asarray(fmvs, np.int32))
        b = np.ascontiguousarray(np.asarray(bmvs, np.int32))
        h = C.c_void_p()
        self._check(self.lib.vc1me_predict_bi_frame(self.ctx, past.handle, future.handle, C.byref(params),
                                                    f.ctypes.data, b.ctypes.data, C.byref(h)))
        return Frame(self, h.value, past.width, past.height)

    def predict_p_frame_padded(self, ref: Frame, params: abi.Params, result: abi.ResultArrays) -> Frame:
        """vc1me_predict_p_frame_padded: the padded picture (the macroblock
        raster of params) as a new resident frame."""
        h = C.c_void_p()
        self._check(self.lib.vc1me_predict_p_frame_padded(self.ctx, ref.handle, C.byref(params),
                                                          C.byref(result.struct), C.byref(h)))
        mbw, mbh = abi.mb_grid(params.width, params.height)
        return Frame(self, h.value, mbw * 16, mbh * 16)

    def predict_bi_frame_padded(self, past: Frame, future: Frame, params: abi.Params, fmvs, bmvs) -> Frame:
        """vc1me_predict_bi_frame_padded: a new resident frame of the
        macroblock raster of params."""
        f = np.ascontiguousarray(np.asarray(fmvs, np.int32))
        b = np.ascontiguousarray(np.asarray(bmvs, np.int32))
        h = C.c_void_p()
        self._check(self.lib.vc1me_predict_bi_frame_padded(self.ctx, past.handle, future.handle, C.byref(params),
                                                           f.ctypes.data, b.ctypes.data, C.byref(h)))
        mbw, mbh = abi.mb_grid(params.width, params.height)
        return Frame(self, h.value, mbw * 16, mbh * 16)

    def intensity_compensated(self, ref: Frame, ic: tuple) -> Frame:
        """vc1me_frame_intensity_compensated: a new resident frame."""
        h = C.c_void_p()
        self._check(self.lib.vc1me_frame_intensity_compensated(self.ctx, ref.handle, C.byref(abi.Intensity(*ic)),
                                                               C.byref(h)))
        return Frame(self, h.value, ref.width, ref.height)

    def analyze_b(self, cur: Frame, past: Frame, future: Frame, params: abi.Params,
                  binp: abi.BInputs) -> abi.BResultArrays:
        """vc1me_analyze_b_frame: params.mode is the B picture's 1-MV mode."""
        r = abi.BResultArrays(params.width, params.height)
        self._check(self.lib.vc1me_analyze_b_frame(self.ctx, cur.handle, past.handle, future.handle, C.byref(params),
                                                   C.byref(binp.struct), C.byref(r.struct)))
        return r

    def analyze_b_step(self, items) -> list:
        """vc1me_analyze_b_step over items of (cur, past, future, params,
        BInputs); one BResultArrays each."""
        n = len(items)
        jobs = (abi.BJob * n)()
        outs = []
        for i, (cur, past, fut, p, binp) in enumerate(items):
            r = abi.BResultArrays(p.width, p.height)
            outs.append(r)
            jobs[i].cur, jobs[i].past, jobs[i].future = cur.handle, past.handle, fut.handle
            jobs[i].params = C.pointer(p)
            jobs[i].b = C.pointer(binp.struct)
            jobs[i].out = C.pointer(r.struct)
        self._check(self.lib.vc1me_analyze_b_step(self.ctx, jobs, n))
        return outs

    def has_good_motion_match(self, cur: Frame, ref: Frame, params: abi.Params, fade: bool = True,
                              want_field: bool = False):
        """vc1me_has_good_motion_match: the answer, or (answer, field) with
        every macroblock's refined vector."""
        out = C.c_int()
        mbw, mbh = abi.mb_grid(cur.width, cur.height)
        field = np.zeros((mbw * mbh, 2), np.int32) if want_field else None
        self._check(self.lib.vc1me_has_good_motion_match(self.ctx, cur.handle, ref.handle, C.byref(params), int(fade),
                                                         C.byref(out), field.ctypes.data if want_field else None))
        return (bool(out.value), field) if want_field else bool(out.value)

    def scene_change_score(self, cur: Frame, ref: Frame, params: abi.Params, fade: bool = True,
                           threshold: float = 28.0) -> float:
        """vc1me_scene_change_score."""
        out = C.c_double()
        self._check(self.lib.vc1me_scene_change_score(self.ctx, cur.handle, ref.handle, C.byref(params), int(fade),
                                                      float(threshold), C.byref(out)))
        return out.value

    def analyze(self, cur: Frame, ref: Frame, params: abi.Params) -> abi.ResultArrays:
        r = abi.ResultArrays(params.width, params.height)
        self._check(self.lib.vc1me_analyze_p_frame(self.ctx, cur.handle, ref.handle, C.byref(params), C.byref(r.struct)))
        return r

    @staticmethod
    def _jobs(pairs, params_list, want_results: bool):
        n = len(pairs)
        jobs = (abi.Job * n)()
        results = []
        for i, ((cur, ref), p) in enumerate(zip(pairs, params_list)):
            jobs[i].cur = cur.handle
            jobs[i].ref = ref.handle
            jobs[i].params = C.pointer(p)
            if want_results:
                r = abi.ResultArrays(p.width, p.height)
                results.append(r)
                jobs[i].out = C.pointer(r.struct)
            else:
                jobs[i].out = None
        return jobs, results

    def analyze_step(self, pairs, params_list, want_results: bool = True):
        jobs, results = self._jobs(pairs, params_list, want_results)
        self._check(self.lib.vc1me_analyze_step(self.ctx, jobs, len(pairs)))
        return results

    def analyze_seeded(self, cur: Frame, ref: Frame, params: abi.Params, seed: abi.ResultArrays) -> abi.ResultArrays:
        """The seeded core (vc1me_analyze_p_frame_seeded): `seed` is a
        result of the same raster, normally the 1-MV quarter-pel pass."""
        r = abi.ResultArrays(params.width, params.height)
        self._check(self.lib.vc1me_analyze_p_frame_seeded(self.ctx, cur.handle, ref.handle, C.byref(params),
                                                          C.byref(seed.struct), C.byref(r.struct)))
        return r

    def analyze_step_seeded(self, pairs, params_list, seeds, want_results: bool = True):
        """analyze_step of the seeded core, seeds[i] seeding pair i."""
        jobs, results = self._jobs(pairs, params_list, want_results)
        n = len(pairs)
        arr = (C.POINTER(abi.PResult) * n)(*[C.pointer(s.struct) for s in seeds])
        self._check(self.lib.vc1me_analyze_step_seeded(self.ctx, jobs, arr, n))
        return results

    def analyze_step_learned(self, net: "ProposalNet", pairs, params_list, threshold: float = 0.5,
                             want_results: bool = True):
        """analyze_step with the learned seeding of `net`."""
        jobs, results = self._jobs(pairs, params_list, want_results)
        self._check(self.lib.vc1me_analyze_step_learned(self.ctx, net.handle, jobs, len(pairs), threshold))
        return results

    def analyze_step_proposals(self, pairs, params_list, proposals, threshold: float = 0.5,
                               want_results: bool = True):
        """analyze_step seeded by caller-given proposals, one (mbh, mbw, 5)
        float array per pair in the network's output layout."""
        jobs, results = self._jobs(pairs, params_list, want_results)
        n = len(pairs)
        props = [np.ascontiguousarray(p, np.float32) for p in proposals]
        for p, prm in zip(props, params_list):
            mbw, mbh = abi.mb_grid(prm.width, prm.height)
            assert p.size == mbw * mbh * 5, (p.shape, mbw, mbh)
        arr = (C.POINTER(C.c_float) * n)(*[p.ctypes.data_as(C.POINTER(C.c_float)) for p in props])
        self._check(self.lib.vc1me_analyze_step_proposals(self.ctx, jobs, n, arr, threshold))
        return results

    def set_learned_candidates(self, on: bool) -> None:
        """Use the proposals' vectors in learned seeding (vc1me_set_learned_candidates)."""
        self._check(self.lib.vc1me_set_learned_candidates(self.ctx, 1 if on else 0))

    def set_graphs(self, on: bool) -> None:
        """Replay plain steps as captured HIP graphs (vc1me_set_graphs)."""
        self._check(self.lib.vc1me_set_graphs(self.ctx, 1 if on else 0))

    def set_schedule(self, schedule: int) -> None:
        """Macroblock schedule of later steps (abi.SCHEDULE_DIAGONAL or
        abi.SCHEDULE_RELAX; vc1me_set_schedule)."""
        self._check(self.lib.vc1me_set_schedule(self.ctx, int(schedule)))

    def relax_passes(self):
        """[(analyses, device ms or -1)] per relaxation pass of the last step."""
        cap = 128
        a = (C.c_int64 * cap)()
        m = (C.c_double * cap)()
        n = self.lib.vc1me_last_relax_passes(self.ctx, a, m, cap)
        if n < 0:
            self._check(n)
        return [(int(a[k]), float(m[k])) for k in range(min(n, cap))]

    def set_cost_report(self, on: bool) -> None:
        self._check(self.lib.vc1me_set_cost_report(self.ctx, 1 if on else 0))

    def last_mb_costs(self, job: int, mbs: int):
        """(cost20 (mbs,) uint64, 1-MV vectors (mbs, 2) int32) of job `job`
        of the last step; needs set_cost_report(True) before that step."""
        c = np.zeros(mbs, np.uint64)
        v = np.zeros((mbs, 2), np.int32)
        self._check(self.lib.vc1me_last_mb_costs(self.ctx, job, c.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                 v.ctypes.data_as(C.POINTER(C.c_int32))))
        return c, v

    def score_batch(self, cur: np.ndarray, ref: np.ndarray, x0: int, y0: int, bw: int, bh: int,
                    cands: np.ndarray, sample_stride: int = 1, bilinear: int = 0, rnd: int = 0,
                    metric: int = abi.METRIC_SAD) -> np.ndarray:
        cur = np.ascontiguousarray(cur, np.uint8)
        ref = np.ascontiguousarray(ref, np.uint8)
        h, w = cur.shape
        cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
        out = np.zeros(len(cands), np.uint64)
        self._check(self.lib.vc1me_score_batch(
            self.ctx, cur.ctypes.data, ref.ctypes.data, w, h, x0, y0, bw, bh, sample_stride, bilinear, rnd,
            metric, cands.ctypes.data_as(C.POINTER(abi.Candidate)), len(cands),
            out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def score_batch_frames(self, cur: Frame, ref: Frame, x0: int, y0: int, bw: int, bh: int, cands: np.ndarray,
                           sample_stride: int = 1, bilinear: int = 0, rnd: int = 0, metric: int = abi.METRIC_SAD,
                           noalloc: bool = False) -> np.ndarray:
        """vc1me_score_batch_frames (or its allocation-free form) on resident
        frames."""
        cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
        out = np.zeros(len(cands), np.uint64)
        fn = self.lib.vc1me_score_batch_frames_noalloc if noalloc else self.lib.vc1me_score_batch_frames
        self._check(fn(self.ctx, cur.handle, ref.handle, x0, y0, bw, bh, sample_stride, bilinear, rnd, metric,
                       cands.ctypes.data_as(C.POINTER(abi.Candidate)), len(cands),
                       out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def score_batch_gen(self, cur: np.ndarray, cur_gen: int, ref: np.ndarray, ref_gen: int, x0: int, y0: int,
                        bw: int, bh: int, cands: np.ndarray, sample_stride: int = 1, bilinear: int = 0, rnd: int = 0,
                        metric: int = abi.METRIC_SAD) -> np.ndarray:
        """vc1me_score_batch_gen: the planes' device copies keyed by the
        caller's generation numbers."""
        cur = np.ascontiguousarray(cur, np.uint8)
        ref = np.ascontiguousarray(ref, np.uint8)
        h, w = cur.shape
        cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
        out = np.zeros(len(cands), np.uint64)
        self._check(self.lib.vc1me_score_batch_gen(
            self.ctx, cur.ctypes.data, cur_gen, ref.ctypes.data, ref_gen, w, h, x0, y0, bw, bh, sample_stride,
            bilinear, rnd, metric, cands.ctypes.data_as(C.POINTER(abi.Candidate)), len(cands),
            out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def block_costs(self, cur: Frame, ref: Frame, params: abi.Params, x0: int, y0: int, bw: int, bh: int,
                    mode: int, metric: int, cands: np.ndarray) -> np.ndarray:
        cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
        out = np.zeros(len(cands), np.uint64)
        self._check(self.lib.vc1me_block_costs(
            self.ctx, cur.handle, ref.handle, C.byref(params), x0, y0, bw, bh, mode, metric,
            cands.ctypes.data_as(C.POINTER(abi.Candidate)), len(cands), out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def set_profiling(self, on: bool) -> None:
        self._check(self.lib.vc1me_set_profiling(self.ctx, 1 if on else 0))

    def timing(self) -> abi.Timing:
        t = abi.Timing()
        self._check(self.lib.vc1me_last_timing(self.ctx, C.byref(t)))
        return t

    def work(self) -> abi.Work:
        w = abi.Work()
        self._check(self.lib.vc1me_last_work(self.ctx, C.byref(w)))
        return w

    def counters(self) -> abi.Counters:
        c = abi.Counters()
        self._check(self.lib.vc1me_counters_get(self.ctx, C.byref(c)))
        return c


class ProposalNet:
    """The learned-seeding proposal network of a Backend (vc1me_nn_*)."""

    def __init__(self, backend: Backend):
        self.backend = backend
        self.lib = backend.lib
        h = C.c_void_p()
        backend._check(self.lib.vc1me_nn_create(backend.ctx, C.byref(h)))
        self.handle = h
        self.nparams = int(self.lib.vc1me_nn_param_count())

    def close(self) -> None:
        if self.handle:
            self.lib.vc1me_nn_destroy(self.handle)
            self.handle = None

    def set_params(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, np.float32)
        assert params.size == self.nparams
        self.backend._check(self.lib.vc1me_nn_set_params(self.handle, params.ctypes.data_as(C.POINTER(C.c_float))))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        self.backend._check(self.lib.vc1me_nn_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    @staticmethod
    def grid(width: int, height: int):
        h4, w4 = (height + 3) // 4, (width + 3) // 4
        return h4, w4, (h4 + 3) // 4, (w4 + 3) // 4

    @staticmethod
    def _globals(globals_, n):
        if globals_ is None:
            return None, None
        g = np.ascontiguousarray(globals_, np.int32).reshape(n, 2)
        return g, g.ctypes.data_as(C.POINTER(C.c_int32))

    def forward_frames(self, pairs, globals_=None) -> np.ndarray:
        """Outputs (n, mbh, mbw, 5); globals_ (n, 2) integer pixels or None."""
        n = len(pairs)
        curs = (C.c_void_p * n)(*[c.handle for c, _ in pairs])
        refs = (C.c_void_p * n)(*[r.handle for _, r in pairs])
        _, _, mbh, mbw = self.grid(pairs[0][0].width, pairs[0][0].height)
        out = np.zeros((n, mbh, mbw, 5), np.float32)
        g, gp = self._globals(globals_, n)
        self.backend._check(self.lib.vc1me_nn_forward_frames(self.handle, curs, refs, gp, n,
                                                            out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def forward_backward(self, x: np.ndarray, dout: np.ndarray):
        """x: (n, h4, w4, 2) float; dout: (n, mbh, mbw, 5).  Returns (out, grads)."""
        x = np.ascontiguousarray(x, np.float32)
        dout = np.ascontiguousarray(dout, np.float32)
        n, h4, w4, _ = x.shape
        out = np.zeros_like(dout)
        grads = np.zeros(self.nparams, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self.backend._check(self.lib.vc1me_nn_forward_backward(self.handle, fp(x), n, h4, w4, fp(dout), fp(out),
                                                              fp(grads)))
        return out, grads

    def train_step(self, pairs, targets: np.ndarray, valid: np.ndarray, lr: float, globals_=None) -> float:
        n = len(pairs)
        curs = (C.c_void_p * n)(*[c.handle for c, _ in pairs])
        refs = (C.c_void_p * n)(*[r.handle for _, r in pairs])
        targets = np.ascontiguousarray(targets, np.int32)
        valid = np.ascontiguousarray(valid, np.uint8)
        g, gp = self._globals(globals_, n)
        loss = C.c_float()
        self.backend._check(self.lib.vc1me_nn_train_step(
            self.handle, curs, refs, gp, targets.ctypes.data_as(C.POINTER(C.c_int32)),
            valid.ctypes.data_as(C.POINTER(C.c_uint8)), n, lr, C.byref(loss)))
        return float(loss.value)

