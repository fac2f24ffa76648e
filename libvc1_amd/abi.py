"""ctypes mirror of include/vc1me.h (the C ABI of the motion-estimation backend).

The structs here are byte-for-byte the ones declared in include/vc1me.h; the
same definitions serve any shared library that speaks that ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

VC1ME_OK = 0
VC1ME_EINVAL = -1
VC1ME_EDEVICE = -2
VC1ME_ENOMEM = -3
VC1ME_ELIMIT = -4
VC1ME_EINJECTED = -5
VC1ME_EKERNEL = -6
VC1ME_NOALLOC_MAX_CANDIDATES = 65536   # include/vc1me.h

METRIC_SAD = 0
METRIC_SATD = 1

MODE_ONE_MV_QPEL = 0
MODE_ONE_MV_HPEL = 1
MODE_ONE_MV_HPEL_BILINEAR = 2
MODE_MIXED_MV = 3

# macroblock schedules (vc1me_set_schedule)
SCHEDULE_DIAGONAL = 0
SCHEDULE_RELAX = 1


class Candidate(C.Structure):
    _fields_ = [("xq", C.c_int32), ("yq", C.c_int32)]


class Params(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("display_width", C.c_int32), ("display_height", C.c_int32),
        ("pqindex", C.c_int32),
        ("motion_search_range", C.c_int32), ("motion_local_search_range", C.c_int32),
        ("me_algorithm", C.c_int32), ("me_quality", C.c_int32),
        ("subpixel_refinement", C.c_int32), ("fast_subpixel_refinement", C.c_int32),
        ("subpixel_refinement_range", C.c_int32),
        ("long_range_search_mode", C.c_int32),
        ("mode", C.c_int32), ("allow_intra", C.c_int32), ("debug_disable_p_intra", C.c_int32),
        ("distant_match_max_mae", C.c_double), ("inter_intra_threshold", C.c_double),
    ]


class PResult(C.Structure):
    _fields_ = [
        ("mvs", C.POINTER(C.c_int32)),
        ("block_mvs", C.POINTER(C.c_int32)),
        ("use_4mv", C.POINTER(C.c_uint8)),
        ("use_intra", C.POINTER(C.c_uint8)),
        ("screen_sads", C.POINTER(C.c_uint64)),
        ("inter_intra_cost_ratio", C.POINTER(C.c_double)),
        ("distant_local_mae", C.POINTER(C.c_double)),
        ("distant_candidate_mae", C.POINTER(C.c_double)),
        ("distant_match_decision", C.POINTER(C.c_uint8)),
        ("mean_abs_residual", C.c_double),
        ("rate_complexity", C.c_double),
        ("motion_sad", C.c_uint64),
        ("moved_macroblocks", C.c_uint64),
        ("fractional_chroma_macroblocks", C.c_uint64),
        ("four_mv_macroblocks", C.c_uint64),
        ("intra_macroblocks", C.c_uint64),
        ("global_dx", C.c_int32),
        ("global_dy", C.c_int32),
        ("global_sad", C.c_uint64),
    ]


class Planes(C.Structure):
    _fields_ = [("y", C.POINTER(C.c_uint8)), ("u", C.POINTER(C.c_uint8)), ("v", C.POINTER(C.c_uint8)),
                ("width", C.c_int32), ("height", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [("dispatches", C.c_uint64), ("candidates", C.c_uint64),
                ("frame_uploads", C.c_uint64), ("frame_upload_bytes", C.c_uint64),
                ("submit_wait_nanoseconds", C.c_uint64),
                ("frames_analyzed", C.c_uint64), ("macroblocks_analyzed", C.c_uint64)]


class Timing(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("mb_kernel_ms_total", C.c_double),
                ("mb_kernel_launches", C.c_uint64), ("global_kernel_ms_total", C.c_double),
                ("global_kernel_launches", C.c_uint64), ("graph_replay", C.c_uint64),
                ("schedule", C.c_uint64), ("relax_passes", C.c_uint64), ("relax_analyses", C.c_uint64),
                ("prepass_ms", C.c_double)]


class Work(C.Structure):
    _fields_ = [("sad_px", C.c_uint64), ("interp_taps", C.c_uint64), ("satd_ops", C.c_uint64),
                ("sig_ops", C.c_uint64), ("frame_px", C.c_uint64), ("seeded_macroblocks", C.c_uint64)]


class Intensity(C.Structure):
    """vc1me_intensity: LUMSCALE / LUMSHIFT as coded, enabled 0 = none."""
    _fields_ = [("enabled", C.c_int32), ("lumscale", C.c_int32), ("lumshift", C.c_int32)]

    def as_tuple(self):
        return (int(self.enabled), int(self.lumscale), int(self.lumshift))


B_FORWARD, B_BACKWARD, B_INTERPOLATED, B_DIRECT, B_INTRA = 0, 1, 2, 3, 4


class BParams(C.Structure):
    _fields_ = [("future_anchor_mvs", C.POINTER(C.c_int32)), ("future_anchor_4mv", C.POINTER(C.c_uint8)),
                ("forward_intensity", Intensity), ("fraction_num", C.c_int32), ("fraction_den", C.c_int32),
                ("allow_intra", C.c_int32), ("zero_motion_shortcut", C.c_int32)]


class BResult(C.Structure):
    _fields_ = [
        ("modes", C.POINTER(C.c_uint8)), ("forward_mvs", C.POINTER(C.c_int32)),
        ("backward_mvs", C.POINTER(C.c_int32)), ("inter_intra_cost_ratio", C.POINTER(C.c_double)),
        ("forward_distant_local_mae", C.POINTER(C.c_double)),
        ("forward_distant_candidate_mae", C.POINTER(C.c_double)),
        ("forward_distant_match_decision", C.POINTER(C.c_uint8)),
        ("backward_distant_local_mae", C.POINTER(C.c_double)),
        ("backward_distant_candidate_mae", C.POINTER(C.c_double)),
        ("backward_distant_match_decision", C.POINTER(C.c_uint8)),
        ("forward_macroblocks", C.c_uint64), ("backward_macroblocks", C.c_uint64),
        ("interpolated_macroblocks", C.c_uint64), ("direct_macroblocks", C.c_uint64),
        ("intra_macroblocks", C.c_uint64), ("moved_macroblocks", C.c_uint64),
        ("fractional_chroma_macroblocks", C.c_uint64),
        ("mean_abs_residual", C.c_double), ("rate_complexity", C.c_double),
        ("motion_sad", C.c_uint64), ("motion_bits", C.c_uint64),
    ]


class BJob(C.Structure):
    _fields_ = [("cur", C.c_void_p), ("past", C.c_void_p), ("future", C.c_void_p), ("params", C.POINTER(Params)),
                ("b", C.POINTER(BParams)), ("out", C.POINTER(BResult))]


class BInputs:
    """numpy-backed vc1me_b_params: the future anchor's fields and the rest."""

    def __init__(self, anchor_mvs, anchor_4mv, intensity=(0, 32, 0), fraction=(1, 2), allow_intra=1,
                 zero_motion_shortcut=0):
        self.mvs = np.ascontiguousarray(np.asarray(anchor_mvs, np.int32).reshape(-1, 2))
        self.four = np.ascontiguousarray(np.asarray(anchor_4mv, np.uint8).reshape(-1))
        self.struct = BParams(self.mvs.ctypes.data_as(C.POINTER(C.c_int32)),
                              self.four.ctypes.data_as(C.POINTER(C.c_uint8)), Intensity(*intensity),
                              fraction[0], fraction[1], allow_intra, zero_motion_shortcut)


class BResultArrays:
    """numpy-backed storage for one vc1me_b_result."""

    ARRAYS = ("modes", "forward_mvs", "backward_mvs", "inter_intra_cost_ratio", "forward_distant_local_mae",
              "forward_distant_candidate_mae", "forward_distant_match_decision", "backward_distant_local_mae",
              "backward_distant_candidate_mae", "backward_distant_match_decision")
    SCALARS = ("forward_macroblocks", "backward_macroblocks", "interpolated_macroblocks", "direct_macroblocks",
               "intra_macroblocks", "moved_macroblocks", "fractional_chroma_macroblocks", "mean_abs_residual",
               "rate_complexity", "motion_sad", "motion_bits")

    def __init__(self, width: int, height: int):
        mbw, mbh = mb_grid(width, height)
        n = mbw * mbh
        self.mbs = n
        self.struct = BResult()
        for k in self.ARRAYS:
            if k.endswith("mvs"):
                a = np.zeros((n, 2), np.int32)
            elif k in ("modes",) or k.endswith("decision"):
                a = np.zeros(n, np.uint8)
            else:
                a = np.zeros(n, np.float64)
            setattr(self, k, a)
            ct = {np.int32: C.c_int32, np.uint8: C.c_uint8, np.float64: C.c_double}[a.dtype.type]
            setattr(self.struct, k, a.ctypes.data_as(C.POINTER(ct)))

    def scalars(self) -> dict:
        return {k: getattr(self.struct, k) for k in self.SCALARS}

    def diff(self, other: "BResultArrays") -> list[str]:
        """Names of fields that are not bit-identical (doubles compared as bits)."""
        bad = []
        for k in self.ARRAYS:
            a, b = getattr(self, k), getattr(other, k)
            if a.dtype == np.float64:
                if not np.array_equal(a.view(np.uint64), b.view(np.uint64)):
                    bad.append(k)
            elif not np.array_equal(a, b):
                bad.append(k)
        for k in self.SCALARS:
            a, b = getattr(self.struct, k), getattr(other.struct, k)
            if isinstance(a, float):
                if np.float64(a).view(np.uint64) != np.float64(b).view(np.uint64):
                    bad.append(k)
            elif a != b:
                bad.append(k)
        return bad


class Job(C.Structure):
    _fields_ = [("cur", C.c_void_p), ("ref", C.c_void_p), ("params", C.POINTER(Params)),
                ("out", C.POINTER(PResult))]


def default_params(width: int, height: int, **over) -> Params:
    """The reference's EncoderConfig defaults for the ME fields
    (src/encoder_internal.h:243-253, :289-302); keyword arguments override."""
    p = Params(width=width, height=height, display_width=0, display_height=0, pqindex=9,
               motion_search_range=1024, motion_local_search_range=32, me_algorithm=0,
               me_quality=2, subpixel_refinement=1, fast_subpixel_refinement=0,
               subpixel_refinement_range=1024, long_range_search_mode=0,
               mode=MODE_ONE_MV_QPEL, allow_intra=1, debug_disable_p_intra=0,
               distant_match_max_mae=255.0, inter_intra_threshold=0.80)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def mb_grid(width: int, height: int) -> tuple[int, int]:
    return (width + 15) // 16, (height + 15) // 16


class ResultArrays:
    """numpy-backed storage for one vc1me_p_result."""

    def __init__(self, width: int, height: int):
        mbw, mbh = mb_grid(width, height)
        n = mbw * mbh
        self.mbs = n
        self.mvs = np.zeros((n, 2), np.int32)
        self.block_mvs = np.zeros((n, 4, 2), np.int32)
        self.use_4mv = np.zeros(n, np.uint8)
        self.use_intra = np.zeros(n, np.uint8)
        self.screen_sads = np.zeros(n, np.uint64)
        self.inter_intra_cost_ratio = np.zeros(n, np.float64)
        self.distant_local_mae = np.zeros(n, np.float64)
        self.distant_candidate_mae = np.zeros(n, np.float64)
        self.distant_match_decision = np.zeros(n, np.uint8)
        self.struct = PResult()
        s = self.struct
        s.mvs = self.mvs.ctypes.data_as(C.POINTER(C.c_int32))
        s.block_mvs = self.block_mvs.ctypes.data_as(C.POINTER(C.c_int32))
        s.use_4mv = self.use_4mv.ctypes.data_as(C.POINTER(C.c_uint8))
        s.use_intra = self.use_intra.ctypes.data_as(C.POINTER(C.c_uint8))
        s.screen_sads = self.screen_sads.ctypes.data_as(C.POINTER(C.c_uint64))
        s.inter_intra_cost_ratio = self.inter_intra_cost_ratio.ctypes.data_as(C.POINTER(C.c_double))
        s.distant_local_mae = self.distant_local_mae.ctypes.data_as(C.POINTER(C.c_double))
        s.distant_candidate_mae = self.distant_candidate_mae.ctypes.data_as(C.POINTER(C.c_double))
        s.distant_match_decision = self.distant_match_decision.ctypes.data_as(C.POINTER(C.c_uint8))

    ARRAYS = ("mvs", "block_mvs", "use_4mv", "use_intra", "screen_sads", "inter_intra_cost_ratio",
              "distant_local_mae", "distant_candidate_mae", "distant_match_decision")
    SCALARS = ("mean_abs_residual", "rate_complexity", "motion_sad", "moved_macroblocks",
               "fractional_chroma_macroblocks", "four_mv_macroblocks", "intra_macroblocks",
               "global_dx", "global_dy", "global_sad")

    def scalars(self) -> dict:
        return {k: getattr(self.struct, k) for k in self.SCALARS}

    def diff(self, other: "ResultArrays") -> list[str]:
        """Names of fields that are not bit-identical (doubles compared as bits)."""
        bad = []
        for k in self.ARRAYS:
            a, b = getattr(self, k), getattr(other, k)
            if a.dtype == np.float64:
                if not np.array_equal(a.view(np.uint64), b.view(np.uint64)):
                    bad.append(k)
            elif not np.array_equal(a, b):
                bad.append(k)
        for k in self.SCALARS:
            a, b = getattr(self.struct, k), getattr(other.struct, k)
            if isinstance(a, float):
                if np.float64(a).view(np.uint64) != np.float64(b).view(np.uint64):
                    bad.append(k)
            elif a != b:
                bad.append(k)
        return bad


class PlaneSet:
    """Keeps numpy planes alive behind a vc1me_planes struct."""

    def __init__(self, y: np.ndarray, u: np.ndarray | None = None, v: np.ndarray | None = None):
        self.y = np.ascontiguousarray(y, dtype=np.uint8)
        h, w = self.y.shape
        self.u = np.ascontiguousarray(u if u is not None else np.full((h // 2, w // 2), 128, np.uint8), dtype=np.uint8)
        self.v = np.ascontiguousarray(v if v is not None else np.full((h // 2, w // 2), 128, np.uint8), dtype=np.uint8)
        self.struct = Planes(self.y.ctypes.data_as(C.POINTER(C.c_uint8)),
                             self.u.ctypes.data_as(C.POINTER(C.c_uint8)),
                             self.v.ctypes.data_as(C.POINTER(C.c_uint8)), w, h)

    @classmethod
    def of(cls, pic) -> "PlaneSet":
        return cls(pic.y, pic.u, pic.v)
