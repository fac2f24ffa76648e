"""The allocation-free batch scorer (vc1me_score_batch_frames_noalloc)
against the allocating one (vc1me_score_batch_frames) through the C ABI: two
calls in a row on the same frames, the second with fewer candidates, must
give the same costs bit for bit; a list above the limit is refused before
any device work."""
import ctypes as C

import numpy as np
import pytest

from libvc1_amd import abi, synth

pytestmark = pytest.mark.gpu


def _score(backend, fn, cur, ref, x0, y0, bw, bh, stride, bil, metric, cands):
    cands = np.ascontiguousarray(cands, np.int32).reshape(-1, 2)
    out = np.full(len(cands), 0xDEADBEEF, np.uint64)
    rc = fn(backend.ctx, cur.handle, ref.handle, x0, y0, bw, bh, stride, bil, 0, metric,
            cands.ctypes.data_as(C.POINTER(abi.Candidate)), len(cands), out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, out


def test_noalloc_scorer_equals_the_allocating_one(backend):
    d = synth.pan_rectangles(41, 3, width=208, height=120, max_pan_px=8, max_obj_px=16, frame_indices=[0, 2])
    fc, fr = backend.upload(d[2]), backend.upload(d[0])
    lib = backend.lib
    rng = np.random.default_rng(3)
    try:
        # (x0, y0, bw, bh, sample_stride, bilinear, metric, n): a large list,
        # then a smaller one that leaves stale entries of the first behind
        for x0, y0, bw, bh, stride, bil, metric, n in ((32, 48, 16, 16, 1, 0, abi.METRIC_SAD, 5000),
                                                      (200, 112, 8, 8, 1, 1, abi.METRIC_SATD, 37),
                                                      (96, 16, 16, 16, 4, 0, abi.METRIC_SAD, 1)):
            cands = rng.integers(-160, 161, size=(n, 2))
            if stride > 1 or bil:
                cands &= ~1 if bil else ~3        # even (half-pel) or integer vectors, as the callers send
            print(f"[case] noalloc scorer n={n} block {bw}x{bh} at ({x0},{y0}) stride {stride} "
                  f"bilinear {bil} metric {metric}", flush=True)
            rc, want = _score(backend, lib.vc1me_score_batch_frames, fc, fr, x0, y0, bw, bh, stride, bil, metric,
                              cands)
            assert rc == abi.VC1ME_OK, lib.vc1me_last_error()
            rc, got = _score(backend, lib.vc1me_score_batch_frames_noalloc, fc, fr, x0, y0, bw, bh, stride, bil,
                             metric, cands)
            assert rc == abi.VC1ME_OK, lib.vc1me_last_error()
            np.testing.assert_array_equal(got, want)
        big = np.zeros((abi.VC1ME_NOALLOC_MAX_CANDIDATES + 1, 2), np.int32)
        rc, _ = _score(backend, lib.vc1me_score_batch_frames_noalloc, fc, fr, 0, 0, 16, 16, 1, 0, 0, big)
        assert rc == abi.VC1ME_ELIMIT
    finally:
        fc.release()
        fr.release()
