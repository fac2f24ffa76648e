"""HIP backend vs the CPU oracle: bit-exact P-frame motion estimation.

Every test calls the product through the C ABI (libvc1_amd/libvc1me.so) and
compares every PAnalysis field, doubles included, bit for bit with the
oracle's restatement of Vc1Encoder::analyze_p_picture_core
(src/encoder_motion.cpp:47-442) on the same seeded inputs.
"""
import numpy as np
import pytest

from libvc1_amd import abi, synth
from _oracle import describe_diff, oracle_analyze, oracle_score_batch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _check_frame(backend, cur, ref, params, label):
    _say(f"[case] {label} {params.width}x{params.height} mode={params.mode} R={params.motion_search_range} "
         f"local={params.motion_local_search_range} algo={params.me_algorithm} q={params.me_quality}")
    want = oracle_analyze(params, cur, ref)
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        got = backend.analyze(fc, fr, params)
    finally:
        fc.release()
        fr.release()
    assert not got.diff(want), f"{label}\n" + describe_diff(got, want, params.width)
    return got


def test_aaa_block_cost_primitive_first(backend):
    test_block_cost_primitive_fuzz(backend)


def test_score_batch_matches_kernel_law(backend):
    rng = np.random.default_rng(11)
    for (w, h) in ((48, 40), (352, 288)):
        cur = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
        ref = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
        for metric in (0, 1):
            for stride in ((1, 2, 4) if metric == 0 else (1,)):
                for bil in (0, 1):
                    for (x0, y0, bw, bh) in ((16, 16, 16, 16), (w - 16, h - 8, 16, 16), (8, 8, 8, 8), (w - 8, h - 8, 8, 8)):
                        cands = rng.integers(-40, 41, size=(37, 2)).astype(np.int32)
                        if bil:
                            cands &= ~1
                        _say(f"[case] score_batch {w}x{h} metric={metric} stride={stride} bil={bil} blk=({x0},{y0},{bw},{bh})")
                        got = backend.score_batch(cur, ref, x0, y0, bw, bh, cands, stride, bil, 0, metric)
                        want = oracle_score_batch(cur, ref, x0, y0, bw, bh, cands, stride, bil, 0, metric)
                        assert np.array_equal(got, want)


def test_score_batch_generations(backend):
    """vc1me_score_batch_gen keys the planes' device copies by the caller's
    generation numbers: the kernel law as before; a plane written in place
    under a new generation is scored afresh, and the checksum-keyed
    vc1me_score_batch notices the write by itself."""
    rng = np.random.default_rng(12)
    cur = rng.integers(0, 256, size=(96, 160), dtype=np.uint8)
    ref = rng.integers(0, 256, size=(96, 160), dtype=np.uint8)
    cands = rng.integers(-40, 41, size=(36, 2)).astype(np.int32)
    _say("[case] score_batch_gen")
    want = oracle_score_batch(cur, ref, 32, 16, 16, 16, cands, 1, 0, 0, 0)
    assert np.array_equal(backend.score_batch_gen(cur, 1, ref, 1, 32, 16, 16, 16, cands), want)
    ref[:, :] = 255 - ref                       # written in place
    want2 = oracle_score_batch(cur, ref, 32, 16, 16, 16, cands, 1, 0, 0, 0)
    assert not np.array_equal(want, want2)
    assert np.array_equal(backend.score_batch_gen(cur, 1, ref, 2, 32, 16, 16, 16, cands), want2)
    assert np.array_equal(backend.score_batch(cur, ref, 32, 16, 16, 16, cands), want2)


def test_score_batch_empty_and_limits(backend):
    cur = np.zeros((32, 32), np.uint8)
    out = backend.score_batch(cur, cur, 0, 0, 16, 16, np.zeros((0, 2), np.int32))
    assert out.shape == (0,)


def test_cif_translation_mixed_mv(backend):
    seq = synth.cif_translation(1, frames=3)
    p = abi.default_params(352, 288, pqindex=8, motion_search_range=64, motion_local_search_range=64,
                           mode=abi.MODE_MIXED_MV)
    _check_frame(backend, seq[1], seq[0], p, "cif mixed")


def test_pan_long_range_hidden_padding(backend):
    d = synth.pan_rectangles(3, 8, width=640, height=360, frame_indices=[0, 3])
    p = abi.default_params(640, 360, pqindex=8, motion_search_range=512, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    got = _check_frame(backend, d[3], d[0], p, "pan 640x360 long range")
    assert got.use_4mv.sum() > 0


@pytest.mark.parametrize("over", [
    dict(me_algorithm=1),
    dict(fast_subpixel_refinement=1),
    dict(me_quality=1),
    dict(long_range_search_mode=1),
    dict(long_range_search_mode=2),
    dict(motion_search_range=0),
    dict(motion_search_range=24, motion_local_search_range=24),
    dict(motion_search_range=7, motion_local_search_range=7),
    dict(mode=abi.MODE_ONE_MV_HPEL),
    dict(mode=abi.MODE_ONE_MV_HPEL_BILINEAR),
    dict(mode=abi.MODE_ONE_MV_QPEL, allow_intra=0),
    dict(subpixel_refinement_range=2),
    dict(pqindex=2),
    dict(pqindex=31),
])
def test_option_matrix(backend, over):
    d = synth.pan_rectangles(5, 4, width=320, height=176, max_pan_px=6, max_obj_px=20, frame_indices=[0, 2])
    base = dict(pqindex=8, motion_search_range=128, motion_local_search_range=32, mode=abi.MODE_MIXED_MV)
    base.update(over)
    p = abi.default_params(320, 176, **base)
    _check_frame(backend, d[2], d[0], p, f"options {over}")


def test_odd_display_and_small_frames(backend):
    rng = np.random.default_rng(3)
    for (w, h, dw, dh) in ((48, 40, 47, 39), (64, 48, 0, 0), (96, 64, 90, 60), (16, 16, 0, 0)):
        base = np.clip(rng.normal(128, 40, size=(h + 8, w + 8)), 0, 255).astype(np.uint8)
        cy, ry = base[4:4 + h, 3:3 + w].copy(), base[2:2 + h, 5:5 + w].copy()
        cu = np.full((h // 2, w // 2), 120, np.uint8)
        cur = synth.Picture(cy, cu, cu.copy())
        ref = synth.Picture(ry, cu.copy(), cu.copy())
        p = abi.default_params(w, h, display_width=dw, display_height=dh, pqindex=8, motion_search_range=64,
                               motion_local_search_range=16, mode=abi.MODE_MIXED_MV)
        _check_frame(backend, cur, ref, p, f"small {w}x{h} display {dw}x{dh}")


def test_batched_step_equals_single_frames(backend):
    jobs = []
    for i, (w, h) in enumerate(((320, 176), (176, 144), (352, 288))):
        d = synth.pan_rectangles(20 + i, 3, width=w, height=h, max_pan_px=8, max_obj_px=16, frame_indices=[0, 1])
        p = abi.default_params(w, h, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                               mode=abi.MODE_MIXED_MV)
        jobs.append((d[1], d[0], p))
    frames = [(backend.upload(c), backend.upload(r)) for c, r, _ in jobs]
    _say("[case] batched step of 3 streams")
    try:
        got = backend.analyze_step(frames, [p for _, _, p in jobs])
    finally:
        for a, b in frames:
            a.release()
            b.release()
    for (c, r, p), g in zip(jobs, got):
        want = oracle_analyze(p, c, r)
        assert not g.diff(want), describe_diff(g, want, p.width)


def _golden_files():
    import glob
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return sorted(glob.glob(os.path.join(here, "*.npz")))


@pytest.mark.parametrize("path", _golden_files(), ids=lambda p: p.rsplit("/", 1)[-1][:-4])
def test_gpu_matches_reference_fixture(backend, path):
    """The GPU result against outputs written by the compiled reference itself."""
    import json
    z = np.load(path, allow_pickle=False)
    cur = synth.Picture(z["cur_y"], z["cur_u"], z["cur_v"])
    ref = synth.Picture(z["ref_y"], z["ref_u"], z["ref_v"])
    h, w = cur.y.shape
    p = abi.default_params(w, h, **json.loads(str(z["params"])))
    _say(f"[case] golden {path.rsplit('/', 1)[-1]}")
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        got = backend.analyze(fc, fr, p)
    finally:
        fc.release()
        fr.release()
    want = abi.ResultArrays(w, h)
    for k in abi.ResultArrays.ARRAYS:
        getattr(want, k)[...] = z[f"out_{k}"]
    for k, v in json.loads(str(z["scalars"])).items():
        setattr(want.struct, k, v)
    assert not got.diff(want), describe_diff(got, want, w)


def test_work_counters_match_oracle_op_counts(backend):
    """The kernel's algorithmic op counters equal the oracle's count of the
    reference algorithm's operations on the same inputs."""
    from _oracle import OracleOps
    d = synth.pan_rectangles(31, 3, width=320, height=176, max_pan_px=8, max_obj_px=24, frame_indices=[0, 2])
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    ops = OracleOps()
    oracle_analyze(p, d[2], d[0], ops)
    fc, fr = backend.upload(d[2]), backend.upload(d[0])
    try:
        backend.analyze(fc, fr, p)
        w = backend.work()
    finally:
        fc.release()
        fr.release()
    assert (w.sad_px, w.interp_taps, w.satd_ops, w.sig_ops, w.frame_px) == \
        (ops.sad_px, ops.interp_taps, ops.satd_ops, ops.sig_ops, ops.frame_px)


def test_multi_stream_step(backend):
    """One analysis step over two streams of different rasters agrees with the
    oracle on each."""
    jobs = []
    for i, (w, h) in enumerate(((640, 360), (352, 288))):
        d = synth.pan_rectangles(40 + i, 4, width=w, height=h, frame_indices=[0, 3])
        p = abi.default_params(w, h, pqindex=8, motion_search_range=512, motion_local_search_range=32,
                               mode=abi.MODE_MIXED_MV)
        jobs.append((d[3], d[0], p))
    frames = [(backend.upload(c), backend.upload(r)) for c, r, _ in jobs]
    _say("[case] multi-stream step 640x360 + 352x288")
    try:
        got = backend.analyze_step(frames, [p for _, _, p in jobs])
    finally:
        for a, b in frames:
            a.release()
            b.release()
    for (c, r, p), g in zip(jobs, got):
        want = oracle_analyze(p, c, r)
        assert not g.diff(want), describe_diff(g, want, p.width)


def test_block_cost_primitive_fuzz(backend):
    """The sub-pel device primitive (interpolated SAD / SATD of 16x16 and 8x8
    blocks, shared and per-candidate windows) against the oracle's
    block_sad_motion / block_satd_motion on far, near, clustered and edge
    candidates."""
    import ctypes as C
    from _oracle import oracle
    rng = np.random.default_rng(77)
    d = synth.pan_rectangles(3, 8, width=640, height=360, frame_indices=[0, 3])
    cur, ref = d[3], d[0]
    p = abi.default_params(640, 360, pqindex=8, motion_search_range=512, motion_local_search_range=32)
    cs, rs = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        for bw in (16, 8):
            blocks = [(0, 0), (624 if bw == 16 else 632, 352), (320, 176), (16 * 7 + (8 if bw == 8 else 0), 32 + (8 if bw == 8 else 0)),
                      (0, 352), (624 if bw == 16 else 632, 0)]
            for (x0, y0) in blocks:
                for mode in (0, 2, 3):
                    for metric in (0, 1):
                        centre = rng.integers(-1600, 1600), rng.integers(-900, 900)
                        near = [(int(centre[0]) + dx, int(centre[1]) + dy) for dy in (-2, 0, 2) for dx in (-2, 0, 2)]
                        rand = [(int(a), int(b)) for a, b in zip(rng.integers(-2000, 2000, 20), rng.integers(-1000, 1000, 20))]
                        small = [(int(a), int(b)) for a, b in zip(rng.integers(-9, 10, 12), rng.integers(-9, 10, 12))]
                        cands = np.array(near + rand + small, np.int32)
                        cands[:, 0] = np.clip(cands[:, 0], -4096, 4095)
                        cands[:, 1] = np.clip(cands[:, 1], -1024, 1023)
                        if mode == 2:
                            cands &= ~1
                        _say(f"[case] block costs bw={bw} at ({x0},{y0}) mode={mode} metric={metric}")
                        got = backend.block_costs(fc, fr, p, x0, y0, bw, bw, mode, metric, cands)
                        want = np.array([oracle().oracle_block_cost(C.byref(p), C.byref(cs.struct), C.byref(rs.struct),
                                                                    x0, y0, bw, bw, int(a), int(b), mode, metric)
                                         for a, b in cands], np.uint64)
                        bad = np.nonzero(got != want)[0]
                        assert not len(bad), f"bw={bw} ({x0},{y0}) mode={mode} metric={metric}: " + \
                            ", ".join(f"{tuple(cands[i])}: {got[i]} vs {want[i]}" for i in bad[:6])
    finally:
        fc.release()
        fr.release()


def test_block_costs_on_noise_at_every_phase(backend):
    """Interpolated SADs on white noise, where the bicubic filter sums leave
    [0, 255 << shift] on both sides for many samples (the clamps of every
    phase), at every quarter-pel phase of a 25x25 candidate grid for blocks
    in the corners, on the edges and inside a 176x144 frame, bicubic and
    bilinear, against the oracle's block_sad_motion."""
    import ctypes as C
    from _oracle import oracle
    rng = np.random.default_rng(4242)
    w, h = 176, 144
    cur = synth.Picture(rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
                      rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))
    ref = synth.Picture(rng.integers(0, 256, (h, w), dtype=np.uint8), rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8),
                      rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8))
    p = abi.default_params(w, h, pqindex=8, motion_search_range=64, motion_local_search_range=16)
    cs, rs = abi.PlaneSet.of(cur), abi.PlaneSet.of(ref)
    grid = np.array([(x, y) for y in range(-12, 13) for x in range(-12, 13)], np.int32)
    fc, fr = backend.upload(cur), backend.upload(ref)
    try:
        for (x0, y0) in ((0, 0), (160, 128), (80, 64), (0, 64), (160, 0), (96, 128)):
            for mode in (0, 2):
                cands = grid & ~1 if mode == 2 else grid
                _say(f"[case] block costs on noise at ({x0},{y0}) mode={mode}")
                got = backend.block_costs(fc, fr, p, x0, y0, 16, 16, mode, 0, cands)
                want = np.array([oracle().oracle_block_cost(C.byref(p), C.byref(cs.struct), C.byref(rs.struct),
                                                            x0, y0, 16, 16, int(a), int(b), mode, 0)
                                 for a, b in cands], np.uint64)
                bad = np.nonzero(got != want)[0]
                assert not len(bad), f"({x0},{y0}) mode={mode}: " + \
                    ", ".join(f"{tuple(cands[i])}: {got[i]} vs {want[i]}" for i in bad[:6])
    finally:
        fc.release()
        fr.release()


@pytest.mark.parametrize("mode,rng_", [(abi.MODE_ONE_MV_QPEL, 512), (abi.MODE_MIXED_MV, 32), (abi.MODE_MIXED_MV, 128)],
                         ids=["qpel_r512", "mixed_r32", "mixed_r128"])
def test_pan_variants(backend, mode, rng_):
    d = synth.pan_rectangles(3, 8, width=640, height=360, frame_indices=[0, 3])
    p = abi.default_params(640, 360, pqindex=8, motion_search_range=rng_, motion_local_search_range=32, mode=mode)
    _check_frame(backend, d[3], d[0], p, f"pan variant mode={mode} R={rng_}")


def test_graph_replay_matches_eager(backend):
    """Steps replayed from a captured HIP graph give the launch-by-launch
    results, over several replays with other steps in between."""
    shapes = ((352, 288), (320, 176))
    sets = []
    for k in range(3):
        jobs = []
        for i, (w, h) in enumerate(shapes):
            d = synth.pan_rectangles(80 + 10 * k + i, 3, width=w, height=h, frame_indices=[0, 2])
            p = abi.default_params(w, h, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                                   mode=abi.MODE_MIXED_MV)
            jobs.append((d[2], d[0], p))
        sets.append(jobs)
    up = [[(backend.upload(c), backend.upload(r)) for c, r, _ in s] for s in sets]
    params = [p for _, _, p in sets[0]]
    try:
        eager = []
        for k in range(3):
            _say(f"[case] eager step {k}")
            eager.append(backend.analyze_step(up[k], params))
        backend.set_graphs(True)
        flags = []
        for k in (0, 1, 2, 1, 0):
            _say(f"[case] graph step on set {k}")
            got = backend.analyze_step(up[k], params)
            flags.append(backend.timing().graph_replay)
            for g, e, p in zip(got, eager[k], params):
                assert not g.diff(e), describe_diff(g, e, p.width)
            backend.set_graphs(False)
            backend.analyze_step(up[(k + 1) % 3], params, want_results=False)   # eager work in between
            backend.set_graphs(True)
        assert flags[0] == 2 and all(f == 1 for f in flags[1:]), flags
    finally:
        backend.set_graphs(False)
        for s in up:
            for a, b in s:
                a.release()
                b.release()
    for (c, r, p), e in zip(sets[1], eager[1]):
        want = oracle_analyze(p, c, r)
        assert not e.diff(want), describe_diff(e, want, p.width)


def test_batched_streams_runner_matches_oracle(backend):
    """Streams of uneven raster and length through the batching runner (steps
    of 3 pairs; the anti-diagonal sweep from a graph for one raster, the
    relaxation schedule for the other) give each pair the oracle's result."""
    from libvc1_amd import batching
    specs = [(300, 176, 144, 4), (301, 320, 176, 3), (302, 176, 144, 2)]
    streams, pics = [], []
    for seed, w, h, n in specs:
        d = synth.pan_rectangles(seed, n, width=w, height=h, max_pan_px=8, max_obj_px=16)
        p = batching.stream_params(w, h)
        pics.append([d[k] for k in range(n)])
        streams.append(batching.Stream(w, h, [backend.upload(x) for x in pics[-1]], p, seed))
    # one raster as a table written before the schedule knob (sweep + graphs),
    # the other under the relaxation schedule
    table = {"settings": {"176x144": {"jobs_per_step": 3, "graphs": True},
                          "320x176": {"jobs_per_step": 3, "schedule": "relax", "graphs": False}}}
    _say("[case] batching runner, 3 streams")
    try:
        out = batching.run(backend, streams, table, collect=True)
    finally:
        for s in streams:
            for f in s.frames:
                f.release()
    assert out["macroblocks"] == sum((n - 1) * ((w + 15) // 16) * ((h + 15) // 16) for _, w, h, n in specs)
    for si, s in enumerate(streams):
        for k, (c, r) in enumerate(s.pairs):
            want = oracle_analyze(s.params, pics[si][c], pics[si][r])
            got = out["results"][(si, k)]
            assert not got.diff(want), describe_diff(got, want, s.width)


def test_runner_with_steps_in_flight_matches_oracle(backend):
    """Steps kept in flight on three contexts (host threads, one stream each)
    over frames uploaded by the first: every pair equals the oracle's, under
    the captured relaxation step and the sweep."""
    from libvc1_amd import Backend, batching
    specs = [(310, 176, 144, 6), (311, 320, 176, 5)]
    streams, pics = [], []
    for seed, w, h, n in specs:
        d = synth.pan_rectangles(seed, n, width=w, height=h, max_pan_px=8, max_obj_px=16)
        pics.append([d[k] for k in range(n)])
        streams.append(batching.Stream(w, h, [backend.upload(x) for x in pics[-1]], batching.stream_params(w, h), seed))
    table = {"settings": {"176x144": {"jobs_per_step": 2, "schedule": "relax", "graphs": True, "in_flight": 3},
                          "320x176": {"jobs_per_step": 1, "schedule": "diagonal", "graphs": False, "in_flight": 2}}}
    lanes = [backend, Backend(0), Backend(0)]
    _say("[case] batching runner, steps in flight on 3 contexts")
    try:
        out = batching.run(lanes, streams, table, collect=True)
    finally:
        for lane in lanes[1:]:
            lane.close()
        for st in streams:
            for f in st.frames:
                f.release()
    for si, st in enumerate(streams):
        for k, (c, r) in enumerate(st.pairs):
            want = oracle_analyze(st.params, pics[si][c], pics[si][r])
            got = out["results"][(si, k)]
            assert not got.diff(want), describe_diff(got, want, st.width)
