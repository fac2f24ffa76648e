"""Macroblock schedules: the relaxation schedule (every macroblock at once,
then worklist passes until nothing changes) against the anti-diagonal sweep
and the CPU oracle, bit for bit, through the C ABI.

The reference analyses macroblocks one by one in raster order
(src/encoder_motion.cpp:113-428); both schedules must give that result.
"""
import os

import numpy as np
import pytest

from libvc1_amd import abi, synth
from _oracle import describe_diff, oracle_analyze

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(backend, jobs, schedule):
    backend.set_schedule(schedule)
    frames = [(backend.upload(c), backend.upload(r)) for c, r, _ in jobs]
    try:
        got = backend.analyze_step(frames, [p for _, _, p in jobs])
        t = backend.timing()
    finally:
        for a, b in frames:
            a.release()
            b.release()
        backend.set_schedule(abi.SCHEDULE_RELAX)
    return got, t


def _cases():
    out = []
    d = synth.pan_rectangles(3, 8, width=640, height=360, frame_indices=[0, 3])
    out.append(("pan 640x360 R512", d[3], d[0], abi.default_params(
        640, 360, pqindex=8, motion_search_range=512, motion_local_search_range=32, mode=abi.MODE_MIXED_MV)))
    seq = synth.cif_translation(1, frames=3)
    out.append(("cif mixed R64", seq[1], seq[0], abi.default_params(
        352, 288, pqindex=8, motion_search_range=64, motion_local_search_range=64, mode=abi.MODE_MIXED_MV)))
    d = synth.pan_rectangles(9, 3, width=320, height=176, max_pan_px=8, max_obj_px=24, frame_indices=[0, 2])
    for label, over in (("diamond", dict(me_algorithm=1)), ("hpel bilinear", dict(mode=abi.MODE_ONE_MV_HPEL_BILINEAR)),
                        ("zero range", dict(motion_search_range=0)), ("no intra", dict(allow_intra=0)),
                        ("long-range mode 2", dict(long_range_search_mode=2))):
        kw = dict(pqindex=8, motion_search_range=256, motion_local_search_range=32, mode=abi.MODE_MIXED_MV)
        kw.update(over)
        out.append((label, d[2], d[0], abi.default_params(320, 176, **kw)))
    return out


@pytest.mark.parametrize("case", range(7), ids=lambda i: f"case{i}")
def test_relax_equals_diagonal_and_oracle(backend, case):
    label, cur, ref, p = _cases()[case]
    _say(f"[case] schedules {label} {p.width}x{p.height}")
    want = oracle_analyze(p, cur, ref)
    relax, tr = _run(backend, [(cur, ref, p)], abi.SCHEDULE_RELAX)
    diag, td = _run(backend, [(cur, ref, p)], abi.SCHEDULE_DIAGONAL)
    assert not relax[0].diff(want), f"relax {label}\n" + describe_diff(relax[0], want, p.width)
    assert not diag[0].diff(want), f"diagonal {label}\n" + describe_diff(diag[0], want, p.width)
    mbs = relax[0].mbs
    assert tr.schedule == abi.SCHEDULE_RELAX and tr.relax_passes >= 1 and tr.relax_analyses >= mbs
    assert td.schedule == abi.SCHEDULE_DIAGONAL and td.relax_passes == 0
    _say(f"   relax: {tr.relax_passes} passes, {tr.relax_analyses} analyses for {mbs} macroblocks")


def test_relax_multi_raster_step_and_counters(backend):
    """A step of three streams of different rasters: identical results and
    identical op counters under both schedules (the counters keep only each
    macroblock's final analysis)."""
    jobs = []
    for i, (w, h) in enumerate(((640, 360), (176, 144), (352, 288))):
        d = synth.pan_rectangles(50 + i, 4, width=w, height=h, frame_indices=[0, 3])
        jobs.append((d[3], d[0], abi.default_params(w, h, pqindex=8, motion_search_range=512,
                                                    motion_local_search_range=32, mode=abi.MODE_MIXED_MV)))
    _say("[case] relax multi-raster step")
    relax, _ = _run(backend, jobs, abi.SCHEDULE_RELAX)
    wr = backend.work()
    diag, _ = _run(backend, jobs, abi.SCHEDULE_DIAGONAL)
    wd = backend.work()
    for (c, r, p), a, b in zip(jobs, relax, diag):
        assert not a.diff(b), describe_diff(a, b, p.width)
    assert (wr.sad_px, wr.interp_taps, wr.satd_ops, wr.sig_ops, wr.frame_px) == \
        (wd.sad_px, wd.interp_taps, wd.satd_ops, wd.sig_ops, wd.frame_px)
    c, r, p = jobs[0]
    want = oracle_analyze(p, c, r)
    assert not relax[0].diff(want), describe_diff(relax[0], want, p.width)


def test_relax_pass_limit_falls_back_to_the_sweep(backend):
    """With the pass limit at 1 the relaxation cannot settle a textured pan;
    the step finishes with the diagonal sweep and stays exact."""
    d = synth.pan_rectangles(5, 4, width=320, height=176, frame_indices=[0, 3])
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    want = oracle_analyze(p, d[3], d[0])
    _say("[case] relax pass limit 1")
    os.environ["VC1ME_RELAX_MAX_PASSES"] = "1"
    try:
        got, t = _run(backend, [(d[3], d[0], p)], abi.SCHEDULE_RELAX)
    finally:
        del os.environ["VC1ME_RELAX_MAX_PASSES"]
    assert not got[0].diff(want), describe_diff(got[0], want, p.width)
    assert t.relax_passes == 1
    assert t.schedule == abi.SCHEDULE_DIAGONAL    # the sweep finished the step


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graph"])
def test_relax_continues_when_the_launched_passes_do_not_settle(graphs):
    """A new launch shape starts with a fixed number of device-driven passes
    (here forced to 2); when the last launched pass still changed something,
    the step launches more passes until it settles, exact either way, and
    the shape's later steps start with enough passes."""
    from libvc1_amd import Backend
    d = synth.pan_rectangles(6, 4, width=320, height=176, frame_indices=[0, 3])
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    want = oracle_analyze(p, d[3], d[0])
    _say(f"[case] relax with 2 first passes, graphs {graphs}")
    os.environ["VC1ME_RELAX_FIRST_PASSES"] = "2"
    os.environ["VC1ME_RELAX_PASS0_PHASES"] = "3"
    os.environ["VC1ME_RELAX_GUESS_PHASES"] = "2"
    be = Backend(0)       # a fresh context: no pass count known for the shape
    try:
        be.set_graphs(graphs)
        fc, fr = be.upload(d[3]), be.upload(d[0])
        try:
            first = be.analyze_step([(fc, fr)], [p])[0]
            t1 = be.timing()
            again = be.analyze_step([(fc, fr)], [p])[0]
            t2 = be.timing()
        finally:
            fc.release()
            fr.release()
    finally:
        del os.environ["VC1ME_RELAX_FIRST_PASSES"]
        del os.environ["VC1ME_RELAX_PASS0_PHASES"]
        del os.environ["VC1ME_RELAX_GUESS_PHASES"]
        be.close()
    assert t1.schedule == abi.SCHEDULE_RELAX and t1.relax_passes > 2
    assert t2.relax_passes == t1.relax_passes
    # in the graph: the guess pass as two launches by rows, pass 0 as three,
    # then the other passes the shape needed plus two
    assert t2.mb_kernel_launches == 2 + t1.relax_passes + 2 + 2 if graphs else True
    for got in (first, again):
        assert not got.diff(want), describe_diff(got, want, p.width)
    if graphs:
        assert t1.graph_replay == 2 and t2.graph_replay == 2    # recaptured with more passes


@pytest.mark.parametrize("phases,guess_phases", [(1, 1), (2, 3), (5, 2), (16, 16)])
def test_relax_pass0_by_rows_is_exact(phases, guess_phases):
    """Pass 0 launched by rows (launch r analyses the rows y with
    y % phases == r; a changed macroblock queues its bottom dependents only
    when their row came first) relaxes to the oracle's field for every phase
    count, including more phases than the picture has macroblock rows, and
    whatever the guess pass's own launches by rows left as starting state."""
    from libvc1_amd import Backend
    d = synth.pan_rectangles(7, 4, width=320, height=176, frame_indices=[0, 3])
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=256, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    want = oracle_analyze(p, d[3], d[0])
    _say(f"[case] relax, pass 0 in {phases} launches by rows, guess in {guess_phases}")
    os.environ["VC1ME_RELAX_PASS0_PHASES"] = str(phases)
    os.environ["VC1ME_RELAX_GUESS_PHASES"] = str(guess_phases)
    be = Backend(0)
    try:
        fc, fr = be.upload(d[3]), be.upload(d[0])
        try:
            got = be.analyze_step([(fc, fr)], [p])[0]
            t = be.timing()
        finally:
            fc.release()
            fr.release()
    finally:
        del os.environ["VC1ME_RELAX_PASS0_PHASES"]
        del os.environ["VC1ME_RELAX_GUESS_PHASES"]
        be.close()
    assert t.schedule == abi.SCHEDULE_RELAX
    assert not got.diff(want), describe_diff(got, want, p.width)


def test_relax_learned_step_equals_diagonal(backend):
    """The proposal-seeded step runs the same macroblock analysis; both
    schedules give the same seeded result."""
    from libvc1_amd import learned
    from libvc1_amd.backend import ProposalNet
    d = synth.pan_rectangles(77, 4, width=640, height=360, frame_indices=[0, 3])
    p = abi.default_params(640, 360, pqindex=8, motion_search_range=512, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    fc, fr = backend.upload(d[3]), backend.upload(d[0])
    _say("[case] learned step under both schedules")
    try:
        backend.set_schedule(abi.SCHEDULE_RELAX)
        a = backend.analyze_step_learned(net, [(fc, fr)], [p], 0.5)
        backend.set_schedule(abi.SCHEDULE_DIAGONAL)
        b = backend.analyze_step_learned(net, [(fc, fr)], [p], 0.5)
    finally:
        backend.set_schedule(abi.SCHEDULE_RELAX)
        fc.release()
        fr.release()
        net.close()
    assert not a[0].diff(b[0]), describe_diff(a[0], b[0], p.width)


def test_set_schedule_rejects_unknown(backend):
    from libvc1_amd.backend import BackendError
    with pytest.raises(BackendError):
        backend.set_schedule(7)


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graph"])
def test_relax_step_of_255_jobs(backend, graphs):
    """The largest step (255 jobs) under the relaxation schedule: the
    work-queue passes carry entries of jobs 128..254 (the entry's top bit
    set), and every job equals the oracle (five distinct pairs, repeated)."""
    pics = [synth.pan_rectangles(400 + k, 3, width=96, height=64, max_pan_px=6, max_obj_px=20, frame_indices=[0, 2])
            for k in range(5)]
    p = abi.default_params(96, 64, pqindex=8, motion_search_range=128, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV)
    want = [oracle_analyze(p, d[2], d[0]) for d in pics]
    frames = [(backend.upload(d[2]), backend.upload(d[0])) for d in pics]
    _say(f"[case] relaxation, 255 jobs of 96x64, graphs {graphs}")
    backend.set_graphs(graphs)
    try:
        for _ in range(2 if graphs else 1):   # capture, then replay
            got = backend.analyze_step([frames[j % 5] for j in range(255)], [p] * 255)
            passes = backend.relax_passes()
            assert backend.timing().schedule == abi.SCHEDULE_RELAX
            assert len(passes) > 1 and passes[1][0] > 0, passes     # work-queue passes ran
            for j, g in enumerate(got):
                assert not g.diff(want[j % 5]), (j, describe_diff(g, want[j % 5], 96))
    finally:
        backend.set_graphs(False)
        for a, b in frames:
            a.release()
            b.release()
