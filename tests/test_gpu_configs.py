"""The benchmark configurations at their full sizes on the GPU, through the C
ABI (cases in tests/_config_cases.py):

  cfg2  1080p IBBP P pairs of the benchmark sequence, R 512: pair (3, 0)
        against the CPU oracle field by field under both macroblock
        schedules, and both pairs against digests written by the compiled
        reference
  cfg3  one stream per raster of the 64-stream workload (720x480 to
        3840x2160, R 256) through the batching runner, every pair against
        reference-written digests
  cfg4  learned seeding at 1080p on two held-out sequences: at most 1% rise
        of the summed SAD + lambda*R of the final 1-MV vectors against the
        plain search, and what the trained network gains over the same
        seeded search without it
  cfg5  one GPU's share (4 of 16 streams) of the 2160p hard-motion workload
        (200 px/frame, cuts, duplicates, R 1024) as one learned and batched
        step: the plain step's pairs across a cut, on a duplicate and after
        the cut against reference-written digests, and the learned step
        within 1% of the plain cost
"""
import json
import os

import numpy as np
import pytest

import _config_cases as cc
from libvc1_amd import abi, batching, learned
from libvc1_amd.backend import ProposalNet
from _oracle import describe_diff, oracle_analyze

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WORKERS = 4


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(HERE, "golden", "config_digests.json")) as f:
        return json.load(f)


def _expect(digests, key, res):
    want = digests[key]
    got = cc.digest(res)
    assert got == want["sha256"], f"{key}: digest differs; scalars {res.scalars()} vs reference {want['scalars']}"


# ---------------------------------------------------------------- cfg2

def test_cfg2_1080p_pairs(backend, digests):
    d = cc.cfg2_frames()
    p = cc.cfg2_params()
    frames = {k: backend.upload(v) for k, v in d.items()}
    try:
        jobs = [(frames[c], frames[r]) for c, r in cc.CFG2_PAIRS]
        for sched, name in ((abi.SCHEDULE_RELAX, "relax"), (abi.SCHEDULE_DIAGONAL, "diagonal")):
            _say(f"[case] cfg2 1920x1080 R=512 pairs {cc.CFG2_PAIRS} schedule {name}")
            backend.set_schedule(sched)
            try:
                res = backend.analyze_step(jobs, [p] * len(jobs))
            finally:
                backend.set_schedule(abi.SCHEDULE_RELAX)
            for (c, r), g in zip(cc.CFG2_PAIRS, res):
                _expect(digests, f"cfg2/{c}-{r}", g)
            if sched == abi.SCHEDULE_RELAX:
                first = res[0]
        c, r = cc.CFG2_PAIRS[0]
        want = oracle_analyze(p, d[c], d[r])
        assert not first.diff(want), describe_diff(first, want, p.width)
        assert first.use_4mv.any() and first.struct.global_dx != 0
    finally:
        for f in frames.values():
            f.release()


# ---------------------------------------------------------------- cfg3

def test_cfg3_rasters_through_the_runner(backend, digests):
    gen = cc.generate_parallel(cc.cfg3_stream, list(range(len(cc.CFG3_RASTERS))), WORKERS)
    streams = [batching.Stream(w, h, [backend.upload(pic) for pic in pics], cc.cfg3_params(w, h), k)
               for k, ((w, h), pics) in enumerate(zip(cc.CFG3_RASTERS, gen))]
    _say(f"[case] cfg3 runner over {[f'{s.width}x{s.height}' for s in streams]} R=256")
    try:
        out = batching.run(backend, streams, collect=True)
        for si, s in enumerate(streams):
            for k, (c, r) in enumerate(s.pairs):
                _expect(digests, f"cfg3/{s.width}x{s.height}/{c}-{r}", out["results"][(si, k)])
    finally:
        for s in streams:
            for f in s.frames:
                f.release()


# ---------------------------------------------------------------- cfg4

def _seed_everywhere_at_global(mbw, mbh, n):
    """Proposals of the seeded search without a network: both candidates on
    the global vector and every macroblock confident."""
    return [np.concatenate([np.zeros((mbh, mbw, 4), np.float32), np.full((mbh, mbw, 1), 30.0, np.float32)], axis=2)
            for _ in range(n)]


def test_cfg4_learned_1080p_held_out(backend):
    seqs = cc.generate_parallel(cc.cfg4_frames, cc.CFG4_SEEDS, WORKERS)
    p = cc.cfg2_params()
    mbw, mbh = abi.mb_grid(p.width, p.height)
    uploaded = [{k: backend.upload(v) for k, v in d.items()} for d in seqs]
    jobs = [(fr[c], fr[r]) for fr in uploaded for c, r in cc.CFG4_PAIRS]
    plist = [p] * len(jobs)
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    _say(f"[case] cfg4 learned 1080p held-out seeds {cc.CFG4_SEEDS}, {len(jobs)} pairs")
    try:
        backend.set_cost_report(True)
        try:
            backend.analyze_step(jobs, plist, want_results=False)
            plain_work = backend.work()
            c_plain = np.concatenate([backend.last_mb_costs(j, mbw * mbh)[0] for j in range(len(jobs))])
            backend.analyze_step_learned(net, jobs, plist, 0.5, want_results=False)
            learned_work = backend.work()
            c_learn = np.concatenate([backend.last_mb_costs(j, mbw * mbh)[0] for j in range(len(jobs))])
            backend.analyze_step_proposals(jobs, plist, _seed_everywhere_at_global(mbw, mbh, len(jobs)), 0.5,
                                           want_results=False)
            c_nonet = np.concatenate([backend.last_mb_costs(j, mbw * mbh)[0] for j in range(len(jobs))])
        finally:
            backend.set_cost_report(False)
    finally:
        net.close()
        for fr in uploaded:
            for f in fr.values():
                f.release()
    cp, cl, cn = (x.astype(np.float64).sum() for x in (c_plain, c_learn, c_nonet))
    rise, rise_nonet = cl / cp - 1.0, cn / cp - 1.0
    seeded = learned_work.seeded_macroblocks / c_plain.size
    _say(f"   learned cost rise {rise:+.4%} (seeded {seeded:.1%}, worse {np.mean(c_learn > c_plain):.3%}); "
         f"without the network {rise_nonet:+.4%}; integer SAD work {learned_work.sad_px / plain_work.sad_px:.3f}x")
    assert rise <= 0.01
    # the network's gain: most macroblocks skip the plain integer search...
    assert seeded >= 0.3
    assert learned_work.sad_px < plain_work.sad_px
    # ...at a lower cost than the same seeded search driven without it
    assert cl < cn and rise_nonet - rise >= 0.001


# ---------------------------------------------------------------- cfg5

def test_cfg5_2160p_hard_motion_share(backend, digests):
    gen = cc.generate_parallel(cc.cfg5_stream, list(range(cc.CFG5_STREAMS)), WORKERS)
    p = cc.cfg5_params()
    uploaded, jobs, keys = [], [], []
    for seed, pairs in gen:
        for label, (c, r) in pairs.items():
            fc, fr = backend.upload(c), backend.upload(r)
            uploaded += [fc, fr]
            jobs.append((fc, fr))
            keys.append(f"cfg5/{seed}/{label}")
    plist = [p] * len(jobs)
    _say(f"[case] cfg5 2160p R=1024, {cc.CFG5_STREAMS} streams x {len(jobs) // cc.CFG5_STREAMS} pairs in one step")
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    try:
        res = backend.analyze_step(jobs, plist)
        checked = 0
        for key, r in zip(keys, res):
            if key in digests:
                _expect(digests, key, r)
                checked += 1
            if key.endswith("duplicate"):
                assert not r.mvs.any() and not r.screen_sads.any()
        assert checked >= 3
        cmp = learned.cost_compare(backend, net, jobs, plist, 0.5)
    finally:
        net.close()
        for f in uploaded:
            f.release()
    _say(f"   learned vs plain: {cmp}")
    assert cmp["macroblocks"] == len(jobs) * 240 * 135
    assert cmp["cost_rise"] <= 0.01
