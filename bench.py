#!/usr/bin/env python3
"""Benchmark: 1080p P-frame motion estimation, MB searches per second.

Workload (the reference's 1080p IBBP configuration): 1920x1080 8-bit 4:2:0,
60 frames of blurred-noise background under a camera pan of up to 16 px/frame
with eight textured rectangles moving at up to 48 px/frame (quarter-pel),
noise sigma 2; IBBP with GOP 15, so the P pictures are frames 3, 6, 9, 12 of
each GOP, each predicted from the anchor three frames back: 16 P frames,
16 x 8160 = 130,560 macroblocks per step.  ME parameters: search range 512
(extended MV range, long-range search on), local range 32, Mixed-MV (16x16 +
4-MV 8x8), quarter-pel, SATD finalists, PQINDEX 8, P-intra decision on.

One step = vc1me_analyze_step over all 16 (cur, ref) pairs with the frames
already resident in HBM, under the default relaxation schedule (a guess
pass, every macroblock at once, then worklist passes until nothing changes;
bit-exact with the raster-order reference), captured as one HIP graph; the
MV fields stay on the device (their download is timed separately and
reported as `value_with_download`).  `value` keeps several steps in flight
(`steps_in_flight`: one context with its own stream and device state per
host thread, all on the same resident frames), so that the latency-bound end
of one step overlaps the next; their number is tuned on this workload before
the timed steps (two rounds of K steps at each of 1, 2 and 3,
`schedule.in_flight_trial`)
unless --in-flight sets it; `value_one_in_flight` is the same steps one at a
time.  The same steps under the anti-diagonal sweep
are timed for comparison (`schedule.diagonal_value`).

The learned-seeded form (`learned`) runs the same steps through
vc1me_analyze_step_learned with the trained weights of
libvc1_amd/weights/proposal_net.bin (network forward inside the timed step)
and is judged by the per-macroblock cost report: the summed SAD + lambda*R of
the final 1-MV vectors against the plain form's, the fraction of macroblocks
that got worse and the mean excess.  Two ablations take the network's
contribution away (every macroblock seeded from the global vector alone; the
network's confidence gate kept but its candidates removed).

The seeded leg (`seeded_core`) runs the seeded mode passes of the
reference's P-picture analysis on the same pairs: the unseeded 1-MV
quarter-pel seed pass, the three seeded mode candidates as one 48-job
vc1me_analyze_step_seeded (its MB/s is `value`), and the final seeded
Mixed-MV pass with intra.

The B-picture leg (`b_pictures`) runs the B pictures between each GOP's I
picture and its first P anchor (frames g+1 and g+2 between g and g+3, 8
pictures) as one vc1me_analyze_b_step, and one by one through
vc1me_analyze_b_frame, timed per picture.

The batched leg (`batched`) runs the 64-stream workload of uneven rasters
and lengths (lengths scaled to keep the run short) through the batching
runner with the tuned table; it is reported beside the headline, not as it.

Also reported: the roofline of the dominant kernels (the macroblock
analysis: k_mb_guess, then k_mb_relax0 and k_mb_relaxq, one launch per
relaxation pass) from HIP events recorded around each of their launches in
the same steps run one at a time, and the reference's own multithreaded CPU
path timed on this host on a bounded sample of the same frames.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--in-flight N] [--no-check] [--no-cpu] [--no-learned]
                    [--no-compare] [--no-seeded] [--no-bframes] [--no-batched] [--batched-scale S]
    python bench.py --copy-calibration     # 256 MiB device copy, for HBM counter scaling

`roofline.traffic` (HBM bytes per launch of the macroblock kernel) comes from
the FETCH_SIZE / WRITE_SIZE counter runs summarised in
profiles/<kernel>_1080p_ibbp16_pmc.json, each scaled by what the counter reads
for the calibration copy; it is null when that summary is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "1080p P-frame ME MBs/s: plain vs learned-seeded at ≤1% SAD+λR cost rise"
SEED = 20241203
WIDTH, HEIGHT, FRAMES = 1920, 1080, 60


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_params():
    from libvc1_amd import abi
    return abi.default_params(WIDTH, HEIGHT, pqindex=8, motion_search_range=512, motion_local_search_range=32,
                              mode=abi.MODE_MIXED_MV, me_quality=2, allow_intra=1)


def make_frames():
    from libvc1_amd import synth
    pairs = synth.ibbp_p_pairs(FRAMES, gop=15, spacing=3)
    # the P pairs' frames, plus the two B pictures after each GOP's I picture
    needed = sorted({i for p in pairs for i in p} | {1, 2, 16, 17, 31, 32, 46, 47})
    t0 = time.time()
    pics = synth.pan_rectangles(SEED, FRAMES, width=WIDTH, height=HEIGHT, max_pan_px=16, max_obj_px=48,
                                frame_indices=needed)
    log(f"generated {len(needed)} frames in {time.time() - t0:.1f}s")
    return pairs, pics


def host_description() -> dict:
    """The host the CPU baseline runs on: logical CPUs the process may use,
    logical CPUs of the machine, physical cores, the CPU model and the SIMD
    extensions the reference's auto-selection can pick from."""
    out = {"nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count(),
           "logical_cpus": os.cpu_count()}
    try:
        info = open("/proc/cpuinfo").read()
        model = next((ln.split(":", 1)[1].strip() for ln in info.splitlines() if ln.startswith("model name")), None)
        flags = next((ln.split(":", 1)[1].split() for ln in info.splitlines() if ln.startswith("flags")), [])
        cores = {(blk.split("physical id")[1].split("\n")[0], blk.split("core id")[1].split("\n")[0])
                 for blk in info.split("\n\n") if "core id" in blk and "physical id" in blk}
        out["model"] = model
        out["physical_cores"] = len(cores) or None
        out["isa"] = [f for f in ("sse4_2", "avx2", "fma", "bmi2", "avx512f", "avx512bw", "avx512vl", "avx512_vnni")
                      if f in flags]
    except OSError:
        pass
    return out


def cpu_baseline(pairs, pics, params):
    """Reference CPU path (oracle/_ref) on `threads` independent pairs, one
    whole frame pair per thread, SIMD tier auto-selected by the reference."""
    import ctypes as C

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from libvc1_amd import abi
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libvc1ref.so")
    host = host_description()
    threads = max(1, min(16, host.get("nproc") or 1))
    n = min(len(pairs), threads)
    curs = [abi.PlaneSet.of(pics[c]) for c, _ in pairs[:n]]
    refs = [abi.PlaneSet.of(pics[r]) for _, r in pairs[:n]]
    ca = (abi.Planes * n)(*[p.struct for p in curs])
    ra = (abi.Planes * n)(*[p.struct for p in refs])
    mbs = ((WIDTH + 15) // 16) * ((HEIGHT + 15) // 16)
    if os.path.exists(ref_so):
        lib = C.CDLL(ref_so)
        lib.ref_time_p_frames.argtypes = [C.POINTER(abi.Params), C.c_int, C.POINTER(abi.Planes),
                                          C.POINTER(abi.Planes), C.c_int, C.c_int]
        lib.ref_time_p_frames.restype = C.c_double
        secs = lib.ref_time_p_frames(C.byref(params), n, ca, ra, threads, 1)
        secs1 = lib.ref_time_p_frames(C.byref(params), 1, ca, ra, 1, 1)
        if secs > 0:
            out = {"value": round(n * mbs / secs, 1), "unit": "MB/s", "cores": threads, "kind": "reference",
                   "threads": threads, "host": host,
                   "sample": f"{n} of the 16 1080p P-frame pairs, one pair per thread on {threads} threads "
                             f"(of {host.get('nproc')} usable logical CPUs), reference SIMD auto-selected",
                   "seconds": round(secs, 3),
                   "value_1_thread": round(mbs / secs1, 1) if secs1 > 0 else None}
            if all(k in pics for k in (0, 1, 3)):
                # one B picture (frame 1 between anchors 0 and 3, 1-MV qpel) on one thread
                lib.ref_analyze_p_core.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.Planes),
                                                   C.POINTER(abi.Planes), C.POINTER(abi.PResult), C.c_int]
                lib.ref_time_b_frame.argtypes = [C.POINTER(abi.Params), C.POINTER(abi.Planes), C.POINTER(abi.Planes),
                                                 C.POINTER(abi.Planes), C.POINTER(abi.BParams)]
                lib.ref_time_b_frame.restype = C.c_double
                p0, p1, p3 = (abi.PlaneSet.of(pics[k]) for k in (0, 1, 3))
                anchor = abi.ResultArrays(WIDTH, HEIGHT)
                ap = abi.default_params(WIDTH, HEIGHT, pqindex=params.pqindex,
                                        motion_search_range=params.motion_search_range,
                                        motion_local_search_range=params.motion_local_search_range,
                                        mode=abi.MODE_MIXED_MV, allow_intra=1)
                lib.ref_analyze_p_core(C.byref(ap), C.byref(p3.struct), C.byref(p0.struct), C.byref(anchor.struct), 1)
                bp = abi.default_params(WIDTH, HEIGHT, pqindex=params.pqindex,
                                        motion_search_range=params.motion_search_range,
                                        motion_local_search_range=params.motion_local_search_range,
                                        mode=abi.MODE_ONE_MV_QPEL)
                binp = abi.BInputs(anchor.mvs, anchor.use_4mv, fraction=(1, 3))
                bs = lib.ref_time_b_frame(C.byref(bp), C.byref(p1.struct), C.byref(p0.struct), C.byref(p3.struct),
                                          C.byref(binp.struct))
                if bs > 0:
                    out["b_picture_ms_1_thread"] = round(bs * 1e3, 2)
            return out
    # the C restatement, single thread, on one pair
    from _oracle import oracle_analyze
    t0 = time.time()
    oracle_analyze(params, pics[pairs[0][0]], pics[pairs[0][1]])
    secs = time.time() - t0
    return {"value": round(mbs / secs, 1), "unit": "MB/s", "cores": 1, "kind": "port",
            "sample": "1 of the 16 1080p P-frame pairs, scalar C oracle", "seconds": round(secs, 3)}


def measured_traffic(kernel: str):
    """HBM bytes per launch of the macroblock kernel from the committed
    counter summary of its schedule (profiles/<kernel>_1080p_ibbp16_pmc.json:
    FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the calibration copy)."""
    try:
        with open(os.path.join(ROOT, "profiles", f"{kernel}_1080p_ibbp16_pmc.json")) as f:
            d = json.load(f)
        return d[kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration() -> int:
    """A plain device-to-device copy of known size (256 MiB, three times):
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 256 << 20
    x = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
    return 0


def device_peak_ops():
    """VALU int32 lane-op peak: CUs x 64 lanes x clock."""
    cus, clk = 256, 2.4e9
    try:
        import torch
        props = torch.cuda.get_device_properties(0)
        cus = int(getattr(props, "multi_processor_count", cus)) or cus
        rate = getattr(props, "clock_rate", 0)
        if rate:
            clk = float(rate) * 1e3
    except Exception:  # plumbing only: nominal MI355X figures otherwise
        pass
    return cus, clk, cus * 64 * clk


def warm(lanes, fn, n: int = 1) -> None:
    """n steps on every context, one context after the other (run_in_flight
    deals steps to whichever context is free, so a warm-up through it can
    leave a context without its first step, graph capture included, until
    the timed steps)."""
    for ln in lanes:
        for k in range(max(1, n)):
            fn(ln, k)


def learned_leg(lanes, jobs, plist, plain_value, plain_value_one, steps, warmup, mbs_per_step, threshold=0.5,
                in_flight=0):
    """The learned-seeded form on the same step: timed like the plain form
    (network forward inside each captured step; its number of steps in
    flight tuned the same way on the contexts of `lanes`, unless `in_flight`
    fixes it), then judged by the cost report against the plain form, with
    two ablations that remove the network's contribution."""
    import numpy as np
    from libvc1_amd import abi, batching, learned
    from libvc1_amd.backend import ProposalNet

    pool = lanes
    be = pool[0]
    params = learned.load_weights()
    nets = []
    for ln in pool:
        net = ProposalNet(ln)
        net.set_params(params)
        nets.append(net)
        ln.set_graphs(True)
    idx = {id(ln): i for i, ln in enumerate(pool)}

    def learned_step(ln, k):
        ln.analyze_step_learned(nets[idx[id(ln)]], jobs, plist, threshold, want_results=False)

    # steps in flight tuned for this form as for the plain one (two rounds
    # of `steps` steps at each count, each count's better round)
    trial = {}
    if in_flight:
        lanes = pool[:in_flight]
    else:
        warm(pool[:max(batching.IN_FLIGHT_CANDIDATES)], learned_step)
        for _ in range(2):
            for k in batching.IN_FLIGHT_CANDIDATES:
                t0 = time.perf_counter()
                batching.run_in_flight(pool[:k], steps, learned_step)
                trial[k] = max(trial.get(k, 0.0), mbs_per_step * steps / (time.perf_counter() - t0))
        lanes = pool[:max(trial, key=trial.get)]
        log("learned, steps in flight, trial MB/s: " + ", ".join(f"{k}: {v:,.0f}" for k, v in trial.items()) +
            f" -> {len(lanes)}")

    def timed(fn):
        warm(lanes, fn, warmup)
        t0 = time.perf_counter()
        batching.run_in_flight(lanes, steps, fn)
        return time.perf_counter() - t0

    elapsed = timed(learned_step)
    value = mbs_per_step * steps / elapsed
    t0 = time.perf_counter()
    for k in range(steps):
        learned_step(be, k)
    value_one = mbs_per_step * steps / (time.perf_counter() - t0)
    log(f"learned: {elapsed * 1e3 / steps:.1f} ms/step, {value:,.0f} MB/s ({value / plain_value:.3f}x plain; "
        f"one in flight {value_one / plain_value_one:.3f}x)")
    for ln in pool:
        ln.set_graphs(False)
    net = nets[0]
    judged = learned.cost_compare(be, net, jobs, plist, threshold)
    log(f"learned cost: {judged}")

    # ablations: the same seeded search without the network's candidates
    # (timed one step at a time through vc1me_analyze_step_proposals, which
    # uploads the proposals from the host: compared with the plain form one
    # step at a time)
    mbw, mbh = abi.mb_grid(plist[0].width, plist[0].height)
    res = be.analyze_step(jobs, plist)
    glob = np.array([[r.struct.global_dx, r.struct.global_dy] for r in res], np.int32)
    outs = net.forward_frames(jobs, glob)
    abl = {}
    for name, props in (
            ("no_network_all_seeded", [np.concatenate([np.zeros((mbh, mbw, 4), np.float32),
                                                       np.full((mbh, mbw, 1), 30.0, np.float32)], axis=2)
                                       for _ in jobs]),
            ("network_gate_without_candidates", [np.concatenate([np.zeros((mbh, mbw, 4), np.float32),
                                                                 o[..., 4:5]], axis=2) for o in outs])):
        be.analyze_step_proposals(jobs, plist, props, threshold, want_results=False)
        t0 = time.perf_counter()
        for _ in range(steps):
            be.analyze_step_proposals(jobs, plist, props, threshold, want_results=False)
        v = mbs_per_step * steps / (time.perf_counter() - t0)
        be.set_cost_report(True)
        try:
            be.analyze_step(jobs, plist, want_results=False)
            cp = np.concatenate([be.last_mb_costs(j, mbw * mbh)[0] for j in range(len(jobs))]).astype(np.float64)
            be.analyze_step_proposals(jobs, plist, props, threshold, want_results=False)
            ca = np.concatenate([be.last_mb_costs(j, mbw * mbh)[0] for j in range(len(jobs))]).astype(np.float64)
        finally:
            be.set_cost_report(False)
        abl[name] = {"value_one_in_flight": round(v, 1), "speedup_vs_plain_one_in_flight": round(v / plain_value_one, 4),
                     "cost_rise": round(float(ca.sum() / cp.sum() - 1), 6),
                     "frac_worse": round(float(np.mean(ca > cp)), 6)}
        log(f"ablation {name}: {abl[name]}")
    # the network as a confidence gate only (vc1me_set_learned_candidates(0))
    for ln in lanes:
        ln.set_learned_candidates(False)
        ln.set_graphs(True)
    try:
        gv = mbs_per_step * steps / timed(learned_step)
        for ln in lanes:
            ln.set_graphs(False)
        gj = learned.cost_compare(be, net, jobs, plist, threshold)
    finally:
        for ln in lanes:
            ln.set_learned_candidates(True)
            ln.set_graphs(False)
    gate_only = {"value": round(gv, 1), "speedup_vs_plain": round(gv / plain_value, 4),
                 "cost_rise": round(gj["cost_rise"], 6), "frac_worse": round(gj["frac_worse"], 6),
                 "seeded_fraction": round(gj["seeded_fraction"], 4)}
    log(f"learned, gate only: {gate_only}")
    for n in nets:
        n.close()
    meta_path = learned.WEIGHTS + ".json"
    sha = json.load(open(meta_path)).get("sha256") if os.path.exists(meta_path) else None
    return {
        "value": round(value, 1), "unit": "MB/s", "speedup_vs_plain": round(value / plain_value, 4),
        "ms_per_step": round(elapsed * 1e3 / steps, 3), "steps_in_flight": len(lanes),
        "in_flight_trial": {str(k): round(v, 1) for k, v in trial.items()},
        "value_one_in_flight": round(value_one, 1),
        "speedup_vs_plain_one_in_flight": round(value_one / plain_value_one, 4), "threshold": threshold,
        "cost_rise": round(judged["cost_rise"], 6), "frac_worse": round(judged["frac_worse"], 6),
        "mean_excess_per_mb": round(judged["mean_excess"], 3), "seeded_fraction": round(judged["seeded_fraction"], 4),
        "cost": "sum over MBs of SAD + lambda*R of the final 1-MV vector (exact integers 20*C), learned vs plain",
        "weights_sha256": sha, "ablation": abl, "gate_only": gate_only,
    }


def seeded_leg(be, jobs, plist, plain_value, steps, warmup, mbs_per_step, check_pair=None):
    """The seeded mode passes of the reference's P-picture analysis
    (Vc1Encoder::analyze_p_picture, src/encoder_motion.cpp:524-571) on the
    same 16 pairs: the unseeded 1-MV quarter-pel pass without intra (the
    shared seed), the three seeded mode candidates (1-MV half-pel, 1-MV
    half-pel bilinear, Mixed-MV; no intra) as one 48-job step, and the final
    seeded pass of the chosen mode with intra (Mixed-MV here: the choice is
    the sampled transform RD, which is not part of this backend).  Reported:
    the seeded steps' MB/s, the whole multi-pass time per P frame, and the
    final pass's SAD + lambda*R against the single plain pass."""
    import numpy as np
    from libvc1_amd import abi

    base = plist[0]
    kw = dict(pqindex=base.pqindex, motion_search_range=base.motion_search_range,
              motion_local_search_range=base.motion_local_search_range, me_quality=base.me_quality)
    qpel = abi.default_params(base.width, base.height, mode=abi.MODE_ONE_MV_QPEL, allow_intra=0, **kw)
    cand_modes = (abi.MODE_ONE_MV_HPEL, abi.MODE_ONE_MV_HPEL_BILINEAR, abi.MODE_MIXED_MV)
    cand_params = [abi.default_params(base.width, base.height, mode=m, allow_intra=0, **kw) for m in cand_modes]
    final = abi.default_params(base.width, base.height, mode=abi.MODE_MIXED_MV, allow_intra=1, **kw)
    n = len(jobs)
    cjobs = [j for _ in cand_modes for j in jobs]
    cplist = [p for p in cand_params for _ in jobs]

    def seed_step():
        return be.analyze_step(jobs, [qpel] * n)

    def cand_step(seeds):
        return be.analyze_step_seeded(cjobs, cplist, seeds * len(cand_modes))

    def final_step(cands, want=False):
        return be.analyze_step_seeded(jobs, [final] * n, cands[2 * n:3 * n], want_results=want)

    seeds = seed_step()
    cands = cand_step(seeds)
    for _ in range(warmup):
        final_step(cands)
    t_seed = t_cand = t_final = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        seeds = seed_step()
        t1 = time.perf_counter()
        cands = cand_step(seeds)
        t2 = time.perf_counter()
        final_step(cands)
        t3 = time.perf_counter()
        t_seed += t1 - t0
        t_cand += t2 - t1
        t_final += t3 - t2
    cand_value = len(cand_modes) * mbs_per_step * steps / t_cand
    final_value = mbs_per_step * steps / t_final
    per_frame_ms = (t_seed + t_cand + t_final) * 1e3 / (steps * n)
    log(f"seeded mode passes: seed pass {t_seed * 1e3 / steps:.1f} ms, 3 candidates {t_cand * 1e3 / steps:.1f} ms "
        f"({cand_value:,.0f} MB/s), final {t_final * 1e3 / steps:.1f} ms ({final_value:,.0f} MB/s); "
        f"{per_frame_ms:.2f} ms per P frame")
    mbw, mbh = abi.mb_grid(base.width, base.height)
    be.set_cost_report(True)
    try:
        be.analyze_step(jobs, plist, want_results=False)
        cp = np.concatenate([be.last_mb_costs(j, mbw * mbh)[0] for j in range(n)]).astype(np.float64)
        res = final_step(cands, want=True)
        cs = np.concatenate([be.last_mb_costs(j, mbw * mbh)[0] for j in range(n)]).astype(np.float64)
    finally:
        be.set_cost_report(False)
    parity = None
    if check_pair is not None:
        from _oracle import oracle_analyze_seeded
        cur, ref = check_pair
        t1 = time.time()
        bad = cands[2 * n].diff(oracle_analyze_seeded(cand_params[2], cur, ref, seeds[0]))
        bad += res[0].diff(oracle_analyze_seeded(final, cur, ref, cands[2 * n]))
        parity = "bit-exact" if not bad else f"MISMATCH in {bad}"
        log(f"seeded passes of pair 0 vs oracle (Mixed-MV candidate, final): {parity} ({time.time() - t1:.1f}s)")
    intensity = intensity_part(be, jobs, res, check_pair)
    # the prediction pictures of the final pass (vc1me_predict_p_frame)
    be.predict_p_frame(jobs[0][1], final, res[0]).release()
    t0 = time.perf_counter()
    for (c, r), rr in zip(jobs, res):
        be.predict_p_frame(r, final, rr).release()
    pred_ms = (time.perf_counter() - t0) * 1e3 / n
    log(f"prediction pictures: {pred_ms:.3f} ms per frame (host-timed, allocation included)")
    return {"value": round(cand_value, 1), "unit": "MB/s",
            "value_final_pass": round(final_value, 1), "speedup_vs_plain": round(cand_value / plain_value, 4),
            "ms_seed_pass": round(t_seed * 1e3 / steps, 3), "ms_candidate_step": round(t_cand * 1e3 / steps, 3),
            "ms_final_pass": round(t_final * 1e3 / steps, 3), "ms_per_p_frame_all_passes": round(per_frame_ms, 3),
            "final_cost_vs_plain": round(float(cs.sum() / cp.sum() - 1), 6),
            "passes": "1-MV qpel seed (unseeded, no intra); 48-job step of seeded Hpel / HpelBilinear / Mixed-MV "
                      "candidates (no intra); final seeded Mixed-MV with intra",
            "parity": parity, "intensity": intensity, "prediction_ms_per_frame": round(pred_ms, 4)}


def b_leg(be, pics, frames, steps, check):
    """B pictures of the same IBBP workload: for each GOP's first anchor pair
    (P at frame 3 predicted from frame 0), the two B pictures at 1/3 and 2/3
    (frames 1 and 2) through vc1me_analyze_b_frame in 1-MV quarter-pel mode,
    the anchor's Mixed-MV result supplying the direct-mode vectors.  Times
    the whole B analysis (two core passes, costs, decision) per picture."""
    import numpy as np
    from libvc1_amd import abi

    base = dict(pqindex=8, motion_search_range=512, motion_local_search_range=32, me_quality=2)
    groups = [(g0, g0 + 3) for g0 in (0, 15, 30, 45) if g0 + 3 in frames and g0 + 1 in frames and g0 + 2 in frames]
    if not groups:
        return None
    anchors = {}
    for past, fut in groups:
        anchors[fut] = be.analyze(frames[fut], frames[past],
                                  abi.default_params(WIDTH, HEIGHT, mode=abi.MODE_MIXED_MV, allow_intra=1, **base))
    p = abi.default_params(WIDTH, HEIGHT, mode=abi.MODE_ONE_MV_QPEL, **base)
    work = [(past + k, past, fut, abi.BInputs(anchors[fut].mvs, anchors[fut].use_4mv, fraction=(k, 3)))
            for past, fut in groups for k in (1, 2)]
    be.analyze_b(frames[work[0][0]], frames[work[0][1]], frames[work[0][2]], p, work[0][3])
    t0 = time.perf_counter()
    n = 0
    for _ in range(max(1, steps // 2)):
        for cur, past, fut, binp in work:
            last = be.analyze_b(frames[cur], frames[past], frames[fut], p, binp)
            n += 1
    secs_single = (time.perf_counter() - t0) / n
    items = [(frames[cur], frames[past], frames[fut], p, binp) for cur, past, fut, binp in work]
    be.analyze_b_step(items)
    t0 = time.perf_counter()
    n = 0
    for _ in range(max(1, steps // 2)):
        last = be.analyze_b_step(items)[-1]
        n += len(items)
    secs = (time.perf_counter() - t0) / n
    mbs = ((WIDTH + 15) // 16) * ((HEIGHT + 15) // 16)
    parity = None
    if check:
        from _oracle import oracle_analyze_b
        cur, past, fut, binp = work[-1]
        bad = last.diff(oracle_analyze_b(p, pics[cur], pics[past], pics[fut], binp))
        parity = "bit-exact" if not bad else f"MISMATCH in {bad}"
        log(f"B picture {cur} vs oracle: {parity}")
    modes = {k: int(getattr(last.struct, k)) for k in ("direct_macroblocks", "forward_macroblocks",
                                                        "backward_macroblocks", "interpolated_macroblocks",
                                                        "intra_macroblocks")}
    log(f"B pictures: {secs * 1e3:.2f} ms per picture in one step of {len(items)} ({mbs / secs:,.0f} MB/s), "
        f"{secs_single * 1e3:.2f} ms one by one; last picture {modes}")
    return {"ms_per_b_picture": round(secs * 1e3, 3), "value": round(mbs / secs, 1), "unit": "MB/s",
            "ms_per_b_picture_one_by_one": round(secs_single * 1e3, 3), "pictures_per_step": len(items),
            "pictures": len(work), "mode": "1-MV qpel, anchor Mixed-MV, fractions 1/3 and 2/3", "parity": parity,
            "last_picture_modes": modes}


def intensity_part(be, jobs, final_results, check_pair):
    """The fade path's two kernels on the same pairs (the reference runs them
    when fade compensation is on, src/encoder_motion.cpp:573-576):
    estimate_intensity_comp from the final pass's vectors, and the
    compensated reference.  Host-timed per frame (launches, copies and the
    host fit included); the remap's bytes are read + written once."""
    import numpy as np
    from libvc1_amd import abi

    mvs = [r.mvs.copy() for r in final_results]
    ics = [be.estimate_intensity(c, r, m) for (c, r), m in zip(jobs, mvs)]
    t0 = time.perf_counter()
    for (c, r), m in zip(jobs, mvs):
        be.estimate_intensity(c, r, m)
    t_est = (time.perf_counter() - t0) / len(jobs)
    code = (1, 19, 10)
    for c, r in jobs[:2]:
        be.intensity_compensated(r, code).release()
    t0 = time.perf_counter()
    made = [be.intensity_compensated(r, code) for _, r in jobs]
    t_map = (time.perf_counter() - t0) / len(jobs)
    for f in made:
        f.release()
    w, h = jobs[0][0].width, jobs[0][0].height
    frame_bytes = w * h * 3 // 2
    parity = None
    if check_pair is not None:
        from _oracle import oracle_estimate_intensity, oracle_intensity_compensate
        cur, ref = check_pair
        g = be.intensity_compensated(jobs[0][1], code)
        try:
            got = be.download(g)
        finally:
            g.release()
        ok = ics[0] == oracle_estimate_intensity(cur, ref, mvs[0]) and \
            all(np.array_equal(a, b) for a, b in zip(got, oracle_intensity_compensate(ref, code)))
        parity = "bit-exact" if ok else "MISMATCH"
        log(f"intensity estimate and remap of pair 0 vs oracle: {parity}")
    log(f"intensity: estimate {t_est * 1e3:.3f} ms/frame, compensated reference {t_map * 1e3:.3f} ms/frame "
        f"({2 * frame_bytes / t_map / 1e9:.1f} GB/s host-timed), estimates {sorted(set(ics))}")
    return {"estimate_ms_per_frame": round(t_est * 1e3, 4), "compensate_ms_per_frame": round(t_map * 1e3, 4),
            "compensate_gbps_host_timed": round(2 * frame_bytes / t_map / 1e9, 2), "parity": parity,
            "estimates": sorted({str(x) for x in ics})}


def scorer_leg(be, pics, frames):
    """Per-call latency of the drop-in batch scorer (the compute slot's
    score_motion_batch) on the 1080p frames at the reference's batch sizes:
    9 (a sub-pel stencil), 36 (UMH finalists), 289 (the rescue's coarse grid)
    and 1089 (the dense grid), for vc1me_score_batch (host planes keyed by a
    checksum), vc1me_score_batch_gen (keyed by generation numbers),
    vc1me_score_batch_frames (resident frames) and its allocation-free form;
    host-timed per call, every call synchronous as the slot's contract is."""
    import numpy as np
    rng = np.random.default_rng(3)
    cur, ref = pics[3].y, pics[0].y
    fc, fr = frames[3], frames[0]
    forms = {
        "score_batch": lambda c: be.score_batch(cur, ref, 960, 540, 16, 16, c),
        "score_batch_gen": lambda c: be.score_batch_gen(cur, 3, ref, 0, 960, 540, 16, 16, c),
        "score_batch_frames": lambda c: be.score_batch_frames(fc, fr, 960, 540, 16, 16, c),
        "score_batch_frames_noalloc": lambda c: be.score_batch_frames(fc, fr, 960, 540, 16, 16, c, noalloc=True),
    }
    out = {}
    for n in (9, 36, 289, 1089):
        cands = (rng.integers(-32, 33, size=(n, 2)) * 4).astype(np.int32)
        row = {}
        for name, fn in forms.items():
            fn(cands)
            reps = 50 if name == "score_batch" else 200
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(cands)
            row[name] = round((time.perf_counter() - t0) / reps * 1e6, 1)
        out[str(n)] = row
    log(f"scorer: microseconds per call {out}")
    return {"us_per_call": out, "block": "16x16 at (960, 540) of the 1080p pair (3, 0), integer candidates within "
                                          "+-32 px, SAD, stride 1"}


HARD_STREAMS, HARD_SEED = 16, 9_500_000
HARD_STEPS = (("mid-scene", 10), ("scene cut", 30))


def _hard_stream(seed):
    from libvc1_amd import synth
    need = sorted({t - k for _, t in HARD_STEPS for k in (0, 1)})
    d = synth.pan_rectangles(seed, 120, width=3840, height=2160, max_pan_px=200, max_obj_px=200, cut_every=30,
                             dup_fraction=0.05, frame_indices=need)
    return seed, d


def hard_frames():
    """The 2160p hard-motion workload's frames (velocities up to 200 px per
    frame, a cut to a fresh scene every 30 frames, 5% exact duplicates; one
    seed per stream) around two frame steps of 16 streams, generated in worker
    processes before the GPU is touched."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    t0 = time.time()
    seeds = [HARD_SEED + i for i in range(HARD_STREAMS)]
    with ProcessPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)), mp_context=mp.get_context("fork")) as pool:
        out = list(pool.map(_hard_stream, seeds, chunksize=1))
    log(f"hard motion: generated {sum(len(d) for _, d in out)} 2160p pictures of {len(out)} streams in "
        f"{time.time() - t0:.1f}s")
    return out


def hard_leg(lanes, gen, steps):
    """The 2160p hard-motion workload: 16 concurrent streams, one frame step
    of all of them per analysis step (16 pairs at 3840x2160, R 1024), plain
    (one captured graph per step) and learned-seeded, at a mid-scene step and
    at the step across the scene cut, with the learned form's cost rise."""
    import numpy as np
    from libvc1_amd import abi, batching, learned
    from libvc1_amd.backend import ProposalNet
    p = abi.default_params(3840, 2160, pqindex=8, motion_search_range=1024, motion_local_search_range=32,
                           mode=abi.MODE_MIXED_MV, me_quality=2, allow_intra=1)
    mbs = 240 * 135
    be = lanes[0]
    nets = []
    for ln in lanes:
        nets.append(ProposalNet(ln))
        nets[-1].set_params(learned.load_weights())
    net = nets[0]
    idx = {id(ln): i for i, ln in enumerate(lanes)}
    res = {}
    uploaded = []
    try:
        for ln in lanes:
            ln.set_graphs(True)
        for label, t in HARD_STEPS:
            jobs = []
            for _, d in gen:
                fc, fr = be.upload(d[t]), be.upload(d[t - 1])
                uploaded += [fc, fr]
                jobs.append((fc, fr))
            plist = [p] * len(jobs)

            def plain(ln, k):
                ln.analyze_step(jobs, plist, want_results=False)

            def seeded(ln, k):
                ln.analyze_step_learned(nets[idx[id(ln)]], jobs, plist, 0.5, want_results=False)
            # the same frame step repeated, `len(lanes)` steps in flight
            secs = {}
            for name, fn in (("plain", plain), ("learned", seeded)):
                warm(lanes, fn)
                t0 = time.perf_counter()
                batching.run_in_flight(lanes, steps, fn)
                secs[name] = (time.perf_counter() - t0) / steps
            plain_s, learned_s = secs["plain"], secs["learned"]
            for ln in lanes:
                ln.set_graphs(False)
            cmp = learned.cost_compare(be, net, jobs, plist, 0.5)
            for ln in lanes:
                ln.set_graphs(True)
            n_mbs = len(jobs) * mbs
            res[label] = {"frame": t, "ms_per_step": round(plain_s * 1e3, 2), "value": round(n_mbs / plain_s, 1),
                          "learned_ms_per_step": round(learned_s * 1e3, 2),
                          "learned_value": round(n_mbs / learned_s, 1),
                          "learned_speedup": round(plain_s / learned_s, 4),
                          "learned_cost_rise": round(cmp["cost_rise"], 6),
                          "learned_seeded_fraction": round(cmp["seeded_fraction"], 4)}
            log(f"hard motion {label}: {res[label]}")
    finally:
        for ln in lanes:
            ln.set_graphs(False)
        for n in nets:
            n.close()
        for f in uploaded:
            f.release()
    return {"streams": HARD_STREAMS, "seed": HARD_SEED, "unit": "MB/s", "steps": res, "steps_in_flight": len(lanes),
            "config": "3840x2160, 16 streams in one step per frame step, R 1024, local 32, Mixed-MV qpel SATD, "
                      "velocities up to 200 px/frame, cuts every 30 frames, 5% duplicates"}


BATCHED_STREAMS, BATCHED_SEED = 64, 5


def batched_frames(scale: float):
    """The 64-stream batched workload's pictures (resolutions uniform over
    720x480, 1280x720, 1920x1080, 3840x2160; lengths uniform in 8..120
    frames times `scale`, at least 2; content as the 1080p workload with
    per-stream seeds), generated in worker processes before the GPU is
    touched."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from libvc1_amd import batching
    specs = batching.stream_specs(BATCHED_STREAMS, BATCHED_SEED, scale)
    t0 = time.time()
    with ProcessPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)), mp_context=mp.get_context("fork")) as pool:
        out = list(pool.map(batching._gen, specs, chunksize=1))
    log(f"batched: generated {sum(len(p) for *_, p in out)} pictures of {len(out)} streams in {time.time() - t0:.1f}s")
    return out


def batched_leg(lanes, generated, scale: float, check: bool):
    """Every P pair of the 64 streams through the batching runner (steps,
    modes and steps in flight from the tuned table, frames resident), timed
    once after a warm-up run; one 720x480 pair is checked against the
    oracle."""
    from libvc1_amd import batching
    be = lanes[0]
    streams = [batching.Stream(w, h, [be.upload(p) for p in pics], batching.stream_params(w, h), seed)
               for seed, w, h, pics in generated]
    try:
        batching.run(lanes, streams)                   # warm-up (graph capture, allocations)
        out = batching.run(lanes, streams)
        parity = None
        if check:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from _oracle import oracle_analyze
            si = next(i for i, s in enumerate(streams) if (s.width, s.height) == (720, 480))
            s = streams[si]
            got = be.analyze(s.frames[1], s.frames[0], s.params)
            want = oracle_analyze(s.params, generated[si][3][1], generated[si][3][0])
            bad = got.diff(want)
            parity = f"stream {si} pair (1, 0) bit-exact" if not bad else f"MISMATCH in {bad}"
    finally:
        for s in streams:
            for f in s.frames:
                f.release()
    res = {"value": round(out["mbs_per_s"], 1), "unit": "MB/s", "seconds": round(out["seconds"], 4),
           "macroblocks": out["macroblocks"], "pairs": sum(len(s.pairs) for s in streams), "streams": len(streams),
           "config": f"{BATCHED_STREAMS} streams (seed {BATCHED_SEED}), rasters uniform over 720x480/1280x720/"
                     f"1920x1080/3840x2160, lengths uniform 8-120 frames x {scale} (>= 2), P-only, R 256, local 32, "
                     "Mixed-MV qpel SATD; steps and modes from the tuned table",
           "per_resolution": {k: {"pairs": v["pairs"], "macroblocks": v["macroblocks"], "seconds": v["seconds"],
                                  "mode": f"{v['setting']['schedule']}{'+graph' if v['setting']['graphs'] else ''}"
                                          f" x{v['setting']['jobs_per_step']}, {v['setting'].get('in_flight', 1)}"
                                          " in flight"}
                              for k, v in out["per_resolution"].items()},
           "parity": parity}
    log(f"batched: {res['value']:,.0f} MB/s over {res['pairs']} pairs; {res['per_resolution']}; parity {parity}")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the diagonal-sweep comparison steps")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-stream batched leg")
    ap.add_argument("--no-hard", action="store_true", help="skip the 2160p hard-motion leg")
    ap.add_argument("--no-scorer", action="store_true", help="skip the batch-scorer latency leg")
    ap.add_argument("--no-seeded", action="store_true", help="skip the seeded-core leg")
    ap.add_argument("--no-bframes", action="store_true", help="skip the B-picture leg")
    ap.add_argument("--in-flight", type=int, default=0,
                    help="steps kept in flight, one context and host thread each (0: tuned on this workload before the timed steps)")
    ap.add_argument("--batched-scale", type=float, default=0.1,
                    help="stream lengths (8..120 frames) are scaled by this in the batched leg")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(streams and frame pairs shard across GPUs without communication)", file=sys.stderr)
        return 2

    from libvc1_amd import Backend, abi, batching

    params = workload_params()
    pairs, pics = make_frames()
    batched_streams = None if args.no_batched else batched_frames(args.batched_scale)
    hard_streams = None if args.no_hard else hard_frames()
    be = Backend(0)
    log(f"device {be.device_name}")
    frames = {k: be.upload(v) for k, v in pics.items()}
    jobs = [(frames[c], frames[r]) for c, r in pairs]
    plist = [params] * len(jobs)
    mbs_per_step = len(jobs) * ((WIDTH + 15) // 16) * ((HEIGHT + 15) // 16)
    # steps in flight: one context (own stream and device state) per host
    # thread, all on the frames uploaded by the first; the timed steps: each
    # step one captured HIP graph (relaxation passes driven from the device,
    # every per-macroblock field and the frame aggregates computed; nothing
    # copied back but the step summary)
    pool = [be] + [Backend(0) for _ in range(max(args.in_flight, max(batching.IN_FLIGHT_CANDIDATES)) - 1)]
    for ln in pool:
        ln.set_graphs(True)

    def plain_step(ln, k):
        ln.analyze_step(jobs, plist, want_results=False)
        t = ln.timing()
        return int(t.relax_passes), int(t.relax_analyses), int(t.graph_replay), int(t.schedule)
    flight = args.in_flight
    trial = {}
    if not flight:
        # the number in flight, tuned on this workload before the timed
        # steps (the table's trials run other clips and parameters): two
        # interleaved rounds at each candidate count, as many steps as the
        # timed region (fewer, and the start and the tail, when not every
        # context has a step, weigh more the more contexts there are), after
        # one warm-up step per context; each count's better round
        warm(pool[:max(batching.IN_FLIGHT_CANDIDATES)], plain_step)
        for _ in range(2):
            for k in batching.IN_FLIGHT_CANDIDATES:
                t0 = time.perf_counter()
                batching.run_in_flight(pool[:k], args.steps, plain_step)
                trial[k] = max(trial.get(k, 0.0), mbs_per_step * args.steps / (time.perf_counter() - t0))
        flight = max(trial, key=trial.get)
        log("steps in flight, trial MB/s: " + ", ".join(f"{k}: {v:,.0f}" for k, v in trial.items()) + f" -> {flight}")
    lanes = pool[:flight]
    warm(lanes, plain_step, args.warmup)
    log(f"warmup: {max(1, args.warmup)} step(s) on each of {len(lanes)} context(s)")
    t0 = time.perf_counter()
    recs = batching.run_in_flight(lanes, args.steps, plain_step)
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = mbs_per_step * args.steps / elapsed
    # the same steps one at a time on one context
    t0 = time.perf_counter()
    for k in range(args.steps):
        plain_step(be, k)
    value_one = mbs_per_step * args.steps / (time.perf_counter() - t0)
    # download-inclusive rate (every field of every pair back to the host)
    t1 = time.perf_counter()
    batching.run_in_flight(lanes, args.steps, lambda ln, k: ln.analyze_step(jobs, plist, want_results=True))
    with_dl = mbs_per_step * args.steps / (time.perf_counter() - t1)
    results = be.analyze_step(jobs, plist, want_results=True)
    for ln in pool:
        ln.set_graphs(False)
    schedule = {"kind": "relax" if recs[-1][3] == abi.SCHEDULE_RELAX else "diagonal",
                "launch": "one captured HIP graph per step" if recs[-1][2] == 1 else "eager",
                "steps_in_flight": len(lanes),
                "in_flight_trial": {str(k): round(v, 1) for k, v in trial.items()},
                "passes_per_step": [r[0] for r in recs], "analyses_per_step": [r[1] for r in recs]}
    log(f"timed {args.steps} steps, {len(lanes)} in flight: {ms_per_step:.1f} ms/step, {value:,.0f} MB/s "
        f"(one in flight {value_one:,.0f}), with download {with_dl:,.0f} MB/s, schedule {schedule}")

    # the same steps launched one by one with HIP events around every launch
    # of the macroblock kernel: its launch durations for the roofline
    be.set_profiling(True)
    mb_ms = 0.0
    mb_launches = 0
    work = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        be.analyze_step(jobs, plist, want_results=False)
        tp = be.timing()
        mb_ms += tp.mb_kernel_ms_total
        mb_launches += tp.mb_kernel_launches
        work = be.work()
    eager_value = mbs_per_step * args.steps / (time.perf_counter() - t0)
    be.set_profiling(False)
    schedule["eager_profiled_value"] = round(eager_value, 1)
    schedule["relax_pass_ms"] = [round(ms, 3) for _, ms in be.relax_passes()]

    # the anti-diagonal sweep on the same steps, for comparison
    if not args.no_compare:
        be.set_schedule(abi.SCHEDULE_DIAGONAL)
        be.analyze_step(jobs, plist, want_results=False)
        t0 = time.perf_counter()
        for i in range(args.steps):
            be.analyze_step(jobs, plist, want_results=False)
        schedule["diagonal_value"] = round(mbs_per_step * args.steps / (time.perf_counter() - t0), 1)
        be.set_schedule(abi.SCHEDULE_RELAX)
        log(f"diagonal sweep: {schedule['diagonal_value']:,.0f} MB/s")

    parity = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import oracle_analyze
        c, r = pairs[0]
        t2 = time.time()
        want = oracle_analyze(params, pics[c], pics[r])
        bad = results[0].diff(want)
        parity = "bit-exact" if not bad else f"MISMATCH in {bad}"
        log(f"parity of pair {pairs[0]} vs oracle: {parity} ({time.time() - t2:.1f}s)")

    # roofline of the macroblock analysis kernels (relaxation: k_mb_guess for
    # the starting state, which also runs the extended long-range searches
    # the passes read back, k_mb_relax0 for pass 0, k_mb_relaxq for the
    # work-queue passes; profiled under one tag)
    mb_kernel = "k_mb_analysis" if schedule["kind"] == "relax" else "k_mb_diagonal"
    cus, clk, peak = device_peak_ops()
    mb_ops = work.sad_px + work.interp_taps + work.satd_ops + work.sig_ops
    mb_sec_per_step = (mb_ms / args.steps) / 1e3
    achieved = mb_ops / mb_sec_per_step if mb_sec_per_step > 0 else 0.0
    avg_launch_us = mb_ms * 1e3 / max(1, mb_launches)
    roofline = {
        "bound": "valu",
        "achieved": round(achieved / 1e12, 4),
        "peak": round(peak / 1e12, 2),
        "unit": "TOP/s",
        "frac": round(achieved / peak, 5),
        "traffic": measured_traffic(mb_kernel),
        "kernel": "k_mb_guess + k_mb_relax0 + k_mb_relaxq (the guess pass, then one launch per relaxation pass)"
                  if mb_kernel == "k_mb_analysis" else mb_kernel,
        "avg_launch_us": round(avg_launch_us, 2),
        "launches_per_step": mb_launches // max(1, args.steps),
        "ops_per_step": {"sad_px": work.sad_px, "interp_taps": work.interp_taps, "satd_ops": work.satd_ops,
                         "sig_ops": work.sig_ops},
        "peak_basis": f"{cus} CU x 64 lanes x {clk / 1e9:.2f} GHz int32 lane-ops",
    }
    # algorithmic HBM bytes: each macroblock's current and reference luma once (2 x 256 B)
    launches = max(1, mb_launches // max(1, args.steps))
    alg_bytes_launch = (mbs_per_step / launches) * 512.0
    launch_s = avg_launch_us / 1e6
    roofline_hbm = {"bound": "hbm", "achieved": round(alg_bytes_launch / launch_s / 1e9, 3) if launch_s else 0.0,
                    "peak": 8000.0, "unit": "GB/s",
                    "frac": round(alg_bytes_launch / launch_s / 8e12, 7) if launch_s else 0.0,
                    "traffic": measured_traffic(mb_kernel), "algorithmic_bytes_per_launch": round(alg_bytes_launch),
                    "kernel": mb_kernel}

    seeded = None
    if not args.no_seeded:
        seeded = seeded_leg(be, jobs, plist, value, args.steps, args.warmup, mbs_per_step,
                            None if args.no_check else (pics[pairs[0][0]], pics[pairs[0][1]]))

    bpics = None
    if not args.no_bframes:
        bpics = b_leg(be, pics, frames, args.steps, not args.no_check)

    learned_out = None
    if not args.no_learned:
        learned_out = learned_leg(pool, jobs, plist, value, value_one, args.steps, args.warmup, mbs_per_step,
                                  in_flight=args.in_flight)

    scorer = None if args.no_scorer else scorer_leg(be, pics, frames)

    hard = None
    if hard_streams is not None:
        hard = hard_leg(lanes, hard_streams, max(2, args.steps // 4))
        hard_streams = None

    batched = None
    if batched_streams is not None:
        batched = batched_leg(pool, batched_streams, args.batched_scale, not args.no_check)
        batched_streams = None

    cpu = None
    if not args.no_cpu:
        log("timing the reference CPU path ...")
        cpu = cpu_baseline(pairs, pics, params)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "1080p IBBP GOP15, 16 P-frame pairs per step, R=512 local 32, Mixed-MV qpel SATD, "
                               "PQINDEX 8, plain GPU ME (bit-exact)", "width": WIDTH, "height": HEIGHT,
                   "p_frames_per_step": len(jobs), "macroblocks_per_step": mbs_per_step, "seed": SEED},
        "form": "plain",
        "steps_in_flight": len(lanes),
        "value_one_in_flight": round(value_one, 1),
        "schedule": schedule,
        "value_with_download": round(with_dl, 1),
        "parity": parity,
        "roofline": roofline,
        "roofline_hbm": roofline_hbm,
        "cpu_baseline": cpu,
        "learned": learned_out,
        "seeded_core": seeded,
        "b_pictures": bpics,
        "batched": batched,
        "hard_motion_2160p": hard,
        "scorer": scorer,
    }
    for f in frames.values():
        f.release()
    for ln in lanes[1:]:
        ln.close()
    be.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
