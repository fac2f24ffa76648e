"""Batched plain motion estimation over many streams of uneven resolution and
length, with an autotuned step composition and HIP-graph replay.

Every P pair of a stream (frame k against frame k-1) is independent of the
others once the frames exist, so the runner flattens the pairs of all
streams, groups them by raster, and runs each group as analysis steps of
`jobs_per_step` pairs: one launch per anti-diagonal covers every macroblock
of the step (vc1me_analyze_step) under one of the two macroblock schedules
(vc1me_set_schedule): the relaxation schedule (every macroblock at once, then
worklist passes until nothing changes) or the anti-diagonal sweep, which may
be replayed from a captured HIP graph (vc1me_set_graphs).  Grouping by raster
keeps a step's launch count equal to its own frames' diagonal count, so small
frames never wait on large ones.

Steps may also be kept in flight: with k contexts (each with its own stream
and device state) driven by k host threads, the latency-bound end of one
step (the relaxation's last passes, each a chain of dependent analyses; the
global search's dependent stages) overlaps the full-occupancy passes of the
next.

The autotuner measures, per raster, macroblock searches per second for each
candidate step size under each launch mode (relaxation; sweep; sweep replayed
from a graph) and each number of steps in flight, and persists the winners in
a JSON table per device (``tuning/<device>.json``).  The design has one wave
per macroblock by construction, so the tuned knobs are the step composition,
the schedule, the launch mode and the steps in flight; every setting gives
the same results, so the table changes speed only.

    python -m libvc1_amd.batching tune  [--resolutions 720x480,1280x720,1920x1080,3840x2160]
    python -m libvc1_amd.batching bench [--streams 64] [--seed 5] [--length-scale 1.0]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import abi, synth
from .backend import Backend

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
RESOLUTIONS = ((720, 480), (1280, 720), (1920, 1080), (3840, 2160))
STEP_CANDIDATES = (4, 8, 16, 32, 64, 128, 255)
IN_FLIGHT_CANDIDATES = (1, 2, 3)
DEFAULT_SETTING = {"jobs_per_step": 16, "schedule": "relax", "graphs": False, "in_flight": 1}
# launch modes the tuner tries: (schedule, graphs)
MODES = (("relax", True), ("diagonal", False), ("diagonal", True))
# a trial key: step size, "r" for the relaxation, "g" for graph replay, "xK"
# for K steps in flight (absent: one)
TRIAL_KEY = re.compile(r"(\d+)(r?)(g?)(?:x(\d+))?")


def log(msg: str) -> None:
    print(f"[batching] {msg}", file=sys.stderr, flush=True)


def stream_params(width: int, height: int, search_range: int = 256) -> abi.Params:
    """P-only streams at range 256 (Mixed-MV, qpel, SATD finalists, PQINDEX 8)."""
    return abi.default_params(width, height, pqindex=8, motion_search_range=search_range,
                              motion_local_search_range=32, mode=abi.MODE_MIXED_MV, me_quality=2, allow_intra=1)


@dataclass
class Stream:
    width: int
    height: int
    frames: list                     # resident vc1me frames
    params: abi.Params
    seed: int = 0

    @property
    def pairs(self):
        return [(k, k - 1) for k in range(1, len(self.frames))]


def _gen(args):
    seed, width, height, length = args
    d = synth.pan_rectangles(seed, length, width=width, height=height, max_pan_px=16, max_obj_px=48)
    return seed, width, height, [d[k] for k in range(length)]


def stream_specs(count: int, seed: int, length_scale: float = 1.0, resolutions=RESOLUTIONS):
    """(seed, width, height, length) of `count` streams: resolution uniform
    over `resolutions`, length uniform in 8..120 frames (times length_scale,
    at least 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        n = int(rng.integers(8, 121))
        out.append((seed * 1000 + i, w, h, max(2, int(round(n * length_scale)))))
    return out


def _device_key(name: str) -> str:
    return re.sub(r"[^A-Za-z0-9_.-]+", "_", name.strip()) or "device"


def table_path(backend: Backend) -> str:
    return os.path.join(TUNING_DIR, _device_key(backend.device_name.split("(")[-1].split(":")[0]) + ".json")


def load_table(backend: Backend) -> dict:
    p = table_path(backend)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def run_in_flight(backends, count: int, fn) -> list:
    """fn(backend, k) for k in range(count), with one host thread per backend
    (a context with its own stream), each taking the next k as soon as its
    previous call returned: up to len(backends) steps in flight on the device.
    Returns fn's results in k order; the first exception is re-raised once
    every thread has stopped."""
    import threading
    backends = list(backends)
    if len(backends) <= 1 or count <= 1:
        return [fn(backends[0], k) for k in range(count)]
    out = [None] * count
    state = {"next": 0}
    errors = []
    lock = threading.Lock()

    def lane(be):
        while True:
            with lock:
                if errors or state["next"] >= count:
                    return
                k = state["next"]
                state["next"] += 1
            try:
                out[k] = fn(be, k)
            except BaseException as e:  # noqa: BLE001 - re-raised on the calling thread
                with lock:
                    errors.append(e)
                return

    threads = [threading.Thread(target=lane, args=(be,)) for be in backends]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return out


def setting_for(table: dict, width: int, height: int, pairs: int | None = None) -> dict:
    """The tuned setting of a raster.  With `pairs` (the pairs the runner has
    for the raster) and the table's trials, the best measured mode among step
    sizes the pairs can fill: a raster with few pairs gets the mode that was
    fastest at that size, not the one tuned for full steps."""
    raster = f"{width}x{height}"
    trials = table.get("trials", {}).get(raster)
    if pairs is not None and trials:
        best = None
        sizes = sorted({int(re.match(r"\d+", k).group()) for k in trials})
        for k, rate in trials.items():
            m = TRIAL_KEY.fullmatch(k)
            if not m:
                continue
            b, flight = int(m.group(1)), int(m.group(4) or 1)
            # the pairs must fill `flight` steps of the trial's size
            if b * flight > pairs and (b, flight) != (sizes[0], 1):
                continue
            if best is None or rate > best[0]:
                best = (rate, b, "relax" if m.group(2) else "diagonal", bool(m.group(3)), flight)
        if best is not None:
            # every trial grew with the step size: spread the pairs over the
            # steps in flight (at most 255 per step) in the mode measured best
            per = max(best[1], min(255, -(-pairs // best[4])))
            return {"jobs_per_step": per, "schedule": best[2], "graphs": best[3], "in_flight": best[4],
                    "mbs_per_s": best[0]}
    st = dict(table.get("settings", {}).get(raster, DEFAULT_SETTING))
    st.setdefault("schedule", "diagonal")      # tables written before the relaxation schedule
    st.setdefault("in_flight", 1)              # tables written before steps in flight
    return st


def best_in_flight(table: dict, width: int, height: int, per_step: int, default: int = 2) -> int:
    """Steps in flight measured best for captured relaxation steps of
    `per_step` pairs of a raster (trials "<n>rg" and "<n>rgx<k>"), else
    `default`."""
    trials = table.get("trials", {}).get(f"{width}x{height}", {})
    best = None
    for k, rate in trials.items():
        m = TRIAL_KEY.fullmatch(k)
        if m and int(m.group(1)) == per_step and m.group(2) and m.group(3):
            flight = int(m.group(4) or 1)
            if best is None or rate > best[0]:
                best = (rate, flight)
    return best[1] if best else default


def _schedule_code(name: str) -> int:
    return abi.SCHEDULE_RELAX if name == "relax" else abi.SCHEDULE_DIAGONAL


def _run_steps(backends, jobs, params, per_step: int, graphs: bool, schedule: str = "relax",
               collect: bool = False) -> list:
    """The jobs as steps of per_step pairs, as many in flight as there are
    backends (a single Backend: one at a time)."""
    backends = [backends] if isinstance(backends, Backend) else list(backends)
    for be in backends:
        be.set_graphs(graphs)
        be.set_schedule(_schedule_code(schedule))
    chunks = [jobs[i:i + per_step] for i in range(0, len(jobs), per_step)]
    try:
        res = run_in_flight(backends, len(chunks),
                            lambda be, k: be.analyze_step(chunks[k], [params] * len(chunks[k]), want_results=collect))
    finally:
        for be in backends:
            be.set_graphs(False)
            be.set_schedule(abi.SCHEDULE_RELAX)
    return [r for part in res for r in part]


def autotune(backend: Backend, resolutions=RESOLUTIONS, candidates=STEP_CANDIDATES, frames: int = 5,
             seed: int = 77, repeats: int = 6, flights=IN_FLIGHT_CANDIDATES) -> dict:
    """Measures MB searches/s per raster for every step size, launch mode
    (MODES) and number of steps in flight on `frames`-frame synthetic clips,
    `repeats` timed steps per context (enough that the first and last steps,
    which overlap nothing, do not decide the number in flight); returns the
    table (not yet saved).  Steps in flight run on extra
    contexts of the same device, created here and closed at the end."""
    extra = [Backend(backend.device) for _ in range(max(flights) - 1)]
    settings, trials = {}, {}
    try:
        for (w, h) in resolutions:
            pics = synth.pan_rectangles(seed + w, frames, width=w, height=h)
            fr = [backend.upload(pics[k]) for k in range(frames)]
            base = [(fr[k], fr[k - 1]) for k in range(1, frames)]
            params = stream_params(w, h)
            mbs = int(np.prod(abi.mb_grid(w, h)))
            best, res = None, {}
            for b in candidates:
                jobs = (base * ((b + len(base) - 1) // len(base)))[:b]   # b independent jobs
                for schedule, graphs in MODES:
                    for k in flights:
                        lanes = [backend] + extra[:k - 1]
                        steps = jobs * (repeats * k)
                        # one step on every context (and its capture): dealt
                        # through run_in_flight, a context could miss it and
                        # capture inside the timed steps
                        for ln in lanes:
                            _run_steps(ln, jobs, params, b, graphs, schedule)
                        t0 = time.perf_counter()
                        _run_steps(lanes, steps, params, b, graphs, schedule)
                        rate = len(steps) * mbs / (time.perf_counter() - t0)
                        key = f"{b}{'r' if schedule == 'relax' else ''}{'g' if graphs else ''}{f'x{k}' if k > 1 else ''}"
                        res[key] = round(rate, 1)
                        if best is None or rate > best[0]:
                            best = (rate, b, graphs, schedule, k)
            for f in fr:
                f.release()
            settings[f"{w}x{h}"] = {"jobs_per_step": best[1], "schedule": best[3], "graphs": best[2],
                                    "in_flight": best[4], "mbs_per_s": round(best[0], 1)}
            trials[f"{w}x{h}"] = res
            log(f"{w}x{h}: best {settings[f'{w}x{h}']} of {res}")
    finally:
        for be in extra:
            be.close()
    return {"device": backend.device_name, "settings": settings, "trials": trials,
            "candidates": list(candidates), "in_flight": list(flights),
            "created": time.strftime("%Y-%m-%d %H:%M:%S")}


def save_table(backend: Backend, table: dict) -> str:
    os.makedirs(TUNING_DIR, exist_ok=True)
    p = table_path(backend)
    with open(p, "w") as f:
        json.dump(table, f, indent=1)
    return p


def run(backend, streams: list[Stream], table: dict | None = None, collect: bool = False) -> dict:
    """Analyses every P pair of every stream; returns counts and timing, and
    with collect=True the results as {(stream index, pair index): result}.
    `backend` is one Backend or a list of contexts of one device (the first
    owns the frames); a raster whose setting keeps k steps in flight uses the
    first k of them."""
    backends = [backend] if isinstance(backend, Backend) else list(backend)
    table = load_table(backends[0]) if table is None else table
    groups: dict = {}
    keys: dict = {}
    for si, s in enumerate(streams):
        groups.setdefault((s.width, s.height), []).extend((s.frames[c], s.frames[r]) for c, r in s.pairs)
        keys.setdefault((s.width, s.height), []).extend((si, k) for k in range(len(s.pairs)))
    total_mbs, per, results = 0, {}, {}
    t0 = time.perf_counter()
    for (w, h), jobs in sorted(groups.items()):
        st = setting_for(table, w, h, len(jobs))
        t1 = time.perf_counter()
        lanes = backends[:max(1, min(len(backends), int(st.get("in_flight", 1))))]
        res = _run_steps(lanes, jobs, stream_params(w, h), int(st["jobs_per_step"]), bool(st["graphs"]),
                         str(st["schedule"]), collect)
        results.update(zip(keys[(w, h)], res))
        mbs = len(jobs) * int(np.prod(abi.mb_grid(w, h)))
        per[f"{w}x{h}"] = {"pairs": len(jobs), "macroblocks": mbs, "seconds": round(time.perf_counter() - t1, 4),
                          "setting": st}
        total_mbs += mbs
    elapsed = time.perf_counter() - t0
    out = {"macroblocks": total_mbs, "seconds": elapsed, "mbs_per_s": total_mbs / elapsed, "per_resolution": per}
    if collect:
        out["results"] = results
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m libvc1_amd.batching")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("tune")
    t.add_argument("--resolutions", default=",".join(f"{w}x{h}" for w, h in RESOLUTIONS))
    t.add_argument("--out", default=None)
    b = sub.add_parser("bench")
    b.add_argument("--streams", type=int, default=64)
    b.add_argument("--seed", type=int, default=5)
    b.add_argument("--length-scale", type=float, default=1.0)
    b.add_argument("--workers", type=int, default=16)
    b.add_argument("--table", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "tune":
        be = Backend(0)
        res = [tuple(int(v) for v in r.split("x")) for r in a.resolutions.split(",")]
        table = autotune(be, res)
        path = a.out or table_path(be)
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(table, f, indent=1)
        log(f"saved {path}")
        be.close()
        return 0
    specs = stream_specs(a.streams, a.seed, a.length_scale)
    pool = None
    if a.workers > 1:   # frame generation in worker processes started before any HIP call
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        pool = ProcessPoolExecutor(max_workers=a.workers, mp_context=mp.get_context("fork"))
        list(pool.map(int, range(a.workers)))
    be = Backend(0)
    table = json.load(open(a.table)) if a.table else load_table(be)
    streams = []
    t0 = time.time()
    gen = pool.map(_gen, specs, chunksize=1) if pool else map(_gen, specs)
    for seed, w, h, pics in gen:
        streams.append(Stream(w, h, [be.upload(p) for p in pics], stream_params(w, h), seed))
    log(f"{len(streams)} streams, {sum(len(s.frames) for s in streams)} frames resident ({time.time() - t0:.0f}s)")
    if pool:
        pool.shutdown()
    flight = max([int(v.get("in_flight", 1)) for v in table.get("settings", {}).values()] + [1])
    lanes = [be] + [Backend(0) for _ in range(flight - 1)]
    run(lanes, streams[:4], table)                    # warm up every code path once
    out = run(lanes, streams, table)
    out.update({"streams": len(streams), "pairs": sum(len(s.pairs) for s in streams),
                "length_scale": a.length_scale, "seed": a.seed,
                "resolutions": {f"{w}x{h}": sum(1 for s in streams if (s.width, s.height) == (w, h))
                                for w, h in RESOLUTIONS}})
    for s in streams:
        for f in s.frames:
            f.release()
    for lane in lanes[1:]:
        lane.close()
    be.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
