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

The autotuner measures, per raster, macroblock searches per second for each
candidate step size under each launch mode (relaxation; sweep; sweep replayed
from a graph), and persists the winners in a JSON table per device
(``tuning/<device>.json``).  The design has one wave per macroblock by
construction, so the tuned knobs are the step composition, the schedule and
the launch mode; every mode gives the same results, so the table changes
speed only.

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
DEFAULT_SETTING = {"jobs_per_step": 16, "schedule": "relax", "graphs": False}
# launch modes the tuner tries: (schedule, graphs)
MODES = (("relax", False), ("diagonal", False), ("diagonal", True))


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
        fill = [b for b in sizes if b <= pairs] or sizes[:1]
        for k, rate in trials.items():
            m = re.fullmatch(r"(\d+)(r?)(g?)", k)
            if not m or int(m.group(1)) not in fill:
                continue
            if best is None or rate > best[0]:
                best = (rate, int(m.group(1)), "relax" if m.group(2) else "diagonal", bool(m.group(3)))
        if best is not None:
            # every trial grew with the step size: take all the pairs (at most
            # 255 per step) in the mode measured best at the filled size
            return {"jobs_per_step": max(best[1], min(pairs, 255)), "schedule": best[2], "graphs": best[3],
                    "mbs_per_s": best[0]}
    st = dict(table.get("settings", {}).get(raster, DEFAULT_SETTING))
    st.setdefault("schedule", "diagonal")      # tables written before the relaxation schedule
    return st


def _schedule_code(name: str) -> int:
    return abi.SCHEDULE_RELAX if name == "relax" else abi.SCHEDULE_DIAGONAL


def _run_steps(backend: Backend, jobs, params, per_step: int, graphs: bool, schedule: str = "relax",
               collect: bool = False) -> list:
    backend.set_graphs(graphs)
    backend.set_schedule(_schedule_code(schedule))
    out = []
    try:
        for i in range(0, len(jobs), per_step):
            chunk = jobs[i:i + per_step]
            out += backend.analyze_step(chunk, [params] * len(chunk), want_results=collect)
    finally:
        backend.set_graphs(False)
        backend.set_schedule(abi.SCHEDULE_RELAX)
    return out


def autotune(backend: Backend, resolutions=RESOLUTIONS, candidates=STEP_CANDIDATES, frames: int = 5,
             seed: int = 77, repeats: int = 2) -> dict:
    """Measures MB searches/s per raster for every step size and launch mode
    (MODES) on `frames`-frame synthetic clips; returns the table (not yet
    saved)."""
    settings, trials = {}, {}
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
                _run_steps(backend, jobs, params, b, graphs, schedule)          # warm up (and capture)
                t0 = time.perf_counter()
                for _ in range(repeats):
                    _run_steps(backend, jobs, params, b, graphs, schedule)
                rate = repeats * b * mbs / (time.perf_counter() - t0)
                res[f"{b}{'r' if schedule == 'relax' else ''}{'g' if graphs else ''}"] = round(rate, 1)
                if best is None or rate > best[0]:
                    best = (rate, b, graphs, schedule)
        for f in fr:
            f.release()
        settings[f"{w}x{h}"] = {"jobs_per_step": best[1], "schedule": best[3], "graphs": best[2],
                                "mbs_per_s": round(best[0], 1)}
        trials[f"{w}x{h}"] = res
        log(f"{w}x{h}: best {settings[f'{w}x{h}']} of {res}")
    return {"device": backend.device_name, "settings": settings, "trials": trials,
            "candidates": list(candidates), "created": time.strftime("%Y-%m-%d %H:%M:%S")}


def save_table(backend: Backend, table: dict) -> str:
    os.makedirs(TUNING_DIR, exist_ok=True)
    p = table_path(backend)
    with open(p, "w") as f:
        json.dump(table, f, indent=1)
    return p


def run(backend: Backend, streams: list[Stream], table: dict | None = None, collect: bool = False) -> dict:
    """Analyses every P pair of every stream; returns counts and timing, and
    with collect=True the results as {(stream index, pair index): result}."""
    table = load_table(backend) if table is None else table
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
        res = _run_steps(backend, jobs, stream_params(w, h), int(st["jobs_per_step"]), bool(st["graphs"]),
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
    run(be, streams[:4], table)                       # warm up every code path once
    out = run(be, streams, table)
    out.update({"streams": len(streams), "pairs": sum(len(s.pairs) for s in streams),
                "length_scale": a.length_scale, "seed": a.seed,
                "resolutions": {f"{w}x{h}": sum(1 for s in streams if (s.width, s.height) == (w, h))
                                for w, h in RESOLUTIONS}})
    for s in streams:
        for f in s.frames:
            f.release()
    be.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
