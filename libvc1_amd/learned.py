"""Training, weights and evaluation of the learned-seeding proposal network.

The network (vc1me_nn_* in include/vc1me.h) is trained on the GPU by the
library's own HIP forward, backward and Adam kernels (vc1me_nn_train_step).
Its targets are the final 1-MV vectors that the plain GPU search itself finds
on synthetic sequences (``synth.pan_rectangles``); training seeds start at
``TRAIN_SEED_BASE`` and evaluation seeds at ``EVAL_SEED_BASE``, two disjoint
ranges that also exclude the benchmark's seed.

Weights are stored as a versioned binary: 8-byte magic ``VC1MEPN1``, u32
format version, u32 parameter count, 32-byte SHA-256 of the parameter bytes,
then the float32 parameters (little endian) in the layout vc1me_nn_set_params
takes.  A JSON sidecar (same name, ``.json``) records how they were trained.

    python -m libvc1_amd.learned train [--sequences 200] [--frames 30] [--steps 1500] [--out PATH]
    python -m libvc1_amd.learned eval [--weights PATH] [--sequences 4]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import struct
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import abi, synth
from .backend import Backend, ProposalNet

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposal_net.bin")
MAGIC = b"VC1MEPN1"
FORMAT_VERSION = 1
TRAIN_SEED_BASE = 7_000_000
EVAL_SEED_BASE = 9_000_000
CHANNELS, LAYER_CIN = 32, (2, 32, 32, 32)


def log(msg: str) -> None:
    print(f"[learned] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ weights

def param_count() -> int:
    return sum(CHANNELS * 9 * c + CHANNELS for c in LAYER_CIN) + 5 * CHANNELS + 5


def init_params(seed: int) -> np.ndarray:
    """He-normal convolution weights, zero biases, a small head whose
    confidence starts below the threshold."""
    rng = np.random.default_rng(seed)
    parts = []
    for cin in LAYER_CIN:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * cin)), CHANNELS * 9 * cin))
        parts.append(np.zeros(CHANNELS))
    parts.append(rng.normal(0.0, 0.01, 5 * CHANNELS))
    parts.append(np.array([0.0, 0.0, 0.0, 0.0, -2.0]))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == param_count()
    return p


def save_weights(path: str, params: np.ndarray, meta: dict) -> str:
    params = np.ascontiguousarray(params, "<f4")
    body = params.tobytes()
    digest = hashlib.sha256(body).digest()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<II", FORMAT_VERSION, params.size) + digest + body)
    meta = dict(meta, sha256=digest.hex(), params=int(params.size), format_version=FORMAT_VERSION)
    with open(path + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    return digest.hex()


def load_weights(path: str = WEIGHTS) -> np.ndarray:
    """Parameters from a weights file; raises ValueError if the header, the
    size or the hash does not match."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 48 or data[:8] != MAGIC:
        raise ValueError(f"{path}: not a proposal-network weights file")
    version, count = struct.unpack("<II", data[8:16])
    if version != FORMAT_VERSION or count != param_count():
        raise ValueError(f"{path}: format {version} with {count} parameters, expected {FORMAT_VERSION} / {param_count()}")
    body = data[48:]
    if len(body) != 4 * count or hashlib.sha256(body).digest() != data[16:48]:
        raise ValueError(f"{path}: parameter bytes do not match the stored SHA-256")
    return np.frombuffer(body, "<f4").astype(np.float32)


# ------------------------------------------------------------------ data

def workload_params(width: int, height: int) -> abi.Params:
    """The 1080p IBBP search configuration (R 512, local 32, Mixed-MV, SATD
    finalists, PQINDEX 8, intra decision on) at any raster."""
    return abi.default_params(width, height, pqindex=8, motion_search_range=512, motion_local_search_range=32,
                              mode=abi.MODE_MIXED_MV, me_quality=2, allow_intra=1)


def _sequence(args):
    seed, frames, width, height = args
    pics = synth.pan_rectangles(seed, frames, width=width, height=height, max_pan_px=16, max_obj_px=48)
    return seed, [pics[k] for k in range(frames)]


def sequences(seed_base: int, count: int, frames: int, width: int, height: int, pool=None):
    """Yields (seed, [Picture] * frames) for seeds seed_base .. seed_base+count-1,
    generated in `pool` (a process pool started before the GPU is touched)
    or in this process."""
    jobs = [(seed_base + i, frames, width, height) for i in range(count)]
    if pool is None:
        for j in jobs:
            yield _sequence(j)
        return
    yield from pool.map(_sequence, jobs, chunksize=1)


def start_pool(workers: int):
    """A fork-based process pool whose workers exist before any HIP call of
    this process (workers only run numpy)."""
    if workers <= 1:
        return None
    import multiprocessing as mp
    pool = ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork"))
    list(pool.map(int, range(workers)))   # start every worker now
    return pool


class Dataset:
    """Frame pairs resident on the device with the plain search's final
    1-MV vectors as targets (quarter-pel) and a mask of inter macroblocks."""

    def __init__(self, backend: Backend, width: int, height: int):
        self.backend = backend
        self.width, self.height = width, height
        self.params = workload_params(width, height)
        self.mbs = int(np.prod(abi.mb_grid(width, height)))
        self.frames: list = []
        self.pairs: list[tuple[int, int]] = []
        self.targets: list[np.ndarray] = []
        self.valid: list[np.ndarray] = []
        self.globals: list[tuple[int, int]] = []

    def add_sequence(self, pics, gap: int) -> None:
        base = len(self.frames)
        self.frames += [self.backend.upload(p) for p in pics]
        new = [(base + k, base + k - gap) for k in range(gap, len(pics))]
        be = self.backend
        be.set_cost_report(True)
        try:
            for i in range(0, len(new), 16):
                chunk = new[i:i + 16]
                res = be.analyze_step([(self.frames[c], self.frames[r]) for c, r in chunk], [self.params] * len(chunk))
                for j, r in enumerate(res):
                    _, one = be.last_mb_costs(j, self.mbs)
                    self.targets.append(one.copy())
                    self.valid.append((r.use_intra == 0).astype(np.uint8))
                    self.globals.append((r.struct.global_dx, r.struct.global_dy))
        finally:
            be.set_cost_report(False)
        self.pairs += new

    def batch(self, idx):
        pairs = [(self.frames[self.pairs[i][0]], self.frames[self.pairs[i][1]]) for i in idx]
        return (pairs, np.stack([self.targets[i] for i in idx]), np.stack([self.valid[i] for i in idx]),
                np.array([self.globals[i] for i in idx], np.int32))

    def release(self) -> None:
        for f in self.frames:
            f.release()
        self.frames = []


# ------------------------------------------------------------------ training

def train(backend: Backend, *, sequences_n: int = 200, frames: int = 30, width: int = 1920, height: int = 1080,
          gap: int = 3, steps: int = 1500, batch: int = 8, lr: float = 2e-3, seed: int = 1, pool=None):
    """Trains a network from scratch; returns (params, meta)."""
    t0 = time.time()
    data = Dataset(backend, width, height)
    for n, (s, pics) in enumerate(sequences(TRAIN_SEED_BASE, sequences_n, frames, width, height, pool)):
        data.add_sequence(pics, gap)
        if (n + 1) % 10 == 0 or n + 1 == sequences_n:
            log(f"sequences {n + 1}/{sequences_n}: {len(data.pairs)} pairs ({time.time() - t0:.0f}s)")
    t_data = time.time() - t0
    net = ProposalNet(backend)
    net.set_params(init_params(seed))
    rng = np.random.default_rng(seed)
    hist = []
    t1 = time.time()
    ema = None
    for step in range(steps):
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps)) + 1e-5
        idx = rng.choice(len(data.pairs), size=batch, replace=False)
        pairs, tgt, valid, glob = data.batch(idx)
        loss = net.train_step(pairs, tgt, valid, rate, glob)
        ema = loss if ema is None else 0.98 * ema + 0.02 * loss
        if step % 100 == 0 or step + 1 == steps:
            hist.append((step, float(loss), float(ema)))
            log(f"step {step}/{steps}: loss {loss:.4f} (ema {ema:.4f}) lr {rate:.2e} ({time.time() - t1:.0f}s)")
    t_train = time.time() - t1
    params = net.get_params()
    net.close()
    data.release()
    meta = dict(train_seeds=[TRAIN_SEED_BASE, TRAIN_SEED_BASE + sequences_n - 1], frames=frames, width=width,
                height=height, pair_gap=gap, pairs=len(data.pairs), steps=steps, batch=batch, lr=lr,
                schedule="cosine", optimizer="adam(0.9, 0.999, 1e-8)", init_seed=seed,
                seconds_data=round(t_data, 1), seconds_train=round(t_train, 1), loss_history=hist,
                targets="final 1-MV vectors of the plain GPU search as residuals from its global vector; inter "
                        "macroblocks only")
    return params, meta


# ------------------------------------------------------------------ evaluation

def cost_compare(backend: Backend, net: ProposalNet, jobs, params_list, threshold: float = 0.5) -> dict:
    """Plain vs learned on the same jobs: per-macroblock 20*(SAD + lambda*R)
    of the final 1-MV vector from the cost report, summed and compared."""
    mbs = [int(np.prod(abi.mb_grid(p.width, p.height))) for p in params_list]
    be = backend
    be.set_cost_report(True)
    try:
        be.analyze_step(jobs, params_list, want_results=False)
        plain = [be.last_mb_costs(j, m)[0] for j, m in enumerate(mbs)]
        be.analyze_step_learned(net, jobs, params_list, threshold, want_results=False)
        learned = [be.last_mb_costs(j, m)[0] for j, m in enumerate(mbs)]
        seeded = be.work().seeded_macroblocks
    finally:
        be.set_cost_report(False)
    cp = np.concatenate(plain).astype(np.float64)
    cl = np.concatenate(learned).astype(np.float64)
    return {
        "cost_rise": float(cl.sum() / cp.sum() - 1.0),
        "frac_worse": float(np.mean(cl > cp)),
        "mean_excess": float(np.mean(cl - cp) / 20.0),
        "seeded_fraction": float(seeded / cp.size),
        "macroblocks": int(cp.size),
    }


def evaluate(backend: Backend, params: np.ndarray, *, sequences_n: int = 4, frames: int = 16, width: int = 1920,
             height: int = 1080, gap: int = 3, threshold: float = 0.5, pool=None) -> dict:
    net = ProposalNet(backend)
    net.set_params(params)
    p = workload_params(width, height)
    out = []
    for s, pics in sequences(EVAL_SEED_BASE, sequences_n, frames, width, height, pool):
        fr = [backend.upload(x) for x in pics]
        jobs = [(fr[k], fr[k - gap]) for k in range(gap, len(fr))]
        r = cost_compare(backend, net, jobs, [p] * len(jobs), threshold)
        r["seed"] = s
        out.append(r)
        log(f"eval seed {s}: {r}")
        for f in fr:
            f.release()
    net.close()
    tot = sum(r["macroblocks"] for r in out)
    agg = {k: sum(r[k] * r["macroblocks"] for r in out) / tot for k in ("frac_worse", "mean_excess", "seeded_fraction")}
    agg["cost_rise_mean_over_sequences"] = float(np.mean([r["cost_rise"] for r in out]))
    return {"per_sequence": out, "aggregate": agg}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m libvc1_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--sequences", type=int, default=200)
    t.add_argument("--frames", type=int, default=30)
    t.add_argument("--steps", type=int, default=1500)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--width", type=int, default=1920)
    t.add_argument("--height", type=int, default=1080)
    t.add_argument("--workers", type=int, default=16)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--eval-sequences", type=int, default=2)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--sequences", type=int, default=4)
    e.add_argument("--frames", type=int, default=16)
    e.add_argument("--threshold", type=float, default=0.5)
    e.add_argument("--workers", type=int, default=16)
    a = ap.parse_args(argv)
    pool = start_pool(a.workers)
    be = Backend(0)
    try:
        if a.cmd == "train":
            params, meta = train(be, sequences_n=a.sequences, frames=a.frames, width=a.width, height=a.height,
                                 steps=a.steps, batch=a.batch, lr=a.lr, pool=pool)
            if a.eval_sequences:
                meta["eval"] = evaluate(be, params, sequences_n=a.eval_sequences, width=a.width, height=a.height,
                                        pool=pool)["aggregate"]
            digest = save_weights(a.out, params, meta)
            log(f"saved {a.out} sha256 {digest}")
        else:
            r = evaluate(be, load_weights(a.weights), sequences_n=a.sequences, frames=a.frames,
                         threshold=a.threshold, pool=pool)
            print(json.dumps(r))
    finally:
        be.close()
        if pool is not None:
            pool.shutdown()
    return 0


if __name__ == "__main__":
    sys.exit(main())
