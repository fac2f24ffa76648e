#!/usr/bin/env python3
"""Write the golden fixtures of the P-frame motion-estimation path.

Runs the real libvc1 reference (compiled from its own sources into
oracle/_ref/libvc1ref.so by `make -C oracle ref`) on small deterministic
inputs and stores inputs and every PAnalysis output field as .npz fixtures,
plus SHA-256 digests of the outputs of larger synthetic cases
(analyze_digests.json), the seeded core's fixtures (seeded/*.npz: the
inputs, the seed's vectors and distant fields, every output field), and the
intensity-compensation fixtures (intensity.json: estimate_intensity_comp on
generated fades with their vector fields, and a digest of the luma and
chroma maps of every LUMSCALE / LUMSHIFT code), and the B-picture fixtures
(bframes/*.npz: the three pictures, the future anchor's fields, the B
parameters and every BAnalysis motion output of analyze_b_picture_mode).
The fixtures are data only.

    python tests/golden/make_golden.py              # everything
    python tests/golden/make_golden.py intensity    # intensity.json only
    python tests/golden/make_golden.py bframes      # bframes/*.npz only
    python tests/golden/make_golden.py scene        # scene.json only (scene_change_score)
    python tests/golden/make_golden.py motion       # motion_check.json only (has_good_motion_match,
                                                    # refine_qpel_block)
    python tests/golden/make_golden.py predict      # predict.json only (P prediction pictures)
    python tests/golden/make_golden.py predict_padded   # predict_padded.json only (padded pictures)
    python tests/golden/make_golden.py configs      # config_digests.json only (benchmark configurations'
                                                    # pairs at full size, tests/_config_cases.py)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from libvc1_amd import abi, synth  # noqa: E402
from _scene_cases import SCENE_CASES, scene_inputs  # noqa: E402
from _motion_cases import MOTION_CASES, motion_inputs  # noqa: E402
from _oracle import (oracle_analyze, reference_analyze, reference_analyze_b, reference_scene_change_score,  # noqa: E402
                     reference_has_good_motion_match, reference_refine_qpel_block, reference_predict_p_frame,
                     reference_predict_bi_frame,
                     reference_analyze_seeded, reference_estimate_intensity, reference_intensity_compensate)


def formula_pair(w, h, a, b):
    ys, xs = np.mgrid[0:h, 0:w]
    cur = ((xs * a[0] + ys * a[1] + (xs * ys) % a[2] + a[3]) & 255).astype(np.uint8)
    ref = ((xs * b[0] + ys * b[1] + (xs * ys) % b[2] + b[3]) & 255).astype(np.uint8)
    cu = np.full((h // 2, w // 2), 128, np.uint8)
    return synth.Picture(cur, cu, cu.copy()), synth.Picture(ref, cu.copy(), cu.copy())


def hidden_pair():
    # the 160x120 formula frames of the reference's hidden-padding test
    w, h = 160, 120

    def frame(phase):
        ys, xs = np.mgrid[0:h, 0:w]
        y = ((11 * xs + 19 * ys + phase + ((xs ^ ys) & 7) * 21) & 255).astype(np.uint8)
        i = np.arange((w // 2) * (h // 2))
        u = ((i * 13 + phase) & 255).astype(np.uint8).reshape(h // 2, w // 2)
        v = ((i * 7 + phase) & 255).astype(np.uint8).reshape(h // 2, w // 2)
        return synth.Picture(y, u, v)

    return frame(17), frame(9)


def gradient_pair(sx, sy):
    w, h = 64, 48
    ys, xs = np.mgrid[0:h, 0:w]
    ref = (20 + 3 * xs).astype(np.uint8)
    cur = ref[np.clip(ys + sy, 0, h - 1), np.clip(xs + sx, 0, w - 1)]
    cu = np.full((h // 2, w // 2), 128, np.uint8)
    return synth.Picture(cur, cu, cu.copy()), synth.Picture(ref, cu.copy(), cu.copy())


def small_cases():
    out = []
    c, r = formula_pair(48, 40, (13, 29, 37, 17), (7, 19, 53, 61))
    out.append(("formula_48x40_mixed", c, r, dict(pqindex=8, motion_search_range=32, motion_local_search_range=16,
                                                  mode=abi.MODE_MIXED_MV)))
    c, r = hidden_pair()
    out.append(("hidden_160x120_mixed", c, r, dict(pqindex=8, motion_search_range=6, motion_local_search_range=6,
                                                   mode=abi.MODE_MIXED_MV, allow_intra=0)))
    out.append(("hidden_160x120_bilinear", c, r, dict(pqindex=8, motion_search_range=6, motion_local_search_range=6,
                                                      mode=abi.MODE_ONE_MV_HPEL_BILINEAR, allow_intra=0)))
    c, r = gradient_pair(4, 0)
    out.append(("gradient_64x48_diamond", c, r, dict(pqindex=8, motion_search_range=12, motion_local_search_range=12,
                                                     me_algorithm=1, mode=abi.MODE_ONE_MV_QPEL)))
    d = synth.pan_rectangles(101, 3, width=176, height=144, max_pan_px=8, max_obj_px=24, frame_indices=[0, 2])
    out.append(("pan_176x144_long_range", d[2], d[0], dict(pqindex=8, motion_search_range=128,
                                                            motion_local_search_range=32, mode=abi.MODE_MIXED_MV)))
    rng = np.random.default_rng(5)
    base = np.clip(rng.normal(128, 40, size=(58, 74)), 0, 255).astype(np.uint8)
    cu = np.full((25, 33), 120, np.uint8)
    c = synth.Picture(base[4:54, 3:69].copy(), cu, cu.copy())
    r = synth.Picture(base[2:52, 5:71].copy(), cu.copy(), cu.copy())
    out.append(("odd_display_66x50", c, r, dict(display_width=65, display_height=49, pqindex=8,
                                                motion_search_range=64, motion_local_search_range=16,
                                                mode=abi.MODE_MIXED_MV)))
    return out


def seeded_cases():
    """(name, cur, ref, params of the seeded pass, seed) for the seeded core:
    seeds from the reference's own qpel pass without intra (as the
    reference's multi-pass analysis makes them), and one of arbitrary vectors
    partly outside the search bounds."""
    out = []
    d = synth.pan_rectangles(101, 3, width=176, height=144, max_pan_px=8, max_obj_px=24, frame_indices=[0, 2])
    base = dict(pqindex=8, motion_search_range=128, motion_local_search_range=32)
    q = abi.default_params(176, 144, mode=abi.MODE_ONE_MV_QPEL, allow_intra=0, **base)
    seed = reference_analyze(q, d[2], d[0])
    for tag, mode in (("hpel", abi.MODE_ONE_MV_HPEL), ("bilinear", abi.MODE_ONE_MV_HPEL_BILINEAR),
                      ("mixed", abi.MODE_MIXED_MV)):
        out.append((f"seeded_pan_176x144_{tag}", d[2], d[0], dict(base, mode=mode), seed))
    c, r = hidden_pair()
    hp = dict(pqindex=8, motion_search_range=6, motion_local_search_range=6)
    hseed = reference_analyze(abi.default_params(160, 120, mode=abi.MODE_ONE_MV_QPEL, allow_intra=0, **hp), c, r)
    out.append(("seeded_hidden_160x120_mixed", c, r, dict(hp, mode=abi.MODE_MIXED_MV, allow_intra=0), hseed))
    wild = abi.ResultArrays(176, 144)
    rng = np.random.default_rng(9)
    wild.mvs[...] = rng.integers(-700, 701, size=wild.mvs.shape)
    wild.distant_local_mae[...] = rng.uniform(0, 40, size=wild.mbs)
    wild.distant_candidate_mae[...] = -1.0
    wild.distant_match_decision[...] = rng.integers(0, 4, size=wild.mbs)
    out.append(("seeded_wild_176x144_mixed", d[2], d[0], dict(pqindex=8, motion_search_range=64,
                                                              motion_local_search_range=32, mode=abi.MODE_MIXED_MV),
                wild))
    return out


SEED_ARRAYS = ("mvs", "distant_local_mae", "distant_candidate_mae", "distant_match_decision")


def digest(res: abi.ResultArrays) -> str:
    h = hashlib.sha256()
    for k in abi.ResultArrays.ARRAYS:
        h.update(np.ascontiguousarray(getattr(res, k)).tobytes())
    for k in abi.ResultArrays.SCALARS:
        h.update(repr(getattr(res.struct, k)).encode())
    return h.hexdigest()


DIGEST_CASES = [
    # (name, generator, args, params)
    ("cif_translation_seed1_f1", "cif", dict(seed=1, frames=3, cur=1, ref=0),
     dict(pqindex=8, motion_search_range=64, motion_local_search_range=64, mode=abi.MODE_MIXED_MV)),
    ("cif_translation_seed1_f2_qpel", "cif", dict(seed=1, frames=3, cur=2, ref=1),
     dict(pqindex=8, motion_search_range=64, motion_local_search_range=64, mode=abi.MODE_ONE_MV_QPEL)),
    ("pan_640x360_r512", "pan", dict(seed=3, frames=8, width=640, height=360, cur=3, ref=0),
     dict(pqindex=8, motion_search_range=512, motion_local_search_range=32, mode=abi.MODE_MIXED_MV)),
]


def digest_inputs(kind, a):
    if kind == "cif":
        s = synth.cif_translation(a["seed"], frames=a["frames"])
        return s[a["cur"]], s[a["ref"]], 352, 288
    d = synth.pan_rectangles(a["seed"], a["frames"], width=a["width"], height=a["height"],
                             frame_indices=[a["ref"], a["cur"]])
    return d[a["cur"]], d[a["ref"]], a["width"], a["height"]


def intensity_inputs(case: dict):
    """(cur, ref, mvs) of an intensity case: a fade of frame 3 of a pan
    (or of frame 0 for the static kind) against frame 0."""
    w, h = case["size"]
    d = synth.pan_rectangles(case["seed"], 4, width=w, height=h, frame_indices=[0, 3])
    src = d[0] if case["kind"] == "static" else d[3]
    cur = synth.fade(src, case["gain"], case["offset"])
    return cur, d[0], np.asarray(case["mvs"], np.int32).reshape(-1, 2)


def intensity_map_table(compensate) -> str:
    """SHA-256 over the compensated planes of a 32x16 picture holding every
    luma value twice and 128 distinct chroma values per plane, for enabled
    = 0 and 1 and every (LUMSCALE, LUMSHIFT) code."""
    pic = synth.Picture((np.arange(512) % 256).reshape(16, 32).astype(np.uint8),
                        (np.arange(128).reshape(8, 16) * 2).astype(np.uint8),
                        (np.arange(128).reshape(8, 16) * 2 + 1).astype(np.uint8))
    hsh = hashlib.sha256()
    for en in (0, 1):
        for ls in range(64):
            for sh in range(64):
                for plane in compensate(pic, (en, ls, sh)):
                    hsh.update(plane.tobytes())
    return hsh.hexdigest()


def intensity_fixtures() -> dict:
    cases = []
    rng = np.random.default_rng(91)
    for k, (size, gain, offset, kind) in enumerate([
            ((320, 176), 0.8, 10.0, "static"), ((320, 176), 1.25, -20.0, "static"),
            ((320, 176), 1.0, 0.0, "static"), ((320, 176), 0.9, 6.0, "search"),
            ((320, 176), 1.1, -8.0, "search"), ((200, 120), 0.7, 25.0, "search"),
            ((200, 120), 1.3, -30.0, "random"), ((96, 64), 0.3, 0.0, "static"),
            ((96, 64), 1.6, -90.0, "static"), ((176, 144), 1.0, 12.0, "search")]):
        case = {"seed": 300 + k, "size": list(size), "gain": gain, "offset": offset, "kind": kind}
        mbw, mbh = abi.mb_grid(*size)
        case["mvs"] = [0] * (2 * mbw * mbh)
        cur, ref, _ = intensity_inputs(case)
        if kind == "search":
            q = abi.default_params(size[0], size[1], mode=abi.MODE_ONE_MV_QPEL, allow_intra=0, pqindex=8,
                                   motion_search_range=128)
            case["mvs"] = oracle_analyze(q, cur, ref).mvs.reshape(-1).tolist()
        elif kind == "random":
            case["mvs"] = rng.integers(-60, 60, size=2 * mbw * mbh).tolist()
        cur, ref, mvs = intensity_inputs(case)
        case["expected"] = list(reference_estimate_intensity(cur, ref, mvs))
        cases.append(case)
        print(f"intensity case {k}: {kind} {size} gain {gain} offset {offset} -> {case['expected']}")
    return {"estimate": cases,
            "map_sha256": intensity_map_table(lambda pic, ic: reference_intensity_compensate(pic, ic, ic[0]))}


def b_cases():
    """(name, cur, past, future, params overrides, anchor mvs, anchor 4mv,
    B parameters) of the B-picture fixtures."""
    out = []
    cif = synth.cif_translation(41, frames=4, width=176, height=144, max_px=6)
    base = dict(pqindex=8, motion_search_range=64, motion_local_search_range=32)
    anchor = reference_analyze(abi.default_params(176, 144, mode=abi.MODE_MIXED_MV, **base), cif[3], cif[0])
    out.append(("b_cif_qpel_third", cif[1], cif[0], cif[3], dict(base, mode=abi.MODE_ONE_MV_QPEL),
                anchor.mvs, anchor.use_4mv, dict(fraction=(1, 3), allow_intra=1)))
    out.append(("b_cif_bilinear_twothirds_ic", cif[2], cif[0], cif[3], dict(base, mode=abi.MODE_ONE_MV_HPEL_BILINEAR),
                anchor.mvs, anchor.use_4mv, dict(fraction=(2, 3), allow_intra=1, intensity=(1, 30, 4))))
    other = synth.pan_rectangles(42, 2, width=176, height=144, frame_indices=[1])[1]
    out.append(("b_cut_intra_half", other, cif[0], cif[3], dict(base, mode=abi.MODE_ONE_MV_QPEL),
                anchor.mvs, anchor.use_4mv, dict(fraction=(1, 2), allow_intra=1)))
    c, r = hidden_pair()
    rng = np.random.default_rng(43)
    mbs = 10 * 8
    out.append(("b_hidden_hpel_half", c, r, c, dict(base, mode=abi.MODE_ONE_MV_HPEL, display_width=154,
                                                    display_height=114),
                rng.integers(-40, 40, size=(mbs, 2)).astype(np.int32), rng.integers(0, 2, size=mbs).astype(np.uint8),
                dict(fraction=(1, 2), allow_intra=1)))
    # a linear pan at (+7, -5) qpel per frame: the scaled anchor vectors are exact
    rng = np.random.default_rng(45)
    tex = synth.blur_reflect(rng.uniform(0, 255, size=(144, 176)), 2.0)
    tex = np.clip(np.rint(tex), 0, 255).astype(np.uint8)
    grey = np.full((72, 88), 128, np.uint8)
    lin = [synth.Picture(synth.vc1_luma_translate(tex, 7 * k, -5 * k, wrap=True), grey, grey) for k in range(4)]
    lin_anchor = reference_analyze(abi.default_params(176, 144, mode=abi.MODE_ONE_MV_QPEL, allow_intra=0, **base),
                                   lin[3], lin[0])
    out.append(("b_linear_direct_third", lin[1], lin[0], lin[3], dict(base, mode=abi.MODE_ONE_MV_QPEL),
                lin_anchor.mvs, lin_anchor.use_4mv, dict(fraction=(1, 3), allow_intra=1)))
    small = synth.pan_rectangles(44, 4, width=128, height=96, frame_indices=[0, 1, 3])
    out.append(("b_zero_range_third", small[1], small[0], small[3], dict(base, motion_search_range=0),
                np.zeros((48, 2), np.int32), np.zeros(48, np.uint8), dict(fraction=(1, 3), allow_intra=1)))
    return out


def b_fixtures() -> None:
    os.makedirs(os.path.join(HERE, "bframes"), exist_ok=True)
    for name, cur, past, fut, over, amv, a4, bkw in b_cases():
        h, w = cur.y.shape
        p = abi.default_params(w, h, **over)
        binp = abi.BInputs(amv, a4, **bkw)
        res = reference_analyze_b(p, cur, past, fut, binp)
        arrays = {f"out_{k}": getattr(res, k) for k in abi.BResultArrays.ARRAYS}
        np.savez_compressed(os.path.join(HERE, "bframes", f"{name}.npz"),
                            cur_y=cur.y, cur_u=cur.u, cur_v=cur.v, past_y=past.y, past_u=past.u, past_v=past.v,
                            fut_y=fut.y, fut_u=fut.u, fut_v=fut.v, anchor_mvs=np.asarray(amv, np.int32),
                            anchor_4mv=np.asarray(a4, np.uint8), params=json.dumps(over), bparams=json.dumps(bkw),
                            scalars=json.dumps(res.scalars()), **arrays)
        print(f"wrote bframes/{name}: {res.scalars()}")


def scene_fixtures() -> list:
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=128, motion_local_search_range=32)
    out = []
    for name, fade, thr in SCENE_CASES:
        cur, ref = scene_inputs(name)
        v = reference_scene_change_score(p, cur, ref, fade, thr)
        out.append({"case": name, "fade": fade, "threshold": thr, "score": v,
                    "score_bits": int(np.float64(v).view(np.uint64))})
        print(f"scene {name} fade={fade} thr={thr}: {v}")
    return out


def motion_fixtures() -> dict:
    from _motion_cases import W, H
    good = []
    for name, fade, lr, rng in MOTION_CASES:
        cur, ref = motion_inputs(name)
        p = abi.default_params(W, H, pqindex=8, motion_search_range=rng, long_range_search_mode=lr)
        v = reference_has_good_motion_match(p, cur, ref, fade)
        good.append({"case": name, "fade": fade, "long_range_search_mode": lr, "range": rng, "good": bool(v)})
        print(f"motion {name} fade={fade} lr={lr} range={rng}: {v}")
    refine = []
    cur, ref = motion_inputs("pan")
    rs = np.random.default_rng(78)
    for t in range(240):
        fast, sub = t % 2, 0 if t % 11 == 0 else 1
        bw = 8 if t % 3 == 0 else 16
        case = {"fast": fast, "subpel": sub, "subpel_range": int(rs.choice([1024, 8, 2])),
                "display_width": W - 6 if t % 4 == 1 else 0, "x0": int(rs.integers(0, W // bw)) * bw,
                "y0": int(rs.integers(0, H // bw)) * bw, "bw": bw,
                "seed": [int(rs.integers(-90, 90)), int(rs.integers(-90, 90))],
                "range": int(rs.choice([0, 3, 32, 256])),
                "known": 2**64 - 1 if t % 2 else int(rs.integers(0, 6000))}
        p = abi.default_params(W, H, pqindex=8, fast_subpixel_refinement=fast, subpixel_refinement=sub,
                               subpixel_refinement_range=case["subpel_range"], display_width=case["display_width"])
        case["expected"] = list(reference_refine_qpel_block(p, cur, ref, case["x0"], case["y0"], bw, bw,
                                                            tuple(case["seed"]), case["range"], case["known"]))
        refine.append(case)
    return {"good": good, "refine": refine}


def predict_fixtures() -> list:
    """Vector fields from the reference's analysis of generated pairs (some
    4-MV vectors replaced by wild ones) and SHA-256 digests of the
    reference's prediction planes."""
    from _predict_cases import PREDICT_CASES, predict_inputs
    out = []
    for name, w, h, mode in PREDICT_CASES:
        ref, p, res = predict_inputs(name, w, h, mode, reference_analyze)
        y, u, v = reference_predict_p_frame(p, ref, res, 1)
        out.append({"case": name, "width": w, "height": h, "mode": mode, "mvs": res.mvs.reshape(-1).tolist(),
                    "block_mvs": res.block_mvs.reshape(-1).tolist(), "use_4mv": res.use_4mv.tolist(),
                    "sha256": hashlib.sha256(y.tobytes() + u.tobytes() + v.tobytes()).hexdigest()})
        print(f"predict {name} {w}x{h} mode {mode}: 4-MV {int(res.use_4mv.sum())}")
    return out


def predict_bi_fixtures() -> list:
    """Random forward/backward fields and digests of the reference's
    bi-directional prediction pictures."""
    from _predict_cases import BI_CASES, bi_inputs
    out = []
    for name, w, h, mode in BI_CASES:
        past, fut, fmv, bmv = bi_inputs(name, w, h, mode)
        y, u, v = reference_predict_bi_frame(abi.default_params(w, h, mode=mode), past, fut, fmv, bmv, 1)
        out.append({"case": name, "width": w, "height": h, "mode": mode,
                    "sha256": hashlib.sha256(y.tobytes() + u.tobytes() + v.tobytes()).hexdigest()})
        print(f"predict bi {name} {w}x{h} mode {mode}")
    return out


def predict_padded_fixtures() -> list:
    """Digests of the reference's padded (loop-reconstruction) prediction
    pictures, motion_compensate_mb_padded / _4mv_padded and
    motion_compensate_bi_mb_padded, its scalar and SIMD paths agreeing."""
    from _predict_cases import BI_PADDED_CASES, PADDED_CASES, bi_padded_inputs, padded_inputs
    from _oracle import reference_predict_bi_frame_padded, reference_predict_p_frame_padded
    out = []
    for name, w, h, mode, padded in PADDED_CASES:
        ref, res = padded_inputs(name, w, h, mode, padded)
        p = abi.default_params(w, h, mode=mode)
        shas = {hashlib.sha256(b"".join(x.tobytes() for x in reference_predict_p_frame_padded(p, ref, res, simd)))
                .hexdigest() for simd in (0, 1)}
        assert len(shas) == 1, f"{name}: scalar and SIMD paths differ"
        out.append({"case": name, "kind": "p", "width": w, "height": h, "mode": mode, "sha256": shas.pop()})
        print(f"predict padded {name}: ref {ref.width}x{ref.height}")
    for name, w, h, mode, padded in BI_PADDED_CASES:
        past, fut, f, b = bi_padded_inputs(name, w, h, mode, padded)
        p = abi.default_params(w, h, mode=mode)
        shas = {hashlib.sha256(b"".join(x.tobytes() for x in reference_predict_bi_frame_padded(p, past, fut, f, b,
                                                                                                 simd)))
                .hexdigest() for simd in (0, 1)}
        assert len(shas) == 1, f"{name}: scalar and SIMD paths differ"
        out.append({"case": name, "kind": "bi", "width": w, "height": h, "mode": mode, "sha256": shas.pop()})
        print(f"predict padded {name}: past {past.width}x{past.height}, future {fut.width}x{fut.height}")
    return out


def _config_case(job):
    """(key, sha256, scalars) of one benchmark-configuration pair analysed by
    the compiled reference (run in a worker process)."""
    import _config_cases as cc
    kind, arg, label = job
    if kind == "cfg2":
        d = cc.cfg2_frames()
        c, r = d[arg[0]], d[arg[1]]
        p = cc.cfg2_params()
        key = f"cfg2/{arg[0]}-{arg[1]}"
    elif kind == "cfg3":
        st = cc.cfg3_stream(arg)
        w, h = cc.CFG3_RASTERS[arg]
        c, r = st[label[0]], st[label[1]]
        p = cc.cfg3_params(w, h)
        key = f"cfg3/{w}x{h}/{label[0]}-{label[1]}"
    else:
        seed, pairs = cc.cfg5_stream(arg)
        c, r = pairs[label]
        p = cc.cfg5_params()
        key = f"cfg5/{seed}/{label}"
    res = reference_analyze(p, c, r)
    return key, cc.digest(res), res.scalars()


def config_fixtures() -> dict:
    """Digests of the reference's analysis of the benchmark configurations'
    pairs (tests/_config_cases.py), written to config_digests.json."""
    import _config_cases as cc
    jobs = [("cfg2", pr, None) for pr in cc.CFG2_PAIRS]
    jobs += [("cfg3", k, pr) for k in range(len(cc.CFG3_RASTERS)) for pr in ((1, 0), (2, 1))]
    jobs += [("cfg5", 0, "cut"), ("cfg5", 0, "duplicate"), ("cfg5", 1, "after_cut"), ("cfg5", 2, "cut")]
    out = {}
    for key, sha, scal in cc.generate_parallel(_config_case, jobs, int(os.environ.get("GOLDEN_WORKERS", "8"))):
        out[key] = {"sha256": sha, "scalars": scal}
        print(f"config {key}: {sha[:16]} {scal}", flush=True)
    return out


def main() -> int:
    if sys.argv[1:] == ["configs"]:
        with open(os.path.join(HERE, "config_digests.json"), "w") as f:
            json.dump(config_fixtures(), f, indent=1, sort_keys=True)
        return 0
    if sys.argv[1:] == ["predict"]:
        with open(os.path.join(HERE, "predict.json"), "w") as f:
            json.dump(predict_fixtures(), f)
        with open(os.path.join(HERE, "predict_bi.json"), "w") as f:
            json.dump(predict_bi_fixtures(), f, indent=1)
        with open(os.path.join(HERE, "predict_padded.json"), "w") as f:
            json.dump(predict_padded_fixtures(), f, indent=1)
        return 0
    if sys.argv[1:] == ["predict_padded"]:
        with open(os.path.join(HERE, "predict_padded.json"), "w") as f:
            json.dump(predict_padded_fixtures(), f, indent=1)
        return 0
    if sys.argv[1:] == ["motion"]:
        with open(os.path.join(HERE, "motion_check.json"), "w") as f:
            json.dump(motion_fixtures(), f)
        return 0
    if sys.argv[1:] == ["scene"]:
        with open(os.path.join(HERE, "scene.json"), "w") as f:
            json.dump(scene_fixtures(), f, indent=1)
        return 0
    if sys.argv[1:] == ["bframes"]:
        b_fixtures()
        return 0
    if sys.argv[1:] == ["intensity"]:
        with open(os.path.join(HERE, "intensity.json"), "w") as f:
            json.dump(intensity_fixtures(), f)
        return 0
    for name, c, r, over in small_cases():
        h, w = c.y.shape
        p = abi.default_params(w, h, **over)
        res = reference_analyze(p, c, r)
        arrays = {f"out_{k}": getattr(res, k) for k in abi.ResultArrays.ARRAYS}
        scal = res.scalars()
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), cur_y=c.y, cur_u=c.u, cur_v=c.v, ref_y=r.y,
                            ref_u=r.u, ref_v=r.v, params=json.dumps(over), scalars=json.dumps(scal), **arrays)
        print(f"wrote {name}: 4mv={int(res.use_4mv.sum())} intra={int(res.use_intra.sum())}")
    os.makedirs(os.path.join(HERE, "seeded"), exist_ok=True)
    for name, c, r, over, seed in seeded_cases():
        h, w = c.y.shape
        p = abi.default_params(w, h, **over)
        res = reference_analyze_seeded(p, c, r, seed)
        arrays = {f"out_{k}": getattr(res, k) for k in abi.ResultArrays.ARRAYS}
        arrays.update({f"seed_{k}": getattr(seed, k) for k in SEED_ARRAYS})
        np.savez_compressed(os.path.join(HERE, "seeded", f"{name}.npz"), cur_y=c.y, cur_u=c.u, cur_v=c.v, ref_y=r.y,
                            ref_u=r.u, ref_v=r.v, params=json.dumps(over), scalars=json.dumps(res.scalars()),
                            **arrays)
        print(f"wrote seeded/{name}: 4mv={int(res.use_4mv.sum())} intra={int(res.use_intra.sum())}")
    digests = []
    for name, kind, a, over in DIGEST_CASES:
        c, r, w, h = digest_inputs(kind, a)
        p = abi.default_params(w, h, **over)
        res = reference_analyze(p, c, r)
        digests.append({"name": name, "kind": kind, "args": a, "params": over, "sha256": digest(res),
                        "scalars": res.scalars()})
        print(f"digest {name}: {digests[-1]['sha256'][:16]}")
    with open(os.path.join(HERE, "analyze_digests.json"), "w") as f:
        json.dump(digests, f, indent=1)
    with open(os.path.join(HERE, "intensity.json"), "w") as f:
        json.dump(intensity_fixtures(), f)
    b_fixtures()
    with open(os.path.join(HERE, "scene.json"), "w") as f:
        json.dump(scene_fixtures(), f, indent=1)
    with open(os.path.join(HERE, "motion_check.json"), "w") as f:
        json.dump(motion_fixtures(), f)
    with open(os.path.join(HERE, "predict.json"), "w") as f:
        json.dump(predict_fixtures(), f)
    with open(os.path.join(HERE, "predict_bi.json"), "w") as f:
        json.dump(predict_bi_fixtures(), f, indent=1)
    with open(os.path.join(HERE, "predict_padded.json"), "w") as f:
        json.dump(predict_padded_fixtures(), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
