"""Inputs of the motion-failure cases (tests and fixture writer)."""
import numpy as np

from libvc1_amd import synth

W, H = 240, 136


def motion_inputs(name: str):
    """(cur, ref) of a has_good_motion_match case at 240x136."""
    d = synth.pan_rectangles(12, 4, width=W, height=H, frame_indices=[0, 1, 3])
    rng = np.random.default_rng(77)
    grey = np.full((H // 2, W // 2), 128, np.uint8)
    noise = rng.integers(0, 256, size=(H, W)).astype(np.uint8)
    if name == "pan":
        return d[1], d[0]
    if name == "far":
        return d[3], d[0]
    if name == "fade":
        return synth.fade(d[1], 0.5, 60.0), d[0]
    if name == "noise_vs_flat":
        # nothing to match: every macroblock's best SAD is its zero-vector SAD
        return synth.Picture(noise, grey, grey), synth.Picture(np.full((H, W), 128, np.uint8), grey, grey)
    if name == "late_match":
        # unrelated noise except the last macroblock row, which is the
        # reference's content moved by (3, -2) px
        ref = synth.Picture(rng.integers(0, 256, size=(H, W)).astype(np.uint8), grey, grey)
        cur = noise.copy()
        cur[128:136, 0:W - 3] = ref.y[126:134, 3:W]
        return synth.Picture(cur, grey, grey), ref
    raise KeyError(name)


MOTION_CASES = [(n, fade, lr, rng) for n in ("pan", "far", "fade", "noise_vs_flat", "late_match")
                for fade, lr, rng in ((1, 0, 256), (0, 1, 64), (1, 2, 256), (1, 0, 0))]
