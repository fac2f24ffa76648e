"""Inputs of the scene-change-score cases (tests and fixture writer)."""
import numpy as np

from libvc1_amd import synth


def scene_inputs(name: str):
    """(cur, ref) of a scene-score case at 320x176."""
    W, H = 320, 176
    d = synth.pan_rectangles(90, 4, width=W, height=H, frame_indices=[0, 1, 3])
    if name == "consecutive":
        return d[1], d[0]
    if name == "cut":
        return synth.pan_rectangles(91, 2, width=W, height=H, frame_indices=[1])[1], d[0]
    if name == "fade":
        return synth.fade(d[1], 0.6, 30.0), d[0]
    if name == "zoom_sharp":
        rng = np.random.default_rng(2)
        y = np.clip(np.rint(synth.blur_reflect(rng.integers(0, 256, size=(H, W)).astype(float), 1.2)), 0, 255)
        sharp = synth.Picture(y.astype(np.uint8), d[0].u, d[0].v)
        return synth.zoom(sharp, 1.04), sharp
    if name == "zoom_pan":
        return synth.zoom(d[0], 1.05), d[0]
    raise KeyError(name)


SCENE_CASES = [(n, fade, thr) for n in ("consecutive", "cut", "fade", "zoom_sharp", "zoom_pan")
               for fade, thr in ((1, 28.0), (0, 28.0), (1, 6.0))]
