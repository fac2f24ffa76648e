"""Learned seeding with the network as a confidence gate only
(vc1me_set_learned_candidates(0)): the step equals the proposal-seeded step
given the network's own confidences and zero candidate offsets, and the
plain step is untouched by the setting."""
import numpy as np
import pytest

from libvc1_amd import abi, learned, synth
from libvc1_amd.backend import ProposalNet
from _oracle import describe_diff, oracle_analyze

pytestmark = pytest.mark.gpu


def test_gate_only_equals_zero_candidate_proposals(backend):
    d = synth.pan_rectangles(120, 4, width=320, height=176, frame_indices=[0, 3])
    p = abi.default_params(320, 176, pqindex=8, motion_search_range=256, mode=abi.MODE_MIXED_MV)
    fc, fr = backend.upload(d[3]), backend.upload(d[0])
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    try:
        jobs, plist = [(fc, fr)], [p]
        plain = backend.analyze_step(jobs, plist)[0]
        glob = np.array([[plain.struct.global_dx, plain.struct.global_dy]], np.int32)
        outs = net.forward_frames(jobs, glob)
        mbw, mbh = abi.mb_grid(320, 176)
        props = [np.concatenate([np.zeros((mbh, mbw, 4), np.float32), o[..., 4:5]], axis=2) for o in outs]
        print("[case] learned gate only vs zero-candidate proposals", flush=True)
        want = backend.analyze_step_proposals(jobs, plist, props, 0.5)[0]
        backend.set_learned_candidates(False)
        try:
            got = backend.analyze_step_learned(net, jobs, plist, 0.5)[0]
            again = backend.analyze_step(jobs, plist)[0]
        finally:
            backend.set_learned_candidates(True)
        assert not got.diff(want), describe_diff(got, want, 320)
        ref = oracle_analyze(p, d[3], d[0])
        assert not again.diff(ref), describe_diff(again, ref, 320)
    finally:
        net.close()
        fc.release()
        fr.release()
