"""Learned seeding with the network as a confidence gate only
(vc1me_set_learned_candidates(0)): the step equals the proposal-seeded step
given the network's own confidences and zero candidate offsets, and the
plain step is untouched by the setting."""
import numpy as np
import pytest

from libvc1_amd import abi, learned
from libvc1_amd.backend import ProposalNet
import _config_cases as cc
from _oracle import describe_diff, oracle_analyze

pytestmark = pytest.mark.gpu


def test_gate_only_equals_zero_candidate_proposals(backend):
    # the 1080p benchmark pair: the network (trained at 1080p) is confident
    # on most of its macroblocks, unlike on small rasters
    d = cc.cfg2_frames()
    p = cc.cfg2_params()
    fc, fr = backend.upload(d[3]), backend.upload(d[0])
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    try:
        jobs, plist = [(fc, fr)], [p]
        plain = backend.analyze_step(jobs, plist)[0]
        glob = np.array([[plain.struct.global_dx, plain.struct.global_dy]], np.int32)
        outs = net.forward_frames(jobs, glob)
        mbw, mbh = abi.mb_grid(p.width, p.height)
        props = [np.concatenate([np.zeros((mbh, mbw, 4), np.float32), o[..., 4:5]], axis=2) for o in outs]
        # the gate must pass somewhere (logit >= 0 is confidence >= 0.5), or
        # the comparison below would hold trivially
        assert (outs[0][..., 4] >= 0).any(), "no macroblock passes the confidence gate"
        print("[case] learned gate only vs zero-candidate proposals", flush=True)
        want = backend.analyze_step_proposals(jobs, plist, props, 0.5)[0]
        backend.set_learned_candidates(False)
        try:
            got = backend.analyze_step_learned(net, jobs, plist, 0.5)[0]
            seeded = backend.work().seeded_macroblocks
            again = backend.analyze_step(jobs, plist)[0]
        finally:
            backend.set_learned_candidates(True)
        assert seeded > 0, "the gate-only step seeded no macroblock"
        assert not got.diff(want), describe_diff(got, want, p.width)
        ref = oracle_analyze(p, d[3], d[0])
        assert not again.diff(ref), describe_diff(again, ref, p.width)
    finally:
        net.close()
        fc.release()
        fr.release()


def test_proposals_step_ignores_the_gate_only_setting(backend):
    """vc1me_set_learned_candidates governs the learned step only: the
    proposal step uses the caller's vectors whatever the setting."""
    d = cc.cfg2_frames()
    p = cc.cfg2_params()
    fc, fr = backend.upload(d[48]), backend.upload(d[45])
    net = ProposalNet(backend)
    net.set_params(learned.load_weights())
    try:
        jobs, plist = [(fc, fr)], [p]
        plain = backend.analyze_step(jobs, plist)[0]
        glob = np.array([[plain.struct.global_dx, plain.struct.global_dy]], np.int32)
        outs = net.forward_frames(jobs, glob)
        assert (outs[0][..., 4] >= 0).any(), "no macroblock passes the confidence gate"
        print("[case] proposal step under both candidate settings", flush=True)
        on = backend.analyze_step_proposals(jobs, plist, outs, 0.5)[0]
        backend.set_learned_candidates(False)
        try:
            off = backend.analyze_step_proposals(jobs, plist, outs, 0.5)[0]
        finally:
            backend.set_learned_candidates(True)
        assert not off.diff(on), describe_diff(off, on, p.width)
    finally:
        net.close()
        fc.release()
        fr.release()
