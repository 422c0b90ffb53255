"""What the trained proposal network gains over the method without it.

The method without the network is the static shortlist: the K luma and K
chroma strengths (K = AV1PF_SHORTLIST) most often best in the training
labels, with every LR unit searched.  On held-out moving-rectangle frames
(never seen in training):
  - the network's per-frame list keeps more of the per-block CDEF luma gain
    (unfiltered SSE minus best SSE, each block taking its best listed
    strength) than the static list, at K and at 1 strength;
  - the whole learned search loses less than 0.05 dB luma PSNR against the
    exhaustive search, for CDEF and for LR.
Losses are exact: the full 64-strength MSE table prices any decision and the
LR search's exact scores price the unit choices.
"""
import json
import os

import numpy as np
import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_trained_network_beats_static_shortlist(bd):
    from goav1_amd import learned
    from goav1_amd.lib import PostFilterRunner, ProposalNet
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    dc = json.load(open(os.path.join(root, "tests", "golden", "dc_qlookup_8bit.json")))["dc"]
    meta = json.load(open(os.path.splitext(learned.DEFAULT_WEIGHTS)[0] + ".json"))
    static = np.array(meta["static_shortlist"], np.int32)
    rank = np.argsort(-np.asarray(meta["label_counts"][0]), kind="stable")
    runner = PostFilterRunner()
    net = ProposalNet(runner, max_batch=64)
    try:
        net.set_params(np.load(learned.DEFAULT_WEIGHTS))
        frames = learned.held_out(6, 1920, 1080, bd, seed=770000)
        print(f"learned vs static shortlist on 6 held-out 1920x1080 {bd}-bit frames", flush=True)
        res = learned.evaluate(runner, net, frames, lambda q, b: learned.lambda_for(dc, q, b), static, reps=1,
                               static_rank=rank)
    finally:
        net.close()
        runner.close()
    print(json.dumps({k: v for k, v in res.items() if k != "rows"}), flush=True)
    kept = res["luma_block_gain_kept"]
    assert kept["learned"] > kept["static"], kept
    top = res["topk_gain_kept"]
    assert top["learned"][1] > top["static"][1], top
    assert res["cdef_psnr_loss_learned"] < 0.05 and res["lr_psnr_loss_learned"] < 0.05
