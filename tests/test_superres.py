"""Super-resolution: the oracle pinned, then the HIP stage against it.

Oracle pins (CPU):
  - goav1's known-answer rows (superres_test.go, and the libaom-captured row of
    superres_aligned_libaom_test.go whose source span runs past the coded
    width), which fix step, initial phase, taps, rounding and edge clamps;
  - the filter table's invariants: every phase sums to 128, phase 0 is the
    identity, phase 64 - k is phase k mirrored;
  - constant rows stay constant at every bit depth and ratio;
  - the chain's upscaled LR context rows: with CDEF off the deblocked rows are
    the frame rows, so LR after the upscale must equal LR run as a plain
    chain on the upscaled frame; film grain likewise runs on the upscaled frame.
GPU (through the C ABI): av1pf_superres and the chain with super-resolution
(mixed with plain frames in one batch, graph replays) give the oracle's
samples exactly.
"""
import json
import os

import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_vectors.json")))["superres_upscale_row"]


# ---------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: c["name"])
def test_oracle_known_answer_rows(case):
    got = orc.superres_row(case["src"], case["coded_w"], case["dst_w"], case["bd"])
    assert got.tolist() == case["want"]


def test_oracle_filter_table_invariants():
    t = orc.superres_filter().astype(np.int32)
    assert (t.sum(axis=1) == 128).all()
    assert t[0].tolist() == [0, 0, 0, 128, 0, 0, 0, 0]
    for k in range(1, 64):
        assert t[64 - k].tolist() == t[k][::-1].tolist(), k
    # the centre taps dominate and move monotonically from sample 3 to sample 4
    assert (np.diff(t[:, 3]) <= 0).all() and (np.diff(t[1:, 4]) >= 0).all()


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_oracle_constant_rows(bd):
    rng = np.random.default_rng(bd)
    for _ in range(40):
        cw = int(rng.integers(4, 300))
        uw = int(rng.integers(cw + 1, 2 * cw + 1))
        v = int(rng.integers(0, 1 << bd))
        srcw = (cw + 7) & ~7
        assert (orc.superres_row([v] * srcw, cw, uw, bd) == v).all()


def test_oracle_rejects_invalid_geometry():
    with pytest.raises(ValueError):
        orc.superres_row([0] * 8, 8, 8, 8)        # not wider
    with pytest.raises(ValueError):
        orc.superres_row([0] * 8, 8, 16, 7)       # bit depth


def _frame_sr(seed, up_w, h, denom, bd, ss, mono=False, grain=False, stages=None):
    return synth.superres_frame(seed=seed, up_w=up_w, h=h, denom=denom, bd=bd, ssx=ss[0], ssy=ss[1], mono=mono,
                                grain=grain, stages=stages)


@pytest.mark.parametrize("bd", [8, 10])
def test_oracle_chain_lr_rows_match_plain_chain_without_cdef(bd):
    fr = _frame_sr(11 + bd, 200, 136, 13, bd, (1, 1),
                   stages=abi.STAGE_DEBLOCK | abi.STAGE_SUPERRES | abi.STAGE_LR)
    fr.refresh = False   # the partial chains below must leave the coded frame's padding as deblock wrote it
    _, _, up = orc.run_chain(fr, with_upscaled=True)
    coded, _ = orc.run_chain(fr, stages=abi.STAGE_DEBLOCK)
    plain = orc.superres(coded, fr.upscaled)
    plain.lr, plain.stages = fr.lr, abi.STAGE_LR
    want, _ = orc.run_chain(plain)
    assert orc.plane_md5(up) == orc.plane_md5(want)
    # and the LR stage changed something, so the comparison is not vacuous
    assert orc.plane_md5(up) != orc.plane_md5(plain)


def test_oracle_chain_grain_on_upscaled_frame():
    fr = _frame_sr(5, 136, 72, 16, 10, (1, 1), grain=True,
                   stages=abi.STAGE_CDEF | abi.STAGE_SUPERRES | abi.STAGE_FILMGRAIN)
    fr.refresh = False
    out, disp, up = orc.run_chain(fr, with_upscaled=True)
    coded, _ = orc.run_chain(fr, stages=abi.STAGE_CDEF)
    assert orc.plane_md5(out) == orc.plane_md5(coded)   # the coded frame keeps the CDEF output
    plain = orc.superres(coded, fr.upscaled)
    assert orc.plane_md5(up) == orc.plane_md5(plain)
    plain.grain, plain.stages = fr.grain, abi.STAGE_FILMGRAIN
    _, want = orc.run_chain(plain)
    assert orc.plane_md5(disp) == orc.plane_md5(want)


# ---------------------------------------------------------------- GPU parity
def _dev_out(fr, dev):
    out = fr.copy_pixels()
    for pl, a in zip(out.planes, dev.to_numpy()):
        pl.pix[:] = a
    return out


def _assert_same(a, b, what, full=True):
    """Visible samples by per-plane MD5 and, with full=True, the whole
    allocation (padding and edge-replicated borders included)."""
    ma, mb = orc.plane_md5(a), orc.plane_md5(b)
    msgs = []
    for i, (pa, pb) in enumerate(zip(a.planes, b.planes)):
        va, vb = (pa.pix, pb.pix) if full else (pa.pix[:pa.h, :pa.w], pb.pix[:pb.h, :pb.w])
        diff = np.argwhere(va != vb)
        if len(diff):
            y, x = diff[0]
            where = "visible" if y < pa.h and x < pa.w else "padding"
            msgs.append(f"plane {i}: {len(diff)} diffs, first at y={y} x={x} ({where}): gpu={va[y, x]} oracle={vb[y, x]}")
    if msgs or ma != mb:
        raise AssertionError(f"{what}: " + "; ".join(msgs))


@pytest.fixture(scope="module", params=["staged", "fused"])
def runner(request):
    """A handle per pipeline mode: every test of the module runs once with one
    launch per stage and once with the fused chain launch."""
    from goav1_amd.lib import PIPELINE_FUSED, PostFilterRunner
    r = PostFilterRunner()
    if request.param == "fused":
        r.set_pipeline(PIPELINE_FUSED)
    yield r
    r.close()


def _gpu_chain(runner, frames, batch=True):
    """Run each frame's request on the device (one batch or frame by frame);
    returns per frame (coded, upscaled or None, display or None)."""
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    devs, ups, disps, sides = [], [], [], []
    for fr in frames:
        d = DeviceFrame(fr)
        u = DeviceFrame(fr.upscaled) if fr.stages & abi.STAGE_SUPERRES else None
        disp = None
        if fr.stages & abi.STAGE_FILMGRAIN:
            disp = DeviceFrame.empty_like(fr.upscaled if u is not None else fr)
        devs.append(d)
        ups.append(u)
        disps.append(disp)
        sides.append(DeviceSideData(fr, display=disp, upscaled=u))
    if batch:
        runner.apply_batch(devs, sides)
    else:
        for d, s in zip(devs, sides):
            runner.apply(d, s)
    runner.sync()
    torch.cuda.synchronize()
    res = []
    for fr, d, u, disp in zip(frames, devs, ups, disps):
        res.append((_dev_out(fr, d), _dev_out(fr.upscaled, u) if u is not None else None,
                    _dev_out(fr.upscaled if u is not None else fr, disp) if disp is not None else None))
    return res


def _check(frames, got):
    for i, (fr, (coded, up, disp)) in enumerate(zip(frames, got)):
        out, wdisp, wup = orc.run_chain(fr, with_upscaled=True)
        tag = f"frame {i} {fr.w}->{fr.upscaled.w if fr.upscaled else fr.w}x{fr.h} {fr.bd}b stages {fr.stages:#x}"
        _assert_same(coded, out, tag + " coded")
        if wup is not None:
            _assert_same(up, wup, tag + " upscaled")
        if wdisp is not None:
            _assert_same(disp, wdisp, tag + " display", full=False)


SR_GEOMS = [  # (upscaled width, height, denominator, bit depth, (ssx, ssy), mono)
    (200, 136, 9, 8, (1, 1), False), (200, 136, 16, 10, (1, 1), False), (333, 77, 13, 10, (1, 0), False),
    (136, 72, 11, 8, (0, 0), False), (90, 50, 15, 12, (1, 1), False), (24, 20, 16, 8, (1, 1), False),
    (130, 66, 12, 10, (1, 1), True), (1000, 40, 10, 8, (1, 1), False),
]


@pytest.mark.gpu
@pytest.mark.parametrize("g", SR_GEOMS, ids=lambda g: f"{g[0]}x{g[1]}_d{g[2]}_{g[3]}b_ss{g[4][0]}{g[4][1]}{'_mono' if g[5] else ''}")
def test_superres_entry_point(runner, g):
    from goav1_amd.lib import DeviceFrame
    import torch
    up_w, h, denom, bd, ss, mono = g
    fr = _frame_sr(101, up_w, h, denom, bd, ss, mono=mono, stages=abi.STAGE_SUPERRES)
    print(f"superres {fr.w}->{up_w}x{h} {bd}b", flush=True)
    d, u = DeviceFrame(fr), DeviceFrame(fr.upscaled)
    runner.superres(d, u)
    runner.sync()
    torch.cuda.synchronize()
    _assert_same(_dev_out(fr.upscaled, u), orc.superres(fr, fr.upscaled), "superres")


@pytest.mark.gpu
@pytest.mark.parametrize("g", SR_GEOMS, ids=lambda g: f"{g[0]}x{g[1]}_d{g[2]}_{g[3]}b_ss{g[4][0]}{g[4][1]}{'_mono' if g[5] else ''}")
def test_superres_chain(runner, g):
    up_w, h, denom, bd, ss, mono = g
    fr = _frame_sr(202, up_w, h, denom, bd, ss, mono=mono, grain=not mono)
    print(f"superres chain {fr.w}->{up_w}x{h} {bd}b stages {fr.stages:#x}", flush=True)
    _check([fr], _gpu_chain(runner, [fr], batch=False))


@pytest.mark.gpu
@pytest.mark.parametrize("pipeline", [0, 1], ids=["staged", "fused"])
def test_superres_chain_1080p10(pipeline):
    from goav1_amd.lib import PostFilterRunner
    fr = _frame_sr(303, 1920, 1080, 12, 10, (1, 1), grain=True)
    print(f"superres chain 1080p10 {fr.w}->1920 pipeline {pipeline}", flush=True)
    r = PostFilterRunner()
    r.set_pipeline(pipeline)
    try:
        _check([fr], _gpu_chain(r, [fr], batch=False))
    finally:
        r.close()


def _mixed_sr_frames():
    frames = []
    for i, g in enumerate(SR_GEOMS[:6]):
        up_w, h, denom, bd, ss, mono = g
        frames.append(_frame_sr(400 + i, up_w, h, denom, bd, ss, mono=mono, grain=i % 2 == 0))
        rng = np.random.default_rng(500 + i)
        plain = synth.FrameData(up_w, h, bd, ss[0], ss[1], mono)
        synth.fill_blocky(plain, rng)
        synth.build_side_data(plain, rng, lf_random=True)
        plain.stages = abi.STAGES_ALL & ~abi.STAGE_FILMGRAIN
        frames.append(plain)
    # super-resolution without LR (upscale straight into the upscaled frame)
    frames.append(_frame_sr(600, 200, 136, 14, 10, (1, 1),
                            stages=abi.STAGE_DEBLOCK | abi.STAGE_CDEF | abi.STAGE_SUPERRES))
    return frames


@pytest.mark.gpu
def test_superres_mixed_batch(runner):
    frames = _mixed_sr_frames()
    print(f"superres mixed batch of {len(frames)}", flush=True)
    _check(frames, _gpu_chain(runner, frames, batch=True))


@pytest.mark.gpu
@pytest.mark.parametrize("pipeline", [0, 1], ids=["staged", "fused"])
def test_superres_graph_replays(pipeline):
    from goav1_amd.lib import PostFilterRunner
    frames = _mixed_sr_frames()
    r = PostFilterRunner()
    r.set_pipeline(pipeline)
    try:
        r.set_graphs(True)
        for rep in range(3):
            print(f"superres graph replay {rep} pipeline {pipeline}", flush=True)
            _check(frames, _gpu_chain(r, frames, batch=True))
    finally:
        r.close()


@pytest.mark.gpu
def test_superres_validation(runner):
    from goav1_amd.lib import Av1pfError, DeviceFrame, DeviceSideData
    fr = _frame_sr(7, 200, 136, 12, 8, (1, 1))
    d = DeviceFrame(fr)
    with pytest.raises(Av1pfError):   # no upscaled frame
        runner.apply(d, DeviceSideData(fr))
    wide = synth.FrameData(2 * fr.w + 2, fr.h, 8, 1, 1)
    with pytest.raises(Av1pfError):   # more than twice the coded width
        runner.superres(d, DeviceFrame(wide))
    with pytest.raises(Av1pfError):   # the upscaled frame is the coded frame
        runner.superres(d, d)
    tall = synth.FrameData(200, fr.h + 8, 8, 1, 1)
    with pytest.raises(Av1pfError):   # height differs
        runner.superres(d, DeviceFrame(tall))
