"""GPU parity at the benchmark workloads' own sizes, every frame against the
oracle (full allocations for the filtered frames, visible samples for the
display frames):
  - the 1920x1080 10-bit 4:2:0 moving-rectangle sequence, all 8 frames, in one
    batch (the chain + film grain workload);
  - one 3840x2160 4:2:0 frame per bit depth through every stage;
  - the 64-frame mixed batch (426x240 ... 3840x2160, 8/10-bit, 4:2:0/4:4:4),
    autotuned, then captured as a graph and replayed on fresh copies;
  - the learned search at 1920x1080 and 3840x2160 (2040 filter blocks, near
    its 2048 limit): the decision equals the oracle's finish over the full
    MSE table restricted to the network's shortlist, and every LR unit the
    search filters equals the oracle's curated search.
The oracle runs on the host's cores (one frame per thread; the C code
releases the GIL).
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle_all(frames):
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda f: orc.run_chain(f, f.stages), frames))


def _same(got_planes, want, what, full=True):
    for i, (a, pl) in enumerate(zip(got_planes, want.planes)):
        b = pl.pix
        if not full:
            a, b = a[:pl.h, :pl.w], b[:pl.h, :pl.w]
        d = np.argwhere(a != b)
        if len(d):
            y, x = d[0]
            raise AssertionError(f"{what} plane {i}: {len(d)} diffs, first at y={y} x={x}: gpu={a[y, x]} oracle={b[y, x]}")


def _device(frames):
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    devs, disps, sides = [], [], []
    for fr in frames:
        d = DeviceFrame(fr)
        disp = DeviceFrame.empty_like(fr) if fr.stages & abi.STAGE_FILMGRAIN else None
        devs.append(d)
        disps.append(disp)
        sides.append(DeviceSideData(fr, display=disp))
    return devs, disps, sides


def _check(frames, wants, devs, disps, what):
    for i, (fr, (want, wdisp)) in enumerate(zip(frames, wants)):
        tag = f"{what} frame {i} {fr.w}x{fr.h} {fr.bd}b ss{fr.ssx}{fr.ssy}"
        _same(devs[i].to_numpy(), want, tag)
        if disps[i] is not None:
            _same(disps[i].to_numpy(), wdisp, tag + " display", full=False)


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


def test_sequence_1080p10_all_eight_frames_batched(runner):
    import torch
    frames = synth.sequence_1080p10(seed=1, nframes=8)
    print("1080p10 sequence, 8 frames, oracle", flush=True)
    wants = _oracle_all(frames)
    devs, disps, sides = _device(frames)
    print("1080p10 sequence, 8 frames, one batch", flush=True)
    runner.apply_batch(devs, sides)
    runner.sync()
    torch.cuda.synchronize()
    _check(frames, wants, devs, disps, "1080p10")


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_4k_chain(runner, bd):
    import torch
    rng = np.random.default_rng(2160 + bd)
    fr = synth.FrameData(3840, 2160, bd, 1, 1, sb=64)
    fr.refresh = True
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                          lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=int(rng.choice([64, 128, 256])))
    fr.grain = synth.random_grain(rng, bd)
    fr.stages = abi.STAGES_ALL
    print(f"3840x2160 {bd}-bit chain", flush=True)
    want, wdisp = orc.run_chain(fr)
    devs, disps, sides = _device([fr])
    runner.apply(devs[0], sides[0])
    runner.sync()
    torch.cuda.synchronize()
    _check([fr], [(want, wdisp)], devs, disps, "4K")


def test_mixed_batch_64_frames_tuned_and_graph_replays(runner):
    import torch
    frames = synth.batch_mixed(seed=2, n=64)
    sizes = {(f.w, f.h) for f in frames}
    assert (3840, 2160) in sizes and len(sizes) == 5
    assert {f.bd for f in frames} == {8, 10} and {f.ssx for f in frames} == {0, 1}
    print("64-frame mixed batch, oracle", flush=True)
    wants = _oracle_all(frames)
    devs, disps, sides = _device(frames)
    print("64-frame mixed batch, autotune", flush=True)
    modes = runner.autotune(devs, sides)
    runner.sync()
    torch.cuda.synchronize()
    _check(frames, wants, devs, disps, f"autotune {modes}")
    runner.set_graphs(True)
    try:
        for rep in range(3):
            devs, disps, sides = _device(frames)
            print(f"64-frame mixed batch, graph call {rep}", flush=True)
            runner.apply_batch(devs, sides)
            runner.sync()
            torch.cuda.synchronize()
            _check(frames, wants, devs, disps, f"graph call {rep}")
    finally:
        runner.set_graphs(False)


@pytest.mark.parametrize("size", [(1920, 1080, 8), (3840, 2160, 8), (1920, 1080, 10), (3840, 2160, 10)],
                         ids=["1080p8", "4k8", "1080p10", "4k10"])
def test_learned_search_at_size(size):
    import torch
    from goav1_amd import learned
    from goav1_amd.lib import PostFilterRunner, ProposalNet
    w, h, bd = size
    dc = json.load(open(os.path.join(ROOT, "tests", "golden", "dc_qlookup_8bit.json")))["dc"]
    q = 200
    src, rec, noskip, damping = learned.frame_pair(31000 + w, w, h, q, bd)
    lam = learned.lambda_for(dc, q, bd)
    K = abi.SHORTLIST
    r = PostFilterRunner()
    net = ProposalNet(r, max_batch=64)
    try:
        net.set_params(np.load(learned.DEFAULT_WEIGHTS))
        d = learned.Device(src, rec, noskip)
        assert d.nfb == ((w + 63) // 64) * ((h + 63) // 64)
        print(f"learned search {w}x{h} {bd}-bit ({d.nfb} filter blocks)", flush=True)
        learned.run_learned(r, net, d, damping, lam, 0.9)
        r.sync()
        torch.cuda.synchronize()
        logits = net.infer(d.ds, d.dr).cpu().numpy().astype(np.float64)
        cand = d.cand.cpu().numpy()
        dec = d.dec.cpu().numpy()
        idx = d.idx.cpu().numpy()
        lru = d.lru.cpu().numpy()
    finally:
        net.close()
        r.close()
    print(f"learned search {w}x{h}: oracle table", flush=True)
    full = orc.cdef_search_mse(src, rec, noskip, damping)
    act = learned.active_blocks(w, h, noskip)
    skip = (~act).astype(np.uint8)
    masked = np.full_like(full, np.uint64(1 << 40))
    masked[0][:, cand[:K]] = full[0][:, cand[:K]]
    masked[1][:, cand[K:]] = full[1][:, cand[K:]]
    masked[:, skip.astype(bool), :] = 0
    bits, ys, uvs, fbi = orc.cdef_search_finish(masked, skip, lam)
    assert int(dec[:4].view(np.int32)[0]) == bits
    np.testing.assert_array_equal(dec[4:12], ys)
    np.testing.assert_array_equal(dec[12:20], uvs)
    np.testing.assert_array_equal(idx, fbi)
    # LR: units whose co-located filter block has P(NONE) above the threshold
    # are left NONE, the others are the curated search's records (units
    # within 1e-3 of the threshold, where float rounding may differ, are not
    # checked)
    units, _ = orc.lr_search(src, rec, 64)
    _, _, nhfb, nvfb, hz, vt = learned.geometry(w, h)
    z = logits[:, 128:131]
    p_none = np.exp(z[:, 0] - z.max(1)) / np.exp(z - z.max(1, keepdims=True)).sum(1)
    u = np.arange(hz * vt)
    fb = np.minimum(u // hz, nvfb - 1) * nhfb + np.minimum(u % hz, nhfb - 1)
    pu = p_none[fb]
    sure = np.abs(pu - 0.9) > 1e-3
    on, off = sure & (pu <= 0.9), sure & (pu > 0.9)
    np.testing.assert_array_equal(lru[on], units[on])
    assert (lru[off, 0] == abi.LR_NONE).all()
    assert on.sum() > 0 and (units[on, 0] != abi.LR_NONE).any()
