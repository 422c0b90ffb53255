"""GPU parity: libav1pf.so (HIP, gfx950) against the CPU oracle, per-plane MD5.

Each case builds a seeded synthetic frame (pre-filter reconstruction plus legal
per-superblock side data), runs the requested stages through the C ABI on the
device and through the oracle on the CPU, and requires identical per-plane MD5
of the visible samples (bit-exact integer work).  Every device plane sits
between canary-filled guard bands (DeviceFrame), checked each time a frame
is read back: a write outside a stage's planes fails the case.
"""
import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc

pytestmark = pytest.mark.gpu


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


def gpu_run(runner, fr, stages=None):
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    d = DeviceFrame(fr)
    disp = None
    st = fr.stages if stages is None else stages
    if st & abi.STAGE_FILMGRAIN:
        disp = DeviceFrame.empty_like(fr)
    side = DeviceSideData(fr, display=disp, stages=st)
    runner.apply(d, side)
    runner.sync()
    torch.cuda.synchronize()
    out = fr.copy_pixels()
    for pl, a in zip(out.planes, d.to_numpy()):
        pl.pix[:] = a
    dout = None
    if disp is not None:
        dout = fr.copy_pixels()
        for pl, a in zip(dout.planes, disp.to_numpy()):
            pl.pix[:] = a
    return out, dout


def assert_same(a, b, what, full=True):
    """Per-plane MD5 of the visible samples, and with full=True every sample
    of the allocation (MI-aligned padding written by deblock and CDEF, the
    edge-replicated borders of a refreshing frame, untouched padding)."""
    ma, mb = orc.plane_md5(a), orc.plane_md5(b)
    msgs = []
    for i, (pa, pb) in enumerate(zip(a.planes, b.planes)):
        va, vb = (pa.pix, pb.pix) if full else (pa.pix[:pa.h, :pa.w], pb.pix[:pb.h, :pb.w])
        assert va.shape == vb.shape, f"{what}: plane {i} shapes {va.shape} {vb.shape}"
        diff = np.argwhere(va != vb)
        if len(diff):
            y, x = diff[0]
            where = "visible" if y < pa.h and x < pa.w else "padding"
            msgs.append(f"plane {i}: {len(diff)} diffs, first at y={y} x={x} ({where}): "
                        f"gpu={va[y, x]} oracle={vb[y, x]}")
    if msgs or ma != mb:
        raise AssertionError(f"{what}: " + "; ".join(msgs))


def small_frame(seed, w, h, bd, ss, lr_unit=64, grain=False, lr_types=(abi.LR_SWITCHABLE,) * 3):
    rng = np.random.default_rng(seed)
    fr = synth.FrameData(w, h, bd, ss, ss)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                          lr_types=lr_types, lr_unit=lr_unit)
    fr.stages = abi.STAGES_ALL & ~abi.STAGE_FILMGRAIN
    if grain:
        fr.grain = synth.random_grain(rng, bd)
        fr.stages |= abi.STAGE_FILMGRAIN
    return fr


STAGE_SETS = [abi.STAGE_DEBLOCK, abi.STAGE_CDEF, abi.STAGE_LR,
              abi.STAGE_DEBLOCK | abi.STAGE_CDEF | abi.STAGE_LR]
GEOMS = [(200, 136, 8, 1), (200, 136, 10, 1), (136, 72, 8, 0), (136, 72, 10, 0),
         (66, 66, 8, 1), (426, 240, 10, 1), (90, 50, 8, 1), (200, 136, 12, 1), (136, 72, 12, 0)]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: f"{g[0]}x{g[1]}_{g[2]}b_ss{g[3]}")
@pytest.mark.parametrize("stages", STAGE_SETS, ids=["deblock", "cdef", "lr", "chain"])
def test_stage_parity(runner, geom, stages):
    w, h, bd, ss = geom
    fr = small_frame(1000 + w + h + bd + ss, w, h, bd, ss)
    print(f"case {w}x{h} bd={bd} ss={ss} stages={stages:#x}", flush=True)
    want, _ = orc.run_chain(fr, stages)
    got, _ = gpu_run(runner, fr, stages)
    assert_same(got, want, f"{w}x{h} bd{bd} ss{ss} stages {stages:#x}")


@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("ss", [0, 1])
def test_film_grain_parity(runner, bd, ss):
    fr = small_frame(77 + bd + ss, 232, 150, bd, ss, grain=True)
    print(f"grain case bd={bd} ss={ss}", flush=True)
    want, wdisp = orc.run_chain(fr)
    got, gdisp = gpu_run(runner, fr)
    assert_same(got, want, "filtered frame")
    assert_same(gdisp, wdisp, "display frame", full=False)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_chain_422(runner, bd):
    """4:2:2 through every stage (chroma LR units as large as luma units)."""
    rng = np.random.default_rng(4220 + bd)
    fr = synth.FrameData(232, 150, bd, 1, 0)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)), lr_unit=64)
    fr.grain = synth.random_grain(rng, bd)
    fr.stages = abi.STAGES_ALL
    print(f"4:2:2 chain bd={bd}", flush=True)
    want, wdisp = orc.run_chain(fr)
    got, gdisp = gpu_run(runner, fr)
    assert_same(got, want, "4:2:2 filtered frame")
    assert_same(gdisp, wdisp, "4:2:2 display frame", full=False)


@pytest.mark.parametrize("unit", [64, 128, 256])
def test_lr_unit_sizes(runner, unit):
    fr = small_frame(300 + unit, 520, 300, 10, 1, lr_unit=unit)
    want, _ = orc.run_chain(fr)
    got, _ = gpu_run(runner, fr)
    assert_same(got, want, f"lr unit {unit}")


def test_keyframe_720p(runner):
    fr = synth.keyframe_720p(0)
    want, _ = orc.run_chain(fr)
    got, _ = gpu_run(runner, fr)
    assert_same(got, want, "720p keyframe")


def test_sequence_1080p10_first_frames(runner):
    for fr in synth.sequence_1080p10(1, nframes=2):
        want, wdisp = orc.run_chain(fr)
        got, gdisp = gpu_run(runner, fr)
        assert_same(got, want, "1080p10 filtered")
        assert_same(gdisp, wdisp, "1080p10 display", full=False)


def _frames_from(fr, devs):
    out = fr.copy_pixels()
    for pl, a in zip(out.planes, devs.to_numpy()):
        pl.pix[:] = a
    return out


def mixed_frames():
    """14 seeded frames of mixed size, depth, chroma format and stage set."""
    rng = np.random.default_rng(4242)
    sizes = [(200, 136), (136, 72), (66, 66), (426, 240), (90, 50), (320, 180)]
    stage_sets = STAGE_SETS + [abi.STAGES_ALL, abi.STAGE_FILMGRAIN, abi.STAGE_CDEF | abi.STAGE_FILMGRAIN]
    frames = []
    for i in range(14):
        w, h = sizes[i % len(sizes)]
        bd = 8 if i % 3 else 10
        ss = (i // 2) % 2
        fr = small_frame(9000 + i, w, h, bd, ss, lr_unit=int(rng.choice([64, 128])))
        fr.grain = synth.random_grain(rng, bd, lag=int(rng.integers(0, 4)))
        fr.stages = stage_sets[i % len(stage_sets)]
        frames.append(fr)
    return frames


def device_batch(frames):
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    devs, disps, sides = [], [], []
    for fr in frames:
        d = DeviceFrame(fr)
        disp = DeviceFrame.empty_like(fr) if fr.stages & abi.STAGE_FILMGRAIN else None
        devs.append(d)
        disps.append(disp)
        sides.append(DeviceSideData(fr, display=disp, stages=fr.stages))
    return devs, disps, sides


def check_batch(frames, devs, disps, wants, what):
    for i, fr in enumerate(frames):
        want, wdisp = wants[i]
        assert_same(_frames_from(fr, devs[i]), want, f"{what} frame {i} stages {fr.stages:#x}")
        if disps[i] is not None:
            assert_same(_frames_from(fr, disps[i]), wdisp, f"{what} frame {i} display", full=False)


@pytest.fixture(scope="module")
def mixed():
    frames = mixed_frames()
    return frames, [orc.run_chain(fr, fr.stages) for fr in frames]


def test_mixed_batch(runner, mixed):
    """One apply_batch over frames of mixed size, depth, chroma format and
    stage set must equal the oracle frame by frame."""
    import torch
    frames, wants = mixed
    devs, disps, sides = device_batch(frames)
    print(f"mixed batch of {len(frames)} frames", flush=True)
    runner.apply_batch(devs, sides)
    runner.sync()
    torch.cuda.synchronize()
    check_batch(frames, devs, disps, wants, "batch")


@pytest.mark.parametrize("mode", [2, 4, 0])
def test_mixed_batch_work_distribution(runner, mixed, mode):
    """CDEF, LR and film-grain apply with several tiles per workgroup or
    persistent workgroups on the device queue give the same frames."""
    import torch
    frames, wants = mixed
    for st in (abi.STAGE_CDEF, abi.STAGE_LR, abi.STAGE_FILMGRAIN):
        runner.set_stage_mode(st, mode)
    try:
        for rep in range(2):
            devs, disps, sides = device_batch(frames)
            print(f"mixed batch, stage mode {mode}, run {rep}", flush=True)
            runner.apply_batch(devs, sides)
            runner.sync()
            torch.cuda.synchronize()
            check_batch(frames, devs, disps, wants, f"mode {mode} run {rep}")
    finally:
        for st in (abi.STAGE_CDEF, abi.STAGE_LR, abi.STAGE_FILMGRAIN):
            runner.set_stage_mode(st, -1)


def test_mixed_batch_autotune_and_graph_replays(runner, mixed):
    """autotune returns the oracle's frames and picks a mode per stage; then
    the class is captured as a graph and replayed three times on fresh copies,
    with other device work and a single-frame call (another class) between
    replays; every replay equals the oracle."""
    import torch
    frames, wants = mixed
    devs, disps, sides = device_batch(frames)
    print("mixed batch autotune", flush=True)
    modes = runner.autotune(devs, sides)
    runner.sync()
    torch.cuda.synchronize()
    assert set(modes.values()) <= {0, 1, 2, 4}, modes
    check_batch(frames, devs, disps, wants, "autotune")
    runner.set_graphs(True)
    try:
        for rep in range(3):
            devs, disps, sides = device_batch(frames)
            print(f"mixed batch graph replay {rep}", flush=True)
            runner.apply_batch(devs, sides)
            runner.sync()
            torch.cuda.synchronize()
            check_batch(frames, devs, disps, wants, f"graph replay {rep}")
            junk = torch.arange(1 << 16, device="cuda").float().sum()   # eager work between replays
            out, _ = gpu_run(runner, frames[0])                          # another shape class
            assert_same(out, wants[0][0], "single-frame call between replays")
            assert junk.item() > 0
    finally:
        runner.set_graphs(False)


@pytest.mark.parametrize("bd", [8, 10])
def test_stage_entry_points(runner, bd):
    """deblock -> save LR rows -> cdef -> save (after CDEF) -> lr -> film grain
    through the per-stage entry points equals the chain."""
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    fr = small_frame(555 + bd, 264, 168, bd, 1, grain=True)
    d = DeviceFrame(fr)
    disp = DeviceFrame.empty_like(fr)
    side = DeviceSideData(fr, display=disp)
    print(f"stage entry points bd={bd}", flush=True)
    runner.deblock(d, side)
    runner.lr_save_boundaries(d, side, after_cdef=False)
    runner.cdef(d, side)
    runner.lr_save_boundaries(d, side, after_cdef=True)
    runner.lr(d, side)
    runner.filmgrain(d, disp, side)
    runner.sync()
    torch.cuda.synchronize()
    want, wdisp = orc.run_chain(fr)
    assert_same(_frames_from(fr, d), want, "stage-level chain")
    assert_same(_frames_from(fr, disp), wdisp, "stage-level display", full=False)


@pytest.mark.parametrize("lag", [0, 1, 2, 3])
@pytest.mark.parametrize("ss", [(1, 1), (1, 0), (0, 0)], ids=["420", "422", "444"])
def test_film_grain_lags(runner, lag, ss):
    """Every AR lag on every chroma layout (template rows in one and two
    passes of the wavefront)."""
    rng = np.random.default_rng(31 * lag + 7 * ss[0] + ss[1])
    fr = synth.FrameData(150, 94, 10, ss[0], ss[1])
    synth.fill_blocky(fr, rng)
    fr.grain = synth.random_grain(rng, 10, lag=lag)
    fr.stages = abi.STAGE_FILMGRAIN
    print(f"grain lag={lag} ss={ss}", flush=True)
    want, wdisp = orc.run_chain(fr)
    got, gdisp = gpu_run(runner, fr)
    assert_same(gdisp, wdisp, f"display lag {lag} ss {ss}", full=False)


@pytest.mark.parametrize("case", [
    dict(w=97, h=67, bd=10, ss=1, lag=2, restricted=1), dict(w=72, h=40, bd=10, ss=0, lag=1, overlap=0),
    dict(w=66, h=66, bd=8, ss=1, lag=0, csfl=1), dict(w=80, h=36, bd=10, ss=1, lag=3, gss=2, shift=8, arshift=9),
    dict(w=161, h=99, bd=8, ss=0, lag=3, restricted=1, identity=1)], ids=lambda c: "_".join(f"{k}{v}" for k, v in c.items()))
def test_film_grain_variants(runner, case):
    """Restricted range (with the identity-matrix chroma limit), no overlap,
    chroma scaling from luma, grain scale shift and other shifts."""
    c = dict(restricted=0, overlap=1, csfl=0, gss=0, shift=10, arshift=7, identity=0)
    c.update(case)
    rng = np.random.default_rng(900 + c["w"] + c["h"])
    fr = synth.FrameData(c["w"], c["h"], c["bd"], c["ss"], c["ss"])
    synth.fill_blocky(fr, rng)
    fr.matrix_identity = c["identity"]
    g = synth.random_grain(rng, c["bd"], lag=c["lag"], csfl=bool(c["csfl"]))
    g.clip_to_restricted_range, g.overlap = c["restricted"], c["overlap"]
    g.grain_scale_shift, g.scaling_shift, g.ar_coeff_shift = c["gss"], c["shift"], c["arshift"]
    fr.grain = g
    fr.stages = abi.STAGE_FILMGRAIN
    print(f"grain variant {case}", flush=True)
    want, wdisp = orc.run_chain(fr)
    got, gdisp = gpu_run(runner, fr)
    assert_same(gdisp, wdisp, f"display {case}", full=False)


def test_deblock_misaligned_masks_rejected(runner):
    """The deblock kernels read the mask records as 32-bit words: a mask array
    that is not 4-byte aligned is refused by the host before any launch."""
    from goav1_amd.lib import Av1pfError, DeviceFrame, DeviceSideData
    rng = np.random.default_rng(4)
    fr = synth.FrameData(128, 96, 8, 1, 1)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True)
    fr.stages = abi.STAGE_DEBLOCK
    side = DeviceSideData(fr)
    side.req.deblock.contents.masks += 2
    with pytest.raises(Av1pfError):
        runner.apply(DeviceFrame(fr), side)


def test_lr_chroma_unit_below_stripe_rejected(runner):
    """Chroma LR units halve only for 4:2:0: a 32-sample unit on a 4:2:2 plane
    (64-row stripes) is not a legal request and is refused before launch."""
    from goav1_amd.lib import Av1pfError, DeviceFrame, DeviceSideData
    rng = np.random.default_rng(422)
    fr = synth.FrameData(128, 96, 8, 1, 0)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lr_unit=64)
    assert fr.lr["unit_size"] == [64, 64, 64]
    fr.lr["unit_size"] = [64, 32, 32]
    fr.stages = abi.STAGE_LR
    with pytest.raises(Av1pfError):
        runner.apply(DeviceFrame(fr), DeviceSideData(fr))


def test_autotune_validates_requests(runner):
    """autotune validates every request before planning, as apply_batch does:
    a missing CDEF parameter block or a mask grid that does not match the frame
    is refused with INVALID_ARG and nothing is launched."""
    from goav1_amd.lib import Av1pfError, DeviceFrame, DeviceSideData
    rng = np.random.default_rng(8)
    fr = synth.FrameData(128, 96, 8, 1, 1)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True)
    fr.stages = abi.STAGE_DEBLOCK | abi.STAGE_CDEF
    side = DeviceSideData(fr)
    side.req.cdef = None
    with pytest.raises(Av1pfError) as e:
        runner.autotune([DeviceFrame(fr)], [side])
    assert e.value.code == abi.ERR_INVALID_ARG
    side = DeviceSideData(fr)
    side.req.deblock.contents.mi_cols += 2
    with pytest.raises(Av1pfError) as e:
        runner.autotune([DeviceFrame(fr)], [side])
    assert e.value.code == abi.ERR_INVALID_ARG


def test_graph_replay_then_autotune_without_sync(runner, mixed):
    """A graph replay followed at once by autotune (which starts building film
    grain templates on the side stream) and a second replay: the template
    pair the replay reads is not rewritten under it, every result equals the
    oracle."""
    import torch
    frames, wants = mixed
    runner.set_graphs(True)
    try:
        devs, disps, sides = device_batch(frames)
        runner.apply_batch(devs, sides)               # capture
        runner.sync()
        devs, disps, sides = device_batch(frames)
        print("graph replay then autotune", flush=True)
        runner.apply_batch(devs, sides)               # replay, not synchronised
        devs2, disps2, sides2 = device_batch(frames)
        runner.autotune(devs2, sides2)
        devs3, disps3, sides3 = device_batch(frames)
        runner.apply_batch(devs3, sides3)             # capture again (autotune dropped the class)
        runner.sync()
        torch.cuda.synchronize()
        check_batch(frames, devs, disps, wants, "replay before autotune")
        check_batch(frames, devs2, disps2, wants, "autotune after replay")
        check_batch(frames, devs3, disps3, wants, "replay after autotune")
    finally:
        runner.set_graphs(False)


def test_guard_bands_catch_writes_outside_the_planes(runner):
    """The footprint check itself: host writes into either guard band of a
    plane are reported, untouched bands pass after a full chain."""
    from goav1_amd.lib import DeviceFrame
    fr = small_frame(9, 136, 72, 10, 1, grain=True)
    got, _ = gpu_run(runner, fr)  # reads back (and checks) frame and display
    want, _ = orc.run_chain(fr)
    assert_same(got, want, "guarded chain")
    for band in ("before", "after"):
        d = DeviceFrame(fr)
        d.check_canaries()
        buf, g = d._bufs[1], d._guard[1]
        k = g - 1 if band == "before" else buf.numel() - g
        buf[k] = 0
        with pytest.raises(AssertionError, match=band):
            d.to_numpy()
