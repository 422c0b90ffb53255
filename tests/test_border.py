"""Reference-frame border extension (goav1 decoder/postfilter.go:760-771,
Frame.ExtendBordersFrom, frame/buffer.go:293-409).

A frame that refreshes a reference slot is edge-replicated into its whole
allocation (stride / bytes per sample columns, superblock-aligned rows) after
loop restoration and super-resolution, before film grain.  The oracle is
pinned by numpy's edge padding of the visible plane, which is the same
mapping stated independently: sample (y, x) of the allocation is the visible
sample (min(y, h-1), min(x, w-1)).  GPU tests compare the full allocation.
"""
import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc


def _edge_pad(pl):
    v = pl.pix[:pl.h, :pl.w]
    return np.pad(v, ((0, pl.alloc_h - pl.h), (0, pl.alloc_w - pl.w)), mode="edge")


def _frame(seed, w, h, bd, ssx, ssy, sb=64, pad_cols=0, mono=False):
    rng = np.random.default_rng(seed)
    fr = synth.FrameData(w, h, bd, ssx, ssy, mono=mono, sb=sb, pad_cols=pad_cols)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                          lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=64)
    fr.grain = synth.random_grain(rng, bd, mono=mono)
    fr.refresh = True
    return fr


GEOMS = [  # w, h, bd, ssx, ssy, sb, pad_cols, mono
    (200, 136, 8, 1, 1, 64, 0, False), (201, 133, 10, 1, 1, 64, 0, False), (90, 50, 12, 1, 0, 128, 0, False),
    (136, 72, 10, 0, 0, 64, 32, False), (66, 66, 8, 1, 1, 128, 6, False), (130, 70, 10, 1, 1, 64, 0, True),
]
GEOM_IDS = [f"{g[0]}x{g[1]}_{g[2]}b_ss{g[3]}{g[4]}_sb{g[5]}_pad{g[6]}{'_mono' if g[7] else ''}" for g in GEOMS]


@pytest.mark.parametrize("g", GEOMS, ids=GEOM_IDS)
def test_oracle_extension_is_edge_padding(g):
    """With no stage left the chain only extends: every plane of the
    allocation equals the edge padding of its visible samples."""
    w, h, bd, ssx, ssy, sb, pad, mono = g
    fr = _frame(7 + w + h, w, h, bd, ssx, ssy, sb, pad, mono)
    out, _ = orc.run_chain(fr, stages=0)
    for i, pl in enumerate(out.planes):
        assert pl.pix.shape == (pl.alloc_h, pl.alloc_w)
        assert np.array_equal(pl.pix, _edge_pad(fr.planes[i])), f"plane {i}"


@pytest.mark.parametrize("g", GEOMS, ids=GEOM_IDS)
def test_oracle_extension_follows_lr_and_precedes_grain(g):
    """The full chain with refresh equals the chain without it on the visible
    samples and on the display frame (film grain reads visible samples), and
    its padding is the edge padding of the filtered picture."""
    w, h, bd, ssx, ssy, sb, pad, mono = g
    fr = _frame(11 + w + h, w, h, bd, ssx, ssy, sb, pad, mono)
    fr.stages = abi.STAGES_ALL
    out, disp = orc.run_chain(fr)
    fr.refresh = False
    plain, pdisp = orc.run_chain(fr)
    assert orc.plane_md5(out) == orc.plane_md5(plain)
    assert orc.plane_md5(disp) == orc.plane_md5(pdisp)
    for i, pl in enumerate(out.planes):
        assert np.array_equal(pl.pix, _edge_pad(pl)), f"plane {i}"
    # without refresh the padding past the MI-aligned extent is untouched
    for i, (pl, src) in enumerate(zip(plain.planes, fr.planes)):
        ah = ((fr.h + 7) & ~7) >> (fr.ssy if i else 0)
        assert np.array_equal(pl.pix[ah:], src.pix[ah:]), f"plane {i} rows past the MI-aligned height"


def test_oracle_superres_extends_the_upscaled_frame():
    """With super-resolution the upscaled frame is the output that is
    extended; the coded frame keeps its padding."""
    fr = synth.superres_frame(seed=3, up_w=200, h=100, denom=13, bd=10)
    assert fr.refresh and fr.upscaled.planes[0].alloc_h == 128
    out, _, up = orc.run_chain(fr, with_upscaled=True)
    for i, pl in enumerate(up.planes):
        assert np.array_equal(pl.pix, _edge_pad(pl)), f"upscaled plane {i}"
    fr.refresh = False
    out2, _, up2 = orc.run_chain(fr, with_upscaled=True)
    for a, b in zip(out.planes, out2.planes):
        assert np.array_equal(a.pix, b.pix)


# ---------------------------------------------------------------- GPU parity
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


def _gpu(runner, fr, stages):
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    d = DeviceFrame(fr)
    disp = DeviceFrame.empty_like(fr) if stages & abi.STAGE_FILMGRAIN else None
    runner.apply(d, DeviceSideData(fr, display=disp, stages=stages))
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


def _same_full(a, b, what):
    for i, (pa, pb) in enumerate(zip(a.planes, b.planes)):
        d = np.argwhere(pa.pix != pb.pix)
        assert not len(d), f"{what} plane {i}: {len(d)} diffs, first {d[0].tolist()}"


@pytest.mark.gpu
@pytest.mark.parametrize("stages", [0, abi.STAGE_DEBLOCK | abi.STAGE_CDEF, abi.STAGES_ALL],
                         ids=["extend_only", "deblock_cdef", "all"])
@pytest.mark.parametrize("g", GEOMS, ids=GEOM_IDS)
def test_gpu_extension_parity(runner, g, stages):
    w, h, bd, ssx, ssy, sb, pad, mono = g
    fr = _frame(23 + w + h, w, h, bd, ssx, ssy, sb, pad, mono)
    print(f"border case {GEOM_IDS[GEOMS.index(g)]} stages={stages:#x}", flush=True)
    want, wdisp = orc.run_chain(fr, stages=stages)
    got, gdisp = _gpu(runner, fr, stages)
    _same_full(got, want, "frame")
    if wdisp is not None:
        assert orc.plane_md5(gdisp) == orc.plane_md5(wdisp)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_gpu_1080_rows_replicate_row_1079(runner, bd):
    """A 1080-row frame in a 64x64-superblock allocation: after the chain,
    luma rows 1080..1087 equal row 1079 and chroma rows 540..543 equal row
    539, and the whole allocation equals the oracle's."""
    fr = _frame(1080 + bd, 1920, 1080, bd, 1, 1, sb=64)
    fr.stages = abi.STAGES_ALL
    print(f"1080p border bd={bd}", flush=True)
    got, _ = _gpu(runner, fr, fr.stages)
    y = got.planes[0].pix
    assert y.shape == (1088, 1920)
    assert all(np.array_equal(y[r], y[1079]) for r in range(1080, 1088))
    for c in (1, 2):
        p = got.planes[c].pix
        assert p.shape == (544, 960)
        assert all(np.array_equal(p[r], p[539]) for r in range(540, 544))
    want, _ = orc.run_chain(fr)
    _same_full(got, want, "1080p frame")


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10, 12])
def test_gpu_superres_extends_upscaled(runner, bd):
    """1280 -> 1920 x 1080 (denominator 12): the upscaled frame's rows
    1080..1087 equal row 1079 and every sample of both frames equals the
    oracle's."""
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    fr = synth.superres_frame(seed=40 + bd, up_w=1920, h=1080, denom=12, bd=bd, grain=True)
    print(f"superres border bd={bd}", flush=True)
    d, u = DeviceFrame(fr), DeviceFrame(fr.upscaled)
    disp = DeviceFrame.empty_like(fr.upscaled)
    runner.apply(d, DeviceSideData(fr, display=disp, upscaled=u))
    runner.sync()
    torch.cuda.synchronize()
    want, wdisp, wup = orc.run_chain(fr, with_upscaled=True)
    gu = fr.upscaled.copy_pixels()
    for pl, a in zip(gu.planes, u.to_numpy()):
        pl.pix[:] = a
    gc = fr.copy_pixels()
    for pl, a in zip(gc.planes, d.to_numpy()):
        pl.pix[:] = a
    assert all(np.array_equal(gu.planes[0].pix[r], gu.planes[0].pix[1079]) for r in range(1080, 1088))
    _same_full(gu, wup, "upscaled")
    _same_full(gc, want, "coded")
