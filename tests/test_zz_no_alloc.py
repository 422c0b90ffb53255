"""The allocation-free entry point av1pf_apply_noalloc against av1pf_apply.

Two calls in a row on one handle, the second with a smaller frame (so the
static scratch still holds the first call's larger contents), each compared
bit for bit with av1pf_apply on a fresh copy of the same frame: the whole
allocation of the filtered frame and the visible samples of the display
frame.  The small frame is also checked against the oracle.  A request larger
than the scratch is refused before any launch.
"""

import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc

pytestmark = pytest.mark.gpu

def _run(runner, fr, noalloc):
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    import torch
    d = DeviceFrame(fr)
    disp = DeviceFrame.empty_like(fr) if fr.stages & abi.STAGE_FILMGRAIN else None
    side = DeviceSideData(fr, display=disp)
    (runner.apply_noalloc if noalloc else runner.apply)(d, side)
    runner.sync()
    torch.cuda.synchronize()
    return d.to_numpy(), (disp.to_numpy() if disp is not None else None)

def _frames():
    big = synth.sequence_1080p10(seed=17, nframes=1)[0]          # 1920x1080 10-bit 4:2:0, every stage + grain
    rng = np.random.default_rng(71)
    small = synth.FrameData(426, 240, 8, 0, 0, sb=64)           # 8-bit 4:4:4
    synth.fill_blocky(small, rng)
    synth.build_side_data(small, rng, lf_random=True, sharpness=3, lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=64)
    small.grain = synth.random_grain(rng, 8, lag=2)
    small.stages = abi.STAGES_ALL
    small.refresh = True
    return big, small

def test_noalloc_equals_apply_two_calls():
    from goav1_amd.lib import PostFilterRunner
    runner = PostFilterRunner()
    try:
        big, small = _frames()
        for name, fr in (("1920x1080 10-bit 4:2:0", big), ("426x240 8-bit 4:4:4", small)):
            print(f"noalloc case {name}", flush=True)
            want, wdisp = _run(runner, fr, noalloc=False)
            got, gdisp = _run(runner, fr, noalloc=True)
            for i, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), f"{name}: plane {i} differs from av1pf_apply"
            for i, (a, b, pl) in enumerate(zip(gdisp, wdisp, fr.planes)):
                assert np.array_equal(a[:pl.h, :pl.w], b[:pl.h, :pl.w]), f"{name}: display plane {i} differs"
        out, odisp = orc.run_chain(small)
        for i, (a, pl) in enumerate(zip(got, out.planes)):
            assert np.array_equal(a, pl.pix), f"small frame plane {i} differs from the oracle"
    finally:
        runner.close()

def test_noalloc_refuses_requests_past_its_scratch():
    from goav1_amd.lib import Av1pfError, DeviceFrame, DeviceSideData, PostFilterRunner
    runner = PostFilterRunner()
    try:
        rng = np.random.default_rng(5)
        fr = synth.FrameData(3840, 2160, 12, 0, 0)               # 4:4:4 12-bit: ~50 MB of second buffers
        synth.build_side_data(fr, rng, lr_types=None)
        fr.stages = abi.STAGE_CDEF
        with pytest.raises(Av1pfError) as e:
            runner.apply_noalloc(DeviceFrame(fr), DeviceSideData(fr))
        assert e.value.code == abi.ERR_SHORT_BUFFER
        # the entry point still works afterwards
        big, small = _frames()
        want, _ = _run(runner, small, noalloc=False)
        got, _ = _run(runner, small, noalloc=True)
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
    finally:
        runner.close()
