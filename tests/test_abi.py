"""C ABI checks that run without a GPU: the library loads, exports every entry
point include/av1pf.h declares, its structs have the layout the ctypes mirror
assumes, and the host-side validation rejects bad requests before any device
work (the reference validates before mutating, decoder/postfilter.go:598-616)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from goav1_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "av1pf.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(av1pf_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from goav1_amd import lib
    L = lib.load()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), f"libav1pf.so does not export {n}"
    assert set(names) == set(lib.EXPORTS)
    assert L.av1pf_version().startswith(b"av1pf")


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "av1pf.h"
#define P(s) printf(#s " %zu\n", sizeof(s))
#define O(s, f) printf(#s "." #f " %zu\n", offsetof(s, f))
int main(void) {
  P(av1pf_frame); P(av1pf_lf_mask); P(av1pf_deblock_params); P(av1pf_cdef_params);
  P(av1pf_lr_unit); P(av1pf_lr_params); P(av1pf_grain_params); P(av1pf_request); P(av1pf_cdef_decision);
  P(av1pf_lf_record); P(av1pf_lf_build_params);
  O(av1pf_lf_record, split); O(av1pf_lf_record, delta_lf); O(av1pf_lf_build_params, ref_delta);
  O(av1pf_lf_build_params, seg_delta); O(av1pf_request, upscaled);
  O(av1pf_frame, plane); O(av1pf_frame, stride); O(av1pf_deblock_params, level_cache);
  O(av1pf_cdef_params, index); O(av1pf_lr_params, units); O(av1pf_grain_params, ar_coeffs_cr);
  O(av1pf_grain_params, cr_offset); O(av1pf_request, display);
  return 0; }
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    sizes = {"av1pf_frame": abi.Frame, "av1pf_lf_mask": abi.LFMask, "av1pf_deblock_params": abi.DeblockParams,
             "av1pf_cdef_params": abi.CdefParams, "av1pf_lr_unit": abi.LRUnit, "av1pf_lr_params": abi.LRParams,
             "av1pf_grain_params": abi.GrainParams, "av1pf_request": abi.Request,
             "av1pf_cdef_decision": abi.CdefDecision, "av1pf_lf_build_params": abi.LFBuildParams}
    from goav1_amd import synth
    assert int(out["av1pf_lf_record"]) == synth.LF_RECORD.itemsize == abi.LF_RECORD_BYTES
    for f in ("split", "delta_lf"):
        assert int(out[f"av1pf_lf_record.{f}"]) == synth.LF_RECORD.fields[f][1], f
    for name, cls in sizes.items():
        assert int(out[name]) == C.sizeof(cls), name
    offs = {"av1pf_frame.plane": abi.Frame.plane.offset, "av1pf_frame.stride": abi.Frame.stride.offset,
            "av1pf_deblock_params.level_cache": abi.DeblockParams.level_cache.offset,
            "av1pf_cdef_params.index": abi.CdefParams.index.offset,
            "av1pf_lr_params.units": abi.LRParams.units.offset,
            "av1pf_grain_params.ar_coeffs_cr": abi.GrainParams.ar_coeffs_cr.offset,
            "av1pf_grain_params.cr_offset": abi.GrainParams.cr_offset.offset,
            "av1pf_request.display": abi.Request.display.offset,
            "av1pf_request.upscaled": abi.Request.upscaled.offset,
            "av1pf_lf_build_params.ref_delta": abi.LFBuildParams.ref_delta.offset,
            "av1pf_lf_build_params.seg_delta": abi.LFBuildParams.seg_delta.offset}
    for name, off in offs.items():
        assert int(out[name]) == off, name


def test_create_without_device_fails_cleanly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    from goav1_amd import lib
    L = lib.load()
    h = C.c_void_p()
    rc = L.av1pf_create(C.byref(h), -1)
    assert rc == abi.ERR_NO_DEVICE and not h.value


def test_null_context_is_rejected():
    from goav1_amd import lib
    L = lib.load()
    f = abi.Frame()
    r = abi.Request()
    assert L.av1pf_apply(None, C.byref(f), C.byref(r), None) == abi.ERR_INVALID_ARG
    assert L.av1pf_sync(None, None) == abi.ERR_INVALID_ARG
    assert L.av1pf_apply_batch(None, 0, None, None, None) == abi.ERR_INVALID_ARG


def test_product_path_has_no_cpu_fallback():
    """The runner must refuse to run without a device rather than fall back."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    from goav1_amd.lib import PostFilterRunner
    with pytest.raises(RuntimeError):
        PostFilterRunner()
