"""Host binding of libav1pf.so: the post-filter runner behind the C ABI.

`PostFilterRunner` mirrors goav1's FrameWorkPostFilterRunner
(internal/av1/decoder/work.go:153-155): one `apply(frame, request)` per
completed frame, running the remaining post-filter stages in place on the
frame and writing film grain into a separate display frame
(decoder/postfilter.go:589-845).  Device memory comes from torch (plumbing
only); every filter runs in the HIP kernels of libav1pf.so.  There is no CPU
fallback: if the library or a GPU is missing, construction raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libav1pf.so")
_lib = None

EXPORTS = ["av1pf_create", "av1pf_destroy", "av1pf_version", "av1pf_last_error", "av1pf_apply",
           "av1pf_apply_batch", "av1pf_sync", "av1pf_deblock", "av1pf_cdef", "av1pf_lr_save_boundaries",
           "av1pf_lr", "av1pf_filmgrain", "av1pf_cdef_search", "av1pf_cdef_search_finish", "av1pf_cdef_lambda",
           "av1pf_cdef_search_select", "av1pf_lr_search",
           "av1pf_nn_param_count", "av1pf_nn_create", "av1pf_nn_destroy", "av1pf_nn_set_params",
           "av1pf_nn_get_params", "av1pf_nn_tiles", "av1pf_nn_forward", "av1pf_nn_infer", "av1pf_nn_backward",
           "av1pf_nn_adam",
           "av1pf_lr_search_masked", "av1pf_cdef_search_shortlist", "av1pf_search_learned", "av1pf_search_shortlist_size",
           "av1pf_profile", "av1pf_profile_read", "av1pf_set_graphs", "av1pf_autotune", "av1pf_set_stage_mode",
           "av1pf_lr_search_exhaustive", "av1pf_superres", "av1pf_deblock_build",
           "av1pf_cdef_search_apply", "av1pf_apply_noalloc", "av1pf_set_pipeline"]
STAGE_NAMES = ["deblock", "lr_save", "cdef", "lr", "copy", "grain_template", "grain_apply", "cdef_search",
               "cdef_search_select", "lr_search", "nn_forward", "nn_train", "cdef_search_learned",
               "lr_search_exhaustive", "superres", "deblock_build", "extend", "chain"]
PROFILE_STAGES = 18
PIPELINE_STAGED, PIPELINE_FUSED = 0, 1


class Av1pfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"av1pf error {code}: {msg}")
        self.code = code


def load():
    """Load libav1pf.so (built by goav1_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python -m goav1_amd.build")
    lib = C.CDLL(LIB_PATH)
    lib.av1pf_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.av1pf_destroy.argtypes = [C.c_void_p]
    lib.av1pf_destroy.restype = None
    lib.av1pf_version.restype = C.c_char_p
    lib.av1pf_last_error.argtypes = [C.c_void_p]
    lib.av1pf_last_error.restype = C.c_char_p
    lib.av1pf_apply.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Request), C.c_void_p]
    lib.av1pf_apply_noalloc.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Request), C.c_void_p]
    lib.av1pf_apply_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.Frame), C.POINTER(abi.Request), C.c_void_p]
    lib.av1pf_sync.argtypes = [C.c_void_p, C.c_void_p]
    lib.av1pf_set_graphs.argtypes = [C.c_void_p, C.c_int]
    lib.av1pf_lr_search_exhaustive.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_int,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.av1pf_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.Frame), C.POINTER(abi.Request),
                                   C.POINTER(C.c_int32), C.c_void_p]
    lib.av1pf_set_stage_mode.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
    lib.av1pf_set_pipeline.argtypes = [C.c_void_p, C.c_int]
    lib.av1pf_deblock.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.DeblockParams), C.c_void_p]
    lib.av1pf_cdef.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.CdefParams), C.c_void_p]
    lib.av1pf_lr_save_boundaries.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.LRParams), C.c_int, C.c_void_p]
    lib.av1pf_lr.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.LRParams), C.c_void_p]
    lib.av1pf_superres.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_void_p]
    lib.av1pf_cdef_search_apply.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    lib.av1pf_deblock_build.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(abi.LFBuildParams), C.c_void_p, C.c_void_p,
                                        C.c_void_p]
    lib.av1pf_filmgrain.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame),
                                    C.POINTER(abi.GrainParams), C.c_void_p]
    lib.av1pf_cdef_search.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_void_p]
    lib.av1pf_cdef_search_finish.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.POINTER(C.c_int)]
    lib.av1pf_cdef_search_select.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.c_void_p, C.c_void_p, C.c_uint64,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
    lib.av1pf_lr_search.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
    lib.av1pf_nn_param_count.restype = C.c_int64
    lib.av1pf_nn_create.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
    lib.av1pf_nn_destroy.argtypes = [C.c_void_p]
    lib.av1pf_nn_destroy.restype = None
    lib.av1pf_nn_set_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.av1pf_nn_get_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.av1pf_nn_tiles.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_void_p, C.c_void_p]
    lib.av1pf_nn_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.av1pf_nn_infer.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_void_p, C.c_void_p]
    lib.av1pf_nn_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.av1pf_nn_adam.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p]
    lib.av1pf_lr_search_masked.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.av1pf_cdef_search_shortlist.argtypes = [C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame), C.c_void_p,
                                                C.c_int, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p]
    lib.av1pf_search_learned.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(abi.Frame), C.POINTER(abi.Frame),
                                         C.c_void_p, C.c_int, C.c_uint64, C.c_float, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.av1pf_cdef_lambda.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.av1pf_cdef_lambda.restype = C.c_uint64
    lib.av1pf_profile.argtypes = [C.c_void_p, C.c_int]
    lib.av1pf_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    _lib = lib
    return lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the post-filter path runs only on the HIP device")
    return torch


class DeviceFrame:
    """A frame resident in device memory in the ABI's sample layout.

    Each plane lives inside a larger device buffer: guard bands of at least
    4 KiB and two rows (a multiple of 256 bytes, so the plane keeps the
    buffer's alignment) sit before and after it, filled with a canary value.
    No stage may write outside the planes it is given; check_canaries()
    reports any sample of a band that changed, and to_numpy() calls it, so
    every comparison of a device frame also checks its footprint.
    """

    CANARY8 = 0xA5
    CANARY16 = 0x5AA5  # above every 12-bit sample value

    def __init__(self, fr, device="cuda"):
        self.meta = fr
        self.bd = fr.bd
        self._guarded([pl.pix for pl in fr.planes], device)
        self.abi = self._abi()

    @classmethod
    def empty_like(cls, fr, device="cuda"):
        obj = cls.__new__(cls)
        obj.meta, obj.bd = fr, fr.bd
        obj._guarded([np.zeros(pl.pix.shape, np.uint16) for pl in fr.planes], device)
        obj.abi = obj._abi()
        return obj

    def _guarded(self, pixes, device):
        torch = _torch()
        esz = 1 if self.bd == 8 else 2
        canary = self.CANARY8 if self.bd == 8 else self.CANARY16 - 0x10000 * (self.CANARY16 >= 0x8000)
        self.planes, self._bufs, self._guard = [], [], []
        for pix in pixes:
            rows, cols = pix.shape
            g = max(4096, 2 * cols * esz)
            g = (g + 255) // 256 * 256 // esz  # guard length in samples
            host = pix.astype(np.uint8) if self.bd == 8 else pix.view(np.int16)
            buf = torch.full((2 * g + rows * cols,), canary, dtype=torch.uint8 if self.bd == 8 else torch.int16,
                             device=device)
            buf[g:g + rows * cols].copy_(torch.from_numpy(np.ascontiguousarray(host).reshape(-1)))
            self._bufs.append(buf)
            self._guard.append(g)
            self.planes.append(buf[g:g + rows * cols].view(rows, cols))

    def clone(self):
        """A device copy of this frame, guard bands included."""
        c = DeviceFrame.__new__(DeviceFrame)
        c.meta, c.bd = self.meta, self.bd
        c._bufs = [b.clone() for b in self._bufs]
        c._guard = list(self._guard)
        c.planes = [b[g:g + t.numel()].view(t.shape) for b, g, t in zip(c._bufs, c._guard, self.planes)]
        c.abi = c._abi()
        return c

    def check_canaries(self):
        canary = self.CANARY8 if self.bd == 8 else self.CANARY16 - 0x10000 * (self.CANARY16 >= 0x8000)
        for i, (buf, g) in enumerate(zip(self._bufs, self._guard)):
            n = buf.numel()
            for name, band, base in (("before", buf[:g], -g), ("after", buf[n - g:], n - 2 * g)):
                bad = (band != canary).nonzero()
                if bad.numel():
                    k = int(bad[0, 0])
                    raise AssertionError(
                        f"plane {i}: {bad.numel()} guard samples {name} the plane were written, first at "
                        f"sample offset {base + k} from the plane start")

    def _abi(self):
        fr = self.meta
        f = abi.Frame()
        f.width, f.height, f.bit_depth = fr.w, fr.h, fr.bd
        f.ss_x, f.ss_y, f.mono, f.matrix_identity = fr.ssx, fr.ssy, fr.mono, fr.matrix_identity
        f.alloc_height = fr.planes[0].alloc_h
        for i, t in enumerate(self.planes):
            f.plane[i] = t.data_ptr()
            f.stride[i] = t.stride(0) * t.element_size()
        return f

    def to_numpy(self):
        """Planes back on the host as uint16 arrays (synth.FrameData layout),
        after checking that no guard band around them was written."""
        self.check_canaries()
        out = []
        for t in self.planes:
            a = t.cpu().numpy()
            out.append(a.astype(np.uint16) if self.bd == 8 else a.view(np.uint16).copy())
        return out


class DeviceSideData:
    """Per-frame side data uploaded to the device, plus the host request struct."""

    def __init__(self, fr, display=None, device="cuda", stages=None, upscaled=None):
        torch = _torch()
        self.keep = []
        req = abi.Request()
        req.stages = fr.stages if stages is None else stages
        req.flags = abi.REQ_REFRESH if getattr(fr, "refresh", False) else 0

        def dev(arr):
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(device)
            self.keep.append(t)
            return t.data_ptr()

        if fr.deblock is not None:
            d = fr.deblock
            p = abi.DeblockParams()
            p.level_y[0], p.level_y[1] = d['level_y']
            p.level_u, p.level_v, p.sharpness = d['level_u'], d['level_v'], d['sharpness']
            p.mi_cols, p.mi_rows, p.sb128_w, p.sb128_h = d['mi_cols'], d['mi_rows'], d['sb128_w'], d['sb128_h']
            p.level_cache = dev(d['level_cache'])
            p.masks = dev(d['masks'].view(np.int16))
            self.keep.append(p)
            req.deblock = C.pointer(p)
        if fr.cdef is not None:
            c = fr.cdef
            p = abi.CdefParams()
            p.damping, p.strength_count = c['damping'], c['strength_count']
            for i in range(8):
                p.y_strength[i], p.uv_strength[i] = int(c['y_strength'][i]), int(c['uv_strength'][i])
            p.unit_cols, p.unit_rows = c['unit_cols'], c['unit_rows']
            p.index = dev(c['index'])
            p.noskip = dev(c['noskip'].view(np.int64))
            self.keep.append(p)
            req.cdef = C.pointer(p)
        if fr.lr is not None:
            l = fr.lr
            p = abi.LRParams()
            for i in range(3):
                p.frame_type[i] = l['frame_type'][i]
                p.unit_size[i] = l['unit_size'][i]
                u = l['units'][i]
                p.units[i] = dev(u.view(np.uint8)) if u is not None else None
            self.keep.append(p)
            req.lr = C.pointer(p)
        if fr.grain is not None:
            g = abi.GrainParams()
            C.memmove(C.byref(g), C.byref(fr.grain), C.sizeof(g))
            self.keep.append(g)
            req.grain = C.pointer(g)
        if display is not None:
            self.keep.append(display.abi)
            req.display = C.pointer(display.abi)
        if upscaled is not None:
            self.keep.append(upscaled.abi)
            req.upscaled = C.pointer(upscaled.abi)
        self.req = req


class PostFilterRunner:
    """One handle per decoder (the reference runner is not reentrant)."""

    def __init__(self, device=-1):
        _torch()
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.av1pf_create(C.byref(h), device)
        if rc != abi.OK:
            raise Av1pfError(rc, "av1pf_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.av1pf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != abi.OK:
            raise Av1pfError(rc, self.lib.av1pf_last_error(self.h).decode())

    @staticmethod
    def _stream(stream):
        if stream is None:
            import torch
            return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        return C.c_void_p(stream)

    def apply(self, frame, side, stream=None):
        """Run the request's stages on `frame` (DeviceFrame), asynchronously."""
        self._check(self.lib.av1pf_apply(self.h, C.byref(frame.abi), C.byref(side.req), self._stream(stream)))

    def apply_noalloc(self, frame, side, stream=None):
        """apply() through the allocation-free entry point (the library's
        static per-device scratch; requests up to NOALLOC_SCRATCH_BYTES)."""
        self._check(self.lib.av1pf_apply_noalloc(self.h, C.byref(frame.abi), C.byref(side.req), self._stream(stream)))

    def apply_batch(self, frames, sides, stream=None):
        n = len(frames)
        fa = (abi.Frame * n)(*[f.abi for f in frames])
        ra = (abi.Request * n)(*[s.req for s in sides])
        self._check(self.lib.av1pf_apply_batch(self.h, n, fa, ra, self._stream(stream)))

    def set_graphs(self, enable):
        """Capture each batch shape class as a HIP graph and replay it."""
        self._check(self.lib.av1pf_set_graphs(self.h, int(bool(enable))))

    def autotune(self, frames, sides, stream=None):
        """apply_batch that also picks each tunable stage's work distribution
        for this shape class; returns {stage: tiles per workgroup, 0 = queue}."""
        n = len(frames)
        fa = (abi.Frame * n)(*[f.abi for f in frames])
        ra = (abi.Request * n)(*[s.req for s in sides])
        modes = (C.c_int32 * 3)()
        self._check(self.lib.av1pf_autotune(self.h, n, fa, ra, modes, self._stream(stream)))
        return {"cdef": modes[0], "lr": modes[1], "grain_apply": modes[2]}

    def set_stage_mode(self, stage, mode):
        """Force a stage's work distribution (abi.STAGE_CDEF / _LR / _FILMGRAIN;
        1, 2, 4 tiles per workgroup, 0 = persistent queue, -1 = automatic)."""
        self._check(self.lib.av1pf_set_stage_mode(self.h, int(stage), int(mode)))

    def set_pipeline(self, mode):
        """PIPELINE_STAGED (one launch per stage) or PIPELINE_FUSED (deblock,
        LR context save, CDEF, super-resolution and LR in one persistent
        launch with per-band dependency counters); same results."""
        self._check(self.lib.av1pf_set_pipeline(self.h, int(mode)))

    # stage-level entry points (the stage functions the chain composes)
    def deblock(self, frame, side, stream=None):
        self._check(self.lib.av1pf_deblock(self.h, C.byref(frame.abi), side.req.deblock, self._stream(stream)))

    def cdef(self, frame, side, stream=None):
        self._check(self.lib.av1pf_cdef(self.h, C.byref(frame.abi), side.req.cdef, self._stream(stream)))

    def lr_save_boundaries(self, frame, side, after_cdef, stream=None):
        self._check(self.lib.av1pf_lr_save_boundaries(self.h, C.byref(frame.abi), side.req.lr, int(after_cdef),
                                                      self._stream(stream)))

    def lr(self, frame, side, stream=None):
        self._check(self.lib.av1pf_lr(self.h, C.byref(frame.abi), side.req.lr, self._stream(stream)))

    def filmgrain(self, src, dst, side, stream=None):
        self._check(self.lib.av1pf_filmgrain(self.h, C.byref(src.abi), C.byref(dst.abi), side.req.grain,
                                             self._stream(stream)))

    def superres(self, coded, upscaled, stream=None):
        """Upscale every plane of `coded` into `upscaled` (DeviceFrames)."""
        self._check(self.lib.av1pf_superres(self.h, C.byref(coded.abi), C.byref(upscaled.abi), self._stream(stream)))

    def cdef_search_apply(self, recon, noskip_ptr, damping, decision_ptr, fb_index_ptr, stream=None):
        """CDEF on `recon` (DeviceFrame, in place) with a device-side search
        decision and per-block indices."""
        self._check(self.lib.av1pf_cdef_search_apply(self.h, C.byref(recon.abi), C.c_void_p(noskip_ptr), int(damping),
                                                     C.c_void_p(decision_ptr), C.c_void_p(fb_index_ptr),
                                                     self._stream(stream)))

    def deblock_build(self, records_ptr, params, level_cache_ptr, masks_ptr, stream=None):
        """Build the deblock masks and level cache on the device from a record
        map (device pointers; params an abi.LFBuildParams)."""
        self._check(self.lib.av1pf_deblock_build(self.h, C.c_void_p(records_ptr), C.byref(params),
                                                 C.c_void_p(level_cache_ptr), C.c_void_p(masks_ptr),
                                                 self._stream(stream)))

    def sync(self, stream=None):
        self._check(self.lib.av1pf_sync(self.h, self._stream(stream)))

    def cdef_search_select(self, rec, noskip_dev_ptr, mse_dev_ptr, lam, decision_dev_ptr, fb_index_dev_ptr,
                           stream=None):
        """Device strength selection over the MSE table; decision is a 24-byte
        av1pf_cdef_decision in device memory."""
        self._check(self.lib.av1pf_cdef_search_select(self.h, C.byref(rec.abi), noskip_dev_ptr, mse_dev_ptr,
                                                      C.c_uint64(int(lam)), decision_dev_ptr, fb_index_dev_ptr,
                                                      self._stream(stream)))

    def lr_search(self, src, rec, unit_size, units_dev_ptr, sse_dev_ptr=None, stream=None):
        """Curated encoder LR search on the luma plane; units = hz*vt 16-byte records."""
        self._check(self.lib.av1pf_lr_search(self.h, C.byref(src.abi), C.byref(rec.abi), int(unit_size), units_dev_ptr,
                                             sse_dev_ptr, self._stream(stream)))

    def cdef_search_shortlist(self, src, rec, noskip_ptr, damping, lam, cand_ptr, mse_sub_ptr, dec_ptr, idx_ptr,
                              stream=None):
        self._check(self.lib.av1pf_cdef_search_shortlist(self.h, C.byref(src.abi), C.byref(rec.abi), noskip_ptr,
                                                         int(damping), C.c_uint64(int(lam)), cand_ptr, mse_sub_ptr,
                                                         dec_ptr, idx_ptr, self._stream(stream)))

    def search_learned(self, net, src, rec, noskip_ptr, damping, lam, lr_thr, dec_ptr, idx_ptr, lr_units_ptr,
                       cand_ptr, conf_ptr=None, stream=None):
        self._check(self.lib.av1pf_search_learned(self.h, net.h, C.byref(src.abi), C.byref(rec.abi), noskip_ptr,
                                                  int(damping), C.c_uint64(int(lam)), float(lr_thr), dec_ptr,
                                                  idx_ptr, lr_units_ptr, cand_ptr, conf_ptr, self._stream(stream)))

    def lr_search_exhaustive(self, src, rec, unit_size, units_ptr, sse_ptr, fit_ptr, stream=None):
        """Exhaustive superset LR search: unit records, sse[n][28], fit[n][38] int8."""
        self._check(self.lib.av1pf_lr_search_exhaustive(self.h, C.byref(src.abi), C.byref(rec.abi), int(unit_size),
                                                        units_ptr, sse_ptr, fit_ptr, self._stream(stream)))

    def lr_search_masked(self, src, rec, unit_size, mask_ptr, units_ptr, sse_ptr=None, stream=None):
        self._check(self.lib.av1pf_lr_search_masked(self.h, C.byref(src.abi), C.byref(rec.abi), int(unit_size),
                                                    mask_ptr, units_ptr, sse_ptr, self._stream(stream)))

    def profile(self, enable=True):
        self._check(self.lib.av1pf_profile(self.h, int(enable)))

    def profile_read(self):
        """{stage: (total_ms, launches, algorithmic_bytes)} since profile(True)."""
        ms = (C.c_double * PROFILE_STAGES)()
        n = (C.c_int * PROFILE_STAGES)()
        by = (C.c_double * PROFILE_STAGES)()
        self._check(self.lib.av1pf_profile_read(self.h, ms, n, by))
        return {STAGE_NAMES[i]: (ms[i], n[i], by[i]) for i in range(PROFILE_STAGES) if n[i]}

    def cdef_search(self, src, rec, noskip_dev_ptr, damping, mse_dev_ptr, stream=None):
        self._check(self.lib.av1pf_cdef_search(self.h, C.byref(src.abi), C.byref(rec.abi), noskip_dev_ptr,
                                               damping, mse_dev_ptr, self._stream(stream)))


def cdef_search_finish(mse, skip_fb, lam):
    """Host finish of the CDEF search over an mse[2][nfb][64] uint64 table:
    returns (bits, y_strength[8], uv_strength[8], fb_index[nfb])."""
    lib = load()
    m = np.ascontiguousarray(mse, dtype=np.uint64)
    nfb = m.shape[1]
    sk = np.ascontiguousarray(skip_fb, dtype=np.uint8)
    ys = np.zeros(8, np.uint8)
    uvs = np.zeros(8, np.uint8)
    idx = np.zeros(max(nfb, 1), np.uint8)
    bits = C.c_int()
    rc = lib.av1pf_cdef_search_finish(m.ctypes.data, nfb, sk.ctypes.data, C.c_uint64(int(lam)), ys.ctypes.data,
                                      uvs.ctypes.data, idx.ctypes.data, C.byref(bits))
    if rc != abi.OK:
        raise Av1pfError(rc, "av1pf_cdef_search_finish failed")
    return bits.value, ys, uvs, idx[:nfb]


def cdef_lambda(dc_quant, key_frame=True, gf_update=False, tl_factor=128):
    return int(load().av1pf_cdef_lambda(int(dc_quant), int(key_frame), int(gf_update), int(tl_factor)))


class ProposalNet:
    """The learned-proposal CNN on the device (parameters, workspace, passes)."""

    def __init__(self, runner, max_batch=256):
        self.runner = runner
        self.lib = runner.lib
        h = C.c_void_p()
        runner._check(self.lib.av1pf_nn_create(runner.h, int(max_batch), C.byref(h)))
        self.h = h
        self.max_batch = max_batch
        self.nparams = int(self.lib.av1pf_nn_param_count())

    def close(self):
        if self.h:
            self.lib.av1pf_nn_destroy(self.h)
            self.h = None

    def set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        assert P.size == self.nparams
        self.runner._check(self.lib.av1pf_nn_set_params(self.h, P.ctypes.data))

    def get_params(self):
        P = np.zeros(self.nparams, np.float32)
        self.runner._check(self.lib.av1pf_nn_get_params(self.h, P.ctypes.data))
        return P

    def tiles(self, src, rec, out=None, stream=None):
        """BF16 tiles [nfb][64][64][2] (torch int16 tensor) of a DeviceFrame pair."""
        torch = _torch()
        nfb = ((rec.meta.w + 63) // 64) * ((rec.meta.h + 63) // 64)
        if out is None:
            out = torch.empty((nfb, 64, 64, 2), dtype=torch.int16, device="cuda")
        self.runner._check(self.lib.av1pf_nn_tiles(self.runner.h, C.byref(src.abi), C.byref(rec.abi), out.data_ptr(),
                                                   self.runner._stream(stream)))
        return out

    def forward(self, tiles, n=None, out=None, stream=None):
        torch = _torch()
        n = tiles.shape[0] if n is None else n
        if out is None:
            out = torch.empty((n, 131), dtype=torch.float32, device="cuda")
        self.runner._check(self.lib.av1pf_nn_forward(self.h, tiles.data_ptr(), int(n), out.data_ptr(),
                                                     self.runner._stream(stream)))
        return out

    def infer(self, src, rec, out=None, stream=None):
        """Logits [nfb][131] of a DeviceFrame pair in one fused launch."""
        torch = _torch()
        nfb = ((rec.meta.w + 63) // 64) * ((rec.meta.h + 63) // 64)
        if out is None:
            out = torch.empty((nfb, 131), dtype=torch.float32, device="cuda")
        self.runner._check(self.lib.av1pf_nn_infer(self.h, C.byref(src.abi), C.byref(rec.abi), out.data_ptr(),
                                                   self.runner._stream(stream)))
        return out

    def backward(self, tiles, targets, grads, loss, n=None, stream=None):
        n = tiles.shape[0] if n is None else n
        self.runner._check(self.lib.av1pf_nn_backward(self.h, tiles.data_ptr(), int(n), targets.data_ptr(),
                                                      grads.data_ptr(), loss.data_ptr(), self.runner._stream(stream)))

    def adam(self, grads, lr, weight_decay=0.0, stream=None):
        self.runner._check(self.lib.av1pf_nn_adam(self.h, grads.data_ptr(), float(lr), float(weight_decay),
                                                  self.runner._stream(stream)))
