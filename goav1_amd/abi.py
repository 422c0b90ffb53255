"""ctypes mirror of include/av1pf.h (the C ABI of libav1pf.so).

These are the plain-data structures that cross the drop-in boundary.  They
mirror goav1's post-filter request types:
  - av1pf_frame            frame.Frame / frame.Layout      (internal/av1/frame/buffer.go:18-139)
  - av1pf_lf_mask          lfmask.FilterMask                (internal/av1/lfmask/mask.go:38-41)
  - av1pf_deblock_params   LoopFilterParams + level cache   (parser/loopfilter.go:14-32,
                                                            threading/loop_filter_masks.go:20-46)
  - av1pf_cdef_params      CDEFParams + CDEF index map      (parser/cdef.go:14-21,
                                                            threading/cdef_index.go:13-18)
  - av1pf_lr_unit/params   RestorationUnit / Params         (tile/restoration.go:22-28,
                                                            parser/restoration.go:15-34)
  - av1pf_grain_params     FilmGrainParams                  (parser/film_grain.go:21-56)
  - av1pf_request          FrameWorkPostFilterRequest       (decoder/postfilter.go:278-284)
"""
import ctypes as C

OK = 0
ERR_INVALID_ARG = -1
ERR_UNSUPPORTED = -2
ERR_DEVICE = -3
ERR_SHORT_BUFFER = -4
ERR_NO_DEVICE = -5

STAGE_DEBLOCK = 0x01
STAGE_CDEF = 0x02
STAGE_SUPERRES = 0x04
STAGE_LR = 0x08
STAGE_FILMGRAIN = 0x10
STAGES_ALL = STAGE_DEBLOCK | STAGE_CDEF | STAGE_LR | STAGE_FILMGRAIN

REQ_REFRESH = 0x01
NOALLOC_SCRATCH_BYTES = 32 << 20

LR_NONE, LR_WIENER, LR_SGRPROJ, LR_SWITCHABLE = 0, 1, 2, 3


class Frame(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32),
        ("ss_x", C.c_int32), ("ss_y", C.c_int32), ("mono", C.c_int32),
        ("matrix_identity", C.c_int32), ("alloc_height", C.c_int32),
        ("plane", C.c_void_p * 3), ("stride", C.c_int64 * 3),
    ]


class LFMask(C.Structure):
    _fields_ = [("y", C.c_uint16 * (2 * 32 * 3 * 2)), ("uv", C.c_uint16 * (2 * 32 * 2 * 2))]


class DeblockParams(C.Structure):
    _fields_ = [
        ("level_y", C.c_uint8 * 2), ("level_u", C.c_uint8), ("level_v", C.c_uint8),
        ("sharpness", C.c_uint8), ("pad", C.c_uint8 * 3),
        ("mi_cols", C.c_int32), ("mi_rows", C.c_int32),
        ("sb128_w", C.c_int32), ("sb128_h", C.c_int32),
        ("level_cache", C.c_void_p), ("masks", C.c_void_p),
    ]


class CdefParams(C.Structure):
    _fields_ = [
        ("damping", C.c_uint8), ("strength_count", C.c_uint8),
        ("y_strength", C.c_uint8 * 8), ("uv_strength", C.c_uint8 * 8), ("pad", C.c_uint8 * 6),
        ("unit_cols", C.c_int32), ("unit_rows", C.c_int32),
        ("index", C.c_void_p), ("noskip", C.c_void_p),
    ]


class LRUnit(C.Structure):
    _fields_ = [
        ("type", C.c_uint8), ("sgr_set", C.c_uint8), ("sgr_xqd", C.c_int8 * 2),
        ("wiener_v", C.c_int8 * 3), ("wiener_h", C.c_int8 * 3), ("pad", C.c_uint8 * 6),
    ]


class LRParams(C.Structure):
    _fields_ = [
        ("frame_type", C.c_uint8 * 3), ("pad0", C.c_uint8),
        ("unit_size", C.c_uint16 * 3), ("pad1", C.c_uint16),
        ("units", C.c_void_p * 3),
    ]


class GrainParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint16),
        ("num_y_points", C.c_uint8), ("num_cb_points", C.c_uint8), ("num_cr_points", C.c_uint8),
        ("chroma_scaling_from_luma", C.c_uint8), ("scaling_shift", C.c_uint8),
        ("ar_coeff_lag", C.c_uint8), ("ar_coeff_shift", C.c_uint8), ("grain_scale_shift", C.c_uint8),
        ("overlap", C.c_uint8), ("clip_to_restricted_range", C.c_uint8),
        ("y_points", C.c_uint8 * 28), ("cb_points", C.c_uint8 * 20), ("cr_points", C.c_uint8 * 20),
        ("ar_coeffs_y", C.c_int8 * 24), ("ar_coeffs_cb", C.c_int8 * 25), ("ar_coeffs_cr", C.c_int8 * 25),
        ("cb_mult", C.c_uint8), ("cb_luma_mult", C.c_uint8), ("cb_offset", C.c_uint16),
        ("cr_mult", C.c_uint8), ("cr_luma_mult", C.c_uint8), ("cr_offset", C.c_uint16),
    ]


class LFBuildParams(C.Structure):
    _fields_ = [("mi_cols", C.c_int32), ("mi_rows", C.c_int32), ("ss_x", C.c_int32), ("ss_y", C.c_int32),
                ("mono", C.c_int32), ("level_y", C.c_uint8 * 2), ("level_u", C.c_uint8), ("level_v", C.c_uint8),
                ("mode_ref_deltas", C.c_uint8), ("delta_lf_multi", C.c_uint8), ("seg_enabled", C.c_uint8),
                ("pad", C.c_uint8), ("ref_delta", C.c_int8 * 8), ("mode_delta", C.c_int8 * 2),
                ("pad2", C.c_int8 * 2), ("seg_delta", (C.c_int16 * 4) * 8)]


LF_VALID, LF_SKIP, LF_INTRA, LF_HAS_UV = 0x01, 0x02, 0x04, 0x08
LF_RECORD_BYTES = 20


class Request(C.Structure):
    _fields_ = [
        ("stages", C.c_uint32), ("flags", C.c_uint32),
        ("deblock", C.POINTER(DeblockParams)), ("cdef", C.POINTER(CdefParams)),
        ("lr", C.POINTER(LRParams)), ("grain", C.POINTER(GrainParams)),
        ("display", C.POINTER(Frame)), ("upscaled", C.POINTER(Frame)),
    ]


assert C.sizeof(LFMask) == 1280
assert C.sizeof(LRUnit) == 16
assert C.sizeof(Frame) == 32 + 24 + 24


# strengths per plane in the learned search's shortlist (AV1PF_SHORTLIST)
SHORTLIST = 4


class CdefDecision(C.Structure):
    """av1pf_cdef_decision: the device strength selection's result."""
    _fields_ = [("cdef_bits", C.c_int32), ("y_strength", C.c_uint8 * 8), ("uv_strength", C.c_uint8 * 8),
                ("pad", C.c_uint8 * 4)]
