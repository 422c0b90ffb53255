// Internal declarations shared by the gfx950 post-filter kernels and the host
// side of libav1pf.  Not part of the public ABI (include/av1pf.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "av1pf.h"

namespace pf {

// One plane as the kernels see it.  Samples are uint8 (8-bit) or uint16.
struct PlaneView {
    char   *base;      // sample (0,0)
    int64_t stride;    // bytes
    int     w, h;      // visible size
    int     aw, ah;    // MI-aligned extent (what deblock / CDEF read and write)
};

struct FrameView {
    PlaneView p[3];
    int nplanes;
    int bd;            // 8, 10, 12
    int ssx, ssy;
    int w, h;          // luma visible size
};

template <typename T>
__device__ __forceinline__ T *row_ptr(const PlaneView &pv, int y) {
    return reinterpret_cast<T *>(pv.base + (int64_t)y * pv.stride);
}

// ---- deblocking --------------------------------------------------------
struct DeblockArgs {
    const uint8_t *level_cache;   // device
    const av1pf_lf_mask *masks;   // device
    int mi_cols, sb128_w;
    uint8_t E[64], I[64];         // per-level limits for the frame sharpness
    int plane_mask;               // which planes are filtered
};
hipError_t launch_deblock(const FrameView &f, const DeblockArgs &a, hipStream_t s);

// ---- CDEF --------------------------------------------------------------
struct CdefArgs {
    const int8_t *index;          // device, per unit
    const uint64_t *noskip;       // device, per unit (may be null = all)
    int unit_cols, unit_rows;
    int damping;
    int strength_count;
    uint8_t y_strength[8], uv_strength[8];
    int yfilt, uvfilt;            // any nonzero luma / chroma strength
};
// src -> dst (out of place); every sample of the aligned extent of dst is written.
hipError_t launch_cdef(const FrameView &src, const FrameView &dst, const CdefArgs &a, hipStream_t s, int *err);

// ---- loop restoration ---------------------------------------------------
struct LrPlaneArgs {
    int type;                     // frame type (0 = none)
    int unit_size, horz, vert;    // unit grid
    const av1pf_lr_unit *units;   // device
    uint16_t *bnd;                // device: saved deblocked rows [stripes][4][pw] (above0, above1, below0, below1)
    int stripes;
};
struct LrArgs { LrPlaneArgs p[3]; };
// save the deblocked context rows of every stripe (called on the deblocked frame)
hipError_t launch_lr_save(const FrameView &f, const LrArgs &a, hipStream_t s);
// src (post-CDEF) -> dst
hipError_t launch_lr(const FrameView &src, const FrameView &dst, const LrArgs &a, hipStream_t s);

// ---- film grain ---------------------------------------------------------
struct GrainArgs {
    av1pf_grain_params g;
    int active[3];
    int16_t *templates;           // device: 3 x 73 x 82
    uint16_t jump_luma[16];       // GF(2) LFSR jump matrices (columns) for one template row
    uint16_t jump_chroma[16];
    uint8_t luts[3][256];
};
hipError_t launch_grain_templates(const GrainArgs &a, int bd, int ssx, int ssy, hipStream_t s);
hipError_t launch_grain_apply(const FrameView &src, const FrameView &dst, const GrainArgs &a,
                              int matrix_identity, hipStream_t s);

// ---- encoder CDEF search -------------------------------------------------
hipError_t launch_cdef_search(const FrameView &src, const FrameView &rec, const uint64_t *noskip,
                              int damping, uint64_t *mse, int nhfb, int nvfb, hipStream_t s);

// plain copy of the aligned extent of every plane
hipError_t launch_copy(const FrameView &src, const FrameView &dst, hipStream_t s);

}  // namespace pf
