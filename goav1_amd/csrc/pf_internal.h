// Internal declarations shared by the gfx950 post-filter kernels and the host
// side of libav1pf.  Not part of the public ABI (include/av1pf.h is).
//
// Every stage kernel is batched: one launch covers any number of frames of any
// size, bit depth and chroma format.  The host uploads one FrameDesc per frame
// (device memory) and, per launch, a segment table: segment k covers the
// workgroups [base_k, base_{k+1}) and names the frame and plane they work on.
// A workgroup finds its segment by binary search, so the grid is one flat
// range over all frames and the GPU stays full even for small frames.
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
    int     cw, ch;    // allocation: stride / bytes per sample columns, allocated rows
    int     al;        // alignment in bytes (power of two, at most 16) of base and stride: every
                       // row start is al-aligned, so vector accesses of al bytes at aligned columns are legal
};

struct FrameView {
    PlaneView p[3];
    int nplanes;
    int bd;            // 8, 10, 12
    int ssx, ssy;
    int w, h;          // luma visible size
};

// A pointer into device global memory, as the compiler's global address space
// (the cast lets it use global rather than flat accesses; flat accesses also
// count as LDS traffic, so every LDS wait would wait for them too).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> gmem(T *p)
{
    return (gptr<T>)p;
}

// Plane memory is device global memory: the cast lets the compiler use global
// (not flat) loads and stores for everything addressed from a row pointer, so
// waits on LDS traffic do not also wait on frame accesses.
template <typename T>
__device__ __forceinline__ T *row_ptr(const PlaneView &pv, int y) {
    auto g = (__attribute__((address_space(1))) char *)pv.base;
    return (T *)(g + (int64_t)y * pv.stride);
}

struct DbDesc {
    const uint8_t *lc;             // level cache, 4 bytes per MI
    const av1pf_lf_mask *masks;
    int mi_cols, sbw;
    int sharp;                     // sharpness 0..7 (the kernels derive the limits of a level from it)
};

struct CdefDesc {
    const int8_t *index;
    const uint64_t *noskip;        // may be null (all blocks)
    int unit_cols, unit_rows;
    int damping, strength_count;
    int yfilt, uvfilt;
    uint8_t y_strength[8], uv_strength[8];
    // encoder apply: strengths from a device-side search decision and the
    // unit index from its per-block indices (then index and the strengths
    // above are unused)
    const av1pf_cdef_decision *dec;
    const uint8_t *fb_index;
    int rows8, cols8;              // 8x8 blocks inside the frame
};

// Super-resolution geometry of one plane (the position arithmetic is in
// superres_taps.h).
struct SrPlane {
    int srcw;      // MI-aligned coded width: source samples the filter may read
    int codedw;    // coded plane width (sets step and phase)
    int dstw;      // upscaled plane width
    int rows;      // rows upscaled (the plane's visible height)
    int step, x0;  // position step and initial subpel position, 1/16384 sample
};

struct LrPlaneDesc {
    int type;                      // frame type, 0 = none (plane copied)
    int sr;                        // boundary rows are upscaled from the coded frame (SrPlane of the plane)
    int copy_pad;                  // copy the MI-aligned padding from the LR input (input and output are views
                                   // of one frame; not with super-resolution, whose output keeps its own)
    int us, hz, vt;                // unit size and grid
    const av1pf_lr_unit *units;
    uint16_t *bnd;                 // [stripes][4][bp]: rows y0-2, y0-1, y1, y1+1 of the deblocked frame
    int bp;                        // row pitch of bnd in samples (pw rounded up to 8: 16-byte rows)
    int pw, ph, procw, sh, off, xtiles, stripes;
};

// 32-bit copies of the grain parameters the apply reads (one scalar load)
enum { GK_OVERLAP, GK_CSFL, GK_CLIP, GK_SHIFT, GK_MID, GK_CB_MULT, GK_CB_LMULT, GK_CB_OFF, GK_CR_MULT,
       GK_CR_LMULT, GK_CR_OFF, GK_ACT0, GK_ACT1, GK_ACT2, GK_OCOLS, GK_N };

constexpr int GR_SCALE = 4096;     // expanded scaling table bytes per plane (12-bit samples)
// film-grain working memory per frame: templates, then the expanded scaling tables
constexpr size_t GR_TMPL_BYTES = ((size_t)3 * 73 * 82 * 2 + 255) / 256 * 256;
constexpr size_t GR_FRAME_BYTES = GR_TMPL_BYTES + 3 * GR_SCALE;

struct GrainDesc {
    int32_t k[16];                 // GK_* (GK_MID: matrix identity)
    av1pf_grain_params g;
    int active[3];
    uint16_t row_seed[3][73];      // LFSR state at the start of each template row
    alignas(16) uint8_t luts[3][256];
    int16_t *templates;            // 3 x 73 x 82
    uint8_t *scale;                // 3 x GR_SCALE: the scaling function of every sample value
    uint8_t *offsets;              // [orows][ocols] per-32x32-block random offsets
    int orows, ocols;
    int matrix_identity;
};

// Views of a frame: A the frame (coded size), B scratch of A's shape, D the
// film-grain display frame; with super-resolution U is the caller's upscaled
// frame and S scratch of U's shape.
enum { VIEW_A = 0, VIEW_B = 1, VIEW_D = 2, VIEW_U = 3, VIEW_S = 4, NVIEWS = 5 };

struct FrameDesc {
    FrameView v[NVIEWS];
    int lr_src, lr_dst;            // views the LR stage reads / writes
    int sr_src, sr_dst;            // views the super-resolution stage reads / writes
    int fin;                       // view holding the filtered picture (film grain reads it)
    int ext;                       // view whose borders are edge-replicated (refreshing frames)
    SrPlane sr[3];
    DbDesc db;
    CdefDesc cd;
    LrPlaneDesc lr[3];
    GrainDesc gr;
    int *err;                      // device validation flag
};

struct Seg {
    int frame, plane;              // frame index, plane (or -1)
    int nx;                        // tiles per row of this segment
    int base;                      // first workgroup of the segment
};

// Workgroup -> segment (segs sorted by base; segs[n].base = total).
__device__ __forceinline__ int find_seg(const Seg *__restrict__ segs, int n, int wg)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].base <= wg) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---- work distribution of a stage launch ---------------------------------
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  Neighbouring tiles share halo
// rows and columns, so neighbours should run on one XCD:
//   per_wg > 0: the tiles are cut into chunks of kChunk workgroups' worth of
//     consecutive tiles, chunk c going to XCD c % 8; the i-th workgroup of
//     XCD x takes workgroup slot i % kChunk of chunk (i / kChunk) * 8 + x
//     (grid = ceil(total / per_wg) rounded up to 8 * kChunk; the extra
//     workgroups find no tile);
//   per_wg == 0: a persistent grid of resident workgroups; a workgroup takes
//     tiles one at a time from its XCD's queue (8 counters, zero before the
//     launch, over 8 contiguous runs of the tiles) and then from the others.
// Every workgroup leaves the loop on the same shared value, so every wave
// reaches the exit.
constexpr int kXcds = 8;
constexpr int kChunk = 8;
struct StageLaunch {
    int per_wg;
    unsigned int *queue;   // kXcds counters (persistent mode)
};

template <typename Body>
__device__ __forceinline__ void for_each_unit(int total, const StageLaunch &L, Body &&body)
{
    __shared__ int next_unit;
    const int b = blockIdx.x, x = b % kXcds;
    if (L.per_wg > 0) {
        const int i = b / kXcds;
        const int lw = ((i / kChunk) * kXcds + x) * kChunk + i % kChunk;   // logical workgroup
        const int u0 = lw * L.per_wg;
        for (int k = 0; k < L.per_wg; k++) {
            const int u = u0 + k;
            if (u >= total) return;
            if (k) __syncthreads();   // the previous unit is done with the LDS
            [[clang::always_inline]] body(u);   // inlined even when the tile body has other callers
        }
    } else {
        for (int q = 0; q < kXcds; q++) {
            const int qq = (x + q) % kXcds;
            const int lo = (int)((int64_t)total * qq / kXcds), hi = (int)((int64_t)total * (qq + 1) / kXcds);
            for (;;) {
                __syncthreads();           // the previous unit is done with the LDS and next_unit
                if (threadIdx.x == 0) next_unit = lo + (int)atomicAdd(L.queue + qq, 1u);
                __syncthreads();
                const int u = next_unit;
                if (u >= hi) break;
                [[clang::always_inline]] body(u);
            }
        }
    }
}

// Grid of a stage launch: ceil(total / per_wg) rounded up to 8 * kChunk, or
// the resident workgroups of the kernel (occupancy x compute units, at most
// total) when persistent.
int stage_grid(const void *kernel, int total, int per_wg);

// ---- launches (all batched) ------------------------------------------------
constexpr int DB_ROWS = 8;         // rows of 4x4 units per deblock workgroup (64 units wide)
hipError_t launch_deblock(const FrameDesc *d, const Seg *segv, int nv, int totv, const Seg *segh, int nh, int toth,
                          hipStream_t s);
hipError_t launch_lr_save(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s);
hipError_t launch_cdef(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L, hipStream_t s);
hipError_t launch_lr(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L, hipStream_t s);
// src < 0: each frame's fin view.  Segments tile the aligned extent in
// COPY_COLS x COPY_ROWS blocks.
constexpr int COPY_COLS = 512;
constexpr int COPY_ROWS = 4;
hipError_t launch_copy(const FrameDesc *d, const Seg *segs, int nseg, int total, int src, int dst, hipStream_t s);
hipError_t launch_superres(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s);
// Border extension of each frame's ext view.  Segment plane p < 3: the right
// padding of plane p's visible rows; p >= 3: the rows of plane p - 3 below the
// visible height.  Tiles of EXT_COLS columns x EXT_ROWS rows.
constexpr int EXT_COLS = 64;
constexpr int EXT_ROWS = 4;
hipError_t launch_extend(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s);
hipError_t launch_deblock_build(const av1pf_lf_record *rec, const av1pf_lf_build_params &p, uint8_t *level_cache,
                                av1pf_lf_mask *masks, hipStream_t s);
constexpr int SR_COLS = 512;       // output columns per super-resolution tile
constexpr int SR_ROWS = 8;         // rows per super-resolution chunk
constexpr int SR_CHUNKS = 1;       // chunks per super-resolution tile
hipError_t launch_grain_templates(const FrameDesc *d, const Seg *luma, int nl, const Seg *chroma, int nc, hipStream_t s);
hipError_t launch_grain_apply(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L,
                              hipStream_t s);
hipError_t launch_zero_u32(unsigned int *p, int n, hipStream_t s);

// ---- encoder CDEF search (single frame) ----------------------------------
hipError_t launch_cdef_search(const FrameView &src, const FrameView &rec, const uint64_t *noskip,
                              int damping, uint64_t *mse, int nhfb, int nvfb, hipStream_t s);

// One launch (every workgroup resident; refused when the grid does not fit the
// device); epoch counts the
// launches on this state from 1 (the state is zero before the first); bit 1
// of err is set if a barrier did not complete.
hipError_t launch_cdef_select(const uint64_t *mse, const uint64_t *noskip, int nhfb, int nvfb, int mi_cols, int mi_rows,
                              uint64_t lambda, void *state, unsigned int epoch, av1pf_cdef_decision *dec,
                              uint8_t *fb_index, int *err, hipStream_t s);
size_t finish_state_bytes();
hipError_t launch_lr_search(const PlaneView &src, const PlaneView &rec, int bd, int us, int hz, int vt,
                            const uint8_t *mask, av1pf_lr_unit *out, uint64_t *sse, hipStream_t s);
hipError_t launch_lr_search_exhaustive(const PlaneView &src, const PlaneView &rec, int us, int hz, int vt,
                                       av1pf_lr_unit *out, uint64_t *sse, int8_t *fit, hipStream_t s);
hipError_t launch_lr_mask(const float *logits, int nhfb, int nvfb, int hz, int vt, float thr, uint8_t *mask,
                          hipStream_t s);
constexpr int kMaxSearchBlocks = 4096;   // filter blocks the device selection holds

// ---- fused chain (chain.hip) ------------------------------------------------
// Stages of the fused launch, in dependency order; every stage but LR keeps a
// completion counter per frame and 64-luma-row band.
enum { CS_DBV = 0, CS_DBH, CS_SAVE, CS_CDEF, CS_SR, CS_LR, CS_N };
// One tile of the queue: the stage's tile (x, y) of plane `plane` (deblock
// band, context-row chunk and stripe, CDEF unit column and row,
// super-resolution tile, LR processing unit and stripe).
struct ChainItem {
    uint16_t frame;
    uint8_t stage, plane;
    uint16_t x, y;
};
// Per frame: band count, the first of its counters ([CS_N][rows + 1], stage
// major) and the stages whose output super-resolution and LR read (-1: an
// input of the call).
struct ChainFrame {
    int rows, cbase;
    int sr_src, lr_src;
};
// The queue is kXcds parts, part q = items[qbase[q] .. qbase[q + 1]); head
// holds a take counter per part, then the abort word, all zero at launch, as
// are the counters cnt (targets tgt, same layout).
hipError_t launch_chain(const FrameDesc *d, const ChainFrame *frames, const ChainItem *items, int nitems,
                        const unsigned *qbase, unsigned *head, unsigned *cnt, const unsigned *tgt, hipStream_t s);

// tile geometry helpers shared by host and device
__host__ __device__ inline int lr_stripes(int ph, int sh, int off) { return (ph + off + sh - 1) / sh; }

}  // namespace pf
