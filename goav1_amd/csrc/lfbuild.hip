// Deblocking side data built on gfx950 from the loop-filter record map:
// goav1's order-independent mask build (FrameWorkLoopFilterMasks.BuildFromMap,
// threading/loop_filter_masks.go:452-546, with the lfmask builders of
// lfmask/build.go and lfmask/mask.go) and the level cache of the decode-time
// build (LevelCache.fill, build.go:232, with blockLevels,
// loop_filter_masks.go:88-123, and CalcLFValues, lfmask/values.go:57-119).
//
// The sequential build visits blocks and ORs their edges into the masks.  Its
// result has a per-cell form, used here so that every mask word is computed
// independently (no atomics, every word written):
//   - a luma cell has a vertical edge on its left side when it is in its
//     block's first column, or when a transform leaf starts at its column
//     and the block is not an inter skip block; the filter class is the
//     smaller of the clamped log2 widths (0..2) of the leaves on the two
//     sides, the left side at the frame edge counting as 2.  Leaves come from
//     tiling the block with its transform and, for inter blocks, splitting
//     leaves by the split words (depth 0 and 1, bit yOff * 4 + xOff);
//   - chroma cells take the block carrying the shared chroma (the block that
//     covers the bottom-right luma cell of the cell's luma group, the dav1d
//     has_chroma gate); block edges take the class of the record left of the
//     block's first luma column (or above its first row) at the chroma cell's
//     bottom-right luma row (column), 1 at the frame edge; inner edges lie on
//     multiples of the chroma transform size;
//   - horizontal edges likewise with heights and the cell above.
// Four workgroups per 128x128 region (luma and chroma, each direction): a
// thread classifies 4 cells and ORs them into the region's words in LDS.
#include "pf_internal.h"

namespace pf {

namespace {

// tile.BlockSize (dav1d order): width and height in 4x4 units as log2
__constant__ uint8_t kBlkL2W[22] = { 5, 5, 4, 4, 4, 4, 3, 3, 3, 3, 2, 2, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0 };
__constant__ uint8_t kBlkL2H[22] = { 5, 4, 5, 4, 3, 2, 4, 3, 2, 1, 4, 3, 2, 1, 0, 3, 2, 1, 0, 2, 1, 0 };
// tile.TransformSize: log2 width / height in 4x4 units, size after one split
__constant__ uint8_t kTxL2W[19] = { 0, 1, 2, 3, 4, 0, 1, 1, 2, 2, 3, 3, 4, 0, 2, 1, 3, 2, 4 };
__constant__ uint8_t kTxL2H[19] = { 0, 1, 2, 3, 4, 1, 0, 2, 1, 3, 2, 4, 3, 2, 0, 3, 1, 4, 2 };
__constant__ uint8_t kTxSub[19] = { 0, 0, 1, 2, 3, 0, 0, 1, 1, 2, 2, 3, 3, 5, 6, 7, 8, 9, 10 };

struct LfMap {
    const av1pf_lf_record *rec;
    int cols, rows, ssx, ssy, chroma;
};

__device__ __forceinline__ const av1pf_lf_record &cell(const LfMap &M, int x, int y)
{
    return M.rec[(int64_t)y * M.cols + x];
}

__device__ __forceinline__ bool rec_ok(const av1pf_lf_record &r)
{
    return (r.flags & AV1PF_LF_VALID) && r.size < 22 && r.tx_y < 19 && r.tx_uv < 19;
}

// The luma transform leaf covering cell (lx, ly) of block r: clamped width /
// height classes and whether the leaf starts at that column / row.
struct Leaf {
    int cw, ch;
    bool sx, sy;
};

__device__ __forceinline__ Leaf leaf_of(const av1pf_lf_record &r, int lx, int ly)
{
    int tx = r.tx_y;
    int l2w = kTxL2W[tx], l2h = kTxL2H[tx];
    int ux = lx >> l2w, uy = ly >> l2h;
    int x0 = ux << l2w, y0 = uy << l2h;
    if (!(r.flags & AV1PF_LF_INTRA)) {
        int xoff = ux, yoff = uy;
#pragma unroll
        for (int depth = 0; depth < 2; depth++) {
            if (tx == 0 || xoff > 3 || yoff > 3 || !((r.split[depth] >> (yoff * 4 + xoff)) & 1)) break;
            const bool hs = l2w >= l2h, vs = l2h >= l2w;   // split across the width / the height
            const int dx = hs && (lx - x0) >= (1 << (l2w - 1));
            const int dy = vs && (ly - y0) >= (1 << (l2h - 1));
            x0 += dx << (l2w - 1);
            y0 += dy << (l2h - 1);
            xoff = xoff * 2 + dx;
            yoff = yoff * 2 + dy;
            tx = kTxSub[tx];
            l2w = kTxL2W[tx];
            l2h = kTxL2H[tx];
        }
    }
    return { min(l2w, 2), min(l2h, 2), lx == x0, ly == y0 };
}

// class of the leaf that covers cell (x, y), -1 without a valid block there
__device__ __forceinline__ int nb_class(const LfMap &M, int x, int y, bool vertical)
{
    const av1pf_lf_record &r = cell(M, x, y);
    if (!rec_ok(r)) return -1;
    const int lx = x - r.mi_col, ly = y - r.mi_row;
    if (lx < 0 || ly < 0 || lx >= (1 << kBlkL2W[r.size]) || ly >= (1 << kBlkL2H[r.size])) return -1;
    const Leaf f = leaf_of(r, lx, ly);
    return vertical ? f.cw : f.ch;
}

// Filter class of the luma edge on the left (dir 0) or top (dir 1) side of
// cell (x, y), -1 for no edge.
__device__ int luma_edge(const LfMap &M, int dir, int x, int y)
{
    if (x >= M.cols || y >= M.rows) return -1;
    const av1pf_lf_record &r = cell(M, x, y);
    if (!rec_ok(r)) return -1;
    const int lx = x - r.mi_col, ly = y - r.mi_row;
    if (lx < 0 || ly < 0 || lx >= (1 << kBlkL2W[r.size]) || ly >= (1 << kBlkL2H[r.size])) return -1;
    const Leaf f = leaf_of(r, lx, ly);
    const bool v = dir == 0;
    const int own = v ? f.cw : f.ch;
    if (v ? lx == 0 : ly == 0) {   // block edge
        int ctx = 2;
        if (v ? x > 0 : y > 0) {
            const int c = v ? nb_class(M, x - 1, y, true) : nb_class(M, x, y - 1, false);
            if (c >= 0) ctx = c;
        }
        return min(own, ctx);
    }
    const bool inner = !(r.flags & AV1PF_LF_SKIP) || (r.flags & AV1PF_LF_INTRA);
    if (!inner || !(v ? f.sx : f.sy)) return -1;
    const Leaf g = v ? leaf_of(r, lx - 1, ly) : leaf_of(r, lx, ly - 1);
    return min(own, v ? g.cw : g.ch);
}

__device__ __forceinline__ int uv_class(int tx, bool vertical) { return (vertical ? kTxL2W[tx] : kTxL2H[tx]) != 0; }

__device__ __forceinline__ bool carries_chroma(const LfMap &M, const av1pf_lf_record &r)
{
    return rec_ok(r) && (r.flags & AV1PF_LF_HAS_UV) && ((1 << kBlkL2W[r.size]) > M.ssx || (r.mi_col & 1)) &&
           ((1 << kBlkL2H[r.size]) > M.ssy || (r.mi_row & 1));
}

// the block carrying chroma cell (cx, cy), or null
__device__ __forceinline__ const av1pf_lf_record *chroma_owner(const LfMap &M, int cx, int cy)
{
    const int lx = min((cx << M.ssx) + M.ssx, M.cols - 1), ly = min((cy << M.ssy) + M.ssy, M.rows - 1);
    const av1pf_lf_record &r = cell(M, lx, ly);
    return carries_chroma(M, r) ? &r : nullptr;
}

// Filter class (0: 4-tap, 1: 6-tap) of the chroma edge on the left / top side
// of chroma cell (cx, cy), -1 for no edge.
__device__ int chroma_edge(const LfMap &M, int dir, int cx, int cy)
{
    const int ccols = (M.cols + M.ssx) >> M.ssx, crows = (M.rows + M.ssy) >> M.ssy;
    if (cx >= ccols || cy >= crows) return -1;
    const av1pf_lf_record *rp = chroma_owner(M, cx, cy);
    if (!rp) return -1;
    const av1pf_lf_record &r = *rp;
    if (r.mi_col >= M.cols || r.mi_row >= M.rows) return -1;
    const int cbx = r.mi_col >> M.ssx, cby = r.mi_row >> M.ssy;
    const int lx = cx - cbx, ly = cy - cby;
    if (lx < 0 || ly < 0) return -1;
    const bool v = dir == 0;
    const int own = uv_class(r.tx_uv, v);
    if (v ? lx == 0 : ly == 0) {
        int ctx = 1;
        if (v ? r.mi_col > 0 : r.mi_row > 0) {
            const av1pf_lf_record &n = v ? cell(M, r.mi_col - 1, min((cy << M.ssy) + M.ssy, M.rows - 1))
                                         : cell(M, min((cx << M.ssx) + M.ssx, M.cols - 1), r.mi_row - 1);
            if (rec_ok(n) && (n.flags & AV1PF_LF_HAS_UV)) ctx = uv_class(n.tx_uv, v);
        }
        return min(own, ctx);
    }
    const bool inner = !(r.flags & AV1PF_LF_SKIP) || (r.flags & AV1PF_LF_INTRA);
    const int step = v ? kTxL2W[r.tx_uv] : kTxL2H[r.tx_uv];
    if (!inner || ((v ? lx : ly) & ((1 << step) - 1))) return -1;
    return own;
}

// One workgroup per (region, kind): kind 0/1 luma vertical / horizontal
// edges, 2/3 chroma.  Each thread classifies 4 cells of the region's 32 x 32
// (luma) grid and ORs the bits into per-(position, class) words in LDS; then
// the words are stored, every word of the kind written.
__global__ __launch_bounds__(256) void lf_mask_kernel(LfMap M, av1pf_lf_mask *masks, int sbw)
{
    __shared__ uint32_t w[32][3];
    const int region = blockIdx.x >> 2, kind = blockIdx.x & 3;
    const int rx = region % sbw, ry = region / sbw;
    av1pf_lf_mask &m = masks[region];
    const int t = threadIdx.x;
    if (t < 96) (&w[0][0])[t] = 0;
    __syncthreads();
    const int dir = kind & 1;
    const int pos = t & 31;
    if (kind < 2) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int along = (t >> 5) * 4 + k;
            const int x = rx * 32 + (dir == 0 ? pos : along), y = ry * 32 + (dir == 0 ? along : pos);
            const int c = luma_edge(M, dir, x, y);
            if (c >= 0) atomicOr(&w[pos][c], 1u << along);
        }
        __syncthreads();
        if (t < 96) {
            const int p = t / 3, c = t % 3;
            m.y[dir][p][c][0] = (uint16_t)(w[p][c] & 0xffffu);
            m.y[dir][p][c][1] = (uint16_t)(w[p][c] >> 16);
        }
    } else {
        const int npos = 32 >> (dir == 0 ? M.ssx : M.ssy);
        const int nalong = 32 >> (dir == 0 ? M.ssy : M.ssx), hn = nalong >> 1;
        if (M.chroma && pos < npos) {
            const int cx0 = (rx * 32) >> M.ssx, cy0 = (ry * 32) >> M.ssy;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int along = (t >> 5) * 4 + k;
                if (along >= nalong) break;
                const int cx = cx0 + (dir == 0 ? pos : along), cy = cy0 + (dir == 0 ? along : pos);
                const int c = chroma_edge(M, dir, cx, cy);
                if (c >= 0) atomicOr(&w[pos][c], 1u << along);
            }
        }
        __syncthreads();
        if (t < 64) {
            const int p = t >> 1, c = t & 1;
            m.uv[dir][p][c][0] = (uint16_t)(w[p][c] & ((1u << hn) - 1));
            m.uv[dir][p][c][1] = (uint16_t)(w[p][c] >> hn);
        }
    }
}

struct LfLevels {
    uint8_t level_y[2], level_u, level_v;
    uint8_t mode_ref_deltas, delta_lf_multi, seg_enabled, pad;
    int8_t ref_delta[8], mode_delta[2];
    int16_t seg_delta[8][4];
};

__device__ __forceinline__ int lf_value(int base_lvl, int lf_delta, int seg_delta, int ref, int mode, const LfLevels &p)
{
    const int base = min(max(min(max(base_lvl + lf_delta, 0), 63) + seg_delta, 0), 63);
    if (!p.mode_ref_deltas) return base;
    const int sh = base >= 32;
    const int d = ref == 0 ? p.ref_delta[0] : p.mode_delta[mode] + p.ref_delta[ref];
    return min(max(base + d * (1 << sh), 0), 63);
}

// bytes 0..3 of a block's level cell: Y vertical, Y horizontal, U, V
__device__ uint32_t block_levels(const LfLevels &p, const av1pf_lf_record &r)
{
    if (p.level_y[0] == 0 && p.level_y[1] == 0) return 0;
    const int seg = p.seg_enabled ? (r.segment & 7) : 0;
    const int ref = r.ref & 7, mode = r.ref ? (r.mode & 1) : 0;
    const int d0 = r.delta_lf[0];
    const int d1 = p.delta_lf_multi ? r.delta_lf[1] : d0;
    const int d2 = p.delta_lf_multi ? r.delta_lf[2] : d0;
    const int d3 = p.delta_lf_multi ? r.delta_lf[3] : d0;
    const int16_t *sd = p.seg_delta[seg];
    const int s0 = p.seg_enabled ? sd[0] : 0, s1 = p.seg_enabled ? sd[1] : 0;
    const int s2 = p.seg_enabled ? sd[2] : 0, s3 = p.seg_enabled ? sd[3] : 0;
    const uint32_t a = lf_value(p.level_y[0], d0, s0, ref, mode, p);
    const uint32_t b = lf_value(p.level_y[1], d1, s1, ref, mode, p);
    const uint32_t c = p.level_u ? lf_value(p.level_u, d2, s2, ref, mode, p) : 0;
    const uint32_t d = p.level_v ? lf_value(p.level_v, d3, s3, ref, mode, p) : 0;
    return a | (b << 8) | (c << 16) | (d << 24);
}

__global__ __launch_bounds__(256) void lf_level_kernel(LfMap M, LfLevels P, uint32_t *lc)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M.cols * M.rows) return;
    const int x = i % M.cols, y = i / M.cols;
    uint32_t v = lc[i];
    const av1pf_lf_record &r = cell(M, x, y);
    if (rec_ok(r)) v = (v & 0xffff0000u) | (block_levels(P, r) & 0xffffu);
    if (M.chroma && x < ((M.cols + M.ssx) >> M.ssx) && y < ((M.rows + M.ssy) >> M.ssy)) {
        const av1pf_lf_record *o = chroma_owner(M, x, y);
        if (o) v = (v & 0xffffu) | (block_levels(P, *o) & 0xffff0000u);
    }
    lc[i] = v;
}

}  // namespace

hipError_t launch_deblock_build(const av1pf_lf_record *rec, const av1pf_lf_build_params &p, uint8_t *level_cache,
                                av1pf_lf_mask *masks, hipStream_t s)
{
    LfMap M{ rec, p.mi_cols, p.mi_rows, p.ss_x, p.ss_y, !p.mono };
    LfLevels L{};
    L.level_y[0] = p.level_y[0];
    L.level_y[1] = p.level_y[1];
    L.level_u = p.level_u;
    L.level_v = p.level_v;
    L.mode_ref_deltas = p.mode_ref_deltas;
    L.delta_lf_multi = p.delta_lf_multi;
    L.seg_enabled = p.seg_enabled;
    for (int i = 0; i < 8; i++) {
        L.ref_delta[i] = p.ref_delta[i];
        for (int k = 0; k < 4; k++) L.seg_delta[i][k] = p.seg_delta[i][k];
    }
    L.mode_delta[0] = p.mode_delta[0];
    L.mode_delta[1] = p.mode_delta[1];
    const int sbw = (p.mi_cols + 31) >> 5, sbh = (p.mi_rows + 31) >> 5;
    hipLaunchKernelGGL(lf_mask_kernel, dim3(sbw * sbh * 4), dim3(256), 0, s, M, masks, sbw);
    const int n = p.mi_cols * p.mi_rows;
    hipLaunchKernelGGL(lf_level_kernel, dim3((n + 255) / 256), dim3(256), 0, s, M, L, reinterpret_cast<uint32_t *>(level_cache));
    return hipGetLastError();
}

}  // namespace pf
