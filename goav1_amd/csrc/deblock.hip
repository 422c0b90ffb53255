// AV1 deblocking on gfx950, driven by decode-time edge bitmasks.
//
// Semantics follow goav1's mask apply (decoder/postfilter_loopfilter_masks_sb.go:25-190,
// loopfilter/sb.go:99-151,348-446): one 4-sample edge unit per set mask bit, the
// width class picked 16 > 8/6 > 4, the level read from the 4-byte-per-MI level
// cache with the left (vertical edges) or above (horizontal edges) cell as the
// fallback for a zero level, the frame's left column / top row never filtered,
// and units outside the MI(4)-aligned crop masked out.
//
// Parallel decomposition.  The reference walks 128-px region rows, vertical
// edges of every plane then horizontal edges.  Because decoder masks obey AV1's
// transform-spacing rule, the edges of one direction never read a sample another
// edge of that direction writes, so the whole vertical pass is one launch with a
// thread per filtered line (across all planes), then the horizontal pass.  Each
// line is independent inside dav1d's loop_filter(), so a thread per line is the
// reference's per-unit loop split four ways.  The pass is HBM bound: a thread
// touches at most 14 samples of its own row (vertical) or column (horizontal).
#include "pf_internal.h"

namespace pf {

__device__ __forceinline__ int iabs_(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ int iclip_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One line across the edge: dav1d loop_filter() body (loopfilter/sb.go:348-446).
// p points at q0; st steps across the edge (elements).
template <typename T>
__device__ __forceinline__ void lpf_line(T *p, int64_t st, int bd, int E, int I, int H, int wd)
{
    const int s = bd - 8;
    const int F = 1 << s;
    E <<= s; I <<= s; H <<= s;
    const unsigned twoI = 2u * (unsigned)I, twoF = 2u * (unsigned)F, twoH = 2u * (unsigned)H;
    const int maxPix = (1 << bd) - 1;
    const int dmin = -(128 << s), dmax = (128 << s) - 1;
    int p1 = p[-2 * st], p0 = p[-st], q0 = p[0], q1 = p[st];
    if ((unsigned)(p1 - p0 + I) > twoI || (unsigned)(q1 - q0 + I) > twoI ||
        iabs_(p0 - q0) * 2 + (iabs_(p1 - q1) >> 1) > E)
        return;
    int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
    if (wd > 4) {
        p2 = p[-3 * st]; q2 = p[2 * st];
        if ((unsigned)(p2 - p1 + I) > twoI || (unsigned)(q2 - q1 + I) > twoI) return;
        if (wd > 6) {
            p3 = p[-4 * st]; q3 = p[3 * st];
            if ((unsigned)(p3 - p2 + I) > twoI || (unsigned)(q3 - q2 + I) > twoI) return;
        }
    }
    bool flat8in = false;
    if (wd >= 6) {
        flat8in = (unsigned)(p2 - p0 + F) <= twoF && (unsigned)(p1 - p0 + F) <= twoF &&
                  (unsigned)(q1 - q0 + F) <= twoF && (unsigned)(q2 - q0 + F) <= twoF;
        if (wd >= 8 && flat8in)
            flat8in = (unsigned)(p3 - p0 + F) <= twoF && (unsigned)(q3 - q0 + F) <= twoF;
    }
    if (flat8in) {
        if (wd >= 16) {
            int p6 = p[-7 * st], p5 = p[-6 * st], p4 = p[-5 * st];
            int q4 = p[4 * st], q5 = p[5 * st], q6 = p[6 * st];
            if ((unsigned)(p6 - p0 + F) <= twoF && (unsigned)(p5 - p0 + F) <= twoF &&
                (unsigned)(p4 - p0 + F) <= twoF && (unsigned)(q4 - q0 + F) <= twoF &&
                (unsigned)(q5 - q0 + F) <= twoF && (unsigned)(q6 - q0 + F) <= twoF) {
                p[-6 * st] = (T)((p6 * 7 + p5 * 2 + p4 * 2 + p3 + p2 + p1 + p0 + q0 + 8) >> 4);
                p[-5 * st] = (T)((p6 * 5 + p5 * 2 + p4 * 2 + p3 * 2 + p2 + p1 + p0 + q0 + q1 + 8) >> 4);
                p[-4 * st] = (T)((p6 * 4 + p5 + p4 * 2 + p3 * 2 + p2 * 2 + p1 + p0 + q0 + q1 + q2 + 8) >> 4);
                p[-3 * st] = (T)((p6 * 3 + p5 + p4 + p3 * 2 + p2 * 2 + p1 * 2 + p0 + q0 + q1 + q2 + q3 + 8) >> 4);
                p[-2 * st] = (T)((p6 * 2 + p5 + p4 + p3 + p2 * 2 + p1 * 2 + p0 * 2 + q0 + q1 + q2 + q3 + q4 + 8) >> 4);
                p[-st] = (T)((p6 + p5 + p4 + p3 + p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + q2 + q3 + q4 + q5 + 8) >> 4);
                p[0] = (T)((p5 + p4 + p3 + p2 + p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + q3 + q4 + q5 + q6 + 8) >> 4);
                p[st] = (T)((p4 + p3 + p2 + p1 + p0 + q0 * 2 + q1 * 2 + q2 * 2 + q3 + q4 + q5 + q6 * 2 + 8) >> 4);
                p[2 * st] = (T)((p3 + p2 + p1 + p0 + q0 + q1 * 2 + q2 * 2 + q3 * 2 + q4 + q5 + q6 * 3 + 8) >> 4);
                p[3 * st] = (T)((p2 + p1 + p0 + q0 + q1 + q2 * 2 + q3 * 2 + q4 * 2 + q5 + q6 * 4 + 8) >> 4);
                p[4 * st] = (T)((p1 + p0 + q0 + q1 + q2 + q3 * 2 + q4 * 2 + q5 * 2 + q6 * 5 + 8) >> 4);
                p[5 * st] = (T)((p0 + q0 + q1 + q2 + q3 + q4 * 2 + q5 * 2 + q6 * 7 + 8) >> 4);
                return;
            }
        }
        if (wd >= 8) {
            p[-3 * st] = (T)((p3 * 3 + p2 * 2 + p1 + p0 + q0 + 4) >> 3);
            p[-2 * st] = (T)((p3 * 2 + p2 + p1 * 2 + p0 + q0 + q1 + 4) >> 3);
            p[-st] = (T)((p3 + p2 + p1 + p0 * 2 + q0 + q1 + q2 + 4) >> 3);
            p[0] = (T)((p2 + p1 + p0 + q0 * 2 + q1 + q2 + q3 + 4) >> 3);
            p[st] = (T)((p1 + p0 + q0 + q1 * 2 + q2 + q3 * 2 + 4) >> 3);
            p[2 * st] = (T)((p0 + q0 + q1 + q2 * 2 + q3 * 3 + 4) >> 3);
            return;
        }
        p[-2 * st] = (T)((p2 * 3 + p1 * 2 + p0 * 2 + q0 + 4) >> 3);
        p[-st] = (T)((p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + 4) >> 3);
        p[0] = (T)((p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + 4) >> 3);
        p[st] = (T)((p0 + q0 * 2 + q1 * 2 + q2 * 3 + 4) >> 3);
        return;
    }
    const bool hev = (unsigned)(p1 - p0 + H) > twoH || (unsigned)(q1 - q0 + H) > twoH;
    int f = hev ? iclip_(iclip_(p1 - q1, dmin, dmax) + 3 * (q0 - p0), dmin, dmax)
                : iclip_(3 * (q0 - p0), dmin, dmax);
    const int f1 = min(f + 4, dmax) >> 3;
    const int f2 = min(f + 3, dmax) >> 3;
    p[-st] = (T)iclip_(p0 + f2, 0, maxPix);
    p[0] = (T)iclip_(q0 - f1, 0, maxPix);
    if (!hev) {
        f = (f1 + 1) >> 1;
        p[-2 * st] = (T)iclip_(p1 + f, 0, maxPix);
        p[st] = (T)iclip_(q1 - f, 0, maxPix);
    }
}

// Per-plane constants of one deblock segment.
struct DbPlaneK {
    int w4, h4;       // MI(4)-aligned position extent, 4x4 units
    int rcw, rch;     // region size in 4x4 units (32 >> ss)
    int lrcw, lrch;   // their log2
    int lane_v, lane_h;
    int comp_v, comp_h;
    bool luma;
};

__device__ __forceinline__ DbPlaneK db_plane(const FrameView &f, int pl)
{
    DbPlaneK k;
    const int sx = pl ? f.ssx : 0, sy = pl ? f.ssy : 0;
    const int pw = (f.w + sx) >> sx, ph = (f.h + sy) >> sy;
    k.w4 = ((pw + 3) & ~3) >> 2;
    k.h4 = ((ph + 3) & ~3) >> 2;
    k.rcw = 32 >> sx;
    k.rch = 32 >> sy;
    k.lrcw = 5 - sx;
    k.lrch = 5 - sy;
    k.luma = pl == 0;
    k.lane_v = pl ? (16 >> sy) : 16;
    k.lane_h = pl ? (16 >> sx) : 16;
    k.comp_v = pl ? pl + 1 : 0;
    k.comp_h = pl ? pl + 1 : 1;
    return k;
}

// Width (4/6/8/16) from the three class bits, 0 = no edge.
__device__ __forceinline__ int unit_wd(uint32_t m0, uint32_t m1, uint32_t m2, bool luma)
{
    if (m2) return 16;
    if (m1) return luma ? 8 : 6;
    if (m0) return 4;
    return 0;
}

__device__ __forceinline__ uint32_t mask_bit(gptr<const uint16_t> halves, int n, int lane)
{
    return n < lane ? (halves[0] >> n) & 1u : (halves[1] >> (n - lane)) & 1u;
}

// Vertical edges: a 64 x 4 workgroup covers 64 edge columns (x4) by 4 rows of
// 4x4 units (y4); a thread looks up its unit's mask and level once and filters
// the unit's 4 pixel rows.
__global__ __launch_bounds__(256) void deblock_v_kernel(const FrameDesc *__restrict__ descs,
                                                        const Seg *__restrict__ segs, int nseg)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const FrameView &f = D.v[VIEW_A];
    const int pl = sg.plane;
    const DbPlaneK P = db_plane(f, pl);
    const int t = blockIdx.x - sg.base;
    const int tx = t % sg.nx, ty = t / sg.nx;
    const int x4 = tx * 64 + threadIdx.x;
    const int y4 = ty * 4 + threadIdx.y;
    if (x4 == 0 || x4 >= P.w4 || y4 >= P.h4) return;
    const int rr = y4 >> P.lrch, rc = x4 >> P.lrcw;
    const int pos = x4 & (P.rcw - 1), n = y4 & (P.rch - 1);
    const auto &m = gmem(D.db.masks)[rr * D.db.sbw + rc];
    int wd;
    if (P.luma)
        wd = unit_wd(mask_bit(m.y[0][pos][0], n, 16), mask_bit(m.y[0][pos][1], n, 16), mask_bit(m.y[0][pos][2], n, 16), true);
    else
        wd = unit_wd(mask_bit(m.uv[0][pos][0], n, P.lane_v), mask_bit(m.uv[0][pos][1], n, P.lane_v), 0, false);
    if (!wd) return;
    const auto lc = gmem(D.db.lc);
    const int li = y4 * D.db.mi_cols + x4;
    int L = lc[4 * li + P.comp_v];
    if (L == 0) L = lc[4 * (li - 1) + P.comp_v];
    if (L == 0) return;
    const PlaneView &pv = f.p[pl];
    const int E = D.db.E[L], I = D.db.I[L], Hh = L >> 4;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        if (f.bd == 8) lpf_line<uint8_t>(row_ptr<uint8_t>(pv, y4 * 4 + r) + x4 * 4, 1, 8, E, I, Hh, wd);
        else lpf_line<uint16_t>(row_ptr<uint16_t>(pv, y4 * 4 + r) + x4 * 4, 1, f.bd, E, I, Hh, wd);
    }
}

// Horizontal edges: a 64 x 4 workgroup covers 64 columns of 4x4 units (x4) by 4
// edge rows (y4); a thread looks up its unit's mask and level once and filters
// the unit's 4 pixel columns.
__global__ __launch_bounds__(256) void deblock_h_kernel(const FrameDesc *__restrict__ descs,
                                                        const Seg *__restrict__ segs, int nseg)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const FrameView &f = D.v[VIEW_A];
    const int pl = sg.plane;
    const DbPlaneK P = db_plane(f, pl);
    const int t = blockIdx.x - sg.base;
    const int tx = t % sg.nx, ty = t / sg.nx;
    const int x4 = tx * 64 + threadIdx.x;
    const int y4 = ty * 4 + threadIdx.y;
    if (y4 == 0 || y4 >= P.h4 || x4 >= P.w4) return;
    const int rr = y4 >> P.lrch, rc = x4 >> P.lrcw;
    const int pos = y4 & (P.rch - 1), n = x4 & (P.rcw - 1);
    const auto &m = gmem(D.db.masks)[rr * D.db.sbw + rc];
    int wd;
    if (P.luma)
        wd = unit_wd(mask_bit(m.y[1][pos][0], n, 16), mask_bit(m.y[1][pos][1], n, 16), mask_bit(m.y[1][pos][2], n, 16), true);
    else
        wd = unit_wd(mask_bit(m.uv[1][pos][0], n, P.lane_h), mask_bit(m.uv[1][pos][1], n, P.lane_h), 0, false);
    if (!wd) return;
    const auto lc = gmem(D.db.lc);
    const int li = y4 * D.db.mi_cols + x4;
    int L = lc[4 * li + P.comp_h];
    if (L == 0) L = lc[4 * (li - D.db.mi_cols) + P.comp_h];
    if (L == 0) return;
    const PlaneView &pv = f.p[pl];
    const int E = D.db.E[L], I = D.db.I[L], Hh = L >> 4;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (f.bd == 8) lpf_line<uint8_t>(row_ptr<uint8_t>(pv, y4 * 4) + x4 * 4 + c, pv.stride, 8, E, I, Hh, wd);
        else lpf_line<uint16_t>(row_ptr<uint16_t>(pv, y4 * 4) + x4 * 4 + c, pv.stride / 2, f.bd, E, I, Hh, wd);
    }
}

hipError_t launch_deblock(const FrameDesc *d, const Seg *segv, int nv, int totv, const Seg *segh, int nh, int toth,
                          hipStream_t s)
{
    if (totv) hipLaunchKernelGGL(deblock_v_kernel, dim3(totv), dim3(64, 4), 0, s, d, segv, nv);
    if (toth) hipLaunchKernelGGL(deblock_h_kernel, dim3(toth), dim3(64, 4), 0, s, d, segh, nh);
    return hipGetLastError();
}

}  // namespace pf
