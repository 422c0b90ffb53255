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
// edge of that direction writes, so the whole vertical pass is one launch
// (across all planes and frames), then the horizontal pass.  A workgroup takes
// 64 x DB_ROWS 4x4 units of a plane: it gathers the units with an edge into one
// LDS list per filter width, then filters the lists with a thread per unit
// (its 4 lines; every line is independent inside dav1d's loop_filter()), so no
// lane idles on a unit without an edge and a wave's lanes share the width.
#include "pf_internal.h"

namespace pf {

__device__ __forceinline__ int iabs_(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ int iclip_(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One line across the edge: dav1d loop_filter() body (loopfilter/sb.go:348-446)
// on values already in registers: a[0..7] = p3 p2 p1 p0 q0 q1 q2 q3, w[0..5] =
// p6 p5 p4 q4 q5 q6 (read only for wd 16).  The new values go back into a and
// w; the return value says which samples changed: 0 none, 2 p0 q0, 4 p1..q1,
// 6 p2..q2, 12 p5..q5.
__device__ __forceinline__ int lpf_vals(int *a, int *w, int bd, int E, int I, int H, int wd)
{
    const int s = bd - 8;
    const int F = 1 << s;
    E <<= s; I <<= s; H <<= s;
    const unsigned twoI = 2u * (unsigned)I, twoF = 2u * (unsigned)F, twoH = 2u * (unsigned)H;
    const int maxPix = (1 << bd) - 1;
    const int dmin = -(128 << s), dmax = (128 << s) - 1;
    const int p3 = a[0], p2 = a[1], p1 = a[2], p0 = a[3], q0 = a[4], q1 = a[5], q2 = a[6], q3 = a[7];
    if ((unsigned)(p1 - p0 + I) > twoI || (unsigned)(q1 - q0 + I) > twoI ||
        iabs_(p0 - q0) * 2 + (iabs_(p1 - q1) >> 1) > E)
        return 0;
    if (wd > 4) {
        if ((unsigned)(p2 - p1 + I) > twoI || (unsigned)(q2 - q1 + I) > twoI) return 0;
        if (wd > 6 && ((unsigned)(p3 - p2 + I) > twoI || (unsigned)(q3 - q2 + I) > twoI)) return 0;
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
            const int p6 = w[0], p5 = w[1], p4 = w[2], q4 = w[3], q5 = w[4], q6 = w[5];
            if ((unsigned)(p6 - p0 + F) <= twoF && (unsigned)(p5 - p0 + F) <= twoF &&
                (unsigned)(p4 - p0 + F) <= twoF && (unsigned)(q4 - q0 + F) <= twoF &&
                (unsigned)(q5 - q0 + F) <= twoF && (unsigned)(q6 - q0 + F) <= twoF) {
                w[1] = (p6 * 7 + p5 * 2 + p4 * 2 + p3 + p2 + p1 + p0 + q0 + 8) >> 4;
                w[2] = (p6 * 5 + p5 * 2 + p4 * 2 + p3 * 2 + p2 + p1 + p0 + q0 + q1 + 8) >> 4;
                a[0] = (p6 * 4 + p5 + p4 * 2 + p3 * 2 + p2 * 2 + p1 + p0 + q0 + q1 + q2 + 8) >> 4;
                a[1] = (p6 * 3 + p5 + p4 + p3 * 2 + p2 * 2 + p1 * 2 + p0 + q0 + q1 + q2 + q3 + 8) >> 4;
                a[2] = (p6 * 2 + p5 + p4 + p3 + p2 * 2 + p1 * 2 + p0 * 2 + q0 + q1 + q2 + q3 + q4 + 8) >> 4;
                a[3] = (p6 + p5 + p4 + p3 + p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + q2 + q3 + q4 + q5 + 8) >> 4;
                a[4] = (p5 + p4 + p3 + p2 + p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + q3 + q4 + q5 + q6 + 8) >> 4;
                a[5] = (p4 + p3 + p2 + p1 + p0 + q0 * 2 + q1 * 2 + q2 * 2 + q3 + q4 + q5 + q6 * 2 + 8) >> 4;
                a[6] = (p3 + p2 + p1 + p0 + q0 + q1 * 2 + q2 * 2 + q3 * 2 + q4 + q5 + q6 * 3 + 8) >> 4;
                a[7] = (p2 + p1 + p0 + q0 + q1 + q2 * 2 + q3 * 2 + q4 * 2 + q5 + q6 * 4 + 8) >> 4;
                w[3] = (p1 + p0 + q0 + q1 + q2 + q3 * 2 + q4 * 2 + q5 * 2 + q6 * 5 + 8) >> 4;
                w[4] = (p0 + q0 + q1 + q2 + q3 + q4 * 2 + q5 * 2 + q6 * 7 + 8) >> 4;
                return 12;
            }
        }
        if (wd >= 8) {
            a[1] = (p3 * 3 + p2 * 2 + p1 + p0 + q0 + 4) >> 3;
            a[2] = (p3 * 2 + p2 + p1 * 2 + p0 + q0 + q1 + 4) >> 3;
            a[3] = (p3 + p2 + p1 + p0 * 2 + q0 + q1 + q2 + 4) >> 3;
            a[4] = (p2 + p1 + p0 + q0 * 2 + q1 + q2 + q3 + 4) >> 3;
            a[5] = (p1 + p0 + q0 + q1 * 2 + q2 + q3 * 2 + 4) >> 3;
            a[6] = (p0 + q0 + q1 + q2 * 2 + q3 * 3 + 4) >> 3;
            return 6;
        }
        a[2] = (p2 * 3 + p1 * 2 + p0 * 2 + q0 + 4) >> 3;
        a[3] = (p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + 4) >> 3;
        a[4] = (p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + 4) >> 3;
        a[5] = (p0 + q0 * 2 + q1 * 2 + q2 * 3 + 4) >> 3;
        return 4;
    }
    const bool hev = (unsigned)(p1 - p0 + H) > twoH || (unsigned)(q1 - q0 + H) > twoH;
    int f = hev ? iclip_(iclip_(p1 - q1, dmin, dmax) + 3 * (q0 - p0), dmin, dmax)
                : iclip_(3 * (q0 - p0), dmin, dmax);
    const int f1 = min(f + 4, dmax) >> 3;
    const int f2 = min(f + 3, dmax) >> 3;
    a[3] = iclip_(p0 + f2, 0, maxPix);
    a[4] = iclip_(q0 - f1, 0, maxPix);
    if (hev) return 2;
    f = (f1 + 1) >> 1;
    a[2] = iclip_(p1 + f, 0, maxPix);
    a[5] = iclip_(q1 - f, 0, maxPix);
    return 4;
}

// Write back samples [LO, 7 - LO] of a (and p5 p4 q4 q5 from w when LO is 0)
// for all 4 lines, unconditionally and in address order (the 4 lines of a
// sample position first when the lines are adjacent, else a line's samples
// first), so neighbouring samples merge into wide stores.
template <int LO, bool ACROSS, int NL, typename T>
__device__ __forceinline__ void lpf_store(gptr<T> p, int64_t lstep, int64_t st, const int (*a)[8], const int (*w)[6])
{
    auto put = [&](int i, int k, int v) { p[i * lstep + k * st] = (T)v; };
    if constexpr (ACROSS) {
        if constexpr (LO == 0) {
#pragma unroll
            for (int i = 0; i < NL; i++) put(i, -6, w[i][1]);
#pragma unroll
            for (int i = 0; i < NL; i++) put(i, -5, w[i][2]);
        }
#pragma unroll
        for (int j = LO; j <= 7 - LO; j++)
#pragma unroll
            for (int i = 0; i < NL; i++) put(i, j - 4, a[i][j]);
        if constexpr (LO == 0) {
#pragma unroll
            for (int i = 0; i < NL; i++) put(i, 4, w[i][3]);
#pragma unroll
            for (int i = 0; i < NL; i++) put(i, 5, w[i][4]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < NL; i++) {
            if constexpr (LO == 0) {
                put(i, -6, w[i][1]);
                put(i, -5, w[i][2]);
            }
#pragma unroll
            for (int j = LO; j <= 7 - LO; j++) put(i, j - 4, a[i][j]);
            if constexpr (LO == 0) {
                put(i, 4, w[i][3]);
                put(i, 5, w[i][4]);
            }
        }
    }
}

// NL lines of one edge unit: q0 of line i at p + i * lstep, samples across
// the edge st apart (elements).  Every line's p3..q3 is loaded before any is
// used, and p6..p4 / q4..q6 of every line together when the edge is 16 wide.
// The lines write back the widest span any of them changed: a line that changed less rewrites values it read, and no other edge
// of the same direction writes a sample this edge reads (transform spacing), so
// the rewrite is safe in parallel.
template <bool ACROSS, int NL, typename T>
__device__ __forceinline__ void lpf_lines(gptr<T> p, int64_t lstep, int64_t st, int bd, int E, int I, int H, int wd)
{
    int a[NL][8], w[NL][6];
#pragma unroll
    for (int i = 0; i < NL; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) a[i][j] = p[i * lstep + (j - 4) * st];
    if (wd >= 16) {
#pragma unroll
        for (int i = 0; i < NL; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                w[i][j] = p[i * lstep + (j - 7) * st];
                w[i][3 + j] = p[i * lstep + (j + 4) * st];
            }
    }
    int n = 0;
#pragma unroll
    for (int i = 0; i < NL; i++) n = max(n, lpf_vals(a[i], w[i], bd, E, I, H, wd));
    if (n == 12) lpf_store<0, ACROSS, NL, T>(p, lstep, st, a, w);
    else if (n == 6) lpf_store<1, ACROSS, NL, T>(p, lstep, st, a, w);
    else if (n == 4) lpf_store<2, ACROSS, NL, T>(p, lstep, st, a, w);
    else if (n == 2) lpf_store<3, ACROSS, NL, T>(p, lstep, st, a, w);
}

// A 16-wide unit reads 14 samples a line: its lines go in two pairs, so the
// samples in flight stay within the registers of the narrower widths.
template <bool ACROSS, typename T>
__device__ __forceinline__ void lpf_unit(gptr<T> p, int64_t lstep, int64_t st, int bd, int E, int I, int H, int wd)
{
    if (wd >= 16) {
        lpf_lines<ACROSS, 2, T>(p, lstep, st, bd, E, I, H, wd);
        lpf_lines<ACROSS, 2, T>(p + 2 * lstep, lstep, st, bd, E, I, H, wd);
    } else {
        lpf_lines<ACROSS, 4, T>(p, lstep, st, bd, E, I, H, wd);
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

// Bit n of a lane whose two 16-bit halves (the second starting at bit
// `lane`) are one 32-bit word.
__device__ __forceinline__ uint32_t mask_bit(uint32_t halves, int n, int lane)
{
    return n < lane ? (halves >> n) & 1u : (halves >> (16 + n - lane)) & 1u;
}

// Limits of a level (loopfilter thresholds: the interior limit shifted by the
// sharpness and capped, the edge limit 2 * (level + 2) + interior).
__device__ __forceinline__ void db_limits(int L, int sharp, int &E, int &I)
{
    int lim = L >> ((sharp > 0) + (sharp > 4));
    if (sharp > 0 && lim > 9 - sharp) lim = 9 - sharp;
    I = max(lim, 1);
    E = 2 * (L + 2) + I;
}

// Width class and level of an edge unit: the unit's mask words and both level
// candidates (the unit's cell, then the previous cell across the edge) are
// loaded before either is used.
__device__ __forceinline__ int unit_edge(const DbDesc &db, const DbPlaneK &P, int dir, int rr, int rc, int pos,
                                         int n, int li, int prev, int comp, int &L)
{
    const auto m = gmem(db.masks) + rr * db.sbw + rc;
    const int lane = dir == 0 ? P.lane_v : P.lane_h;
    uint32_t w0, w1, w2 = 0;
    if (P.luma) {
        const auto q = (gptr<const uint32_t>)&m->y[dir][pos][0][0];
        w0 = q[0];
        w1 = q[1];
        w2 = q[2];
    } else {
        const auto q = (gptr<const uint32_t>)&m->uv[dir][pos][0][0];
        w0 = q[0];
        w1 = q[1];
    }
    const auto lc = gmem(db.lc);
    const int l0 = lc[4 * li + comp], l1 = lc[4 * (li - prev) + comp];
    L = l0 ? l0 : l1;
    return unit_wd(mask_bit(w0, n, lane), mask_bit(w1, n, lane), P.luma ? mask_bit(w2, n, lane) : 0, P.luma);
}

// LDS of a deblock tile: the edge units of each width class, and their counts.
struct DbLds {
    uint32_t list[3][64 * DB_ROWS];
    int cnt[3];
};

// Append one entry per lane with p set to an LDS list: one LDS atomic per
// wave, each lane's slot from the count of set lanes below it.  Returns the
// lane's slot (-1 when p is false).
__device__ __forceinline__ int wave_append(int *cnt, bool p)
{
    const uint64_t b = __ballot(p);
    if (!b) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)b) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __popcll(b));
    base = __shfl(base, leader);
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
    return p ? base + below : -1;
}

// One direction of a plane's edges (DIR 0: vertical edges, filtered along
// rows; 1: horizontal edges, filtered along columns).  A workgroup covers 64
// columns by DB_ROWS rows of 4x4 units.  First every unit's mask bits and level
// are read (a thread per unit column, DB_ROWS / 4 rows per thread, all loads
// in flight together) and the units with an edge to filter are appended to an
// LDS list per filter width; then the lists are filtered width by width, a
// thread per unit, so every lane has an edge and the lanes of a wave share the
// width (same loads, same filter branches).
// The LDS the tile uses (DbLds) is passed in, so the stage kernels and the
// chain kernel share the body.
template <int DIR>
__device__ __forceinline__ void deblock_tile(const FrameDesc &D, int pl, int tx, int ty, DbLds &S)
{
    uint32_t (*list)[64 * DB_ROWS] = S.list;
    int *cnt = S.cnt;
    const FrameView &f = D.v[VIEW_A];
    const DbPlaneK P = db_plane(f, pl);
    const int tid = threadIdx.x;
    if (tid < 3) cnt[tid] = 0;
    __syncthreads();
    constexpr int K = DB_ROWS / 4;
    int wd[K], L[K];
    const int lx = tid & 63;
    const int x4 = tx * 64 + lx;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int ly = (tid >> 6) + 4 * k;
        const int y4 = ty * DB_ROWS + ly;
        wd[k] = 0;
        L[k] = 0;
        const bool in = x4 < P.w4 && y4 < P.h4 && (DIR == 0 ? x4 > 0 : y4 > 0);
        if (in) {
            if (DIR == 0)
                wd[k] = unit_edge(D.db, P, 0, y4 >> P.lrch, x4 >> P.lrcw, x4 & (P.rcw - 1), y4 & (P.rch - 1),
                                  y4 * D.db.mi_cols + x4, 1, P.comp_v, L[k]);
            else
                wd[k] = unit_edge(D.db, P, 1, y4 >> P.lrch, x4 >> P.lrcw, y4 & (P.rch - 1), x4 & (P.rcw - 1),
                                  y4 * D.db.mi_cols + x4, D.db.mi_cols, P.comp_h, L[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int ly = (tid >> 6) + 4 * k;
        const bool on = wd[k] && L[k];
        const int cls = wd[k] >= 16 ? 0 : wd[k] >= 6 ? 1 : 2;
        const uint32_t e = (uint32_t)lx | ((uint32_t)ly << 6) | ((uint32_t)L[k] << 10);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int slot = wave_append(&cnt[c], on && cls == c);
            if (slot >= 0) list[c][slot] = e;
        }
    }
    __syncthreads();
    const PlaneView &pv = f.p[pl];
    // the three lists laid end to end from wave boundaries: one pass covers
    // all of them when they fit the workgroup's lanes, and every wave still
    // filters a single width
    const int n0 = cnt[0], n1 = cnt[1], n2 = cnt[2];
    const int o1 = (n0 + 63) & ~63, o2 = o1 + ((n1 + 63) & ~63), tot = o2 + ((n2 + 63) & ~63);
    for (int i0 = tid; i0 < tot; i0 += 256) {
        const int c = i0 < o1 ? 0 : i0 < o2 ? 1 : 2;   // uniform over a wave
        const int i = i0 - (c == 0 ? 0 : c == 1 ? o1 : o2);
        const int w = c == 0 ? 16 : c == 1 ? (P.luma ? 8 : 6) : 4;
        if (i < (c == 0 ? n0 : c == 1 ? n1 : n2)) {
            const uint32_t e = list[c][i];
            const int ex = tx * 64 + (int)(e & 63), ey = ty * DB_ROWS + (int)((e >> 6) & 15), el = (int)(e >> 10);
            int E, I;
            db_limits(el, D.db.sharp, E, I);
            if (f.bd == 8) {
                const auto q = gmem(row_ptr<uint8_t>(pv, ey * 4) + ex * 4);
                if (DIR == 0) lpf_unit<false, uint8_t>(q, pv.stride, 1, 8, E, I, el >> 4, w);
                else lpf_unit<true, uint8_t>(q, 1, pv.stride, 8, E, I, el >> 4, w);
            } else {
                const auto q = gmem(row_ptr<uint16_t>(pv, ey * 4) + ex * 4);
                if (DIR == 0) lpf_unit<false, uint16_t>(q, pv.stride / 2, 1, f.bd, E, I, el >> 4, w);
                else lpf_unit<true, uint16_t>(q, 1, pv.stride / 2, f.bd, E, I, el >> 4, w);
            }
        }
    }
}

template <int DIR>
__device__ __forceinline__ void deblock_dir(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs, int nseg)
{
    __shared__ DbLds S;
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const int t = blockIdx.x - sg.base;
    deblock_tile<DIR>(descs[sg.frame], sg.plane, t % sg.nx, t / sg.nx, S);
}

__global__ __launch_bounds__(256) void deblock_v_kernel(const FrameDesc *__restrict__ descs,
                                                        const Seg *__restrict__ segs, int nseg)
{
    deblock_dir<0>(descs, segs, nseg);
}

__global__ __launch_bounds__(256) void deblock_h_kernel(const FrameDesc *__restrict__ descs,
                                                        const Seg *__restrict__ segs, int nseg)
{
    deblock_dir<1>(descs, segs, nseg);
}

hipError_t launch_deblock(const FrameDesc *d, const Seg *segv, int nv, int totv, const Seg *segh, int nh, int toth,
                          hipStream_t s)
{
    if (totv) hipLaunchKernelGGL(deblock_v_kernel, dim3(totv), dim3(256), 0, s, d, segv, nv);
    if (toth) hipLaunchKernelGGL(deblock_h_kernel, dim3(toth), dim3(256), 0, s, d, segh, nh);
    return hipGetLastError();
}

}  // namespace pf
