// Loop-restoration filters over an LDS tile, shared by the decoder kernel
// (lr.hip) and the encoder search (lr_search.hip).
//
// The tile holds a w x h block (w, h <= 64) plus a 3-sample halo on every
// side, LR_TH rows at pitch LR_TW, sample (0, 0) at row 3, column LR_OX
// (O points there): column LR_OX = 4 puts every fourth sample column at an
// 8-byte boundary, so four adjacent samples are one aligned 64-bit LDS read.
// Each filter hands its outputs to a sink, so the decoder writes the frame and
// the search accumulates squared errors from the same arithmetic:
//   - Wiener (restoration/wiener.go:132-216): H pass with offset 2^(bd+6),
//     round0, clamp to 2^(bd+1+7-round0)-1; V pass with round1;
//   - self-guided (restoration/selfguided.go): box sums, uint32 p/z/A/B math,
//     6/5 (r=2, every other row) and 4/3 (r=1) weighted sums, projection with
//     an int16 wrap before the clip.
// Block size 256 threads; the caller synchronises before the call.  LDS the
// caller provides: the tile, a 70 x 64 uint16 Wiener intermediate, and the
// self-guided A (uint16: at most 256) and B (int32) planes, LR_AB x LR_AB; the
// Wiener intermediate may share memory with A and B.
#pragma once
#include "pf_internal.h"

namespace pf {

#define LR_TW 72          // tile pitch: 4 + 64 + 3, rounded to 8 bytes
#define LR_TH 70          // tile rows: 64 + 6
#define LR_OX 4           // tile column of sample x = 0
#define LR_AB 66          // 64 + 2
#define LR_SGR_ROWS 16    // rows of a self-guided quarter tile (the decoder kernel)

__device__ __forceinline__ int32_t r2s(int32_t v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ uint32_t r2u(uint32_t v, int n) { return n <= 0 ? v : (v + (1u << (n - 1))) >> n; }

static __device__ __constant__ int8_t kSgrR[16][2] = {
    {2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{0,1},{0,1},{0,1},{0,1},{2,0},{2,0}};
static __device__ __constant__ int16_t kSgrS[16][2] = {
    {140,3236},{112,2158},{93,1618},{80,1438},{70,1295},{58,1177},{47,1079},{37,996},
    {30,925},{25,863},{-1,2589},{-1,1618},{-1,1177},{-1,925},{56,-1},{22,-1}};

// AV1's x/(x+1) Q8 table from its closed form floor((513 z + 1) / (2 z + 2)).
// The quotient is never an integer for 1 <= z <= 254 (2z + 2 would have to
// divide 511 - z, i.e. 2m + 1 would divide 512), so it lies at least 1/510
// from one; a float product with the hardware reciprocal (relative error
// under 2^-22, absolute under 1e-4 here) floors to the exact value.
__device__ __forceinline__ int32_t x_by_xplus1(uint32_t z)
{
    const float q = (float)(513u * z + 1u) * __builtin_amdgcn_rcpf((float)(2u * z + 2u));
    return z == 0 ? 1 : (z >= 255 ? 256 : (int32_t)q);
}

// Samples s[K0..K1] of an LDS row as separate 16-bit loads: the odd ones go
// through an opaque copy of the base, so no two loads are adjacent to the load
// merger (a merged 32-bit read is misaligned for half the lanes and stalls).
template <int K0, int K1>
__device__ __forceinline__ void lds_run(const uint16_t *s, int32_t (&v)[K1 - K0 + 1])
{
    const uint16_t *so = s;
    asm volatile("" : "+v"(so));
#pragma unroll
    for (int k = K0; k <= K1; k++) v[k - K0] = (k & 1) ? so[k] : s[k];
}

// Lanes per row for a region w samples wide (a power of two >= w, at most 64):
// lane x of a row group handles column x, row groups step by 256 / lanes.
__device__ __forceinline__ int lanes_for(int w) { return w > 32 ? 64 : w > 16 ? 32 : w > 8 ? 16 : 8; }

// Four samples from one aligned 64-bit LDS word.
__device__ __forceinline__ void unpack4(uint2 v, int32_t *t)
{
    t[0] = (int32_t)(v.x & 0xffffu);
    t[1] = (int32_t)(v.x >> 16);
    t[2] = (int32_t)(v.y & 0xffffu);
    t[3] = (int32_t)(v.y >> 16);
}

// Wiener over the tile.  A lane computes four adjacent columns c..c+3 (c a
// multiple of 4): the H pass reads the twelve samples c-4..c+7 of a row as
// three aligned 64-bit words and writes four intermediates as one word of
// tmp (pitch 64, rows -3..h+2); the V pass reads seven words of tmp.
// sink(y, x, v[4], n) receives the outputs of columns x..x+n-1 (n < 4 only
// at the tile's right edge).
template <typename Sink4>
__device__ __forceinline__ void wiener_tile(const uint16_t *O, int w, int h, int bd, const int8_t *vt, const int8_t *ht,
                                            uint16_t *tmp, Sink4 sink)
{
    const int tid = threadIdx.x;
    const int th = h + 6;
    const int32_t mx = (1 << bd) - 1;
    int round0 = 3, round1 = 11;
    const int ibr = bd + 7 - round0 + 2;
    if (ibr > 16) { round0 += ibr - 16; round1 -= ibr - 16; }
    const int32_t h0 = ht[0], h1 = ht[1], h2 = ht[2];
    const int32_t h3 = -2 * (h0 + h1 + h2);
    const int32_t v0 = vt[0], v1 = vt[1], v2 = vt[2];
    const int32_t v3 = -2 * (v0 + v1 + v2);
    const int32_t limit = 1 << (bd + 1 + 7 - round0), hoff = 1 << (bd + 6);
    const int gpr = lanes_for(w) >> 2, c = (tid & (gpr - 1)) * 4, rstep = 256 / gpr;
    if (c < w)
        for (int rr = tid / gpr; rr < th; rr += rstep) {
            const uint2 *sw = reinterpret_cast<const uint2 *>(O + (rr - 3) * LR_TW + c - 4);
            int32_t t[12];
            unpack4(sw[0], t);
            unpack4(sw[1], t + 4);
            unpack4(sw[2], t + 8);
            uint32_t o[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int32_t sum = (t[k + 4] << 7) + hoff;
                sum += h0 * (t[k + 1] + t[k + 7]) + h1 * (t[k + 2] + t[k + 6]) + h2 * (t[k + 3] + t[k + 5]) +
                       h3 * t[k + 4];
                o[k] = (uint32_t)min(max(r2s(sum, round0), 0), limit - 1);
            }
            *reinterpret_cast<uint2 *>(tmp + rr * 64 + c) = make_uint2(o[0] | (o[1] << 16), o[2] | (o[3] << 16));
        }
    __syncthreads();
    const int32_t voff = 1 << (bd + round1 - 1);
    if (c < w)
        for (int r = tid / gpr; r < h; r += rstep) {
            int32_t s[7][4];
#pragma unroll
            for (int k = 0; k < 7; k++) unpack4(*reinterpret_cast<const uint2 *>(tmp + (r + k) * 64 + c), s[k]);
            int v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int32_t sum = (s[3][j] << 7) - voff;
                sum += v0 * (s[0][j] + s[6][j]) + v1 * (s[1][j] + s[5][j]) + v2 * (s[2][j] + s[4][j]) + v3 * s[3][j];
                v[j] = min(max(r2s(sum, round1), 0), mx);
            }
            sink(r, c, v, min(4, w - c));
        }
}

// Box sum and sum of squares of the (2R+1)-wide row segment centred at s.
template <int R>
__device__ __forceinline__ void row_box(const uint16_t *s, int32_t &sum, int32_t &sq)
{
    int32_t t[2 * R + 1];
    lds_run<-R, R>(s, t);
    sum = 0;
    sq = 0;
#pragma unroll
    for (int i = 0; i <= 2 * R; i++) {
        sum += t[i];
        sq += t[i] * t[i];
    }
}

// A and B of one self-guided pass (radius R, every STEP-th row from -1) for
// the positions x = -1..w: one lane per column, the rows split among
// 256 / (w + 2) row groups; the box sums slide down the column, each row's
// horizontal box sum computed once.
template <int R, int STEP>
__device__ __forceinline__ void sgr_ab(const uint16_t *O, int w, int h, int bd, uint32_t s, uint16_t *Ab, int32_t *Bb)
{
    constexpr int N = (2 * R + 1) * (2 * R + 1);
    const uint32_t recip = (uint32_t)((8192 + N) / (2 * N));
    const int tid = threadIdx.x;
    const int ncols = w + 2;                                   // x = -1 .. w
    const int nrows = STEP == 2 ? (h + 3) / 2 : h + 2;          // y = -1, -1 + STEP, ...
    const int groups = 256 / ncols;                             // row groups (3 for 64-wide tiles)
    const int col = tid % ncols, grp = tid / ncols;
    if (grp >= groups) return;
    const int per = (nrows + groups - 1) / groups;
    const int k0 = grp * per, k1 = min(k0 + per, nrows);
    if (k0 >= k1) return;
    const int x = col - 1;
    int32_t S[2 * R + 1], Q[2 * R + 1];
    {
        const int y = -1 + k0 * STEP;
#pragma unroll
        for (int i = 0; i <= 2 * R; i++) row_box<R>(O + (y - R + i) * LR_TW + x, S[i], Q[i]);
    }
    for (int k = k0; k < k1; k++) {
        const int y = -1 + k * STEP;
        if (k > k0) {
#pragma unroll
            for (int i = 0; i + STEP <= 2 * R; i++) { S[i] = S[i + STEP]; Q[i] = Q[i + STEP]; }
#pragma unroll
            for (int i = 2 * R + 1 - STEP; i <= 2 * R; i++) row_box<R>(O + (y - R + i) * LR_TW + x, S[i], Q[i]);
        }
        int32_t sum = 0, sq = 0;
#pragma unroll
        for (int i = 0; i <= 2 * R; i++) { sum += S[i]; sq += Q[i]; }
        const uint32_t a = (uint32_t)r2s(sq, 2 * (bd - 8));
        const uint32_t b = (uint32_t)r2s(sum, bd - 8);
        const uint32_t p = a * (uint32_t)N >= b * b ? a * (uint32_t)N - b * b : 0u;
        uint32_t z = r2u(p * s, 20);
        z = min(z, 255u);
        const int32_t aa = x_by_xplus1(z);
        Ab[(y + 1) * LR_AB + x + 1] = (uint16_t)aa;
        Bb[(y + 1) * LR_AB + x + 1] = (int32_t)r2u((uint32_t)(256 - aa) * (uint32_t)sum * recip, 12);
    }
}

// Lane layout of the per-pixel filter stages: lane x of row group g handles
// column x of the consecutive rows [ry, end) (at most SGR_PPT; ry even, so a
// row's parity is fixed by its step j).
#define SGR_PPT 16
struct LaneRows {
    int x, ry, end;
};
__device__ __forceinline__ LaneRows lane_rows(int w, int h)
{
    const int lpr = lanes_for(w), groups = 256 / lpr;
    const int rpg = (((h + groups - 1) / groups) + 1) & ~1;   // rows per group, even, <= SGR_PPT for h <= 64
    const int ry = ((int)threadIdx.x / lpr) * rpg;
    return { (int)threadIdx.x & (lpr - 1), ry, min(ry + rpg, h) };
}

// The self-guided filter outputs of set `set` for this lane's pixels (Q4),
// pass by pass (pass 0: r = 2, pass 1: r = 1; a pass of radius 0 is skipped):
// use(pass, j, u, f) gets row ry + j's output f and its centre sample u.
// A lane walks its rows downwards, so the weighted 3x3 sums of A and B
// (selfguided.go) are built from per-row sums kept across steps: with H the
// sum of a row's three values around x and C the centre one,
//   r = 2 (A, B on rows -1, 1, 3, ...): P = 6 C + 5 (left + right) = 5 H + C,
//     an even row takes P(y - 1) + P(y + 1) (shift 9), an odd row P(y) (shift 8);
//   r = 1 (every row): 4 (cross) + 3 (diagonals) = 4 H(y) + Q(y - 1) + Q(y + 1),
//     Q = 3 H + C;
// the same integer sums as the stencils, regrouped.
template <int PPT = SGR_PPT, typename Use>
__device__ __forceinline__ void sgr_passes(const uint16_t *O, int w, int h, int bd, int set, uint16_t *Ab, int32_t *Bb,
                                           Use use)
{
    const LaneRows L = lane_rows(w, h);
    for (int pass = 0; pass < 2; pass++) {
        const int r = kSgrR[set][pass];
        if (r == 0) continue;
        const uint32_t s = (uint32_t)kSgrS[set][pass];
        __syncthreads();
        if (pass == 0) sgr_ab<2, 2>(O, w, h, bd, s, Ab, Bb);
        else sgr_ab<1, 1>(O, w, h, bd, s, Ab, Bb);
        __syncthreads();
        if (L.x >= w || L.ry >= L.end) continue;
        const int x = L.x;
        // the centre A value is read through a second base the load merger
        // cannot relate to the first (an opaque zero offset), so the three
        // 16-bit reads of a row stay separate: merged, they would be
        // misaligned for half the lanes
        int oz = 0;
        asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
        const uint16_t *Az = Ab + oz;
        // H and C of A and B on row y (the tables are offset by one row and column)
        auto row = [&](int y, int32_t &ha, int32_t &ca, int32_t &hb, int32_t &cb) {
            const int i = (y + 1) * LR_AB + x + 1;
            const uint16_t *A = Ab + i;
            const int32_t *B = Bb + i;
            ca = Az[i];
            ha = (int32_t)A[-1] + ca + A[1];
            cb = B[0];
            hb = B[-1] + B[0] + B[1];
        };
        if (pass == 0) {
            int32_t ha, ca, hb, cb;
            row(L.ry - 1, ha, ca, hb, cb);
            int32_t pa = 5 * ha + ca, pb = 5 * hb + cb;
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const int y = L.ry + j;
                if (y >= L.end) break;
                const int32_t uu = O[y * LR_TW + x];
                if ((j & 1) == 0) {
                    row(y + 1, ha, ca, hb, cb);
                    const int32_t na = 5 * ha + ca, nb = 5 * hb + cb;
                    use(0, j, uu, r2s((pa + na) * uu + (pb + nb), 9));
                    pa = na;
                    pb = nb;
                } else {
                    use(0, j, uu, r2s(pa * uu + pb, 8));
                }
            }
        } else {
            int32_t ha0, ca0, hb0, cb0, ha1, ca1, hb1, cb1;
            row(L.ry - 1, ha0, ca0, hb0, cb0);
            row(L.ry, ha1, ca1, hb1, cb1);
            int32_t qa0 = 3 * ha0 + ca0, qb0 = 3 * hb0 + cb0;
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const int y = L.ry + j;
                if (y >= L.end) break;
                int32_t ha2, ca2, hb2, cb2;
                row(y + 1, ha2, ca2, hb2, cb2);
                const int32_t qa2 = 3 * ha2 + ca2, qb2 = 3 * hb2 + cb2;
                const int32_t uu = O[y * LR_TW + x];
                use(1, j, uu, r2s((4 * ha1 + qa0 + qa2) * uu + (4 * hb1 + qb0 + qb2), 9));
                qa0 = 3 * ha1 + ca1;
                qb0 = 3 * hb1 + cb1;
                ha1 = ha2;
                ca1 = ca2;
                hb1 = hb2;
                cb1 = cb2;
            }
        }
    }
}

// The two self-guided outputs of set `set` for this lane's pixels (Q4; zero
// for a pass of radius 0): f0 from the r = 2 pass, f1 from the r = 1 pass.
template <int PPT = SGR_PPT>
__device__ __forceinline__ void sgr_filters(const uint16_t *O, int w, int h, int bd, int set, uint16_t *Ab, int32_t *Bb,
                                            int32_t (&f0)[PPT], int32_t (&f1)[PPT])
{
#pragma unroll
    for (int j = 0; j < PPT; j++) { f0[j] = 0; f1[j] = 0; }
    sgr_passes<PPT>(O, w, h, bd, set, Ab, Bb, [&](int pass, int j, int32_t, int32_t f) {
        if (pass == 0) f0[j] = f;
        else f1[j] = f;
    });
}

// xq of a set from its coded xqd (restoration/selfguided.go:111-126).
__device__ __forceinline__ void sgr_xq(int set, const int8_t *xqd, int &xq0, int &xq1)
{
    if (kSgrR[set][0] == 0) { xq0 = 0; xq1 = 128 - xqd[1]; }
    else if (kSgrR[set][1] == 0) { xq0 = xqd[0]; xq1 = 0; }
    else { xq0 = xqd[0]; xq1 = 128 - xq0 - xqd[1]; }
}

// The projection of one pixel: u + xq0 (f0 - u) + xq1 (f1 - u) in Q11, the
// int16 wrap, then the clip.
__device__ __forceinline__ int sgr_project(int u, int32_t f0, int32_t f1, int xq0, int xq1, int32_t mx)
{
    const int32_t uu = (int32_t)u << 4;
    int32_t v = uu << 7;
    v += xq0 * (f0 - uu);
    v += xq1 * (f1 - uu);
    const int32_t wv = (int16_t)r2s(v, 11);
    return min(max(wv, 0), mx);
}

// PPT: rows per lane, at least the rows lane_rows gives a lane for h (16 for
// h <= 64, 8 for h <= 32).
template <int PPT = SGR_PPT, typename Sink>
__device__ __forceinline__ void sgr_tile(const uint16_t *O, int w, int h, int bd, int set, const int8_t *xqd,
                                         uint16_t *Ab, int32_t *Bb, Sink sink)
{
    int xq0, xq1;
    sgr_xq(set, xqd, xq0, xq1);
    int32_t f0[PPT], f1[PPT];
    sgr_filters<PPT>(O, w, h, bd, set, Ab, Bb, f0, f1);
    const LaneRows L = lane_rows(w, h);
    const int32_t mx = (1 << bd) - 1;
    if (L.x < w) {
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            const int y = L.ry + j;
            if (y >= L.end) break;
            sink(y, L.x, sgr_project(O[y * LR_TW + L.x], f0[j], f1[j], xq0, xq1, mx));
        }
    }
}


// The decoder's self-guided filter, LR_SGR_ROWS rows at a time: both
// passes' A and B planes of the rows (Ab0/Bb0 for r = 2, Ab1/Bb1 for r = 1,
// each (LR_SGR_ROWS + 2) x LR_AB) are built first, then one walk per lane
// forms both outputs of a pixel from the same per-row sums as sgr_passes and
// projects it at once: no per-lane output arrays, and the A/B planes of a
// part of the tile only (a position's A and B depend only on the tile's
// samples, so the rows either side of a seam come out the same).
template <typename Sink>
__device__ __forceinline__ void sgr_tile_parts(const uint16_t *O, int w, int h, int bd, int set, const int8_t *xqd,
                                               uint16_t *Ab0, int32_t *Bb0, uint16_t *Ab1, int32_t *Bb1, Sink sink)
{
    int xq0, xq1;
    sgr_xq(set, xqd, xq0, xq1);
    const int r0 = kSgrR[set][0], r1 = kSgrR[set][1];
    const uint32_t s0 = (uint32_t)kSgrS[set][0], s1 = (uint32_t)kSgrS[set][1];
    const int32_t mx = (1 << bd) - 1;
#pragma unroll 1
    for (int ya = 0; ya < h; ya += LR_SGR_ROWS) {
        const int hh = min(LR_SGR_ROWS, h - ya);
        const uint16_t *Oq = O + ya * LR_TW;
        __syncthreads();   // the previous part's walk has read its planes
        if (r0) sgr_ab<2, 2>(Oq, w, hh, bd, s0, Ab0, Bb0);
        if (r1) sgr_ab<1, 1>(Oq, w, hh, bd, s1, Ab1, Bb1);
        __syncthreads();
        const LaneRows L = lane_rows(w, hh);   // at most 4 rows per lane, ry even
        if (L.x >= w || L.ry >= L.end) continue;
        const int x = L.x;
        int oz = 0;   // opaque zero: keeps the three 16-bit reads of a row separate (see sgr_passes)
        asm volatile("v_mov_b32 %0, 0" : "=v"(oz));
        auto row = [&](const uint16_t *Ab, const int32_t *Bb, int y, int32_t &ha, int32_t &ca, int32_t &hb,
                       int32_t &cb) {
            const int i = (y + 1) * LR_AB + x + 1;
            const uint16_t *A = Ab + i;
            const int32_t *B = Bb + i;
            ca = (Ab + oz)[i];
            ha = (int32_t)A[-1] + ca + A[1];
            cb = B[0];
            hb = B[-1] + B[0] + B[1];
        };
        // r = 2: P = 5 H + C of the odd rows; r = 1: H of the current row and
        // Q = 3 H + C of the rows above and below
        int32_t pa = 0, pb = 0, ha1 = 0, ca1 = 0, hb1 = 0, cb1 = 0, qa0 = 0, qb0 = 0;
        if (r0) {
            int32_t ha, ca, hb, cb;
            row(Ab0, Bb0, L.ry - 1, ha, ca, hb, cb);
            pa = 5 * ha + ca;
            pb = 5 * hb + cb;
        }
        if (r1) {
            int32_t ha0, ca0, hb0, cb0;
            row(Ab1, Bb1, L.ry - 1, ha0, ca0, hb0, cb0);
            row(Ab1, Bb1, L.ry, ha1, ca1, hb1, cb1);
            qa0 = 3 * ha0 + ca0;
            qb0 = 3 * hb0 + cb0;
        }
#pragma unroll
        for (int j = 0; j < LR_SGR_ROWS / 4; j++) {
            const int y = L.ry + j;
            if (y >= L.end) break;
            const int32_t uu = Oq[y * LR_TW + x];
            int32_t f0 = 0, f1 = 0;
            if (r0) {
                if ((j & 1) == 0) {
                    int32_t ha, ca, hb, cb;
                    row(Ab0, Bb0, y + 1, ha, ca, hb, cb);
                    const int32_t na = 5 * ha + ca, nb = 5 * hb + cb;
                    f0 = r2s((pa + na) * uu + (pb + nb), 9);
                    pa = na;
                    pb = nb;
                } else {
                    f0 = r2s(pa * uu + pb, 8);
                }
            }
            if (r1) {
                int32_t ha2, ca2, hb2, cb2;
                row(Ab1, Bb1, y + 1, ha2, ca2, hb2, cb2);
                const int32_t qa2 = 3 * ha2 + ca2, qb2 = 3 * hb2 + cb2;
                f1 = r2s((4 * ha1 + qa0 + qa2) * uu + (4 * hb1 + qb0 + qb2), 9);
                qa0 = 3 * ha1 + ca1;
                qb0 = 3 * hb1 + cb1;
                ha1 = ha2;
                ca1 = ca2;
                hb1 = hb2;
                cb1 = cb2;
            }
            sink(ya + y, x, sgr_project(uu, f0, f1, xq0, xq1, mx));
        }
    }
}

}  // namespace pf
