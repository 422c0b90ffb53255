// Loop-restoration filters over an LDS tile, shared by the decoder kernel
// (lr.hip) and the encoder search (lr_search.hip).
//
// The tile holds a w x h block (w, h <= 64) plus a 3-sample halo on every
// side at pitch LR_TW; O points at sample (0, 0).  Each filter hands every
// output sample to sink(y, x, value), so the decoder writes the frame and the
// search accumulates squared errors from the same arithmetic:
//   - Wiener (restoration/wiener.go:132-216): H pass with offset 2^(bd+6),
//     round0, clamp to 2^(bd+1+7-round0)-1; V pass with round1;
//   - self-guided (restoration/selfguided.go): box sums, uint32 p/z/A/B math,
//     6/5 (r=2, every other row) and 4/3 (r=1) weighted sums, projection with
//     an int16 wrap before the clip.
// Block size 256 threads; the caller synchronises before the call.  LDS the
// caller provides: the tile, a 70 x 64 uint16 Wiener intermediate, and the
// self-guided A and B planes (int32, LR_AB x LR_AB); the Wiener intermediate
// may share memory with A and B.
#pragma once
#include "pf_internal.h"

namespace pf {

#define LR_TW 70          // 64 + 6
#define LR_AB 66          // 64 + 2

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

template <typename Sink>
__device__ __forceinline__ void wiener_tile(const uint16_t *O, int w, int h, int bd, const int8_t *vt, const int8_t *ht,
                                            uint16_t *tmp, Sink sink)
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
    const int lpr = lanes_for(w), c = tid & (lpr - 1), rstep = 256 / lpr;
    if (c < w)
        for (int rr = tid / lpr; rr < th; rr += rstep) {
            const int r = rr - 3;
            int32_t t[7];
            lds_run<-3, 3>(O + r * LR_TW + c, t);
            int32_t sum = (t[3] << 7) + hoff;
            sum += h0 * (t[0] + t[6]) + h1 * (t[1] + t[5]) + h2 * (t[2] + t[4]) + h3 * t[3];
            tmp[rr * 64 + c] = (uint16_t)min(max(r2s(sum, round0), 0), limit - 1);
        }
    __syncthreads();
    const int32_t voff = 1 << (bd + round1 - 1);
    if (c < w)
        for (int r = tid / lpr; r < h; r += rstep) {
            const uint16_t *s = tmp + (r + 3) * 64 + c;
            int32_t sum = ((int32_t)s[0] << 7) - voff;
            sum += v0 * ((int32_t)s[-3 * 64] + s[3 * 64]) + v1 * ((int32_t)s[-2 * 64] + s[2 * 64]) +
                   v2 * ((int32_t)s[-64] + s[64]) + v3 * (int32_t)s[0];
            sink(r, c, min(max(r2s(sum, round1), 0), mx));
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
__device__ __forceinline__ void sgr_ab(const uint16_t *O, int w, int h, int bd, uint32_t s, int32_t *Ab, int32_t *Bb)
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
        Ab[(y + 1) * LR_AB + x + 1] = aa;
        Bb[(y + 1) * LR_AB + x + 1] = (int32_t)r2u((uint32_t)(256 - aa) * (uint32_t)sum * recip, 12);
    }
}

// Lane layout of the per-pixel filter stages: lane x of row group ry handles
// column x of rows ry, ry + rstep, ... (at most SGR_PPT rows).
#define SGR_PPT 16
struct LaneRows {
    int x, ry, rstep;
};
__device__ __forceinline__ LaneRows lane_rows(int w)
{
    const int lpr = lanes_for(w);
    return { (int)threadIdx.x & (lpr - 1), (int)threadIdx.x / lpr, 256 / lpr };
}

// The two self-guided outputs of set `set` for this lane's pixels (Q4; zero
// for a pass of radius 0): f0 from the r = 2 pass, f1 from the r = 1 pass.
__device__ __forceinline__ void sgr_filters(const uint16_t *O, int w, int h, int bd, int set, int32_t *Ab, int32_t *Bb,
                                            int32_t (&f0)[SGR_PPT], int32_t (&f1)[SGR_PPT])
{
    const LaneRows L = lane_rows(w);
#pragma unroll
    for (int j = 0; j < SGR_PPT; j++) { f0[j] = 0; f1[j] = 0; }
    for (int pass = 0; pass < 2; pass++) {
        const int r = kSgrR[set][pass];
        if (r == 0) continue;
        const uint32_t s = (uint32_t)kSgrS[set][pass];
        __syncthreads();
        if (pass == 0) sgr_ab<2, 2>(O, w, h, bd, s, Ab, Bb);
        else sgr_ab<1, 1>(O, w, h, bd, s, Ab, Bb);
        __syncthreads();
        if (L.x < w) {
#pragma unroll
            for (int j = 0; j < SGR_PPT; j++) {
                const int y = L.ry + L.rstep * j, x = L.x;
                if (y >= h) break;
                const int32_t uu = O[y * LR_TW + x];
                const int32_t *A = Ab + (y + 1) * LR_AB + x + 1;
                const int32_t *B = Bb + (y + 1) * LR_AB + x + 1;
                int32_t a, b;
                if (pass == 0) {
                    if ((y & 1) == 0) {
                        a = (A[-LR_AB] + A[LR_AB]) * 6 + (A[-LR_AB - 1] + A[-LR_AB + 1] + A[LR_AB - 1] + A[LR_AB + 1]) * 5;
                        b = (B[-LR_AB] + B[LR_AB]) * 6 + (B[-LR_AB - 1] + B[-LR_AB + 1] + B[LR_AB - 1] + B[LR_AB + 1]) * 5;
                        f0[j] = r2s(a * uu + b, 9);
                    } else {
                        a = A[0] * 6 + (A[-1] + A[1]) * 5;
                        b = B[0] * 6 + (B[-1] + B[1]) * 5;
                        f0[j] = r2s(a * uu + b, 8);
                    }
                } else {
                    a = (A[0] + A[-1] + A[1] + A[-LR_AB] + A[LR_AB]) * 4 +
                        (A[-LR_AB - 1] + A[LR_AB - 1] + A[-LR_AB + 1] + A[LR_AB + 1]) * 3;
                    b = (B[0] + B[-1] + B[1] + B[-LR_AB] + B[LR_AB]) * 4 +
                        (B[-LR_AB - 1] + B[LR_AB - 1] + B[-LR_AB + 1] + B[LR_AB + 1]) * 3;
                    f1[j] = r2s(a * uu + b, 9);
                }
            }
        }
    }
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

template <typename Sink>
__device__ __forceinline__ void sgr_tile(const uint16_t *O, int w, int h, int bd, int set, const int8_t *xqd,
                                         int32_t *Ab, int32_t *Bb, Sink sink)
{
    int xq0, xq1;
    sgr_xq(set, xqd, xq0, xq1);
    int32_t f0[SGR_PPT], f1[SGR_PPT];
    sgr_filters(O, w, h, bd, set, Ab, Bb, f0, f1);
    const LaneRows L = lane_rows(w);
    const int32_t mx = (1 << bd) - 1;
    if (L.x < w) {
#pragma unroll
        for (int j = 0; j < SGR_PPT; j++) {
            const int y = L.ry + L.rstep * j;
            if (y >= h) break;
            sink(y, L.x, sgr_project(O[y * LR_TW + L.x], f0[j], f1[j], xq0, xq1, mx));
        }
    }
}


}  // namespace pf
