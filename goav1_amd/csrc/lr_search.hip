// Encoder loop-restoration search on gfx950: the curated search of goav1's
// keyframe encoder, luma, 8-bit (encoder/restoration.go:149-262,
// encoder/keyframe_lossy.go:2310-2386), and the same search on 10-bit frames
// (an extension with no reference counterpart: estimates clipped to the
// 10-bit range, the decoder filters at 10 bits, the acceptance margin w*h
// scaled by 2^(2 * (bd - 8))):
//   - units of the luma restoration grid (8-row offset, the last unit absorbs
//     the remainder, tile/restoration_geometry.go:54-93);
//   - baseline SSE; six Wiener tap triples scored with an H-only 7-tap sum
//     over frame-clamped columns, (sum + 64) >> 7 clipped to 0..255 (the
//     reference's approximation carries no identity tap); a candidate wins
//     when sse + w*h < the running best;
//   - the Wiener winner re-scored with the exact decoder Wiener (V = H taps)
//     on the recon bordered by 3 edge-clamped samples around the frame;
//   - four SGR presets scored with the exact decoder SGR on the same bordered
//     recon, same threshold; SGR is chosen when it beats the exact Wiener SSE;
//   - units wider or taller than 64 stay NONE (the reference's apply rejects
//     them).
// One 256-thread workgroup per unit: the unit and its 3-sample border in LDS,
// the filters from lr_filters.h (the decoder's own), squared errors reduced
// across the workgroup.  Integer-VALU bound: ~4 SGR + 1 Wiener evaluations
// per pixel.
#include "pf_internal.h"
#include "lr_filters.h"

namespace pf {

__device__ __constant__ int8_t kWienerCand[6][3] = { { 0, -3, 12 }, { 0, -5, 18 }, { 0, -2, 8 },
                                                     { 0, 2, 10 },  { 1, 4, 14 },  { 0, -1, 4 } };
__device__ __constant__ int8_t kSgrPreset[4] = { 0, 2, 10, 14 };
__device__ __constant__ int8_t kSgrXqd[4][2] = { { 8, 24 }, { 16, 40 }, { 0, 48 }, { 24, 72 } };

// Sum of one value per thread over the 256-thread workgroup.
__device__ __forceinline__ uint64_t wg_sum(uint64_t v, uint64_t *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// N sums over the workgroup at once: the waves' shuffle reductions of the N
// values overlap, and one pair of barriers serves all of them (red: 4 N words)
template <int N>
__device__ __forceinline__ void wg_sum_n(int64_t (&v)[N], uint64_t *red)
{
    uint64_t u[N];
#pragma unroll
    for (int k = 0; k < N; k++) u[k] = (uint64_t)v[k];
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < N; k++) u[k] += __shfl_xor(u[k], o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < N; k++) red[(threadIdx.x >> 6) * N + k] = u[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = (int64_t)(red[k] + red[N + k] + red[2 * N + k] + red[3 * N + k]);
}

template <typename T>
__global__ __launch_bounds__(256) void lr_search_kernel(PlaneView S, PlaneView R, int bd, int us, int hz, int vt,
                                                        const uint8_t *mask, av1pf_lr_unit *out, uint64_t *sse_out)
{
    __shared__ __align__(16) uint16_t tile[LR_TW * LR_TH];
    __shared__ uint16_t srcT[64 * 64];
    __shared__ __align__(16) char work[LR_AB * LR_AB * 6];   // Wiener intermediate, then SGR A/B
    __shared__ uint64_t red[4];
    int32_t *Bb = reinterpret_cast<int32_t *>(work);
    uint16_t *Ab = reinterpret_cast<uint16_t *>(work + LR_AB * LR_AB * 4);
    uint16_t *tmp = reinterpret_cast<uint16_t *>(work);   // SGR runs after the Wiener re-score
    const int tid = threadIdx.x;
    const int u = blockIdx.x;
    const int col = u % hz, row = u / hz;
    const int W = R.w, H = R.h;
    const int x0 = col * us, x1 = col == hz - 1 ? W : x0 + us;
    int y0 = row * us, y1 = row == vt - 1 ? H : y0 + us;
    if (y0 > 0) y0 -= 8;
    if (y1 < H) y1 -= 8;
    const int w = x1 - x0, h = y1 - y0;
    const int maxv = (1 << bd) - 1;
    const uint64_t margin = (uint64_t)(w * h) << (2 * (bd - 8));   // a candidate must beat the best by this
    uint64_t e[12];
#pragma unroll
    for (int i = 0; i < 12; i++) e[i] = ~0ull;
    if (mask && !mask[u]) {   // unit left out of the search: NONE, nothing scored
        if (tid == 0) {
            av1pf_lr_unit none;
            for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&none)[i] = 0;
            out[u] = none;
            if (sse_out)
                for (int i = 0; i < 12; i++) sse_out[(size_t)u * 12 + i] = e[i];
        }
        return;
    }

    // a unit of at most 64 x 64 (the exact filters' limit) is staged once,
    // with its 3-sample border edge-clamped at the frame, and the source
    // beside it, every load in flight before the first LDS store; the
    // baseline and the estimates then read LDS (the border columns are the
    // frame-clamped columns the estimate reads)
    const bool small = w <= 64 && h <= 64;
    if (small) {
        const int tw = w + 6, th = h + 6;
        constexpr int KT = (LR_TH * 70 + 255) / 256, KS = 64 * 64 / 256;
        uint16_t v[KT], sv[KS];
#pragma unroll
        for (int k = 0; k < KT; k++) {
            const int i = tid + k * 256, r = i / tw, c = i - r * tw;
            const int y = min(max(y0 + r - 3, 0), H - 1), x = min(max(x0 + c - 3, 0), W - 1);
            v[k] = i < tw * th ? row_ptr<T>(R, y)[x] : 0;
        }
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int i = tid + k * 256, r = i >> 6, c = i & 63;
            sv[k] = r < h && c < w ? row_ptr<T>(S, y0 + r)[x0 + c] : 0;
        }
#pragma unroll
        for (int k = 0; k < KT; k++) {
            const int i = tid + k * 256, r = i / tw, c = i - r * tw;
            if (i < tw * th) tile[r * LR_TW + c + LR_OX - 3] = v[k];
        }
#pragma unroll
        for (int k = 0; k < KS; k++) srcT[tid + k * 256] = sv[k];
        __syncthreads();
    }

    // ---- baseline and the H-only Wiener approximations (any unit size)
    uint64_t b = 0, a[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = tid; i < w * h; i += 256) {
        const int yy = i / w, xx = i % w;
        int svv, px[7];
        if (small) {
            const uint16_t *rr = tile + (yy + 3) * LR_TW + LR_OX + xx;
            svv = srcT[yy * 64 + xx];
#pragma unroll
            for (int k = 0; k < 7; k++) px[k] = rr[k - 3];
        } else {
            const T *rr = row_ptr<T>(R, y0 + yy);
            svv = row_ptr<T>(S, y0 + yy)[x0 + xx];
#pragma unroll
            for (int k = 0; k < 7; k++) px[k] = rr[min(max(x0 + xx + k - 3, 0), W - 1)];
        }
        const int d0 = svv - px[3];
        b += (uint64_t)(d0 * d0);
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const int t0 = kWienerCand[c][0], t1 = kWienerCand[c][1], t2 = kWienerCand[c][2];
            const int hs = t0 * (px[0] + px[6]) + t1 * (px[1] + px[5]) + t2 * (px[2] + px[4]) - 2 * (t0 + t1 + t2) * px[3];
            const int pred = min(max((hs + 64) >> 7, 0), maxv);
            const int d = svv - pred;
            a[c] += (uint64_t)(d * d);
        }
    }
    const uint64_t base = wg_sum(b, red);
    e[0] = base;
    av1pf_lr_unit res;
    for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&res)[i] = 0;
    res.type = AV1PF_LR_NONE;
    if (base != 0) {
        uint64_t best = base;
        int wbest = -1;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            e[1 + c] = wg_sum(a[c], red);
            if (e[1 + c] + margin < best) { best = e[1 + c]; wbest = c; }
        }
        if (small) {
            const uint16_t *O = tile + 3 * LR_TW + LR_OX;
            uint64_t acc = 0;
            auto err = [&](int y, int x, int v) {
                const int d = (int)srcT[y * 64 + x] - v;
                acc += (uint64_t)(d * d);
            };
            auto err4 = [&](int y, int x, const int *v, int n) {
                for (int k = 0; k < n; k++) err(y, x + k, v[k]);
            };
            uint64_t best_sse = ~0ull;
            if (wbest >= 0) {
                wiener_tile(O, w, h, bd, kWienerCand[wbest], kWienerCand[wbest], tmp, err4);
                e[7] = best_sse = wg_sum(acc, red);
                res.type = AV1PF_LR_WIENER;
                for (int k = 0; k < 3; k++) res.wiener_v[k] = res.wiener_h[k] = kWienerCand[wbest][k];
            }
            uint64_t sbest = base;
            int pbest = -1;
#pragma unroll
            for (int p = 0; p < 4; p++) {
                acc = 0;
                __syncthreads();
                sgr_tile(O, w, h, bd, kSgrPreset[p], kSgrXqd[p], Ab, Bb, err);
                e[8 + p] = wg_sum(acc, red);
                if (e[8 + p] + margin < sbest) { sbest = e[8 + p]; pbest = p; }
            }
            if (pbest >= 0 && sbest < best_sse) {
                for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&res)[i] = 0;
                res.type = AV1PF_LR_SGRPROJ;
                res.sgr_set = (uint8_t)kSgrPreset[pbest];
                res.sgr_xqd[0] = kSgrXqd[pbest][0];
                res.sgr_xqd[1] = kSgrXqd[pbest][1];
            }
        }
    }
    if (tid == 0) {
        out[u] = res;
        if (sse_out)
            for (int i = 0; i < 12; i++) sse_out[(size_t)u * 12 + i] = e[i];
    }
}

// ---- exhaustive superset mode ------------------------------------------------
// No counterpart in the reference (its search is the curated one above).  Per
// unit of at most 64x64: the baseline; the six curated Wiener triples scored
// exactly; a least-squares Wiener fit (symmetric Q7 taps summing to 128,
// alternating: horizontal taps given the vertical ones, then vertical given
// horizontal, twice from the identity; each 3x3 system from int64 statistics
// of the 7x7 windows, solved by Cramer's rule in double, rounded
// floor(x + 0.5) and clamped to [-5,10], [-23,8], [-17,46]); the four curated
// SGR presets; and all 16 SGR sets with a least-squares projection (the xq
// minimising sum (2^7 (16 s - 16 u) - xq0 (f0 - 16 u) - xq1 (f1 - 16 u))^2,
// coded as xqd in [-96,31] x [-32,95] as libaom codes it), every candidate
// scored exactly with the decoder's filters.  Candidates in that order replace
// the running best when sse + w*h < best, so the decision is never worse than
// the curated one.  sse[unit][28]; fit[unit][38] = Wiener v[3], h[3], then the
// fitted xqd[2] of each set.
__device__ __constant__ int8_t kTapMin[3] = { -5, -23, -17 };
__device__ __constant__ int8_t kTapMax[3] = { 10, 8, 46 };

__device__ __forceinline__ int round_clamp(double x, int lo, int hi)
{
    const int v = (int)floor(x + 0.5);
    return min(max(v, lo), hi);
}

// 3x3 symmetric system (m00 m01 m02 m11 m12 m22) by Cramer's rule; the same
// operations in the same order as the CPU restatement, without contraction.
__device__ void solve_taps(const int64_t *M, const int64_t *R, int8_t *out)
{
#pragma clang fp contract(off)
    const double a = (double)M[0], b = (double)M[1], c = (double)M[2], d = (double)M[3], e = (double)M[4],
                 f = (double)M[5];
    const double r0 = (double)R[0], r1 = (double)R[1], r2 = (double)R[2];
    const double c00 = d * f - e * e, c01 = b * f - e * c, c02 = b * e - d * c;
    const double det = a * c00 - b * c01 + c * c02;
    if (det == 0.0) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    const double x0 = (r0 * c00 - b * (r1 * f - e * r2) + c * (r1 * e - d * r2)) / det;
    const double x1 = (a * (r1 * f - e * r2) - r0 * c01 + c * (b * r2 - r1 * c)) / det;
    const double x2 = (a * (d * r2 - r1 * e) - b * (b * r2 - r1 * c) + r0 * c02) / det;
    out[0] = (int8_t)round_clamp(x0, kTapMin[0], kTapMax[0]);
    out[1] = (int8_t)round_clamp(x1, kTapMin[1], kTapMax[1]);
    out[2] = (int8_t)round_clamp(x2, kTapMin[2], kTapMax[2]);
}

__device__ void sgr_fit(int set, const int64_t *st, int8_t *xqd)
{
#pragma clang fp contract(off)
    const int r0 = kSgrR[set][0], r1 = kSgrR[set][1];
    double xq0 = 0.0, xq1 = 0.0;
    if (r0 && r1) {
        const double a = (double)st[0], b = (double)st[1], d = (double)st[2], c0 = (double)st[3], c1 = (double)st[4];
        const double det = a * d - b * b;
        if (det != 0.0) {
            xq0 = (c0 * d - c1 * b) / det;
            xq1 = (c1 * a - c0 * b) / det;
        }
    } else if (r0) {
        if (st[0]) xq0 = (double)st[3] / (double)st[0];
    } else {
        if (st[2]) xq1 = (double)st[4] / (double)st[2];
    }
    if (r0 && r1) {
        xqd[0] = (int8_t)round_clamp(xq0, -96, 31);
        xqd[1] = (int8_t)round_clamp(128.0 - (double)xqd[0] - floor(xq1 + 0.5), -32, 95);
    } else if (r0) {
        xqd[0] = (int8_t)round_clamp(xq0, -96, 31);
        xqd[1] = (int8_t)min(max(128 - xqd[0], -32), 95);
    } else {
        xqd[0] = 0;
        xqd[1] = (int8_t)round_clamp(128.0 - floor(xq1 + 0.5), -32, 95);
    }
}

// The exhaustive mode in LR_EXH_PARTS workgroups per unit, each scoring its
// own candidates (the scores and fits do not depend on each other): part 0
// the Wiener least-squares fit and its score, part 1 the six curated Wiener
// triples, part 2 + s SGR set s (its preset, when it has one, then its fit).
// lr_exh_decide_kernel then takes the decision from the scores in the
// sequential order (lr_search_decide).
constexpr int LR_EXH_PARTS = 18;

__global__ __launch_bounds__(256) void lr_search_exh_kernel(PlaneView S, PlaneView R, int us, int hz, int vt,
                                                            uint64_t *sse_out, int8_t *fit_out)
{
    __shared__ __align__(16) uint16_t tile[LR_TW * LR_TH];
    __shared__ uint8_t srcT[64 * 64];
    __shared__ __align__(16) char work[LR_AB * LR_AB * 6];   // Wiener intermediate, then SGR A/B
    __shared__ uint64_t red[4 * 9];
    __shared__ int8_t taps_s[2][3];
    __shared__ int8_t xqd_s[2];
    int32_t *Bb = reinterpret_cast<int32_t *>(work);
    uint16_t *Ab = reinterpret_cast<uint16_t *>(work + LR_AB * LR_AB * 4);
    uint16_t *tmp = reinterpret_cast<uint16_t *>(work);
    const int tid = threadIdx.x;
    const int u = blockIdx.x, part = blockIdx.y;
    const int col = u % hz, row = u / hz;
    const int W = R.w, H = R.h;
    const int x0 = col * us, x1 = col == hz - 1 ? W : x0 + us;
    int y0 = row * us, y1 = row == vt - 1 ? H : y0 + us;
    if (y0 > 0) y0 -= 8;
    if (y1 < H) y1 -= 8;
    const int w = x1 - x0, h = y1 - y0;
    uint64_t *e = sse_out + (size_t)u * 28;
    int8_t *ft = fit_out + (size_t)u * 38;
    // each part's own slots: no score (~0) and no fit (0) until it finds them
    if (tid == 0) {
        if (part == 0) {
            e[7] = ~0ull;
            for (int k = 0; k < 6; k++) ft[k] = 0;
        } else if (part == 1) {
            for (int c = 0; c < 6; c++) e[1 + c] = ~0ull;
        } else {
            const int set = part - 2;
            e[12 + set] = ~0ull;
            for (int q = 0; q < 4; q++)
                if (kSgrPreset[q] == set) e[8 + q] = ~0ull;
            ft[6 + 2 * set] = ft[7 + 2 * set] = 0;
        }
    }
    uint64_t b = 0;
    for (int i = tid; i < w * h; i += 256) {
        const int y = y0 + i / w, x = x0 + i % w;
        const int d = (int)row_ptr<uint8_t>(S, y)[x] - (int)row_ptr<uint8_t>(R, y)[x];
        b += (uint64_t)(d * d);
    }
    const uint64_t base = wg_sum(b, red);
    if (part == 0 && tid == 0) e[0] = base;
    if (base != 0 && w <= 64 && h <= 64) {
        const int tw = w + 6, th = h + 6;
        for (int i = tid; i < tw * th; i += 256) {
            const int r = i / tw, c = i - r * tw;
            const int y = min(max(y0 + r - 3, 0), H - 1), x = min(max(x0 + c - 3, 0), W - 1);
            tile[r * LR_TW + c + LR_OX - 3] = row_ptr<uint8_t>(R, y)[x];
        }
        for (int i = tid; i < w * h; i += 256) {
            const int r = i / w, c = i - r * w;
            srcT[r * 64 + c] = row_ptr<uint8_t>(S, y0 + r)[x0 + c];
        }
        __syncthreads();
        const uint16_t *O = tile + 3 * LR_TW + LR_OX;
        const uint64_t wh = (uint64_t)(w * h);
        const LaneRows L = lane_rows(w, h);
        // ---- Wiener fit: alternating least squares over this lane's pixels
        int8_t fv[3] = { 0, 0, 0 }, fh[3] = { 0, 0, 0 };
        for (int half = 0; part == 0 && half < 4; half++) {
            const bool horiz = (half & 1) == 0;   // solve H given V, then V given H
            const int8_t *fx = horiz ? fv : fh;
            const int ftp[7] = { fx[0], fx[1], fx[2], 128 - 2 * (fx[0] + fx[1] + fx[2]), fx[2], fx[1], fx[0] };
            int64_t st[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
            if (L.x < w) {
                for (int j = 0; j < SGR_PPT; j++) {
                    const int y = L.ry + j, x = L.x;
                    if (y >= L.end) break;
                    // 32-bit sums are exact: 8-bit samples and taps within
                    // the fit's clamps (outer taps -5..10, -23..8, -17..46,
                    // centre 0..218: 376 in absolute sum), so |Y| < 2^17,
                    // |z| < 2^19 and |t| < 2^24; the products need 64 bits
                    int32_t Y[7];
#pragma unroll
                    for (int k = 0; k < 7; k++) {
                        int32_t acc = 0;
#pragma unroll
                        for (int q = 0; q < 7; q++) {
                            const int yy = horiz ? y + q - 3 : y + k - 3, xx = horiz ? x + k - 3 : x + q - 3;
                            acc += ftp[q] * (int32_t)O[yy * LR_TW + xx];
                        }
                        Y[k] = acc;
                    }
                    const int32_t z0 = Y[0] + Y[6] - 2 * Y[3], z1 = Y[1] + Y[5] - 2 * Y[3], z2 = Y[2] + Y[4] - 2 * Y[3];
                    const int32_t t = ((int32_t)srcT[y * 64 + x] << 14) - 128 * Y[3];
                    st[0] += (int64_t)z0 * z0; st[1] += (int64_t)z0 * z1; st[2] += (int64_t)z0 * z2;
                    st[3] += (int64_t)z1 * z1; st[4] += (int64_t)z1 * z2; st[5] += (int64_t)z2 * z2;
                    st[6] += (int64_t)z0 * t; st[7] += (int64_t)z1 * t; st[8] += (int64_t)z2 * t;
                }
            }
            int64_t tot[9];
#pragma unroll
            for (int k = 0; k < 9; k++) tot[k] = st[k];
            wg_sum_n(tot, red);
            if (tid == 0) solve_taps(tot, tot + 6, taps_s[horiz ? 1 : 0]);
            __syncthreads();
            int8_t *dstt = horiz ? fh : fv;
            for (int k = 0; k < 3; k++) dstt[k] = taps_s[horiz ? 1 : 0][k];
            __syncthreads();
        }
        if (part == 0 && tid == 0)
            for (int k = 0; k < 3; k++) { ft[k] = fv[k]; ft[3 + k] = fh[k]; }
        // ---- Wiener: six curated triples, then the fit, all exact
        uint64_t acc = 0;
        auto err = [&](int y, int x, int v) {
            const int d = (int)srcT[y * 64 + x] - v;
            acc += (uint64_t)(d * d);
        };
        auto err4 = [&](int y, int x, const int *v, int n) {
            for (int k = 0; k < n; k++) err(y, x + k, v[k]);
        };
        for (int c = part == 0 ? 6 : 0; part <= 1 && c < (part == 0 ? 7 : 6); c++) {
            const int8_t *vt = c < 6 ? kWienerCand[c] : fv;
            const int8_t *ht = c < 6 ? kWienerCand[c] : fh;
            acc = 0;
            __syncthreads();
            wiener_tile(O, w, h, 8, vt, ht, tmp, err4);
            const uint64_t sc = wg_sum(acc, red);
            if (tid == 0) e[1 + c] = sc;
        }
        // ---- SGR: per set the curated preset (when it has one), then the fit
        for (int set = part - 2; set >= 0 && set == part - 2; set++) {
            int32_t f0[SGR_PPT], f1[SGR_PPT];
            __syncthreads();
            sgr_filters(O, w, h, 8, set, Ab, Bb, f0, f1);
            const int r0 = kSgrR[set][0], r1 = kSgrR[set][1];
            int64_t st[5] = { 0, 0, 0, 0, 0 };
            if (L.x < w) {
#pragma unroll
                for (int j = 0; j < SGR_PPT; j++) {
                    const int y = L.ry + j, x = L.x;
                    if (y >= L.end) break;
                    const int32_t uu = (int32_t)O[y * LR_TW + x] << 4;
                    const int64_t d0 = r0 ? f0[j] - uu : 0, d1 = r1 ? f1[j] - uu : 0;
                    const int64_t ee = (((int64_t)srcT[y * 64 + x] << 4) - uu) * 128;
                    st[0] += d0 * d0; st[1] += d0 * d1; st[2] += d1 * d1; st[3] += d0 * ee; st[4] += d1 * ee;
                }
            }
            int64_t tot[5];
#pragma unroll
            for (int k = 0; k < 5; k++) tot[k] = st[k];
            wg_sum_n(tot, red);
            if (tid == 0) sgr_fit(set, tot, xqd_s);
            __syncthreads();
            int8_t fx[2] = { xqd_s[0], xqd_s[1] };
            if (tid == 0) { ft[6 + 2 * set] = fx[0]; ft[7 + 2 * set] = fx[1]; }
            int p = -1;
            for (int q = 0; q < 4; q++)
                if (kSgrPreset[q] == set) p = q;
            for (int k = p >= 0 ? 0 : 1; k < 2; k++) {
                const int8_t *xq = k == 0 ? kSgrXqd[p] : fx;
                int xq0, xq1;
                sgr_xq(set, xq, xq0, xq1);
                uint64_t a2 = 0;
                if (L.x < w) {
#pragma unroll
                    for (int j = 0; j < SGR_PPT; j++) {
                        const int y = L.ry + j, x = L.x;
                        if (y >= L.end) break;
                        const int d = (int)srcT[y * 64 + x] - sgr_project(O[y * LR_TW + x], f0[j], f1[j], xq0, xq1, 255);
                        a2 += (uint64_t)(d * d);
                    }
                }
                const uint64_t sc = wg_sum(a2, red);
                if (tid == 0) e[k == 0 ? 8 + p : 12 + set] = sc;
            }
        }
    }
}

// The decision of the exhaustive mode from its 28 scores, in the sequential
// order of the single-pass search: base, the six curated Wiener triples, the
// Wiener fit, then per SGR set its preset (when it has one) and its fit; a
// candidate replaces the running best when its SSE plus w * h is below it.
__global__ void lr_exh_decide_kernel(int W, int H, int us, int hz, int vt, const uint64_t *sse, const int8_t *fit,
                                     av1pf_lr_unit *out)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= hz * vt) return;
    const int col = u % hz, row = u / hz;
    const int x0 = col * us, x1 = col == hz - 1 ? W : x0 + us;
    int y0 = row * us, y1 = row == vt - 1 ? H : y0 + us;
    if (y0 > 0) y0 -= 8;
    if (y1 < H) y1 -= 8;
    const int w = x1 - x0, h = y1 - y0;
    const uint64_t *e = sse + (size_t)u * 28;
    const int8_t *ft = fit + (size_t)u * 38;
    av1pf_lr_unit res;
    for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&res)[i] = 0;
    res.type = AV1PF_LR_NONE;
    const uint64_t base = e[0];
    if (base != 0 && w <= 64 && h <= 64) {
        const uint64_t wh = (uint64_t)(w * h);
        uint64_t best = base;
        for (int c = 0; c < 7; c++) {
            const uint64_t sc = e[1 + c];
            if (sc + wh < best) {
                best = sc;
                for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&res)[i] = 0;
                res.type = AV1PF_LR_WIENER;
                for (int k = 0; k < 3; k++) {
                    res.wiener_v[k] = c < 6 ? kWienerCand[c][k] : ft[k];
                    res.wiener_h[k] = c < 6 ? kWienerCand[c][k] : ft[3 + k];
                }
            }
        }
        for (int set = 0; set < 16; set++) {
            int p = -1;
            for (int q = 0; q < 4; q++)
                if (kSgrPreset[q] == set) p = q;
            for (int k = p >= 0 ? 0 : 1; k < 2; k++) {
                const uint64_t sc = e[k == 0 ? 8 + p : 12 + set];
                if (sc + wh < best) {
                    best = sc;
                    for (int i = 0; i < 16; i++) reinterpret_cast<uint8_t *>(&res)[i] = 0;
                    res.type = AV1PF_LR_SGRPROJ;
                    res.sgr_set = (uint8_t)set;
                    res.sgr_xqd[0] = k == 0 ? kSgrXqd[p][0] : ft[6 + 2 * set];
                    res.sgr_xqd[1] = k == 0 ? kSgrXqd[p][1] : ft[7 + 2 * set];
                }
            }
        }
    }
    out[u] = res;
}

hipError_t launch_lr_search_exhaustive(const PlaneView &src, const PlaneView &rec, int us, int hz, int vt,
                                       av1pf_lr_unit *out, uint64_t *sse, int8_t *fit, hipStream_t s)
{
    hipLaunchKernelGGL(lr_search_exh_kernel, dim3(hz * vt, LR_EXH_PARTS), dim3(256), 0, s, src, rec, us, hz, vt, sse,
                       fit);
    hipLaunchKernelGGL(lr_exh_decide_kernel, dim3((hz * vt + 63) / 64), dim3(64), 0, s, rec.w, rec.h, us, hz, vt, sse,
                       fit, out);
    return hipGetLastError();
}

hipError_t launch_lr_search(const PlaneView &src, const PlaneView &rec, int bd, int us, int hz, int vt,
                            const uint8_t *mask, av1pf_lr_unit *out, uint64_t *sse, hipStream_t s)
{
    if (bd == 8)
        hipLaunchKernelGGL(lr_search_kernel<uint8_t>, dim3(hz * vt), dim3(256), 0, s, src, rec, bd, us, hz, vt, mask,
                           out, sse);
    else
        hipLaunchKernelGGL(lr_search_kernel<uint16_t>, dim3(hz * vt), dim3(256), 0, s, src, rec, bd, us, hz, vt, mask,
                           out, sse);
    return hipGetLastError();
}

// Units the network expects to stay NONE (P(NONE) above the threshold, from
// the tile of the co-located 64x64 filter block) are left out of the search.
__global__ void lr_mask_kernel(const float *logits, int nhfb, int nvfb, int hz, int vt, float thr, uint8_t *mask)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= hz * vt) return;
    const int fb = min(u / hz, nvfb - 1) * nhfb + min(u % hz, nhfb - 1);
    const float *z = logits + (size_t)fb * 131 + 128;
    const float mx = fmaxf(z[0], fmaxf(z[1], z[2]));
    const float e0 = __expf(z[0] - mx), se = e0 + __expf(z[1] - mx) + __expf(z[2] - mx);
    mask[u] = e0 / se > thr ? 0 : 1;
}

hipError_t launch_lr_mask(const float *logits, int nhfb, int nvfb, int hz, int vt, float thr, uint8_t *mask,
                          hipStream_t s)
{
    hipLaunchKernelGGL(lr_mask_kernel, dim3((hz * vt + 255) / 256), dim3(256), 0, s, logits, nhfb, nvfb, hz, vt, thr, mask);
    return hipGetLastError();
}

}  // namespace pf
