// Encoder loop-restoration search on gfx950: the curated search of goav1's
// keyframe encoder, luma, 8-bit (encoder/restoration.go:149-262,
// encoder/keyframe_lossy.go:2310-2386):
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

__global__ __launch_bounds__(256) void lr_search_kernel(PlaneView S, PlaneView R, int us, int hz, int vt,
                                                        const uint8_t *mask, av1pf_lr_unit *out, uint64_t *sse_out)
{
    __shared__ uint16_t tile[LR_TW * LR_TW];
    __shared__ uint8_t srcT[64 * 64];
    __shared__ __align__(16) char work[LR_AB * LR_AB * 8];   // Wiener intermediate, then SGR A/B
    __shared__ uint64_t red[4];
    int32_t *Bb = reinterpret_cast<int32_t *>(work);
    int32_t *Ab = reinterpret_cast<int32_t *>(work + LR_AB * LR_AB * 4);
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

    // ---- baseline and the H-only Wiener approximations (any unit size)
    uint64_t b = 0, a[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = tid; i < w * h; i += 256) {
        const int y = y0 + i / w, x = x0 + i % w;
        const uint8_t *rr = row_ptr<uint8_t>(R, y);
        const int sv = row_ptr<uint8_t>(S, y)[x];
        int px[7];
#pragma unroll
        for (int k = 0; k < 7; k++) px[k] = rr[min(max(x + k - 3, 0), W - 1)];
        const int d0 = sv - px[3];
        b += (uint64_t)(d0 * d0);
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const int t0 = kWienerCand[c][0], t1 = kWienerCand[c][1], t2 = kWienerCand[c][2];
            const int hs = t0 * (px[0] + px[6]) + t1 * (px[1] + px[5]) + t2 * (px[2] + px[4]) - 2 * (t0 + t1 + t2) * px[3];
            const int pred = min(max((hs + 64) >> 7, 0), 255);
            const int d = sv - pred;
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
            if (e[1 + c] + (uint64_t)(w * h) < best) { best = e[1 + c]; wbest = c; }
        }
        if (w <= 64 && h <= 64) {
            // ---- the unit with its 3-sample border, edge-clamped at the frame
            const int tw = w + 6, th = h + 6;
            for (int i = tid; i < tw * th; i += 256) {
                const int r = i / tw, c = i - r * tw;
                const int y = min(max(y0 + r - 3, 0), H - 1), x = min(max(x0 + c - 3, 0), W - 1);
                tile[r * LR_TW + c] = row_ptr<uint8_t>(R, y)[x];
            }
            for (int i = tid; i < w * h; i += 256) {
                const int r = i / w, c = i - r * w;
                srcT[r * 64 + c] = row_ptr<uint8_t>(S, y0 + r)[x0 + c];
            }
            __syncthreads();
            const uint16_t *O = tile + 3 * LR_TW + 3;
            uint64_t acc = 0;
            auto err = [&](int y, int x, int v) {
                const int d = (int)srcT[y * 64 + x] - v;
                acc += (uint64_t)(d * d);
            };
            uint64_t best_sse = ~0ull;
            if (wbest >= 0) {
                wiener_tile(O, w, h, 8, kWienerCand[wbest], kWienerCand[wbest], tmp, err);
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
                sgr_tile(O, w, h, 8, kSgrPreset[p], kSgrXqd[p], Ab, Bb, err);
                e[8 + p] = wg_sum(acc, red);
                if (e[8 + p] + (uint64_t)(w * h) < sbest) { sbest = e[8 + p]; pbest = p; }
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

hipError_t launch_lr_search(const PlaneView &src, const PlaneView &rec, int us, int hz, int vt, const uint8_t *mask,
                            av1pf_lr_unit *out, uint64_t *sse, hipStream_t s)
{
    hipLaunchKernelGGL(lr_search_kernel, dim3(hz * vt), dim3(256), 0, s, src, rec, us, hz, vt, mask, out, sse);
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
