// AV1 film-grain synthesis on gfx950.
//
// Semantics follow goav1 (filmgrain/*.go, decoder/postfilter_filmgrain.go):
//   - 16-bit LFSR (random.go:51-62), plane seeds (^0xb524 / ^0x49d8) and per
//     32-line stripe seeds (random.go:26-48);
//   - 82x73 luma and 44x38 / 82x73 chroma templates from the normative Gaussian
//     table, rounded by 12 - bd + grain_scale_shift, then the auto-regressive
//     filter of lag 0..3 with the luma-average term for chroma
//     (luma.go:24-76, chroma.go:29-90,223-241);
//   - per 32x32 block random offsets, 2-sample overlap blends (27/17, 23/22),
//     piecewise-linear scaling (interpolated for 10/12-bit), chroma scaling index
//     from the pre-grain luma, clip to full or restricted range
//     (apply.go:72-310,480-549, blend.go, sample.go, scaling.go:55-77).
//
// Template kernel: one workgroup.  Each lane fills one template row of
// Gaussian samples after jumping the LFSR to the row start with a GF(2) matrix
// power (the LFSR step is linear over GF(2)); the AR recursion, sequential in
// the reference, runs as a diagonal wavefront: cell (y, x) only needs cells of
// earlier rows within 3 columns and the 3 cells to its left, so every cell with
// x + 4y = t is independent and the recursion takes ~350 barrier steps.
// Apply kernel: one workgroup per 32x32 luma block and its chroma blocks (the
// offsets of block (row, col) are shared by all planes); scaling LUTs in LDS.
#include "pf_internal.h"
#include "av1pf_gaussian_table.h"

namespace pf {

#define GW 82
#define GH 73

__device__ __constant__ int16_t kGauss[2048] = { AV1PF_GAUSSIAN_VALUES };

__device__ __forceinline__ int r2i(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }
__device__ __forceinline__ int clipi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ uint16_t lfsr_step(uint16_t r)
{
    const uint16_t bit = (uint16_t)((r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1);
    return (uint16_t)((r >> 1) | (bit << 15));
}

__device__ __forceinline__ uint16_t gf2_apply(const uint16_t *J, uint16_t s)
{
    uint16_t o = 0;
#pragma unroll
    for (int b = 0; b < 16; b++)
        if ((s >> b) & 1) o ^= J[b];
    return o;
}

// AR value of one cell (luma.go:60-74, chroma.go:69-87); g is the template in LDS.
__device__ __forceinline__ void ar_cell(int16_t *g, int y, int x, int lag, const int8_t *coef, int shift,
                                        int gmin, int gmax, bool chroma, bool luma_term, const int16_t *lg,
                                        int ssx, int ssy)
{
    int sum = 0, pos = 0;
    for (int dr = -lag; dr <= 0; dr++) {
        for (int dc = -lag; dc <= lag; dc++) {
            if (dr == 0 && dc == 0) {
                if (chroma && luma_term) {
                    const int lx = ((x - 3) << ssx) + 3, ly = ((y - 3) << ssy) + 3;
                    int s = 0;
                    for (int dy = 0; dy <= ssy; dy++)
                        for (int dx = 0; dx <= ssx; dx++) s += lg[(ly + dy) * GW + lx + dx];
                    sum += r2i(s, ssx + ssy) * coef[pos];
                }
                goto done;
            }
            sum += coef[pos] * g[(y + dr) * GW + x + dc];
            pos++;
        }
    }
done:
    g[y * GW + x] = (int16_t)clipi(g[y * GW + x] + ((sum + (1 << (shift - 1))) >> shift), gmin, gmax);
}

__global__ __launch_bounds__(256) void grain_template_kernel(GrainArgs a, int bd, int ssx, int ssy)
{
    __shared__ int16_t T[3][GH * GW];
    const int tid = threadIdx.x;
    const av1pf_grain_params &g = a.g;
    const int shift = 12 - bd + g.grain_scale_shift;
    const int cw = ssx ? 44 : GW, ch = ssy ? 38 : GH;
    for (int i = tid; i < 3 * GH * GW; i += 256) (&T[0][0])[i] = 0;
    __syncthreads();
    // ---- Gaussian fill: lane -> (plane, row)
    {
        int pl = -1, row = 0;
        if (tid < GH) { pl = 0; row = tid; }
        else if (tid < GH + ch) { pl = 1; row = tid - GH; }
        else if (tid < GH + 2 * ch) { pl = 2; row = tid - GH - ch; }
        bool on = pl == 0 ? g.num_y_points != 0 : pl > 0;
        if (on) {
            uint16_t s = pl == 0 ? g.seed : (uint16_t)(g.seed ^ (pl == 1 ? 0xb524 : 0x49d8));
            const uint16_t *J = pl == 0 ? a.jump_luma : a.jump_chroma;
            for (int k = 0; k < row; k++) s = gf2_apply(J, s);
            const int w = pl == 0 ? GW : cw;
            for (int x = 0; x < w; x++) {
                s = lfsr_step(s);
                T[pl][row * GW + x] = (int16_t)r2i(kGauss[s >> 5], shift);
            }
        }
    }
    __syncthreads();
    const int lag = g.ar_coeff_lag;
    const int gmin = -(1 << (bd - 1)), gmax = (1 << (bd - 1)) - 1;
    // ---- luma AR (diagonal wavefront, cell (y,x) at step x + 4y)
    if (g.num_y_points && lag) {
        const int y = 3 + tid;
        for (int t = 3 + 4 * 3; t <= (GW - 4) + 4 * (GH - 1); t++) {
            if (tid < GH - 3) {
                const int x = t - 4 * y;
                if (x >= 3 && x < GW - 3)
                    ar_cell(T[0], y, x, lag, g.ar_coeffs_y, g.ar_coeff_shift, gmin, gmax, false, false, nullptr, 0, 0);
            }
            __syncthreads();
        }
    }
    // ---- chroma AR: Cb on lanes 0..127, Cr on lanes 128..255
    if (lag || g.num_y_points) {
        const int pl = 1 + (tid >> 7);
        const int y = 3 + (tid & 127);
        const int8_t *coef = pl == 1 ? g.ar_coeffs_cb : g.ar_coeffs_cr;
        for (int t = 3 + 4 * 3; t <= (cw - 4) + 4 * (ch - 1); t++) {
            if (y < ch) {
                const int x = t - 4 * y;
                if (x >= 3 && x < cw - 3)
                    ar_cell(T[pl], y, x, lag, coef, g.ar_coeff_shift, gmin, gmax, true, g.num_y_points != 0, T[0], ssx, ssy);
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < 3 * GH * GW; i += 256) a.templates[i] = (&T[0][0])[i];
}

__device__ __forceinline__ int scale_lut(const uint8_t *lut, int index, int bd)
{
    index = clipi(index, 0, (1 << bd) - 1);
    const int sh = bd - 8;
    const int x = index >> sh;
    if (sh == 0 || x == 255) return lut[x];
    const int rem = index - (x << sh);
    const int start = lut[x], end = lut[x + 1];
    return (uint8_t)(start + r2i((end - start) * rem, sh));
}

__device__ __forceinline__ int blendw(int prev, int cur, int wp, int wc, int bd)
{
    return clipi(r2i(prev * wp + cur * wc, 5), -(1 << (bd - 1)), (1 << (bd - 1)) - 1);
}

template <typename T>
__global__ __launch_bounds__(256) void grain_apply_kernel(FrameView src, FrameView dst, GrainArgs a, int matrix_identity)
{
    __shared__ uint8_t lut[3][256];
    __shared__ uint8_t offs[2][2];
    const int tid = threadIdx.x;
    const av1pf_grain_params &g = a.g;
    const int bcol = blockIdx.x, brow = blockIdx.y;
    const int bd = src.bd;
    for (int i = tid; i < 3 * 256; i += 256) (&lut[0][0])[i] = a.luts[i >> 8][i & 255];
    if (tid == 0) {
        const int nr = g.overlap && brow > 0 ? 2 : 1;
        for (int i = 0; i < 2; i++) { offs[0][i] = 0; offs[1][i] = 0; }
        for (int i = 0; i < nr; i++) {
            const int num = ((brow - i) * 32) >> 5;
            uint16_t reg = g.seed;
            reg ^= (uint16_t)(((num * 37 + 178) & 255) << 8);
            reg ^= (uint16_t)((num * 173 + 105) & 255);
            uint8_t prev = 0, cur = 0;
            for (int c = 0; c <= bcol; c++) {
                reg = lfsr_step(reg);
                prev = cur;
                cur = (uint8_t)(reg >> 8);
            }
            offs[0][i] = cur;
            offs[1][i] = (g.overlap && bcol > 0) ? prev : 0;
        }
    }
    __syncthreads();
    const int scaling_shift = g.scaling_shift;
    for (int pl = 0; pl < src.nplanes; pl++) {
        if (!a.active[pl]) continue;
        const int sx = pl ? src.ssx : 0, sy = pl ? src.ssy : 0;
        const PlaneView &sp = src.p[pl];
        const PlaneView &dp = dst.p[pl];
        const int bw = 32 >> sx, bh = 32 >> sy;
        const int px0 = bcol * bw, py0 = brow * bh;
        const int pw = sp.w, ph = sp.h;
        if (px0 >= pw || py0 >= ph) continue;
        const int bwid = min(bw, pw - px0), height = min(bh, ph - py0);
        const int ys = (g.overlap && brow > 0) ? min(height, 2 >> sy) : 0;
        const int xs = (g.overlap && bcol > 0) ? min(bwid, 2 >> sx) : 0;
        const int16_t *tmpl = a.templates + pl * GH * GW;
        int minv = 0, maxv = (256 << (bd - 8)) - 1;
        if (g.clip_to_restricted_range) {
            minv = 16 << (bd - 8);
            maxv = (pl == 0 ? 235 : (matrix_identity ? 235 : 240)) << (bd - 8);
        }
        const bool odd_luma = pl && sx && (src.p[0].w & 1);
        const int mult = pl == 1 ? g.cb_mult : g.cr_mult;
        const int lmult = pl == 1 ? g.cb_luma_mult : g.cr_luma_mult;
        const int coff = pl == 1 ? g.cb_offset : g.cr_offset;
        auto goff = [&](int o, bool row) -> int {
            const int n = row ? (o & 15) : (o >> 4);
            const int s = row ? sy : sx;
            return 3 + (2 >> s) * (3 + n);
        };
        for (int i = tid; i < bwid * height; i += 256) {
            const int y = i / bwid, x = i - y * bwid;
#define GS(o, bc, br) tmpl[(goff((o), true) + y + bh * (br)) * GW + goff((o), false) + x + bw * (bc)]
            int gv = GS(offs[0][0], 0, 0);
            const bool csub_x = pl && sx, csub_y = pl && sy;
            if (x < xs) {
                const int left = GS(offs[1][0], 1, 0);
                gv = csub_x ? blendw(left, gv, 23, 22, bd) : blendw(left, gv, x == 1 ? 17 : 27, x == 1 ? 27 : 17, bd);
            }
            if (y < ys) {
                int top = GS(offs[0][1], 0, 1);
                if (x < xs) {
                    const int tl = GS(offs[1][1], 1, 1);
                    top = csub_x ? blendw(tl, top, 23, 22, bd) : blendw(tl, top, x == 1 ? 17 : 27, x == 1 ? 27 : 17, bd);
                }
                gv = csub_y ? blendw(top, gv, 23, 22, bd) : blendw(top, gv, y == 1 ? 17 : 27, y == 1 ? 27 : 17, bd);
            }
#undef GS
            const int px = px0 + x, py = py0 + y;
            const int s = row_ptr<T>(sp, py)[px];
            int idx;
            if (pl == 0) {
                idx = s;
            } else {
                const PlaneView &lp = src.p[0];
                const int lx = px << sx, ly = py << sy;
                int avg = row_ptr<T>(lp, ly)[lx];
                if (sx && !(odd_luma && px == pw - 1)) avg = (avg + row_ptr<T>(lp, ly)[lx + 1] + 1) >> 1;
                if (g.chroma_scaling_from_luma) idx = avg;
                else {
                    const int combined = avg * (lmult - 128) + s * (mult - 128);
                    idx = clipi((combined >> 6) + ((coff << (bd - 8)) - (1 << bd)), 0, (1 << bd) - 1);
                }
            }
            const int scale = scale_lut(lut[pl], idx, bd);
            const int noise = r2i(scale * gv, scaling_shift);
            row_ptr<T>(dp, py)[px] = (T)clipi(s + noise, minv, maxv);
        }
    }
}

hipError_t launch_grain_templates(const GrainArgs &a, int bd, int ssx, int ssy, hipStream_t s)
{
    hipLaunchKernelGGL(grain_template_kernel, dim3(1), dim3(256), 0, s, a, bd, ssx, ssy);
    return hipGetLastError();
}

hipError_t launch_grain_apply(const FrameView &src, const FrameView &dst, const GrainArgs &a,
                              int matrix_identity, hipStream_t s)
{
    dim3 grid((src.w + 31) / 32, (src.h + 31) / 32);
    if (src.bd == 8) hipLaunchKernelGGL(grain_apply_kernel<uint8_t>, grid, dim3(256), 0, s, src, dst, a, matrix_identity);
    else hipLaunchKernelGGL(grain_apply_kernel<uint16_t>, grid, dim3(256), 0, s, src, dst, a, matrix_identity);
    return hipGetLastError();
}

}  // namespace pf
