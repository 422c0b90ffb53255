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
// Templates: one wave per template.  Each lane fills template rows of Gaussian
// samples from the LFSR state at its row start (the host jumps the LFSR with a
// GF(2) matrix power; the LFSR step is linear over GF(2)).  The AR recursion,
// sequential in the reference, runs as a diagonal wavefront: cell (y, x) only
// needs earlier rows within LAG columns and the LAG cells to its left, so all
// cells with x + (LAG+1)*y = t are independent.  A lane owns a row and keeps
// the windows of the rows above in registers, fed by DPP from the lane above,
// so a step has no LDS round trip and no barrier.  The lag is a template
// parameter, so the taps are unrolled with the coefficients in registers.
// Templates depend only on the grain parameters, so the host runs them on a
// side stream overlapped with deblock / CDEF / LR.
// The luma template kernel also draws the per-block random offsets (one LFSR
// walk per 32-line stripe) into a small table.
// The luma template kernel also expands each plane's scaling function to all
// 1 << bd sample values (one byte each) next to the templates.
// Apply: one workgroup per 32x32 luma block and its chroma blocks (the offsets
// of block (row, col) are shared by all planes); a wave per 16-row band of a
// plane, the plane's expanded scaling table in the wave's LDS slice.  Planes
// without grain are copied, so the display frame is complete.
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

// Gaussian fill of one template (rows [0,h), cols [0,w)) by the workgroup;
// thread r starts from the LFSR state at the start of row r (host-computed)
// and reads the Gaussian sequence from its copy in LDS (gz), eight lookups
// in flight at a time.
__device__ __forceinline__ void gauss_fill(int16_t *T, const int16_t *gz, const uint16_t *row_seed, int w, int h,
                                           int shift)
{
    for (int row = threadIdx.x; row < h; row += blockDim.x) {
        uint16_t s = row_seed[row];
        int x = 0;
        for (; x + 8 <= w; x += 8) {
            int v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                s = lfsr_step(s);
                v[k] = gz[s >> 5];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) T[row * GW + x + k] = (int16_t)r2i(v[k], shift);
        }
        for (; x < w; x++) {
            s = lfsr_step(s);
            T[row * GW + x] = (int16_t)r2i(gz[s >> 5], shift);
        }
    }
}

// Template kernels run a workgroup of GT_THREADS: every thread shares the
// fills and copies, and the first wave runs the autoregressive wavefront (the
// other waves walk the same loop with no live rows, so that the passes'
// barriers are reached by all of them).
constexpr int GT_THREADS = 256;

// The Gaussian sequence and (luma kernel) the three scaling tables in LDS.
__device__ __forceinline__ void stage_tables(int16_t *gz)
{
    for (int i = threadIdx.x; i < 2048; i += GT_THREADS) gz[i] = kGauss[i];
}

// A template (GH x GW int16) between LDS and device memory, two samples per
// 32-bit access (both ends are 4-byte aligned).
__device__ __forceinline__ void tmpl_store(gptr<int16_t> out, const int16_t *T)
{
    const auto o = (gptr<uint32_t>)out;
    const uint32_t *t = reinterpret_cast<const uint32_t *>(T);
    for (int i = threadIdx.x; i < GH * GW / 2; i += GT_THREADS) o[i] = t[i];
}

// Lane i receives lane i-1's value (DPP wave_shr:1); lane 0 is overwritten by
// the caller.
__device__ __forceinline__ int from_prev_lane(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}

// Auto-regressive filter (luma.go:60-74, chroma.go:69-87) over rows
// [y0, y0 + nrows), nrows <= 64, by one wave.  Lane l owns row y0 + l and at
// step t works on column x = t - (LAG+1)*l: every cell it needs from the rows
// above was finished at an earlier step, and its own row's left cells are its
// own previous results.  The (2*LAG+1)-wide windows of the LAG rows above live
// in registers and advance one column per step; the new right-hand column
// comes from the lane above through DPP (its newest result and the centres of
// its own windows), and lane 0 reads it from the rows above the pass in LDS,
// which are final.  T holds the Gaussian noise on entry and the filtered
// template on exit for the pass's rows.
template <int LAG, bool CHROMA>
__device__ void ar_rows(int16_t *T, const int16_t *L, int w, int y0, int nrows, const int *c, int shift, int gmin,
                        int gmax, bool luma_term, int ssx, int ssy)
{
    constexpr int K = LAG + 1, WIN = 2 * LAG + 1;
    const int lane = threadIdx.x;
    const int y = y0 + lane;
    const bool live = lane < nrows;
    // The windows are circular: steps go in blocks of WIN, and at phase u of a
    // block (t = t0 + u) column x - LAG + j of row y - 1 - d sits in
    // A[d][(u + j) % WIN] and this row's result of step t' in O[phase of t'],
    // so a step only writes its new column over the oldest one (all indices
    // compile-time constants: registers, no shifting moves).
    int A[LAG][WIN];
    int O[WIN];
#pragma unroll
    for (int j = 0; j < WIN; j++) {
        O[j] = 0;
#pragma unroll
        for (int d = 0; d < LAG; d++) A[d][j] = 0;
    }
    const int t_end = (w - 4) + K * (nrows - 1);
    for (int t0 = -K; t0 <= t_end; t0 += WIN) {
#pragma unroll
        for (int u = 0; u < WIN; u++) {
            const int t = t0 + u;
            if (t > t_end) break;
            const int x = t - K * lane;
            int v = (live && x >= 0 && x < w) ? T[y * GW + x] : 0;
            if (live && x >= 3 && x < w - 3) {
                int sum = 0, pos = 0;
#pragma unroll
                for (int dr = LAG; dr >= 1; dr--) {
#pragma unroll
                    for (int j = 0; j < WIN; j++) sum += __mul24(c[pos++], A[dr - 1][(u + j) % WIN]);
                }
#pragma unroll
                for (int j = 0; j < LAG; j++) sum += __mul24(c[pos++], O[(u - LAG + j + WIN) % WIN]);
                if (CHROMA && luma_term) {
                    const int lx = ((x - 3) << ssx) + 3, ly = ((y - 3) << ssy) + 3;
                    int a = L[ly * GW + lx];
                    if (ssx) a += L[ly * GW + lx + 1];
                    if (ssy) {
                        a += L[(ly + 1) * GW + lx];
                        if (ssx) a += L[(ly + 1) * GW + lx + 1];
                    }
                    sum += r2i(a, ssx + ssy) * c[pos];
                }
                v = clipi(v + ((sum + (1 << (shift - 1))) >> shift), gmin, gmax);
                T[y * GW + x] = (int16_t)v;
            }
            // right-hand column (x + 1 + LAG) of the windows for the next step:
            // the lane above's newest result and the centres of its windows
            int nv = from_prev_lane(v);
            int na[LAG > 1 ? LAG - 1 : 1];
#pragma unroll
            for (int d = 0; d + 1 < LAG; d++) na[d] = from_prev_lane(A[d][(u + LAG) % WIN]);
            if (lane == 0) {
                const int xx = x + 1 + LAG;
                const bool ok = xx >= 0 && xx < w;
                nv = ok ? T[(y - 1) * GW + xx] : 0;
#pragma unroll
                for (int d = 0; d + 1 < LAG; d++) na[d] = ok ? T[(y - 2 - d) * GW + xx] : 0;
            }
            A[0][u] = nv;
#pragma unroll
            for (int d = 1; d < LAG; d++) A[d][u] = na[d - 1];
            O[u] = v;
        }
    }
}

// Lag 0 (chroma with only the luma term): every cell is independent.
__device__ void ar_lag0(int16_t *T, const int16_t *L, int w, int h, int c0, int shift, int gmin, int gmax, int ssx,
                        int ssy)
{
    const int cols = w - 6;
    for (int i = threadIdx.x; i < (h - 3) * cols; i += blockDim.x) {
        const int y = 3 + i / cols, x = 3 + i % cols;
        const int lx = ((x - 3) << ssx) + 3, ly = ((y - 3) << ssy) + 3;
        int a = L[ly * GW + lx];
        if (ssx) a += L[ly * GW + lx + 1];
        if (ssy) {
            a += L[(ly + 1) * GW + lx];
            if (ssx) a += L[(ly + 1) * GW + lx + 1];
        }
        const int sum = r2i(a, ssx + ssy) * c0;
        T[y * GW + x] = (int16_t)clipi(T[y * GW + x] + ((sum + (1 << (shift - 1))) >> shift), gmin, gmax);
    }
}

// Rows 3 .. h-1 in passes of up to 64 rows (luma: 3..66, then 67..72).
template <int LAG, bool CHROMA>
__device__ void ar_template(int16_t *T, const int16_t *L, int w, int h, const int8_t *coef, int shift, int gmin,
                            int gmax, bool luma_term, int ssx, int ssy)
{
    constexpr int NC = 2 * LAG * (LAG + 1) + (CHROMA ? 1 : 0);
    int c[NC > 0 ? NC : 1];
#pragma unroll
    for (int i = 0; i < NC; i++) c[i] = coef[i];
    for (int y0 = 3; y0 < h; y0 += 64) {
        ar_rows<LAG, CHROMA>(T, L, w, y0, min(64, h - y0), c, shift, gmin, gmax, luma_term, ssx, ssy);
        __syncthreads();
    }
}

__device__ void expand_scaling(const GrainDesc &G, const uint8_t (*luts)[256], int bd);

__global__ __launch_bounds__(GT_THREADS) void grain_luma_template_kernel(const FrameDesc *__restrict__ descs,
                                                                         const Seg *__restrict__ segs)
{
    __shared__ __align__(16) int16_t T[GH * GW];
    __shared__ int16_t gz[2048];
    __shared__ uint8_t luts[3][256];
    // a serial dependency chain shares its SIMD with the streaming kernels of
    // the main stream: let it issue first
    __builtin_amdgcn_s_setprio(3);
    const FrameDesc &D = descs[segs[blockIdx.x].frame];
    const GrainDesc &G = D.gr;
    const int bd = D.v[VIEW_A].bd;
    stage_tables(gz);
    for (int i = threadIdx.x; i < 3 * 256; i += GT_THREADS) luts[i >> 8][i & 255] = G.luts[i >> 8][i & 255];
    for (int i = threadIdx.x; i < GH * GW; i += GT_THREADS) T[i] = 0;
    // per-block random offsets: one LFSR per 32-line stripe, one draw per
    // 32-column block (apply.go:43-70, random.go:26-48)
    for (int r = threadIdx.x; r < G.orows; r += GT_THREADS) {
        uint16_t reg = G.g.seed;
        reg ^= (uint16_t)(((r * 37 + 178) & 255) << 8);
        reg ^= (uint16_t)((r * 173 + 105) & 255);
        const auto o = gmem(G.offsets) + (size_t)r * G.ocols;
        for (int c = 0; c < G.ocols; c++) {
            reg = lfsr_step(reg);
            o[c] = (uint8_t)(reg >> 8);
        }
    }
    __syncthreads();
    expand_scaling(G, luts, bd);
    if (G.g.num_y_points) {
        gauss_fill(T, gz, G.row_seed[0], GW, GH, 12 - bd + G.g.grain_scale_shift);
        __syncthreads();
        const int gmin = -(1 << (bd - 1)), gmax = (1 << (bd - 1)) - 1;
        const int8_t *cf = G.g.ar_coeffs_y;
        const int sh = G.g.ar_coeff_shift;
        switch (G.g.ar_coeff_lag) {
        case 1: ar_template<1, false>(T, nullptr, GW, GH, cf, sh, gmin, gmax, false, 0, 0); break;
        case 2: ar_template<2, false>(T, nullptr, GW, GH, cf, sh, gmin, gmax, false, 0, 0); break;
        case 3: ar_template<3, false>(T, nullptr, GW, GH, cf, sh, gmin, gmax, false, 0, 0); break;
        default: break;
        }
    }
    __syncthreads();
    tmpl_store(gmem(G.templates), T);
}

__global__ __launch_bounds__(GT_THREADS) void grain_chroma_template_kernel(const FrameDesc *__restrict__ descs,
                                                                           const Seg *__restrict__ segs)
{
    __shared__ __align__(16) int16_t T[GH * GW];
    __shared__ __align__(16) int16_t L[GH * GW];
    __shared__ int16_t gz[2048];
    __builtin_amdgcn_s_setprio(3);
    const Seg sg = segs[blockIdx.x];
    const FrameDesc &D = descs[sg.frame];
    const GrainDesc &G = D.gr;
    const FrameView &f = D.v[VIEW_A];
    const int bd = f.bd, ssx = f.ssx, ssy = f.ssy;
    const int pl = sg.plane;
    const int cw = ssx ? 44 : GW, ch = ssy ? 38 : GH;
    const bool luma_term = G.g.num_y_points != 0;
    stage_tables(gz);
    {
        const auto lin = (gptr<const uint32_t>)gmem(G.templates);
        uint32_t *l32 = reinterpret_cast<uint32_t *>(L), *t32 = reinterpret_cast<uint32_t *>(T);
        for (int i = threadIdx.x; i < GH * GW / 2; i += GT_THREADS) {
            t32[i] = 0u;
            l32[i] = luma_term ? lin[i] : 0u;
        }
    }
    __syncthreads();
    gauss_fill(T, gz, G.row_seed[pl], cw, ch, 12 - bd + G.g.grain_scale_shift);
    __syncthreads();
    const int lag = G.g.ar_coeff_lag;
    if (lag || luma_term) {
        const int8_t *cf = pl == 1 ? G.g.ar_coeffs_cb : G.g.ar_coeffs_cr;
        const int gmin = -(1 << (bd - 1)), gmax = (1 << (bd - 1)) - 1;
        const int sh = G.g.ar_coeff_shift;
        switch (lag) {
        case 0: ar_lag0(T, L, cw, ch, cf[0], sh, gmin, gmax, ssx, ssy); break;
        case 1: ar_template<1, true>(T, L, cw, ch, cf, sh, gmin, gmax, luma_term, ssx, ssy); break;
        case 2: ar_template<2, true>(T, L, cw, ch, cf, sh, gmin, gmax, luma_term, ssx, ssy); break;
        default: ar_template<3, true>(T, L, cw, ch, cf, sh, gmin, gmax, luma_term, ssx, ssy); break;
        }
    }
    __syncthreads();
    tmpl_store(gmem(G.templates) + pl * GH * GW, T);
}

hipError_t launch_grain_templates(const FrameDesc *d, const Seg *luma, int nl, const Seg *chroma, int nc, hipStream_t s)
{
    if (nl) hipLaunchKernelGGL(grain_luma_template_kernel, dim3(nl), dim3(GT_THREADS), 0, s, d, luma);
    if (nc) hipLaunchKernelGGL(grain_chroma_template_kernel, dim3(nc), dim3(GT_THREADS), 0, s, d, chroma);
    return hipGetLastError();
}

// Products below are of values under 2^12 and weights under 2^8: 24-bit
// multiplies (full rate) give the same results.
__device__ __forceinline__ int scale_lut(const uint8_t *lut, int index, int bd)
{
    index = clipi(index, 0, (1 << bd) - 1);
    const int sh = bd - 8;
    const int x = index >> sh;
    if (sh == 0 || x == 255) return lut[x];
    const int rem = index - (x << sh);
    const int start = lut[x], end = lut[x + 1];
    return (uint8_t)(start + r2i(__mul24(end - start, rem), sh));
}

__device__ __forceinline__ int blendw(int prev, int cur, int wp, int wc, int bd)
{
    return clipi(r2i(__mul24(prev, wp) + __mul24(cur, wc), 5), -(1 << (bd - 1)), (1 << (bd - 1)) - 1);
}

// Scaling function of every sample value (1 << bd entries per plane), from the
// 256-entry piecewise-linear table with the 10/12-bit interpolation: the apply
// then looks a sample's scale up with one LDS byte read.
__device__ void expand_scaling(const GrainDesc &G, const uint8_t (*luts)[256], int bd)
{
    const auto out = gmem(G.scale);
    for (int i = threadIdx.x; i < 3 << bd; i += GT_THREADS) {
        const int pl = i >> bd, v = i & ((1 << bd) - 1);
        out[pl * GR_SCALE + v] = (uint8_t)scale_lut(luts[pl], v, bd);
    }
}

// One task of the apply: 16 rows of one plane of a block, lane (r, g) owning
// row r and the C adjacent columns from g * C (C = block width / 4).  The
// lane's samples, template values and (chroma) luma samples are read with
// one vector load each; the overlap neighbours' template values only by the
// lanes in the 2-sample overlap band.  A group reaching past the plane's
// right edge reads sample by sample with clamped positions.  The plane's
// expanded scaling table is loaded by the whole wave together with the task's
// own loads and written to the wave's LDS slice before the first lookup, so a
// task waits for memory once.
template <typename T, int C, bool CHROMA>
__device__ __forceinline__ void grain_task(const FrameView &src, const FrameView &dst, const GrainDesc &G,
                                           const int32_t *K, int pl, int bcol, int brow, int rg, uint8_t *lut,
                                           const uint8_t *offs, int lane)
{
    const int bd = src.bd;
    const int sx = CHROMA ? src.ssx : 0, sy = CHROMA ? src.ssy : 0;
    const PlaneView &sp = src.p[pl];
    const PlaneView &dp = dst.p[pl];
    const int bw = 32 >> sx, bh = 32 >> sy;
    const int px0 = bcol * bw, py0 = brow * bh;
    const int pw = sp.w, ph = sp.h;
    const int height = min(bh, ph - py0), bwid = min(bw, pw - px0);
    const int y = rg * 16 + (lane >> 2), x = (lane & 3) * C;
    const bool act = y < height && x < bwid;
    const int px = px0 + x, py = py0 + y;
    const bool full = px + C <= pw;
    const auto srow = gmem(reinterpret_cast<const T *>(sp.base + (int64_t)py * sp.stride));
    const auto drow = gmem(reinterpret_cast<T *>(dp.base + (int64_t)py * dp.stride));
    if (!K[GK_ACT0 + pl]) {   // plane without grain: the display gets the filtered samples
        if (!act) return;
        int sa[C];
        if (full) {
#pragma unroll
            for (int j = 0; j < C; j++) sa[j] = srow[px + j];
#pragma unroll
            for (int j = 0; j < C; j++) drow[px + j] = (T)sa[j];
        } else {
#pragma unroll
            for (int j = 0; j < C; j++) sa[j] = srow[min(px + j, pw - 1)];
#pragma unroll
            for (int j = 0; j < C; j++)
                if (px + j < pw) drow[px + j] = (T)sa[j];
        }
        return;
    }
    // the first 1024 bytes of the scaling table (all of it below 12 bits)
    const auto sc = reinterpret_cast<gptr<const uint32_t>>(gmem(G.scale) + pl * GR_SCALE);
    const int n16 = (1 << bd) >> 4;
    uint4 s0 = make_uint4(0, 0, 0, 0);
    if (lane < n16) {
        const auto q = sc + 4 * lane;
        s0 = make_uint4(q[0], q[1], q[2], q[3]);
    }
    const int ys = (K[GK_OVERLAP] && brow > 0) ? min(height, 2 >> sy) : 0;
    const int xs = (K[GK_OVERLAP] && bcol > 0) ? min(bwid, 2 >> sx) : 0;
    const bool left = act && x < xs, top = act && y < ys;
    constexpr int NL = CHROMA ? 2 * C : 1;   // luma samples under a chroma group: 2C or C
    int sa[C], gv[C], la[NL], tp[C];
    int gl0 = 0, gl1 = 0, gu0 = 0, gu1 = 0;
    int lw = 0;
    const auto tmpl = gmem(G.templates) + pl * GH * GW;
    if (act) {
        if (full) {
#pragma unroll
            for (int j = 0; j < C; j++) sa[j] = srow[px + j];
        } else {
#pragma unroll
            for (int j = 0; j < C; j++) sa[j] = srow[min(px + j, pw - 1)];
        }
        if constexpr (CHROMA) {
            const PlaneView &lp = src.p[0];
            lw = lp.w;
            const auto lrow = gmem(reinterpret_cast<const T *>(lp.base + (int64_t)(py << sy) * lp.stride));
            const int lx = px << sx, nl = C << sx;
            if (lx + nl <= lw) {
#pragma unroll
                for (int j = 0; j < NL; j++)
                    if (j < nl) la[j] = lrow[lx + j];
            } else {
#pragma unroll
                for (int j = 0; j < NL; j++)
                    if (j < nl) la[j] = lrow[min(lx + j, lw - 1)];
            }
        }
        // template values (sample.go:9-77 offsets)
        const int ox0 = 3 + (2 >> sx) * (3 + (offs[0] >> 4)), oy0 = 3 + (2 >> sy) * (3 + (offs[0] & 15));
        const auto t = tmpl + (oy0 + y) * GW + ox0 + x;
#pragma unroll
        for (int j = 0; j < C; j++) gv[j] = t[j];
        if (left) {
            const int ox1 = 3 + (2 >> sx) * (3 + (offs[1] >> 4)), oy1 = 3 + (2 >> sy) * (3 + (offs[1] & 15));
            const auto u = tmpl + (oy1 + y) * GW + ox1 + bw;
            gl0 = u[0];
            gl1 = u[1];
        }
        if (top) {
            const int ox2 = 3 + (2 >> sx) * (3 + (offs[2] >> 4)), oy2 = 3 + (2 >> sy) * (3 + (offs[2] & 15));
            const auto u = tmpl + (oy2 + y + bh) * GW + ox2 + x;
#pragma unroll
            for (int j = 0; j < C; j++) tp[j] = u[j];
            if (left) {
                const int ox3 = 3 + (2 >> sx) * (3 + (offs[3] >> 4)), oy3 = 3 + (2 >> sy) * (3 + (offs[3] & 15));
                const auto v = tmpl + (oy3 + y + bh) * GW + ox3 + bw;
                gu0 = v[0];
                gu1 = v[1];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();   // the previous task of this wave is done with the table
    if (lane < n16) reinterpret_cast<uint4 *>(lut)[lane] = s0;
    for (int k = 1; k < GR_SCALE / 16 / 64; k++)   // 12-bit: the other 3 KB
        if (lane + 64 * k < n16) {
            const auto q = sc + 4 * (lane + 64 * k);
            reinterpret_cast<uint4 *>(lut)[lane + 64 * k] = make_uint4(q[0], q[1], q[2], q[3]);
        }
    __builtin_amdgcn_wave_barrier();   // LDS operations of a wave complete in order
    if (!act) return;
    // horizontal overlap weights of columns 0 and 1 (blend.go)
    const int wl0 = sx ? 23 : 27, wc0 = sx ? 22 : 17;
    if (left) {
        gv[0] = blendw(gl0, gv[0], wl0, wc0, bd);
        if (xs > 1) gv[1] = blendw(gl1, gv[1], 17, 27, bd);
    }
    if (top) {
        if (left) {
            tp[0] = blendw(gu0, tp[0], wl0, wc0, bd);
            if (xs > 1) tp[1] = blendw(gu1, tp[1], 17, 27, bd);
        }
        const int wt = sy ? 23 : (y == 1 ? 17 : 27), wb = sy ? 22 : (y == 1 ? 27 : 17);
#pragma unroll
        for (int j = 0; j < C; j++) gv[j] = blendw(tp[j], gv[j], wt, wb, bd);
    }
    int minv = 0, maxv = (256 << (bd - 8)) - 1;
    if (K[GK_CLIP]) {
        minv = 16 << (bd - 8);
        maxv = (!CHROMA ? 235 : (K[GK_MID] ? 235 : 240)) << (bd - 8);
    }
    const int imax = (1 << bd) - 1;
    const int shift = K[GK_SHIFT], rnd = 1 << (shift - 1);
    int mult = 0, lmult = 0, cbase = 0;
    if constexpr (CHROMA) {
        mult = (pl == 1 ? K[GK_CB_MULT] : K[GK_CR_MULT]) - 128;
        lmult = (pl == 1 ? K[GK_CB_LMULT] : K[GK_CR_LMULT]) - 128;
        cbase = ((pl == 1 ? K[GK_CB_OFF] : K[GK_CR_OFF]) << (bd - 8)) - (1 << bd);
    }
    int out[C];
#pragma unroll
    for (int j = 0; j < C; j++) {
        int idx = sa[j];
        if constexpr (CHROMA) {
            int avg = la[j];
            if (sx) {
                // the last chroma column of an odd-width luma plane has one luma sample
                avg = (((px + j) << 1) + 1 < lw) ? (la[2 * j] + la[2 * j + 1] + 1) >> 1 : la[2 * j];
            }
            if (K[GK_CSFL]) idx = avg;
            else idx = clipi(((__mul24(avg, lmult) + __mul24(sa[j], mult)) >> 6) + cbase, 0, imax);
        }
        const int scale = lut[min((unsigned)idx, (unsigned)imax)];
        const int noise = (__mul24(scale, gv[j]) + rnd) >> shift;
        out[j] = clipi(sa[j] + noise, minv, maxv);
    }
    if (full) {
#pragma unroll
        for (int j = 0; j < C; j++) drow[px + j] = (T)out[j];
    } else {
#pragma unroll
        for (int j = 0; j < C; j++)
            if (px + j < pw) drow[px + j] = (T)out[j];
    }
}

// A unit is one 32x32 luma block position with its chroma blocks.  Its tasks
// (16-row bands of a plane: 2 of luma, 2 >> ssy of each chroma plane) go to
// the workgroup's 4 waves in turn; each wave stages its plane's expanded
// scaling table into its own LDS slice (LDS operations of one wave complete in
// order, so no barrier is needed) and reads the block offsets of the block and
// of its left / top / top-left neighbours itself.
__global__ __launch_bounds__(256) void grain_apply_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs,
                                                          int nseg, int total, StageLaunch L)
{
    __shared__ __align__(16) uint8_t slut[4][GR_SCALE];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint8_t *lut = slut[wave];
    for_each_unit(total, L, [&](int u) {
        const Seg sg = segs[find_seg(segs, nseg, u)];
        const FrameDesc &D = descs[sg.frame];
        const GrainDesc &G = D.gr;
        const int t = u - sg.base;
        const int bcol = t % sg.nx, brow = t / sg.nx;
        int32_t K[GK_N];
#pragma unroll
        for (int i = 0; i < GK_N; i++) K[i] = G.k[i];
        const FrameView &src = D.v[D.fin];
        const FrameView &dst = D.v[VIEW_D];
        const int ncb = src.nplanes > 1 ? 2 >> src.ssy : 0;   // chroma tasks per plane
        const int ntask = 2 + 2 * ncb;
        // offsets of this block and its left / top / top-left neighbours
        // (those outside the frame or without overlap are not read: 0)
        uint8_t offs[4];
        {
            const auto o = gmem(G.offsets) + (size_t)brow * K[GK_OCOLS] + bcol;
            const bool ov = K[GK_OVERLAP];
            const bool hl = ov && bcol > 0, ht = ov && brow > 0;
            const uint8_t o0 = o[0];
            const uint8_t o1 = o[hl ? -1 : 0];
            const uint8_t o2 = o[ht ? -K[GK_OCOLS] : 0];
            const uint8_t o3 = o[(hl && ht) ? -K[GK_OCOLS] - 1 : 0];
            offs[0] = o0;
            offs[1] = hl ? o1 : 0;
            offs[2] = ht ? o2 : 0;
            offs[3] = (hl && ht) ? o3 : 0;
        }
        const int bd = src.bd;
        for (int task = wave; task < ntask; task += 4) {
            const int pl = task < 2 ? 0 : 1 + (task - 2) / ncb;
            const int rg = task < 2 ? task : (task - 2) % ncb;
            if (bd == 8) {
                if (pl == 0) grain_task<uint8_t, 8, false>(src, dst, G, K, 0, bcol, brow, rg, lut, offs, lane);
                else if (src.ssx) grain_task<uint8_t, 4, true>(src, dst, G, K, pl, bcol, brow, rg, lut, offs, lane);
                else grain_task<uint8_t, 8, true>(src, dst, G, K, pl, bcol, brow, rg, lut, offs, lane);
            } else {
                if (pl == 0) grain_task<uint16_t, 8, false>(src, dst, G, K, 0, bcol, brow, rg, lut, offs, lane);
                else if (src.ssx) grain_task<uint16_t, 4, true>(src, dst, G, K, pl, bcol, brow, rg, lut, offs, lane);
                else grain_task<uint16_t, 8, true>(src, dst, G, K, pl, bcol, brow, rg, lut, offs, lane);
            }
        }
    });
}

hipError_t launch_grain_apply(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L,
                              hipStream_t s)
{
    if (total)
        hipLaunchKernelGGL(grain_apply_kernel, dim3(stage_grid((const void *)grain_apply_kernel, total, L.per_wg)),
                           dim3(256), 0, s, d, segs, nseg, total, L);
    return hipGetLastError();
}

}  // namespace pf
