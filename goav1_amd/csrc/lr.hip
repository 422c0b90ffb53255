// AV1 loop restoration (Wiener and self-guided) on gfx950.
//
// Semantics follow goav1's striped libaom flow (tile/restoration_apply.go:526-570,
// tile/restoration_boundary.go, restoration/wiener.go, restoration/selfguided.go):
//   - the plane is the visible (rounded) plane; units are offset 8 rows up
//     (restoration_geometry.go:54-93) and filtered in 64-row stripes
//     (:97-143) and 64-column processing units (:163-236);
//   - inside a stripe the filter reads the post-CDEF frame, except the 3 rows
//     above / below: interior seams read rows saved from the DEBLOCKED frame
//     (above: y0-2, y0-2, y0-1; below: y1, y1+1, y1+1 with a single saved row
//     duplicated), the frame top / bottom read the frame edge row
//     (restoration_boundary.go:192-255, 553-614);
//   - columns past the plane read the edge sample (3-sample extension of the
//     data, 4-sample extension of the saved rows: both are edge replication);
//   - Wiener: H pass with offset 2^(bd+6), round0, clamp to 2^(bd+1+7-round0)-1,
//     V pass with round1 (wiener.go:132-216);
//   - SGR: box sums, uint32 p/z/A/B math with LUTs, 6/5 (r=2, every other row)
//     and 4/3 (r=1) weighted sums, projection with an int16 wrap (selfguided.go).
//
// Decomposition: one workgroup per (frame, plane, stripe, 64-column processing
// unit); every such tile lies inside exactly one restoration unit, so the
// record is looked up per tile.  The tile and its 3-sample halo are staged in
// LDS once.  The kernel is out of place (post-CDEF frame -> output frame); NONE
// units and NONE planes copy.
#include <type_traits>
#include "pf_internal.h"
#include "lr_filters.h"
#include "superres_taps.h"

namespace pf {

// Four samples of a row starting at column x (a multiple of 4) as packed
// 16-bit pairs: one vector load when the four lie inside [0, pw) and the row
// allows al-byte accesses, else edge-clamped sample loads.
template <typename T>
__device__ __forceinline__ uint2 load4(const T *row, int x, int pw, bool vec)
{
    if (vec && x >= 0 && x + 4 <= pw) {
        if constexpr (sizeof(T) == 2) {
            return *(const uint2 *)(row + x);
        } else {
            const uint32_t v = *(const uint32_t *)(row + x);
            return make_uint2((v & 0xffu) | ((v & 0xff00u) << 8), ((v >> 16) & 0xffu) | ((v >> 8) & 0xff0000u));
        }
    }
    uint32_t s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = row[min(max(x + k, 0), pw - 1)];
    return make_uint2(s[0] | (s[1] << 16), s[2] | (s[3] << 16));
}

// Four output samples at column x of a row (a multiple of 4), n of them valid.
template <typename T>
__device__ __forceinline__ void store4(T *row, int x, const int *v, int n, bool vec)
{
    if (vec && n == 4) {
        if constexpr (sizeof(T) == 2)
            *reinterpret_cast<uint2 *>(row + x) = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                                                             (uint32_t)v[2] | ((uint32_t)v[3] << 16));
        else
            *reinterpret_cast<uint32_t *>(row + x) =
                (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
        return;
    }
    for (int k = 0; k < n; k++) row[x + k] = (T)v[k];
}

// Copy of the tile's w x h samples from sp to dp (a NONE unit): eight
// samples per lane and row when both planes allow it.
template <typename T>
__device__ __forceinline__ void copy_tile(const PlaneView &sp, const PlaneView &dp, int x0, int y0, int w, int h)
{
    const int tid = threadIdx.x;
    constexpr int VB = 8 * (int)sizeof(T);   // bytes of eight samples
    if (sp.al >= VB && dp.al >= VB) {
        using V = typename std::conditional<sizeof(T) == 2, uint4, uint2>::type;
        const int full = w >> 3, tail = w & 7, per = full + (tail ? 1 : 0);
        for (int i = tid; i < per * h; i += 256) {
            const int r = i / per, k = i - r * per;
            const T *src = row_ptr<T>(sp, y0 + r) + x0 + 8 * k;
            T *dst = row_ptr<T>(dp, y0 + r) + x0 + 8 * k;
            if (k < full) {
                *reinterpret_cast<V *>(dst) = *reinterpret_cast<const V *>(src);
            } else {
                for (int j = 0; j < tail; j++) dst[j] = src[j];
            }
        }
        return;
    }
    const int lpr = lanes_for(w), c = tid & (lpr - 1), step = 256 / lpr;
    if (c < w)
        for (int r = tid / lpr; r < h; r += step) row_ptr<T>(dp, y0 + r)[x0 + c] = row_ptr<T>(sp, y0 + r)[x0 + c];
}

// The unit record of tile (k, t) of a plane (NONE for a NONE plane).
__device__ __forceinline__ av1pf_lr_unit lr_tile_unit(const LrPlaneDesc &g, int k, int t)
{
    av1pf_lr_unit u;
    u.type = AV1PF_LR_NONE;
    if (g.type != AV1PF_LR_NONE) {
        const int x0 = k * g.procw, y0 = max(0, t * g.sh - g.off);
        const int uc = min(x0 / g.us, g.hz - 1);
        const int ur = min((y0 + g.off) / g.us, g.vt - 1);
        u = g.units[ur * g.hz + uc];
    }
    return u;
}

template <typename T>
__device__ __forceinline__ void lr_tile(const PlaneView &sp, const PlaneView &dp, const LrPlaneDesc &g, int bd,
                                        int k, int t, uint16_t *tile, uint16_t *tmp, uint16_t *Ab0, int32_t *Bb0,
                                        uint16_t *Ab1, int32_t *Bb1)
{
    if (k >= g.xtiles || t >= g.stripes) return;
    const int x0 = k * g.procw;
    const int w = min(g.procw, g.pw - x0);
    const int y0 = max(0, t * g.sh - g.off);
    const int y1 = min((t + 1) * g.sh - g.off, g.ph);
    if (w <= 0 || y0 >= y1) return;
    const int h = y1 - y0;
    const int tid = threadIdx.x;
    if (g.copy_pad) {
        // the reference restores in place and writes only the visible plane:
        // the MI-aligned padding keeps what the input holds there
        const int padw = sp.aw - g.pw;
        if (k == g.xtiles - 1 && padw > 0)
            for (int i = tid; i < padw * h; i += 256) {
                const int y = y0 + i / padw, x = g.pw + i % padw;
                row_ptr<T>(dp, y)[x] = row_ptr<T>(sp, y)[x];
            }
        const int padh = sp.ah - g.ph, cw = (k == g.xtiles - 1 ? sp.aw : x0 + w) - x0;
        if (t == g.stripes - 1 && padh > 0)
            for (int i = tid; i < padh * cw; i += 256) {
                const int y = g.ph + i / cw, x = x0 + i % cw;
                row_ptr<T>(dp, y)[x] = row_ptr<T>(sp, y)[x];
            }
    }

    const av1pf_lr_unit u = lr_tile_unit(g, k, t);
    if (u.type != AV1PF_LR_WIENER && u.type != AV1PF_LR_SGRPROJ) {
        copy_tile<T>(sp, dp, x0, y0, w, h);
        return;
    }
    // ---- stage the tile with its 3-sample halo: four samples per load, the
    // loads of a lane all in flight before its LDS writes.  Tile column
    // LR_OX + 4i - 4 holds frame column x0 - 4 + 4i.
    const int nominal = g.sh - (y0 == 0 ? g.off : 0);
    const bool copy_above = y0 != 0, copy_below = y0 + nominal < g.ph;
    const int th = h + 6, nch = (w + 6) / 4 + 1, items = th * nch;
    // the saved rows are global memory: reached through a global pointer as row_ptr does
    const uint16_t *bnd = (const uint16_t *)(gmem(g.bnd) + (int64_t)t * 4 * g.bp);
    const bool fvec = sp.al >= 4 * (int)sizeof(T);
    constexpr int KMAX = (LR_TH * 18 + 255) / 256;
    uint2 v[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        const int it = tid + j * 256;
        if (it >= items) break;
        const int rr = it / nch, ci = it - rr * nch;
        const int r = rr - 3, x = x0 - 4 + 4 * ci;
        if (r < 0 && copy_above) {
            v[j] = load4<uint16_t>(bnd + (r == -1 ? 1 : 0) * g.bp, x, g.pw, true);
        } else if (r >= h && copy_below) {
            v[j] = load4<uint16_t>(bnd + (2 + min(r - h, 1)) * g.bp, x, g.pw, true);
        } else {
            v[j] = load4<T>(row_ptr<T>(sp, min(max(y0 + r, 0), g.ph - 1)), x, g.pw, fvec);
        }
    }
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        const int it = tid + j * 256;
        if (it >= items) break;
        const int rr = it / nch, ci = it - rr * nch;
        *reinterpret_cast<uint2 *>(tile + rr * LR_TW + 4 * ci) = v[j];
    }
    __syncthreads();
    const uint16_t *O = tile + 3 * LR_TW + LR_OX;   // tile origin
    if (u.type == AV1PF_LR_WIENER) {
        const bool dvec = dp.al >= 4 * (int)sizeof(T);
        wiener_tile(O, w, h, bd, u.wiener_v, u.wiener_h, tmp, [&](int y, int x, const int *o, int n) {
            store4<T>(row_ptr<T>(dp, y0 + y), x0 + x, o, n, dvec);
        });
    } else {
        sgr_tile_parts(O, w, h, bd, u.sgr_set, u.sgr_xqd, Ab0, Bb0, Ab1, Bb1,
                       [&](int y, int x, int o) { row_ptr<T>(dp, y0 + y)[x0 + x] = (T)o; });
    }
}

// LDS of an LR tile: the stripe tile with its halo, and one buffer that holds
// the Wiener intermediate or the self-guided A/B planes of both passes for a
// part of the tile.
constexpr int LR_ABN = LR_AB * (LR_SGR_ROWS + 2);
struct __align__(16) LrLds {
    uint16_t tile[LR_TW * LR_TH];
    char work[LR_ABN * 12];
};
static_assert(LR_TH * 64 * 2 <= LR_ABN * 12, "Wiener intermediate fits the shared buffer");

// Processing unit k of stripe st of plane pl of a frame.
__device__ __forceinline__ void lr_tile_of(const FrameDesc &D, int pl, int k, int st, LrLds &S)
{
    int32_t *Bb0 = reinterpret_cast<int32_t *>(S.work);
    int32_t *Bb1 = Bb0 + LR_ABN;
    uint16_t *Ab0 = reinterpret_cast<uint16_t *>(S.work + LR_ABN * 8);
    uint16_t *Ab1 = Ab0 + LR_ABN;
    uint16_t *tmp = reinterpret_cast<uint16_t *>(S.work);
    const FrameView &src = D.v[D.lr_src];
    const FrameView &dst = D.v[D.lr_dst];
    if (src.bd == 8)
        lr_tile<uint8_t>(src.p[pl], dst.p[pl], D.lr[pl], 8, k, st, S.tile, tmp, Ab0, Bb0, Ab1, Bb1);
    else
        lr_tile<uint16_t>(src.p[pl], dst.p[pl], D.lr[pl], src.bd, k, st, S.tile, tmp, Ab0, Bb0, Ab1, Bb1);
}

// 6 workgroups per CU: 23.8 KB of LDS each, and at most 80 VGPRs
__global__ __launch_bounds__(256, 6) void lr_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs, int nseg,
                                                 int total, StageLaunch L)
{
    __shared__ LrLds S;
    for_each_unit(total, L, [&](int u) {
        const Seg sg = segs[find_seg(segs, nseg, u)];
        const int t = u - sg.base;
        lr_tile_of(descs[sg.frame], sg.plane, t % sg.nx, t / sg.nx, S);
    });
}

// Deblocked context rows of stripe t, column x.  With super-resolution the
// rows are read from the coded frame and upscaled to the LR plane width
// (restoration_boundary.go:257-345): sr is the plane's upscale geometry.
template <typename T>
__device__ __forceinline__ void lr_save_cols(const PlaneView &pv, const LrPlaneDesc &g, const SrPlane *sr, int maxv,
                                             int x, int t)
{
    if (x >= g.pw || t >= g.stripes) return;
    const int y0 = max(0, t * g.sh - g.off);
    const int y1 = min((t + 1) * g.sh - g.off, g.ph);
    const auto b = gmem(g.bnd) + (int64_t)t * 4 * g.bp;
    auto px = [&](int y) { return sr ? sr_sample<T>(row_ptr<T>(pv, y), *sr, x, maxv) : (int)row_ptr<T>(pv, y)[x]; };
    if (t > 0) {
        b[x] = px(y0 - 2);
        b[g.bp + x] = px(y0 - 1);
    }
    if (y1 < g.ph) {
        b[2 * g.bp + x] = px(y1);
        b[3 * g.bp + x] = px(y1 + 1 < g.ph ? y1 + 1 : y1);
    }
}

// 256 columns (chunk xc) of the context rows of stripe st of plane pl.
__device__ __forceinline__ void lr_save_tile(const FrameDesc &D, int pl, int xc, int st)
{
    const int x = xc * 256 + threadIdx.x;
    const FrameView &f = D.v[VIEW_A];
    const LrPlaneDesc &g = D.lr[pl];
    const SrPlane *sr = g.sr ? &D.sr[pl] : nullptr;
    const int maxv = (1 << f.bd) - 1;
    if (f.bd == 8) lr_save_cols<uint8_t>(f.p[pl], g, sr, maxv, x, st);
    else lr_save_cols<uint16_t>(f.p[pl], g, sr, maxv, x, st);
}

__global__ __launch_bounds__(256) void lr_save_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs, int nseg)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const int t = blockIdx.x - sg.base;
    lr_save_tile(descs[sg.frame], sg.plane, t % sg.nx, t / sg.nx);
}

hipError_t launch_lr_save(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(lr_save_kernel, dim3(total), dim3(256), 0, s, d, segs, nseg);
    return hipGetLastError();
}

hipError_t launch_lr(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L, hipStream_t s)
{
    if (total)
        hipLaunchKernelGGL(lr_kernel, dim3(stage_grid((const void *)lr_kernel, total, L.per_wg)), dim3(256), 0, s, d, segs,
                           nseg, total, L);
    return hipGetLastError();
}

__global__ void zero_u32_kernel(unsigned int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(unsigned int *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

int stage_grid(const void *kernel, int total, int per_wg)
{
    if (per_wg > 0) {
        const int g = (total + per_wg - 1) / per_wg, step = kXcds * kChunk;
        return (g + step - 1) / step * step;
    }
    static int cus = 0;
    int dev = 0, per_cu = 0;
    if (!cus && (hipGetDevice(&dev) != hipSuccess ||
                 hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess))
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return std::min(total, cus * per_cu);
}

// Plain copy of one plane's aligned extent.  A workgroup copies COPY_COLS
// samples of COPY_ROWS rows; a lane moves eight adjacent samples with one
// 16-byte (8-bit: 8-byte) load and store when both planes allow it.
template <typename T>
__device__ __forceinline__ void copy_cols(const PlaneView &s, const PlaneView &d, int x, int y)
{
    if (x >= s.aw || y >= s.ah) return;
    const T *src = row_ptr<T>(s, y) + x;
    T *dst = row_ptr<T>(d, y) + x;
    constexpr int VB = 8 * (int)sizeof(T);
    if (x + 8 <= s.aw && s.al >= VB && d.al >= VB) {
        using V = typename std::conditional<sizeof(T) == 2, uint4, uint2>::type;
        *(V *)dst = *(const V *)src;
        return;
    }
    for (int k = 0; k < 8 && x + k < s.aw; k++) dst[k] = src[k];
}

__global__ __launch_bounds__(256) void copy_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs,
                                                   int nseg, int src, int dst)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int x = (t % sg.nx) * COPY_COLS + (threadIdx.x % (COPY_COLS / 8)) * 8;
    const int y = (t / sg.nx) * COPY_ROWS + threadIdx.x / (COPY_COLS / 8);
    const FrameView &S = D.v[src < 0 ? D.fin : src];
    if (S.bd == 8) copy_cols<uint8_t>(S.p[sg.plane], D.v[dst].p[sg.plane], x, y);
    else copy_cols<uint16_t>(S.p[sg.plane], D.v[dst].p[sg.plane], x, y);
}

hipError_t launch_copy(const FrameDesc *d, const Seg *segs, int nseg, int total, int src, int dst, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(copy_kernel, dim3(total), dim3(256), 0, s, d, segs, nseg, src, dst);
    return hipGetLastError();
}

}  // namespace pf
