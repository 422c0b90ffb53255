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
#include "pf_internal.h"
#include "lr_filters.h"
#include "superres_taps.h"

namespace pf {

template <typename T>
__device__ __forceinline__ int read_px(const PlaneView &pv, int y, int x)
{
    return row_ptr<T>(pv, y)[x];
}

template <typename T>
__device__ __forceinline__ void lr_tile(const PlaneView &sp, const PlaneView &dp, const LrPlaneDesc &g, int bd,
                                        int k, int t, uint16_t *tile, uint16_t *tmp, uint16_t *Ab, int32_t *Bb)
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

    av1pf_lr_unit u;
    u.type = AV1PF_LR_NONE;
    if (g.type != AV1PF_LR_NONE) {
        const int uc = min(x0 / g.us, g.hz - 1);
        const int ur = min((y0 + g.off) / g.us, g.vt - 1);
        u = g.units[ur * g.hz + uc];
    }
    if (u.type != AV1PF_LR_WIENER && u.type != AV1PF_LR_SGRPROJ) {
        // copy: every load of the lane's rows in flight before its stores
        const int lpr = lanes_for(w), c = tid & (lpr - 1), step = 256 / lpr;
        if (c < w)
            for (int r0 = tid / lpr; r0 < h; r0 += 8 * step) {
                T v[8];
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (r0 + j * step < h) v[j] = row_ptr<T>(sp, y0 + r0 + j * step)[x0 + c];
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (r0 + j * step < h) row_ptr<T>(dp, y0 + r0 + j * step)[x0 + c] = v[j];
            }
        return;
    }
    // ---- stage the tile with its 3-sample halo; a wave's rows are uniform,
    // and a lane issues the loads of 6 rows before writing them to LDS
    const int nominal = g.sh - (y0 == 0 ? g.off : 0);
    const bool copy_above = y0 != 0, copy_below = y0 + nominal < g.ph;
    const int tw = w + 6, th = h + 6;
    const auto bnd = gmem(g.bnd) + (int64_t)t * 4 * g.pw;
    for (int cc = tid & 63; cc < tw; cc += 64) {
        const int c = cc - 3;
        const int x = min(max(x0 + c, 0), g.pw - 1);
        for (int rr0 = tid >> 6; rr0 < th; rr0 += 24) {
            int v[6];
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const int rr = rr0 + 4 * j;
                if (rr >= th) break;
                const int r = rr - 3;
                if (r < 0) {
                    if (copy_above) v[j] = bnd[(r == -1 ? 1 : 0) * g.pw + x];
                    else v[j] = read_px<T>(sp, max(y0 + r, 0), x);
                } else if (r >= h) {
                    if (copy_below) v[j] = bnd[(2 + min(r - h, 1)) * g.pw + x];
                    else v[j] = read_px<T>(sp, min(y0 + r, g.ph - 1), x);
                } else {
                    v[j] = read_px<T>(sp, y0 + r, x);
                }
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const int rr = rr0 + 4 * j;
                if (rr >= th) break;
                tile[rr * LR_TW + cc] = (uint16_t)v[j];
            }
        }
    }
    __syncthreads();
    const uint16_t *O = tile + 3 * LR_TW + 3;   // tile origin
    const int8_t *vt = u.wiener_v, *ht = u.wiener_h;
    auto put = [&](int y, int x, int v) { row_ptr<T>(dp, y0 + y)[x0 + x] = (T)v; };
    if (u.type == AV1PF_LR_WIENER) wiener_tile(O, w, h, bd, vt, ht, tmp, put);
    else sgr_tile(O, w, h, bd, u.sgr_set, u.sgr_xqd, Ab, Bb, put);
}

// 4 workgroups per CU: 35.9 KB of LDS each, and at most 128 VGPRs
__global__ __launch_bounds__(256, 4) void lr_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs, int nseg,
                                                 int total, StageLaunch L)
{
    __shared__ uint16_t tile[LR_TW * LR_TW];
    // Wiener intermediate and self-guided A/B planes share one buffer
    __shared__ __align__(16) char work[LR_AB * LR_AB * 6];
    int32_t *Bb = reinterpret_cast<int32_t *>(work);
    uint16_t *Ab = reinterpret_cast<uint16_t *>(work + LR_AB * LR_AB * 4);
    uint16_t *tmp = reinterpret_cast<uint16_t *>(work);
    static_assert(LR_TW * 64 * 2 <= LR_AB * LR_AB * 6, "Wiener intermediate fits the shared buffer");
    for_each_unit(total, L, [&](int u) {
        const Seg sg = segs[find_seg(segs, nseg, u)];
        const FrameDesc &D = descs[sg.frame];
        const int t = u - sg.base;
        const int k = t % sg.nx, st = t / sg.nx;
        const FrameView &src = D.v[D.lr_src];
        const FrameView &dst = D.v[D.lr_dst];
        const int pl = sg.plane;
        if (src.bd == 8)
            lr_tile<uint8_t>(src.p[pl], dst.p[pl], D.lr[pl], 8, k, st, tile, tmp, Ab, Bb);
        else
            lr_tile<uint16_t>(src.p[pl], dst.p[pl], D.lr[pl], src.bd, k, st, tile, tmp, Ab, Bb);
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
    const auto b = gmem(g.bnd) + (int64_t)t * 4 * g.pw;
    auto px = [&](int y) { return sr ? sr_sample<T>(row_ptr<T>(pv, y), *sr, x, maxv) : (int)row_ptr<T>(pv, y)[x]; };
    if (t > 0) {
        b[x] = px(y0 - 2);
        b[g.pw + x] = px(y0 - 1);
    }
    if (y1 < g.ph) {
        b[2 * g.pw + x] = px(y1);
        b[3 * g.pw + x] = px(y1 + 1 < g.ph ? y1 + 1 : y1);
    }
}

__global__ __launch_bounds__(256) void lr_save_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs, int nseg)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int x = (t % sg.nx) * 256 + threadIdx.x, st = t / sg.nx;
    const FrameView &f = D.v[VIEW_A];
    const LrPlaneDesc &g = D.lr[sg.plane];
    const SrPlane *sr = g.sr ? &D.sr[sg.plane] : nullptr;
    const int maxv = (1 << f.bd) - 1;
    if (f.bd == 8) lr_save_cols<uint8_t>(f.p[sg.plane], g, sr, maxv, x, st);
    else lr_save_cols<uint16_t>(f.p[sg.plane], g, sr, maxv, x, st);
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

// Plain copy of one plane's aligned extent: workgroup = 256 samples of a row.
template <typename T>
__device__ __forceinline__ void copy_cols(const PlaneView &s, const PlaneView &d, int x, int y)
{
    if (x >= s.aw || y >= s.ah) return;
    row_ptr<T>(d, y)[x] = row_ptr<T>(s, y)[x];
}

__global__ __launch_bounds__(256) void copy_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs,
                                                   int nseg, int src, int dst)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int x = (t % sg.nx) * 256 + threadIdx.x, y = t / sg.nx;
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
