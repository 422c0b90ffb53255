// CDEF on gfx950: one workgroup per 64x64 unit, all planes of the unit in the
// same workgroup so chroma reuses the luma directions straight from LDS.
//
// Semantics follow goav1's snapshot walk (decoder/postfilter_cdef.go:525-1031):
//   unit strength / plane activity / direction-only luma units (:936-958, :1041),
//   the 144-stride input with VeryLarge sentinels and last-row/column replication
//   for non-final units (:1101-1250), the MI-aligned plane extent (:1390-1430),
//   8x8 direction search (cdef/direction.go:37-58,432-502), variance-adjusted
//   luma primary strength (cdef/filter.go:603-612), 4:2:2 / 4:4:0 direction
//   remap (:614-636) and the constrained primary/secondary filter with the
//   VeryLarge-masked clip (cdef/filter.go:79-188, filter_u8.go:124-134).
//
// Layout: the unit plus its 2-row / 8-column halo is staged once per plane in
// LDS as uint16 (so the VeryLarge sentinel fits).  The direction search runs one
// lane per 8x8 block with the 8x15 partial sums in registers; the filter runs
// 16 (luma) or 4..16 (chroma) pixels per lane with all thirteen taps read from
// LDS.  The kernel is out of place (src -> dst): every sample of the plane's
// MI-aligned extent is written, filtered or copied, so neighbours always read
// pre-CDEF samples exactly as the reference's snapshot does.
#include "pf_internal.h"

namespace pf {

#define CDEF_VL 0x4000
#define TPITCH 80   // 64 + 2*8
#define TROWS 68    // 64 + 2*2

__device__ __constant__ int kDivTable[9] = { 0, 840, 420, 280, 210, 168, 140, 120, 105 };

// direction offsets (cdef/filter.go:53-66) as (dy*TPITCH + dx) for the LDS tile
__device__ __constant__ int kDirOff[12][2] = {
    { 1 * TPITCH, 2 * TPITCH },         { 1 * TPITCH, 2 * TPITCH - 1 },
    { -1 * TPITCH + 1, -2 * TPITCH + 2 }, { 1, -1 * TPITCH + 2 },
    { 1, 2 },                           { 1, 1 * TPITCH + 2 },
    { 1 * TPITCH + 1, 2 * TPITCH + 2 }, { 1 * TPITCH, 2 * TPITCH + 1 },
    { 1 * TPITCH, 2 * TPITCH },         { 1 * TPITCH, 2 * TPITCH - 1 },
    { -1 * TPITCH + 1, -2 * TPITCH + 2 }, { 1, -1 * TPITCH + 2 },
};

__device__ __forceinline__ int constrain(int diff, int t, int sh)
{
    const int a = diff < 0 ? -diff : diff;
    const int l = min(max(t - (a >> sh), 0), a);
    return diff < 0 ? -l : l;
}

__device__ __forceinline__ int ilog2_(unsigned v) { return 31 - __clz(v); }

__device__ __forceinline__ int adjust_strength(int strength, int32_t var)
{
    if (var == 0) return 0;
    int i = 0;
    const int32_t v = var >> 6;
    if (v != 0) i = min(ilog2_((unsigned)v), 12);
    return (strength * (4 + i) + 8) >> 4;
}

// 8x8 direction search on an LDS tile (cdef/direction.go:37-58, :432-502).
__device__ void find_dir(const uint16_t *img, int pitch, int cs, int *dir, int32_t *var)
{
    int32_t pa[8][15];
#pragma unroll
    for (int d = 0; d < 8; d++)
#pragma unroll
        for (int k = 0; k < 15; k++) pa[d][k] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int32_t x = (int32_t)(img[i * pitch + j] >> cs) - 128;
            pa[0][i + j] += x;
            pa[1][i + (j >> 1)] += x;
            pa[2][i] += x;
            pa[3][3 + i - (j >> 1)] += x;
            pa[4][7 + i - j] += x;
            pa[5][3 - (i >> 1) + j] += x;
            pa[6][j] += x;
            pa[7][(i >> 1) + j] += x;
        }
    }
    uint32_t c[8];
#pragma unroll
    for (int d = 0; d < 8; d += 2) {
        // d even: 0,4 diagonal; 2,6 straight
        uint32_t s = 0;
        if (d == 2 || d == 6) {
#pragma unroll
            for (int k = 0; k < 8; k++) s += (uint32_t)(pa[d][k] * pa[d][k]);
            s *= 105u;
        } else {
#pragma unroll
            for (int k = 0; k < 7; k++)
                s += (uint32_t)(pa[d][k] * pa[d][k] + pa[d][14 - k] * pa[d][14 - k]) * (uint32_t)kDivTable[k + 1];
            s += (uint32_t)(pa[d][7] * pa[d][7]) * 105u;
        }
        c[d] = s;
    }
#pragma unroll
    for (int d = 1; d < 8; d += 2) {
        uint32_t s = 0;
#pragma unroll
        for (int k = 3; k < 8; k++) s += (uint32_t)(pa[d][k] * pa[d][k]);
        s *= 105u;
#pragma unroll
        for (int k = 0; k < 3; k++)
            s += (uint32_t)(pa[d][k] * pa[d][k] + pa[d][10 - k] * pa[d][10 - k]) * (uint32_t)kDivTable[2 * k + 2];
        c[d] = s;
    }
    int best = 0;
    int32_t bc = (int32_t)c[0];
#pragma unroll
    for (int d = 1; d < 8; d++)
        if ((int32_t)c[d] > bc) { bc = (int32_t)c[d]; best = d; }
    *dir = best;
    *var = (int32_t)((uint32_t)bc - c[(best + 4) & 7]) >> 10;
}

template <typename T>
__device__ __forceinline__ void copy_unit(const PlaneView &s, const PlaneView &d, int ux, int uy, int uw, int uh, int tid)
{
    const int n = uw * uh;
    for (int i = tid; i < n; i += 256) {
        const int y = i / uw, x = i - y * uw;
        row_ptr<T>(d, uy + y)[ux + x] = row_ptr<T>(s, uy + y)[ux + x];
    }
}

// Stage the unit input (postfilter_cdef.go:1101-1250) into an LDS tile with
// origin at tile[2*pitch + 8].
template <typename T>
__device__ __forceinline__ void load_tile(uint16_t *tile, int pitch, const PlaneView &s, int ux, int uy, int uw, int uh, int tid)
{
    const int W = s.aw, H = s.ah;
    const int sy1 = min(uy + uh + 2, H), sx1 = min(ux + uw + 8, W);
    const bool row_ext = sy1 < uy + uh + 2 && sy1 > 0 && uy + uh < H;
    const bool col_ext = sx1 < ux + uw + 8 && sx1 > 0 && ux + uw < W;
    const int rows = uh + 4, cols = uw + 16;
    for (int i = tid; i < rows * cols; i += 256) {
        const int r = i / cols, c = i - r * cols;
        const int y = uy - 2 + r, x = ux - 8 + c;
        const bool vy = y >= 0 && (y < sy1 || row_ext);
        const bool vx = x >= 0 && (x < sx1 || col_ext);
        uint16_t v = CDEF_VL;
        if (vy && vx) v = row_ptr<T>(s, min(y, sy1 - 1))[min(x, sx1 - 1)];
        tile[r * pitch + c] = v;
    }
}

__device__ __forceinline__ int str_pri(uint8_t p) { return p >> 2; }
__device__ __forceinline__ int str_sec(uint8_t p) { const int s = p & 3; return s == 3 ? 4 : s; }

template <typename T>
__device__ __forceinline__ void cdef_unit(const FrameView &src, const FrameView &dst, const CdefDesc &a, int *err,
                                          int uc, int ur, uint16_t *tile, uint8_t *bdir, int32_t *bvar, uint8_t *bst)
{
    const int tid = threadIdx.x;
    const int ui = ur * a.unit_cols + uc;
    int idx = a.index[ui];
    if (idx >= a.strength_count) {
        if (tid == 0) atomicOr(err, 1);
        idx = -1;
    }
    const uint64_t m = a.noskip ? a.noskip[ui] : ~0ull;
    const int cs = src.bd - 8;
    const bool chroma_filtering = a.uvfilt != 0;

    for (int pl = 0; pl < src.nplanes; pl++) {
        const int xd = pl ? src.ssx : 0, yd = pl ? src.ssy : 0;
        const PlaneView &sp = src.p[pl];
        const PlaneView &dp = dst.p[pl];
        const int ux = (uc * 64) >> xd, uy = (ur * 64) >> yd;
        const int uw = min(64 >> xd, sp.aw - ux), uh = min(64 >> yd, sp.ah - uy);
        if (uw <= 0 || uh <= 0) continue;
        const bool process = pl == 0 ? (a.yfilt || chroma_filtering) : a.uvfilt;
        bool active = false, dir_only = false, need_dir = false;
        uint8_t packed = 0;
        if (process && idx >= 0) {
            packed = pl ? a.uv_strength[idx] : a.y_strength[idx];
            const int uvpri = str_pri(a.uv_strength[idx]);
            need_dir = pl == 0 ? (str_pri(a.y_strength[idx]) != 0 || (chroma_filtering && uvpri != 0)) : uvpri != 0;
            dir_only = pl == 0 && chroma_filtering && packed == 0 && uvpri != 0;
            active = packed != 0 || dir_only;
        }
        const int bw = 8 >> xd, bh = 8 >> yd, bwl = 3 - xd, bhl = 3 - yd;
        const int bcols = (uw + bw - 1) / bw, brows = (uh + bh - 1) / bh;
        uint64_t lm = 0;
        for (int by = 0; by < brows; by++) lm |= ((m >> (by * 8)) & ((1ull << bcols) - 1)) << (by * 8);
        if (!active || lm == 0) {
            copy_unit<T>(sp, dp, ux, uy, uw, uh, tid);
            continue;
        }
        __syncthreads();   // previous plane finished with the tile / block tables
        load_tile<T>(tile, TPITCH, sp, ux, uy, uw, uh, tid);
        __syncthreads();
        const uint16_t *org = tile + 2 * TPITCH + 8;
        const int pri = str_pri(packed) << cs;
        const int sec = str_sec(packed) << cs;
        const int damping = a.damping + cs - (pl ? 1 : 0);
        if (tid < 64) {
            const int by = tid >> 3, bx = tid & 7;
            const bool listed = (lm >> tid) & 1;
            if (pl == 0 && need_dir && listed) {
                int d;
                int32_t v;
                find_dir(org + (by << bhl) * TPITCH + (bx << bwl), TPITCH, cs, &d, &v);
                bdir[tid] = (uint8_t)d;
                bvar[tid] = v;
            }
        }
        __syncthreads();
        if (tid < 64) {
            const bool listed = (lm >> tid) & 1;
            int st = 0xff;
            if (listed) {
                const bool luma_adjust = pl == 0 && (need_dir || pri != 0);
                st = luma_adjust ? adjust_strength(pri, bvar[tid]) : pri;
            }
            bst[tid] = (uint8_t)min(st, 0xff);
            if (pl == 1 && need_dir && listed && xd != yd) {
                const uint8_t c422[8] = { 7, 0, 2, 4, 5, 6, 6, 6 };
                const uint8_t c440[8] = { 1, 2, 2, 2, 3, 4, 6, 0 };
                bdir[tid] = xd ? c422[bdir[tid]] : c440[bdir[tid]];
            }
        }
        __syncthreads();
        if (dir_only) {
            copy_unit<T>(sp, dp, ux, uy, uw, uh, tid);
            continue;
        }
        const int psh_base = damping;
        const int n = uw * uh;
        for (int i = tid; i < n; i += 256) {
            const int y = i / uw, x = i - y * uw;
            const int b = ((y >> bhl) << 3) + (x >> bwl);
            const uint16_t *c = org + y * TPITCH + x;
            const int xv = c[0];
            int out = xv;
            const int st = bst[b];
            if (st != 0xff && (st != 0 || sec != 0)) {
                const int dir = pri ? bdir[b] : 0;
                const int psh = st ? max(psh_base - ilog2_((unsigned)st), 0) : 0;
                const int ssh = sec ? max(psh_base - ilog2_((unsigned)sec), 0) : 0;
                const int t0 = ((st >> cs) & 1) ? 3 : 4, t1 = ((st >> cs) & 1) ? 3 : 2;
                const int o0 = kDirOff[dir + 2][0], o1 = kDirOff[dir + 2][1];
                const int s0a = kDirOff[dir + 4][0], s0b = kDirOff[dir + 4][1];
                const int s1a = kDirOff[dir][0], s1b = kDirOff[dir][1];
                int sum = 0;
                if (st && sec) {
                    const int a0 = c[o0], a1 = c[-o0], a2 = c[o1], a3 = c[-o1];
                    const int b0 = c[s0a], b1 = c[-s0a], b2 = c[s1a], b3 = c[-s1a];
                    const int b4 = c[s0b], b5 = c[-s0b], b6 = c[s1b], b7 = c[-s1b];
                    sum = t0 * (constrain(a0 - xv, st, psh) + constrain(a1 - xv, st, psh)) +
                          t1 * (constrain(a2 - xv, st, psh) + constrain(a3 - xv, st, psh));
                    sum += 2 * (constrain(b0 - xv, sec, ssh) + constrain(b1 - xv, sec, ssh) +
                                constrain(b2 - xv, sec, ssh) + constrain(b3 - xv, sec, ssh));
                    sum += constrain(b4 - xv, sec, ssh) + constrain(b5 - xv, sec, ssh) +
                           constrain(b6 - xv, sec, ssh) + constrain(b7 - xv, sec, ssh);
                    const int mk = CDEF_VL - 1;
                    int hi = max(max(max(xv, a0 & mk), max(a1 & mk, a2 & mk)), max(a3 & mk, b0 & mk));
                    hi = max(max(max(hi, b1 & mk), max(b2 & mk, b3 & mk)), max(max(b4 & mk, b5 & mk), max(b6 & mk, b7 & mk)));
                    int lo = min(min(min(xv, a0), min(a1, a2)), min(a3, b0));
                    lo = min(min(min(lo, b1), min(b2, b3)), min(min(b4, b5), min(b6, b7)));
                    out = xv + ((8 + sum - (sum < 0)) >> 4);
                    out = min(max(out, lo), hi);
                } else if (st) {
                    sum = t0 * (constrain(c[o0] - xv, st, psh) + constrain(c[-o0] - xv, st, psh)) +
                          t1 * (constrain(c[o1] - xv, st, psh) + constrain(c[-o1] - xv, st, psh));
                    out = xv + ((8 + sum - (sum < 0)) >> 4);
                } else {
                    sum = 2 * (constrain(c[s0a] - xv, sec, ssh) + constrain(c[-s0a] - xv, sec, ssh) +
                               constrain(c[s1a] - xv, sec, ssh) + constrain(c[-s1a] - xv, sec, ssh)) +
                          constrain(c[s0b] - xv, sec, ssh) + constrain(c[-s0b] - xv, sec, ssh) +
                          constrain(c[s1b] - xv, sec, ssh) + constrain(c[-s1b] - xv, sec, ssh);
                    out = xv + ((8 + sum - (sum < 0)) >> 4);
                }
            }
            row_ptr<T>(dp, uy + y)[ux + x] = (T)out;
        }
    }
}

__global__ __launch_bounds__(256) void cdef_kernel(const FrameDesc *__restrict__ descs,
                                                   const Seg *__restrict__ segs, int nseg)
{
    __shared__ uint16_t tile[TROWS * TPITCH];
    __shared__ uint8_t bdir[64];
    __shared__ int32_t bvar[64];
    __shared__ uint8_t bst[64];    // per-block primary strength (after luma adjust), 0xff = not listed
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int uc = t % sg.nx, ur = t / sg.nx;
    if (D.v[VIEW_A].bd == 8)
        cdef_unit<uint8_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, tile, bdir, bvar, bst);
    else
        cdef_unit<uint16_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, tile, bdir, bvar, bst);
}

hipError_t launch_cdef(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(cdef_kernel, dim3(total), dim3(256), 0, s, d, segs, nseg);
    return hipGetLastError();
}

}  // namespace pf
