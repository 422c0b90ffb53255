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
// LDS as uint16 (so the VeryLarge sentinel fits; pitch 82 so that the rows of
// different 8x8 blocks fall in different banks).  The direction search runs on
// all four waves: wave w evaluates directions w and w + 4 for the block of
// each lane, the winner is taken per block from the eight costs in LDS.  The
// per-block filter parameters (tap offsets, adjusted strength, shifts) are
// then packed once into a 16-byte LDS record, and the filter walks the plane
// row by row with one lane per pair of columns (coalesced global stores),
// reading the record of its block and the thirteen taps of both samples from
// LDS, the arithmetic packed two samples to a register.  The kernel is out of
// place (src -> dst): every sample of the plane's MI-aligned extent is
// written, filtered or copied, so neighbours always read pre-CDEF samples
// exactly as the reference's snapshot does.
#include "pf_internal.h"

namespace pf {

#define CDEF_VL 0x4000
#define TPITCH 82   // 64 + 2*8, plus 2 to spread the block rows over the banks
#define TROWS 68    // 64 + 2*2
#define TSHIFT (TROWS * TPITCH + 1)   // odd offset of the tile's second copy (samples)

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

// Line of sample (i, j) of an 8x8 block in direction D, and the cost of the
// direction from its line sums: squared sums weighted by 840 / line length
// (cdef/direction.go:37-58, :432-502).
template <int D>
__device__ __forceinline__ constexpr int dir_line(int i, int j)
{
    return D == 0 ? i + j : D == 1 ? i + (j >> 1) : D == 2 ? i : D == 3 ? 3 + i - (j >> 1)
         : D == 4 ? 7 + i - j : D == 5 ? 3 - (i >> 1) + j : D == 6 ? j : (i >> 1) + j;
}
template <int D>
__device__ __forceinline__ uint32_t dir_finish(const int32_t (&pa)[15])
{
    uint32_t s = 0;
    if constexpr (D == 2 || D == 6) {
#pragma unroll
        for (int k = 0; k < 8; k++) s += (uint32_t)(pa[k] * pa[k]);
        s *= 105u;
    } else if constexpr (D == 0 || D == 4) {
#pragma unroll
        for (int k = 0; k < 7; k++)
            s += (uint32_t)(pa[k] * pa[k] + pa[14 - k] * pa[14 - k]) * (uint32_t)kDivTable[k + 1];
        s += (uint32_t)(pa[7] * pa[7]) * 105u;
    } else {
#pragma unroll
        for (int k = 3; k < 8; k++) s += (uint32_t)(pa[k] * pa[k]);
        s *= 105u;
#pragma unroll
        for (int k = 0; k < 3; k++)
            s += (uint32_t)(pa[k] * pa[k] + pa[10 - k] * pa[10 - k]) * (uint32_t)kDivTable[2 * k + 2];
    }
    return s;
}

// Costs of directions D0 and D1 for the 8x8 block at img: the block is read
// row by row (samples (x >> cs) - 128) into both directions' line sums, so
// only a row of samples and the two sets of sums are live.
template <int D0, int D1>
__device__ __forceinline__ void dir_cost2(const uint16_t *img, int cs, uint32_t &c0, uint32_t &c1)
{
    int32_t p0[15], p1[15];
#pragma unroll
    for (int k = 0; k < 15; k++) { p0[k] = 0; p1[k] = 0; }
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int32_t x = (int32_t)(img[i * TPITCH + j] >> cs) - 128;
            p0[dir_line<D0>(i, j)] += x;
            p1[dir_line<D1>(i, j)] += x;
        }
        __builtin_amdgcn_sched_barrier(0);   // one row's reads in flight, not the block's 64
    }
    c0 = dir_finish<D0>(p0);
    c1 = dir_finish<D1>(p1);
}

// A lane's loads of up to 8 rows are in flight before its stores.
template <typename T>
__device__ __forceinline__ void copy_unit(const PlaneView &s, const PlaneView &d, int ux, int uy, int uw, int uh, int tid)
{
    const int tpr = uw > 32 ? 64 : uw > 16 ? 32 : uw > 8 ? 16 : 8;
    const int x = tid & (tpr - 1), step = 256 / tpr;
    if (x >= uw) return;
    for (int y0 = tid / tpr; y0 < uh; y0 += 8 * step) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = row_ptr<T>(s, uy + min(y0 + j * step, uh - 1))[ux + x];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (y0 + j * step < uh) row_ptr<T>(d, uy + y0 + j * step)[ux + x] = v[j];
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
    for (int c = tid & 63; c < cols; c += 64) {
        const int x = ux - 8 + c;
        const bool vx = x >= 0 && (x < sx1 || col_ext);
        const int cx = min(x, sx1 - 1);
        // a lane's loads of 6 rows are in flight before its LDS writes
        for (int r0 = tid >> 6; r0 < rows; r0 += 24) {
            // loads from clamped (always valid) positions, unconditionally,
            // so that none waits at a branch join; VeryLarge chosen after
            uint16_t v[6];
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const int y = uy - 2 + r0 + 4 * j;
                v[j] = row_ptr<T>(s, min(max(y, 0), sy1 - 1))[min(max(cx, 0), sx1 - 1)];
            }
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const int r = r0 + 4 * j;
                const int y = uy - 2 + r;
                const bool vy = y >= 0 && (y < sy1 || row_ext);
                if (r < rows) {
                    const uint16_t val = (vy && vx) ? v[j] : (uint16_t)CDEF_VL;
                    tile[r * pitch + c] = val;
                    tile[TSHIFT + r * pitch + c] = val;   // the copy at an odd offset
                }
            }
        }
    }
}

__device__ __forceinline__ int str_pri(uint8_t p) { return p >> 2; }
__device__ __forceinline__ int str_sec(uint8_t p) { const int s = p & 3; return s == 3 ? 4 : s; }

// Two samples in the halves of a register: packed 16-bit arithmetic.
typedef short v2s __attribute__((ext_vector_type(2)));
typedef unsigned short v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2s splat2(int v) { return (v2s){ (short)v, (short)v }; }
__device__ __forceinline__ v2u splatu2(int v) { return (v2u){ (unsigned short)v, (unsigned short)v }; }
__device__ __forceinline__ v2s vmax(v2s a, v2s b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ v2s vmin(v2s a, v2s b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ v2s as2(uint32_t v) { return __builtin_bit_cast(v2s, v); }
// constrain(d, t, sh) = sign(d) min(|d|, max(0, t - (|d| >> sh))) of two
// samples, written as the clamp of d to [-c, c] with c = max(0, t - (|d| >> sh)).
__device__ __forceinline__ v2s cst2(v2s a, v2s x, v2u t, v2u sh)
{
    const v2s d = a - x;
    const v2u ad = __builtin_bit_cast(v2u, vmax(d, -d));
    const v2s c = __builtin_bit_cast(v2s, __builtin_elementwise_sub_sat(t, ad >> sh));
    return vmin(vmax(d, -c), c);
}

// Per-block filter record: for the 12 taps (+-primary 0, +-primary 1, +-dir+2
// secondary 0, +-dir-2 secondary 0, +-dir+2 secondary 1, +-dir-2 secondary 1)
// the 32-bit-word offset of the tap pair from the pair at the filtered
// samples, then strength, primary shift, secondary shift and the filter flag
// (1: filter, 2: clip to the taps' range).  The filtered pair starts at an even
// sample; a tap at an even sample offset o reads the tile at o, one at an odd
// offset reads the copy TSHIFT (odd) samples further on, so both are 4-byte
// aligned single reads.
struct __align__(16) BlockRec { int16_t e[12]; int p; int pad[1]; };
__device__ __forceinline__ int tap_word(int o) { return ((o & 1) ? TSHIFT + o : o) >> 1; }

template <typename T>
__device__ __forceinline__ void cdef_unit(const FrameView &src, const FrameView &dst, const CdefDesc &a, int *err,
                                          int uc, int ur, uint16_t *tile, uint8_t *bdir, int32_t *bvar,
                                          uint32_t (*dcost)[8], BlockRec *brec)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ui = ur * a.unit_cols + uc;
    const uint64_t m = a.noskip ? gmem(a.noskip)[ui] : ~0ull;
    int idx, count, yfilt, uvfilt;
    gptr<const uint8_t> ystr, uvstr;
    if (a.dec) {
        // encoder apply (cdef_apply.go:202-228, the noskip path): a unit is
        // read when one of its 8x8 blocks inside the frame has residual
        const auto dec = gmem(a.dec);
        const int bits = dec->cdef_bits;
        // cdef_bits is 0..3 (8 strengths); anything else is flagged and read as 3
        if (bits < 0 || bits > 3) {
            if (tid == 0) atomicOr(err, 1);
        }
        count = 1 << min(max(bits, 0), 3);
        ystr = dec->y_strength;
        uvstr = dec->uv_strength;
        yfilt = uvfilt = 0;
        for (int k = 0; k < count; k++) {
            yfilt |= ystr[k] != 0;
            uvfilt |= src.nplanes > 1 && uvstr[k] != 0;
        }
        const int rb = min(8, a.rows8 - ur * 8), cb = min(8, a.cols8 - uc * 8);
        uint64_t in = 0;
        for (int by = 0; by < rb; by++) in |= ((1ull << cb) - 1) << (by * 8);
        idx = (m & in) ? gmem(a.fb_index)[ui] : -1;
    } else {
        count = a.strength_count;
        ystr = gmem(a.y_strength);
        uvstr = gmem(a.uv_strength);
        yfilt = a.yfilt;
        uvfilt = a.uvfilt;
        idx = gmem(a.index)[ui];
    }
    if (idx >= count) {
        if (tid == 0) atomicOr(err, 1);
        idx = -1;
    }
    const int cs = src.bd - 8;
    const bool chroma_filtering = uvfilt != 0;

    for (int pl = 0; pl < src.nplanes; pl++) {
        const int xd = pl ? src.ssx : 0, yd = pl ? src.ssy : 0;
        const PlaneView &sp = src.p[pl];
        const PlaneView &dp = dst.p[pl];
        const int ux = (uc * 64) >> xd, uy = (ur * 64) >> yd;
        const int uw = min(64 >> xd, sp.aw - ux), uh = min(64 >> yd, sp.ah - uy);
        if (uw <= 0 || uh <= 0) continue;
        const bool process = pl == 0 ? (yfilt || chroma_filtering) : uvfilt;
        bool active = false, dir_only = false, need_dir = false;
        uint8_t packed = 0;
        if (process && idx >= 0) {
            packed = pl ? uvstr[idx] : ystr[idx];
            const int uvpri = str_pri(uvstr[idx]);
            need_dir = pl == 0 ? (str_pri(ystr[idx]) != 0 || (chroma_filtering && uvpri != 0)) : uvpri != 0;
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
        if (pl == 0 && need_dir) {
            // wave w: directions w and w + 4 of the block of this lane
            if ((lm >> lane) & 1) {
                const uint16_t *img = org + (lane >> 3) * 8 * TPITCH + (lane & 7) * 8;
                uint32_t c0, c1;
                switch (wv) {
                case 0: dir_cost2<0, 4>(img, cs, c0, c1); break;
                case 1: dir_cost2<1, 5>(img, cs, c0, c1); break;
                case 2: dir_cost2<2, 6>(img, cs, c0, c1); break;
                default: dir_cost2<3, 7>(img, cs, c0, c1); break;
                }
                dcost[lane][wv] = c0;
                dcost[lane][wv + 4] = c1;
            }
            __syncthreads();
            if (tid < 64 && ((lm >> tid) & 1)) {   // first maximum, as a signed comparison
                int best = 0;
                int32_t bc = (int32_t)dcost[tid][0];
#pragma unroll
                for (int d = 1; d < 8; d++)
                    if ((int32_t)dcost[tid][d] > bc) { bc = (int32_t)dcost[tid][d]; best = d; }
                bdir[tid] = (uint8_t)best;
                bvar[tid] = (int32_t)((uint32_t)bc - dcost[tid][(best + 4) & 7]) >> 10;
            }
        }
        __syncthreads();
        if (tid < 64) {
            const bool listed = (lm >> tid) & 1;
            int dir = 0;
            if (need_dir && listed) {
                dir = bdir[tid];
                if (pl == 1 && xd != yd) {
                    const uint8_t c422[8] = { 7, 0, 2, 4, 5, 6, 6, 6 };
                    const uint8_t c440[8] = { 1, 2, 2, 2, 3, 4, 6, 0 };
                    dir = xd ? c422[dir] : c440[dir];
                    bdir[tid] = (uint8_t)dir;
                }
            }
            int st = 0;
            if (listed) {
                const bool luma_adjust = pl == 0 && (need_dir || pri != 0);
                st = luma_adjust ? adjust_strength(pri, bvar[tid]) : pri;
            }
            const int fd = pri ? dir : 0;
            const int psh = st ? max(damping - ilog2_((unsigned)st), 0) : 0;
            const int ssh = sec ? max(damping - ilog2_((unsigned)sec), 0) : 0;
            const int flag = (listed && (st != 0 || sec != 0) ? 1 : 0) | (st && sec ? 2 : 0);
            BlockRec r;
            const int tap[6] = { kDirOff[fd + 2][0], kDirOff[fd + 2][1], kDirOff[fd + 4][0], kDirOff[fd][0],
                                 kDirOff[fd + 4][1], kDirOff[fd][1] };
#pragma unroll
            for (int k = 0; k < 6; k++) {
                r.e[2 * k] = (int16_t)tap_word(tap[k]);
                r.e[2 * k + 1] = (int16_t)tap_word(-tap[k]);
            }
            r.p = min(st, 255) | (psh << 8) | (ssh << 16) | (flag << 24);
            r.pad[0] = 0;
            brec[tid] = r;
        }
        __syncthreads();
        if (dir_only) {
            copy_unit<T>(sp, dp, ux, uy, uw, uh, tid);
            continue;
        }
        // the filter walks horizontal sample pairs, one pair per lane: both
        // samples of a pair lie in one block (blocks are at least 4 wide), so
        // they share the record, and every tap is one 32-bit LDS read and
        // packed 16-bit arithmetic (values stay below 2^15: samples < 2^12,
        // the VeryLarge sentinel 2^14, the weighted sum below 2^12)
        const int np = (uw + 1) >> 1;
        const int tpr = np > 16 ? 32 : np > 8 ? 16 : np > 4 ? 8 : 4;   // lanes per row, a power of two >= np
        const int x = 2 * (tid & (tpr - 1));
        if (x < uw) {
            for (int y = tid / tpr; y < uh; y += 256 / tpr) {
                const BlockRec r = brec[((y >> bhl) << 3) + (x >> bwl)];
                const uint32_t *W = reinterpret_cast<const uint32_t *>(tile) + ((org - tile + y * TPITCH + x) >> 1);
                const v2s xv = as2(W[0]);
                v2s out = xv;
                const int flag = r.p >> 24;
                if (flag & 1) {
                    const int st = r.p & 255, psh = (r.p >> 8) & 255, ssh = (r.p >> 16) & 255;
                    const v2s a0 = as2(W[r.e[0]]), a1 = as2(W[r.e[1]]), a2 = as2(W[r.e[2]]), a3 = as2(W[r.e[3]]);
                    const v2s b0 = as2(W[r.e[4]]), b1 = as2(W[r.e[5]]), b2 = as2(W[r.e[6]]), b3 = as2(W[r.e[7]]);
                    const v2s b4 = as2(W[r.e[8]]), b5 = as2(W[r.e[9]]), b6 = as2(W[r.e[10]]), b7 = as2(W[r.e[11]]);
                    v2s sum = splat2(0);
                    if (pri) {   // st may be 0 for a flat block: its primary terms vanish
                        const v2u t = splatu2(st), sh = splatu2(psh);
                        const v2s t0 = splat2(((st >> cs) & 1) ? 3 : 4), t1 = splat2(((st >> cs) & 1) ? 3 : 2);
                        sum = t0 * (cst2(a0, xv, t, sh) + cst2(a1, xv, t, sh)) + t1 * (cst2(a2, xv, t, sh) + cst2(a3, xv, t, sh));
                    }
                    if (sec) {
                        const v2u t = splatu2(sec), sh = splatu2(ssh);
                        sum += splat2(2) * (cst2(b0, xv, t, sh) + cst2(b1, xv, t, sh) + cst2(b2, xv, t, sh) +
                                            cst2(b3, xv, t, sh)) +
                               cst2(b4, xv, t, sh) + cst2(b5, xv, t, sh) + cst2(b6, xv, t, sh) + cst2(b7, xv, t, sh);
                    }
                    // (8 + sum - (sum < 0)) >> 4; a true vector comparison is -1
                    out = xv + ((splat2(8) + sum + (v2s)(sum < splat2(0))) >> splat2(4));
                    if (flag & 2) {   // clip to the taps' range, VeryLarge excluded from the maximum
                        const v2s mk = splat2(CDEF_VL - 1);
                        v2s hi = vmax(vmax(vmax(xv, a0 & mk), vmax(a1 & mk, a2 & mk)), vmax(a3 & mk, b0 & mk));
                        hi = vmax(vmax(vmax(hi, b1 & mk), vmax(b2 & mk, b3 & mk)), vmax(vmax(b4 & mk, b5 & mk), vmax(b6 & mk, b7 & mk)));
                        v2s lo = vmin(vmin(vmin(xv, a0), vmin(a1, a2)), vmin(a3, b0));
                        lo = vmin(vmin(vmin(lo, b1), vmin(b2, b3)), vmin(vmin(b4, b5), vmin(b6, b7)));
                        out = vmin(vmax(out, lo), hi);
                    }
                }
                const auto o = gmem(row_ptr<T>(dp, uy + y) + ux + x);
                o[0] = (T)(uint16_t)out.x;
                if (x + 1 < uw) o[1] = (T)(uint16_t)out.y;
            }
        }
    }
}

// 6 workgroups per CU: 26.7 KB of LDS each, and at most 80 VGPRs
__global__ __launch_bounds__(256, 6) void cdef_kernel(const FrameDesc *__restrict__ descs,
                                                   const Seg *__restrict__ segs, int nseg, int total, StageLaunch L)
{
    __shared__ __align__(16) uint16_t tile[TSHIFT + TROWS * TPITCH + 1];   // the tile, then its copy at TSHIFT
    __shared__ uint8_t bdir[64];
    __shared__ int32_t bvar[64];
    __shared__ uint32_t dcost[64][8];   // direction costs per block
    __shared__ BlockRec brec[64];
    for_each_unit(total, L, [&](int u) {
        const Seg sg = segs[find_seg(segs, nseg, u)];
        const FrameDesc &D = descs[sg.frame];
        const int t = u - sg.base;
        const int uc = t % sg.nx, ur = t / sg.nx;
        if (D.v[VIEW_A].bd == 8)
            cdef_unit<uint8_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, tile, bdir, bvar, dcost, brec);
        else
            cdef_unit<uint16_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, tile, bdir, bvar, dcost, brec);
    });
}

hipError_t launch_cdef(const FrameDesc *d, const Seg *segs, int nseg, int total, const StageLaunch &L, hipStream_t s)
{
    if (total)
        hipLaunchKernelGGL(cdef_kernel, dim3(stage_grid((const void *)cdef_kernel, total, L.per_wg)), dim3(256), 0, s,
                           d, segs, nseg, total, L);
    return hipGetLastError();
}

}  // namespace pf
