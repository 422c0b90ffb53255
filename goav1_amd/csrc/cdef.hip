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

// Outside-the-frame sentinel of the LDS tile.  goav1 uses VeryLarge = 0x4000
// (cdef/constants.go:17); here 0xC000 stands in for it: a tap at the
// sentinel still differs from every sample by more than 2^(damping + 1), so
// its constrained term is 0 as with 0x4000, and it is the smallest int16 and
// the largest uint16, so it drops out of the signed maximum and the unsigned
// minimum of the clip without masking.
#define CDEF_VL 0xC000
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
// origin at tile[2*TPITCH + 8], in two steps so that the global loads of a
// plane can be issued early and all be in flight at once: tile_fetch loads a
// lane's share -- up to TCH chunks of 4 horizontally adjacent samples, one
// 4-sample vector load per chunk that lies inside the plane (clamped single
// loads for chunks at the plane's edges) -- into registers; tile_store
// replaces the samples outside the frame by the sentinel (or replicates the
// last row / column for a unit that is not the last one), and writes the
// tile and its copy at the odd offset TSHIFT.
constexpr int TCH = 6;   // chunks per lane: (64 + 4) rows x (64 + 16) / 4 chunks <= 6 x 256
struct TileGeo {
    int rows, nch, sy1, sx1, ux, uy;
    bool row_ext, col_ext;
};
__device__ __forceinline__ TileGeo tile_geo(const PlaneView &s, int ux, int uy, int uw, int uh)
{
    TileGeo g;
    const int W = s.aw, H = s.ah;
    g.sy1 = min(uy + uh + 2, H);
    g.sx1 = min(ux + uw + 8, W);
    g.row_ext = g.sy1 < uy + uh + 2 && g.sy1 > 0 && uy + uh < H;
    g.col_ext = g.sx1 < ux + uw + 8 && g.sx1 > 0 && ux + uw < W;
    g.rows = uh + 4;
    g.nch = (uw + 16) >> 2;   // uw is a multiple of 4 (MI-aligned extent)
    g.ux = ux;
    g.uy = uy;
    return g;
}
template <typename T>
struct TileRegs {
    uint32_t v[TCH][sizeof(T)];   // 4 samples: 2 words (16-bit) or 1 word (8-bit)
};
template <typename T>
__device__ __forceinline__ void tile_fetch(TileRegs<T> &R, const PlaneView &s, const TileGeo &g, int tid)
{
    const bool vec = s.al >= 4 * (int)sizeof(T);
    // chunk i = tid + 256 j is (row r, chunk k) of the tile: one division, then steps
    int r = tid / g.nch, k = tid - r * g.nch;
    const int dr = 256 / g.nch, dk = 256 - dr * g.nch;
#pragma unroll
    for (int j = 0; j < TCH; j++) {
        if (j) {
            r += dr;
            k += dk;
            if (k >= g.nch) { k -= g.nch; r++; }
        }
        if (r < g.rows) {
            const int y = min(max(g.uy - 2 + r, 0), g.sy1 - 1);
            const int x = g.ux - 8 + 4 * k;
            const T *row = row_ptr<T>(s, y);
            if (vec && x >= 0 && x + 3 < g.sx1) {
                if (sizeof(T) == 2) {
                    const uint2 w = *(const uint2 *)(row + x);
                    R.v[j][0] = w.x;
                    R.v[j][sizeof(T) - 1] = w.y;
                } else {
                    R.v[j][0] = *(const uint32_t *)(row + x);
                }
            } else {
                uint32_t e[4];
#pragma unroll
                for (int q = 0; q < 4; q++) e[q] = row[min(max(x + q, 0), g.sx1 - 1)];
                if (sizeof(T) == 2) {
                    R.v[j][0] = e[0] | (e[1] << 16);
                    R.v[j][sizeof(T) - 1] = e[2] | (e[3] << 16);
                } else {
                    R.v[j][0] = e[0] | (e[1] << 8) | (e[2] << 16) | (e[3] << 24);
                }
            }
        }
    }
}
template <typename T>
__device__ __forceinline__ void tile_store(uint16_t *tile, const TileRegs<T> &R, const TileGeo &g, int tid)
{
    // a unit whose whole input lies inside the plane stores its samples as
    // loaded: no sentinel, no replication
    const bool inner = g.ux >= 8 && g.uy >= 2 && g.sx1 == g.ux + 4 * g.nch - 8 && g.sy1 == g.uy + g.rows - 2;
    int r = tid / g.nch, k = tid - r * g.nch;
    const int dr = 256 / g.nch, dk = 256 - dr * g.nch;
#pragma unroll
    for (int j = 0; j < TCH; j++) {
        if (j) {
            r += dr;
            k += dk;
            if (k >= g.nch) { k -= g.nch; r++; }
        }
        if (r < g.rows) {
            uint32_t *t = reinterpret_cast<uint32_t *>(tile + r * TPITCH + 4 * k);
            uint16_t *c = tile + TSHIFT + r * TPITCH + 4 * k;
            if (inner) {
                uint32_t w0, w1;
                if (sizeof(T) == 2) {
                    w0 = R.v[j][0];
                    w1 = R.v[j][sizeof(T) - 1];
                } else {   // bytes 0..3 to half words
                    w0 = __builtin_amdgcn_perm(0u, R.v[j][0], 0x0c010c00u);
                    w1 = __builtin_amdgcn_perm(0u, R.v[j][0], 0x0c030c02u);
                }
                t[0] = w0;
                t[1] = w1;
                c[0] = (uint16_t)w0;
                *reinterpret_cast<uint32_t *>(c + 1) = __builtin_amdgcn_alignbit(w1, w0, 16);
                c[3] = (uint16_t)(w1 >> 16);
                continue;
            }
            const int y = g.uy - 2 + r;
            const int x = g.ux - 8 + 4 * k;
            const bool vy = y >= 0 && (y < g.sy1 || g.row_ext);
            uint32_t e[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t v = sizeof(T) == 2 ? (R.v[j][q >> 1] >> (16 * (q & 1))) & 0xffffu
                                                  : (R.v[j][0] >> (8 * q)) & 0xffu;
                const bool vx = x + q >= 0 && (x + q < g.sx1 || g.col_ext);
                e[q] = (vy && vx) ? v : (uint32_t)CDEF_VL;
            }
            // the tile row starts 4-byte aligned (even pitch), the chunk at an even column
            t[0] = e[0] | (e[1] << 16);
            t[1] = e[2] | (e[3] << 16);
            // the copy starts at an odd sample: one half word, a word, one half word
            c[0] = (uint16_t)e[0];
            *reinterpret_cast<uint32_t *>(c + 1) = e[1] | (e[2] << 16);
            c[3] = (uint16_t)e[3];
        }
    }
}

// The eight packed strengths of a table (4-byte aligned) as one value, byte i = strength i.
__device__ __forceinline__ uint64_t str8(gptr<const uint8_t> p)
{
    const auto w = (gptr<const uint32_t>)p;
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
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
__device__ __forceinline__ v2u asu(v2s v) { return __builtin_bit_cast(v2u, v); }
__device__ __forceinline__ v2u vminu(v2u a, v2u b) { return __builtin_elementwise_min(a, b); }
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
    // plane 0's tile loads go out first: they depend only on the unit's
    // position, so they overlap the loads of the unit's parameters below
    TileRegs<T> R;
    const int uw0 = min(64, src.p[0].aw - uc * 64), uh0 = min(64, src.p[0].ah - ur * 64);
    const TileGeo g0 = tile_geo(src.p[0], uc * 64, ur * 64, uw0, uh0);
    if (uw0 > 0 && uh0 > 0) tile_fetch<T>(R, src.p[0], g0, tid);
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
        const uint64_t ys = str8(ystr), uvs = str8(uvstr);
        const uint64_t used = count == 8 ? ~0ull : (1ull << (8 * count)) - 1;
        yfilt = (ys & used) != 0;
        uvfilt = src.nplanes > 1 && (uvs & used) != 0;
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
    // all eight strengths of each kind as two words each, read without
    // waiting for the index, then picked by it
    const uint64_t ysw = str8(ystr), uvsw = str8(uvstr);
    const int ypk = idx >= 0 ? (int)((ysw >> (8 * idx)) & 0xff) : 0;
    const int uvpk = idx >= 0 ? (int)((uvsw >> (8 * idx)) & 0xff) : 0;

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
            packed = (uint8_t)(pl ? uvpk : ypk);
            const int uvpri = str_pri((uint8_t)uvpk);
            need_dir = pl == 0 ? (str_pri((uint8_t)ypk) != 0 || (chroma_filtering && uvpri != 0)) : uvpri != 0;
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
        const TileGeo g = pl == 0 ? g0 : tile_geo(sp, ux, uy, uw, uh);
        if (pl > 0) tile_fetch<T>(R, sp, g, tid);   // in flight across the barrier
        __syncthreads();   // previous plane finished with the tile / block tables
        tile_store<T>(tile, R, g, tid);
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
        // one lane per (block, pair of columns): both samples of a pair lie
        // in one block (blocks are at least 4 wide) and share its record, so
        // the lane reads the record and forms the 12 tap addresses once, then
        // walks the block's rows, each tap one 32-bit LDS read at an
        // immediate row offset.  Packed 16-bit arithmetic (values stay below
        // 2^15 in magnitude: samples < 2^12, the sentinel -2^14, the weighted
        // sum below 2^12).  Lanes of a wave cover whole rows of a block row,
        // so the stores coalesce.  Blocks 4 wide (chroma subsampled across)
        // give 128 items: two lanes take each, one half of its rows apiece.
        const int P = bw >> 1;               // pairs per block row: 2 or 4
        const int per_brow = 8 * P;          // items per block row (8 block columns)
        const int hr = P == 2 ? bh >> 1 : bh;                  // rows per lane
        const int it = P == 2 ? (tid & 127) : tid;
        const int r0 = P == 2 ? (tid >> 7) * hr : 0;
        if (it < brows * per_brow) {
            const int br = it / per_brow, c = it - br * per_brow;
            const int bc = c / P;
            const int x = bc * bw + 2 * (c - bc * P);
            const int y0 = br * bh + r0;
            const int rows = min(hr, uh - y0);
            if (x < uw && rows > 0) {
                const BlockRec r = brec[br * 8 + bc];
                const uint32_t *Wt = reinterpret_cast<const uint32_t *>(tile);
                const int w0 = (int)((org - tile + y0 * TPITCH + x) >> 1);   // word of the pair in row y0
                const int flag = r.p >> 24;
                const int st = r.p & 255, psh = (r.p >> 8) & 255, ssh = (r.p >> 16) & 255;
                const v2u tp = splatu2(st), shp = splatu2(psh), ts = splatu2(sec), shs = splatu2(ssh);
                const v2s t0 = splat2(((st >> cs) & 1) ? 3 : 4), t1 = splat2(((st >> cs) & 1) ? 3 : 2);
                int A[12];   // word index of each tap pair in row y0
#pragma unroll
                for (int k = 0; k < 12; k++) A[k] = w0 + r.e[k];
                const bool pair = x + 1 < uw;
                const bool wide = pair && dp.al >= (int)(2 * sizeof(T));   // one store for the pair
                auto row_out = gmem(row_ptr<T>(dp, uy + y0) + ux + x);
#pragma unroll
                for (int yy = 0; yy < 8; yy++) {
                    if (yy < rows) {
                        constexpr int RW = TPITCH / 2;   // words per tile row
                        const v2s xv = as2(Wt[w0 + yy * RW]);
                        v2s out = xv;
                        if (flag & 1) {
                            const v2s a0 = as2(Wt[A[0] + yy * RW]), a1 = as2(Wt[A[1] + yy * RW]);
                            const v2s a2 = as2(Wt[A[2] + yy * RW]), a3 = as2(Wt[A[3] + yy * RW]);
                            const v2s b0 = as2(Wt[A[4] + yy * RW]), b1 = as2(Wt[A[5] + yy * RW]);
                            const v2s b2 = as2(Wt[A[6] + yy * RW]), b3 = as2(Wt[A[7] + yy * RW]);
                            const v2s b4 = as2(Wt[A[8] + yy * RW]), b5 = as2(Wt[A[9] + yy * RW]);
                            const v2s b6 = as2(Wt[A[10] + yy * RW]), b7 = as2(Wt[A[11] + yy * RW]);
                            v2s sum = splat2(0);
                            if (pri)   // st may be 0 for a flat block: its primary terms vanish
                                sum = t0 * (cst2(a0, xv, tp, shp) + cst2(a1, xv, tp, shp)) +
                                      t1 * (cst2(a2, xv, tp, shp) + cst2(a3, xv, tp, shp));
                            if (sec)
                                sum += splat2(2) * (cst2(b0, xv, ts, shs) + cst2(b1, xv, ts, shs) + cst2(b2, xv, ts, shs) +
                                                    cst2(b3, xv, ts, shs)) +
                                       cst2(b4, xv, ts, shs) + cst2(b5, xv, ts, shs) + cst2(b6, xv, ts, shs) +
                                       cst2(b7, xv, ts, shs);
                            // (8 + sum - (sum < 0)) >> 4, with sum >> 15 the -1 of a negative sum
                            out = xv + ((splat2(8) + sum + (sum >> splat2(15))) >> splat2(4));
                            if (flag & 2) {
                                // clip to the taps' range: the sentinel (0xC000) is the
                                // smallest signed value, so it never sets the maximum,
                                // and the largest unsigned one, so it never sets the minimum
                                v2s hi = vmax(vmax(vmax(xv, a0), vmax(a1, a2)), vmax(a3, b0));
                                hi = vmax(vmax(vmax(hi, b1), vmax(b2, b3)), vmax(vmax(b4, b5), vmax(b6, b7)));
                                const v2u ux0 = __builtin_bit_cast(v2u, xv);
                                v2u lo = vminu(vminu(vminu(ux0, asu(a0)), vminu(asu(a1), asu(a2))), vminu(asu(a3), asu(b0)));
                                lo = vminu(vminu(vminu(lo, asu(b1)), vminu(asu(b2), asu(b3))),
                                           vminu(vminu(asu(b4), asu(b5)), vminu(asu(b6), asu(b7))));
                                out = vmin(vmax(out, __builtin_bit_cast(v2s, lo)), hi);
                            }
                        }
                        const auto o = row_out + (int64_t)yy * (dp.stride / (int64_t)sizeof(T));
                        if (wide && sizeof(T) == 2) {
                            *(gptr<uint32_t>)o = __builtin_bit_cast(uint32_t, out);
                        } else if (wide) {
                            *(gptr<uint16_t>)o = (uint16_t)((uint16_t)out.x | ((uint16_t)out.y << 8));
                        } else {
                            o[0] = (T)(uint16_t)out.x;
                            if (pair) o[1] = (T)(uint16_t)out.y;
                        }
                    }
                }
            }
        }
    }
}

// LDS of a CDEF unit: the tile, then its copy at TSHIFT; per 8x8 block the
// direction, variance, direction costs and filter record.
struct __align__(16) CdefLds {
    uint16_t tile[TSHIFT + TROWS * TPITCH + 1];
    uint8_t bdir[64];
    int32_t bvar[64];
    uint32_t dcost[64][8];
    BlockRec brec[64];
};

// Unit (uc, ur) of a frame: the deblocked frame A into its scratch B.
__device__ __forceinline__ void cdef_tile(const FrameDesc &D, int uc, int ur, CdefLds &S)
{
    if (D.v[VIEW_A].bd == 8)
        cdef_unit<uint8_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, S.tile, S.bdir, S.bvar, S.dcost, S.brec);
    else
        cdef_unit<uint16_t>(D.v[VIEW_A], D.v[VIEW_B], D.cd, D.err, uc, ur, S.tile, S.bdir, S.bvar, S.dcost, S.brec);
}

// 6 workgroups per CU: 26.7 KB of LDS each, and at most 80 VGPRs
__global__ __launch_bounds__(256, 6) void cdef_kernel(const FrameDesc *__restrict__ descs,
                                                   const Seg *__restrict__ segs, int nseg, int total, StageLaunch L)
{
    __shared__ CdefLds S;
    for_each_unit(total, L, [&](int u) {
        const Seg sg = segs[find_seg(segs, nseg, u)];
        const int t = u - sg.base;
        cdef_tile(descs[sg.frame], t % sg.nx, t / sg.nx, S);
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
