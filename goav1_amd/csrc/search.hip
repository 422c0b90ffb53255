// Encoder CDEF strength search (SVT search level 1) on gfx950.
//
// Semantics follow goav1's cdefSearcher (encoder/cdef_search.go:580-749) and
// cdef.FilterBlocksU8SubsampledTrusted (cdef/filter_search.go:20-74):
//   - candidate gi 0..15 = primary gi with secondary 0, then for each primary
//     p the secondaries 1..3 (level-1 list, cdef_search.go:72-110);
//   - blocks = 8x8 blocks of the 64x64 filter block with a non-skip MI, in
//     raster order (computeCDEFList, :410-450);
//   - the filter input has VeryLarge outside the VISIBLE plane (loadCDEFInput,
//     :479-505); luma directions come from the reconstruction; chroma takes the
//     luma direction of the co-located block;
//   - a trial's luma primary is variance-adjusted per block only when it is
//     nonzero; a block whose strengths are both zero keeps the reconstruction;
//   - mse[0][fb][gi] = luma SSE, mse[1][fb][gi] = U SSE + V SSE (uint64), row
//     step 1 (level 1 subsampling factor).
//
// Decomposition: one workgroup per filter block.  Each lane owns 16 luma (then
// 8 chroma) pixels, loads the centre and its taps once from LDS, and evaluates
// all 64 trials against them with 64 register accumulators (the batched form
// of cdef/filter_search_batch.go); the accumulators are reduced across the
// workgroup with wave shuffles and a 4-way LDS sum.  The work is integer-VALU
// bound (64 trials x ~70 ops per pixel).
#include <map>
#include <mutex>
#include <type_traits>
#include "pf_internal.h"
#include "nn_internal.h"

namespace pf {

#define SVL 0x4000
#define SP 80   // LDS tile pitch

__device__ __constant__ uint8_t kLevel1[64] = {
    0, 4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 52, 56, 60,
    1, 2, 3, 5, 6, 7, 9, 10, 11, 13, 14, 15, 17, 18, 19, 21, 22, 23, 25, 26, 27,
    29, 30, 31, 33, 34, 35, 37, 38, 39, 41, 42, 43, 45, 46, 47, 49, 50, 51, 53, 54, 55,
    57, 58, 59, 61, 62, 63 };

__device__ __constant__ int kSOff[12][2] = {
    { 1 * SP, 2 * SP },         { 1 * SP, 2 * SP - 1 },
    { -1 * SP + 1, -2 * SP + 2 }, { 1, -1 * SP + 2 },
    { 1, 2 },                   { 1, 1 * SP + 2 },
    { 1 * SP + 1, 2 * SP + 2 }, { 1 * SP, 2 * SP + 1 },
    { 1 * SP, 2 * SP },         { 1 * SP, 2 * SP - 1 },
    { -1 * SP + 1, -2 * SP + 2 }, { 1, -1 * SP + 2 },
};
__device__ __constant__ int kDiv[9] = { 0, 840, 420, 280, 210, 168, 140, 120, 105 };

__device__ __forceinline__ int s_constrain(int diff, int t, int sh)
{
    const int a = diff < 0 ? -diff : diff;
    const int l = min(max(t - (a >> sh), 0), a);
    return diff < 0 ? -l : l;
}
__device__ __forceinline__ int s_ilog2(unsigned v) { return 31 - __clz(v); }

// CS = bit depth - 8: the direction search reads samples shifted down by CS
template <int CS>
__device__ void s_find_dir(const uint16_t *img, int *dir, int32_t *var)
{
    int32_t pa[8][15];
#pragma unroll
    for (int d = 0; d < 8; d++)
#pragma unroll
        for (int k = 0; k < 15; k++) pa[d][k] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int32_t x = (int32_t)(img[i * SP + j] >> CS) - 128;
            pa[0][i + j] += x; pa[1][i + (j >> 1)] += x; pa[2][i] += x; pa[3][3 + i - (j >> 1)] += x;
            pa[4][7 + i - j] += x; pa[5][3 - (i >> 1) + j] += x; pa[6][j] += x; pa[7][(i >> 1) + j] += x;
        }
    uint32_t c[8];
#pragma unroll
    for (int d = 0; d < 8; d++) {
        uint32_t s = 0;
        if (d == 2 || d == 6) {
#pragma unroll
            for (int k = 0; k < 8; k++) s += (uint32_t)(pa[d][k] * pa[d][k]);
            s *= 105u;
        } else if (d == 0 || d == 4) {
#pragma unroll
            for (int k = 0; k < 7; k++) s += (uint32_t)(pa[d][k] * pa[d][k] + pa[d][14 - k] * pa[d][14 - k]) * (uint32_t)kDiv[k + 1];
            s += (uint32_t)(pa[d][7] * pa[d][7]) * 105u;
        } else {
#pragma unroll
            for (int k = 3; k < 8; k++) s += (uint32_t)(pa[d][k] * pa[d][k]);
            s *= 105u;
#pragma unroll
            for (int k = 0; k < 3; k++) s += (uint32_t)(pa[d][k] * pa[d][k] + pa[d][10 - k] * pa[d][10 - k]) * (uint32_t)kDiv[2 * k + 2];
        }
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

// The reconstruction tile (ROWS rows of SP columns: the block with its
// 2-row / 8-column halo, VeryLarge outside [0,pw)x[0,ph); the whole tile is
// staged whatever the block's size, so a lane's positions are fixed) and the
// filter block's source samples (w x h of at most S x S) into srcT (pitch
// 64).  BATCH: every load in flight before the first LDS store (for the
// kernel whose occupancy its registers already set); otherwise a plain loop
// that keeps the register count low.
template <typename T, int ROWS, int S, bool BATCH = true>
__device__ __forceinline__ void s_stage(uint16_t *tile, const PlaneView &rp, uint16_t *srcT, const PlaneView &sp,
                                        int x0, int y0, int w, int h, int tid)
{
    constexpr int N = ROWS * SP, K = (N + 255) / 256, KS = S * S / 256;
    auto rec_at = [&](int i) -> uint16_t {
        const int r = i / SP, c = i - r * SP;
        const int y = y0 - 2 + r, x = x0 - 8 + c;
        return (y >= 0 && y < rp.h && x >= 0 && x < rp.w) ? (uint16_t)row_ptr<T>(rp, y)[x] : (uint16_t)SVL;
    };
    auto src_at = [&](int i) -> uint16_t {
        const int r = i / S, c = i % S;
        return (r < h && c < w) ? (uint16_t)row_ptr<T>(sp, y0 + r)[x0 + c] : (uint16_t)0;
    };
    if constexpr (BATCH) {
        uint16_t v[K], u[KS];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * 256;
            v[k] = i < N ? rec_at(i) : (uint16_t)SVL;
        }
#pragma unroll
        for (int k = 0; k < KS; k++) u[k] = src_at(tid + k * 256);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * 256;
            if (i < N) tile[i] = v[k];
        }
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int i = tid + k * 256;
            srcT[(i / S) * 64 + i % S] = u[k];
        }
    } else {
        for (int i = tid; i < N; i += 256) tile[i] = rec_at(i);
        for (int i = tid; i < S * S; i += 256) srcT[(i / S) * 64 + i % S] = src_at(i);
    }
}

// The 8x8 luma blocks of the filter block that have a non-skip MI, in raster
// order (one wave: a lane per block position, compacted by ballot).
__device__ __forceinline__ void s_block_list(uint64_t m, int nhb, int nvb, uint8_t *blist, int *nblk, int tid)
{
    if (tid < 64) {
        const bool on = 2 * (tid >> 3) < nvb && 2 * (tid & 7) < nhb && ((m >> tid) & 1);
        const uint64_t vote = __ballot(on);
        if (on) blist[__popcll(vote & ((1ull << tid) - 1))] = (uint8_t)tid;
        if (tid == 0) *nblk = __popcll(vote);
    }
}

// Evaluate the 64 trials for one pixel and add its squared errors to acc.
// CS = bit depth - 8: the decoder's high-bit-depth CDEF scales both strengths
// by 2^CS (the primary after the luma variance adjustment reads the scaled
// strength), picks the primary tap set by the unscaled strength's parity, and
// takes a damping (here passed in) CS larger.
template <bool LUMA, int CS>
__device__ __forceinline__ void s_pixel(uint32_t acc[64], const uint16_t *c, int srcv, int dir, int adj_i,
                                        bool has_var, int damping)
{
    const int x = c[0];
    const int p0 = kSOff[dir + 2][0], p1 = kSOff[dir + 2][1];
    const int a0 = c[p0], a1 = c[-p0], a2 = c[p1], a3 = c[-p1];
    // secondary taps for the block direction (primary trials) and direction 0 (primary-0 trials)
    int sb[8], sz[8];
    {
        const int o[4] = { kSOff[dir + 4][0], kSOff[dir][0], kSOff[dir + 4][1], kSOff[dir][1] };
#pragma unroll
        for (int k = 0; k < 4; k++) { sb[2 * k] = c[o[k]]; sb[2 * k + 1] = c[-o[k]]; }
        const int z[4] = { kSOff[4][0], kSOff[0][0], kSOff[4][1], kSOff[0][1] };
#pragma unroll
        for (int k = 0; k < 4; k++) { sz[2 * k] = c[z[k]]; sz[2 * k + 1] = c[-z[k]]; }
    }
    const int mk = SVL - 1;
    int hi = max(max(max(x, a0 & mk), max(a1 & mk, a2 & mk)), a3 & mk);
    int lo = min(min(min(x, a0), min(a1, a2)), a3);
#pragma unroll
    for (int k = 0; k < 8; k++) { hi = max(hi, sb[k] & mk); lo = min(lo, sb[k]); }
#pragma unroll
    for (int gi = 0; gi < 64; gi++) {
        const int packed = gi < 16 ? 4 * gi : 4 * ((gi - 16) / 3) + (gi - 16) % 3 + 1;   // compile-time
        const int pri = packed >> 2;
        const int sec = ((packed & 3) == 3 ? 4 : (packed & 3)) << CS;
        int st = pri << CS;
        if (LUMA && pri) st = has_var ? (((pri << CS) * (4 + adj_i) + 8) >> 4) : 0;
        int out = x;
        if (st || sec) {
            const int ssh = sec ? max(damping - s_ilog2((unsigned)sec), 0) : 0;
            int sum = 0;
            if (pri) {
                const int psh = st ? max(damping - s_ilog2((unsigned)st), 0) : 0;
                const int t0 = ((st >> CS) & 1) ? 3 : 4, t1 = ((st >> CS) & 1) ? 3 : 2;
                if (st)
                    sum = t0 * (s_constrain(a0 - x, st, psh) + s_constrain(a1 - x, st, psh)) +
                          t1 * (s_constrain(a2 - x, st, psh) + s_constrain(a3 - x, st, psh));
                if (sec)
                    sum += 2 * (s_constrain(sb[0] - x, sec, ssh) + s_constrain(sb[1] - x, sec, ssh) +
                                s_constrain(sb[2] - x, sec, ssh) + s_constrain(sb[3] - x, sec, ssh)) +
                           s_constrain(sb[4] - x, sec, ssh) + s_constrain(sb[5] - x, sec, ssh) +
                           s_constrain(sb[6] - x, sec, ssh) + s_constrain(sb[7] - x, sec, ssh);
                out = x + ((8 + sum - (sum < 0)) >> 4);
                if (st && sec) out = min(max(out, lo), hi);
            } else {
                sum = 2 * (s_constrain(sz[0] - x, sec, ssh) + s_constrain(sz[1] - x, sec, ssh) +
                           s_constrain(sz[2] - x, sec, ssh) + s_constrain(sz[3] - x, sec, ssh)) +
                      s_constrain(sz[4] - x, sec, ssh) + s_constrain(sz[5] - x, sec, ssh) +
                      s_constrain(sz[6] - x, sec, ssh) + s_constrain(sz[7] - x, sec, ssh);
                out = x + ((8 + sum - (sum < 0)) >> 4);
            }
        }
        const int e = srcv - out;
        acc[gi] += (uint32_t)(e * e);
    }
}

__device__ __forceinline__ void s_reduce(uint32_t acc[64], uint32_t (*part)[64], int tid)
{
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int gi = 0; gi < 64; gi++) {
        uint32_t v = acc[gi];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) part[wv][gi] = v;
    }
}

// CS = bit depth - 8 (0: 8-bit samples, 2: 10-bit); damping is the frame's
// (the decoder damping of the trials is damping + CS, one less for chroma).
// Two workgroups per filter block (blockIdx.z): 0 the luma row of the table,
// 1 the chroma row; both search the luma directions, which chroma filters with.
template <int CS>
__global__ __launch_bounds__(256) void cdef_search_kernel(FrameView src, FrameView rec, const uint64_t *noskip,
                                                          int damping, uint64_t *mse, int nhfb, int nvfb)
{
    using T = typename std::conditional<CS == 0, uint8_t, uint16_t>::type;
    __shared__ uint16_t tile[68 * SP];
    __shared__ uint16_t srcT[64 * 64];
    __shared__ uint8_t bdir[64];
    __shared__ int8_t badj[64];      // adjust index i, -1 = zero variance
    __shared__ uint8_t blist[64];
    __shared__ int nblk_s;
    __shared__ uint32_t part[4][64];
    const int tid = threadIdx.x;
    const int fbc = blockIdx.x, fbr = blockIdx.y, chroma = blockIdx.z;
    const int fb = fbr * nhfb + fbc;
    const int nfb = nhfb * nvfb;
    const int mi_cols = ((rec.w + 7) >> 3) << 1, mi_rows = ((rec.h + 7) >> 3) << 1;
    const int nhb = min(16, mi_cols - 16 * fbc), nvb = min(16, mi_rows - 16 * fbr);
    const uint64_t m = noskip ? noskip[fb] : ~0ull;
    s_block_list(m, nhb, nvb, blist, &nblk_s, tid);
    __syncthreads();
    const int nblk = nblk_s;
    uint64_t *m0 = mse + (size_t)fb * 64, *m1 = mse + ((size_t)nfb + fb) * 64;
    if (nblk == 0) {
        if (tid < 64) (chroma ? m1 : m0)[tid] = 0;
        return;
    }
    uint32_t acc[64];
    // ---- luma
    {
        const int x0 = fbc * 64, y0 = fbr * 64, w = nhb * 4, h = nvb * 4;
        s_stage<T, 68, 64>(tile, rec.p[0], srcT, src.p[0], x0, y0, w, h, tid);
        __syncthreads();
        if (tid < nblk) {
            const int b = blist[tid];
            int d;
            int32_t v;
            s_find_dir<CS>(tile + (2 + (b >> 3) * 8) * SP + 8 + (b & 7) * 8, &d, &v);
            bdir[b] = (uint8_t)d;
            int i = -1;
            if (v != 0) {
                i = 0;
                const int32_t vv = v >> 6;
                if (vv != 0) i = min(s_ilog2((unsigned)vv), 12);
            }
            badj[b] = (int8_t)i;
        }
        __syncthreads();
        if (!chroma) {
#pragma unroll
            for (int gi = 0; gi < 64; gi++) acc[gi] = 0;
            const int npx = nblk * 64;
            for (int p = tid; p < npx; p += 256) {
                const int b = blist[p >> 6], q = p & 63;
                const int by = b >> 3, bx = b & 7, yy = by * 8 + (q >> 3), xx = bx * 8 + (q & 7);
                const uint16_t *c = tile + (2 + yy) * SP + 8 + xx;
                s_pixel<true, CS>(acc, c, srcT[yy * 64 + xx], bdir[b], badj[b], badj[b] >= 0, damping + CS);
            }
            s_reduce(acc, part, tid);
            __syncthreads();
            if (tid < 64) m0[tid] = (uint64_t)part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
            return;
        }
    }
    // ---- chroma (U + V)
#pragma unroll
    for (int gi = 0; gi < 64; gi++) acc[gi] = 0;
    for (int pl = 1; pl <= 2; pl++) {
        const int x0 = fbc * 32, y0 = fbr * 32, w = nhb * 2, h = nvb * 2;
        __syncthreads();
        s_stage<T, 36, 32>(tile, rec.p[pl], srcT, src.p[pl], x0, y0, w, h, tid);
        __syncthreads();
        const int npx = nblk * 16;
        for (int p = tid; p < npx; p += 256) {
            const int b = blist[p >> 4], q = p & 15;
            const int by = b >> 3, bx = b & 7, yy = by * 4 + (q >> 2), xx = bx * 4 + (q & 3);
            const uint16_t *c = tile + (2 + yy) * SP + 8 + xx;
            s_pixel<false, CS>(acc, c, srcT[yy * 64 + xx], bdir[b], 0, true, damping - 1 + CS);
        }
    }
    __syncthreads();
    s_reduce(acc, part, tid);
    __syncthreads();
    if (tid < 64) m1[tid] = (uint64_t)part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// ---------------------------------------------------------------------------
// Wave exchanges and reductions.  Lane exchanges (lane l reads lane l ^ X):
// DPP quad permutations for distances 1 and 2, swizzles within 32 lanes for
// 4..16, a permute for 32.
template <int X>
__device__ __forceinline__ uint32_t xchg32(uint32_t v)
{
    if constexpr (X == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false);
    else if constexpr (X == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false);
    else if constexpr (X == 4) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x101F);
    else if constexpr (X == 8) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x201F);
    else if constexpr (X == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
    else return (uint32_t)__shfl_xor((int)v, 32, 64);
}
template <int X>
__device__ __forceinline__ uint64_t xchg64(uint64_t v)
{
    return (uint64_t)xchg32<X>((uint32_t)v) | ((uint64_t)xchg32<X>((uint32_t)(v >> 32)) << 32);
}
// Largest 64-bit value over the wave (every lane ends with it).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    v = max(v, xchg64<1>(v));
    v = max(v, xchg64<2>(v));
    v = max(v, xchg64<4>(v));
    v = max(v, xchg64<8>(v));
    v = max(v, xchg64<16>(v));
    return max(v, xchg64<32>(v));
}

// ---------------------------------------------------------------------------
// Strength selection on the device in one launch (finish /
// jointStrengthSearchDual / searchOneDual, cdef_search.go:780-945).
//
// The four cdef_bits levels are independent searches, so they run
// concurrently: level 0 on workgroup 0, level L >= 1 on SEL_G workgroups,
// one per luma strength j.  A level's step sequence is fixed (nb = 1 << L
// greedy steps filling slots 0..nb-1, then 4 nb refinements that drop the
// oldest pair and fill the last slot again); only the chosen pairs depend on
// the data.
//  - A step with no prior pair (slot 0) chooses luma and chroma separately
//    (first minimum of the column sums over the active blocks); every
//    workgroup of the level computes it, so it needs no exchange.
//  - Any other step: every workgroup holds the active blocks' best cost over
//    the prior pairs (best_i).  Wave w takes a contiguous share of the
//    blocks; a lane per block forms the room max(0, best_i - a_ij) (a block
//    with no room saves nothing for any k, as the reference skips it), and
//    the lanes k sum the savings max(0, room_i - b_ik) over the blocks (with
//    every value below 2^32 as saturating 32-bit differences, summed in 64
//    bits).  The workgroup's best pair becomes a
//    key (gain << 12 | 4095 - (64 j + k): the largest gain, then the first
//    pair in (j, k) order) that it publishes in its own slot of the step, and
//    every workgroup reads the level's SEL_G slots and keeps the largest key,
//    so all of them hold the same lists.
//  - Each chosen pair's cost column (a_ij + b_ik of every block) stays in LDS
//    when it fits, so the best cost over the pairs the next refinement keeps
//    is a minimum over LDS columns, computed while the exchange is in flight.
//  - The last level to finish (a ticket) takes the cheapest level by the RD
//    cost and writes the decision and each block's index.
// The chroma rows of the active blocks are read as 32-bit values: from LDS
// when they fit there (TABLES >= 1; TABLES = 2 also holds the luma rows),
// else from a 32-bit copy in the state buffer that the workgroups write at
// the start, behind a grid-wide barrier.  A value of 2^32 or more switches a
// workgroup to the 64-bit rows in device memory.
// The slots are written by device-scope atomic stores and polled by
// device-scope atomic loads, which are performed where every XCD sees them,
// so the exchange needs no cache write-back or invalidation; the keys are the
// only data exchanged.  Each member's slots sit in their own 4 KiB page, so a
// level's polls spread over the memory channels instead of queueing on a few
// cache lines.  The slots alternate between two halves by launch parity, each
// launch clearing the half the next one uses; the ticket and barrier
// counters only grow.  A wait that runs past a bounded spin marks the launch
// aborted (every workgroup then leaves) and sets bit 1 of the error flag.
constexpr int SEL_G = 64;                  // workgroups per level >= 1 (one per luma strength)
static_assert(SEL_G == 64, "one wave reads a level's slots, one lane each");
// threads per workgroup: 8 waves when a table sits in LDS, else 16 (the
// chroma rows come from device memory: more waves hide its latency)
__host__ __device__ constexpr int sel_threads(int tables) { return tables == 0 ? 1024 : 512; }
constexpr int SEL_GRID = 1 + 3 * SEL_G;    // level 0 on workgroup 0
constexpr int SEL_STEPS = 39;              // exchanged steps of level 3 (8 greedy + 32 refinements, less the first)
constexpr int SEL_HALF = 3 * SEL_STEPS;    // slots of one launch parity in a member's page
constexpr int SEL_PAGE = 512;              // 64-bit words per member page (4 KiB)
static_assert(2 * SEL_HALF <= SEL_PAGE, "a member's slots fit its page");
constexpr int SEL_TP = 65;                 // LDS row pitch of the tables (words): lanes reading one
                                           // column of consecutive rows hit distinct banks
// dynamic LDS; the kernel's own variables take less than the rest of 160 KiB
__host__ __device__ constexpr size_t sel_lds_budget(int tables) { return 160 * 1024 - (tables == 0 ? 10240 : 6144); }

struct FinishState {
    unsigned long long slot[SEL_G][SEL_PAGE];   // published keys, bit 63 = present
    uint32_t c32[4096 * 64];               // 32-bit chroma rows of the active blocks (TABLES == 0)
    unsigned int done;                     // levels finished (monotonic, 4 per launch)
    unsigned int conv;                     // workgroups past the chroma copy (monotonic, SEL_GRID per launch
                                           // of any table variant)
    unsigned int c32bad;                   // epoch whose copy holds a value of 2^32 or more
    unsigned int abort;                    // epoch of an aborted launch
    int ly[4][8], lc[4][8];                // each level's final pairs
    unsigned long long cost[4];
};

// LDS: active block list, each block's best cost, the member's luma column,
// the chosen pairs' cost columns (when they fit), then the chroma and luma
// tables.
struct SelLayout {
    size_t act, best, lcol, cc, ct, lt, total;
};

__host__ __device__ inline SelLayout sel_layout(int nfb, int tables, bool cc)
{
    SelLayout L;
    size_t o = 0;
    L.act = o;
    o += ((size_t)nfb * 4 + 15) / 16 * 16;
    L.best = o;
    o += (size_t)nfb * 8;
    L.lcol = o;
    o += ((size_t)nfb * 4 + 15) / 16 * 16;
    L.cc = o;
    o += cc ? ((size_t)8 * nfb * 4 + 15) / 16 * 16 : 0;
    L.ct = o;
    o += tables >= 1 ? (size_t)nfb * SEL_TP * 4 : 0;
    L.lt = o;
    o += tables == 2 ? (size_t)nfb * SEL_TP * 4 : 0;
    L.total = o;
    return L;
}

// bounded wait on a monotonic counter; false when the launch was aborted
__device__ inline bool sel_wait_count(unsigned int *cnt, unsigned int target, FinishState *st, unsigned int epoch,
                                      int *err)
{
    for (int spin = 0;; spin++) {
        if (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        if (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch || spin > (1 << 21)) {
            atomicExch(&st->abort, epoch);
            atomicOr(err, 2);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template <int TABLES>
__global__ __launch_bounds__(sel_threads(TABLES)) void cdef_select_kernel(const uint64_t *__restrict__ mse, const uint64_t *noskip,
                                                            int nhfb, int nvfb, int mi_cols, int mi_rows,
                                                            unsigned long long lambda, FinishState *st,
                                                            unsigned int epoch, int use_cc, av1pf_cdef_decision *dec,
                                                            uint8_t *fb_index, int *err)
{
    constexpr int SEL_T = sel_threads(TABLES), SEL_W = SEL_T / 64;
    extern __shared__ __align__(16) char smem[];
    __shared__ unsigned long long part[SEL_W][64];
    __shared__ unsigned long long amask[64];
    __shared__ int aoff[64];
    __shared__ unsigned long long s_key, s_dist, s_min[2];
    __shared__ int ly[8], lc[8], rp[8];
    __shared__ int s_nact, s_bad, s_abort, s_wide, s_pos, s_fpos;
    __shared__ bool s_last;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nfb = nhfb * nvfb;
    const int b = blockIdx.x;
    const int level = b == 0 ? 0 : 1 + (b - 1) / SEL_G;
    const int j = b == 0 ? 0 : (b - 1) % SEL_G;
    const bool cc = use_cc != 0;
    const SelLayout Lo = sel_layout(nfb, TABLES, cc);
    int *act = reinterpret_cast<int *>(smem + Lo.act);
    unsigned long long *best = reinterpret_cast<unsigned long long *>(smem + Lo.best);
    uint32_t *lcol = reinterpret_cast<uint32_t *>(smem + Lo.lcol);   // a_ij, 32-bit
    uint32_t *costc = reinterpret_cast<uint32_t *>(smem + Lo.cc);    // [8][nfb], saturated at 2^32 - 1
    uint32_t *ct = reinterpret_cast<uint32_t *>(smem + Lo.ct);       // chroma rows
    uint32_t *lt = reinterpret_cast<uint32_t *>(smem + Lo.lt);       // luma rows (TABLES == 2)
    const uint64_t *Lm = mse, *Cm = mse + (size_t)nfb * 64;
    // ---- active blocks (a non-skip 8x8 block inside the frame), ascending:
    // a ballot per 64 blocks, the chunks' offsets by one wave's scan, then
    // every block writes its own entry
    for (int c0 = wv; c0 * 64 < nfb; c0 += SEL_W) {
        const int fb = c0 * 64 + lane;
        bool a = false;
        if (fb < nfb) {
            const int fbr = fb / nhfb, fbc = fb % nhfb;
            const int nhb = min(16, mi_cols - 16 * fbc), nvb = min(16, mi_rows - 16 * fbr);
            // the 8x8 blocks inside the frame: (nhb + 1) / 2 columns of the
            // first (nvb + 1) / 2 rows of the 8 x 8 bit grid
            const uint64_t cols = (1ull << ((nhb + 1) >> 1)) - 1;
            const int rows = (nvb + 1) >> 1;
            const uint64_t valid = (cols * 0x0101010101010101ull) & (rows >= 8 ? ~0ull : (1ull << (8 * rows)) - 1);
            a = ((noskip ? noskip[fb] : ~0ull) & valid) != 0;
        }
        const uint64_t vote = __ballot(a);
        if (lane == 0) amask[c0] = vote;
    }
    __syncthreads();
    if (wv == 0) {
        const int nch = (nfb + 63) >> 6;   // at most 64 (4096 blocks)
        const int cnt = lane < nch ? __popcll(amask[lane]) : 0;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        aoff[lane] = incl - cnt;
        if (lane == 63) {
            s_nact = incl;
            s_bad = 0;
            s_abort = 0;
            s_wide = 0;
        }
    }
    __syncthreads();
    for (int c0 = wv; c0 * 64 < nfb; c0 += SEL_W) {
        const uint64_t m = amask[c0];
        if ((m >> lane) & 1) act[aoff[c0] + __popcll(m & ((1ull << lane) - 1))] = c0 * 64 + lane;
    }
    // the next launch's slots start empty
    if (b == 0)
        for (int i = t; i < SEL_G * SEL_HALF; i += SEL_T)
            st->slot[i / SEL_HALF][((epoch + 1) & 1) * SEL_HALF + i % SEL_HALF] = 0ull;
    __syncthreads();
    const int nact = s_nact;
    bool bad = false;
    if (TABLES >= 1) {
        // the MSE rows of the active blocks in LDS, eight rows' loads in
        // flight per thread before any store
        if (b != 0)
            for (int e0 = t; e0 < nact * 64; e0 += 8 * SEL_T) {
                uint64_t c[8], l[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = e0 + u * SEL_T;
                    const int fb = act[min(e, nact * 64 - 1) >> 6], g = e & 63;
                    c[u] = Cm[(size_t)fb * 64 + g];
                    l[u] = TABLES == 2 ? Lm[(size_t)fb * 64 + g] : 0;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int e = e0 + u * SEL_T;
                    if (e < nact * 64) {
                        ct[(e >> 6) * SEL_TP + (e & 63)] = (uint32_t)c[u];
                        if (TABLES == 2) lt[(e >> 6) * SEL_TP + (e & 63)] = (uint32_t)l[u];
                        bad |= ((c[u] | l[u]) >> 32) != 0;
                    }
                }
            }
        // every launch arrives at the barrier counter, so its target is
        // SEL_GRID per launch whichever kernel ran before
        if (t == 0) __hip_atomic_fetch_add(&st->conv, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        // the 32-bit chroma copy: every workgroup writes its share, then all
        // wait for all (the copy is read from every XCD: a release of this
        // XCD's L2 before the arrival, an acquire after the wait)
        bool cbad = false;
        for (int e0 = b * SEL_T + t; e0 < nact * 64; e0 += 4 * SEL_GRID * SEL_T) {
            uint64_t c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = min(e0 + u * SEL_GRID * SEL_T, nact * 64 - 1);
                c[u] = Cm[(size_t)act[e >> 6] * 64 + (e & 63)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * SEL_GRID * SEL_T;
                if (e < nact * 64) {
                    st->c32[e] = (uint32_t)c[u];
                    cbad |= (c[u] >> 32) != 0;
                }
            }
        }
        if (cbad) atomicExch(&st->c32bad, epoch);
        // each wave's stores complete before the barrier (it does not wait
        // for them), so the one write-back below covers them
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (t == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            // the write-back must be complete before the arrival (the
            // compiler does not count it as an outstanding store)
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __hip_atomic_fetch_add(&st->conv, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!sel_wait_count(&st->conv, SEL_GRID * epoch, st, epoch, err)) s_abort = 1;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (__hip_atomic_load(&st->c32bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) s_bad = 1;
        }
    }
    // the member's luma column, and every best cost "none yet"
    if (b != 0)
        for (int i0 = t; i0 < nact; i0 += 4 * SEL_T) {
            uint64_t v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = Lm[(size_t)act[min(i0 + u * SEL_T, nact - 1)] * 64 + j];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int i = i0 + u * SEL_T;
                if (i < nact) {
                    lcol[i] = (uint32_t)v[u];
                    best[i] = ~0ull;
                    bad |= (v[u] >> 32) != 0;
                }
            }
        }
    if (bad) s_bad = 1;
    __syncthreads();
    const bool lok = b != 0 && !s_bad;   // lcol and the 32-bit chroma rows hold exact values
    const bool ctab = TABLES >= 1 && lok, ltab = TABLES == 2 && lok;
    const uint32_t *c32 = st->c32;
    auto luma = [&](int i, int g) -> unsigned long long {
        return ltab ? (unsigned long long)lt[i * SEL_TP + g] : Lm[(size_t)act[i] * 64 + g];
    };
    auto chroma = [&](int i, int k) -> unsigned long long {
        return ctab ? (unsigned long long)ct[i * SEL_TP + k]
             : (TABLES == 0 && lok) ? (unsigned long long)c32[i * 64 + k]
                                    : Cm[(size_t)act[i] * 64 + k];
    };
    const int nb = 1 << level;
    const int steps = s_abort ? 0 : level == 0 ? 1 : 5 * nb;   // level 0 repeats one independent choice: once
    // wave w owns the blocks [lo, hi) (a multiple of 8 blocks each): it folds
    // their best costs, forms their rooms and sums their savings, so the waves
    // meet only around the exchange
    const int per = ((nact + SEL_W - 1) / SEL_W + 7) & ~7;
    const int lo = min(wv * per, nact), hi = min(lo + per, nact);
    bool wwide = false;   // a cost of 2^32 or more among the wave's blocks (sticky; s_wide: any block)
    // a new pair: its cost column, and the wave's blocks' best costs lowered
    // to it (best_i is "none yet" before the first pair, and before a
    // refinement's new pair the minimum over the pairs that refinement keeps)
    auto fold = [&](int py, int pc, int pos) {
        for (int i = lo + lane; i < hi; i += 64) {
            const unsigned long long c = luma(i, py) + chroma(i, pc);
            const unsigned long long m = c < best[i] ? c : best[i];
            best[i] = m;
            if (cc) costc[pos * nfb + i] = (c >> 32) ? 0xffffffffu : (uint32_t)c;
            wwide |= (c >> 32) != 0;
        }
        wwide = __ballot(wwide) != 0;
        if (wwide) s_wide = 1;   // for the recompute, which reads every wave's blocks
    };
    int gstep = 0;
    for (int s = 0; s < steps; s++) {
        const int slot = s < nb ? s : nb - 1;
        if (slot == 0) {
            // column sums: SEL_T / 128 threads per column over interleaved blocks
            const int col = t & 127, sub = t >> 7;
            unsigned long long cs = 0;
            for (int i = sub; i < nact; i += SEL_T / 128) cs += col < 64 ? luma(i, col) : chroma(i, col - 64);
            reinterpret_cast<unsigned long long *>(part)[sub * 128 + col] = cs;   // part as [SEL_T / 128][128]
            __syncthreads();
            if (t < 128) {
                // wave 0 holds the luma sums, wave 1 the chroma sums, lane k
                // strength k: the first minimum by a shuffle reduction
                const unsigned long long *pp = reinterpret_cast<const unsigned long long *>(part);
                unsigned long long v = 0;
#pragma unroll
                for (int q = 0; q < SEL_T / 128; q++) v += pp[128 * q + t];
                int k = lane;
                for (int o = 32; o > 0; o >>= 1) {
                    const unsigned long long ov = __shfl_xor(v, o);
                    const int ok = __shfl_xor(k, o);
                    if (ov < v || (ov == v && ok < k)) {
                        v = ov;
                        k = ok;
                    }
                }
                if (lane == 0) {
                    s_min[wv] = v;
                    if (wv == 0) ly[0] = nact ? k : 0;
                    else lc[0] = nact ? k : 0;
                }
            }
            __syncthreads();
            if (level == 0) {
                if (t == 0) s_dist = nact ? s_min[0] + s_min[1] : 0ull;
                continue;
            }
            const int py = ly[0], pc = lc[0];
            // the list holds pair 0 in cost column 0; the next step's pair goes to column 1
            __syncthreads();
            if (t == 0) {
                rp[0] = 0;
                s_pos = 1;
            }
            fold(py, pc, 0);
            continue;
        }
        // savings of the pairs (j, lane) over this wave's blocks: a lane per
        // block forms its room, then the rows are read eight at a time (one
        // address, immediate offsets; rows past the wave's own up to the
        // table's end carry no room) and each row's room is read from its lane
        unsigned long long gain = 0;
        if (lok && !wwide) {
            for (int base = lo; base < hi; base += 64) {
                const int n = min(64, hi - base);
                const uint32_t room =
                    lane < n ? __builtin_elementwise_sub_sat((uint32_t)best[base + lane], lcol[base + lane]) : 0u;
                const int n8 = base + ((n + 7) & ~7) <= nact ? (n + 7) & ~7 : n & ~7;
                int u0 = 0;
                for (; u0 < n8; u0 += 8) {
                    uint32_t c[8];
                    if constexpr (TABLES >= 1) {
                        const uint32_t *row = ct + (base + u0) * SEL_TP + lane;
#pragma unroll
                        for (int u = 0; u < 8; u++) c[u] = row[u * SEL_TP];
                    } else {
                        const uint32_t *row = c32 + (size_t)(base + u0) * 64 + lane;
#pragma unroll
                        for (int u = 0; u < 8; u++) c[u] = row[u * 64];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        gain += __builtin_elementwise_sub_sat(
                            (uint32_t)__builtin_amdgcn_readlane((int)room, u0 + u), c[u]);
                }
                for (; u0 < n; u0++) {
                    const uint32_t c = TABLES >= 1 ? ct[(base + u0) * SEL_TP + lane] : c32[(size_t)(base + u0) * 64 + lane];
                    gain += __builtin_elementwise_sub_sat((uint32_t)__builtin_amdgcn_readlane((int)room, u0), c);
                }
            }
        } else {
            auto save = [&](unsigned long long bi, unsigned long long a, unsigned long long c) {
                const unsigned long long room = a < bi ? bi - a : 0ull;
                return c < room ? room - c : 0ull;
            };
            for (int i = lo; i < hi; i += 8) {
                unsigned long long bi[8], a[8], c[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int ii = min(i + u, hi - 1);
                    bi[u] = best[ii];
                    a[u] = lok ? (unsigned long long)lcol[ii] : Lm[(size_t)act[ii] * 64 + j];
                    c[u] = chroma(ii, lane);
                }
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (i + u < hi) gain += save(bi[u], a[u], c[u]);
            }
        }
        part[wv][lane] = gain;
        const bool next_refines = s + 1 < steps && s + 1 >= nb;
        __syncthreads();
        if (wv == 0) {
            unsigned long long g = 0;
#pragma unroll
            for (int w = 0; w < SEL_W; w++) g += part[w][lane];
            unsigned long long key = wave_max_u64((g << 12) | (unsigned long long)(4095 - (j * 64 + lane)));
            const int at = (epoch & 1) * SEL_HALF + (level - 1) * SEL_STEPS + gstep;
            // a store that is not waited for: the wave goes on to poll
            if (lane == 0)
                __hip_atomic_store(&st->slot[j][at], key | (1ull << 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // every workgroup of the level reads the level's slots of this
            // step; two reads of a slot in flight (alternating registers), so
            // a key is seen about one memory latency after it lands (the
            // abort word every 64 rounds)
            const unsigned long long *sp = &st->slot[lane][at];
            unsigned long long v, pa = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned long long pb = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int spin = 0;; spin++) {
                if (pa >> 63) {
                    v = pa;
                    break;
                }
                pa = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (pb >> 63) {
                    v = pb;
                    break;
                }
                pb = __hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((spin & 63) == 63 &&
                    (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch ||
                     spin > (1 << 21))) {
                    atomicExch(&st->abort, epoch);
                    atomicOr(err, 2);
                    s_abort = 1;
                    v = 0;
                    break;
                }
            }
            key = wave_max_u64(v & ~(1ull << 63));
            if (lane == 0) {
                s_key = key;
                s_fpos = s_pos;
            }
        } else if (next_refines) {
            // while the exchange is in flight: the next step (a refinement,
            // which drops list entry 0) keeps entries 1 .. slot-1 besides the
            // pair this step adds; their minimum per block (every gain has
            // read best: a barrier lies between)
            for (int i = t - 64; i < nact; i += SEL_T - 64) {
                unsigned long long m = ~0ull;
                if (cc && !s_wide) {
#pragma unroll
                    for (int g = 1; g < 8; g++)
                        if (g < slot) {
                            const unsigned long long v = costc[rp[g] * nfb + i];
                            m = v < m ? v : m;
                        }
                } else {
#pragma unroll
                    for (int g = 1; g < 8; g++)
                        if (g < slot) {
                            const unsigned long long v = luma(i, ly[g]) + chroma(i, lc[g]);
                            m = v < m ? v : m;
                        }
                }
                best[i] = m;
            }
        }
        __syncthreads();
        if (s_abort) break;
        const int pair = 4095 - (int)(s_key & 4095ull);
        if (wv == SEL_W - 1) {
            // the lists after this step (the new pair in `slot`, its cost in
            // column s_fpos), then, when the next step refines, without entry
            // 0, whose column the next pair reuses; read by the next step's
            // recompute, behind its first barrier (the last wave has the
            // fewest blocks to fold)
            const int e = next_refines ? min(lane + 1, 7) : min(lane, 7);
            int y = ly[e], c = lc[e], r = rp[e];
            const int r0 = rp[0];
            if (e == slot) {
                y = pair >> 6;
                c = pair & 63;
                r = s_fpos;
            }
            asm volatile("" ::: "memory");   // every lane reads before any lane writes
            if (next_refines ? lane < nb - 1 : lane == slot) {
                ly[lane] = y;
                lc[lane] = c;
                rp[lane] = r;
            }
            if (lane == 0) s_pos = next_refines ? (slot == 0 ? s_fpos : r0) : s + 1;
        }
        gstep++;
        fold(pair >> 6, pair & 63, s_fpos);
    }
    __syncthreads();
    // ---- level result, then the last level to finish decides
    if (b != 0 && j != 0) return;
    if (b != 0) {
        // the level's distortion: every block at its best listed pair
        unsigned long long ps = 0;
        for (int i = t; i < nact; i += SEL_T) ps += best[i];
        for (int o = 32; o > 0; o >>= 1) ps += __shfl_down(ps, o);
        if (lane == 0) part[wv][0] = ps;
        __syncthreads();
        if (t == 0) {
            unsigned long long v = 0;
            for (int w = 0; w < SEL_W; w++) v += part[w][0];
            s_dist = nact ? v : 0ull;
        }
    }
    if (t == 0) {
        if (!s_abort) {
            const long long rate = (long long)(nact * level + nb * 12) << 9;
            st->cost[level] =
                (unsigned long long)((rate * (long long)lambda + (1 << 8)) >> 9) + ((s_dist * 16) << 7);
            for (int i = 0; i < 8; i++) {
                st->ly[level][i] = i < nb ? ly[i] : 0;
                st->lc[level][i] = i < nb ? lc[i] : 0;
            }
        }
        __threadfence();
        s_last = atomicAdd(&st->done, 1u) == 4u * epoch - 1u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return;
    __shared__ int ys[8], us[8], bits;
    if (t == 0) {
        int bb = 0;
        for (int l = 1; l < 4; l++)
            if (st->cost[l] < st->cost[bb]) bb = l;
        bits = bb;
        for (int i = 0; i < 8; i++) {
            ys[i] = st->ly[bb][i];
            us[i] = st->lc[bb][i];
        }
    }
    __syncthreads();
    const int nbb = 1 << bits;
    for (int fb = t; fb < nfb; fb += SEL_T) fb_index[fb] = 0;
    __syncthreads();
    for (int i = t; i < nact; i += SEL_T) {
        const int fb = act[i];
        unsigned long long bm = 1ull << 63;
        int bg = 0;
        for (int g = 0; g < nbb; g++) {
            const unsigned long long c = Lm[(size_t)fb * 64 + ys[g]] + Cm[(size_t)fb * 64 + us[g]];
            if (c < bm) {
                bm = c;
                bg = g;
            }
        }
        fb_index[fb] = (uint8_t)bg;
    }
    if (t == 0) {
        dec->cdef_bits = bits;
        for (int i = 0; i < 8; i++) {
            dec->y_strength[i] = i < nbb ? kLevel1[ys[i]] : 0;
            dec->uv_strength[i] = i < nbb ? kLevel1[us[i]] : 0;
        }
    }
}

size_t finish_state_bytes() { return sizeof(FinishState); }
size_t small_state_bytes() { return 1024; }

hipError_t launch_cdef_select(const uint64_t *mse, const uint64_t *noskip, int nhfb, int nvfb, int mi_cols, int mi_rows,
                              uint64_t lambda, void *state, unsigned int epoch, av1pf_cdef_decision *dec,
                              uint8_t *fb_index, int *err, hipStream_t s)
{
    FinishState *st = static_cast<FinishState *>(state);
    unsigned long long lam = lambda;
    const int nfb = nhfb * nvfb;
    if (nfb > 4096) return hipErrorInvalidValue;
    // the most that fits in LDS: both tables, the chroma table, neither; the
    // cost columns when they still fit
    int tables = 0, use_cc = 0;
    bool found = false;
    for (int tb = 2; tb >= 0 && !found; tb--)
        for (int c = 1; c >= 0 && !found; c--)
            if (sel_layout(nfb, tb, c != 0).total <= sel_lds_budget(tb)) {
                tables = tb;
                use_cc = c;
                found = true;
            }
    if (!found) return hipErrorInvalidValue;
    const void *fn = tables == 2 ? (const void *)cdef_select_kernel<2>
                   : tables == 1 ? (const void *)cdef_select_kernel<1>
                                 : (const void *)cdef_select_kernel<0>;
    const size_t lds = sel_layout(nfb, tables, use_cc != 0).total;
    // the attribute calls once per device and variant (they cost host time
    // between the launches of a frame's search)
    static std::mutex mu;
    static std::map<std::pair<int, int>, bool> ready;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!ready[{ dev, tables }]) {
            hipFuncAttributes fa;
            e = hipFuncGetAttributes(&fa, fn);
            if (e != hipSuccess) return e;
            if (fa.sharedSizeBytes + sel_lds_budget(tables) > 160 * 1024) return hipErrorInvalidValue;
            e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sel_lds_budget(tables));
            if (e != hipSuccess) return e;
            // every workgroup must be resident at once (a workgroup waits for
            // the slots of the other workgroups of its level, and for the
            // chroma copy of all): the device holds the grid at the largest
            // LDS request, or the launch is refused here
            int cus = 0, per_cu = 0;
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (e != hipSuccess) return e;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, sel_threads(tables), sel_lds_budget(tables));
            if (e != hipSuccess) return e;
            if ((long)cus * per_cu < SEL_GRID) return hipErrorCooperativeLaunchTooLarge;
            ready[{ dev, tables }] = true;
        }
    }
    void *args[] = { (void *)&mse, (void *)&noskip, (void *)&nhfb, (void *)&nvfb, (void *)&mi_cols, (void *)&mi_rows,
                     (void *)&lam, (void *)&st, (void *)&epoch, (void *)&use_cc, (void *)&dec, (void *)&fb_index,
                     (void *)&err };
    // an ordinary launch (the grid fits the device, checked above): the
    // cooperative queue's hand-off cost 15 us per call.  Workgroups that other
    // kernels keep from starting are waited for; every wait is bounded and
    // flags the launch aborted (bit 1 of the error flag) when it runs out.
    return hipLaunchKernel(fn, dim3(SEL_GRID), dim3(sel_threads(tables)), args, lds, s);
}

// ---------------------------------------------------------------------------
// Learned-proposal search.  The network proposes a frame shortlist of
// NCAND luma and NCAND chroma strengths; only those are evaluated exactly,
// and the same greedy selection runs over the NCAND x NCAND pairs.

// Exact SSE of the shortlisted strengths.  mse[2][nfb][NCAND].  Per pixel the
// taps are read once and turned into |d| and a signed weight (primary: +-1,
// secondary: +-2 for the near and +-1 for the far taps); a strength's
// contribution is then sum w_k clamp(s - (|d_k| >> shift), 0, |d_k|), the
// constrain function with its sign folded into the weight.  The secondary
// sums depend only on the secondary strength, so each one present in the
// shortlist (`need`: bits 0..2 for strengths 1, 2, 4 along the block
// direction, bits 3..5 along direction 0 for primary-0 trials) is computed
// once per pixel and shared by the trials.
// clamp(v, 0, a) for a >= 0, one v_med3_i32
__device__ __forceinline__ int s_clamp0(int v, int a)
{
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "v"(a));
    return r;
}

template <int N>
__device__ __forceinline__ int s_taps_sum(const int (&ad)[N], const int (&sw)[N], int st, int sh)
{
    int s = 0;
#pragma unroll
    for (int k = 0; k < N; k++) s += __mul24(sw[k], s_clamp0(st - (ad[k] >> sh), ad[k]));
    return s;
}

// CS and damping as in s_pixel (damping includes CS)
template <bool LUMA, int CS>
__device__ __forceinline__ void s_pixel_sub(uint32_t acc[NCAND], const uint16_t *c, int srcv, int dir, int adj_i,
                                            bool has_var, int damping, const int *pk, unsigned need)
{
    const int x = c[0];
    const int p0 = kSOff[dir + 2][0], p1 = kSOff[dir + 2][1];
    const int a[4] = { c[p0], c[-p0], c[p1], c[-p1] };
    int sb[8];
    {
        const int o[4] = { kSOff[dir + 4][0], kSOff[dir][0], kSOff[dir + 4][1], kSOff[dir][1] };
#pragma unroll
        for (int k = 0; k < 4; k++) { sb[2 * k] = c[o[k]]; sb[2 * k + 1] = c[-o[k]]; }
    }
    const int mk = SVL - 1;
    int hi = max(max(max(x, a[0] & mk), max(a[1] & mk, a[2] & mk)), a[3] & mk);
    int lo = min(min(min(x, a[0]), min(a[1], a[2])), a[3]);
#pragma unroll
    for (int k = 0; k < 8; k++) { hi = max(hi, sb[k] & mk); lo = min(lo, sb[k]); }
    int ap[4], wp[4], as[8], ws[8];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int d = a[k] - x;
        ap[k] = d < 0 ? -d : d;
        wp[k] = d < 0 ? -1 : 1;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int d = sb[k] - x, w = k < 4 ? 2 : 1;
        as[k] = d < 0 ? -d : d;
        ws[k] = d < 0 ? -w : w;
    }
    int S[3] = { 0, 0, 0 }, Z[3] = { 0, 0, 0 };
#pragma unroll
    for (int i = 0; i < 3; i++)
        if ((need >> i) & 1) S[i] = s_taps_sum(as, ws, (1 << i) << CS, max(damping - i - CS, 0));
    if (need & 0x38) {
        const int z[4] = { kSOff[4][0], kSOff[0][0], kSOff[4][1], kSOff[0][1] };
        int az[8], wz[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int d = c[(k & 1) ? -z[k >> 1] : z[k >> 1]] - x, w = k < 4 ? 2 : 1;
            az[k] = d < 0 ? -d : d;
            wz[k] = d < 0 ? -w : w;
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
            if ((need >> (3 + i)) & 1) Z[i] = s_taps_sum(az, wz, (1 << i) << CS, max(damping - i - CS, 0));
    }
#pragma unroll
    for (int q = 0; q < NCAND; q++) {
        const int packed = pk[q];
        const int pri = packed >> 2, sc = packed & 3;   // secondary code 0..3 = strength 0, 1, 2, 4
        int out;
        if (pri) {
            int st = pri << CS;
            if (LUMA) st = has_var ? (((pri << CS) * (4 + adj_i) + 8) >> 4) : 0;
            int sum = sc == 0 ? 0 : sc == 1 ? S[0] : sc == 2 ? S[1] : S[2];
            if (st) {
                const int psh = max(damping - s_ilog2((unsigned)st), 0);
                const int t0 = ((st >> CS) & 1) ? 3 : 4, t1 = ((st >> CS) & 1) ? 3 : 2;
                const int m01 = __mul24(wp[0], s_clamp0(st - (ap[0] >> psh), ap[0])) +
                                __mul24(wp[1], s_clamp0(st - (ap[1] >> psh), ap[1]));
                const int m23 = __mul24(wp[2], s_clamp0(st - (ap[2] >> psh), ap[2])) +
                                __mul24(wp[3], s_clamp0(st - (ap[3] >> psh), ap[3]));
                sum += __mul24(t0, m01) + __mul24(t1, m23);
            }
            out = x + ((8 + sum - (sum < 0)) >> 4);
            if (st && sc) out = min(max(out, lo), hi);
        } else {
            const int sum = sc == 0 ? 0 : sc == 1 ? Z[0] : sc == 2 ? Z[1] : Z[2];
            out = x + ((8 + sum - (sum < 0)) >> 4);
        }
        const int e = srcv - out;
        acc[q] += (uint32_t)__mul24(e, e);
    }
}

// Which shared secondary sums a shortlist needs (see s_pixel_sub).
__device__ __forceinline__ unsigned s_need(const int *pk)
{
    unsigned need = 0;
    for (int q = 0; q < NCAND; q++) {
        const int pri = pk[q] >> 2, sc = pk[q] & 3;
        if (sc) need |= 1u << ((sc == 3 ? 2 : sc - 1) + (pri ? 0 : 3));
    }
    return need;
}

template <int CS>
__global__ __launch_bounds__(256) void cdef_search_subset_kernel(FrameView src, FrameView rec, const uint64_t *noskip,
                                                                 int damping, const int *cand, uint64_t *mse, int nhfb,
                                                                 int nvfb)
{
    using T = typename std::conditional<CS == 0, uint8_t, uint16_t>::type;
    __shared__ uint16_t tile[68 * SP];
    __shared__ uint16_t srcT[64 * 64];
    __shared__ uint8_t bdir[64];
    __shared__ int8_t badj[64];
    __shared__ uint8_t blist[64];
    __shared__ int nblk_s;
    __shared__ uint32_t part[4][NCAND];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int fbc = blockIdx.x, fbr = blockIdx.y;
    const int fb = fbr * nhfb + fbc;
    const int nfb = nhfb * nvfb;
    const int mi_cols = ((rec.w + 7) >> 3) << 1, mi_rows = ((rec.h + 7) >> 3) << 1;
    const int nhb = min(16, mi_cols - 16 * fbc), nvb = min(16, mi_rows - 16 * fbr);
    const uint64_t m = noskip ? noskip[fb] : ~0ull;
    int pk[2][NCAND];
#pragma unroll
    for (int q = 0; q < NCAND; q++) {
        pk[0][q] = kLevel1[cand[q]];
        pk[1][q] = kLevel1[cand[NCAND + q]];
    }
    const unsigned need[2] = { s_need(pk[0]), s_need(pk[1]) };
    s_block_list(m, nhb, nvb, blist, &nblk_s, tid);
    __syncthreads();
    const int nblk = nblk_s;
    uint64_t *m0 = mse + (size_t)fb * NCAND, *m1 = mse + ((size_t)nfb + fb) * NCAND;
    if (nblk == 0) {
        if (tid < NCAND) { m0[tid] = 0; m1[tid] = 0; }
        return;
    }
    uint32_t acc[NCAND];
    for (int pl = 0; pl < 3; pl++) {
        const int sh = pl ? 1 : 0;
        const int x0 = (fbc * 64) >> sh, y0 = (fbr * 64) >> sh, w = (nhb * 4) >> sh, h = (nvb * 4) >> sh;
        if (pl < 2) {
#pragma unroll
            for (int q = 0; q < NCAND; q++) acc[q] = 0;
        }
        __syncthreads();
        if (pl == 0) s_stage<T, 68, 64, false>(tile, rec.p[pl], srcT, src.p[pl], x0, y0, w, h, tid);
        else s_stage<T, 36, 32, false>(tile, rec.p[pl], srcT, src.p[pl], x0, y0, w, h, tid);
        __syncthreads();
        if (pl == 0) {
            if (tid < nblk) {
                const int b = blist[tid];
                int d;
                int32_t v;
                s_find_dir<CS>(tile + (2 + (b >> 3) * 8) * SP + 8 + (b & 7) * 8, &d, &v);
                bdir[b] = (uint8_t)d;
                int i = -1;
                if (v != 0) {
                    i = 0;
                    const int32_t vv = v >> 6;
                    if (vv != 0) i = min(s_ilog2((unsigned)vv), 12);
                }
                badj[b] = (int8_t)i;
            }
            __syncthreads();
            // a wave covers one 8x8 block (64 pixels): its direction and
            // strength adjustment are wave-uniform
            for (int p = tid; p < nblk * 64; p += 256) {
                const int b = __builtin_amdgcn_readfirstlane(blist[p >> 6]), q = p & 63;
                const int d = __builtin_amdgcn_readfirstlane(bdir[b]), adj = __builtin_amdgcn_readfirstlane(badj[b]);
                const int yy = (b >> 3) * 8 + (q >> 3), xx = (b & 7) * 8 + (q & 7);
                s_pixel_sub<true, CS>(acc, tile + (2 + yy) * SP + 8 + xx, srcT[yy * 64 + xx], d, adj, adj >= 0,
                                      damping + CS, pk[0], need[0]);
            }
        } else {
            for (int p = tid; p < nblk * 16; p += 256) {
                const int b = blist[p >> 4], q = p & 15;
                const int yy = (b >> 3) * 4 + (q >> 2), xx = (b & 7) * 4 + (q & 3);
                s_pixel_sub<false, CS>(acc, tile + (2 + yy) * SP + 8 + xx, srcT[yy * 64 + xx], bdir[b], 0, true,
                                       damping - 1 + CS, pk[1], need[1]);
            }
        }
        if (pl == 0 || pl == 2) {
#pragma unroll
            for (int q = 0; q < NCAND; q++) {
                uint32_t v = acc[q];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) part[wv][q] = v;
            }
            __syncthreads();
            if (tid < NCAND) {
                const uint64_t sum = (uint64_t)part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
                if (pl == 0) m0[tid] = sum; else m1[tid] = sum;
            }
        }
    }
}

// Greedy selection over the NCAND x NCAND shortlist pairs.  The four levels
// (cdef_bits 0..3) are independent greedy runs, so workgroup b of 4 runs
// level b, all of its steps in this one launch.  Thread t keeps the active
// blocks t, t + 512, ... (NPER of them) with their 2 * NCAND shortlist SSEs in
// registers, exact as int32: an 8-bit 64x64 luma block's SSE is at most
// 4096 * 255^2 and its two 32x32 chroma blocks' 2048 * 255^2, so a pair's
// cost stays below 2^29.  A step computes each block's best cost among the
// chosen pairs; the gains max(0, best - y_j - uv_k) of the NP = NCAND^2 pairs
// are summed by pair lanes: lane l of a wave serves pair l % NP for the
// (64 / NP)-th part l / NP of the wave's 64 blocks and holds y_j + uv_k of
// those blocks, the blocks' best costs come through LDS (broadcast reads),
// and the parts are added across lanes.  Every wave adds the waves' totals
// (written by step parity, one barrier per step) and takes the first maximum
// as one 64-bit key (gain, 63 - pair).  Each level's
// cost and pairs go to `lv`; the last workgroup to finish (ticket) keeps the
// cheapest level (the lowest on ties, as a sequential scan with a strict
// comparison would) and writes the decision.
struct SmallLevel {
    unsigned long long cost;
    int ys[8], us[8];
};
struct SmallState {
    SmallLevel lv[4];
    unsigned int ticket;
};
static_assert(sizeof(SmallState) <= 1024, "small selection state");
#define SMALL_MAX_FB 2048
#define SMALL_T 512

// A block's NCAND shortlist SSEs of one plane as one register vector:
// indexing it with a wave-uniform (scalar) index stays in registers.
typedef int cand8 __attribute__((ext_vector_type(NCAND)));


// The greedy selection over the shortlist pairs with 64-bit costs, for
// tables whose SSEs need more than 32 bits (10-bit frames: a 64x64 luma
// block's SSE reaches 4096 * 1023^2).  Same steps, the same first-maximum
// and first-minimum rules, the same RD cost and ticket as fin_small_kernel,
// which runs it instead of its own steps when a table needs it;
// thread t keeps the active blocks t, t + SMALL_T, ... (NPER of them) with
// their shortlist SSEs in registers and sums every pair's gain over them.
template <int NPER>
__device__ __noinline__ void fin_small_wide(const uint64_t *mse, const uint64_t *noskip, int nhfb, int nvfb,
                                               int mi_cols, int mi_rows, unsigned long long lambda, const int *cand,
                                               SmallState *ss, av1pf_cdef_decision *dec, uint8_t *fb_index)
{
    constexpr int NW = SMALL_T / 64;
    constexpr int NP = NCAND * NCAND;
    __shared__ uint16_t act[SMALL_MAX_FB];
    __shared__ uint8_t flag[SMALL_MAX_FB];
    __shared__ unsigned long long wpart[NW][NP + 1];   // per-wave pair gains (or column sums) and base sum
    __shared__ int pick_s, nact_s, bits_s;
    __shared__ unsigned sel_s[2];
    __shared__ bool last;
    __shared__ unsigned long long dist_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nfb = nhfb * nvfb;
    for (int fb = t; fb < nfb; fb += SMALL_T) flag[fb] = fb_active(noskip, fb, nhfb, mi_cols, mi_rows) ? 1 : 0;
    __syncthreads();
    if (t < 64) {
        int n = 0;
        for (int base = 0; base < nfb; base += 64) {
            const bool a = base + lane < nfb && flag[base + lane];
            const uint64_t vote = __ballot(a);
            if (a) act[n + __popcll(vote & ((1ull << lane) - 1))] = (uint16_t)(base + lane);
            n += __popcll(vote);
        }
        if (lane == 0) nact_s = n;
    }
    __syncthreads();
    const int nact = nact_s;
    // blocks past nact hold zeros: best cost 0, no gain, nothing in a column sum
    unsigned long long Y[NPER][NCAND], U[NPER][NCAND];
#pragma unroll
    for (int r = 0; r < NPER; r++) {
        const int i = t + r * SMALL_T;
        const size_t fb = i < nact ? act[i] : 0;
#pragma unroll
        for (int q = 0; q < NCAND; q++) {
            Y[r][q] = i < nact ? mse[fb * NCAND + q] : 0;
            U[r][q] = i < nact ? mse[((size_t)nfb + fb) * NCAND + q] : 0;
        }
    }
    const int bits = blockIdx.x;
    const int nb = 1 << bits;
    unsigned sy = 0, su = 0;
    for (int step = 0; step < 5 * nb; step++) {
        const int slot = step < nb ? step : nb - 1;
        const bool final_step = step == 5 * nb - 1;
        if (slot == 0) {
#pragma unroll
            for (int q = 0; q < 2 * NCAND; q++) {
                unsigned long long v = 0;
#pragma unroll
                for (int r = 0; r < NPER; r++) v += q < NCAND ? Y[r][q] : U[r][q - NCAND];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) wpart[wv][q] = v;
            }
            __syncthreads();
            if (t == 0) {
                unsigned long long cs[2 * NCAND];
                for (int q = 0; q < 2 * NCAND; q++) {
                    cs[q] = 0;
                    for (int w = 0; w < NW; w++) cs[q] += wpart[w][q];
                }
                int by = 0, bc = 0;
                for (int q = 1; q < NCAND; q++) {
                    if (cs[q] < cs[by]) by = q;
                    if (cs[NCAND + q] < cs[NCAND + bc]) bc = q;
                }
                pick_s = by * NCAND + bc;
                dist_s = nact ? cs[by] + cs[NCAND + bc] : 0;
            }
        } else {
            unsigned long long b[NPER], bsum = 0;
#pragma unroll
            for (int r = 0; r < NPER; r++) {
                b[r] = Y[r][sy & 15] + U[r][su & 15];
                for (int g = 1; g < slot; g++)
                    b[r] = min(b[r], Y[r][(sy >> (4 * g)) & 15] + U[r][(su >> (4 * g)) & 15]);
                bsum += b[r];
            }
#pragma unroll
            for (int p = 0; p < NP; p++) {
                unsigned long long g = 0;
#pragma unroll
                for (int r = 0; r < NPER; r++) {
                    const unsigned long long c = Y[r][p / NCAND] + U[r][p % NCAND];
                    g += b[r] > c ? b[r] - c : 0;
                }
                for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o, 64);
                if (lane == 0) wpart[wv][p] = g;
            }
            if (final_step) {
                for (int o = 32; o > 0; o >>= 1) bsum += __shfl_xor(bsum, o, 64);
                if (lane == 0) wpart[wv][NP] = bsum;
            }
            __syncthreads();
            if (t == 0) {   // first maximum over the pairs
                unsigned long long bestg = 0;
                int bp = 0;
                for (int p = 0; p < NP; p++) {
                    unsigned long long g = 0;
                    for (int w = 0; w < NW; w++) g += wpart[w][p];
                    if (p == 0 || g > bestg) { bestg = g; bp = p; }
                }
                pick_s = bp;
                if (final_step) {
                    unsigned long long base = 0;
                    for (int w = 0; w < NW; w++) base += wpart[w][NP];
                    dist_s = base - bestg;
                }
            }
        }
        __syncthreads();
        const unsigned pk = (unsigned)__builtin_amdgcn_readfirstlane(pick_s);
        sy = (sy & ~(15u << (4 * slot))) | ((pk / NCAND) << (4 * slot));
        su = (su & ~(15u << (4 * slot))) | ((pk % NCAND) << (4 * slot));
        if (final_step) {
            if (t == 0) {
                const long long rate = (long long)(nact * bits + nb * 12) << 9;
                SmallLevel &L = ss->lv[bits];
                L.cost = (unsigned long long)((rate * (long long)lambda + (1 << 8)) >> 9) + ((dist_s * 16) << 7);
                for (int i = 0; i < 8; i++) {
                    L.ys[i] = i < nb ? (int)((sy >> (4 * i)) & 15) : 0;
                    L.us[i] = i < nb ? (int)((su >> (4 * i)) & 15) : 0;
                }
            }
        } else if (step + 1 >= nb) {
            sy >>= 4;
            su >>= 4;
        }
        __syncthreads();   // pick_s and the partials are rewritten by the next step
    }
    if (t == 0) {
        __threadfence();
        last = atomicAdd(&ss->ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    if (t == 0) {
        __threadfence();
        volatile SmallLevel *V = ss->lv;
        int best_bits = 0;
        unsigned long long best_cost = V[0].cost;
        for (int b = 1; b < 4; b++)
            if (V[b].cost < best_cost) { best_cost = V[b].cost; best_bits = b; }
        const int n = 1 << best_bits;
        unsigned a = 0, c = 0;
        for (int i = 0; i < n; i++) { a |= (unsigned)V[best_bits].ys[i] << (4 * i); c |= (unsigned)V[best_bits].us[i] << (4 * i); }
        ss->ticket = 0;
        dec->cdef_bits = best_bits;
        for (int i = 0; i < 8; i++) {
            dec->y_strength[i] = i < n ? kLevel1[cand[V[best_bits].ys[i]]] : 0;
            dec->uv_strength[i] = i < n ? kLevel1[cand[NCAND + V[best_bits].us[i]]] : 0;
        }
        bits_s = best_bits;
        sel_s[0] = a;
        sel_s[1] = c;
    }
    for (int fb = t; fb < nfb; fb += SMALL_T) fb_index[fb] = 0;
    __syncthreads();
    const int n = 1 << bits_s;
    const unsigned fy = (unsigned)__builtin_amdgcn_readfirstlane(sel_s[0]), fu = (unsigned)__builtin_amdgcn_readfirstlane(sel_s[1]);
#pragma unroll
    for (int r = 0; r < NPER; r++) {
        const int i = t + r * SMALL_T;
        if (i >= nact) continue;
        unsigned long long bm = Y[r][fy & 15] + U[r][fu & 15];
        int bg = 0;
        for (int g = 1; g < n; g++) {
            const unsigned long long c = Y[r][(fy >> (4 * g)) & 15] + U[r][(fu >> (4 * g)) & 15];
            if (c < bm) { bm = c; bg = g; }
        }
        fb_index[act[i]] = (uint8_t)bg;
    }
}

template <int NPER, bool WIDE_CHECK>
__global__ __launch_bounds__(SMALL_T) void fin_small_kernel(const uint64_t *mse, const uint64_t *noskip, int nhfb,
                                                            int nvfb, int mi_cols, int mi_rows,
                                                            unsigned long long lambda, const int *cand,
                                                            SmallState *ss, av1pf_cdef_decision *dec,
                                                            uint8_t *fb_index)
{
    constexpr int NW = SMALL_T / 64;
    constexpr int NP = NCAND * NCAND, BPL = NP;   // pairs; blocks per lane = 64 / (64 / NP) lanes per pair
    static_assert(64 % NP == 0 && BPL % 8 == 0, "pair lanes need NCAND^2 to divide 64 and 8 | NCAND^2");
    __shared__ uint16_t act[SMALL_MAX_FB];
    __shared__ unsigned long long amask[SMALL_MAX_FB / 64];
    __shared__ int aoff[SMALL_MAX_FB / 64];
    // per-wave pair gains (or column sums) and base sum, by step parity: every
    // wave reads all of them after the step's one barrier, and a wave can be
    // one step ahead of another
    __shared__ unsigned long long wpart[2][NW][65];
    __shared__ __align__(16) uint32_t bw[NPER * SMALL_T];   // each block's best cost, for the pair lanes
    __shared__ __align__(16) uint32_t yu[NPER * SMALL_T][2 * NCAND];   // the blocks' shortlist SSEs
    __shared__ int pick_s, nact_s, bits_s;
    __shared__ unsigned sel_s[2];
    __shared__ bool last;
    __shared__ unsigned long long dist_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nfb = nhfb * nvfb;
    // ascending compaction: a ballot per 64 blocks, the chunks' offsets by one
    // wave's scan, then every block writes its own entry
    for (int c0 = wv; c0 * 64 < nfb; c0 += NW) {
        const int fb = c0 * 64 + lane;
        const uint64_t vote = __ballot(fb < nfb && fb_active(noskip, fb, nhfb, mi_cols, mi_rows));
        if (lane == 0) amask[c0] = vote;
    }
    __syncthreads();
    if (wv == 0) {
        const int nch = (nfb + 63) >> 6;
        const int cnt = lane < nch ? __popcll(amask[lane]) : 0;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane < SMALL_MAX_FB / 64) aoff[lane] = incl - cnt;
        if (lane == 63) nact_s = incl;
    }
    __syncthreads();
    for (int c0 = wv; c0 * 64 < nfb; c0 += NW) {
        const uint64_t m = amask[c0];
        if ((m >> lane) & 1) act[aoff[c0] + __popcll(m & ((1ull << lane) - 1))] = (uint16_t)(c0 * 64 + lane);
    }
    __syncthreads();
    const int nact = nact_s;
    // blocks past nact hold zeros: best cost 0, no gain, nothing in a column sum
    // every block's shortlist SSEs, a block's loads in flight before use
    cand8 Y[NPER], U[NPER];
    bool big = false;
#pragma unroll
    for (int r = 0; r < NPER; r++) {
        const int i = t + r * SMALL_T;
        const size_t fb = i < nact ? act[i] : 0;
        uint64_t my[NCAND], mu[NCAND];
#pragma unroll
        for (int q = 0; q < NCAND; q++) {
            my[q] = mse[fb * NCAND + q];
            mu[q] = mse[((size_t)nfb + fb) * NCAND + q];
        }
#pragma unroll
        for (int q = 0; q < NCAND; q++) {
            const uint64_t a = i < nact ? my[q] : 0, c = i < nact ? mu[q] : 0;
            Y[r][q] = (int)a;
            U[r][q] = (int)c;
            big |= ((a | c) >> 28) != 0;
            yu[i][q] = (uint32_t)a;
            yu[i][NCAND + q] = (uint32_t)c;
        }
    }
    // the 32-bit arithmetic below is exact when every SSE is below 2^28 (a
    // pair's cost below 2^29, eight gains below 2^32): always for 8-bit
    // frames; a 10-bit table with a larger value takes the 64-bit steps
    if (WIDE_CHECK) {
        if (__syncthreads_or(big)) {
            fin_small_wide<NPER>(mse, noskip, nhfb, nvfb, mi_cols, mi_rows, lambda, cand, ss, dec, fb_index);
            return;
        }
    }
    // pair lanes: lane l of wave w serves pair pr = 8 j + k (j, k < NCAND) for
    // the blocks w * 64 + part * BPL + m (+ 512 r), m < BPL, and holds their
    // y_j + uv_k
    const int pr = lane % NP, part = lane / NP;
    __syncthreads();   // the shortlist SSEs of every block are in LDS
    uint32_t YU[NPER][BPL];
#pragma unroll
    for (int r = 0; r < NPER; r++)
#pragma unroll
        for (int m = 0; m < BPL; m++) {
            const int i = r * SMALL_T + wv * 64 + part * BPL + m;
            YU[r][m] = yu[i][pr / NCAND] + yu[i][NCAND + pr % NCAND];   // zeros past nact
        }
    const int bits = blockIdx.x;
    const int nb = 1 << bits;
    // the chosen pairs, 4 bits per slot, kept by every thread (wave-uniform)
    unsigned sy = 0, su = 0;
    for (int step = 0; step < 5 * nb; step++) {
        const int slot = step < nb ? step : nb - 1;
        const bool final_step = step == 5 * nb - 1;
        if (slot == 0) {
            // column sums, one wave reduction per shortlist column
#pragma unroll
            for (int q = 0; q < 2 * NCAND; q++) {
                unsigned long long v = 0;
#pragma unroll
                for (int r = 0; r < NPER; r++) v += (unsigned)(q < NCAND ? Y[r][q] : U[r][q - NCAND]);
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) wpart[0][wv][q] = v;
            }
            __syncthreads();
            if (t == 0) {
                unsigned long long cs[2 * NCAND];
                for (int q = 0; q < 2 * NCAND; q++) {
                    cs[q] = 0;
                    for (int w = 0; w < NW; w++) cs[q] += wpart[0][w][q];
                }
                int by = 0, bc = 0;
                for (int q = 1; q < NCAND; q++) {
                    if (cs[q] < cs[by]) by = q;
                    if (cs[NCAND + q] < cs[NCAND + bc]) bc = q;
                }
                pick_s = by * NCAND + bc;
                dist_s = nact ? cs[by] + cs[NCAND + bc] : 0;
            }
        } else {
            // the prior pairs as a wave-uniform bit set over the NP pairs;
            // every block's best cost is the minimum over that set, read
            // with constant register indices
            unsigned pm = 0;
            for (int g = 0; g < slot; g++) pm |= 1u << (((sy >> (4 * g)) & 15) * NCAND + ((su >> (4 * g)) & 15));
            pm = (unsigned)__builtin_amdgcn_readfirstlane((int)pm);
            unsigned b[NPER];
            unsigned long long bsum = 0;
#pragma unroll
            for (int r = 0; r < NPER; r++) {
                b[r] = ~0u;
#pragma unroll
                for (int q = 0; q < NP; q++)
                    if ((pm >> q) & 1) b[r] = min(b[r], (unsigned)(Y[r][q / NCAND] + U[r][q % NCAND]));
                bsum += b[r];
            }
#pragma unroll
            for (int r = 0; r < NPER; r++) bw[r * SMALL_T + t] = b[r];
            __builtin_amdgcn_wave_barrier();   // the wave reads back only its own blocks' costs
            unsigned long long gain = 0;
#pragma unroll
            for (int r = 0; r < NPER; r++)
#pragma unroll
                for (int m = 0; m < BPL; m += 8) {   // 8 gains below 2^29 each fit 32 bits
                    const uint32_t *bp = &bw[r * SMALL_T + wv * 64 + part * BPL + m];
                    const uint4 b0 = *reinterpret_cast<const uint4 *>(bp);
                    const uint4 b1 = *reinterpret_cast<const uint4 *>(bp + 4);
                    uint32_t g = __builtin_elementwise_sub_sat(b0.x, YU[r][m]) +
                                 __builtin_elementwise_sub_sat(b0.y, YU[r][m + 1]) +
                                 __builtin_elementwise_sub_sat(b0.z, YU[r][m + 2]) +
                                 __builtin_elementwise_sub_sat(b0.w, YU[r][m + 3]);
                    g += __builtin_elementwise_sub_sat(b1.x, YU[r][m + 4]) +
                         __builtin_elementwise_sub_sat(b1.y, YU[r][m + 5]) +
                         __builtin_elementwise_sub_sat(b1.z, YU[r][m + 6]) +
                         __builtin_elementwise_sub_sat(b1.w, YU[r][m + 7]);
                    gain += g;
                }
            // add the parts: the lanes of a pair are NP apart
            if constexpr (NP <= 4) gain += xchg64<4>(gain);
            if constexpr (NP <= 8) gain += xchg64<8>(gain);
            if constexpr (NP <= 16) gain += xchg64<16>(gain);
            if constexpr (NP <= 32) gain += xchg64<32>(gain);
            const int ph = step & 1;
            if (final_step) {   // the frame distortion is needed at the last step only
                for (int o = 32; o > 0; o >>= 1) bsum += __shfl_xor(bsum, o, 64);
                if (lane == 0) wpart[ph][wv][64] = bsum;
            }
            wpart[ph][wv][lane] = gain;
            __syncthreads();
            // every wave: the first maximum over the pairs, the largest
            // (gain, 63 - pair) key
            unsigned long long v = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) v += wpart[ph][w][lane];
            // the lanes of a pair hold the same total: the maximum over the
            // first NP lanes' neighbourhoods is the wave's
            unsigned long long key = (v << 6) | (unsigned long long)(63 - pr);
            key = max(key, xchg64<1>(key));
            key = max(key, xchg64<2>(key));
            if constexpr (NP > 4) key = max(key, xchg64<4>(key));
            if constexpr (NP > 8) key = max(key, xchg64<8>(key));
            if constexpr (NP > 16) key = max(key, xchg64<16>(key));
            if constexpr (NP > 32) key = max(key, xchg64<32>(key));
            if (t == 0 && final_step) {
                unsigned long long base = 0;
                for (int w = 0; w < NW; w++) base += wpart[ph][w][64];
                dist_s = base - (key >> 6);
            }
            const unsigned pk0 = 63 - (unsigned)(__builtin_amdgcn_readfirstlane((int)(uint32_t)key) & 63);
            sy = (sy & ~(15u << (4 * slot))) | ((pk0 / NCAND) << (4 * slot));
            su = (su & ~(15u << (4 * slot))) | ((pk0 % NCAND) << (4 * slot));
        }
        if (slot == 0) {
            __syncthreads();   // pick_s of the column sums
            const unsigned pk = (unsigned)__builtin_amdgcn_readfirstlane(pick_s);
            sy = (sy & ~15u) | (pk / NCAND);
            su = (su & ~15u) | (pk % NCAND);
        }
        if (final_step) {
            if (t == 0) {
                const long long rate = (long long)(nact * bits + nb * 12) << 9;
                SmallLevel &L = ss->lv[bits];
                L.cost = (unsigned long long)((rate * (long long)lambda + (1 << 8)) >> 9) + ((dist_s * 16) << 7);
                for (int i = 0; i < 8; i++) {
                    L.ys[i] = i < nb ? (int)((sy >> (4 * i)) & 15) : 0;
                    L.us[i] = i < nb ? (int)((su >> (4 * i)) & 15) : 0;
                }
            }
        } else if (step + 1 >= nb) {   // refinement: the next step drops the oldest pair
            sy >>= 4;
            su >>= 4;
        }
        // the partials alternate by step parity, and a wave passes the next
        // step's barrier only when every wave has read this step's
    }
    if (t == 0) {
        __threadfence();
        last = atomicAdd(&ss->ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    if (t == 0) {
        __threadfence();
        volatile SmallLevel *V = ss->lv;
        int best_bits = 0;
        unsigned long long best_cost = V[0].cost;
        for (int b = 1; b < 4; b++)
            if (V[b].cost < best_cost) { best_cost = V[b].cost; best_bits = b; }
        const int n = 1 << best_bits;
        unsigned a = 0, c = 0;
        for (int i = 0; i < n; i++) { a |= (unsigned)V[best_bits].ys[i] << (4 * i); c |= (unsigned)V[best_bits].us[i] << (4 * i); }
        ss->ticket = 0;
        dec->cdef_bits = best_bits;
        for (int i = 0; i < 8; i++) {
            dec->y_strength[i] = i < n ? kLevel1[cand[V[best_bits].ys[i]]] : 0;
            dec->uv_strength[i] = i < n ? kLevel1[cand[NCAND + V[best_bits].us[i]]] : 0;
        }
        bits_s = best_bits;
        sel_s[0] = a;
        sel_s[1] = c;
    }
    for (int fb = t; fb < nfb; fb += SMALL_T) fb_index[fb] = 0;
    __syncthreads();
    const int n = 1 << bits_s;
    const unsigned fy = (unsigned)__builtin_amdgcn_readfirstlane(sel_s[0]), fu = (unsigned)__builtin_amdgcn_readfirstlane(sel_s[1]);
#pragma unroll
    for (int r = 0; r < NPER; r++) {
        const int i = t + r * SMALL_T;
        if (i >= nact) continue;
        int bm = Y[r][fy & 15] + U[r][fu & 15], bg = 0;
        for (int g = 1; g < n; g++) {
            const int c = Y[r][(fy >> (4 * g)) & 15] + U[r][(fu >> (4 * g)) & 15];
            if (c < bm) { bm = c; bg = g; }
        }
        fb_index[act[i]] = (uint8_t)bg;
    }
}

hipError_t launch_search_shortlist(const FrameView &src, const FrameView &rec, const uint64_t *noskip, int damping,
                                   uint64_t lambda, const int *cand, uint64_t *mse_sub, av1pf_cdef_decision *dec,
                                   uint8_t *fb_index, void *small_state, hipStream_t s)
{
    const int mi_cols = ((rec.w + 7) >> 3) << 1, mi_rows = ((rec.h + 7) >> 3) << 1;
    const int nhfb = (mi_cols + 15) / 16, nvfb = (mi_rows + 15) / 16;
    if (rec.bd == 8)
        hipLaunchKernelGGL(cdef_search_subset_kernel<0>, dim3(nhfb, nvfb), dim3(256), 0, s, src, rec, noskip, damping,
                           cand, mse_sub, nhfb, nvfb);
    else
        hipLaunchKernelGGL(cdef_search_subset_kernel<2>, dim3(nhfb, nvfb), dim3(256), 0, s, src, rec, noskip, damping,
                           cand, mse_sub, nhfb, nvfb);
    const int nfb = nhfb * nvfb;
    SmallState *ss = static_cast<SmallState *>(small_state);
    const unsigned long long lam = lambda;
    if (rec.bd == 8) {
        auto fin = nfb <= SMALL_T ? fin_small_kernel<1, false> : nfb <= 2 * SMALL_T ? fin_small_kernel<2, false>
                                                                                    : fin_small_kernel<4, false>;
        hipLaunchKernelGGL(fin, dim3(4), dim3(SMALL_T), 0, s, mse_sub, noskip, nhfb, nvfb, mi_cols, mi_rows, lam, cand,
                           ss, dec, fb_index);
    } else {
        // 10-bit: the 32-bit steps when every SSE allows them, else the 64-bit ones
        auto fin = nfb <= SMALL_T ? fin_small_kernel<1, true> : nfb <= 2 * SMALL_T ? fin_small_kernel<2, true>
                                                                                   : fin_small_kernel<4, true>;
        hipLaunchKernelGGL(fin, dim3(4), dim3(SMALL_T), 0, s, mse_sub, noskip, nhfb, nvfb, mi_cols, mi_rows, lam, cand,
                           ss, dec, fb_index);
    }
    return hipGetLastError();
}

hipError_t launch_cdef_search(const FrameView &src, const FrameView &rec, const uint64_t *noskip,
                              int damping, uint64_t *mse, int nhfb, int nvfb, hipStream_t s)
{
    if (rec.bd == 8)
        hipLaunchKernelGGL(cdef_search_kernel<0>, dim3(nhfb, nvfb, 2), dim3(256), 0, s, src, rec, noskip, damping, mse,
                           nhfb, nvfb);
    else
        hipLaunchKernelGGL(cdef_search_kernel<2>, dim3(nhfb, nvfb, 2), dim3(256), 0, s, src, rec, noskip, damping, mse,
                           nhfb, nvfb);
    return hipGetLastError();
}

}  // namespace pf
