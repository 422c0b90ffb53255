// AV1 super-resolution upscale on gfx950 (superres/superres.go:38-107,
// decoder/postfilter_superres.go:198-264).
//
// Every plane row is upscaled independently from the coded width to the
// upscaled width with the normative 8-tap polyphase filter; the source span is
// the MI-aligned coded width, indexes clamped to it.  One workgroup takes a
// tile of SR_COLS output columns by SR_ROWS * SR_CHUNKS rows.  The source
// samples the tile reads (at most SR_COLS + 7 per row, since the step is below
// one sample) are staged in LDS with 16-byte loads, every load of the tile in
// flight before the first LDS write; each thread then filters two adjacent
// output columns in every row from LDS with their phases' taps in registers
// (packed 16-bit dot products) and stores the pair as one word.  The pass is HBM bound: it reads the coded
// plane once and writes the upscaled plane once.
#include "pf_internal.h"
#include "superres_taps.h"

namespace pf {

constexpr int SR_TROWS = SR_ROWS * SR_CHUNKS;   // rows per tile

typedef short sr_pair __attribute__((ext_vector_type(2)));   // two int16 lanes of a word
__device__ __forceinline__ sr_pair sr_as_pair(uint32_t v) { return __builtin_bit_cast(sr_pair, v); }
__device__ __forceinline__ sr_pair sr_as_pair(int32_t v) { return __builtin_bit_cast(sr_pair, v); }
constexpr int SR_SPAN = 544;                // staged samples per row: the <= 520 the tile reads + alignment slack

// Stage rows [y0, y0 + rows) of the tile's source span into LDS, starting at
// sample lb (lb <= s0, lb * bps 16-byte aligned relative to the row start).
// Every thread issues its 16-byte loads for all rows before writing LDS.  A
// 16-byte chunk wholly inside [0, srcw) is one vector load; chunks crossing an
// edge read their samples one by one with the clamp.
template <typename T>
__device__ __forceinline__ void sr_stage(const PlaneView &s, const SrPlane &g, int y0, int rows, int lb, int nchunk,
                                         uint16_t (*L)[SR_SPAN])
{
    constexpr int per = 16 / sizeof(T);                  // samples per chunk
    constexpr int kMax = (SR_TROWS * SR_SPAN / per + 255) / 256;
    const int total = rows * nchunk;
    uint4 v[kMax];
#pragma unroll
    for (int k = 0; k < kMax; k++) {
        const int i = threadIdx.x + k * 256;
        if (i < total) {
            const int r = i / nchunk, q = lb + (i - r * nchunk) * per;
            const T *row = row_ptr<T>(s, y0 + r);
            if (q >= 0 && q + per <= g.srcw) {
                v[k] = *reinterpret_cast<const uint4 *>(row + q);
            } else {
                uint32_t w[4] = { 0, 0, 0, 0 };
#pragma unroll
                for (int j = 0; j < per; j++)
                    w[j / (per / 4)] |= (uint32_t)row[min(max(q + j, 0), g.srcw - 1)] << (8 * sizeof(T) * (j % (per / 4)));
                v[k] = make_uint4(w[0], w[1], w[2], w[3]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kMax; k++) {
        const int i = threadIdx.x + k * 256;
        if (i < total) {
            const int r = i / nchunk, c = (i - r * nchunk) * per;
            if constexpr (sizeof(T) == 2) {
                *reinterpret_cast<uint4 *>(&L[r][c]) = v[k];
            } else {
                const uint32_t w[4] = { v[k].x, v[k].y, v[k].z, v[k].w };
                uint32_t o[8];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    o[2 * j] = (w[j] & 0xffu) | ((w[j] << 8) & 0xff0000u);
                    o[2 * j + 1] = ((w[j] >> 16) & 0xffu) | ((w[j] >> 8) & 0xff0000u);
                }
                *reinterpret_cast<uint4 *>(&L[r][c]) = make_uint4(o[0], o[1], o[2], o[3]);
                *reinterpret_cast<uint4 *>(&L[r][c + 8]) = make_uint4(o[4], o[5], o[6], o[7]);
            }
        }
    }
}

// A tile is SR_COLS output columns by SR_TROWS rows.  Thread t filters output
// columns X0 + 2t and X0 + 2t + 1 in every row from LDS (taps in registers)
// and stores both samples together when the destination row allows it.
template <typename T>
__device__ __forceinline__ void sr_tile(const PlaneView &s, const PlaneView &d, const SrPlane &g, int maxv, int tx,
                                        int ty, uint16_t (*L)[SR_SPAN])
{
    constexpr int per = 16 / sizeof(T);
    const int X0 = tx * SR_COLS, Y0 = ty * SR_TROWS;
    const int rows = min(SR_TROWS, g.rows - Y0);
    const int p0 = -(1 << kSrScaleBits) + g.x0 + X0 * g.step;   // position of column X0
    const int s0 = (p0 >> kSrScaleBits) - kSrTapOffset;          // first source sample of the tile
    const int f0 = p0 & kSrScaleMask;
    const int span = ((f0 + (SR_COLS - 1) * g.step) >> kSrScaleBits) + 8;
    // first staged sample: s0 rounded down to a whole 16-byte chunk when every
    // row start is 16-byte aligned (the vector loads need it); a plane with
    // unaligned rows is staged from s0 sample by sample
    const bool aligned = ((reinterpret_cast<uintptr_t>(s.base) | (uintptr_t)s.stride) & 15) == 0;
    const int lb = aligned ? s0 - ((s0 % per) + per) % per : s0;
    const int nchunk = (s0 + span - lb + per - 1) / per;
    if (aligned) sr_stage<T>(s, g, Y0, rows, lb, nchunk, L);
    else {
        for (int i = threadIdx.x; i < rows * nchunk * per; i += 256) {
            const int r = i / (nchunk * per), c = i - r * nchunk * per;
            L[r][c] = row_ptr<T>(s, Y0 + r)[min(max(lb + c, 0), g.srcw - 1)];
        }
    }
    __syncthreads();
    const int xa = X0 + 2 * threadIdx.x;
    if (xa >= g.dstw) return;
    // The two columns' 8-tap windows start at samples oa and ob = oa or oa + 1
    // (the step is below one sample); both lie in the 5 words from word oa / 2
    // of the staged row.  Each window is 4 sample pairs, cut from neighbouring
    // words by a 0- or 16-bit funnel shift, and a pair of taps times a pair of
    // samples is one packed 16-bit dot product.
    const int ra = f0 + 2 * threadIdx.x * g.step, rb = ra + g.step;
    const int oa = (ra >> kSrScaleBits) + (s0 - lb), ob = (rb >> kSrScaleBits) + (s0 - lb);
    const int ka = oa >> 1;
    const uint32_t sha = (oa & 1) * 16, shb = (ob & 1) * 16;
    const bool db = (ob >> 1) != ka;
    const int32_t *ta = kSrTaps[(ra & kSrScaleMask) >> kSrPhaseShift];
    const int32_t *tb = kSrTaps[(rb & kSrScaleMask) >> kSrPhaseShift];
    const sr_pair a0 = sr_as_pair(ta[0]), a1 = sr_as_pair(ta[1]), a2 = sr_as_pair(ta[2]), a3 = sr_as_pair(ta[3]);
    const sr_pair b0 = sr_as_pair(tb[0]), b1 = sr_as_pair(tb[1]), b2 = sr_as_pair(tb[2]), b3 = sr_as_pair(tb[3]);
    const bool both = xa + 1 < g.dstw;
    const bool pair = both && ((reinterpret_cast<uintptr_t>(d.base) | (uintptr_t)d.stride) % (2 * sizeof(T))) == 0;
    for (int r = 0; r < rows; r++) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(&L[r][0]) + ka;
        const uint32_t D0 = w[0], D1 = w[1], D2 = w[2], D3 = w[3], D4 = w[4];
        int sa = __builtin_amdgcn_sdot2(a0, sr_as_pair(__builtin_amdgcn_alignbit(D1, D0, sha)), 0, false);
        sa = __builtin_amdgcn_sdot2(a1, sr_as_pair(__builtin_amdgcn_alignbit(D2, D1, sha)), sa, false);
        sa = __builtin_amdgcn_sdot2(a2, sr_as_pair(__builtin_amdgcn_alignbit(D3, D2, sha)), sa, false);
        sa = __builtin_amdgcn_sdot2(a3, sr_as_pair(__builtin_amdgcn_alignbit(D4, D3, sha)), sa, false);
        const uint32_t E0 = db ? D1 : D0, E1 = db ? D2 : D1, E2 = db ? D3 : D2, E3 = db ? D4 : D3;
        int sb = __builtin_amdgcn_sdot2(b0, sr_as_pair(__builtin_amdgcn_alignbit(E1, E0, shb)), 0, false);
        sb = __builtin_amdgcn_sdot2(b1, sr_as_pair(__builtin_amdgcn_alignbit(E2, E1, shb)), sb, false);
        sb = __builtin_amdgcn_sdot2(b2, sr_as_pair(__builtin_amdgcn_alignbit(E3, E2, shb)), sb, false);
        sb = __builtin_amdgcn_sdot2(b3, sr_as_pair(__builtin_amdgcn_alignbit(D4, E3, shb)), sb, false);
        T *out = row_ptr<T>(d, Y0 + r) + xa;
        const int va = sr_round_clamp(sa, maxv), vb = sr_round_clamp(sb, maxv);
        if (pair) {
            if constexpr (sizeof(T) == 2) *reinterpret_cast<uint32_t *>(out) = (uint32_t)va | ((uint32_t)vb << 16);
            else *reinterpret_cast<uint16_t *>(out) = (uint16_t)(va | (vb << 8));
        } else {
            out[0] = (T)va;
            if (both) out[1] = (T)vb;
        }
    }
}

// LDS of a super-resolution tile: its source rows.
struct __align__(16) SrLds {
    uint16_t rows[SR_TROWS][SR_SPAN];
};

// Tile (tx, ty) of plane pl of a frame.
__device__ __forceinline__ void sr_tile_of(const FrameDesc &D, int pl, int tx, int ty, SrLds &L)
{
    const FrameView &S = D.v[D.sr_src];
    const FrameView &O = D.v[D.sr_dst];
    const int maxv = (1 << S.bd) - 1;
    if (S.bd == 8) sr_tile<uint8_t>(S.p[pl], O.p[pl], D.sr[pl], maxv, tx, ty, L.rows);
    else sr_tile<uint16_t>(S.p[pl], O.p[pl], D.sr[pl], maxv, tx, ty, L.rows);
}

__global__ __launch_bounds__(256) void superres_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs,
                                                       int nseg)
{
    __shared__ SrLds L;
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const int t = blockIdx.x - sg.base;
    sr_tile_of(descs[sg.frame], sg.plane, t % sg.nx, t / sg.nx, L);
}

hipError_t launch_superres(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(superres_kernel, dim3(total), dim3(256), 0, s, d, segs, nseg);
    return hipGetLastError();
}

}  // namespace pf
