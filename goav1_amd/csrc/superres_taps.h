// AV1 super-resolution: the normative 8-tap upscale filter and the position
// arithmetic of one output sample (superres/superres.go:38-60, 78-107).
//
// Output column x of a plane upscaled from coded width cw to width uw reads
// source position p(x) = -(1 << 14) + x0 + x * step in 1/16384 samples, with
//   step = ((cw << 14) + uw / 2) / uw,
//   x0   = ((-((uw - cw) << 13) + uw / 2) / uw + (1 << 7) - (uw * step - (cw << 14)) / 2) & 0x3fff;
// the 8 taps start at sample (p >> 14) - 3, phase (p & 0x3fff) >> 8 picks the
// filter, and source indexes clamp to [0, srcw - 1] where srcw is the
// MI-aligned coded width (the coded frame carries real samples up to it).
// The output is round2(sum, 7) clamped to the sample range.
#pragma once
#include <stdint.h>
#include "pf_internal.h"

namespace pf {

constexpr int kSrScaleBits = 14;
constexpr int kSrScaleMask = (1 << kSrScaleBits) - 1;
constexpr int kSrPhaseShift = 8;   // scale bits - filter bits
constexpr int kSrTapOffset = 3;

// Per phase: 8 taps summing to 128; phase 64 - k is phase k mirrored.  The
// values are AV1's normative upscale filter (the table superres.go:144-180
// holds), packed two taps per int32 as (hi << 16) | (lo & 0xffff).
#define SR_PAIR(a, b) (int32_t)(((uint32_t)(uint16_t)(int16_t)(b) << 16) | (uint16_t)(int16_t)(a))
#define SR_ROW(a, b, c, d, e, f, g, h) { SR_PAIR(a, b), SR_PAIR(c, d), SR_PAIR(e, f), SR_PAIR(g, h) }
static __device__ const int32_t kSrTaps[64][4] = {
    SR_ROW(0, 0, 0, 128, 0, 0, 0, 0),       SR_ROW(0, 0, -1, 128, 2, -1, 0, 0),
    SR_ROW(0, 1, -3, 127, 4, -2, 1, 0),     SR_ROW(0, 1, -4, 127, 6, -3, 1, 0),
    SR_ROW(0, 2, -6, 126, 8, -3, 1, 0),     SR_ROW(0, 2, -7, 125, 11, -4, 1, 0),
    SR_ROW(-1, 2, -8, 125, 13, -5, 2, 0),   SR_ROW(-1, 3, -9, 124, 15, -6, 2, 0),
    SR_ROW(-1, 3, -10, 123, 18, -6, 2, -1), SR_ROW(-1, 3, -11, 122, 20, -7, 3, -1),
    SR_ROW(-1, 4, -12, 121, 22, -8, 3, -1), SR_ROW(-1, 4, -13, 120, 25, -9, 3, -1),
    SR_ROW(-1, 4, -14, 118, 28, -9, 3, -1), SR_ROW(-1, 4, -15, 117, 30, -10, 4, -1),
    SR_ROW(-1, 5, -16, 116, 32, -11, 4, -1), SR_ROW(-1, 5, -16, 114, 35, -12, 4, -1),
    SR_ROW(-1, 5, -17, 112, 38, -12, 4, -1), SR_ROW(-1, 5, -18, 111, 40, -13, 5, -1),
    SR_ROW(-1, 5, -18, 109, 43, -14, 5, -1), SR_ROW(-1, 6, -19, 107, 45, -14, 5, -1),
    SR_ROW(-1, 6, -19, 105, 48, -15, 5, -1), SR_ROW(-1, 6, -19, 103, 51, -16, 5, -1),
    SR_ROW(-1, 6, -20, 101, 53, -16, 6, -1), SR_ROW(-1, 6, -20, 99, 56, -17, 6, -1),
    SR_ROW(-1, 6, -20, 97, 58, -17, 6, -1), SR_ROW(-1, 6, -20, 95, 61, -18, 6, -1),
    SR_ROW(-2, 7, -20, 93, 64, -18, 6, -2), SR_ROW(-2, 7, -20, 91, 66, -19, 6, -1),
    SR_ROW(-2, 7, -20, 88, 69, -19, 6, -1), SR_ROW(-2, 7, -20, 86, 71, -19, 6, -1),
    SR_ROW(-2, 7, -20, 84, 74, -20, 7, -2), SR_ROW(-2, 7, -20, 81, 76, -20, 7, -1),
    SR_ROW(-2, 7, -20, 79, 79, -20, 7, -2), SR_ROW(-1, 7, -20, 76, 81, -20, 7, -2),
    SR_ROW(-2, 7, -20, 74, 84, -20, 7, -2), SR_ROW(-1, 6, -19, 71, 86, -20, 7, -2),
    SR_ROW(-1, 6, -19, 69, 88, -20, 7, -2), SR_ROW(-1, 6, -19, 66, 91, -20, 7, -2),
    SR_ROW(-2, 6, -18, 64, 93, -20, 7, -2), SR_ROW(-1, 6, -18, 61, 95, -20, 6, -1),
    SR_ROW(-1, 6, -17, 58, 97, -20, 6, -1), SR_ROW(-1, 6, -17, 56, 99, -20, 6, -1),
    SR_ROW(-1, 6, -16, 53, 101, -20, 6, -1), SR_ROW(-1, 5, -16, 51, 103, -19, 6, -1),
    SR_ROW(-1, 5, -15, 48, 105, -19, 6, -1), SR_ROW(-1, 5, -14, 45, 107, -19, 6, -1),
    SR_ROW(-1, 5, -14, 43, 109, -18, 5, -1), SR_ROW(-1, 5, -13, 40, 111, -18, 5, -1),
    SR_ROW(-1, 4, -12, 38, 112, -17, 5, -1), SR_ROW(-1, 4, -12, 35, 114, -16, 5, -1),
    SR_ROW(-1, 4, -11, 32, 116, -16, 5, -1), SR_ROW(-1, 4, -10, 30, 117, -15, 4, -1),
    SR_ROW(-1, 3, -9, 28, 118, -14, 4, -1), SR_ROW(-1, 3, -9, 25, 120, -13, 4, -1),
    SR_ROW(-1, 3, -8, 22, 121, -12, 4, -1), SR_ROW(-1, 3, -7, 20, 122, -11, 3, -1),
    SR_ROW(-1, 2, -6, 18, 123, -10, 3, -1), SR_ROW(0, 2, -6, 15, 124, -9, 3, -1),
    SR_ROW(0, 2, -5, 13, 125, -8, 2, -1),   SR_ROW(0, 1, -4, 11, 125, -7, 2, 0),
    SR_ROW(0, 1, -3, 8, 126, -6, 2, 0),     SR_ROW(0, 1, -3, 6, 127, -4, 1, 0),
    SR_ROW(0, 1, -2, 4, 127, -3, 1, 0),     SR_ROW(0, 0, -1, 2, 128, -1, 0, 0),
};
#undef SR_ROW
#undef SR_PAIR

// sum of a tap pair (packed lo/hi int16) times two samples
__device__ __forceinline__ int sr_mac2(int acc, int32_t pair, int a, int b)
{
    return acc + (int)(int16_t)(pair & 0xffff) * a + (pair >> 16) * b;
}

__device__ __forceinline__ int sr_round_clamp(int sum, int maxv)
{
    return min(max((sum + 64) >> 7, 0), maxv);
}

// One output sample straight from a source row (the sparse boundary-row path).
template <typename T>
__device__ __forceinline__ int sr_sample(const T *row, const SrPlane &g, int x, int maxv)
{
    const int p = -(1 << kSrScaleBits) + g.x0 + x * g.step;
    const int b = (p >> kSrScaleBits) - kSrTapOffset;
    const int32_t *taps = kSrTaps[(p & kSrScaleMask) >> kSrPhaseShift];
    int v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = row[min(max(b + k, 0), g.srcw - 1)];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) sum = sr_mac2(sum, taps[k], v[2 * k], v[2 * k + 1]);
    return sr_round_clamp(sum, maxv);
}

}  // namespace pf
