// AV1 super-resolution upscale on gfx950 (superres/superres.go:38-107,
// decoder/postfilter_superres.go:198-264).
//
// Every plane row is upscaled independently from the coded width to the
// upscaled width with the normative 8-tap polyphase filter; the source span is
// the MI-aligned coded width, indexes clamped to it.  One workgroup takes a
// tile of 256 output columns by SR_ROWS rows: the source samples the tile's
// columns read (at most 264 per row, since the step is below one sample) are
// staged in LDS row by row with the clamp applied, then each thread filters its
// column in every row from LDS with the phase's taps held in registers.  Output
// rows are stored as 256 consecutive samples.  The pass is HBM bound: it reads
// the coded plane once and writes the upscaled plane once.
#include "pf_internal.h"
#include "superres_taps.h"

namespace pf {

constexpr int SR_SPAN = 272;   // staged samples per row (>= 264, the widest span)

template <typename T>
__device__ __forceinline__ void sr_tile(const PlaneView &s, const PlaneView &d, const SrPlane &g, int maxv, int tx,
                                        int ty, uint16_t (*lds)[SR_SPAN])
{
    const int X0 = tx * 256, y0 = ty * SR_ROWS;
    const int rows = min(SR_ROWS, g.rows - y0);
    const int p0 = -(1 << kSrScaleBits) + g.x0 + X0 * g.step;   // position of column X0
    const int s0 = (p0 >> kSrScaleBits) - kSrTapOffset;          // first source sample of the tile
    const int f0 = p0 & kSrScaleMask;
    const int span = ((f0 + 255 * g.step) >> kSrScaleBits) + 8;
    const int tid = threadIdx.x;
    for (int r = 0; r < rows; r++) {
        const T *row = row_ptr<T>(s, y0 + r);
        for (int i = tid; i < span; i += 256) lds[r][i] = row[min(max(s0 + i, 0), g.srcw - 1)];
    }
    __syncthreads();
    const int x = X0 + tid;
    if (x < g.dstw) {
        const int rel = f0 + tid * g.step;
        const int o = rel >> kSrScaleBits;
        const int32_t *tp = kSrTaps[(rel & kSrScaleMask) >> kSrPhaseShift];
        const int32_t t0 = tp[0], t1 = tp[1], t2 = tp[2], t3 = tp[3];
        for (int r = 0; r < rows; r++) {
            const uint16_t *w = &lds[r][o];
            int sum = sr_mac2(0, t0, w[0], w[1]);
            sum = sr_mac2(sum, t1, w[2], w[3]);
            sum = sr_mac2(sum, t2, w[4], w[5]);
            sum = sr_mac2(sum, t3, w[6], w[7]);
            row_ptr<T>(d, y0 + r)[x] = (T)sr_round_clamp(sum, maxv);
        }
    }
}

__global__ __launch_bounds__(256) void superres_kernel(const FrameDesc *__restrict__ descs, const Seg *__restrict__ segs,
                                                       int nseg)
{
    __shared__ uint16_t lds[SR_ROWS][SR_SPAN];
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int pl = sg.plane;
    const FrameView &S = D.v[D.sr_src];
    const FrameView &O = D.v[D.sr_dst];
    const int maxv = (1 << S.bd) - 1;
    if (S.bd == 8) sr_tile<uint8_t>(S.p[pl], O.p[pl], D.sr[pl], maxv, t % sg.nx, t / sg.nx, lds);
    else sr_tile<uint16_t>(S.p[pl], O.p[pl], D.sr[pl], maxv, t % sg.nx, t / sg.nx, lds);
}

hipError_t launch_superres(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(superres_kernel, dim3(total), dim3(256), 0, s, d, segs, nseg);
    return hipGetLastError();
}

}  // namespace pf
