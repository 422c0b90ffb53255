// Reference-frame border extension on gfx950.
//
// goav1 edge-replicates every plane of a frame that refreshes a reference slot
// into its whole allocation after loop restoration and super-resolution and
// before film grain (decoder/postfilter.go:760-771, Frame.ExtendBordersFrom /
// Plane.extendBorders / extendRight, frame/buffer.go:293-409):
//   - columns [width, stride / bytes per sample) of each visible row take the
//     row's last visible sample;
//   - rows [height, allocated rows) are copies of the extended last visible
//     row, so sample (y, x) there is row height-1 at column min(x, width-1).
// The second rule is applied in that closed form, so no tile depends on
// another tile's writes and one launch covers both regions of every plane.
// A workgroup is EXT_COLS x EXT_ROWS samples; a wave is one row of the tile.
#include "pf_internal.h"

namespace pf {

template <typename T>
__device__ __forceinline__ void extend_tile(const PlaneView &pv, int part, int tx, int ty)
{
    const int lane = threadIdx.x % EXT_COLS, r = threadIdx.x / EXT_COLS;
    if (part == 0) {
        // right padding of visible row y
        const int x = pv.w + tx * EXT_COLS + lane, y = ty * EXT_ROWS + r;
        if (x >= pv.cw || y >= pv.h) return;
        T *row = row_ptr<T>(pv, y);
        row[x] = row[pv.w - 1];
    } else {
        // a row below the visible height: the extended last visible row
        const int x = tx * EXT_COLS + lane, y = pv.h + ty * EXT_ROWS + r;
        if (x >= pv.cw || y >= pv.ch) return;
        row_ptr<T>(pv, y)[x] = row_ptr<T>(pv, pv.h - 1)[min(x, pv.w - 1)];
    }
}

__global__ __launch_bounds__(EXT_COLS *EXT_ROWS) void extend_kernel(const FrameDesc *__restrict__ descs,
                                                                     const Seg *__restrict__ segs, int nseg)
{
    const Seg sg = segs[find_seg(segs, nseg, blockIdx.x)];
    const FrameDesc &D = descs[sg.frame];
    const int t = blockIdx.x - sg.base;
    const int pl = sg.plane % 3, part = sg.plane / 3;
    const FrameView &f = D.v[D.ext];
    if (f.bd == 8) extend_tile<uint8_t>(f.p[pl], part, t % sg.nx, t / sg.nx);
    else extend_tile<uint16_t>(f.p[pl], part, t % sg.nx, t / sg.nx);
}

hipError_t launch_extend(const FrameDesc *d, const Seg *segs, int nseg, int total, hipStream_t s)
{
    if (total) hipLaunchKernelGGL(extend_kernel, dim3(total), dim3(EXT_COLS * EXT_ROWS), 0, s, d, segs, nseg);
    return hipGetLastError();
}

}  // namespace pf
