// The fused post-filter chain on gfx950: deblock (vertical, then horizontal
// edges), the loop-restoration context-row save, CDEF, super-resolution and
// loop restoration of every frame of a call in ONE persistent launch.
//
// goav1 runs these stages one whole frame at a time (decoder/postfilter.go:
// 589-782) or, in its superblock-row pipeline, row by row as soon as the rows a
// stage reads are final (decoder/postfilter_sbrow.go:9-37, postfilter_rowpipe.go:
// 6-27).  Here the stages' tiles (the same tiles, and the same device code, as
// the stage kernels in deblock.hip, lr.hip, cdef.hip and superres.hip) form one
// work queue in a dependency order, and resident workgroups take tiles from it
// one at a time.  Each frame keeps a completion counter per stage and 64-row
// band ("row"): a tile adds 1 to its own band's counter when it is done, and
// waits, before it starts, until the counters of the bands it reads or
// overwrites have reached their tile counts.  Rows are 64 luma rows (64 >> ss_y
// rows of a chroma plane), and the dependencies follow from the filter
// footprints (row r of a plane, PR = 64 >> ss_y plane rows per band, off =
// 8 >> ss_y the restoration stripe offset):
//   - a horizontal edge at row e reads rows e-7..e+6 and writes e-6..e+5 (the
//     14-tap filter; chroma filters reach less), so deblock-H of band p reads
//     the deblock-V output of bands p-1..p+1, and row r is final once the
//     horizontal edges within r-5..r+6 are done;
//   - the edges of one direction never read what another edge of that
//     direction writes (AV1's transform-edge spacing), so tiles of one
//     direction never wait for each other;
//   - the context rows of stripe t (rows y0-2, y0-1 with y0 = t*PR-off, and
//     y1, y1+1 with y1 = y0+PR) lie in bands t-1 and t: save(t) waits for
//     deblock-H of bands t-1 and t;
//   - CDEF unit row u reads rows u*PR-2 .. u*PR+PR+1 of the deblocked frame:
//     deblock-H of bands u-1..u+1;
//   - super-resolution of band p reads band p of its source (CDEF output, or
//     the deblocked frame);
//   - restoration stripe t reads its source rows y0..y1-1 (bands t-1 and t,
//     and the MI-aligned padding below the last stripe, also in those bands),
//     reads the context rows of stripe t, and overwrites the deblocked frame
//     rows that CDEF and the context saves of stripes t-1 and t+1 read when it
//     writes back into the frame: it waits for save(t-1..t+1) and for the
//     producer of its source in bands t-1 and t (CDEF, super-resolution, or
//     deblock-H).
// The queue holds the tiles in steps: band p of deblock-V in step p,
// deblock-H in step p+3, save in t+5, CDEF in u+6, super-resolution in p+8,
// restoration in t+10, so that every tile a tile waits for sits at least two
// steps earlier.  The queue is kept in one part per XCD (a band row's tiles
// split into contiguous column chunks, chunk c in part c, so neighbouring
// tiles share an L2), each part in step order; workgroups take from their
// XCD's part, then from the others.  The launch cannot deadlock whatever the
// grid and the occupancy: let m be the earliest step holding an unfinished
// tile; a taken tile of step m waits only for finished tiles, and every part
// whose next tile is of step m holds before it only tiles of steps <= m, which
// are finished or can finish, so a workgroup takes that tile next.  Every wait is bounded: a counter that does not
// reach its count within 2^20 polls (about a second) ends every wait of the
// launch and sets bit 4 of the device error flag, and the tiles go on (the
// call then reports a device error), so the launch always drains.
//
// Memory ordering: the frame planes are written by ordinary stores and read
// by ordinary loads on other compute units and XCDs.  A finished tile makes
// its stores visible device-wide before its counter is raised: every wave
// waits until its own stores have reached its XCD's L2, then (after a
// barrier) one wave issues an agent-scope release fence, which writes the L2
// back.  A waiting tile polls the counters with agent-scope atomic loads and
// then one wave issues an agent-scope acquire fence (the L1 and L2
// invalidation) and waits for it before the workgroup reads.
//
// The stage files are compiled as part of this translation unit so that the
// chain kernel inlines the stage bodies; the LDS of the largest stage (CDEF)
// holds any tile's working set.
#include "deblock.hip"
#include "cdef.hip"
#include "lr.hip"
#include "superres.hip"

namespace pf {

union ChainLds {
    DbLds db;
    CdefLds cd;
    LrLds lr;
    SrLds sr;
};

// Polls of one tile's waits before it gives up (a poll is an atomic load and a
// short sleep, about a microsecond: far beyond any launch's length).
constexpr unsigned kChainSpinLimit = 1u << 20;

__device__ __forceinline__ unsigned chain_load(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until counters [lo, hi] of one stage of a frame reach their targets
// (rows outside 0..nrows-1 are skipped).  Returns false after the poll limit,
// and then raises *abort, which ends every other wait of the launch at once.
__device__ __forceinline__ bool chain_wait(const unsigned *cnt, const unsigned *tgt, int stage_base, int lo, int hi,
                                           int nrows, unsigned &polls, unsigned *abort)
{
    lo = max(lo, 0);
    hi = min(hi, nrows - 1);
    for (int r = lo; r <= hi; r++) {
        const unsigned want = tgt[stage_base + r];
        int nap = 0;   // backs off to about a microsecond between polls
        while (chain_load(cnt + stage_base + r) < want) {
            if (chain_load(abort)) return false;
            if (++polls > kChainSpinLimit) {
                __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            if (nap < 4) {
                __builtin_amdgcn_s_sleep(4);
                nap++;
            } else {
                __builtin_amdgcn_s_sleep(32);
            }
        }
    }
    return true;
}

__device__ __forceinline__ int chain_plane_row(int stage, const FrameDesc &D, int pl, int ty)
{
    const int sy = pl > 0 ? D.v[VIEW_A].ssy : 0;
    if (stage == CS_DBV || stage == CS_DBH) return (ty * DB_ROWS * 4) >> (6 - sy);
    if (stage == CS_SR) return (ty * SR_TROWS) >> (6 - sy);
    return ty;
}

// Thread 0 of the workgroup: wait for every counter the tile depends on.
__device__ __forceinline__ bool chain_deps(const ChainItem &it, const ChainFrame &F, const unsigned *cnt,
                                           const unsigned *tgt, int row, unsigned *abort)
{
    const int R = F.rows, W = F.rows + 1;   // counter rows per stage
    const int b = F.cbase;
    unsigned polls = 0;
    bool ok = true;
    switch (it.stage) {
    case CS_DBH:
        ok = chain_wait(cnt, tgt, b + CS_DBV * W, row - 1, row + 1, R, polls, abort);
        break;
    case CS_SAVE:
        ok = chain_wait(cnt, tgt, b + CS_DBH * W, row - 1, row, R, polls, abort);
        break;
    case CS_CDEF:
        ok = chain_wait(cnt, tgt, b + CS_DBH * W, row - 1, row + 1, R, polls, abort);
        break;
    case CS_SR:
        if (F.sr_src == CS_CDEF) ok = chain_wait(cnt, tgt, b + CS_CDEF * W, row, row, R, polls, abort);
        else if (F.sr_src == CS_DBH) ok = chain_wait(cnt, tgt, b + CS_DBH * W, row - 1, row + 1, R, polls, abort);
        break;
    case CS_LR: {
        // rows of the stripe's source and padding: bands row-1 and row,
        // clamped to the frame (the last stripe may start in the last band)
        const int hi = min(row, R - 1);
        ok = chain_wait(cnt, tgt, b + CS_SAVE * W, row - 1, row + 1, W, polls, abort);
        if (ok && F.lr_src >= 0) ok = chain_wait(cnt, tgt, b + F.lr_src * W, hi - 1, hi, R, polls, abort);
        break;
    }
    default:
        break;
    }
    return ok;
}

// s_waitcnt vmcnt(0): the wave's loads and stores have completed (its stores
// are in the L2 of its XCD), and so has the cache write-back or invalidation
// of an agent-scope fence issued before it.  Written as an opaque instruction
// with a memory clobber: the compiler neither moves memory accesses across it
// nor drops it, whereas its own wait after a fence's L2 write-back is dropped
// as redundant when no store is outstanding, which would let the counter
// increment overtake the write-back.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// 6 workgroups per CU (the CDEF tile's 26.7 KB of LDS), at most 80 VGPRs.
// The queue is split in kXcds step-ordered parts (qbase[q] .. qbase[q + 1]),
// one per XCD: a workgroup takes tiles from its own XCD's part, then from the
// others.  A tile with dependencies: thread 0 waits for the counters, then
// invalidates the caches (agent-scope acquire, one wave) before the workgroup
// reads.  A tile with consumers: every wave waits for its own stores, then
// thread 0 writes the L2 back (agent-scope release, one wave) and raises the
// counter.  LR tiles have no consumers and deblock-V tiles no dependencies.
__global__ __launch_bounds__(256, 6) void chain_kernel(const FrameDesc *__restrict__ descs,
                                                    const ChainFrame *__restrict__ frames,
                                                    const ChainItem *__restrict__ items,
                                                    const unsigned *__restrict__ qbase, unsigned *head, unsigned *cnt,
                                                    const unsigned *__restrict__ tgt)
{
    unsigned *abort = head + kXcds;
    __shared__ ChainLds S;
    __shared__ int next;
    const int tid = threadIdx.x, x = blockIdx.x % kXcds;
    int q = 0;   // thread 0: queues tried so far
    for (;;) {
        if (tid == 0) {
            int i = -1;
            for (; q < kXcds; q++) {
                const int qq = (x + q) % kXcds;
                const unsigned k = qbase[qq] + atomicAdd(head + qq, 1u);
                if (k < qbase[qq + 1]) {
                    i = (int)k;
                    break;
                }
            }
            next = i;
        }
        __syncthreads();
        const int i = next;
        if (i < 0) break;   // the same value for every wave
        const ChainItem it = items[i];
        const ChainFrame &F = frames[it.frame];
        const FrameDesc &D = descs[it.frame];
        const int row = chain_plane_row(it.stage, D, it.plane, it.y);
        if (tid == 0 && it.stage != CS_DBV) {
            if (!chain_deps(it, F, cnt, tgt, row, abort)) atomicOr(D.err, 4);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            wait_vmem();
        }
        __syncthreads();
        switch (it.stage) {
        case CS_DBV: deblock_tile<0>(D, it.plane, it.x, it.y, S.db); break;
        case CS_DBH: deblock_tile<1>(D, it.plane, it.x, it.y, S.db); break;
        case CS_SAVE: lr_save_tile(D, it.plane, it.x, it.y); break;
        case CS_CDEF: cdef_tile(D, it.x, it.y, S.cd); break;
        case CS_SR: sr_tile_of(D, it.plane, it.x, it.y, S.sr); break;
        default: lr_tile_of(D, it.plane, it.x, it.y, S.lr); break;
        }
        if (it.stage != CS_LR) {
            wait_vmem();
            __syncthreads();   // every wave's stores are in its XCD's L2
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                wait_vmem();
                __hip_atomic_fetch_add(cnt + F.cbase + it.stage * (F.rows + 1) + row, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();   // the LDS and `next` are free again
    }
}

hipError_t launch_chain(const FrameDesc *d, const ChainFrame *frames, const ChainItem *items, int nitems,
                        const unsigned *qbase, unsigned *head, unsigned *cnt, const unsigned *tgt, hipStream_t s)
{
    if (nitems) {
        // a multiple of kXcds workgroups, so that every XCD holds its share
        const int g = std::max(kXcds, stage_grid((const void *)chain_kernel, nitems, 0) / kXcds * kXcds);
        hipLaunchKernelGGL(chain_kernel, dim3(g), dim3(256), 0, s, d, frames, items, qbase, head, cnt, tgt);
    }
    return hipGetLastError();
}

}  // namespace pf
