// libav1pf host side: validation, scratch arenas and the stage chain behind the
// C ABI declared in include/av1pf.h.
//
// The chain mirrors decoder/postfilter.go:589-845 (applySupportedPostFiltersInPlace):
//   deblock (in place) -> save LR context rows from the deblocked frame ->
//   CDEF (frame -> scratch) -> LR (scratch -> frame) -> film grain (frame ->
//   display).  Every request is validated on the host before anything is
//   launched (postfilter.go:598-616); the only check that needs device data,
//   a CDEF strength index past strength_count (postfilter_cdef.go:942-944), is
//   flagged by the kernel and reported by av1pf_sync.
//
// Batching: a call covers n frames.  The host writes one FrameDesc per frame
// and one segment table per stage into a pinned staging slot, copies it to the
// device in one transfer, and launches each stage once for all frames.  Slots
// form a ring guarded by events, so back-to-back calls never overwrite
// descriptors a previous call's kernels still read.  Film-grain templates
// depend only on the grain parameters and run on a side stream (which also
// uploads the descriptors) that joins the main stream before the grain apply;
// with double-buffered template memory, the templates of call N+1 are built
// while call N is still filtering.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "pf_internal.h"
#include "nn_internal.h"

using namespace pf;

// Scratch of the allocation-free entry point av1pf_apply_noalloc: every arena
// one call needs (descriptor image, second frame buffers, LR context rows,
// film-grain working memory), AV1PF_NOALLOC_SCRATCH_BYTES per device, held
// until the process ends.  Host code never passes the array's name to a
// kernel or a copy: each device's copy is reached through the address
// hipGetSymbolAddress returns for it (noalloc_scratch below).
static __device__ __align__(256) char scratch[AV1PF_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device `device`, looked up once
// per device; nullptr when the lookup fails.
void *noalloc_scratch(int device, std::string *err)
{
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index past the allocation-free scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

namespace {
constexpr int kSlots = 4;
// Film-grain template sets (and side streams) the calls rotate through: one
// frame's templates (a serial chain of about 220 us for 1080p: luma, then
// chroma) build while the previous frame filters.  Two: the process has four
// hardware queues (GPU_MAX_HW_QUEUES), and streams beyond them share queues,
// which serialises their work with the main stream's.
constexpr int kTmplSets = 2;
constexpr size_t kSlotBytes = size_t(1) << 20;   // ~400 frames of descriptors
// Stages whose work distribution is tuned: CDEF, LR, film-grain apply.
// Mode = units per workgroup (1, 2, 4) or 0 = persistent workgroups on a queue.
constexpr int kTunable = 3;
struct StageModes {
    int m[kTunable] = { 1, 1, 1 };
};
// One captured chain of a shape class: graph, its host image and device tables.
struct GraphInst {
    hipGraphExec_t exec = nullptr;
    char *host = nullptr, *dev = nullptr;
    hipEvent_t done = nullptr;
    bool used = false;
};
struct GraphClass {
    GraphInst inst[2];
    int next = 0;
};
void free_graph_inst(GraphInst &I)
{
    if (I.exec) hipGraphExecDestroy(I.exec);
    if (I.host) hipHostFree(I.host);
    if (I.dev) hipFree(I.dev);
    if (I.done) hipEventDestroy(I.done);
    I = GraphInst();
}
}

struct av1pf_ctx {
    int device = 0;
    std::string err;
    char *scratch = nullptr;      // frame-sized second buffers (CDEF output / LR ping-pong)
    size_t scratch_cap = 0;
    uint16_t *lr_bnd = nullptr;   // LR stripe context rows
    size_t lr_bnd_cap = 0;
    // film-grain templates (3 x 73 x 82 per frame) and block offsets in
    // kTmplSets rotating sets: the side streams fill the sets of the next
    // calls while the grain apply of this call reads its own
    int16_t *templates[kTmplSets] = {};
    size_t templates_cap[kTmplSets] = {};
    uint8_t *grain_offs[kTmplSets] = {};
    size_t grain_offs_cap[kTmplSets] = {};
    hipEvent_t tmpl_free[kTmplSets] = {};  // grain apply done with set p
    bool tmpl_free_used[kTmplSets] = {};
    int tpar = 0;
    hipEvent_t uploaded = nullptr;
    void *finish_state = nullptr;  // device CDEF strength selection state (zeroed once)
    unsigned int select_epoch = 0; // selection launches on finish_state
    char *learn_buf = nullptr;     // learned search: tiles, logits, shortlist table, LR mask
    size_t learn_cap = 0;
    void *small_state = nullptr;   // shortlist selection: per-level results and ticket (zeroed once)
    int *dev_err = nullptr;       // device-side validation flag
    hipStream_t sides[kTmplSets] = {};    // film-grain template stream of each template set
    hipStream_t side = nullptr;   // sides[0]: graphs, allocation-free calls, the learned search's LR
    hipEvent_t join = nullptr;
    hipEvent_t learn_fork = nullptr, learn_join = nullptr;   // learned search: LR search on the side stream
    // shape classes: tuned work distributions and captured graphs
    std::map<std::vector<int>, StageModes> tuned;
    std::map<std::vector<int>, GraphClass> gclass;
    bool graphs = false;
    int forced[kTunable] = { -1, -1, -1 };   // test override of the stage modes
    int pipeline = AV1PF_PIPELINE_STAGED;    // staged launches or the fused chain launch
    hipStream_t cap = nullptr;               // capture stream of the graphs
    hipEvent_t g_fork = nullptr, g_up = nullptr, g_join = nullptr;
    void drop_graphs(const std::vector<int> &key)
    {
        auto it = gclass.find(key);
        if (it == gclass.end()) return;
        hipDeviceSynchronize();
        for (auto &I : it->second.inst) free_graph_inst(I);
        gclass.erase(it);
    }
    // descriptor ring
    char *meta_dev[kSlots] = {};
    char *meta_host[kSlots] = {};
    size_t meta_cap[kSlots] = {};
    hipEvent_t meta_done[kSlots] = {};
    bool meta_used[kSlots] = {};
    int slot = 0;
    // stage-level LR: boundaries saved by av1pf_lr_save_boundaries for this frame
    const void *bnd_frame = nullptr;
    // arenas replaced by larger ones while queued work may still read them
    std::vector<void *> retired;
    // device address of the allocation-free entry point's scratch (nullptr: unavailable)
    void *noalloc = nullptr;
    std::string noalloc_err;
    // encoder CDEF apply: the device-side decision the next single-frame chain uses
    const av1pf_cdef_decision *apply_dec = nullptr;
    const uint8_t *apply_idx = nullptr;
    // per-stage timing
    bool profiling = false;
    struct Rec { int stage; hipEvent_t a, b; double bytes; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double acc_ms[AV1PF_PROFILE_STAGES] = {};
    int acc_n[AV1PF_PROFILE_STAGES] = {};
    double acc_bytes[AV1PF_PROFILE_STAGES] = {};
};

namespace {

enum { ST_DEBLOCK = 0, ST_LR_SAVE, ST_CDEF, ST_LR, ST_COPY, ST_GRAIN_T, ST_GRAIN_A, ST_SEARCH, ST_SEARCH_SELECT,
       ST_LR_SEARCH, ST_NN_FORWARD, ST_NN_TRAIN, ST_SEARCH_LEARNED, ST_LR_SEARCH_EXH, ST_SUPERRES,
       ST_DEBLOCK_BUILD, ST_EXTEND, ST_CHAIN };

int fail(av1pf_ctx *c, int code, const char *msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                       \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);     \
            return AV1PF_ERR_DEVICE;                                            \
        }                                                                       \
    } while (0)

hipEvent_t take_event(av1pf_ctx *c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Brackets one stage launch with events when profiling is on.
struct StageTimer {
    av1pf_ctx *c;
    int stage;
    hipStream_t s;
    double bytes;
    hipEvent_t a = nullptr;
    StageTimer(av1pf_ctx *c_, int st, hipStream_t s_, double by) : c(c_), stage(st), s(s_), bytes(by)
    {
        if (c->profiling) {
            a = take_event(c);
            hipEventRecord(a, s);
        }
    }
    ~StageTimer()
    {
        if (c->profiling && a) {
            hipEvent_t b = take_event(c);
            hipEventRecord(b, s);
            c->recs.push_back({ stage, a, b, bytes });
        }
    }
};

__global__ void zero_int_kernel(int *p) { *p = 0; }

int bps_of(int bd) { return bd == 8 ? 1 : 2; }

double view_bytes(const FrameView &f)
{
    double n = 0;
    for (int i = 0; i < f.nplanes; i++) n += (double)f.p[i].w * f.p[i].h;
    return n * bps_of(f.bd);
}

int validate_frame(av1pf_ctx *c, const av1pf_frame *f)
{
    if (!f) return fail(c, AV1PF_ERR_INVALID_ARG, "null frame");
    if (f->width <= 0 || f->height <= 0 || f->width > 16384 || f->height > 16384)
        return fail(c, AV1PF_ERR_INVALID_ARG, "frame size out of range");
    if (f->bit_depth != 8 && f->bit_depth != 10 && f->bit_depth != 12)
        return fail(c, AV1PF_ERR_INVALID_ARG, "bit depth must be 8, 10 or 12");
    if (f->ss_x < 0 || f->ss_x > 1 || f->ss_y < 0 || f->ss_y > 1 || f->ss_y > f->ss_x)
        return fail(c, AV1PF_ERR_INVALID_ARG, "unsupported chroma subsampling");
    const int np = f->mono ? 1 : 3;
    const int aw = (f->width + 7) & ~7;
    if (f->alloc_height != 0 && (f->alloc_height < ((f->height + 7) & ~7) || f->alloc_height > 16384 + 128))
        return fail(c, AV1PF_ERR_INVALID_ARG, "allocated height below the MI-aligned height or out of range");
    for (int i = 0; i < np; i++) {
        const int sx = i ? f->ss_x : 0;
        if (!f->plane[i]) return fail(c, AV1PF_ERR_INVALID_ARG, "null plane");
        if (f->stride[i] < (int64_t)(aw >> sx) * bps_of(f->bit_depth))
            return fail(c, AV1PF_ERR_SHORT_BUFFER, "plane stride shorter than the MI-aligned width");
    }
    return AV1PF_OK;
}

FrameView view_of(const av1pf_frame *f)
{
    FrameView v{};
    v.nplanes = f->mono ? 1 : 3;
    v.bd = f->bit_depth;
    v.ssx = f->ss_x;
    v.ssy = f->ss_y;
    v.w = f->width;
    v.h = f->height;
    const int aw = (f->width + 7) & ~7, ah = (f->height + 7) & ~7;
    for (int i = 0; i < v.nplanes; i++) {
        const int sx = i ? f->ss_x : 0, sy = i ? f->ss_y : 0;
        v.p[i].base = static_cast<char *>(f->plane[i]);
        v.p[i].stride = f->stride[i];
        v.p[i].w = (f->width + sx) >> sx;
        v.p[i].h = (f->height + sy) >> sy;
        v.p[i].aw = aw >> sx;
        v.p[i].ah = ah >> sy;
        v.p[i].cw = (int)(f->stride[i] / bps_of(f->bit_depth));
        v.p[i].ch = (f->alloc_height ? f->alloc_height : ah) >> sy;
        const uint64_t bits = reinterpret_cast<uintptr_t>(f->plane[i]) | (uint64_t)f->stride[i] | 16u;
        v.p[i].al = (int)(bits & (~bits + 1));
    }
    return v;
}

size_t scratch_bytes(const FrameView &f, int64_t strides[3])
{
    const int bps = bps_of(f.bd);
    size_t need = 0;
    for (int i = 0; i < f.nplanes; i++) {
        strides[i] = ((int64_t)f.p[i].aw * bps + 255) & ~(int64_t)255;
        need += (size_t)strides[i] * f.p[i].ah;
    }
    return (need + 255) & ~(size_t)255;
}

int validate_deblock(av1pf_ctx *c, const av1pf_frame *f, const av1pf_deblock_params *p)
{
    if (!p || !p->level_cache || !p->masks) return fail(c, AV1PF_ERR_INVALID_ARG, "deblock params missing");
    if (reinterpret_cast<uintptr_t>(p->masks) & 3)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock masks must be 4-byte aligned");
    const int mi_cols = ((f->width + 7) >> 3) << 1, mi_rows = ((f->height + 7) >> 3) << 1;
    if (p->mi_cols != mi_cols || p->mi_rows != mi_rows || p->sb128_w != (mi_cols + 31) >> 5 ||
        p->sb128_h != (mi_rows + 31) >> 5)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock mask grid does not match the frame");
    if (p->sharpness > 7 || p->level_y[0] > 63 || p->level_y[1] > 63 || p->level_u > 63 || p->level_v > 63)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock level or sharpness out of range");
    return AV1PF_OK;
}

int validate_cdef(av1pf_ctx *c, const av1pf_frame *f, const av1pf_cdef_params *p)
{
    if (!p || !p->index) return fail(c, AV1PF_ERR_INVALID_ARG, "cdef params missing");
    if (p->unit_cols != (f->width + 63) / 64 || p->unit_rows != (f->height + 63) / 64)
        return fail(c, AV1PF_ERR_INVALID_ARG, "cdef unit grid does not match the frame");
    if (p->strength_count < 1 || p->strength_count > 8 || p->damping < 3 || p->damping > 6)
        return fail(c, AV1PF_ERR_INVALID_ARG, "cdef strength count or damping out of range");
    for (int i = 0; i < 8; i++)
        if (p->y_strength[i] > 63 || p->uv_strength[i] > 63)
            return fail(c, AV1PF_ERR_INVALID_ARG, "cdef strength out of range");
    return AV1PF_OK;
}

int lr_geometry(const av1pf_frame *f, const av1pf_lr_params *p, int pl, int *pw, int *ph, int *hz, int *vt)
{
    const int sx = pl ? f->ss_x : 0, sy = pl ? f->ss_y : 0;
    *pw = (f->width + sx) >> sx;
    *ph = (f->height + sy) >> sy;
    // luma units are 64..256; chroma units are the luma size, halved only for
    // 4:2:0 (parser/restoration.go:70-105), so a unit is never shorter than a
    // 64-luma-row stripe of its plane
    const int us = p->unit_size[pl];
    if (us != 32 && us != 64 && us != 128 && us != 256) return AV1PF_ERR_INVALID_ARG;
    if (us < ((pl && sx && sy) ? 32 : 64)) return AV1PF_ERR_INVALID_ARG;
    *hz = std::max((*pw + (us >> 1)) / us, 1);
    *vt = std::max((*ph + (us >> 1)) / us, 1);
    return AV1PF_OK;
}

int validate_lr(av1pf_ctx *c, const av1pf_frame *f, const av1pf_lr_params *p)
{
    if (!p) return fail(c, AV1PF_ERR_INVALID_ARG, "lr params missing");
    const int np = f->mono ? 1 : 3;
    for (int i = 0; i < np; i++) {
        if (p->frame_type[i] > AV1PF_LR_SWITCHABLE) return fail(c, AV1PF_ERR_INVALID_ARG, "lr frame type out of range");
        if (p->frame_type[i] == AV1PF_LR_NONE) continue;
        int pw, ph, hz, vt;
        if (lr_geometry(f, p, i, &pw, &ph, &hz, &vt)) return fail(c, AV1PF_ERR_INVALID_ARG, "lr unit size invalid");
        if (!p->units[i]) return fail(c, AV1PF_ERR_INVALID_ARG, "lr unit records missing");
    }
    return AV1PF_OK;
}

bool grain_noop(const av1pf_grain_params *g)
{
    return !g->chroma_scaling_from_luma && !g->num_y_points && !g->num_cb_points && !g->num_cr_points;
}

int validate_grain(av1pf_ctx *c, const av1pf_frame *f, const av1pf_grain_params *g, const av1pf_frame *d)
{
    if (!g) return fail(c, AV1PF_ERR_INVALID_ARG, "film grain params missing");
    if (!d) return fail(c, AV1PF_ERR_INVALID_ARG, "film grain needs a display frame");
    if (d->width != f->width || d->height != f->height || d->bit_depth != f->bit_depth ||
        d->ss_x != f->ss_x || d->ss_y != f->ss_y || d->mono != f->mono)
        return fail(c, AV1PF_ERR_INVALID_ARG, "display frame geometry differs from the frame");
    int rc = validate_frame(c, d);
    if (rc) return rc;
    if (g->num_y_points > 14 || g->num_cb_points > 10 || g->num_cr_points > 10 || g->ar_coeff_lag > 3 ||
        g->grain_scale_shift > 3)
        return fail(c, AV1PF_ERR_INVALID_ARG, "film grain params out of range");
    if (grain_noop(g)) return AV1PF_OK;
    if (g->scaling_shift < 8 || g->scaling_shift > 11) return fail(c, AV1PF_ERR_INVALID_ARG, "scaling shift out of range");
    if ((g->ar_coeff_lag || g->num_y_points || g->chroma_scaling_from_luma || g->num_cb_points || g->num_cr_points) &&
        (g->ar_coeff_shift < 6 || g->ar_coeff_shift > 9))
        return fail(c, AV1PF_ERR_INVALID_ARG, "ar coeff shift out of range");
    if (f->mono && (g->chroma_scaling_from_luma || g->num_cb_points || g->num_cr_points))
        return fail(c, AV1PF_ERR_INVALID_ARG, "chroma grain on a monochrome frame");
    if (f->ss_x && f->ss_y && ((g->num_cb_points == 0) != (g->num_cr_points == 0)))
        return fail(c, AV1PF_ERR_INVALID_ARG, "4:2:0 needs both or neither chroma scaling functions");
    const uint8_t(*lists[3])[2] = { (const uint8_t(*)[2])g->y_points, (const uint8_t(*)[2])g->cb_points,
                                    (const uint8_t(*)[2])g->cr_points };
    const int counts[3] = { g->num_y_points, g->num_cb_points, g->num_cr_points };
    for (int l = 0; l < 3; l++)
        for (int i = 1; i < counts[l]; i++)
            if (lists[l][i][0] <= lists[l][i - 1][0])
                return fail(c, AV1PF_ERR_INVALID_ARG, "scaling points must have increasing values");
    return AV1PF_OK;
}

// AV1 scaling LUT (filmgrain/scaling.go:18-51) -- tiny host work.
void build_scaling_lut(uint8_t lut[256], const uint8_t (*pts)[2], int n)
{
    memset(lut, 0, 256);
    if (n == 0) return;
    for (int x = 0; x < pts[0][0]; x++) lut[x] = pts[0][1];
    for (int i = 0; i + 1 < n; i++) {
        const int dy = (int)pts[i + 1][1] - pts[i][1];
        const int dx = (int)pts[i + 1][0] - pts[i][0];
        const int64_t delta = (int64_t)dy * ((65536 + (dx >> 1)) / dx);
        for (int x = 0; x < dx; x++)
            lut[pts[i][0] + x] = (uint8_t)(pts[i][1] + (int)(((int64_t)x * delta + 32768) >> 16));
    }
    for (int x = pts[n - 1][0]; x < 256; x++) lut[x] = pts[n - 1][1];
}

// Columns of the GF(2) matrix of `steps` LFSR steps (the step is linear).
void lfsr_jump(uint16_t J[16], int steps)
{
    for (int b = 0; b < 16; b++) {
        uint16_t r = (uint16_t)(1u << b);
        for (int k = 0; k < steps; k++) {
            const uint16_t bit = (uint16_t)((r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1);
            r = (uint16_t)((r >> 1) | (bit << 15));
        }
        J[b] = r;
    }
}

uint16_t gf2_apply(const uint16_t J[16], uint16_t s)
{
    uint16_t o = 0;
    for (int b = 0; b < 16; b++)
        if ((s >> b) & 1) o ^= J[b];
    return o;
}

// The upscaled frame of a super-resolution request: same format and height as
// the coded frame, wider by a factor in (1, 2] (denominators 9..16 over 8,
// postfilter_superres.go:112-117), planes not overlapping the coded frame's.
int validate_superres(av1pf_ctx *c, const av1pf_frame *f, const av1pf_frame *u)
{
    if (!u) return fail(c, AV1PF_ERR_INVALID_ARG, "superres needs an upscaled frame");
    int rc = validate_frame(c, u);
    if (rc) return rc;
    if (u->bit_depth != f->bit_depth || u->ss_x != f->ss_x || u->ss_y != f->ss_y || u->mono != f->mono ||
        u->height != f->height)
        return fail(c, AV1PF_ERR_INVALID_ARG, "upscaled frame format or height differs from the coded frame");
    if (u->width <= f->width || u->width > 2 * f->width)
        return fail(c, AV1PF_ERR_INVALID_ARG, "upscaled width must be above the coded width and at most twice it");
    const int np = f->mono ? 1 : 3;
    for (int i = 0; i < np; i++)
        for (int j = 0; j < np; j++) {
            const int sy = j ? f->ss_y : 0;
            const char *fa = static_cast<const char *>(f->plane[j]);
            const char *fb = fa + f->stride[j] * (int64_t)((((f->height + 7) & ~7) >> sy));
            const char *ua = static_cast<const char *>(u->plane[i]);
            const char *ub = ua + u->stride[i] * (int64_t)((((u->height + 7) & ~7) >> sy));
            if (ua < fb && fa < ub) return fail(c, AV1PF_ERR_INVALID_ARG, "upscaled frame overlaps the coded frame");
        }
    return AV1PF_OK;
}

// Stages run on the coded frame up to CDEF; with super-resolution, loop
// restoration and film grain run on the upscaled frame
// (ApplyCallerPostFilters, postfilter.go:947-979).
int validate_request(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_request *req)
{
    if (!req) return fail(c, AV1PF_ERR_INVALID_ARG, "null request");
    int rc = validate_frame(c, frame);
    if (rc) return rc;
    const uint32_t st = req->stages;
    if (st & ~(AV1PF_STAGE_DEBLOCK | AV1PF_STAGE_CDEF | AV1PF_STAGE_SUPERRES | AV1PF_STAGE_LR | AV1PF_STAGE_FILMGRAIN))
        return fail(c, AV1PF_ERR_INVALID_ARG, "unknown stage bits");
    if (req->flags & ~AV1PF_REQ_REFRESH) return fail(c, AV1PF_ERR_INVALID_ARG, "unknown request flags");
    const av1pf_frame *post = frame;
    if (st & AV1PF_STAGE_SUPERRES) {
        if ((rc = validate_superres(c, frame, req->upscaled))) return rc;
        post = req->upscaled;
    }
    if ((st & AV1PF_STAGE_DEBLOCK) && (rc = validate_deblock(c, frame, req->deblock))) return rc;
    if ((st & AV1PF_STAGE_CDEF) && (rc = validate_cdef(c, frame, req->cdef))) return rc;
    if ((st & AV1PF_STAGE_LR) && (rc = validate_lr(c, post, req->lr))) return rc;
    if ((st & AV1PF_STAGE_FILMGRAIN) && (rc = validate_grain(c, post, req->grain, req->display))) return rc;
    return AV1PF_OK;
}

// Upscale geometry of plane pl (superres.go:38-48): step and initial phase
// from the coded and upscaled plane widths, source span the MI-aligned coded
// width (postfilter_superres.go:241-252).
SrPlane sr_plane(const av1pf_frame *f, const av1pf_frame *u, int pl)
{
    const int sx = pl ? f->ss_x : 0, sy = pl ? f->ss_y : 0;
    SrPlane g{};
    g.codedw = (f->width + sx) >> sx;
    g.dstw = (u->width + sx) >> sx;
    g.srcw = ((f->width + 7) & ~7) >> sx;
    g.rows = (f->height + sy) >> sy;
    const int64_t cw = g.codedw, uw = g.dstw;
    const int64_t step = ((cw << 14) + uw / 2) / uw;
    const int64_t err = uw * step - (cw << 14);
    int64_t x0 = (-((uw - cw) << 13) + uw / 2) / uw;
    x0 += (1 << 7) - err / 2;
    g.step = (int)step;
    g.x0 = (int)(x0 & 0x3fff);
    return g;
}

struct SegList {
    std::vector<Seg> v;
    int total = 0;
    void add(int frame, int plane, int nx, int ny)
    {
        if (nx <= 0 || ny <= 0) return;
        v.push_back({ frame, plane, nx, total });
        total += nx * ny;
    }
};

struct BatchPlan {
    std::vector<FrameDesc> descs;
    SegList dbv, dbh, save, cdef, lr, copy_ba, copy_ad, tl, tc, grain, sr, ext;
    double bytes_db = 0, bytes_cdef = 0, bytes_lr = 0, bytes_copy = 0, bytes_grain = 0, bytes_sr = 0, bytes_ext = 0;
    size_t scratch = 0, bnd = 0, goffs = 0;
    int ngrain = 0;
};

// Fill descriptors and segment tables; pointers into the arenas are patched
// once the arena bases are known (offsets kept in the descriptors meanwhile).
int plan_batch(av1pf_ctx *c, int n, const av1pf_frame *frames, const av1pf_request *reqs, BatchPlan &P,
               std::vector<size_t> &scr_off, std::vector<size_t> &bnd_off, std::vector<int> &grain_idx,
               std::vector<size_t> &goff, bool skip_lr_save, std::vector<size_t> *sscr_off = nullptr)
{
    goff.assign(n, 0);
    P.descs.assign(n, FrameDesc{});
    scr_off.assign(n, 0);
    if (sscr_off) sscr_off->assign(n, SIZE_MAX);
    bnd_off.assign(n * 3, 0);
    grain_idx.assign(n, -1);
    for (int i = 0; i < n; i++) {
        const av1pf_frame *f = &frames[i];
        const av1pf_request *r = &reqs[i];
        FrameDesc &D = P.descs[i];
        D.v[VIEW_A] = view_of(f);
        const FrameView &A = D.v[VIEW_A];
        int64_t strides[3];
        scr_off[i] = P.scratch;
        P.scratch += scratch_bytes(A, strides);
        D.v[VIEW_B] = A;
        for (int p = 0; p < A.nplanes; p++) {   // base patched later (256-byte aligned planes)
            PlaneView &b = D.v[VIEW_B].p[p];
            b.stride = strides[p];
            b.cw = (int)(strides[p] / bps_of(A.bd));
            b.ch = b.ah;
            b.al = 16;
        }
        const uint32_t st = r->stages;
        const bool sr = (st & AV1PF_STAGE_SUPERRES) != 0;
        const av1pf_frame *pf = sr ? r->upscaled : f;   // the frame LR and film grain see
        D.v[VIEW_U] = sr ? view_of(pf) : A;
        D.v[VIEW_S] = D.v[VIEW_U];
        D.v[VIEW_D] = r->display ? view_of(r->display) : D.v[VIEW_U];
        D.fin = sr ? VIEW_U : VIEW_A;
        if (sr)
            for (int p = 0; p < A.nplanes; p++) D.sr[p] = sr_plane(f, pf, p);
        const FrameView &U = D.v[VIEW_U];
        D.err = c->dev_err;
        int cur = VIEW_A;
        // ---- deblock
        if (st & AV1PF_STAGE_DEBLOCK) {
            const av1pf_deblock_params *p = r->deblock;
            DbDesc &db = D.db;
            db.lc = p->level_cache;
            db.masks = p->masks;
            db.mi_cols = p->mi_cols;
            db.sbw = p->sb128_w;
            db.sharp = p->sharpness;
            const bool luma_zero = p->level_y[0] == 0 && p->level_y[1] == 0;
            for (int pl = 0; pl < A.nplanes; pl++) {
                if (pl && (luma_zero || (pl == 1 ? p->level_u : p->level_v) == 0)) continue;
                const int sx = pl ? A.ssx : 0, sy = pl ? A.ssy : 0;
                const int w4 = ((((f->width + sx) >> sx) + 3) & ~3) >> 2;
                const int h4 = ((((f->height + sy) >> sy) + 3) & ~3) >> 2;
                P.dbv.add(i, pl, (w4 + 63) / 64, (h4 + DB_ROWS - 1) / DB_ROWS);
                P.dbh.add(i, pl, (w4 + 63) / 64, (h4 + DB_ROWS - 1) / DB_ROWS);
            }
            P.bytes_db += 2 * view_bytes(A);
        }
        // ---- LR boundary rows (from the deblocked frame)
        bool lr_any = false;
        if (st & AV1PF_STAGE_LR) {
            const av1pf_lr_params *p = r->lr;
            for (int pl = 0; pl < A.nplanes; pl++) {
                LrPlaneDesc &g = D.lr[pl];
                const int sx = pl ? A.ssx : 0, sy = pl ? A.ssy : 0;
                g.type = p->frame_type[pl];
                g.sr = sr;
                g.copy_pad = !sr;
                g.pw = U.p[pl].w;
                g.ph = U.p[pl].h;
                g.procw = 64 >> sx;
                g.sh = 64 >> sy;
                g.off = 8 >> sy;
                g.xtiles = (g.pw + g.procw - 1) / g.procw;
                g.stripes = lr_stripes(g.ph, g.sh, g.off);
                if (g.type == AV1PF_LR_NONE) continue;
                int pw, ph;
                lr_geometry(pf, p, pl, &pw, &ph, &g.hz, &g.vt);
                g.us = p->unit_size[pl];
                g.units = p->units[pl];
                g.bp = (g.pw + 7) & ~7;
                bnd_off[i * 3 + pl] = P.bnd;
                P.bnd += (size_t)g.stripes * 4 * g.bp;
                lr_any = true;
            }
            if (lr_any && !skip_lr_save)
                for (int pl = 0; pl < A.nplanes; pl++)
                    if (D.lr[pl].type != AV1PF_LR_NONE) P.save.add(i, pl, (D.lr[pl].pw + 255) / 256, D.lr[pl].stripes);
        }
        // ---- CDEF
        if (st & AV1PF_STAGE_CDEF) {
            const av1pf_cdef_params *p = r->cdef;
            CdefDesc &cd = D.cd;
            for (int k = 0; k < p->strength_count && k < 8; k++) {
                if (p->y_strength[k]) cd.yfilt = 1;
                if (!f->mono && p->uv_strength[k]) cd.uvfilt = 1;
            }
            if (i == 0 && c->apply_dec) {
                // strengths and unit indices of a device-side search decision:
                // known only on the device, so the stage always runs
                cd.dec = c->apply_dec;
                cd.fb_index = c->apply_idx;
                cd.rows8 = (f->height + 7) >> 3;
                cd.cols8 = (f->width + 7) >> 3;
                cd.yfilt = cd.uvfilt = 1;
            }
            if (cd.yfilt || cd.uvfilt) {
                cd.index = p->index;
                cd.noskip = p->noskip;
                cd.unit_cols = p->unit_cols;
                cd.unit_rows = p->unit_rows;
                cd.damping = p->damping;
                cd.strength_count = p->strength_count;
                memcpy(cd.y_strength, p->y_strength, 8);
                memcpy(cd.uv_strength, p->uv_strength, 8);
                P.cdef.add(i, -1, p->unit_cols, p->unit_rows);
                P.bytes_cdef += 2 * view_bytes(A);
                cur = VIEW_B;
            }
        }
        // ---- super-resolution: coded picture -> upscaled frame (or its scratch
        // when LR follows); the coded frame keeps the CDEF output
        if (sr) {
            D.sr_src = cur;
            D.sr_dst = lr_any ? VIEW_S : VIEW_U;
            double samples = 0;
            for (int pl = 0; pl < A.nplanes; pl++) {
                const SrPlane &g = D.sr[pl];
                P.sr.add(i, pl, (g.dstw + SR_COLS - 1) / SR_COLS, (g.rows + SR_ROWS * SR_CHUNKS - 1) / (SR_ROWS * SR_CHUNKS));
                samples += (double)(g.codedw + g.dstw) * g.rows;
            }
            P.bytes_sr += samples * bps_of(A.bd);
            if (cur == VIEW_B) {
                for (int pl = 0; pl < A.nplanes; pl++)
                    P.copy_ba.add(i, pl, (A.p[pl].aw + COPY_COLS - 1) / COPY_COLS, (A.p[pl].ah + COPY_ROWS - 1) / COPY_ROWS);
                P.bytes_copy += 2 * view_bytes(A);
            }
            if (lr_any) {
                int64_t strides[3];
                if (sscr_off) (*sscr_off)[i] = P.scratch;
                P.scratch += scratch_bytes(U, strides);
                for (int p = 0; p < U.nplanes; p++) {   // base patched later (256-byte aligned planes)
                    PlaneView &v = D.v[VIEW_S].p[p];
                    v.stride = strides[p];
                    v.cw = (int)(strides[p] / bps_of(U.bd));
                    v.ch = v.ah;
                    v.al = 16;
                }
            }
            cur = D.sr_dst;
        }
        // ---- LR
        if (lr_any) {
            D.lr_src = cur;
            D.lr_dst = sr ? VIEW_U : cur == VIEW_A ? VIEW_B : VIEW_A;
            for (int pl = 0; pl < A.nplanes; pl++) P.lr.add(i, pl, D.lr[pl].xtiles, D.lr[pl].stripes);
            P.bytes_lr += 2 * view_bytes(U);
            cur = D.lr_dst;
        }
        if (cur == VIEW_B) {
            for (int pl = 0; pl < A.nplanes; pl++)
                P.copy_ba.add(i, pl, (A.p[pl].aw + COPY_COLS - 1) / COPY_COLS, (A.p[pl].ah + COPY_ROWS - 1) / COPY_ROWS);
            P.bytes_copy += 2 * view_bytes(A);
        }
        // ---- border extension of a refreshing frame (postfilter.go:760-771):
        // the frame LR and film grain see, once it holds the filtered picture
        D.ext = D.fin;
        if (r->flags & AV1PF_REQ_REFRESH) {
            const FrameView &E = D.v[D.ext];
            for (int pl = 0; pl < E.nplanes; pl++) {
                const PlaneView &pv = E.p[pl];
                const int padw = pv.cw - pv.w, padh = pv.ch - pv.h;
                if (padw > 0) P.ext.add(i, pl, (padw + EXT_COLS - 1) / EXT_COLS, (pv.h + EXT_ROWS - 1) / EXT_ROWS);
                if (padh > 0) P.ext.add(i, 3 + pl, (pv.cw + EXT_COLS - 1) / EXT_COLS, (padh + EXT_ROWS - 1) / EXT_ROWS);
                P.bytes_ext += ((double)std::max(padw, 0) * pv.h + (double)std::max(padh, 0) * pv.cw) * 2 * bps_of(E.bd);
            }
        }
        // ---- film grain
        if (st & AV1PF_STAGE_FILMGRAIN) {
            const av1pf_grain_params *g = r->grain;
            if (grain_noop(g)) {
                for (int pl = 0; pl < U.nplanes; pl++)
                    P.copy_ad.add(i, pl, (U.p[pl].aw + COPY_COLS - 1) / COPY_COLS, (U.p[pl].ah + COPY_ROWS - 1) / COPY_ROWS);
                P.bytes_copy += 2 * view_bytes(U);
            } else {
                GrainDesc &G = D.gr;
                G.g = *g;
                G.active[0] = g->num_y_points != 0;
                G.active[1] = !pf->mono && (g->num_cb_points != 0 || g->chroma_scaling_from_luma);
                G.active[2] = !pf->mono && (g->num_cr_points != 0 || g->chroma_scaling_from_luma);
                build_scaling_lut(G.luts[0], (const uint8_t(*)[2])g->y_points, g->num_y_points);
                if (g->chroma_scaling_from_luma) {
                    memcpy(G.luts[1], G.luts[0], 256);
                    memcpy(G.luts[2], G.luts[0], 256);
                } else {
                    build_scaling_lut(G.luts[1], (const uint8_t(*)[2])g->cb_points, g->num_cb_points);
                    build_scaling_lut(G.luts[2], (const uint8_t(*)[2])g->cr_points, g->num_cr_points);
                }
                uint16_t jl[16], jc[16];
                lfsr_jump(jl, 82);
                lfsr_jump(jc, pf->ss_x ? 44 : 82);
                const uint16_t seeds[3] = { g->seed, (uint16_t)(g->seed ^ 0xb524),
                                            (uint16_t)(g->seed ^ 0x49d8) };
                for (int p = 0; p < 3; p++) {
                    uint16_t st = seeds[p];
                    for (int r = 0; r < 73; r++) {
                        G.row_seed[p][r] = st;
                        st = gf2_apply(p ? jc : jl, st);
                    }
                }
                G.matrix_identity = pf->matrix_identity;
                G.orows = (pf->height + 31) / 32;
                G.ocols = (pf->width + 31) / 32;
                G.k[GK_OVERLAP] = g->overlap;
                G.k[GK_CSFL] = g->chroma_scaling_from_luma;
                G.k[GK_CLIP] = g->clip_to_restricted_range;
                G.k[GK_SHIFT] = g->scaling_shift;
                G.k[GK_MID] = pf->matrix_identity;
                G.k[GK_CB_MULT] = g->cb_mult;
                G.k[GK_CB_LMULT] = g->cb_luma_mult;
                G.k[GK_CB_OFF] = g->cb_offset;
                G.k[GK_CR_MULT] = g->cr_mult;
                G.k[GK_CR_LMULT] = g->cr_luma_mult;
                G.k[GK_CR_OFF] = g->cr_offset;
                G.k[GK_ACT0] = G.active[0];
                G.k[GK_ACT1] = G.active[1];
                G.k[GK_ACT2] = G.active[2];
                G.k[GK_OCOLS] = G.ocols;
                goff[i] = P.goffs;
                P.goffs += (size_t)G.orows * G.ocols;
                grain_idx[i] = P.ngrain++;
                P.tl.add(i, 0, 1, 1);
                if (!pf->mono) {
                    P.tc.add(i, 1, 1, 1);
                    P.tc.add(i, 2, 1, 1);
                }
                P.grain.add(i, -1, (pf->width + 31) / 32, (pf->height + 31) / 32);
                P.bytes_grain += 2 * view_bytes(U);
            }
        }
    }
    return AV1PF_OK;
}

// The fused launch's queue: the tiles of the deblock, context-save, CDEF,
// super-resolution and LR stage lists (the same tiles the stage launches
// cover), placed in steps by band (deblock-V of band p in step p, deblock-H
// p+3, save t+5, CDEF u+6, super-resolution p+8, LR t+10: every counter a tile
// waits for belongs to tiles at least two steps earlier), and the tile count of
// every counter.  The queue is stored as one part per XCD: the tiles of a
// band row of a stage go to the parts in contiguous column chunks, and each
// part is in step order.
struct ChainPlan {
    std::vector<ChainItem> items;
    std::vector<ChainFrame> frames;
    std::vector<uint32_t> targets;
    uint32_t qbase[kXcds + 1];   // the queue's part per XCD
};

int plan_chain(const BatchPlan &P, int n, ChainPlan &C)
{
    static const int kLag[CS_N] = { 0, 3, 5, 6, 8, 10 };
    const SegList *lists[CS_N] = { &P.dbv, &P.dbh, &P.save, &P.cdef, &P.sr, &P.lr };
    C.frames.assign(n, ChainFrame{});
    // bands per frame: the luma rows of the frame and of its LR / upscaled view
    for (int i = 0; i < n; i++) {
        const FrameDesc &D = P.descs[i];
        const int h = std::max(D.v[VIEW_A].p[0].ah, D.v[VIEW_U].p[0].ah);
        C.frames[i].rows = std::max(1, (h + 63) / 64);
    }
    auto tile_row = [&](int st, int fr, int pl, int ty) {
        const int sy = pl > 0 ? P.descs[fr].v[VIEW_A].ssy : 0;
        if (st == CS_DBV || st == CS_DBH) return (ty * DB_ROWS * 4) >> (6 - sy);
        if (st == CS_SR) return (ty * SR_ROWS * SR_CHUNKS) >> (6 - sy);
        return ty;
    };
    // rows of every frame must hold its tiles (the save and LR stripes may
    // use one more: counters are [rows + 1] per stage)
    for (int st = 0; st < CS_N; st++) {
        const SegList &L = *lists[st];
        for (size_t k = 0; k < L.v.size(); k++) {
            const Seg &g = L.v[k];
            const int end = k + 1 < L.v.size() ? L.v[k + 1].base : L.total;
            const int ny = (end - g.base) / g.nx;
            const int last = tile_row(st, g.frame, g.plane, ny - 1);
            ChainFrame &F = C.frames[g.frame];
            if (st == CS_SAVE || st == CS_LR) {
                if (last > F.rows) return AV1PF_ERR_INVALID_ARG;
            } else if (last >= F.rows) {
                F.rows = last + 1;
            }
        }
    }
    int ncnt = 0;
    for (int i = 0; i < n; i++) {
        ChainFrame &F = C.frames[i];
        const FrameDesc &D = P.descs[i];
        F.cbase = ncnt;
        ncnt += CS_N * (F.rows + 1);
        F.sr_src = D.sr_src == VIEW_B ? CS_CDEF : CS_DBH;
        F.lr_src = D.lr_src == VIEW_S ? CS_SR : D.lr_src == VIEW_B ? CS_CDEF : CS_DBH;
    }
    C.targets.assign(ncnt, 0);
    // counting sort by (part, step); inside a step, stage then frame then
    // tile order
    int maxstep = 0;
    for (int i = 0; i < n; i++) maxstep = std::max(maxstep, C.frames[i].rows + 1 + kLag[CS_LR]);
    const int nsteps = maxstep + 1;
    std::vector<int> at(kXcds * nsteps + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
        for (int st = 0; st < CS_N; st++) {
            const SegList &L = *lists[st];
            for (size_t k = 0; k < L.v.size(); k++) {
                const Seg &g = L.v[k];
                const int end = k + 1 < L.v.size() ? L.v[k + 1].base : L.total;
                const int ny = (end - g.base) / g.nx;
                const ChainFrame &F = C.frames[g.frame];
                for (int ty = 0; ty < ny; ty++) {
                    const int row = tile_row(st, g.frame, g.plane, ty);
                    const int step = row + kLag[st];
                    if (pass == 0 && st != CS_LR) C.targets[F.cbase + st * (F.rows + 1) + row] += g.nx;
                    for (int tx = 0; tx < g.nx; tx++) {
                        const int slot = (tx * kXcds / g.nx) * nsteps + step;
                        if (pass == 0)
                            at[slot + 1]++;
                        else
                            C.items[at[slot]++] = ChainItem{ (uint16_t)g.frame, (uint8_t)st,
                                                             (uint8_t)(g.plane < 0 ? 0 : g.plane), (uint16_t)tx,
                                                             (uint16_t)ty };
                    }
                }
            }
        }
        if (pass == 0) {
            for (size_t k = 1; k < at.size(); k++) at[k] += at[k - 1];
            C.items.resize(at.back());
            for (int q = 0; q <= kXcds; q++) C.qbase[q] = (uint32_t)at[q * nsteps];
        }
    }
    return AV1PF_OK;
}

int ensure(av1pf_ctx *c, void **buf, size_t *cap, size_t need)
{
    if (need <= *cap) return AV1PF_OK;
    // work this handle queued earlier may still read the old arena: it is
    // retired, not freed, and released once the handle's streams are idle
    // (av1pf_sync, av1pf_destroy), so growth waits on no stream at all
    if (*buf) c->retired.push_back(*buf);
    *buf = nullptr;
    *cap = 0;
    need = need + need / 4;
    HIPCHK(c, hipMalloc(buf, need));
    *cap = need;
    return AV1PF_OK;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// A caller-provided device region that replaces every arena of one call
// (the allocation-free entry point): descriptor image, second frame buffers,
// LR context rows, film-grain working memory.
struct FixedArena {
    char *base;
    size_t cap;
};

int run_batch(av1pf_ctx *c, int n, const av1pf_frame *frames, const av1pf_request *reqs, hipStream_t s,
              bool skip_lr_save = false, bool tune_modes = false, StageModes *tuned_out = nullptr,
              const FixedArena *fixed = nullptr)
{
    StageModes chosen;
    if (n == 0) return AV1PF_OK;
    BatchPlan P;
    std::vector<size_t> scr_off, bnd_off, goff, sscr_off;
    std::vector<int> grain_idx;
    int rc = plan_batch(c, n, frames, reqs, P, scr_off, bnd_off, grain_idx, goff, skip_lr_save, &sscr_off);
    if (rc) return rc;
    SegList *lists[] = { &P.dbv, &P.dbh, &P.save, &P.cdef, &P.lr, &P.copy_ba, &P.copy_ad, &P.tl, &P.tc, &P.grain, &P.sr,
                         &P.ext };
    const int nl = sizeof(lists) / sizeof(lists[0]);
    size_t off = align256(sizeof(FrameDesc) * n);
    size_t offs[nl];
    for (int k = 0; k < nl; k++) {
        offs[k] = off;
        off = align256(off + sizeof(Seg) * (lists[k]->v.size() + 1));
    }
    const size_t qoff = off;   // work-queue counters of the tunable stages (one per XCD), zero at upload
    off = align256(off + kTunable * kXcds * sizeof(unsigned int));
    // fused chain: queue, per-frame records, counter targets, then the counters
    // and the queue head (zero at upload)
    ChainPlan CP;
    const bool fused = c->pipeline == AV1PF_PIPELINE_FUSED &&
                       (P.dbv.total || P.dbh.total || P.save.total || P.cdef.total || P.sr.total || P.lr.total);
    size_t ch_items = 0, ch_frames = 0, ch_tgt = 0, ch_cnt = 0, ch_heads = 0, ch_zero = 0;
    if (fused) {
        if (n > 65535) return fail(c, AV1PF_ERR_INVALID_ARG, "the fused chain takes at most 65535 frames per call");
        if ((rc = plan_chain(P, n, CP))) return fail(c, rc, "fused chain: a stage's tiles exceed the frame's bands");
        ch_items = off;
        off = align256(off + sizeof(ChainItem) * CP.items.size());
        ch_frames = off;
        off = align256(off + sizeof(ChainFrame) * n);
        ch_tgt = off;   // targets, then the part bases
        off = align256(off + sizeof(uint32_t) * (CP.targets.size() + kXcds + 1));
        // counters, then (in lines of their own) the part heads and the abort word
        ch_cnt = off;
        ch_heads = align256(off + sizeof(uint32_t) * CP.targets.size());
        ch_zero = ch_heads + 256 - ch_cnt;
        off = align256(off + ch_zero);
    }

    const size_t tsz = GR_FRAME_BYTES;
    const bool side = P.tl.total > 0;
    const int par = c->tpar;
    char *scratch_base, *fixed_meta = nullptr;
    uint16_t *bnd_base;
    int16_t *tmpl_base = nullptr;
    uint8_t *goffs_base = nullptr;
    if (fixed) {
        // every arena carved from the caller's region: nothing is allocated
        size_t at = off;   // the descriptor image comes first
        const size_t scr_at = align256(at);
        at = scr_at + align256(P.scratch);
        const size_t bnd_at = at;
        at += align256(P.bnd * 2);
        const size_t tmpl_at = at;
        at += align256(P.ngrain * tsz);
        const size_t goff_at = at;
        at += align256(P.goffs);
        if (at > fixed->cap || off > c->meta_cap[c->slot])
            return fail(c, AV1PF_ERR_SHORT_BUFFER, "request needs more scratch than the allocation-free entry point holds");
        fixed_meta = fixed->base;
        scratch_base = fixed->base + scr_at;
        bnd_base = reinterpret_cast<uint16_t *>(fixed->base + bnd_at);
        tmpl_base = reinterpret_cast<int16_t *>(fixed->base + tmpl_at);
        goffs_base = reinterpret_cast<uint8_t *>(fixed->base + goff_at);
    } else {
        if ((rc = ensure(c, (void **)&c->scratch, &c->scratch_cap, std::max<size_t>(P.scratch, 256)))) return rc;
        if ((rc = ensure(c, (void **)&c->lr_bnd, &c->lr_bnd_cap, std::max<size_t>(P.bnd * 2, 256)))) return rc;
        if (side) {
            c->tpar = (c->tpar + 1) % kTmplSets;
            if ((rc = ensure(c, (void **)&c->templates[par], &c->templates_cap[par],
                             std::max<size_t>(P.ngrain * tsz, 256))))
                return rc;
            if ((rc = ensure(c, (void **)&c->grain_offs[par], &c->grain_offs_cap[par], std::max<size_t>(P.goffs, 256))))
                return rc;
            tmpl_base = c->templates[par];
            goffs_base = c->grain_offs[par];
        }
        scratch_base = c->scratch;
        bnd_base = c->lr_bnd;
    }
    for (int i = 0; i < n; i++) {
        FrameDesc &D = P.descs[i];
        char *base = scratch_base + scr_off[i];
        for (int p = 0; p < D.v[VIEW_B].nplanes; p++) {
            D.v[VIEW_B].p[p].base = base;
            base += (size_t)D.v[VIEW_B].p[p].stride * D.v[VIEW_B].p[p].ah;
        }
        if (sscr_off[i] != SIZE_MAX) {
            base = scratch_base + sscr_off[i];
            for (int p = 0; p < D.v[VIEW_S].nplanes; p++) {
                D.v[VIEW_S].p[p].base = base;
                base += (size_t)D.v[VIEW_S].p[p].stride * D.v[VIEW_S].p[p].ah;
            }
        }
        for (int p = 0; p < 3; p++)
            if (D.lr[p].type != AV1PF_LR_NONE) D.lr[p].bnd = bnd_base + bnd_off[i * 3 + p];
        if (grain_idx[i] >= 0) {
            char *gw = reinterpret_cast<char *>(tmpl_base) + (size_t)grain_idx[i] * GR_FRAME_BYTES;
            D.gr.templates = reinterpret_cast<int16_t *>(gw);
            D.gr.scale = reinterpret_cast<uint8_t *>(gw + GR_TMPL_BYTES);
            D.gr.offsets = goffs_base + goff[i];
        }
    }
    // ---- descriptors + segment tables + stage queues, one host image
    auto fill = [&](char *h) {
        memcpy(h, P.descs.data(), sizeof(FrameDesc) * n);
        for (int k = 0; k < nl; k++) {
            Seg sentinel{ 0, 0, 1, lists[k]->total };
            memcpy(h + offs[k], lists[k]->v.data(), sizeof(Seg) * lists[k]->v.size());
            memcpy(h + offs[k] + sizeof(Seg) * lists[k]->v.size(), &sentinel, sizeof(Seg));
        }
        memset(h + qoff, 0, kTunable * kXcds * sizeof(unsigned int));
        if (fused) {
            memcpy(h + ch_items, CP.items.data(), sizeof(ChainItem) * CP.items.size());
            memcpy(h + ch_frames, CP.frames.data(), sizeof(ChainFrame) * n);
            memcpy(h + ch_tgt, CP.targets.data(), sizeof(uint32_t) * CP.targets.size());
            memcpy(h + ch_tgt + sizeof(uint32_t) * CP.targets.size(), CP.qbase, sizeof(CP.qbase));
            memset(h + ch_cnt, 0, ch_zero);
        }
    };
    // shape class: everything a launch's arguments depend on besides data
    std::vector<int> key;
    key.push_back(n);
    key.push_back(side ? 1 : 0);
    for (int k = 0; k < nl; k++) {
        key.push_back(lists[k]->total);
        key.push_back((int)lists[k]->v.size());
    }
    key.push_back(fused ? 1 : 0);
    key.push_back((int)CP.targets.size());
    StageModes modes;
    if (auto it = c->tuned.find(key); it != c->tuned.end()) modes = it->second;
    for (int i = 0; i < kTunable; i++)
        if (c->forced[i] >= 0) modes.m[i] = c->forced[i];

    // enqueue the chain reading its tables from device memory d (uploaded from h)
    auto enqueue = [&](char *h, char *d, hipStream_t st, bool capture, StageModes md) -> int {
        const FrameDesc *dd = reinterpret_cast<const FrameDesc *>(d);
        auto segp = [&](int k) { return reinterpret_cast<const Seg *>(d + offs[k]); };
        auto nseg = [&](int k) { return (int)lists[k]->v.size(); };
        auto launch_of = [&](int i) {
            StageLaunch L;
            L.per_wg = md.m[i];
            L.queue = reinterpret_cast<unsigned int *>(d + qoff) + i * kXcds;
            return L;
        };
        // tuning: after the stage's real launch, time every work distribution
        // on the same inputs (the stage is out of place, so the result is the same)
        auto tune = [&](int i, auto &&run) -> int {
            if (!tune_modes) return AV1PF_OK;
            static const int cand[] = { 1, 2, 4, 0 };
            float best = 1e30f;
            int pick = md.m[i];
            hipEvent_t e0 = take_event(c), e1 = take_event(c);
            for (int m : cand) {
                StageLaunch L = launch_of(i);
                L.per_wg = m;
                for (int rep = 0; rep < 4; rep++) {
                    if (rep == 1) HIPCHK(c, hipEventRecord(e0, st));
                    if (m == 0) HIPCHK(c, launch_zero_u32(L.queue, kXcds, st));
                    HIPCHK(c, run(L));
                }
                HIPCHK(c, hipEventRecord(e1, st));
                HIPCHK(c, hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; pick = m; }
            }
            c->pool.push_back(e0);
            c->pool.push_back(e1);
            chosen.m[i] = pick;
            return AV1PF_OK;
        };
        hipEvent_t ev_up = capture ? c->g_up : c->uploaded, ev_join = capture ? c->g_join : c->join;
        if (side && fixed) {
            // one template region shared by every call: the upload and the
            // templates follow the stream's previous work (the last call's
            // grain apply included)
            HIPCHK(c, hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, st));
            HIPCHK(c, hipEventRecord(ev_up, st));
            HIPCHK(c, hipStreamWaitEvent(c->side, ev_up, 0));
            {
                StageTimer t(c, ST_GRAIN_T, c->side, 0);
                HIPCHK(c, launch_grain_templates(dd, segp(7), nseg(7), segp(8), nseg(8), c->side));
            }
            HIPCHK(c, hipEventRecord(ev_join, c->side));
        } else if (side) {
            // Film-grain templates depend only on the grain parameters: a side
            // stream uploads the descriptors and builds them without waiting for
            // the previous call's filtering, only for the grain apply that last
            // read this template pair (inside a graph the replay is ordered).
            // Each template pair has its own side stream, so the templates of
            // consecutive calls build concurrently.
            hipStream_t ss = capture ? c->side : c->sides[par];
            if (capture) {
                HIPCHK(c, hipEventRecord(c->g_fork, st));
                HIPCHK(c, hipStreamWaitEvent(ss, c->g_fork, 0));
            } else if (c->tmpl_free_used[par]) {
                HIPCHK(c, hipStreamWaitEvent(ss, c->tmpl_free[par], 0));
            }
            HIPCHK(c, hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, ss));
            HIPCHK(c, hipEventRecord(ev_up, ss));
            HIPCHK(c, hipStreamWaitEvent(st, ev_up, 0));
            {
                StageTimer t(c, ST_GRAIN_T, ss, 0);
                HIPCHK(c, launch_grain_templates(dd, segp(7), nseg(7), segp(8), nseg(8), ss));
            }
            HIPCHK(c, hipEventRecord(ev_join, ss));
        } else {
            HIPCHK(c, hipMemcpyAsync(d, h, off, hipMemcpyHostToDevice, st));
        }
        int rc2;
        if (fused) {
            StageTimer t(c, ST_CHAIN, st, P.bytes_db + P.bytes_cdef + P.bytes_sr + P.bytes_lr);
            unsigned *cnt = reinterpret_cast<unsigned *>(d + ch_cnt);
            const unsigned *tgt = reinterpret_cast<const unsigned *>(d + ch_tgt);
            HIPCHK(c, launch_chain(dd, reinterpret_cast<const ChainFrame *>(d + ch_frames),
                                   reinterpret_cast<const ChainItem *>(d + ch_items), (int)CP.items.size(),
                                   tgt + CP.targets.size(), reinterpret_cast<unsigned *>(d + ch_heads), cnt, tgt,
                                   st));
        }
        if (!fused && (P.dbv.total || P.dbh.total)) {
            StageTimer t(c, ST_DEBLOCK, st, P.bytes_db);
            HIPCHK(c, launch_deblock(dd, segp(0), nseg(0), P.dbv.total, segp(1), nseg(1), P.dbh.total, st));
        }
        if (!fused && P.save.total) {
            StageTimer t(c, ST_LR_SAVE, st, 0);
            HIPCHK(c, launch_lr_save(dd, segp(2), nseg(2), P.save.total, st));
        }
        if (!fused && P.cdef.total) {
            {
                StageTimer t(c, ST_CDEF, st, P.bytes_cdef);
                HIPCHK(c, launch_cdef(dd, segp(3), nseg(3), P.cdef.total, launch_of(0), st));
            }
            if ((rc2 = tune(0, [&](const StageLaunch &L) { return launch_cdef(dd, segp(3), nseg(3), P.cdef.total, L, st); })))
                return rc2;
        }
        if (!fused && P.sr.total) {
            StageTimer t(c, ST_SUPERRES, st, P.bytes_sr);
            HIPCHK(c, launch_superres(dd, segp(10), nseg(10), P.sr.total, st));
        }
        if (!fused && P.lr.total) {
            {
                StageTimer t(c, ST_LR, st, P.bytes_lr);
                HIPCHK(c, launch_lr(dd, segp(4), nseg(4), P.lr.total, launch_of(1), st));
            }
            if ((rc2 = tune(1, [&](const StageLaunch &L) { return launch_lr(dd, segp(4), nseg(4), P.lr.total, L, st); })))
                return rc2;
        }
        if (P.copy_ba.total) {
            StageTimer t(c, ST_COPY, st, P.bytes_copy);
            HIPCHK(c, launch_copy(dd, segp(5), nseg(5), P.copy_ba.total, VIEW_B, VIEW_A, st));
        }
        if (P.ext.total) {
            StageTimer t(c, ST_EXTEND, st, P.bytes_ext);
            HIPCHK(c, launch_extend(dd, segp(11), nseg(11), P.ext.total, st));
        }
        if (P.copy_ad.total) {
            StageTimer t(c, ST_COPY, st, 0);
            HIPCHK(c, launch_copy(dd, segp(6), nseg(6), P.copy_ad.total, -1, VIEW_D, st));
        }
        if (side) HIPCHK(c, hipStreamWaitEvent(st, ev_join, 0));
        if (P.grain.total) {
            {
                StageTimer t(c, ST_GRAIN_A, st, P.bytes_grain);
                HIPCHK(c, launch_grain_apply(dd, segp(9), nseg(9), P.grain.total, launch_of(2), st));
            }
            if ((rc2 = tune(2, [&](const StageLaunch &L) {
                     return launch_grain_apply(dd, segp(9), nseg(9), P.grain.total, L, st);
                 })))
                return rc2;
        }
        return AV1PF_OK;
    };

    if (c->graphs && !c->profiling && !tune_modes && !fixed) {
        // one graph per shape class, two instances used alternately so that the
        // host image of one can be written while the other replays
        GraphClass &G = c->gclass[key];
        GraphInst &I = G.inst[G.next];
        G.next ^= 1;
        if (I.used) HIPCHK(c, hipEventSynchronize(I.done));
        if (!I.exec) {
            HIPCHK(c, hipHostMalloc(&I.host, off, hipHostMallocDefault));
            HIPCHK(c, hipMalloc(&I.dev, off));
            HIPCHK(c, hipEventCreateWithFlags(&I.done, hipEventDisableTiming));
            HIPCHK(c, hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal));
            rc = enqueue(I.host, I.dev, c->cap, true, modes);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(c->cap, &g);
            if (rc) {
                if (g) hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(c, ec);
            const hipError_t ei = hipGraphInstantiate(&I.exec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            HIPCHK(c, ei);
        }
        fill(I.host);
        HIPCHK(c, hipGraphLaunch(I.exec, s));
        HIPCHK(c, hipEventRecord(I.done, s));
        I.used = true;
        if (side) {
            // the replay's grain apply read template pair par: a later call
            // outside the graph waits for it before rewriting the pair
            HIPCHK(c, hipEventRecord(c->tmpl_free[par], s));
            c->tmpl_free_used[par] = true;
        }
        return AV1PF_OK;
    }

    const int slot = c->slot;
    c->slot = (c->slot + 1) % kSlots;
    if (c->meta_used[slot]) HIPCHK(c, hipEventSynchronize(c->meta_done[slot]));
    if (off > c->meta_cap[slot]) {
        if (c->meta_dev[slot]) HIPCHK(c, hipFree(c->meta_dev[slot]));
        if (c->meta_host[slot]) HIPCHK(c, hipHostFree(c->meta_host[slot]));
        const size_t cap = off * 2;
        HIPCHK(c, hipMalloc(&c->meta_dev[slot], cap));
        HIPCHK(c, hipHostMalloc(&c->meta_host[slot], cap, hipHostMallocDefault));
        c->meta_cap[slot] = cap;
    }
    fill(c->meta_host[slot]);
    chosen = modes;
    if ((rc = enqueue(c->meta_host[slot], fixed ? fixed_meta : c->meta_dev[slot], s, false, modes))) return rc;
    if (side && !fixed) {
        HIPCHK(c, hipEventRecord(c->tmpl_free[par], s));
        c->tmpl_free_used[par] = true;
    }
    HIPCHK(c, hipEventRecord(c->meta_done[slot], s));
    c->meta_used[slot] = true;
    if (tune_modes) {
        c->tuned[key] = chosen;
        c->drop_graphs(key);   // graphs captured with the previous modes
        if (tuned_out) *tuned_out = chosen;
    }
    return AV1PF_OK;
}

}  // namespace

extern "C" {

const char *av1pf_version(void) { return "av1pf 0.2 gfx950"; }

const char *av1pf_last_error(const av1pf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int av1pf_create(av1pf_ctx **out, int device)
{
    if (!out) return AV1PF_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return AV1PF_ERR_NO_DEVICE;
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return AV1PF_ERR_NO_DEVICE;
    }
    if (device >= n) return AV1PF_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return AV1PF_ERR_DEVICE;
    av1pf_ctx *c = new av1pf_ctx();
    c->device = device;
    bool ok = hipMalloc(&c->dev_err, sizeof(int)) == hipSuccess &&
              hipEventCreateWithFlags(&c->uploaded, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->join, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->learn_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->learn_join, hipEventDisableTiming) == hipSuccess &&
              hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&c->g_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->g_up, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->g_join, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kTmplSets; i++)
        ok = hipStreamCreateWithFlags(&c->sides[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->tmpl_free[i], hipEventDisableTiming) == hipSuccess;
    if (ok) c->side = c->sides[0];
    // descriptor slots sized for large batches up front, so that no call in
    // steady state allocates
    for (int i = 0; ok && i < kSlots; i++) {
        ok = hipEventCreateWithFlags(&c->meta_done[i], hipEventDisableTiming) == hipSuccess &&
             hipMalloc(&c->meta_dev[i], kSlotBytes) == hipSuccess &&
             hipHostMalloc(&c->meta_host[i], kSlotBytes, hipHostMallocDefault) == hipSuccess;
        if (ok) c->meta_cap[i] = kSlotBytes;
    }
    if (ok) {
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, nullptr, c->dev_err);
        ok = hipDeviceSynchronize() == hipSuccess;
    }
    // the allocation-free entry point's scratch on this device (current here);
    // a failed lookup disables only that entry point
    if (ok) c->noalloc = noalloc_scratch(device, &c->noalloc_err);
    if (!ok) {
        av1pf_destroy(c);
        return AV1PF_ERR_DEVICE;
    }
    *out = c;
    return AV1PF_OK;
}

void av1pf_destroy(av1pf_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    hipDeviceSynchronize();
    for (void *p : c->retired) hipFree(p);
    if (c->scratch) hipFree(c->scratch);
    if (c->lr_bnd) hipFree(c->lr_bnd);
    if (c->dev_err) hipFree(c->dev_err);
    for (int i = 0; i < kSlots; i++) {
        if (c->meta_dev[i]) hipFree(c->meta_dev[i]);
        if (c->meta_host[i]) hipHostFree(c->meta_host[i]);
        if (c->meta_done[i]) hipEventDestroy(c->meta_done[i]);
    }
    if (c->uploaded) hipEventDestroy(c->uploaded);
    if (c->finish_state) hipFree(c->finish_state);
    if (c->learn_buf) hipFree(c->learn_buf);
    if (c->small_state) hipFree(c->small_state);
    if (c->join) hipEventDestroy(c->join);
    if (c->learn_fork) hipEventDestroy(c->learn_fork);
    if (c->learn_join) hipEventDestroy(c->learn_join);
    for (auto &kv : c->gclass)
        for (auto &I : kv.second.inst) free_graph_inst(I);
    if (c->cap) hipStreamDestroy(c->cap);
    for (hipEvent_t e : { c->g_fork, c->g_up, c->g_join })
        if (e) hipEventDestroy(e);
    for (int i = 0; i < kTmplSets; i++) {
        if (c->tmpl_free[i]) hipEventDestroy(c->tmpl_free[i]);
        if (c->templates[i]) hipFree(c->templates[i]);
        if (c->grain_offs[i]) hipFree(c->grain_offs[i]);
    }
    for (hipStream_t ss : c->sides)
        if (ss) hipStreamDestroy(ss);
    for (auto &r : c->recs) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : c->pool) hipEventDestroy(e);
    delete c;
}

int av1pf_apply(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_request *req, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_request(c, frame, req);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (req->stages & AV1PF_STAGE_LR) c->bnd_frame = nullptr;   // the chain's save overwrites saved rows
    return run_batch(c, 1, frame, req, static_cast<hipStream_t>(stream));
}

int av1pf_apply_noalloc(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_request *req, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_request(c, frame, req);
    if (rc) return rc;
    if (!c->noalloc) return fail(c, AV1PF_ERR_DEVICE, ("allocation-free scratch unavailable: " + c->noalloc_err).c_str());
    HIPCHK(c, hipSetDevice(c->device));
    if (req->stages & AV1PF_STAGE_LR) c->bnd_frame = nullptr;
    const FixedArena arena{ static_cast<char *>(c->noalloc), (size_t)AV1PF_NOALLOC_SCRATCH_BYTES };
    return run_batch(c, 1, frame, req, static_cast<hipStream_t>(stream), false, false, nullptr, &arena);
}

int av1pf_apply_batch(av1pf_ctx *c, int n, const av1pf_frame *frames, const av1pf_request *reqs, void *stream)
{
    if (!c || n < 0 || (n && (!frames || !reqs))) return AV1PF_ERR_INVALID_ARG;
    for (int i = 0; i < n; i++) {
        int rc = validate_request(c, &frames[i], &reqs[i]);
        if (rc) return rc;
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (int i = 0; i < n; i++)
        if (reqs[i].stages & AV1PF_STAGE_LR) c->bnd_frame = nullptr;
    return run_batch(c, n, frames, reqs, static_cast<hipStream_t>(stream));
}

int av1pf_set_graphs(av1pf_ctx *c, int enable)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (!enable) {
        HIPCHK(c, hipDeviceSynchronize());
        for (auto &kv : c->gclass)
            for (auto &I : kv.second.inst) free_graph_inst(I);
        c->gclass.clear();
    }
    c->graphs = enable != 0;
    return AV1PF_OK;
}

int av1pf_autotune(av1pf_ctx *c, int n, const av1pf_frame *frames, const av1pf_request *reqs, int32_t *modes,
                   void *stream)
{
    if (!c || n < 0 || (n && (!frames || !reqs))) return AV1PF_ERR_INVALID_ARG;
    for (int i = 0; i < n; i++) {
        int rc = validate_request(c, &frames[i], &reqs[i]);
        if (rc) return rc;
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (int i = 0; i < n; i++)
        if (reqs[i].stages & AV1PF_STAGE_LR) c->bnd_frame = nullptr;
    StageModes chosen;
    const int rc = run_batch(c, n, frames, reqs, static_cast<hipStream_t>(stream), false, true, &chosen);
    if (rc) return rc;
    if (modes)
        for (int i = 0; i < kTunable; i++) modes[i] = chosen.m[i];
    return AV1PF_OK;
}

int av1pf_set_pipeline(av1pf_ctx *c, int mode)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    if (mode != AV1PF_PIPELINE_STAGED && mode != AV1PF_PIPELINE_FUSED)
        return fail(c, AV1PF_ERR_INVALID_ARG, "pipeline mode is AV1PF_PIPELINE_STAGED or AV1PF_PIPELINE_FUSED");
    c->pipeline = mode;
    return AV1PF_OK;
}

int av1pf_set_stage_mode(av1pf_ctx *c, uint32_t stage, int mode)
{
    if (!c || mode < -1 || mode > 8) return AV1PF_ERR_INVALID_ARG;
    const int i = stage == AV1PF_STAGE_CDEF ? 0 : stage == AV1PF_STAGE_LR ? 1 : stage == AV1PF_STAGE_FILMGRAIN ? 2 : -1;
    if (i < 0) return fail(c, AV1PF_ERR_INVALID_ARG, "stage mode applies to CDEF, LR and film grain only");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    c->forced[i] = mode;
    for (auto &kv : c->gclass)   // graphs captured with the previous modes
        for (auto &I : kv.second.inst) free_graph_inst(I);
    c->gclass.clear();
    return AV1PF_OK;
}

int av1pf_sync(av1pf_ctx *c, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    for (hipStream_t ss : c->sides) HIPCHK(c, hipStreamSynchronize(ss));
    for (void *p : c->retired) HIPCHK(c, hipFree(p));   // nothing of this handle is in flight
    c->retired.clear();
    int flag = 0;
    HIPCHK(c, hipMemcpy(&flag, c->dev_err, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), c->dev_err);
        HIPCHK(c, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
        if (flag & 4) return fail(c, AV1PF_ERR_DEVICE, "fused chain: a tile's dependency wait did not complete");
        if (flag & 2) {
            // an aborted selection leaves its barrier counters behind: start over
            HIPCHK(c, hipMemset(c->finish_state, 0, finish_state_bytes()));
            c->select_epoch = 0;
            return fail(c, AV1PF_ERR_DEVICE, "device strength selection: a barrier did not complete");
        }
        return fail(c, AV1PF_ERR_INVALID_ARG, "cdef strength index past strength_count");
    }
    return AV1PF_OK;
}

int av1pf_deblock(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_deblock_params *p, void *stream)
{
    av1pf_request r{};
    r.stages = AV1PF_STAGE_DEBLOCK;
    r.deblock = p;
    return av1pf_apply(c, frame, &r, stream);
}

int av1pf_cdef(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_cdef_params *p, void *stream)
{
    av1pf_request r{};
    r.stages = AV1PF_STAGE_CDEF;
    r.cdef = p;
    return av1pf_apply(c, frame, &r, stream);
}

int av1pf_lr_save_boundaries(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_lr_params *p, int after_cdef, void *stream)
{
    // Frame-edge context is read straight from the post-CDEF frame by the LR
    // kernel, so only the deblocked interior seam rows are stored; the
    // after-CDEF save is a no-op.
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, frame);
    if (rc) return rc;
    if ((rc = validate_lr(c, frame, p))) return rc;
    if (after_cdef) return AV1PF_OK;
    HIPCHK(c, hipSetDevice(c->device));
    BatchPlan P;
    std::vector<size_t> scr_off, bnd_off, goff;
    std::vector<int> gi;
    av1pf_request r{};
    r.stages = AV1PF_STAGE_LR;
    r.lr = p;
    // a request with only the save: run the chain's save segment by planning an
    // LR request and launching just its save list
    if ((rc = plan_batch(c, 1, frame, &r, P, scr_off, bnd_off, gi, goff, false))) return rc;
    if ((rc = ensure(c, (void **)&c->lr_bnd, &c->lr_bnd_cap, std::max<size_t>(P.bnd * 2, 256)))) return rc;
    for (int pl = 0; pl < 3; pl++)
        if (P.descs[0].lr[pl].type != AV1PF_LR_NONE) P.descs[0].lr[pl].bnd = c->lr_bnd + bnd_off[pl];
    const int slot = c->slot;
    c->slot = (c->slot + 1) % kSlots;
    if (c->meta_used[slot]) HIPCHK(c, hipEventSynchronize(c->meta_done[slot]));
    const size_t off = align256(sizeof(FrameDesc));
    const size_t need = off + sizeof(Seg) * (P.save.v.size() + 1);
    if (need > c->meta_cap[slot]) {
        if (c->meta_dev[slot]) HIPCHK(c, hipFree(c->meta_dev[slot]));
        if (c->meta_host[slot]) HIPCHK(c, hipHostFree(c->meta_host[slot]));
        HIPCHK(c, hipMalloc(&c->meta_dev[slot], need * 2));
        HIPCHK(c, hipHostMalloc(&c->meta_host[slot], need * 2, hipHostMallocDefault));
        c->meta_cap[slot] = need * 2;
    }
    memcpy(c->meta_host[slot], P.descs.data(), sizeof(FrameDesc));
    Seg sentinel{ 0, 0, 1, P.save.total };
    memcpy(c->meta_host[slot] + off, P.save.v.data(), sizeof(Seg) * P.save.v.size());
    memcpy(c->meta_host[slot] + off + sizeof(Seg) * P.save.v.size(), &sentinel, sizeof(Seg));
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(c, hipMemcpyAsync(c->meta_dev[slot], c->meta_host[slot], need, hipMemcpyHostToDevice, s));
    {
        StageTimer t(c, ST_LR_SAVE, s, 0);
        HIPCHK(c, launch_lr_save(reinterpret_cast<const FrameDesc *>(c->meta_dev[slot]),
                                 reinterpret_cast<const Seg *>(c->meta_dev[slot] + off), (int)P.save.v.size(),
                                 P.save.total, s));
    }
    HIPCHK(c, hipEventRecord(c->meta_done[slot], s));
    c->meta_used[slot] = true;
    c->bnd_frame = frame->plane[0];
    return AV1PF_OK;
}

int av1pf_lr(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_lr_params *p, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    av1pf_request r{};
    r.stages = AV1PF_STAGE_LR;
    r.lr = p;
    int rc = validate_request(c, frame, &r);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    // boundaries saved by av1pf_lr_save_boundaries for this frame (before CDEF) are used as is
    const bool saved = c->bnd_frame == frame->plane[0];
    c->bnd_frame = nullptr;
    return run_batch(c, 1, frame, &r, static_cast<hipStream_t>(stream), saved);
}

int av1pf_cdef_search_apply(av1pf_ctx *c, const av1pf_frame *recon, const uint64_t *noskip, int damping,
                            const av1pf_cdef_decision *decision, const uint8_t *fb_index, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    if (!decision || !fb_index) return fail(c, AV1PF_ERR_INVALID_ARG, "cdef apply: decision and indices needed");
    int rc = validate_frame(c, recon);
    if (rc) return rc;
    av1pf_cdef_params p{};
    p.damping = (uint8_t)damping;
    p.strength_count = 8;   // the decision's count is read on the device
    p.unit_cols = (recon->width + 63) / 64;
    p.unit_rows = (recon->height + 63) / 64;
    p.index = reinterpret_cast<const int8_t *>(fb_index);
    p.noskip = noskip;
    av1pf_request r{};
    r.stages = AV1PF_STAGE_CDEF;
    r.cdef = &p;
    if ((rc = validate_request(c, recon, &r))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    c->apply_dec = decision;
    c->apply_idx = fb_index;
    rc = run_batch(c, 1, recon, &r, static_cast<hipStream_t>(stream));
    c->apply_dec = nullptr;
    c->apply_idx = nullptr;
    return rc;
}

int av1pf_deblock_build(av1pf_ctx *c, const av1pf_lf_record *records, const av1pf_lf_build_params *p,
                        uint8_t *level_cache, av1pf_lf_mask *masks, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    if (!records || !p || !level_cache || !masks) return fail(c, AV1PF_ERR_INVALID_ARG, "deblock build: null argument");
    if (reinterpret_cast<uintptr_t>(level_cache) & 3)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock build: level cache must be 4-byte aligned");
    // the MI grid of a frame is ((W+7)>>3)<<1 by ((H+7)>>3)<<1: even, at most 16384 pixels
    if (p->mi_cols <= 0 || p->mi_rows <= 0 || (p->mi_cols & 1) || (p->mi_rows & 1) || p->mi_cols > 4096 ||
        p->mi_rows > 4096)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock build: MI grid must be even and at most 4096 cells a side");
    if (p->ss_x < 0 || p->ss_x > 1 || p->ss_y < 0 || p->ss_y > 1 || p->ss_y > p->ss_x)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock build: unsupported chroma subsampling");
    if (p->level_y[0] > 63 || p->level_y[1] > 63 || p->level_u > 63 || p->level_v > 63)
        return fail(c, AV1PF_ERR_INVALID_ARG, "deblock build: level out of range");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_DEBLOCK_BUILD, s, 0);
    HIPCHK(c, launch_deblock_build(records, *p, level_cache, masks, s));
    return AV1PF_OK;
}

int av1pf_superres(av1pf_ctx *c, const av1pf_frame *coded, const av1pf_frame *upscaled, void *stream)
{
    av1pf_request r{};
    r.stages = AV1PF_STAGE_SUPERRES;
    r.upscaled = upscaled;
    return av1pf_apply(c, coded, &r, stream);
}

int av1pf_filmgrain(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *dst, const av1pf_grain_params *p, void *stream)
{
    av1pf_request r{};
    r.stages = AV1PF_STAGE_FILMGRAIN;
    r.grain = p;
    r.display = dst;
    return av1pf_apply(c, src, &r, stream);
}

int av1pf_cdef_search(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon,
                      const uint64_t *noskip, int damping, uint64_t *mse, void *stream)
{
    if (!c || !mse) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != recon->bit_depth ||
        (recon->bit_depth != 8 && recon->bit_depth != 10) || recon->ss_x != 1 || recon->ss_y != 1 || src->ss_x != 1 ||
        src->ss_y != 1 || recon->mono || (recon->width & 7) || (recon->height & 7))
        return fail(c, AV1PF_ERR_INVALID_ARG,
                    "search needs matching 8-bit or 10-bit 4:2:0 frames with sizes a multiple of 8");
    if (damping < 3 || damping > 6) return fail(c, AV1PF_ERR_INVALID_ARG, "damping out of range");
    HIPCHK(c, hipSetDevice(c->device));
    const int mi_cols = ((recon->width + 7) >> 3) << 1, mi_rows = ((recon->height + 7) >> 3) << 1;
    const int nhfb = (mi_cols + 15) / 16, nvfb = (mi_rows + 15) / 16;
    const FrameView sv = view_of(src);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_SEARCH, s, 2 * view_bytes(sv));
    HIPCHK(c, launch_cdef_search(sv, view_of(recon), noskip, damping, mse, nhfb, nvfb, s));
    return AV1PF_OK;
}

int av1pf_cdef_search_select(av1pf_ctx *c, const av1pf_frame *recon, const uint64_t *noskip, const uint64_t *mse,
                             uint64_t lambda, av1pf_cdef_decision *decision, uint8_t *fb_index, void *stream)
{
    if (!c || !mse || !decision || !fb_index) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, recon);
    if (rc) return rc;
    const int mi_cols = ((recon->width + 7) >> 3) << 1, mi_rows = ((recon->height + 7) >> 3) << 1;
    const int nhfb = (mi_cols + 15) / 16, nvfb = (mi_rows + 15) / 16;
    if (nhfb * nvfb > kMaxSearchBlocks) return fail(c, AV1PF_ERR_INVALID_ARG, "more than 4096 filter blocks");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->finish_state) {
        HIPCHK(c, hipMalloc(&c->finish_state, finish_state_bytes()));
        HIPCHK(c, hipMemset(c->finish_state, 0, finish_state_bytes()));
        c->select_epoch = 0;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_SEARCH_SELECT, s, 0);
    HIPCHK(c, launch_cdef_select(mse, noskip, nhfb, nvfb, mi_cols, mi_rows, lambda, c->finish_state, ++c->select_epoch,
                                 decision, fb_index, c->dev_err, s));
    return AV1PF_OK;
}

int av1pf_lr_search_masked(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                           const uint8_t *unit_mask, av1pf_lr_unit *units, uint64_t *sse, void *stream)
{
    if (!c || !units) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != recon->bit_depth ||
        (recon->bit_depth != 8 && recon->bit_depth != 10))
        return fail(c, AV1PF_ERR_INVALID_ARG, "lr search needs two 8-bit or two 10-bit frames of one size");
    if (unit_size != 32 && unit_size != 64 && unit_size != 128 && unit_size != 256)
        return fail(c, AV1PF_ERR_INVALID_ARG, "lr unit size must be 32, 64, 128 or 256");
    HIPCHK(c, hipSetDevice(c->device));
    const int W = recon->width, H = recon->height;
    const int hz = std::max((W + (unit_size >> 1)) / unit_size, 1), vt = std::max((H + (unit_size >> 1)) / unit_size, 1);
    const FrameView sv = view_of(src), rv = view_of(recon);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_LR_SEARCH, s, 2.0 * W * H);
    HIPCHK(c, launch_lr_search(sv.p[0], rv.p[0], rv.bd, unit_size, hz, vt, unit_mask, units, sse, s));
    return AV1PF_OK;
}

int av1pf_lr_search_exhaustive(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                               av1pf_lr_unit *units, uint64_t *sse, int8_t *fit, void *stream)
{
    if (!c || !units || !sse || !fit) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != 8 || recon->bit_depth != 8)
        return fail(c, AV1PF_ERR_INVALID_ARG, "lr search needs two 8-bit frames of one size");
    if (unit_size != 32 && unit_size != 64 && unit_size != 128 && unit_size != 256)
        return fail(c, AV1PF_ERR_INVALID_ARG, "lr unit size must be 32, 64, 128 or 256");
    HIPCHK(c, hipSetDevice(c->device));
    const int W = recon->width, H = recon->height;
    const int hz = std::max((W + (unit_size >> 1)) / unit_size, 1), vt = std::max((H + (unit_size >> 1)) / unit_size, 1);
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_LR_SEARCH_EXH, s, 2.0 * W * H);
    HIPCHK(c, launch_lr_search_exhaustive(view_of(src).p[0], view_of(recon).p[0], unit_size, hz, vt, units, sse, fit, s));
    return AV1PF_OK;
}

// ---------------------------------------------------------------------------
// Learned proposal network (nn.hip): parameters, workspace, passes.
struct av1pf_nn {
    av1pf_ctx *ctx;
    int max_batch;
    NNShape shape;
    float *P = nullptr, *m = nullptr, *v = nullptr;   // FP32 master weights, Adam moments
    uint16_t *Pb = nullptr;                           // BF16 copy read by the GEMMs
    uint16_t *frag = nullptr;                         // conv weights as MFMA B fragments (fused inference)
    char *work = nullptr;
    NNWork ws;
    NNWeights w;
    int step = 0;
};

int64_t av1pf_nn_param_count(void) { return nn_shape().total; }
int av1pf_search_shortlist_size(void) { return AV1PF_SHORTLIST; }

void av1pf_nn_destroy(av1pf_nn *nn)
{
    if (!nn) return;
    hipSetDevice(nn->ctx->device);
    hipDeviceSynchronize();
    hipFree(nn->P);
    hipFree(nn->m);
    hipFree(nn->v);
    hipFree(nn->Pb);
    hipFree(nn->frag);
    hipFree(nn->work);
    delete nn;
}

int av1pf_nn_create(av1pf_ctx *c, int max_batch, av1pf_nn **out)
{
    if (!c || !out || max_batch < 1 || max_batch > 4096) return AV1PF_ERR_INVALID_ARG;
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    av1pf_nn *nn = new av1pf_nn();
    nn->ctx = c;
    nn->max_batch = max_batch;
    nn->shape = nn_shape();
    const size_t P = (size_t)nn->shape.total;
    bool ok = hipMalloc(&nn->P, P * 4) == hipSuccess && hipMalloc(&nn->m, P * 4) == hipSuccess &&
              hipMalloc(&nn->v, P * 4) == hipSuccess && hipMalloc(&nn->Pb, P * 2) == hipSuccess &&
              hipMalloc(&nn->frag, NN_FRAG_ELEMS * 2) == hipSuccess &&
              hipMalloc(&nn->work, nn_work_bytes(max_batch)) == hipSuccess;
    if (!ok) {
        av1pf_nn_destroy(nn);
        return fail(c, AV1PF_ERR_DEVICE, "out of device memory for the network");
    }
    nn_work_carve(nn->work, max_batch, nn->ws);
    for (int l = 0; l < NN_LAYERS; l++) {
        nn->w.wb[l] = nn->Pb + nn->shape.woff[l];
        nn->w.b[l] = nn->P + nn->shape.boff[l];
        nn->w.woff[l] = nn->shape.woff[l];
        nn->w.boff[l] = nn->shape.boff[l];
    }
    *out = nn;
    return AV1PF_OK;
}

int av1pf_nn_set_params(av1pf_nn *nn, const float *params)
{
    if (!nn || !params) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const size_t P = (size_t)nn->shape.total;
    std::vector<uint16_t> b(P);
    for (size_t i = 0; i < P; i++) {
        uint32_t u;
        memcpy(&u, &params[i], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        b[i] = (uint16_t)(u >> 16);
    }
    HIPCHK(c, hipMemcpy(nn->P, params, P * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(nn->Pb, b.data(), P * 2, hipMemcpyHostToDevice));
    std::vector<float> z(P, 0.f);
    HIPCHK(c, hipMemcpy(nn->m, z.data(), P * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(nn->v, z.data(), P * 4, hipMemcpyHostToDevice));
    HIPCHK(c, nn_pack_frag(nn->w, nn->frag, nullptr));
    HIPCHK(c, hipDeviceSynchronize());
    nn->step = 0;
    return AV1PF_OK;
}

int av1pf_nn_get_params(av1pf_nn *nn, float *params)
{
    if (!nn || !params) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    HIPCHK(c, hipMemcpy(params, nn->P, (size_t)nn->shape.total * 4, hipMemcpyDeviceToHost));
    return AV1PF_OK;
}

int av1pf_nn_tiles(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon, uint16_t *tiles, void *stream)
{
    if (!c || !tiles) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != recon->bit_depth)
        return fail(c, AV1PF_ERR_INVALID_ARG, "source and reconstruction differ in geometry");
    HIPCHK(c, hipSetDevice(c->device));
    const int nhfb = (recon->width + 63) / 64, nvfb = (recon->height + 63) / 64;
    const FrameView sv = view_of(src), rv = view_of(recon);
    HIPCHK(c, launch_tiles(sv.p[0], rv.p[0], recon->bit_depth, nhfb, nhfb * nvfb, tiles, static_cast<hipStream_t>(stream)));
    return AV1PF_OK;
}

int av1pf_nn_forward(av1pf_nn *nn, const uint16_t *tiles, int n, float *logits, void *stream)
{
    if (!nn || !tiles || !logits || n < 0) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_NN_FORWARD, s, 0);
    for (int i = 0; i < n; i += nn->max_batch) {
        const int b = std::min(nn->max_batch, n - i);
        HIPCHK(c, nn_forward(nn->w, nn->ws, tiles + (size_t)i * 4096 * 2, b, logits + (size_t)i * NN_OUT, s));
    }
    return AV1PF_OK;
}

namespace {
NNInfer infer_args(const av1pf_nn *nn)
{
    NNInfer p{};
    p.frag = nn->frag;
    for (int l = 0; l < 3; l++) p.bias[l] = nn->w.b[l];
    p.w3 = nn->w.wb[3];
    p.w4 = nn->w.wb[4];
    p.b3 = nn->w.b[3];
    p.b4 = nn->w.b[4];
    return p;
}
}  // namespace

int av1pf_nn_infer(av1pf_nn *nn, const av1pf_frame *src, const av1pf_frame *recon, float *logits, void *stream)
{
    if (!nn || !logits) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != recon->bit_depth ||
        (recon->bit_depth != 8 && recon->bit_depth != 10))
        return fail(c, AV1PF_ERR_INVALID_ARG, "fused inference needs two 8-bit or two 10-bit frames of one geometry");
    HIPCHK(c, hipSetDevice(c->device));
    const int nhfb = (recon->width + 63) / 64, nvfb = (recon->height + 63) / 64;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_NN_FORWARD, s, 0);
    HIPCHK(c, nn_infer(infer_args(nn), view_of(src).p[0], view_of(recon).p[0], recon->bit_depth, nhfb, nhfb * nvfb, logits, nullptr, s));
    return AV1PF_OK;
}

int av1pf_nn_backward(av1pf_nn *nn, const uint16_t *tiles, int n, const float *targets, float *grads, float *loss,
                      void *stream)
{
    if (!nn || !tiles || !targets || !grads || !loss || n < 1 || n > nn->max_batch) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_NN_TRAIN, s, 0);
    HIPCHK(c, nn_zero(loss, 1, s));
    HIPCHK(c, nn_backward(nn->w, nn->ws, tiles, n, targets, grads, loss, s));
    return AV1PF_OK;
}

int av1pf_nn_adam(av1pf_nn *nn, const float *grads, float lr, float weight_decay, void *stream)
{
    if (!nn || !grads || !(lr > 0.f)) return AV1PF_ERR_INVALID_ARG;
    av1pf_ctx *c = nn->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    nn->step++;
    HIPCHK(c, nn_adam(nn->P, nn->Pb, grads, nn->m, nn->v, (int)nn->shape.total, lr, nn->step, weight_decay,
                      static_cast<hipStream_t>(stream)));
    HIPCHK(c, nn_pack_frag(nn->w, nn->frag, static_cast<hipStream_t>(stream)));
    return AV1PF_OK;
}

int av1pf_lr_search(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                    av1pf_lr_unit *units, uint64_t *sse, void *stream)
{
    return av1pf_lr_search_masked(c, src, recon, unit_size, nullptr, units, sse, stream);
}

namespace {
int ensure_small_state(av1pf_ctx *c)
{
    if (c->small_state) return AV1PF_OK;
    HIPCHK(c, hipMalloc(&c->small_state, small_state_bytes()));
    HIPCHK(c, hipMemset(c->small_state, 0, small_state_bytes()));
    return AV1PF_OK;
}

int validate_search_pair(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon)
{
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || src->bit_depth != recon->bit_depth ||
        (recon->bit_depth != 8 && recon->bit_depth != 10) || recon->ss_x != 1 || recon->ss_y != 1 || src->ss_x != 1 ||
        src->ss_y != 1 || recon->mono || (recon->width & 7) || (recon->height & 7))
        return fail(c, AV1PF_ERR_INVALID_ARG,
                    "search needs matching 8-bit or 10-bit 4:2:0 frames with sizes a multiple of 8");
    return AV1PF_OK;
}
}  // namespace

int av1pf_cdef_search_shortlist(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon,
                                const uint64_t *noskip, int damping, uint64_t lambda, const int32_t *cand,
                                uint64_t *mse_sub, av1pf_cdef_decision *decision, uint8_t *fb_index, void *stream)
{
    if (!c || !cand || !mse_sub || !decision || !fb_index) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_search_pair(c, src, recon);
    if (rc) return rc;
    if (damping < 3 || damping > 6) return fail(c, AV1PF_ERR_INVALID_ARG, "damping out of range");
    const int nfb = ((recon->width + 63) / 64) * ((recon->height + 63) / 64);
    if (nfb > 2048) return fail(c, AV1PF_ERR_INVALID_ARG, "more than 2048 filter blocks");
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_SEARCH_LEARNED, s, 0);
    if ((rc = ensure_small_state(c))) return rc;
    HIPCHK(c, launch_search_shortlist(view_of(src), view_of(recon), noskip, damping, lambda, cand, mse_sub, decision,
                                      fb_index, c->small_state, s));
    return AV1PF_OK;
}

int av1pf_search_learned(av1pf_ctx *c, av1pf_nn *nn, const av1pf_frame *src, const av1pf_frame *recon,
                         const uint64_t *noskip, int damping, uint64_t lambda, float lr_none_threshold,
                         av1pf_cdef_decision *decision, uint8_t *fb_index, av1pf_lr_unit *lr_units, int32_t *cand,
                         float *conf, void *stream)
{
    if (!c || !nn || !decision || !fb_index || !lr_units || !cand) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_search_pair(c, src, recon);
    if (rc) return rc;
    if (damping < 3 || damping > 6) return fail(c, AV1PF_ERR_INVALID_ARG, "damping out of range");
    const int nhfb = (recon->width + 63) / 64, nvfb = (recon->height + 63) / 64, nfb = nhfb * nvfb;
    if (nfb > 2048) return fail(c, AV1PF_ERR_INVALID_ARG, "more than 2048 filter blocks");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t logits_b = (size_t)nfb * NN_OUT * 4, probs_b = (size_t)nfb * NN_PROP_STRIDE * 4;
    const size_t msub_b = (size_t)2 * nfb * NCAND * 8, mask_b = (size_t)nfb + 256;
    const size_t part_b = (size_t)((nfb + 63) / 64) * NN_PROP_STRIDE * 4;
    const size_t need = align256(logits_b) + align256(probs_b) + align256(msub_b) + align256(mask_b) +
                        align256(part_b) + 256;
    if ((rc = ensure(c, (void **)&c->learn_buf, &c->learn_cap, need))) return rc;
    char *p = c->learn_buf;
    float *logits = (float *)p; p += align256(logits_b);
    float *probs = (float *)p; p += align256(probs_b);
    uint64_t *msub = (uint64_t *)p; p += align256(msub_b);
    uint8_t *mask = (uint8_t *)p; p += align256(mask_b);
    float *partial = (float *)p; p += align256(part_b);
    float *conf_tmp = (float *)p;
    hipStream_t s = static_cast<hipStream_t>(stream);
    StageTimer t(c, ST_SEARCH_LEARNED, s, 0);
    const FrameView sv = view_of(src), rv = view_of(recon);
    if ((rc = ensure_small_state(c))) return rc;
    NNPropose prop{};
    prop.noskip = noskip;
    prop.mi_cols = ((recon->width + 7) >> 3) << 1;
    prop.mi_rows = ((recon->height + 7) >> 3) << 1;
    prop.probs = probs;
    prop.partial = partial;
    prop.cand = cand;
    prop.conf = conf ? conf : conf_tmp;
    HIPCHK(c, nn_infer(infer_args(nn), sv.p[0], rv.p[0], rv.bd, nhfb, nfb, logits, &prop, s));
    // the LR search depends only on the logits: it runs on the side stream
    // beside the proposal, the CDEF shortlist search and the selection (the
    // main stream's launches are issued first, so nothing on it waits for
    // the host to issue the side stream's)
    HIPCHK(c, hipEventRecord(c->learn_fork, s));
    HIPCHK(c, nn_propose(prop, nfb, s));
    HIPCHK(c, launch_search_shortlist(sv, rv, noskip, damping, lambda, cand, msub, decision, fb_index, c->small_state,
                                      s));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->learn_fork, 0));
    const int W = recon->width, H = recon->height;
    const int hz = std::max((W + 32) / 64, 1), vt = std::max((H + 32) / 64, 1);
    HIPCHK(c, launch_lr_mask(logits, nhfb, nvfb, hz, vt, lr_none_threshold, mask, c->side));
    HIPCHK(c, launch_lr_search(sv.p[0], rv.p[0], rv.bd, 64, hz, vt, mask, lr_units, nullptr, c->side));
    HIPCHK(c, hipEventRecord(c->learn_join, c->side));
    HIPCHK(c, hipStreamWaitEvent(s, c->learn_join, 0));
    return AV1PF_OK;
}

int av1pf_profile(av1pf_ctx *c, int enable)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    for (auto &r : c->recs) { c->pool.push_back(r.a); c->pool.push_back(r.b); }
    c->recs.clear();
    for (int i = 0; i < AV1PF_PROFILE_STAGES; i++) { c->acc_ms[i] = 0; c->acc_n[i] = 0; c->acc_bytes[i] = 0; }
    c->profiling = enable != 0;
    // events are created up front: creating them inside a timed region costs
    // more host time than the launches they bracket
    while (c->profiling && c->pool.size() < 512) {
        hipEvent_t e = nullptr;
        HIPCHK(c, hipEventCreate(&e));
        c->pool.push_back(e);
    }
    return AV1PF_OK;
}

int av1pf_profile_read(av1pf_ctx *c, double ms[AV1PF_PROFILE_STAGES], int count[AV1PF_PROFILE_STAGES],
                       double bytes[AV1PF_PROFILE_STAGES])
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    for (auto &r : c->recs) {
        float e = 0.f;
        HIPCHK(c, hipEventElapsedTime(&e, r.a, r.b));
        c->acc_ms[r.stage] += e;
        c->acc_n[r.stage] += 1;
        c->acc_bytes[r.stage] += r.bytes;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    for (int i = 0; i < AV1PF_PROFILE_STAGES; i++) {
        if (ms) ms[i] = c->acc_ms[i];
        if (count) count[i] = c->acc_n[i];
        if (bytes) bytes[i] = c->acc_bytes[i];
    }
    return AV1PF_OK;
}

}  // extern "C"
