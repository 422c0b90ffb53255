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
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "pf_internal.h"

using namespace pf;

struct av1pf_ctx {
    int device = 0;
    std::string err;
    char *scratch = nullptr;      // second frame buffer (CDEF output / LR ping-pong)
    size_t scratch_cap = 0;
    uint16_t *lr_bnd = nullptr;   // LR stripe context rows
    size_t lr_bnd_cap = 0;
    int16_t *templates = nullptr; // film-grain templates (3 x 73 x 82)
    int *dev_err = nullptr;       // device-side validation flag
    uint64_t *search_tmp = nullptr;
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
enum { ST_DEBLOCK = 0, ST_LR_SAVE, ST_CDEF, ST_LR, ST_COPY, ST_GRAIN_T, ST_GRAIN_A, ST_SEARCH };

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

double frame_bytes(const pf::FrameView &f)
{
    double n = 0;
    for (int i = 0; i < f.nplanes; i++) n += (double)f.p[i].w * f.p[i].h;
    return n * (f.bd == 8 ? 1 : 2);
}
}  // namespace

namespace {

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

__global__ void zero_int_kernel(int *p) { *p = 0; }

int bps_of(int bd) { return bd == 8 ? 1 : 2; }

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
    }
    return v;
}

int ensure_scratch(av1pf_ctx *c, const FrameView &f, FrameView *out)
{
    const int bps = bps_of(f.bd);
    size_t need = 0;
    int64_t strides[3];
    for (int i = 0; i < f.nplanes; i++) {
        strides[i] = ((int64_t)f.p[i].aw * bps + 255) & ~(int64_t)255;
        need += (size_t)strides[i] * f.p[i].ah;
    }
    if (need > c->scratch_cap) {
        if (c->scratch) HIPCHK(c, hipFree(c->scratch));
        c->scratch = nullptr;
        HIPCHK(c, hipMalloc(&c->scratch, need));
        c->scratch_cap = need;
    }
    *out = f;
    size_t off = 0;
    for (int i = 0; i < f.nplanes; i++) {
        out->p[i].base = c->scratch + off;
        out->p[i].stride = strides[i];
        off += (size_t)strides[i] * f.p[i].ah;
    }
    return AV1PF_OK;
}

int validate_deblock(av1pf_ctx *c, const av1pf_frame *f, const av1pf_deblock_params *p)
{
    if (!p || !p->level_cache || !p->masks) return fail(c, AV1PF_ERR_INVALID_ARG, "deblock params missing");
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
    const int us = p->unit_size[pl];
    if (us != 32 && us != 64 && us != 128 && us != 256) return AV1PF_ERR_INVALID_ARG;
    if (us < (64 >> sx)) return AV1PF_ERR_INVALID_ARG;
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
    const bool noop = !g->chroma_scaling_from_luma && !g->num_y_points && !g->num_cb_points && !g->num_cr_points;
    if (noop) return AV1PF_OK;
    if (g->scaling_shift < 8 || g->scaling_shift > 11) return fail(c, AV1PF_ERR_INVALID_ARG, "scaling shift out of range");
    if ((g->ar_coeff_lag || g->num_y_points) && (g->ar_coeff_shift < 6 || g->ar_coeff_shift > 9))
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

// AV1 scaling LUT (filmgrain/scaling.go:18-51) -- tiny host work, LUTs travel as kernel args.
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

int run_grain(av1pf_ctx *c, const av1pf_frame *frame, const FrameView &A, const av1pf_frame *display,
              const av1pf_grain_params *g, hipStream_t s)
{
    FrameView D = view_of(display);
    const bool noop = !g->chroma_scaling_from_luma && !g->num_y_points && !g->num_cb_points && !g->num_cr_points;
    if (noop) {
        HIPCHK(c, launch_copy(A, D, s));
        return AV1PF_OK;
    }
    GrainArgs ga{};
    ga.g = *g;
    ga.active[0] = g->num_y_points != 0;
    ga.active[1] = !frame->mono && (g->num_cb_points != 0 || g->chroma_scaling_from_luma);
    ga.active[2] = !frame->mono && (g->num_cr_points != 0 || g->chroma_scaling_from_luma);
    build_scaling_lut(ga.luts[0], (const uint8_t(*)[2])g->y_points, g->num_y_points);
    if (g->chroma_scaling_from_luma) {
        memcpy(ga.luts[1], ga.luts[0], 256);
        memcpy(ga.luts[2], ga.luts[0], 256);
    } else {
        build_scaling_lut(ga.luts[1], (const uint8_t(*)[2])g->cb_points, g->num_cb_points);
        build_scaling_lut(ga.luts[2], (const uint8_t(*)[2])g->cr_points, g->num_cr_points);
    }
    lfsr_jump(ga.jump_luma, 82);
    lfsr_jump(ga.jump_chroma, frame->ss_x ? 44 : 82);
    if (!c->templates) HIPCHK(c, hipMalloc(&c->templates, 3 * 73 * 82 * sizeof(int16_t)));
    ga.templates = c->templates;
    // copy the frame into the display buffer first (bindFilmGrainPostFilterOutput), then grain
    { StageTimer t(c, ST_COPY, s, 2 * frame_bytes(A)); HIPCHK(c, launch_copy(A, D, s)); }
    { StageTimer t(c, ST_GRAIN_T, s, 0); HIPCHK(c, launch_grain_templates(ga, frame->bit_depth, frame->ss_x, frame->ss_y, s)); }
    { StageTimer t(c, ST_GRAIN_A, s, 2 * frame_bytes(A)); HIPCHK(c, launch_grain_apply(A, D, ga, frame->matrix_identity, s)); }
    return AV1PF_OK;
}

int lr_args(av1pf_ctx *c, const av1pf_frame *f, const av1pf_lr_params *p, LrArgs *out, bool *any)
{
    *out = LrArgs{};
    *any = false;
    const int np = f->mono ? 1 : 3;
    size_t need = 0;
    int stripes[3] = { 0, 0, 0 }, pws[3] = { 0, 0, 0 };
    for (int i = 0; i < np; i++) {
        LrPlaneArgs &a = out->p[i];
        a.type = p->frame_type[i];
        if (a.type == AV1PF_LR_NONE) continue;
        int pw, ph, hz, vt;
        lr_geometry(f, p, i, &pw, &ph, &hz, &vt);
        const int sy = i ? f->ss_y : 0;
        const int sh = 64 >> sy, off = 8 >> sy;
        a.unit_size = p->unit_size[i];
        a.horz = hz;
        a.vert = vt;
        a.units = p->units[i];
        stripes[i] = (ph + off + sh - 1) / sh;
        pws[i] = pw;
        a.stripes = stripes[i];
        need += (size_t)stripes[i] * 4 * pw;
        *any = true;
    }
    if (!*any) return AV1PF_OK;
    if (need > c->lr_bnd_cap) {
        if (c->lr_bnd) HIPCHK(c, hipFree(c->lr_bnd));
        c->lr_bnd = nullptr;
        HIPCHK(c, hipMalloc(&c->lr_bnd, need * sizeof(uint16_t)));
        c->lr_bnd_cap = need;
    }
    size_t off = 0;
    for (int i = 0; i < np; i++) {
        if (out->p[i].type == AV1PF_LR_NONE) continue;
        out->p[i].bnd = c->lr_bnd + off;
        off += (size_t)stripes[i] * 4 * pws[i];
    }
    return AV1PF_OK;
}

bool cdef_has_filtering(const av1pf_cdef_params *p, bool mono, int *yf, int *uvf)
{
    *yf = *uvf = 0;
    for (int i = 0; i < p->strength_count && i < 8; i++) {
        if (p->y_strength[i]) *yf = 1;
        if (!mono && p->uv_strength[i]) *uvf = 1;
    }
    return *yf || *uvf;
}

int validate_request(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_request *req)
{
    if (!req) return fail(c, AV1PF_ERR_INVALID_ARG, "null request");
    int rc = validate_frame(c, frame);
    if (rc) return rc;
    const uint32_t st = req->stages;
    if (st & AV1PF_STAGE_SUPERRES) return fail(c, AV1PF_ERR_UNSUPPORTED, "superres stage is not supported");
    if (st & ~(AV1PF_STAGE_DEBLOCK | AV1PF_STAGE_CDEF | AV1PF_STAGE_LR | AV1PF_STAGE_FILMGRAIN))
        return fail(c, AV1PF_ERR_INVALID_ARG, "unknown stage bits");
    if ((st & AV1PF_STAGE_DEBLOCK) && (rc = validate_deblock(c, frame, req->deblock))) return rc;
    if ((st & AV1PF_STAGE_CDEF) && (rc = validate_cdef(c, frame, req->cdef))) return rc;
    if ((st & AV1PF_STAGE_LR) && (rc = validate_lr(c, frame, req->lr))) return rc;
    if ((st & AV1PF_STAGE_FILMGRAIN) && (rc = validate_grain(c, frame, req->grain, req->display))) return rc;
    return AV1PF_OK;
}

int run_chain(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_request *req, hipStream_t s)
{
    const uint32_t st = req->stages;
    FrameView A = view_of(frame);
    FrameView B;
    int rc = ensure_scratch(c, A, &B);
    if (rc) return rc;
    const FrameView *cur = &A;
    if (st & AV1PF_STAGE_DEBLOCK) {
        const av1pf_deblock_params *p = req->deblock;
        DeblockArgs a{};
        a.level_cache = p->level_cache;
        a.masks = p->masks;
        a.mi_cols = p->mi_cols;
        a.sb128_w = p->sb128_w;
        for (int l = 0; l < 64; l++) {
            int shift = 0;
            if (p->sharpness > 0) shift++;
            if (p->sharpness > 4) shift++;
            int lim = l >> shift;
            if (p->sharpness > 0 && lim > 9 - p->sharpness) lim = 9 - p->sharpness;
            if (lim < 1) lim = 1;
            a.I[l] = (uint8_t)lim;
            a.E[l] = (uint8_t)(2 * (l + 2) + lim);
        }
        const bool luma_zero = p->level_y[0] == 0 && p->level_y[1] == 0;
        a.plane_mask = 1;
        if (!frame->mono && !luma_zero) {
            if (p->level_u) a.plane_mask |= 2;
            if (p->level_v) a.plane_mask |= 4;
        }
        { StageTimer t(c, ST_DEBLOCK, s, 2 * frame_bytes(A)); HIPCHK(c, launch_deblock(A, a, s)); }
    }
    LrArgs la{};
    bool lr_any = false;
    if (st & AV1PF_STAGE_LR) {
        if ((rc = lr_args(c, frame, req->lr, &la, &lr_any))) return rc;
        if (lr_any) { StageTimer t(c, ST_LR_SAVE, s, 0); HIPCHK(c, launch_lr_save(A, la, s)); }
    }
    if (st & AV1PF_STAGE_CDEF) {
        const av1pf_cdef_params *p = req->cdef;
        CdefArgs a{};
        if (cdef_has_filtering(p, frame->mono, &a.yfilt, &a.uvfilt)) {
            a.index = p->index;
            a.noskip = p->noskip;
            a.unit_cols = p->unit_cols;
            a.unit_rows = p->unit_rows;
            a.damping = p->damping;
            a.strength_count = p->strength_count;
            memcpy(a.y_strength, p->y_strength, 8);
            memcpy(a.uv_strength, p->uv_strength, 8);
            { StageTimer t(c, ST_CDEF, s, 2 * frame_bytes(A)); HIPCHK(c, launch_cdef(A, B, a, s, c->dev_err)); }
            cur = &B;
        }
    }
    if (lr_any) {
        const FrameView *other = cur == &A ? &B : &A;
        { StageTimer t(c, ST_LR, s, 2 * frame_bytes(A)); HIPCHK(c, launch_lr(*cur, *other, la, s)); }
        cur = other;
    }
    if (cur != &A) { StageTimer t(c, ST_COPY, s, 2 * frame_bytes(A)); HIPCHK(c, launch_copy(*cur, A, s)); }
    if (st & AV1PF_STAGE_FILMGRAIN)
        if ((rc = run_grain(c, frame, A, req->display, req->grain, s))) return rc;
    return AV1PF_OK;
}

}  // namespace

extern "C" {

const char *av1pf_version(void) { return "av1pf 0.1 gfx950"; }

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
    if (hipMalloc(&c->dev_err, sizeof(int)) != hipSuccess) {
        delete c;
        return AV1PF_ERR_DEVICE;
    }
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, nullptr, c->dev_err);
    if (hipDeviceSynchronize() != hipSuccess) {
        hipFree(c->dev_err);
        delete c;
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
    if (c->scratch) hipFree(c->scratch);
    if (c->lr_bnd) hipFree(c->lr_bnd);
    if (c->templates) hipFree(c->templates);
    if (c->dev_err) hipFree(c->dev_err);
    if (c->search_tmp) hipFree(c->search_tmp);
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
    return run_chain(c, frame, req, static_cast<hipStream_t>(stream));
}

int av1pf_apply_batch(av1pf_ctx *c, int n, const av1pf_frame *frames, const av1pf_request *reqs, void *stream)
{
    if (!c || n < 0 || (n && (!frames || !reqs))) return AV1PF_ERR_INVALID_ARG;
    for (int i = 0; i < n; i++) {
        int rc = validate_request(c, &frames[i], &reqs[i]);
        if (rc) return rc;
    }
    HIPCHK(c, hipSetDevice(c->device));
    for (int i = 0; i < n; i++) {
        int rc = run_chain(c, &frames[i], &reqs[i], static_cast<hipStream_t>(stream));
        if (rc) return rc;
    }
    return AV1PF_OK;
}

int av1pf_sync(av1pf_ctx *c, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    int flag = 0;
    HIPCHK(c, hipMemcpy(&flag, c->dev_err, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), c->dev_err);
        HIPCHK(c, hipStreamSynchronize(static_cast<hipStream_t>(stream)));
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
    // The chain reads frame-edge context straight from the post-CDEF frame, so
    // only the deblocked seam rows are stored; the after-CDEF save is a no-op.
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, frame);
    if (rc) return rc;
    if ((rc = validate_lr(c, frame, p))) return rc;
    if (after_cdef) return AV1PF_OK;
    HIPCHK(c, hipSetDevice(c->device));
    LrArgs la;
    bool any;
    if ((rc = lr_args(c, frame, p, &la, &any))) return rc;
    if (any) HIPCHK(c, launch_lr_save(view_of(frame), la, static_cast<hipStream_t>(stream)));
    return AV1PF_OK;
}

int av1pf_lr(av1pf_ctx *c, const av1pf_frame *frame, const av1pf_lr_params *p, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, frame);
    if (rc) return rc;
    if ((rc = validate_lr(c, frame, p))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    LrArgs la;
    bool any;
    if ((rc = lr_args(c, frame, p, &la, &any))) return rc;
    if (!any) return AV1PF_OK;
    FrameView A = view_of(frame), B;
    if ((rc = ensure_scratch(c, A, &B))) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIPCHK(c, launch_lr(A, B, la, s));
    HIPCHK(c, launch_copy(B, A, s));
    return AV1PF_OK;
}

int av1pf_filmgrain(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *dst, const av1pf_grain_params *p, void *stream)
{
    if (!c) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_grain(c, src, p, dst))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return run_grain(c, src, view_of(src), dst, p, static_cast<hipStream_t>(stream));
}

int av1pf_cdef_search(av1pf_ctx *c, const av1pf_frame *src, const av1pf_frame *recon,
                      const uint64_t *noskip, int damping, uint64_t *mse, void *stream)
{
    if (!c || !mse) return AV1PF_ERR_INVALID_ARG;
    int rc = validate_frame(c, src);
    if (rc) return rc;
    if ((rc = validate_frame(c, recon))) return rc;
    if (src->width != recon->width || src->height != recon->height || recon->bit_depth != 8 ||
        src->bit_depth != 8 || recon->ss_x != 1 || recon->ss_y != 1 || src->ss_x != 1 || src->ss_y != 1 ||
        recon->mono || (recon->width & 7) || (recon->height & 7))
        return fail(c, AV1PF_ERR_INVALID_ARG, "search needs matching 8-bit 4:2:0 frames with sizes a multiple of 8");
    if (damping < 3 || damping > 6) return fail(c, AV1PF_ERR_INVALID_ARG, "damping out of range");
    HIPCHK(c, hipSetDevice(c->device));
    const int mi_cols = ((recon->width + 7) >> 3) << 1, mi_rows = ((recon->height + 7) >> 3) << 1;
    const int nhfb = (mi_cols + 15) / 16, nvfb = (mi_rows + 15) / 16;
    const FrameView sv = view_of(src);
    StageTimer t(c, ST_SEARCH, static_cast<hipStream_t>(stream), 2 * frame_bytes(sv));
    HIPCHK(c, launch_cdef_search(sv, view_of(recon), noskip, damping, mse, nhfb, nvfb,
                                 static_cast<hipStream_t>(stream)));
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
