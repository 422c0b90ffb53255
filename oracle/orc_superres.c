/*
 * CPU oracle: AV1 super-resolution upscale.  Test infrastructure only.
 *
 * Restates superres/superres.go:38-107 (UpscalePlaneCoded, upscaleRowPureGo)
 * and its two callers: the frame upscale of ApplySuperResPostFilter
 * (decoder/postfilter_superres.go:198-264: the source is the MI-aligned coded
 * span, the step and phase come from the coded plane width, the visible rows
 * are upscaled) and the deblocked LR context rows upscaled on save
 * (tile/restoration_boundary.go:257-345).
 */
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* AV1 normative upscale filter, 64 phases x 8 taps (superres.go:144-180). */
static const int16_t kUpscaleFilter[64][8] = {
    { 0, 0, 0, 128, 0, 0, 0, 0 },       { 0, 0, -1, 128, 2, -1, 0, 0 },
    { 0, 1, -3, 127, 4, -2, 1, 0 },     { 0, 1, -4, 127, 6, -3, 1, 0 },
    { 0, 2, -6, 126, 8, -3, 1, 0 },     { 0, 2, -7, 125, 11, -4, 1, 0 },
    { -1, 2, -8, 125, 13, -5, 2, 0 },   { -1, 3, -9, 124, 15, -6, 2, 0 },
    { -1, 3, -10, 123, 18, -6, 2, -1 }, { -1, 3, -11, 122, 20, -7, 3, -1 },
    { -1, 4, -12, 121, 22, -8, 3, -1 }, { -1, 4, -13, 120, 25, -9, 3, -1 },
    { -1, 4, -14, 118, 28, -9, 3, -1 }, { -1, 4, -15, 117, 30, -10, 4, -1 },
    { -1, 5, -16, 116, 32, -11, 4, -1 }, { -1, 5, -16, 114, 35, -12, 4, -1 },
    { -1, 5, -17, 112, 38, -12, 4, -1 }, { -1, 5, -18, 111, 40, -13, 5, -1 },
    { -1, 5, -18, 109, 43, -14, 5, -1 }, { -1, 6, -19, 107, 45, -14, 5, -1 },
    { -1, 6, -19, 105, 48, -15, 5, -1 }, { -1, 6, -19, 103, 51, -16, 5, -1 },
    { -1, 6, -20, 101, 53, -16, 6, -1 }, { -1, 6, -20, 99, 56, -17, 6, -1 },
    { -1, 6, -20, 97, 58, -17, 6, -1 }, { -1, 6, -20, 95, 61, -18, 6, -1 },
    { -2, 7, -20, 93, 64, -18, 6, -2 }, { -2, 7, -20, 91, 66, -19, 6, -1 },
    { -2, 7, -20, 88, 69, -19, 6, -1 }, { -2, 7, -20, 86, 71, -19, 6, -1 },
    { -2, 7, -20, 84, 74, -20, 7, -2 }, { -2, 7, -20, 81, 76, -20, 7, -1 },
    { -2, 7, -20, 79, 79, -20, 7, -2 }, { -1, 7, -20, 76, 81, -20, 7, -2 },
    { -2, 7, -20, 74, 84, -20, 7, -2 }, { -1, 6, -19, 71, 86, -20, 7, -2 },
    { -1, 6, -19, 69, 88, -20, 7, -2 }, { -1, 6, -19, 66, 91, -20, 7, -2 },
    { -2, 6, -18, 64, 93, -20, 7, -2 }, { -1, 6, -18, 61, 95, -20, 6, -1 },
    { -1, 6, -17, 58, 97, -20, 6, -1 }, { -1, 6, -17, 56, 99, -20, 6, -1 },
    { -1, 6, -16, 53, 101, -20, 6, -1 }, { -1, 5, -16, 51, 103, -19, 6, -1 },
    { -1, 5, -15, 48, 105, -19, 6, -1 }, { -1, 5, -14, 45, 107, -19, 6, -1 },
    { -1, 5, -14, 43, 109, -18, 5, -1 }, { -1, 5, -13, 40, 111, -18, 5, -1 },
    { -1, 4, -12, 38, 112, -17, 5, -1 }, { -1, 4, -12, 35, 114, -16, 5, -1 },
    { -1, 4, -11, 32, 116, -16, 5, -1 }, { -1, 4, -10, 30, 117, -15, 4, -1 },
    { -1, 3, -9, 28, 118, -14, 4, -1 }, { -1, 3, -9, 25, 120, -13, 4, -1 },
    { -1, 3, -8, 22, 121, -12, 4, -1 }, { -1, 3, -7, 20, 122, -11, 3, -1 },
    { -1, 2, -6, 18, 123, -10, 3, -1 }, { 0, 2, -6, 15, 124, -9, 3, -1 },
    { 0, 2, -5, 13, 125, -8, 2, -1 },   { 0, 1, -4, 11, 125, -7, 2, 0 },
    { 0, 1, -3, 8, 126, -6, 2, 0 },     { 0, 1, -3, 6, 127, -4, 1, 0 },
    { 0, 1, -2, 4, 127, -3, 1, 0 },     { 0, 0, -1, 2, 128, -1, 0, 0 },
};

const int16_t *orc_superres_filter(int phase)
{
    return (phase >= 0 && phase < 64) ? kUpscaleFilter[phase] : NULL;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* superres.go:38-60, 78-107: one row, the whole 8-tap window clamped to the
 * source span [0, srcw - 1] tap by tap. */
int orc_superres_row(const uint16_t *src, int srcw, int codedw, uint16_t *dst, int dstw, int bd)
{
    if (bd < 8 || bd > 12 || codedw <= 0 || codedw > srcw || dstw <= codedw) return AV1PF_ERR_INVALID_ARG;
    const long step = (((long)codedw << 14) + dstw / 2) / dstw;
    const long err = (long)dstw * step - ((long)codedw << 14);
    long x0 = (-(((long)dstw - codedw) << 13) + dstw / 2) / dstw;
    x0 += (1 << 7) - err / 2;
    x0 &= (1 << 14) - 1;
    const int maxv = (1 << bd) - 1;
    for (int x = 0; x < dstw; x++) {
        const long pos = -(1L << 14) + x0 + (long)x * step;
        const long px = pos >> 14;   /* floor, pos may be negative */
        const int phase = (int)((pos & ((1 << 14) - 1)) >> 8);
        const int16_t *f = kUpscaleFilter[phase];
        long sum = 0;
        for (int k = 0; k < 8; k++) sum += (long)src[clampi((int)(px - 3 + k), 0, srcw - 1)] * f[k];
        long v = (sum + 64) >> 7;
        dst[x] = (uint16_t)clampi((int)v, 0, maxv);
    }
    return AV1PF_OK;
}

/* Plane geometry of an upscale (postfilter_superres.go:241-252,
 * postfilter_cdef.go:1390-1412). */
static void sr_geometry(const orc_frame *c, const orc_frame *u, int pl, int *srcw, int *codedw, int *dstw, int *rows)
{
    const int sx = pl ? c->ssx : 0, sy = pl ? c->ssy : 0;
    *codedw = (c->w + sx) >> sx;
    *dstw = (u->w + sx) >> sx;
    *srcw = ((c->w + 7) & ~7) >> sx;
    if (*srcw > c->stride[pl]) *srcw = c->stride[pl];
    *rows = (c->h + sy) >> sy;
}

int orc_superres(const orc_frame *coded, orc_frame *up)
{
    if (!coded || !up || up->bd != coded->bd || up->ssx != coded->ssx || up->ssy != coded->ssy ||
        up->mono != coded->mono || up->h != coded->h || up->w <= coded->w)
        return AV1PF_ERR_INVALID_ARG;
    for (int pl = 0; pl < (coded->mono ? 1 : 3); pl++) {
        int srcw, codedw, dstw, rows;
        sr_geometry(coded, up, pl, &srcw, &codedw, &dstw, &rows);
        for (int y = 0; y < rows; y++) {
            int rc = orc_superres_row(coded->p[pl] + (size_t)y * coded->stride[pl], srcw, codedw,
                                      up->p[pl] + (size_t)y * up->stride[pl], dstw, coded->bd);
            if (rc) return rc;
        }
    }
    return AV1PF_OK;
}

/* Deblocked LR context rows with super-resolution
 * (SaveRestorationBoundaryLinesFromFramePlaneSuperRes,
 * restoration_boundary.go:280-345): the stripe seams of the upscaled plane
 * take rows of the coded (deblocked) plane, each upscaled to the LR plane
 * width; a single row left is duplicated, then both lines are extended by 4
 * samples on each side. */
int orc_lr_save_superres(const orc_frame *coded, const orc_frame *up, const av1pf_lr_params *p, orc_lr_bounds b[3])
{
    for (int pl = 0; pl < (up->mono ? 1 : 3); pl++) {
        if (p->frame_type[pl] == AV1PF_LR_NONE) continue;
        int pw, ph, hz, vt;
        if (orc_lr_plane_geometry(up, p, pl, &pw, &ph, &hz, &vt)) return AV1PF_ERR_INVALID_ARG;
        int srcw, codedw, dstw, rows;
        sr_geometry(coded, up, pl, &srcw, &codedw, &dstw, &rows);
        if (dstw != pw || rows != ph) return AV1PF_ERR_INVALID_ARG;
        const int sy = pl ? up->ssy : 0;
        const int sh = 64 >> sy, off = 8 >> sy;
        for (int t = 0;; t++) {
            const int y0 = t * sh - off > 0 ? t * sh - off : 0;
            if (y0 >= ph) break;
            const int y1 = (t + 1) * sh - off < ph ? (t + 1) * sh - off : ph;
            for (int side = 0; side < 2; side++) {
                if (side == 0 ? t == 0 : y1 >= ph) continue;
                const int row = side == 0 ? y0 - 2 : y1;
                const int n = ph - row < 2 ? ph - row : 2;
                uint16_t *buf = side == 0 ? b[pl].above : b[pl].below;
                for (int i = 0; i < n; i++) {
                    uint16_t *line = buf + (size_t)(t * 2 + i) * b[pl].stride;
                    orc_superres_row(coded->p[pl] + (size_t)(row + i) * coded->stride[pl], srcw, codedw, line + 4,
                                     pw, coded->bd);
                }
                if (n == 1)
                    memcpy(buf + (size_t)(t * 2 + 1) * b[pl].stride, buf + (size_t)(t * 2) * b[pl].stride,
                           (size_t)b[pl].stride * 2);
                for (int i = 0; i < 2; i++) {
                    uint16_t *l = buf + (size_t)(t * 2 + i) * b[pl].stride;
                    for (int j = 0; j < 4; j++) { l[j] = l[4]; l[4 + pw + j] = l[4 + pw - 1]; }
                }
            }
        }
    }
    return AV1PF_OK;
}
