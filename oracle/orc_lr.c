/*
 * CPU oracle: AV1 loop restoration as goav1 runs it (libaom striped flow).
 * Test infrastructure only (see orc.h).
 *
 * Restates:
 *   restoration/wiener.go:44-47,59-80,132-216   Wiener taps, H pass (offset, round0, clamp to
 *                                               2^(bd+1+7-round0)-1), V pass (round1), rounds
 *   restoration/selfguided.go:38-83             SGR parameter table and LUTs (normative tables,
 *                                               regenerated below from their closed forms)
 *   restoration/selfguided.go:111-126           DecodeSGRXQ
 *   restoration/selfguided.go:132-231           ApplySelfguidedRestoration + sgrWeightedRow (int16 wrap)
 *   restoration/selfguided.go:249-310           selfguidedFast (r=2, 6/5 weights) and selfguided (r=1, 4/3)
 *   restoration/selfguided_rolling.go:19-85     A/B coefficients over rows/cols [-1, size] (uint32 math)
 *   tile/restoration_schedule.go:61-101,400-410 plane grid (rounded plane size, unit counts)
 *   tile/restoration_geometry.go:54-143         UnitRect (8-row offset) and ProcessingStripe
 *   tile/restoration_geometry.go:163-236        ProcessingUnit widths
 *   tile/restoration_boundary.go:11-15,52-94    boundary constants and buffer layout
 *   tile/restoration_boundary.go:126-183        ExtendRestorationFrame (3-sample replication)
 *   tile/restoration_boundary.go:192-255,454-516,879-923  boundary save (deblocked interior
 *                                               seams, post-CDEF frame top/bottom), duplication, 4-px extension
 *   tile/restoration_boundary.go:553-614,651-682 stripe setup / restore
 *   tile/restoration_apply.go:526-570           per record: stripes -> processing units
 *   tile/restoration_frame_bridge.go:101-190    bordered copy, filtered records stored back
 */
#include "orc.h"
#include <stdlib.h>
#include <string.h>

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int32_t clamp32(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int32_t r2s(int32_t v, int n) { return n <= 0 ? v : (int32_t)((v + (1 << (n - 1))) >> n); }
static uint32_t r2u(uint32_t v, int n) { return n <= 0 ? v : (v + (1u << (n - 1))) >> n; }

/* ---- Wiener ------------------------------------------------------------- */
static void wiener_rounds(int bd, int *r0, int *r1)
{
    int round0 = 3, round1 = 2 * 7 - 3;
    int ibr = bd + 7 - round0 + 2;
    if (ibr > 16) { round0 += ibr - 16; round1 -= ibr - 16; }
    *r0 = round0; *r1 = round1;
}

/* src points at the unit's top-left sample; rows -3..h+2 and cols -3..w+2 are read. */
void orc_wiener_unit(const uint16_t *src, int ss, uint16_t *dst, int ds,
                     int w, int h, const int8_t vt[3], const int8_t ht[3], int bd)
{
    int r0, r1;
    wiener_rounds(bd, &r0, &r1);
    int32_t hf[7], vf[7];
    hf[0] = hf[6] = ht[0]; hf[1] = hf[5] = ht[1]; hf[2] = hf[4] = ht[2]; hf[3] = -2 * (ht[0] + ht[1] + ht[2]);
    vf[0] = vf[6] = vt[0]; vf[1] = vf[5] = vt[1]; vf[2] = vf[4] = vt[2]; vf[3] = -2 * (vt[0] + vt[1] + vt[2]);
    const int32_t limit = 1 << (bd + 1 + 7 - r0), hoff = 1 << (bd + 7 - 1);
    uint16_t *tmp = malloc((size_t)w * (h + 6) * sizeof(uint16_t));
    for (int r = -3; r < h + 3; r++) {
        for (int c = 0; c < w; c++) {
            const uint16_t *s = src + (ptrdiff_t)r * ss + c - 3;
            int32_t sum = ((int32_t)s[3] << 7) + hoff;
            for (int k = 0; k < 7; k++) sum += (int32_t)s[k] * hf[k];
            tmp[(r + 3) * w + c] = (uint16_t)clamp32(r2s(sum, r0), 0, limit - 1);
        }
    }
    const int32_t voff = 1 << (bd + r1 - 1), mx = (1 << bd) - 1;
    for (int r = 0; r < h; r++) {
        for (int c = 0; c < w; c++) {
            int32_t sum = ((int32_t)tmp[(r + 3) * w + c] << 7) - voff;
            for (int k = 0; k < 7; k++) sum += (int32_t)tmp[(r + k) * w + c] * vf[k];
            dst[(ptrdiff_t)r * ds + c] = (uint16_t)clamp32(r2s(sum, r1), 0, mx);
        }
    }
    free(tmp);
}

/* ---- self-guided -------------------------------------------------------- */
static const int sgr_radius[16][2] = {
    {2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{2,1},{0,1},{0,1},{0,1},{0,1},{2,0},{2,0}};
static const int sgr_s[16][2] = {
    {140,3236},{112,2158},{93,1618},{80,1438},{70,1295},{58,1177},{47,1079},{37,996},
    {30,925},{25,863},{-1,2589},{-1,1618},{-1,1177},{-1,925},{56,-1},{22,-1}};

/* x/(x+1) in Q8 and 1/n in Q12, AV1's normative tables, from their closed forms:
 * xByXPlus1[z] = round(256 z / (z + 1)) except [0] = 1, [255] = 256;
 * oneByX[n-1] = round(4096 / n). */
static int x_by_xplus1(uint32_t z)
{
    if (z > 255) z = 255;
    if (z == 0) return 1;
    if (z == 255) return 256;
    return (int)((512u * z + (z + 1)) / (2u * (z + 1)));
}
static uint32_t one_by_x(int n) { return (uint32_t)((8192 + n) / (2 * n)); }

void orc_sgr_decode_xq(int set, const int8_t xqd[2], int xq[2])
{
    if (sgr_radius[set][0] == 0) { xq[0] = 0; xq[1] = 128 - xqd[1]; }
    else if (sgr_radius[set][1] == 0) { xq[0] = xqd[0]; xq[1] = 0; }
    else { xq[0] = xqd[0]; xq[1] = 128 - xq[0] - xqd[1]; }
}

/* A/B at rows y (stepping) and cols [-1, w]; ab arrays are (h+2) x (w+2) with origin (+1,+1). */
static void sgr_coeffs(const uint16_t *src, int ss, int w, int h, int bd, int r, int s, int step,
                       int32_t *A, int32_t *B)
{
    const int n = (2 * r + 1) * (2 * r + 1);
    const uint32_t recip = one_by_x(n);
    for (int y = -1; y < h + 1; y += step) {
        for (int x = -1; x < w + 1; x++) {
            int32_t sum = 0, sq = 0;
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++) {
                    int32_t v = src[(ptrdiff_t)(y + dy) * ss + x + dx];
                    sum += v; sq += v * v;
                }
            uint32_t a = (uint32_t)r2s(sq, 2 * (bd - 8));
            uint32_t b = (uint32_t)r2s(sum, bd - 8);
            uint32_t p = a * (uint32_t)n >= b * b ? a * (uint32_t)n - b * b : 0;
            uint32_t z = r2u(p * (uint32_t)s, 20);
            if (z > 255) z = 255;
            int32_t aa = x_by_xplus1(z);
            A[(y + 1) * (w + 2) + x + 1] = aa;
            B[(y + 1) * (w + 2) + x + 1] = (int32_t)r2u((uint32_t)(256 - aa) * (uint32_t)sum * recip, 12);
        }
    }
}

/* The two self-guided filter outputs f0 (r = 2 pass) and f1 (r = 1 pass) of
 * set `set`, [h][w] in Q4, zero for a pass of radius 0. */
void orc_sgr_filters(const uint16_t *src, int ss, int w, int h, int set, int bd, int32_t *f0, int32_t *f1)
{
    int32_t *A = calloc((size_t)(w + 2) * (h + 2), sizeof(int32_t));
    int32_t *B = calloc((size_t)(w + 2) * (h + 2), sizeof(int32_t));
    memset(f0, 0, sizeof(int32_t) * (size_t)w * h);
    memset(f1, 0, sizeof(int32_t) * (size_t)w * h);
    const int W2 = w + 2;
#define AA(y, x) A[((y) + 1) * W2 + (x) + 1]
#define BB(y, x) B[((y) + 1) * W2 + (x) + 1]
    if (sgr_radius[set][0]) {
        sgr_coeffs(src, ss, w, h, bd, 2, sgr_s[set][0], 2, A, B);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t u = src[(ptrdiff_t)y * ss + x], a, b;
                if ((y & 1) == 0) {
                    a = (AA(y - 1, x) + AA(y + 1, x)) * 6 + (AA(y - 1, x - 1) + AA(y - 1, x + 1) + AA(y + 1, x - 1) + AA(y + 1, x + 1)) * 5;
                    b = (BB(y - 1, x) + BB(y + 1, x)) * 6 + (BB(y - 1, x - 1) + BB(y - 1, x + 1) + BB(y + 1, x - 1) + BB(y + 1, x + 1)) * 5;
                    f0[y * w + x] = r2s(a * u + b, 8 + 5 - 4);
                } else {
                    a = AA(y, x) * 6 + (AA(y, x - 1) + AA(y, x + 1)) * 5;
                    b = BB(y, x) * 6 + (BB(y, x - 1) + BB(y, x + 1)) * 5;
                    f0[y * w + x] = r2s(a * u + b, 8 + 4 - 4);
                }
            }
    }
    if (sgr_radius[set][1]) {
        sgr_coeffs(src, ss, w, h, bd, 1, sgr_s[set][1], 1, A, B);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int32_t u = src[(ptrdiff_t)y * ss + x];
                int32_t a = (AA(y, x) + AA(y, x - 1) + AA(y, x + 1) + AA(y - 1, x) + AA(y + 1, x)) * 4 +
                            (AA(y - 1, x - 1) + AA(y + 1, x - 1) + AA(y - 1, x + 1) + AA(y + 1, x + 1)) * 3;
                int32_t b = (BB(y, x) + BB(y, x - 1) + BB(y, x + 1) + BB(y - 1, x) + BB(y + 1, x)) * 4 +
                            (BB(y - 1, x - 1) + BB(y + 1, x - 1) + BB(y - 1, x + 1) + BB(y + 1, x + 1)) * 3;
                f1[y * w + x] = r2s(a * u + b, 8 + 5 - 4);
            }
    }
#undef AA
#undef BB
    free(A); free(B);
}

void orc_sgr_unit(const uint16_t *src, int ss, uint16_t *dst, int ds,
                  int w, int h, int set, const int8_t xqd[2], int bd)
{
    int32_t *f0 = calloc((size_t)w * h, sizeof(int32_t));
    int32_t *f1 = calloc((size_t)w * h, sizeof(int32_t));
    orc_sgr_filters(src, ss, w, h, set, bd, f0, f1);
    int xq[2];
    orc_sgr_decode_xq(set, xqd, xq);
    const int32_t mx = (1 << bd) - 1;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t u = (int32_t)src[(ptrdiff_t)y * ss + x] << 4;
            int32_t v = u << 7;
            v += xq[0] * (f0[y * w + x] - u);
            v += xq[1] * (f1[y * w + x] - u);
            int32_t rr = r2s(v, 7 + 4);
            int32_t wv = (int16_t)rr;
            dst[(ptrdiff_t)y * ds + x] = (uint16_t)clamp32(wv, 0, mx);
        }
    free(f0); free(f1);
}

/* ---- geometry ----------------------------------------------------------- */
int orc_lr_plane_geometry(const orc_frame *f, const av1pf_lr_params *p, int plane,
                          int *pw, int *ph, int *horz, int *vert)
{
    int sx = plane ? f->ssx : 0, sy = plane ? f->ssy : 0;
    *pw = (f->w + sx) >> sx;   /* roundPowerOfTwo(UpscaledWidth, ss) */
    *ph = (f->h + sy) >> sy;
    int us = p->unit_size[plane];
    if (us <= 0) return AV1PF_ERR_INVALID_ARG;
    *horz = imax((*pw + (us >> 1)) / us, 1);
    *vert = imax((*ph + (us >> 1)) / us, 1);
    return AV1PF_OK;
}

static int bstride_for(int pw) { return ((pw + 8) + 31) & ~31; }

int orc_lr_bounds_alloc(const orc_frame *f, const av1pf_lr_params *p, orc_lr_bounds b[3])
{
    memset(b, 0, 3 * sizeof(*b));
    for (int pl = 0; pl < (f->mono ? 1 : 3); pl++) {
        if (p->frame_type[pl] == AV1PF_LR_NONE) continue;
        int pw, ph, hz, vt;
        if (orc_lr_plane_geometry(f, p, pl, &pw, &ph, &hz, &vt)) return AV1PF_ERR_INVALID_ARG;
        int sy = pl ? f->ssy : 0;
        int sh = 64 >> sy, off = 8 >> sy;
        int stripes = (ph + off + sh - 1) / sh;
        b[pl].stride = bstride_for(pw);
        b[pl].rows = stripes * 2;
        b[pl].above = calloc((size_t)b[pl].stride * b[pl].rows, sizeof(uint16_t));
        b[pl].below = calloc((size_t)b[pl].stride * b[pl].rows, sizeof(uint16_t));
    }
    return AV1PF_OK;
}

void orc_lr_bounds_free(orc_lr_bounds b[3])
{
    for (int i = 0; i < 3; i++) { free(b[i].above); free(b[i].below); b[i].above = b[i].below = NULL; }
}

static void save_line(const orc_frame *f, int pl, int pw, int src_row, uint16_t *buf, int stride, int line)
{
    uint16_t *d = buf + (size_t)line * stride;
    memcpy(d + 4, f->p[pl] + (size_t)src_row * f->stride[pl], (size_t)pw * sizeof(uint16_t));
}
static void extend_lines(int pw, uint16_t *buf, int stride, int t)
{
    for (int i = 0; i < 2; i++) {
        uint16_t *l = buf + (size_t)(t * 2 + i) * stride;
        for (int j = 0; j < 4; j++) { l[j] = l[4]; l[4 + pw + j] = l[4 + pw - 1]; }
    }
}

int orc_lr_save(const orc_frame *f, const av1pf_lr_params *p, orc_lr_bounds b[3], int after_cdef)
{
    for (int pl = 0; pl < (f->mono ? 1 : 3); pl++) {
        if (p->frame_type[pl] == AV1PF_LR_NONE) continue;
        int pw, ph, hz, vt;
        orc_lr_plane_geometry(f, p, pl, &pw, &ph, &hz, &vt);
        int sy = pl ? f->ssy : 0;
        int sh = 64 >> sy, off = 8 >> sy;
        for (int t = 0;; t++) {
            int y0 = imax(0, t * sh - off);
            if (y0 >= ph) break;
            int y1 = imin((t + 1) * sh - off, ph);
            int use_above = t > 0, use_below = y1 < ph;
            if (!after_cdef) {
                if (use_above) {
                    int row = y0 - 2, n = imin(2, ph - row);
                    for (int i = 0; i < n; i++) save_line(f, pl, pw, row + i, b[pl].above, b[pl].stride, t * 2 + i);
                    if (n == 1) memcpy(b[pl].above + (size_t)(t * 2 + 1) * b[pl].stride, b[pl].above + (size_t)(t * 2) * b[pl].stride, (size_t)b[pl].stride * 2);
                    extend_lines(pw, b[pl].above, b[pl].stride, t);
                }
                if (use_below) {
                    int row = y1, n = imin(2, ph - row);
                    for (int i = 0; i < n; i++) save_line(f, pl, pw, row + i, b[pl].below, b[pl].stride, t * 2 + i);
                    if (n == 1) memcpy(b[pl].below + (size_t)(t * 2 + 1) * b[pl].stride, b[pl].below + (size_t)(t * 2) * b[pl].stride, (size_t)b[pl].stride * 2);
                    extend_lines(pw, b[pl].below, b[pl].stride, t);
                }
            } else {
                if (!use_above) {
                    for (int i = 0; i < 2; i++) save_line(f, pl, pw, y0, b[pl].above, b[pl].stride, t * 2 + i);
                    extend_lines(pw, b[pl].above, b[pl].stride, t);
                }
                if (!use_below) {
                    for (int i = 0; i < 2; i++) save_line(f, pl, pw, y1 - 1, b[pl].below, b[pl].stride, t * 2 + i);
                    extend_lines(pw, b[pl].below, b[pl].stride, t);
                }
            }
        }
    }
    return AV1PF_OK;
}

/* ---- apply -------------------------------------------------------------- */
#define BORDER 32

int orc_lr_apply(orc_frame *f, const av1pf_lr_params *p, const orc_lr_bounds b[3])
{
    for (int pl = 0; pl < (f->mono ? 1 : 3); pl++) {
        if (p->frame_type[pl] == AV1PF_LR_NONE) continue;
        const av1pf_lr_unit *units = p->units[pl];
        if (!units) return AV1PF_ERR_INVALID_ARG;
        int pw, ph, hz, vt;
        orc_lr_plane_geometry(f, p, pl, &pw, &ph, &hz, &vt);
        int any = 0;
        for (int i = 0; i < hz * vt; i++) {
            if (units[i].type > AV1PF_LR_SGRPROJ) return AV1PF_ERR_INVALID_ARG;
            if (units[i].type != AV1PF_LR_NONE) any = 1;
        }
        if (!any) continue;
        const int sx = pl ? f->ssx : 0, sy = pl ? f->ssy : 0;
        const int us = p->unit_size[pl];
        const int sh = 64 >> sy, off = 8 >> sy, procw = 64 >> sx;
        const int ds = pw + 2 * BORDER;
        uint16_t *data = calloc((size_t)ds * (ph + 2 * BORDER), sizeof(uint16_t));
        uint16_t *dst = calloc((size_t)ds * (ph + 2 * BORDER), sizeof(uint16_t));
        uint16_t *O = data + (size_t)BORDER * ds + BORDER;
        uint16_t *D = dst + (size_t)BORDER * ds + BORDER;
        for (int y = 0; y < ph; y++) memcpy(O + (size_t)y * ds, f->p[pl] + (size_t)y * f->stride[pl], (size_t)pw * 2);
        /* ExtendRestorationFrame with border 3 (restoration_boundary.go:126-183) */
        for (int y = 0; y < ph; y++)
            for (int x = 1; x <= 3; x++) { O[(size_t)y * ds - x] = O[(size_t)y * ds]; O[(size_t)y * ds + pw - 1 + x] = O[(size_t)y * ds + pw - 1]; }
        for (int y = 1; y <= 3; y++) {
            memcpy(O - (ptrdiff_t)y * ds - 3, O - 3, (size_t)(pw + 6) * 2);
            memcpy(O + (ptrdiff_t)(ph - 1 + y) * ds - 3, O + (ptrdiff_t)(ph - 1) * ds - 3, (size_t)(pw + 6) * 2);
        }
        uint16_t save[6][64 * 8 + 8];
        for (int ur = 0; ur < vt; ur++) {
            for (int uc = 0; uc < hz; uc++) {
                const av1pf_lr_unit *u = &units[ur * hz + uc];
                if (u->type == AV1PF_LR_NONE) continue;
                /* UnitRect (restoration_geometry.go:54-93) */
                int x0 = uc * us, x1 = uc + 1 == hz ? pw : x0 + us;
                int y0 = ur * us, y1 = ur + 1 == vt ? ph : y0 + us;
                if (y0 > 0) y0 -= off;
                if (y1 < ph) y1 -= off;
                /* ProcessingStripe (restoration_geometry.go:97-143) */
                for (int y = y0; y < y1;) {
                    int tstripe = (y + off) / sh;
                    int nominal = sh - (y == 0 ? off : 0);
                    int hgt = imin(nominal, y1 - y);
                    int copy_above = y != 0, copy_below = y + nominal < ph;
                    int lw = (x1 - x0) + 8;
                    if (copy_above)
                        for (int i = -3; i < 0; i++) {
                            uint16_t *row = O + (ptrdiff_t)(y + i) * ds + x0 - 4;
                            memcpy(save[i + 3], row, (size_t)lw * 2);
                            memcpy(row, b[pl].above + (size_t)(tstripe * 2 + imax(i + 2, 0)) * b[pl].stride + x0, (size_t)lw * 2);
                        }
                    if (copy_below)
                        for (int i = 0; i < 3; i++) {
                            uint16_t *row = O + (ptrdiff_t)(y + hgt + i) * ds + x0 - 4;
                            memcpy(save[3 + i], row, (size_t)lw * 2);
                            memcpy(row, b[pl].below + (size_t)(tstripe * 2 + imin(i, 1)) * b[pl].stride + x0, (size_t)lw * 2);
                        }
                    for (int x = x0; x < x1; x += procw) {
                        int vw = imin(procw, x1 - x);
                        const uint16_t *s = O + (ptrdiff_t)y * ds + x;
                        uint16_t *d = D + (ptrdiff_t)y * ds + x;
                        if (u->type == AV1PF_LR_WIENER) orc_wiener_unit(s, ds, d, ds, vw, hgt, u->wiener_v, u->wiener_h, f->bd);
                        else orc_sgr_unit(s, ds, d, ds, vw, hgt, u->sgr_set, u->sgr_xqd, f->bd);
                    }
                    if (copy_above)
                        for (int i = -3; i < 0; i++) memcpy(O + (ptrdiff_t)(y + i) * ds + x0 - 4, save[i + 3], (size_t)lw * 2);
                    if (copy_below)
                        for (int i = 0; i < 3; i++) memcpy(O + (ptrdiff_t)(y + hgt + i) * ds + x0 - 4, save[3 + i], (size_t)lw * 2);
                    y += hgt;
                }
                for (int y = y0; y < y1; y++)
                    memcpy(f->p[pl] + (size_t)y * f->stride[pl] + x0, D + (ptrdiff_t)y * ds + x0, (size_t)(x1 - x0) * 2);
            }
        }
        free(data); free(dst);
    }
    return AV1PF_OK;
}

/* exported for the known-answer tests */
int orc_x_by_xplus1(uint32_t z) { return x_by_xplus1(z); }
uint32_t orc_one_by_x(int n) { return one_by_x(n); }
