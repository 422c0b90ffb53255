/*
 * CPU oracle: AV1 deblocking as goav1 runs it from decode-time edge masks.
 * Test infrastructure only (see orc.h).
 *
 * Restates:
 *   loopfilter/threshold.go:18-47       ThresholdsForLevel
 *   loopfilter/sb.go:48-72              NewSBLUT (E = blimit, I = limit, H = level >> 4)
 *   loopfilter/sb.go:99-110             sbUnitWidth (class priority 16 > 8/6 > 4)
 *   loopfilter/sb.go:115-151,316-346    per-unit level resolve + fallback to left/above cell
 *   loopfilter/sb.go:348-446            lpfUnit16 (dav1d loop_filter(), covers 8-bit with bd=8)
 *   decoder/postfilter_loopfilter_masks.go:229-260   pass order: per 128-px region row,
 *                                       vertical edges of Y,U,V then horizontal edges; chroma gate
 *   decoder/postfilter_loopfilter_masks_sb.go:25-190 mask assembly, MI(4)-aligned crop, frame
 *                                       left column / top row skipped
 *   decoder/postfilter_loopfilter.go:2569-2604       plane position extent (MI aligned)
 * The libaom per-edge formulation (loopfilter/filter{4,6,8,14}.go, aom_dsp/loopfilter.c)
 * is restated separately in orc_lf_edge_libaom as an independent cross-check.
 */
#include "orc.h"
#include <stdlib.h>

static int iabs(int v) { return v < 0 ? -v : v; }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int iclip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* loopfilter/threshold.go:18-47 */
void orc_lf_thresholds(int level, int sharpness, int *limit, int *blimit, int *hev)
{
    int shift = 0;
    if (sharpness > 0) shift++;
    if (sharpness > 4) shift++;
    int l = level >> shift;
    if (sharpness > 0 && l > 9 - sharpness) l = 9 - sharpness;
    if (l < 1) l = 1;
    *limit = l;
    *blimit = 2 * (level + 2) + l;
    *hev = level >> 4;
}

/* loopfilter/sb.go:348-446 (lpfUnit16); with bd = 8 it is lpfUnit8 (sb.go:156-246). */
void orc_lpf_unit(uint16_t *pix, ptrdiff_t off, ptrdiff_t stridea, ptrdiff_t strideb,
                  int bd, int E, int I, int H, int wd)
{
    const int s = bd - 8;
    const int F = 1 << s;
    E <<= s; I <<= s; H <<= s;
    const unsigned twoI = 2u * (unsigned)I, twoF = 2u * (unsigned)F, twoH = 2u * (unsigned)H;
    const int maxPix = (1 << bd) - 1;
    const int dmin = -(128 << s), dmax = (128 << s) - 1;
    for (int i = 0; i < 4; i++, off += stridea) {
        int p1 = pix[off - 2 * strideb], p0 = pix[off - strideb];
        int q0 = pix[off], q1 = pix[off + strideb];
        if ((unsigned)(p1 - p0 + I) > twoI || (unsigned)(q1 - q0 + I) > twoI ||
            iabs(p0 - q0) * 2 + (iabs(p1 - q1) >> 1) > E)
            continue;
        int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
        if (wd > 4) {
            p2 = pix[off - 3 * strideb]; q2 = pix[off + 2 * strideb];
            if ((unsigned)(p2 - p1 + I) > twoI || (unsigned)(q2 - q1 + I) > twoI) continue;
            if (wd > 6) {
                p3 = pix[off - 4 * strideb]; q3 = pix[off + 3 * strideb];
                if ((unsigned)(p3 - p2 + I) > twoI || (unsigned)(q3 - q2 + I) > twoI) continue;
            }
        }
        int flat8in = 0;
        if (wd >= 6) {
            flat8in = (unsigned)(p2 - p0 + F) <= twoF && (unsigned)(p1 - p0 + F) <= twoF &&
                      (unsigned)(q1 - q0 + F) <= twoF && (unsigned)(q2 - q0 + F) <= twoF;
            if (wd >= 8 && flat8in)
                flat8in = (unsigned)(p3 - p0 + F) <= twoF && (unsigned)(q3 - q0 + F) <= twoF;
        }
        if (flat8in) {
            if (wd >= 16) {
                int p6 = pix[off - 7 * strideb], p5 = pix[off - 6 * strideb], p4 = pix[off - 5 * strideb];
                int q4 = pix[off + 4 * strideb], q5 = pix[off + 5 * strideb], q6 = pix[off + 6 * strideb];
                if ((unsigned)(p6 - p0 + F) <= twoF && (unsigned)(p5 - p0 + F) <= twoF &&
                    (unsigned)(p4 - p0 + F) <= twoF && (unsigned)(q4 - q0 + F) <= twoF &&
                    (unsigned)(q5 - q0 + F) <= twoF && (unsigned)(q6 - q0 + F) <= twoF) {
                    pix[off - 6 * strideb] = (uint16_t)((p6 * 7 + p5 * 2 + p4 * 2 + p3 + p2 + p1 + p0 + q0 + 8) >> 4);
                    pix[off - 5 * strideb] = (uint16_t)((p6 * 5 + p5 * 2 + p4 * 2 + p3 * 2 + p2 + p1 + p0 + q0 + q1 + 8) >> 4);
                    pix[off - 4 * strideb] = (uint16_t)((p6 * 4 + p5 + p4 * 2 + p3 * 2 + p2 * 2 + p1 + p0 + q0 + q1 + q2 + 8) >> 4);
                    pix[off - 3 * strideb] = (uint16_t)((p6 * 3 + p5 + p4 + p3 * 2 + p2 * 2 + p1 * 2 + p0 + q0 + q1 + q2 + q3 + 8) >> 4);
                    pix[off - 2 * strideb] = (uint16_t)((p6 * 2 + p5 + p4 + p3 + p2 * 2 + p1 * 2 + p0 * 2 + q0 + q1 + q2 + q3 + q4 + 8) >> 4);
                    pix[off - strideb] = (uint16_t)((p6 + p5 + p4 + p3 + p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + q2 + q3 + q4 + q5 + 8) >> 4);
                    pix[off] = (uint16_t)((p5 + p4 + p3 + p2 + p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + q3 + q4 + q5 + q6 + 8) >> 4);
                    pix[off + strideb] = (uint16_t)((p4 + p3 + p2 + p1 + p0 + q0 * 2 + q1 * 2 + q2 * 2 + q3 + q4 + q5 + q6 * 2 + 8) >> 4);
                    pix[off + 2 * strideb] = (uint16_t)((p3 + p2 + p1 + p0 + q0 + q1 * 2 + q2 * 2 + q3 * 2 + q4 + q5 + q6 * 3 + 8) >> 4);
                    pix[off + 3 * strideb] = (uint16_t)((p2 + p1 + p0 + q0 + q1 + q2 * 2 + q3 * 2 + q4 * 2 + q5 + q6 * 4 + 8) >> 4);
                    pix[off + 4 * strideb] = (uint16_t)((p1 + p0 + q0 + q1 + q2 + q3 * 2 + q4 * 2 + q5 * 2 + q6 * 5 + 8) >> 4);
                    pix[off + 5 * strideb] = (uint16_t)((p0 + q0 + q1 + q2 + q3 + q4 * 2 + q5 * 2 + q6 * 7 + 8) >> 4);
                    continue;
                }
            }
            if (wd >= 8) {
                pix[off - 3 * strideb] = (uint16_t)((p3 * 3 + p2 * 2 + p1 + p0 + q0 + 4) >> 3);
                pix[off - 2 * strideb] = (uint16_t)((p3 * 2 + p2 + p1 * 2 + p0 + q0 + q1 + 4) >> 3);
                pix[off - strideb] = (uint16_t)((p3 + p2 + p1 + p0 * 2 + q0 + q1 + q2 + 4) >> 3);
                pix[off] = (uint16_t)((p2 + p1 + p0 + q0 * 2 + q1 + q2 + q3 + 4) >> 3);
                pix[off + strideb] = (uint16_t)((p1 + p0 + q0 + q1 * 2 + q2 + q3 * 2 + 4) >> 3);
                pix[off + 2 * strideb] = (uint16_t)((p0 + q0 + q1 + q2 * 2 + q3 * 3 + 4) >> 3);
                continue;
            }
            pix[off - 2 * strideb] = (uint16_t)((p2 * 3 + p1 * 2 + p0 * 2 + q0 + 4) >> 3);
            pix[off - strideb] = (uint16_t)((p2 + p1 * 2 + p0 * 2 + q0 * 2 + q1 + 4) >> 3);
            pix[off] = (uint16_t)((p1 + p0 * 2 + q0 * 2 + q1 * 2 + q2 + 4) >> 3);
            pix[off + strideb] = (uint16_t)((p0 + q0 * 2 + q1 * 2 + q2 * 3 + 4) >> 3);
            continue;
        }
        int hev = (unsigned)(p1 - p0 + H) > twoH || (unsigned)(q1 - q0 + H) > twoH;
        int f;
        if (hev) f = iclip(iclip(p1 - q1, dmin, dmax) + 3 * (q0 - p0), dmin, dmax);
        else f = iclip(3 * (q0 - p0), dmin, dmax);
        int f1 = imin(f + 4, dmax) >> 3;
        int f2 = imin(f + 3, dmax) >> 3;
        pix[off - strideb] = (uint16_t)iclip(p0 + f2, 0, maxPix);
        pix[off] = (uint16_t)iclip(q0 - f1, 0, maxPix);
        if (!hev) {
            f = (f1 + 1) >> 1;
            pix[off - 2 * strideb] = (uint16_t)iclip(p1 + f, 0, maxPix);
            pix[off + strideb] = (uint16_t)iclip(q1 - f, 0, maxPix);
        }
    }
}

/* ---- libaom per-edge formulation, one line across the edge --------------
 * loopfilter/filter4.go:193-229 (needsFilter4 / filter4Samples),
 * filter6.go, filter8.go, filter14.go (flat masks and wide smoothing), i.e.
 * aom_dsp/loopfilter.c highbd_{lpf,filter}{4,6,8,14}.  Written from the
 * signed-domain formulation so it checks the dav1d-style unit above.        */
static int sclamp(int v, int lo, int hi) { return iclip(v, lo, hi); }

static void lib_filter4(int *p1, int *p0, int *q0, int *q1, int hevth, int s)
{
    const int c = 128 << s, lo = -(128 << s), hi = (128 << s) - 1;
    int ps1 = *p1 - c, ps0 = *p0 - c, qs0 = *q0 - c, qs1 = *q1 - c;
    int hev = iabs(*p1 - *p0) > hevth || iabs(*q1 - *q0) > hevth;
    int filter = hev ? sclamp(ps1 - qs1, lo, hi) : 0;
    filter = sclamp(filter + 3 * (qs0 - ps0), lo, hi);
    int f1 = sclamp(filter + 4, lo, hi) >> 3;
    int f2 = sclamp(filter + 3, lo, hi) >> 3;
    *q0 = sclamp(qs0 - f1, lo, hi) + c;
    *p0 = sclamp(ps0 + f2, lo, hi) + c;
    if (!hev) {
        int o = (f1 + 1) >> 1;
        *q1 = sclamp(qs1 - o, lo, hi) + c;
        *p1 = sclamp(ps1 + o, lo, hi) + c;
    }
}

#define R2(v, n) (((v) + (1 << ((n) - 1))) >> (n))

void orc_lf_edge_libaom(uint16_t *pix, ptrdiff_t off, ptrdiff_t st, int bd,
                        int width, int limit, int blimit, int hevt)
{
    const int s = bd - 8;
    const int lim = limit << s, blim = blimit << s, hev = hevt << s, flat = 1 << s;
    int p[7], q[7];
    int n = width == 4 ? 2 : width == 6 ? 3 : width == 8 ? 4 : 7;
    for (int i = 0; i < n; i++) { p[i] = pix[off - (i + 1) * st]; q[i] = pix[off + i * st]; }
    /* filter mask: every neighbouring step within limit, edge step within blimit */
    int mask = 1;
    int taps = width == 4 ? 2 : width == 6 ? 3 : 4;
    for (int i = 0; i + 1 < taps; i++)
        if (iabs(p[i + 1] - p[i]) > lim || iabs(q[i + 1] - q[i]) > lim) mask = 0;
    if (iabs(p[0] - q[0]) * 2 + iabs(p[1] - q[1]) / 2 > blim) mask = 0;
    if (!mask) return;
    int fl = 0;
    if (width >= 6) {
        fl = 1;
        for (int i = 1; i < taps; i++)
            if (iabs(p[i] - p[0]) > flat || iabs(q[i] - q[0]) > flat) fl = 0;
    }
    int o[14];
    int P0 = p[0], P1 = p[1], P2 = n > 2 ? p[2] : 0, P3 = n > 3 ? p[3] : 0;
    int Q0 = q[0], Q1 = q[1], Q2 = n > 2 ? q[2] : 0, Q3 = n > 3 ? q[3] : 0;
    if (width == 14 && fl) {
        int fl2 = 1;
        for (int i = 4; i < 7; i++)
            if (iabs(p[i] - p[0]) > flat || iabs(q[i] - q[0]) > flat) fl2 = 0;
        if (fl2) {
            int P4 = p[4], P5 = p[5], P6 = p[6], Q4 = q[4], Q5 = q[5], Q6 = q[6];
            o[0] = R2(P6 * 7 + P5 * 2 + P4 * 2 + P3 + P2 + P1 + P0 + Q0, 4);
            o[1] = R2(P6 * 5 + P5 * 2 + P4 * 2 + P3 * 2 + P2 + P1 + P0 + Q0 + Q1, 4);
            o[2] = R2(P6 * 4 + P5 + P4 * 2 + P3 * 2 + P2 * 2 + P1 + P0 + Q0 + Q1 + Q2, 4);
            o[3] = R2(P6 * 3 + P5 + P4 + P3 * 2 + P2 * 2 + P1 * 2 + P0 + Q0 + Q1 + Q2 + Q3, 4);
            o[4] = R2(P6 * 2 + P5 + P4 + P3 + P2 * 2 + P1 * 2 + P0 * 2 + Q0 + Q1 + Q2 + Q3 + Q4, 4);
            o[5] = R2(P6 + P5 + P4 + P3 + P2 + P1 * 2 + P0 * 2 + Q0 * 2 + Q1 + Q2 + Q3 + Q4 + Q5, 4);
            o[6] = R2(P5 + P4 + P3 + P2 + P1 + P0 * 2 + Q0 * 2 + Q1 * 2 + Q2 + Q3 + Q4 + Q5 + Q6, 4);
            o[7] = R2(P4 + P3 + P2 + P1 + P0 + Q0 * 2 + Q1 * 2 + Q2 * 2 + Q3 + Q4 + Q5 + Q6 * 2, 4);
            o[8] = R2(P3 + P2 + P1 + P0 + Q0 + Q1 * 2 + Q2 * 2 + Q3 * 2 + Q4 + Q5 + Q6 * 3, 4);
            o[9] = R2(P2 + P1 + P0 + Q0 + Q1 + Q2 * 2 + Q3 * 2 + Q4 * 2 + Q5 + Q6 * 4, 4);
            o[10] = R2(P1 + P0 + Q0 + Q1 + Q2 + Q3 * 2 + Q4 * 2 + Q5 * 2 + Q6 * 5, 4);
            o[11] = R2(P0 + Q0 + Q1 + Q2 + Q3 + Q4 * 2 + Q5 * 2 + Q6 * 7, 4);
            for (int i = 0; i < 6; i++) pix[off - (6 - i) * st] = (uint16_t)o[i];
            for (int i = 0; i < 6; i++) pix[off + i * st] = (uint16_t)o[6 + i];
            return;
        }
    }
    if ((width == 8 || width == 14) && fl) {
        pix[off - 3 * st] = (uint16_t)R2(P3 + P3 + P3 + 2 * P2 + P1 + P0 + Q0, 3);
        pix[off - 2 * st] = (uint16_t)R2(P3 + P3 + P2 + 2 * P1 + P0 + Q0 + Q1, 3);
        pix[off - st] = (uint16_t)R2(P3 + P2 + P1 + 2 * P0 + Q0 + Q1 + Q2, 3);
        pix[off] = (uint16_t)R2(P2 + P1 + P0 + 2 * Q0 + Q1 + Q2 + Q3, 3);
        pix[off + st] = (uint16_t)R2(P1 + P0 + Q0 + 2 * Q1 + Q2 + Q3 + Q3, 3);
        pix[off + 2 * st] = (uint16_t)R2(P0 + Q0 + Q1 + 2 * Q2 + Q3 + Q3 + Q3, 3);
        return;
    }
    if (width == 6 && fl) {
        pix[off - 2 * st] = (uint16_t)R2(P2 * 3 + P1 * 2 + P0 * 2 + Q0, 3);
        pix[off - st] = (uint16_t)R2(P2 + P1 * 2 + P0 * 2 + Q0 * 2 + Q1, 3);
        pix[off] = (uint16_t)R2(P1 + P0 * 2 + Q0 * 2 + Q1 * 2 + Q2, 3);
        pix[off + st] = (uint16_t)R2(P0 + Q0 * 2 + Q1 * 2 + Q2 * 3, 3);
        return;
    }
    int a = P1, b = P0, c = Q0, d = Q1;
    lib_filter4(&a, &b, &c, &d, hev, s);
    pix[off - 2 * st] = (uint16_t)a; pix[off - st] = (uint16_t)b;
    pix[off] = (uint16_t)c; pix[off + st] = (uint16_t)d;
}

/* ---- frame driver ------------------------------------------------------ */
typedef struct lf_plane {
    uint16_t *pix;
    ptrdiff_t stride;
    int w4, h4;           /* MI(4)-aligned position extent in 4x4 units */
} lf_plane;

static uint32_t lane_mask(int n) { return n >= 32 ? 0xffffffffu : ((1u << n) - 1u); }

static int unit_width(const uint32_t vm[3], int luma, uint32_t bit)
{
    if (vm[2] & bit) return 16;
    if (vm[1] & bit) return luma ? 8 : 6;
    return 4;
}

static void lf_apply_set(const lf_plane *pl, const uint32_t vm[3], int luma, int vertical,
                         int x4, int y4, const uint8_t *lc, int cols, int comp,
                         const int *E, const int *I, int bd)
{
    uint32_t all = vm[0] | vm[1] | vm[2];
    while (all) {
        int n = __builtin_ctz(all);
        uint32_t bit = 1u << n;
        all &= ~bit;
        int ux = vertical ? x4 : x4 + n, uy = vertical ? y4 + n : y4;
        int li = uy * cols + ux;
        int L = lc[4 * li + comp];
        if (L == 0) L = lc[4 * (vertical ? li - 1 : li - cols) + comp];
        if (L == 0) continue;
        ptrdiff_t off = (ptrdiff_t)(uy * 4) * pl->stride + ux * 4;
        if (vertical) orc_lpf_unit(pl->pix, off, pl->stride, 1, bd, E[L], I[L], L >> 4, unit_width(vm, luma, bit));
        else orc_lpf_unit(pl->pix, off, 1, pl->stride, bd, E[L], I[L], L >> 4, unit_width(vm, luma, bit));
    }
}

int orc_deblock(orc_frame *f, const av1pf_deblock_params *p)
{
    if (!f || !p || !p->masks || !p->level_cache) return AV1PF_ERR_INVALID_ARG;
    int mi_cols = ((f->w + 7) >> 3) << 1, mi_rows = ((f->h + 7) >> 3) << 1;
    if (p->mi_cols != mi_cols || p->mi_rows != mi_rows ||
        p->sb128_w != (mi_cols + 31) >> 5 || p->sb128_h != (mi_rows + 31) >> 5)
        return AV1PF_ERR_INVALID_ARG;
    int E[64], I[64], hv;
    for (int l = 0; l < 64; l++) orc_lf_thresholds(l, p->sharpness, &I[l], &E[l], &hv);
    const int luma_zero = p->level_y[0] == 0 && p->level_y[1] == 0;
    const int base[3] = { 1, imin(p->level_u, 63), imin(p->level_v, 63) };
    const int max_plane = f->mono ? 0 : 2;
    lf_plane pl[3];
    for (int i = 0; i <= max_plane; i++) {
        int sx = i ? f->ssx : 0, sy = i ? f->ssy : 0;
        int pw = (f->w + sx) >> sx, ph = (f->h + sy) >> sy;
        pl[i].pix = f->p[i];
        pl[i].stride = f->stride[i];
        pl[i].w4 = ((pw + 3) & ~3) >> 2;
        pl[i].h4 = ((ph + 3) & ~3) >> 2;
    }
    const uint8_t *lc = p->level_cache;
    for (int rRow = 0; rRow < p->sb128_h; rRow++) {
        for (int dir = 0; dir < 2; dir++) {
            for (int plane = 0; plane <= max_plane; plane++) {
                if (plane && (luma_zero || base[plane] == 0)) continue;
                const lf_plane *P = &pl[plane];
                const int luma = plane == 0;
                const int sx = luma ? 0 : f->ssx, sy = luma ? 0 : f->ssy;
                const int rcw = 32 >> sx, rch = 32 >> sy;
                const int comp = luma ? dir : plane + 1;
                const int laneV = luma ? 16 : 16 >> sy, laneH = luma ? 16 : 16 >> sx;
                for (int rCol = 0; rCol < p->sb128_w; rCol++) {
                    const av1pf_lf_mask *m = &p->masks[rRow * p->sb128_w + rCol];
                    int bRow = rRow * rch, bCol = rCol * rcw;
                    int nRows = imin(rch, P->h4 - bRow), nCols = imin(rcw, P->w4 - bCol);
                    if (nRows <= 0 || nCols <= 0) continue;
                    if (dir == 0) {
                        uint32_t rm = lane_mask(nRows);
                        for (int pos = 0; pos < nCols; pos++) {
                            int col = bCol + pos;
                            if (col == 0) continue;
                            uint32_t vm[3] = { 0, 0, 0 };
                            for (int k = 0; k < (luma ? 3 : 2); k++) {
                                uint32_t lo = luma ? m->y[0][pos][k][0] : m->uv[0][pos][k][0];
                                uint32_t hi = luma ? m->y[0][pos][k][1] : m->uv[0][pos][k][1];
                                vm[k] = (lo | (hi << laneV)) & rm;
                            }
                            if (!(vm[0] | vm[1] | vm[2])) continue;
                            lf_apply_set(P, vm, luma, 1, col, bRow, lc, mi_cols, comp, E, I, f->bd);
                        }
                    } else {
                        uint32_t cm = lane_mask(nCols);
                        for (int pos = 0; pos < nRows; pos++) {
                            int row = bRow + pos;
                            if (row == 0) continue;
                            uint32_t vm[3] = { 0, 0, 0 };
                            for (int k = 0; k < (luma ? 3 : 2); k++) {
                                uint32_t lo = luma ? m->y[1][pos][k][0] : m->uv[1][pos][k][0];
                                uint32_t hi = luma ? m->y[1][pos][k][1] : m->uv[1][pos][k][1];
                                vm[k] = (lo | (hi << laneH)) & cm;
                            }
                            if (!(vm[0] | vm[1] | vm[2])) continue;
                            lf_apply_set(P, vm, luma, 0, bCol, row, lc, mi_cols, comp, E, I, f->bd);
                        }
                    }
                }
            }
        }
    }
    (void)imax;
    return AV1PF_OK;
}
