/*
 * CPU oracle: the encoder's curated loop-restoration search (luma).
 * Test infrastructure only (see orc.h).
 * The reference searches 8-bit frames; 10-bit frames are an extension with no
 * reference counterpart: the estimates clip to the bit depth's range, the
 * exact scores use the decoder's filters at that depth, and the acceptance
 * margin w*h scales with the squared sample range (w*h << 2(bd - 8)).
 *
 * Restates:
 *   encoder/restoration.go:149-210      SearchWienerUnitU8: baseline SSE; 6 tap
 *                                       triples scored with an H-only 7-tap sum over
 *                                       frame-clamped columns, (sum + 64) >> 7 clipped
 *                                       to 0..255 (no identity tap); a candidate wins
 *                                       when sse + w*h < the running best
 *   encoder/restoration.go:216-262      SearchSGRProjUnitU8: 4 presets through the exact
 *                                       decoder SGR on the bordered recon, same threshold
 *   encoder/keyframe_lossy.go:2310-2386 driver: recon bordered by 3 edge-clamped samples
 *                                       around the frame; the Wiener winner is re-scored
 *                                       with the exact Wiener (V = H taps); SGR must beat
 *                                       that exact SSE; units wider or taller than 64
 *                                       (restoration/wiener.go:198, selfguided.go:413)
 *                                       stay NONE
 *   tile/restoration_geometry.go:54-93  unit rectangles (8-row offset, last unit absorbs
 *                                       the remainder)
 *
 * sse[unit][12]: 0 baseline, 1..6 Wiener approximations, 7 Wiener exact, 8..11 SGR
 * presets; entries the reference does not compute are UINT64_MAX.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const int8_t kWienerCand[6][3] = { { 0, -3, 12 }, { 0, -5, 18 }, { 0, -2, 8 },
                                          { 0, 2, 10 },  { 1, 4, 14 },  { 0, -1, 4 } };
static const int kSgrPreset[4] = { 0, 2, 10, 14 };
static const int8_t kSgrXqd[4][2] = { { 8, 24 }, { 16, 40 }, { 0, 48 }, { 24, 72 } };

/* The Wiener screening (restoration.go:170-207): candidates in order, each
 * taken when its H-only estimate plus w*h is below the running best, which
 * starts at the baseline; -1 when none is taken. */
static int wiener_pick(const uint64_t e[12], uint64_t wh)
{
    uint64_t best = e[0];
    int pick = -1;
    for (int c = 0; c < 6; c++)
        if (e[1 + c] != UINT64_MAX && e[1 + c] + wh < best) { best = e[1 + c]; pick = c; }
    return pick;
}

/* The unit decision from its scores (restoration.go:149-262 and the
 * keyframe driver keyframe_lossy.go:2338-2386):
 *   - a zero baseline is left NONE (both searches return not found);
 *   - the Wiener screening picks a triple; the unit becomes WIENER with the
 *     exact SSE e[7] as the best so far, whatever e[7] is (the exact score is
 *     not compared with the baseline);
 *   - units wider or taller than 64 samples get neither filter (the exact
 *     filters' scratch refuses them), so they stay NONE;
 *   - the SGR presets run in order from the baseline with the same
 *     "sse + w*h < running best" rule; the found preset replaces the Wiener
 *     choice only when its SSE is strictly below the best so far. */
static void lr_decide(const uint64_t e[12], int w, int h, int cs, av1pf_lr_unit *out)
{
    memset(out, 0, sizeof(*out));
    out->type = AV1PF_LR_NONE;
    const uint64_t base = e[0], wh = ((uint64_t)w * (uint64_t)h) << (2 * cs);
    if (base == 0 || w <= 0 || h <= 0 || w > 64 || h > 64) return;
    uint64_t best_sse = UINT64_MAX;
    const int wbest = wiener_pick(e, wh);
    if (wbest >= 0) {
        best_sse = e[7];
        out->type = AV1PF_LR_WIENER;
        memcpy(out->wiener_v, kWienerCand[wbest], 3);
        memcpy(out->wiener_h, kWienerCand[wbest], 3);
    }
    uint64_t sbest = base;
    int pbest = -1;
    for (int p = 0; p < 4; p++)
        if (e[8 + p] != UINT64_MAX && e[8 + p] + wh < sbest) { sbest = e[8 + p]; pbest = p; }
    if (pbest >= 0 && sbest < best_sse) {
        memset(out, 0, sizeof(*out));
        out->type = AV1PF_LR_SGRPROJ;
        out->sgr_set = (uint8_t)kSgrPreset[pbest];
        out->sgr_xqd[0] = kSgrXqd[pbest][0];
        out->sgr_xqd[1] = kSgrXqd[pbest][1];
    }
}

void orc_lr_decide(const uint64_t e[12], int w, int h, av1pf_lr_unit *out) { lr_decide(e, w, h, 0, out); }

int orc_lr_search_units(int pw, int ph, int us, int *hz, int *vt)
{
    *hz = (pw + (us >> 1)) / us;
    *vt = (ph + (us >> 1)) / us;
    if (*hz < 1) *hz = 1;
    if (*vt < 1) *vt = 1;
    return *hz * *vt;
}

int orc_lr_search(const orc_frame *src, const orc_frame *rec, int us, av1pf_lr_unit *out, uint64_t *sse)
{
    if (!src || !rec || (rec->bd != 8 && rec->bd != 10) || src->bd != rec->bd || src->w != rec->w ||
        src->h != rec->h || us < 32 || us > 256)
        return AV1PF_ERR_INVALID_ARG;
    const int W = rec->w, H = rec->h, pad = 3, bd = rec->bd, cs = bd - 8, maxv = (1 << bd) - 1;
    const int bs = W + 2 * pad;
    uint16_t *bord = malloc(sizeof(uint16_t) * (size_t)bs * (H + 2 * pad));
    uint16_t *tmp = malloc(sizeof(uint16_t) * 64 * 64);
    for (int y = 0; y < H + 2 * pad; y++) {
        int sy = y - pad;
        sy = sy < 0 ? 0 : (sy >= H ? H - 1 : sy);
        const uint16_t *r = rec->p[0] + (size_t)sy * rec->stride[0];
        for (int x = 0; x < bs; x++) {
            int sx = x - pad;
            sx = sx < 0 ? 0 : (sx >= W ? W - 1 : sx);
            bord[(size_t)y * bs + x] = r[sx];
        }
    }
    int hz, vt;
    orc_lr_search_units(W, H, us, &hz, &vt);
    const uint16_t *S = src->p[0], *R = rec->p[0];
    const int ss = src->stride[0], rs = rec->stride[0];
    for (int row = 0; row < vt; row++) {
        for (int col = 0; col < hz; col++) {
            const int u = row * hz + col;
            uint64_t *e = sse + (size_t)u * 12;
            for (int i = 0; i < 12; i++) e[i] = UINT64_MAX;
            memset(&out[u], 0, sizeof(out[u]));
            out[u].type = AV1PF_LR_NONE;
            int x0 = col * us, x1 = col == hz - 1 ? W : x0 + us;
            int y0 = row * us, y1 = row == vt - 1 ? H : y0 + us;
            if (y0 > 0) y0 -= 8;
            if (y1 < H) y1 -= 8;
            const int w = x1 - x0, h = y1 - y0;
            uint64_t base = 0;
            for (int y = y0; y < y1; y++)
                for (int x = x0; x < x1; x++) {
                    const int d = (int)S[(size_t)y * ss + x] - (int)R[(size_t)y * rs + x];
                    base += (uint64_t)(d * d);
                }
            e[0] = base;
            if (base == 0) continue;
            /* Wiener: H-only approximation over the unit */
            for (int c = 0; c < 6; c++) {
                const int t0 = kWienerCand[c][0], t1 = kWienerCand[c][1], t2 = kWienerCand[c][2];
                const int f[7] = { t0, t1, t2, -2 * (t0 + t1 + t2), t2, t1, t0 };
                uint64_t s = 0;
                for (int y = y0; y < y1; y++)
                    for (int x = x0; x < x1; x++) {
                        int hs = 0;
                        for (int k = -3; k <= 3; k++) {
                            int cx = x + k;
                            cx = cx < 0 ? 0 : (cx >= W ? W - 1 : cx);
                            hs += f[k + 3] * (int)R[(size_t)y * rs + cx];
                        }
                        int pred = (hs + 64) >> 7;
                        pred = pred < 0 ? 0 : (pred > maxv ? maxv : pred);
                        const int d = (int)S[(size_t)y * ss + x] - pred;
                        s += (uint64_t)(d * d);
                    }
                e[1 + c] = s;
            }
            const int wbest = wiener_pick(e, (uint64_t)(w * h) << (2 * cs));
            const int valid = w > 0 && h > 0 && w <= 64 && h <= 64;
            const uint16_t *org = bord + (size_t)(y0 + pad) * bs + x0 + pad;
            if (wbest >= 0 && valid) {
                orc_wiener_unit(org, bs, tmp, 64, w, h, kWienerCand[wbest], kWienerCand[wbest], bd);
                uint64_t s = 0;
                for (int y = 0; y < h; y++)
                    for (int x = 0; x < w; x++) {
                        const int d = (int)S[(size_t)(y0 + y) * ss + x0 + x] - (int)tmp[y * 64 + x];
                        s += (uint64_t)(d * d);
                    }
                e[7] = s;
            }
            if (!valid) continue;
            for (int p = 0; p < 4; p++) {
                orc_sgr_unit(org, bs, tmp, 64, w, h, kSgrPreset[p], kSgrXqd[p], bd);
                uint64_t s = 0;
                for (int y = 0; y < h; y++)
                    for (int x = 0; x < w; x++) {
                        const int d = (int)S[(size_t)(y0 + y) * ss + x0 + x] - (int)tmp[y * 64 + x];
                        s += (uint64_t)(d * d);
                    }
                e[8 + p] = s;
            }
            lr_decide(e, w, h, cs, &out[u]);
        }
    }
    free(bord);
    free(tmp);
    return AV1PF_OK;
}

/* ---- exhaustive superset mode (no counterpart in the reference, whose
 * search is the curated one above) ------------------------------------------
 * Per unit of at most 64x64 (larger units stay NONE, as in the curated mode):
 *   sse[0]       baseline
 *   sse[1..6]    the six curated Wiener triples, exact (V = H)
 *   sse[7]       a least-squares Wiener fit, exact
 *   sse[8..11]   the four curated SGR presets, exact
 *   sse[12..27]  SGR set 0..15 with a least-squares projection, exact
 * Decision: candidates in that order, each replacing the running best when
 * sse + w*h < best (the curated search's rule), so the result is never worse
 * than the curated one.
 * Wiener fit: taps symmetric, Q7, summing to 128; alternating least squares,
 * horizontal taps given the vertical ones, then vertical given horizontal,
 * twice, from the identity; each 3x3 system from int64 statistics of the
 * 7x7 windows of the bordered recon, solved by Cramer's rule in double,
 * rounded (floor(x + 0.5)) and clamped to the tap ranges [-5,10], [-23,8],
 * [-17,46].  SGR fit: the projection weights minimising
 * sum (2^7 (s - u) 2^4 - xq0 (f0 - 16u) - xq1 (f1 - 16u))^2 (int64 sums,
 * double solve, floor(x + 0.5)), coded as xqd with the ranges [-96,31] and
 * [-32,95] the way libaom codes them. */
static const int kTapMin[3] = { -5, -23, -17 }, kTapMax[3] = { 10, 8, 46 };
/* SGR radii per set (r = 2 pass, r = 1 pass), the normative parameter table */
static const int kSgrRadius[16][2] = { { 2, 1 }, { 2, 1 }, { 2, 1 }, { 2, 1 }, { 2, 1 }, { 2, 1 }, { 2, 1 }, { 2, 1 },
                                       { 2, 1 }, { 2, 1 }, { 0, 1 }, { 0, 1 }, { 0, 1 }, { 0, 1 }, { 2, 0 }, { 2, 0 } };
static int sgr_radius_of(int set, int pass) { return kSgrRadius[set][pass]; }

static int round_clamp(double x, int lo, int hi)
{
    int v = (int)floor(x + 0.5);
    return v < lo ? lo : (v > hi ? hi : v);
}

/* 3x3 symmetric system M t = r (M given as m00 m01 m02 m11 m12 m22) */
static void solve_taps(const int64_t M[6], const int64_t R[3], int8_t out[3])
{
    const double a = (double)M[0], b = (double)M[1], c = (double)M[2], d = (double)M[3], e = (double)M[4],
                 f = (double)M[5];
    const double r0 = (double)R[0], r1 = (double)R[1], r2 = (double)R[2];
    const double c00 = d * f - e * e, c01 = b * f - e * c, c02 = b * e - d * c;
    const double det = a * c00 - b * c01 + c * c02;
    if (det == 0.0) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    const double x0 = (r0 * c00 - b * (r1 * f - e * r2) + c * (r1 * e - d * r2)) / det;
    const double x1 = (a * (r1 * f - e * r2) - r0 * c01 + c * (b * r2 - r1 * c)) / det;
    const double x2 = (a * (d * r2 - r1 * e) - b * (b * r2 - r1 * c) + r0 * c02) / det;
    out[0] = (int8_t)round_clamp(x0, kTapMin[0], kTapMax[0]);
    out[1] = (int8_t)round_clamp(x1, kTapMin[1], kTapMax[1]);
    out[2] = (int8_t)round_clamp(x2, kTapMin[2], kTapMax[2]);
}

/* Taps of one direction given the other's (fixed) taps: H when horiz, else V.
 * org: unit origin in the bordered recon (stride bs); S: source (stride ss). */
static void wiener_solve_dir(const uint16_t *org, int bs, const uint16_t *S, int ss, int w, int h,
                             const int8_t fixed[3], int horiz, int8_t out[3])
{
    const int ft[7] = { fixed[0], fixed[1], fixed[2], 128 - 2 * (fixed[0] + fixed[1] + fixed[2]), fixed[2], fixed[1],
                        fixed[0] };
    int64_t M[6] = { 0, 0, 0, 0, 0, 0 }, R[3] = { 0, 0, 0 };
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t Y[7];
            for (int k = 0; k < 7; k++) {
                int64_t acc = 0;
                for (int j = 0; j < 7; j++) {
                    /* horiz: Y[k] = vertical filter (fixed) at column x + k - 3 */
                    const int yy = horiz ? y + j - 3 : y + k - 3, xx = horiz ? x + k - 3 : x + j - 3;
                    acc += (int64_t)ft[j] * org[(ptrdiff_t)yy * bs + xx];
                }
                Y[k] = acc;
            }
            const int64_t z0 = Y[0] + Y[6] - 2 * Y[3], z1 = Y[1] + Y[5] - 2 * Y[3], z2 = Y[2] + Y[4] - 2 * Y[3];
            const int64_t t = ((int64_t)S[(ptrdiff_t)y * ss + x] << 14) - 128 * Y[3];
            M[0] += z0 * z0; M[1] += z0 * z1; M[2] += z0 * z2;
            M[3] += z1 * z1; M[4] += z1 * z2; M[5] += z2 * z2;
            R[0] += z0 * t; R[1] += z1 * t; R[2] += z2 * t;
        }
    solve_taps(M, R, out);
}

static void sgr_fit(int set, int64_t H00, int64_t H01, int64_t H11, int64_t C0, int64_t C1, int8_t xqd[2])
{
    const int r0 = sgr_radius_of(set, 0), r1 = sgr_radius_of(set, 1);
    double xq0 = 0.0, xq1 = 0.0;
    if (r0 && r1) {
        const double a = (double)H00, b = (double)H01, d = (double)H11, c0 = (double)C0, c1 = (double)C1;
        const double det = a * d - b * b;
        if (det != 0.0) {
            xq0 = (c0 * d - c1 * b) / det;
            xq1 = (c1 * a - c0 * b) / det;
        }
    } else if (r0) {
        if (H00) xq0 = (double)C0 / (double)H00;
    } else {
        if (H11) xq1 = (double)C1 / (double)H11;
    }
    if (r0 && r1) {
        xqd[0] = (int8_t)round_clamp(xq0, -96, 31);
        xqd[1] = (int8_t)round_clamp(128.0 - (double)xqd[0] - floor(xq1 + 0.5), -32, 95);
    } else if (r0) {
        xqd[0] = (int8_t)round_clamp(xq0, -96, 31);
        xqd[1] = (int8_t)(128 - xqd[0] < -32 ? -32 : (128 - xqd[0] > 95 ? 95 : 128 - xqd[0]));
    } else {
        xqd[0] = 0;
        xqd[1] = (int8_t)round_clamp(128.0 - floor(xq1 + 0.5), -32, 95);
    }
}

static uint64_t sgr_proj_sse(const uint16_t *org, int bs, const uint16_t *S, int ss, int w, int h, const int32_t *f0,
                             const int32_t *f1, int set, const int8_t xqd[2])
{
    int xq[2];
    orc_sgr_decode_xq(set, xqd, xq);
    uint64_t s = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const int32_t u = (int32_t)org[(ptrdiff_t)y * bs + x] << 4;
            int32_t v = u << 7;
            v += xq[0] * (f0[y * w + x] - u);
            v += xq[1] * (f1[y * w + x] - u);
            int32_t o = (int16_t)((v + (1 << 10)) >> 11);
            o = o < 0 ? 0 : (o > 255 ? 255 : o);
            const int d = (int)S[(ptrdiff_t)y * ss + x] - o;
            s += (uint64_t)(d * d);
        }
    return s;
}

int orc_lr_search_exhaustive(const orc_frame *src, const orc_frame *rec, int us, av1pf_lr_unit *out, uint64_t *sse,
                             int8_t *fit)
{
    if (!src || !rec || rec->bd != 8 || src->w != rec->w || src->h != rec->h || us < 32 || us > 256)
        return AV1PF_ERR_INVALID_ARG;
    const int W = rec->w, H = rec->h, pad = 3;
    const int bs = W + 2 * pad;
    uint16_t *bord = malloc(sizeof(uint16_t) * (size_t)bs * (H + 2 * pad));
    uint16_t *tmp = malloc(sizeof(uint16_t) * 64 * 64);
    int32_t *f0 = malloc(sizeof(int32_t) * 64 * 64), *f1 = malloc(sizeof(int32_t) * 64 * 64);
    for (int y = 0; y < H + 2 * pad; y++) {
        int sy = y - pad;
        sy = sy < 0 ? 0 : (sy >= H ? H - 1 : sy);
        const uint16_t *r = rec->p[0] + (size_t)sy * rec->stride[0];
        for (int x = 0; x < bs; x++) {
            int sx = x - pad;
            sx = sx < 0 ? 0 : (sx >= W ? W - 1 : sx);
            bord[(size_t)y * bs + x] = r[sx];
        }
    }
    int hz, vt;
    orc_lr_search_units(W, H, us, &hz, &vt);
    const int ss = src->stride[0];
    for (int row = 0; row < vt; row++) {
        for (int col = 0; col < hz; col++) {
            const int u = row * hz + col;
            uint64_t *e = sse + (size_t)u * 28;
            int8_t *ft = fit + (size_t)u * 38;
            for (int i = 0; i < 28; i++) e[i] = UINT64_MAX;
            memset(ft, 0, 38);
            memset(&out[u], 0, sizeof(out[u]));
            out[u].type = AV1PF_LR_NONE;
            int x0 = col * us, x1 = col == hz - 1 ? W : x0 + us;
            int y0 = row * us, y1 = row == vt - 1 ? H : y0 + us;
            if (y0 > 0) y0 -= 8;
            if (y1 < H) y1 -= 8;
            const int w = x1 - x0, h = y1 - y0;
            const uint16_t *S = src->p[0] + (size_t)y0 * ss + x0;
            const uint16_t *org = bord + (size_t)(y0 + pad) * bs + x0 + pad;
            uint64_t base = 0;
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const int d = (int)S[(size_t)y * ss + x] - (int)org[(size_t)y * bs + x];
                    base += (uint64_t)(d * d);
                }
            e[0] = base;
            if (base == 0 || w > 64 || h > 64) continue;
            const uint64_t wh = (uint64_t)(w * h);
            uint64_t best = base;
            av1pf_lr_unit res;
            memset(&res, 0, sizeof(res));
            res.type = AV1PF_LR_NONE;
            /* Wiener: curated triples, then the fit */
            int8_t taps[7][2][3];
            for (int c = 0; c < 6; c++) { memcpy(taps[c][0], kWienerCand[c], 3); memcpy(taps[c][1], kWienerCand[c], 3); }
            {
                int8_t v[3] = { 0, 0, 0 }, hh[3] = { 0, 0, 0 };
                for (int it = 0; it < 2; it++) {
                    wiener_solve_dir(org, bs, S, ss, w, h, v, 1, hh);
                    wiener_solve_dir(org, bs, S, ss, w, h, hh, 0, v);
                }
                memcpy(taps[6][0], v, 3);
                memcpy(taps[6][1], hh, 3);
                memcpy(ft, v, 3);
                memcpy(ft + 3, hh, 3);
            }
            for (int c = 0; c < 7; c++) {
                orc_wiener_unit(org, bs, tmp, 64, w, h, taps[c][0], taps[c][1], 8);
                uint64_t s = 0;
                for (int y = 0; y < h; y++)
                    for (int x = 0; x < w; x++) {
                        const int d = (int)S[(size_t)y * ss + x] - (int)tmp[y * 64 + x];
                        s += (uint64_t)(d * d);
                    }
                e[1 + c] = s;
                if (s + wh < best) {
                    best = s;
                    memset(&res, 0, sizeof(res));
                    res.type = AV1PF_LR_WIENER;
                    memcpy(res.wiener_v, taps[c][0], 3);
                    memcpy(res.wiener_h, taps[c][1], 3);
                }
            }
            /* SGR: per set the curated preset (when the set has one), then the fit */
            for (int set = 0; set < 16; set++) {
                orc_sgr_filters(org, bs, w, h, set, 8, f0, f1);
                int p = -1;
                for (int q = 0; q < 4; q++)
                    if (kSgrPreset[q] == set) p = q;
                int8_t cxqd[2][2];
                int nc = 0;
                if (p >= 0) { cxqd[nc][0] = kSgrXqd[p][0]; cxqd[nc][1] = kSgrXqd[p][1]; nc++; }
                {
                    int64_t H00 = 0, H01 = 0, H11 = 0, C0 = 0, C1 = 0;
                    for (int y = 0; y < h; y++)
                        for (int x = 0; x < w; x++) {
                            const int32_t uu = (int32_t)org[(size_t)y * bs + x] << 4;
                            const int64_t d0 = sgr_radius_of(set, 0) ? f0[y * w + x] - uu : 0;
                            const int64_t d1 = sgr_radius_of(set, 1) ? f1[y * w + x] - uu : 0;
                            const int64_t ee = (((int64_t)S[(size_t)y * ss + x] << 4) - uu) * 128;
                            H00 += d0 * d0; H01 += d0 * d1; H11 += d1 * d1; C0 += d0 * ee; C1 += d1 * ee;
                        }
                    sgr_fit(set, H00, H01, H11, C0, C1, cxqd[nc]);
                    ft[6 + 2 * set] = cxqd[nc][0];
                    ft[7 + 2 * set] = cxqd[nc][1];
                    nc++;
                }
                for (int k = 0; k < nc; k++) {
                    const uint64_t s = sgr_proj_sse(org, bs, S, ss, w, h, f0, f1, set, cxqd[k]);
                    e[k == nc - 1 ? 12 + set : 8 + p] = s;
                    if (s + wh < best) {
                        best = s;
                        memset(&res, 0, sizeof(res));
                        res.type = AV1PF_LR_SGRPROJ;
                        res.sgr_set = (uint8_t)set;
                        res.sgr_xqd[0] = cxqd[k][0];
                        res.sgr_xqd[1] = cxqd[k][1];
                    }
                }
            }
            out[u] = res;
        }
    }
    free(bord); free(tmp); free(f0); free(f1);
    return AV1PF_OK;
}
