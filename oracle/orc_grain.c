/*
 * CPU oracle: AV1 film-grain synthesis as goav1 applies it to a display copy.
 * Test infrastructure only (see orc.h).
 *
 * Restates:
 *   filmgrain/constants.go:8-32        template sizes, block size 32, overlap 2
 *   filmgrain/random.go:9-72           16-bit LFSR, plane / stripe seeds
 *   filmgrain/gaussian.go:186-188      Gaussian table lookup (normative table, av1pf_gaussian_table.h)
 *   filmgrain/luma.go:24-76            luma template + generic AR recursion
 *   filmgrain/chroma.go:29-90,223-241  chroma template, AR with luma-average term
 *   filmgrain/scaling.go:18-77         BuildScalingLUT / scaleLUT (8/10/12-bit interpolation)
 *   filmgrain/apply.go:72-310          per 32x32 block offsets, overlap blends, apply
 *   filmgrain/apply.go:480-549         chroma overlap weights, chroma scaling index, chroma sample
 *   filmgrain/blend.go:16-37           luma overlap blend and luma sample
 *   filmgrain/sample.go:9-77           template indexing and clip bounds
 *   decoder/postfilter_filmgrain.go:192-263,302-331,516-613   plane activity, LUT choice,
 *                                      chroma before luma (chroma reads pre-grain luma)
 *   decoder/postfilter_filmgrain.go:744-772   chroma row parameters (odd luma width)
 */
#include "orc.h"
#include "av1pf_gaussian_table.h"
#include <stdlib.h>
#include <string.h>

static const short av1pf_gaussian_sequence[2048] = { AV1PF_GAUSSIAN_VALUES };

#define LGW 82
#define LGH 73
#define CGW_SUB 44
#define CGH_SUB 38

static int iclip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static int r2(int v, int n) { return n <= 0 ? v : (v + (1 << (n - 1))) >> n; }

/* random.go:51-62 */
uint16_t orc_fg_random_step(uint16_t *reg, int bits)
{
    uint16_t r = *reg;
    uint16_t bit = (uint16_t)(((r >> 0) ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1);
    r = (uint16_t)((r >> 1) | (bit << 15));
    *reg = r;
    if (bits == 16) return r;
    return (uint16_t)((r >> (16 - bits)) & ((1 << bits) - 1));
}

static uint16_t stripe_seed(uint16_t seed, int luma_line)
{
    int num = luma_line >> 5;
    uint16_t reg = seed;
    reg ^= (uint16_t)(((num * 37 + 178) & 255) << 8);
    reg ^= (uint16_t)((num * 173 + 105) & 255);
    return reg;
}

/* scaling.go:18-51 */
void orc_fg_scaling_lut(uint8_t lut[256], const uint8_t (*pts)[2], int n)
{
    memset(lut, 0, 256);
    if (n == 0) return;
    for (int x = 0; x < pts[0][0]; x++) lut[x] = pts[0][1];
    for (int i = 0; i + 1 < n; i++) {
        int dy = (int)pts[i + 1][1] - pts[i][1];
        int dx = (int)pts[i + 1][0] - pts[i][0];
        int64_t delta = (int64_t)dy * ((65536 + (dx >> 1)) / dx);
        for (int x = 0; x < dx; x++)
            lut[pts[i][0] + x] = (uint8_t)(pts[i][1] + (int)(((int64_t)x * delta + 32768) >> 16));
    }
    for (int x = pts[n - 1][0]; x < 256; x++) lut[x] = pts[n - 1][1];
}

/* scaling.go:55-77 */
static int scale_lut(const uint8_t *lut, int index, int bd)
{
    int mx = (1 << bd) - 1;
    index = iclip(index, 0, mx);
    int sh = bd - 8;
    int x = index >> sh;
    if (bd == 8 || x == 255) return lut[x];
    int rem = index - (x << sh);
    int start = lut[x], end = lut[x + 1];
    return (uint8_t)(start + r2((end - start) * rem, sh));
}

static void ar_filter(int16_t *g, int stride, int w, int h, int lag, const int8_t *coeffs, int shift,
                      int bd, const int16_t *luma, int chroma, int num_y, int ssx, int ssy)
{
    int rounding = 1 << (shift - 1);
    int gmin = -(1 << (bd - 1)), gmax = (1 << (bd - 1)) - 1;
    for (int y = 3; y < h; y++) {
        for (int x = 3; x < w - 3; x++) {
            int sum = 0, pos = 0;
            for (int dr = -lag; dr <= 0; dr++) {
                for (int dc = -lag; dc <= lag; dc++) {
                    if (dr == 0 && dc == 0) {
                        if (chroma && num_y) {
                            /* chromaLumaAverage (chroma.go:223-241) */
                            int lx = ((x - 3) << ssx) + 3, ly = ((y - 3) << ssy) + 3, s = 0;
                            for (int dy = 0; dy <= ssy; dy++)
                                for (int dx = 0; dx <= ssx; dx++) s += luma[(ly + dy) * LGW + lx + dx];
                            sum += r2(s, ssx + ssy) * coeffs[pos];
                        }
                        goto done;
                    }
                    sum += coeffs[pos] * g[(y + dr) * stride + x + dc];
                    pos++;
                }
            }
        done:
            g[y * stride + x] = (int16_t)iclip(g[y * stride + x] + ((sum + rounding) >> shift), gmin, gmax);
        }
    }
}

/* luma.go:24-76 */
int orc_fg_luma_grain(int16_t *grain, const av1pf_grain_params *g, int bd)
{
    memset(grain, 0, LGW * LGH * sizeof(int16_t));
    if (g->num_y_points == 0) return AV1PF_OK;
    uint16_t reg = g->seed;
    int shift = 12 - bd + g->grain_scale_shift;
    for (int y = 0; y < LGH; y++)
        for (int x = 0; x < LGW; x++)
            grain[y * LGW + x] = (int16_t)r2(av1pf_gaussian_sequence[orc_fg_random_step(&reg, 11)], shift);
    if (g->ar_coeff_lag)
        ar_filter(grain, LGW, LGW, LGH, g->ar_coeff_lag, g->ar_coeffs_y, g->ar_coeff_shift, bd, NULL, 0, 0, 0, 0);
    return AV1PF_OK;
}

/* chroma.go:29-90 */
int orc_fg_chroma_grain(int16_t *grain, const int16_t *luma, const av1pf_grain_params *g,
                        int bd, int plane, int ssx, int ssy)
{
    memset(grain, 0, LGW * LGH * sizeof(int16_t));
    uint16_t reg = (uint16_t)(g->seed ^ (plane == 1 ? 0xb524 : 0x49d8));
    int w = ssx ? CGW_SUB : LGW, h = ssy ? CGH_SUB : LGH;
    int shift = 12 - bd + g->grain_scale_shift;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            grain[y * LGW + x] = (int16_t)r2(av1pf_gaussian_sequence[orc_fg_random_step(&reg, 11)], shift);
    if (g->ar_coeff_lag == 0 && g->num_y_points == 0) return AV1PF_OK;
    ar_filter(grain, LGW, w, h, g->ar_coeff_lag, plane == 1 ? g->ar_coeffs_cb : g->ar_coeffs_cr,
              g->ar_coeff_shift, bd, luma, 1, g->num_y_points, ssx, ssy);
    return AV1PF_OK;
}

/* sample.go: template offsets */
static int luma_off(int n) { return 3 + 2 * (3 + n); }
static int chroma_off(int n, int shift) { return 3 + (2 >> shift) * (3 + n); }

static int16_t blend(int prev, int cur, int wp, int wc, int bd)
{
    return (int16_t)iclip(r2(prev * wp + cur * wc, 5), -(1 << (bd - 1)), (1 << (bd - 1)) - 1);
}

typedef struct fg_ctx {
    const av1pf_grain_params *g;
    int bd, ssx, ssy;
    int lumaw, lumah;         /* luma plane size */
    int mat_id;
} fg_ctx;

/* One plane: luma when plane == 0.  src/dst are full planes (dst may alias src
 * for luma); luma is the PRE-grain luma plane. */
static void fg_plane(const fg_ctx *c, int plane, const uint16_t *src, int sstride, uint16_t *dst, int dstride,
                     const uint16_t *luma, int lstride, const int16_t *grain, const uint8_t *lut,
                     int pw, int ph)
{
    const av1pf_grain_params *g = c->g;
    const int bd = c->bd;
    const int sx = plane ? c->ssx : 0, sy = plane ? c->ssy : 0;
    const int bw = 32 >> sx, bh = 32 >> sy;
    const int gw = LGW;
    int minv = 0, maxv = (256 << (bd - 8)) - 1;
    if (g->clip_to_restricted_range) {
        minv = 16 << (bd - 8);
        maxv = (plane == 0 ? 235 : (c->mat_id ? 235 : 240)) << (bd - 8);
    }
    const int odd_luma = plane && sx && (c->lumaw & 1);
    const int mult = plane == 1 ? g->cb_mult : g->cr_mult;
    const int lmult = plane == 1 ? g->cb_luma_mult : g->cr_luma_mult;
    const int coff = plane == 1 ? g->cb_offset : g->cr_offset;
    const int rows = (ph + bh - 1) / bh;
    for (int row = 0; row < rows; row++) {
        int height = ph - row * bh < bh ? ph - row * bh : bh;
        int nr = g->overlap && row > 0 ? 2 : 1;
        uint16_t rnd[2];
        for (int i = 0; i < nr; i++) rnd[i] = stripe_seed(g->seed, (row - i) * 32);
        uint8_t offs[2][2] = { { 0, 0 }, { 0, 0 } };
        for (int bx = 0; bx < pw; bx += bw) {
            int bwid = pw - bx < bw ? pw - bx : bw;
            if (g->overlap && bx > 0)
                for (int i = 0; i < nr; i++) offs[1][i] = offs[0][i];
            for (int i = 0; i < nr; i++) offs[0][i] = (uint8_t)orc_fg_random_step(&rnd[i], 8);
            int ys = g->overlap && row > 0 ? (height < (2 >> sy) ? height : (2 >> sy)) : 0;
            int xs = g->overlap && bx > 0 ? (bwid < (2 >> sx) ? bwid : (2 >> sx)) : 0;
            for (int y = 0; y < height; y++) {
                for (int x = 0; x < bwid; x++) {
#define GS(off, bc, br, xx, yy) grain[(((plane ? chroma_off((off) & 15, sy) : luma_off((off) & 15)) + (yy) + bh * (br)) * gw) + \
                                      (plane ? chroma_off((off) >> 4, sx) : luma_off((off) >> 4)) + (xx) + bw * (bc)]
                    int gv = GS(offs[0][0], 0, 0, x, y);
                    if (x < xs) {
                        int left = GS(offs[1][0], 1, 0, x, y);
                        if (plane && sx) gv = blend(left, gv, 23, 22, bd);
                        else gv = blend(left, gv, x == 1 ? 17 : 27, x == 1 ? 27 : 17, bd);
                    }
                    if (y < ys) {
                        int top = GS(offs[0][1], 0, 1, x, y);
                        if (x < xs) {
                            int tl = GS(offs[1][1], 1, 1, x, y);
                            if (plane && sx) top = blend(tl, top, 23, 22, bd);
                            else top = blend(tl, top, x == 1 ? 17 : 27, x == 1 ? 27 : 17, bd);
                        }
                        if (plane && sy) gv = blend(top, gv, 23, 22, bd);
                        else gv = blend(top, gv, y == 1 ? 17 : 27, y == 1 ? 27 : 17, bd);
                    }
#undef GS
                    int px = bx + x, py = row * bh + y;
                    int s = src[(size_t)py * sstride + px];
                    int idx;
                    if (plane == 0) {
                        idx = s;
                    } else {
                        /* chromaScalingIndex (apply.go:501-523) */
                        int lx = px << sx, ly = py << sy;
                        int avg = luma[(size_t)ly * lstride + lx];
                        if (sx && !(odd_luma && px == pw - 1)) avg = (avg + luma[(size_t)ly * lstride + lx + 1] + 1) >> 1;
                        if (g->chroma_scaling_from_luma) idx = avg;
                        else {
                            int combined = avg * (lmult - 128) + s * (mult - 128);
                            idx = iclip((combined >> 6) + ((coff << (bd - 8)) - (1 << bd)), 0, (1 << bd) - 1);
                        }
                    }
                    int scale = scale_lut(lut, idx, bd);
                    int noise = r2(scale * gv, g->scaling_shift);
                    dst[(size_t)py * dstride + px] = (uint16_t)iclip(s + noise, minv, maxv);
                }
            }
        }
    }
}

int orc_filmgrain(const orc_frame *src, orc_frame *dst, const av1pf_grain_params *g)
{
    const int bd = src->bd;
    /* copy whole frame into the display buffer (bindFilmGrainPostFilterOutput) */
    for (int pl = 0; pl < (src->mono ? 1 : 3); pl++)
        for (int y = 0; y < src->rows[pl] && y < dst->rows[pl]; y++)
            memcpy(dst->p[pl] + (size_t)y * dst->stride[pl], src->p[pl] + (size_t)y * src->stride[pl],
                   (size_t)(src->stride[pl] < dst->stride[pl] ? src->stride[pl] : dst->stride[pl]) * 2);
    if (!g->chroma_scaling_from_luma && !g->num_y_points && !g->num_cb_points && !g->num_cr_points)
        return AV1PF_OK;
    if (g->scaling_shift < 8 || g->scaling_shift > 11) return AV1PF_ERR_INVALID_ARG;
    int active[3] = { g->num_y_points != 0, 0, 0 };
    if (!src->mono) {
        active[1] = g->num_cb_points != 0 || g->chroma_scaling_from_luma;
        active[2] = g->num_cr_points != 0 || g->chroma_scaling_from_luma;
    }
    uint8_t luts[3][256];
    orc_fg_scaling_lut(luts[0], g->y_points, g->num_y_points);
    if (g->chroma_scaling_from_luma) { memcpy(luts[1], luts[0], 256); memcpy(luts[2], luts[0], 256); }
    else {
        orc_fg_scaling_lut(luts[1], g->cb_points, g->num_cb_points);
        orc_fg_scaling_lut(luts[2], g->cr_points, g->num_cr_points);
    }
    int16_t *lg = calloc(LGW * LGH, sizeof(int16_t));
    int16_t *cg = calloc(LGW * LGH, sizeof(int16_t));
    if (active[0]) orc_fg_luma_grain(lg, g, bd);
    fg_ctx c = { g, bd, src->ssx, src->ssy, src->w, src->h, src->matrix_identity };
    int cw = (src->w + src->ssx) >> src->ssx, ch = (src->h + src->ssy) >> src->ssy;
    for (int pl = 1; pl <= 2; pl++) {
        if (!active[pl]) continue;
        orc_fg_chroma_grain(cg, lg, g, bd, pl, src->ssx, src->ssy);
        fg_plane(&c, pl, src->p[pl], src->stride[pl], dst->p[pl], dst->stride[pl],
                 src->p[0], src->stride[0], cg, luts[pl], cw, ch);
    }
    if (active[0])
        fg_plane(&c, 0, src->p[0], src->stride[0], dst->p[0], dst->stride[0], NULL, 0, lg, luts[0], src->w, src->h);
    free(lg); free(cg);
    return AV1PF_OK;
}

/* random.go:39-48 (NewStripeRandom), exported for the known-answer tests */
uint16_t orc_fg_stripe_seed(uint16_t seed, int luma_line) { return stripe_seed(seed, luma_line); }
