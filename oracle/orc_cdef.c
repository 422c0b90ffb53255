/*
 * CPU oracle: CDEF as goav1's decoder runs it (snapshot unit walk).
 * Test infrastructure only (see orc.h).
 *
 * Restates:
 *   cdef/constants.go:9-18              BStride 144, borders 2/8, VeryLarge 0x4000
 *   cdef/strength.go:12-23              DecodeStrength (secondary code 3 -> 4)
 *   cdef/direction.go:37-58,432-502     findDirectionScalar + finishDirection
 *   cdef/filter.go:53-69                direction offsets and tap tables
 *   cdef/filter.go:79-170               filterBlockPureGo (pri+sec with VeryLarge-masked clip)
 *   cdef/filter.go:173-188              constrainShift / constrain
 *   cdef/filter_u8.go:124-134           constrainU8 / cdefRound
 *   cdef/filter.go:373-478              filterFrameBlocks (strength shifts, chroma damping -1,
 *                                       luma direction search, 4:2:2 / 4:4:0 direction remap)
 *   cdef/filter.go:493-560              filterUnitBlocksPureGo (identity copy at zero strength)
 *   cdef/filter.go:603-636              adjustStrength / convertChromaDirections
 *   decoder/postfilter_cdef.go:525-690  plane selection (processPlane, forceLumaDirections)
 *   decoder/postfilter_cdef.go:936-1031 unitStrength + applySnapshot (prefill, directionOnly)
 *   decoder/postfilter_cdef.go:1041-1049 frameWorkCDEFPlaneNeedsPrimaryDirection
 *   decoder/postfilter_cdef.go:1101-1250 sentinel fill + input copy with last row/column
 *                                       replication for non-final units
 *   decoder/postfilter_cdef.go:1390-1430 MI-aligned plane extent, unit grid
 *   decoder/postfilter_cdef.go:1571-1594 block list from the per-unit noskip mask
 */
#include "orc.h"
#include <stdlib.h>
#include <string.h>

#define BSTRIDE 144
#define VBORDER 2
#define HBORDER 8
#define VERY_LARGE 0x4000
#define INPUT_SIZE (BSTRIDE * (128 + 2 * VBORDER))

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int ilog2(unsigned v) { return 31 - __builtin_clz(v); }

static const int div_table[9] = { 0, 840, 420, 280, 210, 168, 140, 120, 105 };

/* cdef/direction.go:467-502: int32 arithmetic (wrapping as Go does). */
static int32_t cost_straight(const int32_t *p)
{
    uint32_t s = 0;
    for (int i = 0; i < 8; i++) s += (uint32_t)p[i] * (uint32_t)p[i];
    return (int32_t)(s * (uint32_t)div_table[8]);
}
static int32_t cost_diag(const int32_t *p)
{
    uint32_t s = 0;
    for (int i = 0; i < 7; i++)
        s += ((uint32_t)p[i] * (uint32_t)p[i] + (uint32_t)p[14 - i] * (uint32_t)p[14 - i]) * (uint32_t)div_table[i + 1];
    s += (uint32_t)p[7] * (uint32_t)p[7] * (uint32_t)div_table[8];
    return (int32_t)s;
}
static int32_t cost_odd(const int32_t *p)
{
    uint32_t s = 0;
    for (int i = 3; i < 8; i++) s += (uint32_t)p[i] * (uint32_t)p[i];
    s *= (uint32_t)div_table[8];
    for (int i = 0; i < 3; i++)
        s += ((uint32_t)p[i] * (uint32_t)p[i] + (uint32_t)p[10 - i] * (uint32_t)p[10 - i]) * (uint32_t)div_table[2 * i + 2];
    return (int32_t)s;
}

/* cdef/direction.go:37-58 and :432-465 */
void orc_cdef_find_dir(const uint16_t *img, int stride, int coeff_shift, int *dir, int32_t *var)
{
    int32_t partial[8][15];
    memset(partial, 0, sizeof(partial));
    for (int i = 0; i < 8; i++) {
        for (int j = 0; j < 8; j++) {
            int32_t x = (int32_t)(img[i * stride + j] >> coeff_shift) - 128;
            partial[0][i + j] += x;
            partial[1][i + (j >> 1)] += x;
            partial[2][i] += x;
            partial[3][3 + i - (j >> 1)] += x;
            partial[4][7 + i - j] += x;
            partial[5][3 - (i >> 1) + j] += x;
            partial[6][j] += x;
            partial[7][(i >> 1) + j] += x;
        }
    }
    int32_t c[8];
    c[0] = cost_diag(partial[0]); c[1] = cost_odd(partial[1]);
    c[2] = cost_straight(partial[2]); c[3] = cost_odd(partial[3]);
    c[4] = cost_diag(partial[4]); c[5] = cost_odd(partial[5]);
    c[6] = cost_straight(partial[6]); c[7] = cost_odd(partial[7]);
    int best = 0;
    int32_t bc = c[0];
    for (int d = 1; d < 8; d++)
        if (c[d] > bc) { bc = c[d]; best = d; }
    *dir = best;
    *var = (int32_t)((uint32_t)bc - (uint32_t)c[(best + 4) & 7]) >> 10;
}

/* cdef/filter.go:53-69 */
static const int dir_off[12][2] = {
    { 1 * BSTRIDE, 2 * BSTRIDE },     { 1 * BSTRIDE, 2 * BSTRIDE - 1 },
    { -1 * BSTRIDE + 1, -2 * BSTRIDE + 2 }, { 1, -1 * BSTRIDE + 2 },
    { 1, 2 },                         { 1, 1 * BSTRIDE + 2 },
    { 1 * BSTRIDE + 1, 2 * BSTRIDE + 2 }, { 1 * BSTRIDE, 2 * BSTRIDE + 1 },
    { 1 * BSTRIDE, 2 * BSTRIDE },     { 1 * BSTRIDE, 2 * BSTRIDE - 1 },
    { -1 * BSTRIDE + 1, -2 * BSTRIDE + 2 }, { 1, -1 * BSTRIDE + 2 },
};
static const int pri_taps[2][2] = { { 4, 2 }, { 3, 3 } };

/* cdef/filter.go:173-179 */
static int constrain_shift(int t, int damping)
{
    if (t == 0) return 0;
    return imax(damping - ilog2((unsigned)t), 0);
}
/* cdef/filter_u8.go:124-129 */
static int constrain(int diff, int t, int shift)
{
    int a = diff < 0 ? -diff : diff;
    int l = imin(imax(t - (a >> shift), 0), a);
    return diff < 0 ? -l : l;
}
/* cdef/filter_u8.go:132-134: x + ((8 + sum - (sum < 0)) >> 4) */
static int cdef_round(int x, int sum) { return x + ((8 + sum - (sum < 0)) >> 4); }

/* cdef/filter.go:79-170; `in` points at the block origin inside a BSTRIDE buffer. */
void orc_cdef_filter_block(uint16_t *dst, int dst_stride, const uint16_t *in,
                           int pri, int sec, int dir, int pri_damp, int sec_damp,
                           int coeff_shift, int w, int h)
{
    const int *pt = pri_taps[(pri >> coeff_shift) & 1];
    const int p0 = dir_off[dir + 2][0], p1 = dir_off[dir + 2][1];
    const int s0a = dir_off[dir + 4][0], s0b = dir_off[dir + 4][1];
    const int s1a = dir_off[dir][0], s1b = dir_off[dir][1];
    const int psh = constrain_shift(pri, pri_damp), ssh = constrain_shift(sec, sec_damp);
    for (int r = 0; r < h; r++) {
        for (int c = 0; c < w; c++) {
            const uint16_t *b = in + r * BSTRIDE + c;
            int x = b[0];
            if (pri && sec) {
                int a0 = b[p0], a1 = b[-p0], a2 = b[p1], a3 = b[-p1];
                int t[8] = { b[s0a], b[-s0a], b[s1a], b[-s1a], b[s0b], b[-s0b], b[s1b], b[-s1b] };
                int sum = pt[0] * (constrain(a0 - x, pri, psh) + constrain(a1 - x, pri, psh)) +
                          pt[1] * (constrain(a2 - x, pri, psh) + constrain(a3 - x, pri, psh));
                sum += 2 * (constrain(t[0] - x, sec, ssh) + constrain(t[1] - x, sec, ssh) +
                            constrain(t[2] - x, sec, ssh) + constrain(t[3] - x, sec, ssh));
                sum += constrain(t[4] - x, sec, ssh) + constrain(t[5] - x, sec, ssh) +
                       constrain(t[6] - x, sec, ssh) + constrain(t[7] - x, sec, ssh);
                const int mk = VERY_LARGE - 1;
                int hi = x, lo = x;
                int all[12] = { a0, a1, a2, a3, t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7] };
                for (int k = 0; k < 12; k++) {
                    hi = imax(hi, all[k] & mk);
                    lo = imin(lo, all[k]);
                }
                dst[r * dst_stride + c] = (uint16_t)imin(imax(cdef_round(x, sum), lo), hi);
            } else if (pri) {
                int sum = pt[0] * (constrain(b[p0] - x, pri, psh) + constrain(b[-p0] - x, pri, psh)) +
                          pt[1] * (constrain(b[p1] - x, pri, psh) + constrain(b[-p1] - x, pri, psh));
                dst[r * dst_stride + c] = (uint16_t)cdef_round(x, sum);
            } else if (sec) {
                int sum = 2 * (constrain(b[s0a] - x, sec, ssh) + constrain(b[-s0a] - x, sec, ssh) +
                               constrain(b[s1a] - x, sec, ssh) + constrain(b[-s1a] - x, sec, ssh)) +
                          constrain(b[s0b] - x, sec, ssh) + constrain(b[-s0b] - x, sec, ssh) +
                          constrain(b[s1b] - x, sec, ssh) + constrain(b[-s1b] - x, sec, ssh);
                dst[r * dst_stride + c] = (uint16_t)cdef_round(x, sum);
            } else {
                dst[r * dst_stride + c] = (uint16_t)x;
            }
        }
    }
}

/* cdef/filter.go:603-612 */
int orc_cdef_adjust_strength(int strength, int32_t var)
{
    if (var == 0) return 0;
    int i = 0;
    int32_t v = var >> 6;
    if (v != 0) i = imin(ilog2((unsigned)v), 12);
    return (strength * (4 + i) + 8) >> 4;
}

/* ---- unit walk ---------------------------------------------------------- */
typedef struct cdef_plane_src {
    const uint16_t *pix; int stride; int w, h;   /* MI-aligned extent of the snapshot */
} cdef_plane_src;

/* decoder/postfilter_cdef.go:1101-1127 */
static void fill_sentinels(uint16_t *in, int uw, int uh)
{
    int fw = uw + 2 * HBORDER, fh = uh + 2 * VBORDER;
    for (int r = 0; r < VBORDER; r++)
        for (int c = 0; c < fw; c++) in[r * BSTRIDE + c] = VERY_LARGE;
    for (int r = VBORDER; r < VBORDER + uh; r++) {
        for (int c = 0; c < HBORDER; c++) in[r * BSTRIDE + c] = VERY_LARGE;
        for (int c = HBORDER + uw; c < fw; c++) in[r * BSTRIDE + c] = VERY_LARGE;
    }
    for (int r = VBORDER + uh; r < fh; r++)
        for (int c = 0; c < fw; c++) in[r * BSTRIDE + c] = VERY_LARGE;
}

/* decoder/postfilter_cdef.go:1151-1250 */
static void copy_input(uint16_t *in, const cdef_plane_src *s, int ux, int uy, int uw, int uh)
{
    int sx0 = imax(ux - HBORDER, 0), sy0 = imax(uy - VBORDER, 0);
    int sx1 = imin(ux + uw + HBORDER, s->w), sy1 = imin(uy + uh + VBORDER, s->h);
    int cw = sx1 - sx0, ch = sy1 - sy0;
    if (cw <= 0 || ch <= 0) return;
    int doff = VBORDER * BSTRIDE + HBORDER + (sy0 - uy) * BSTRIDE + (sx0 - ux);
    for (int r = 0; r < ch; r++)
        for (int c = 0; c < cw; c++) in[doff + r * BSTRIDE + c] = s->pix[(sy0 + r) * s->stride + sx0 + c];
    int want_y1 = uy + uh + VBORDER;
    int row_ext = sy1 < want_y1 && sy1 > 0 && uy + uh < s->h;
    if (row_ext) {
        int last = doff + (ch - 1) * BSTRIDE;
        for (int row = sy1; row < want_y1; row++) {
            int e = doff + (row - sy0) * BSTRIDE;
            memmove(&in[e], &in[last], (size_t)cw * sizeof(uint16_t));
        }
    }
    int want_x1 = ux + uw + HBORDER;
    if (sx1 < want_x1 && sx1 > 0 && ux + uw < s->w) {
        int rows = row_ext ? want_y1 - sy0 : ch;
        for (int r = 0; r < rows; r++) {
            int ro = doff + r * BSTRIDE;
            uint16_t last = in[ro + cw - 1];
            for (int i = ro + cw; i < ro + (want_x1 - sx0); i++) in[i] = last;
        }
    }
}

static int strength_pri(uint8_t packed) { return packed >> 2; }
static int strength_sec(uint8_t packed) { int s = packed & 3; return s == 3 ? 4 : s; }

int orc_cdef(orc_frame *f, const av1pf_cdef_params *p)
{
    if (!f || !p || !p->index) return AV1PF_ERR_INVALID_ARG;
    const int cols = (f->w + 63) / 64, rows = (f->h + 63) / 64;
    if (p->unit_cols != cols || p->unit_rows != rows || p->strength_count < 1 || p->strength_count > 8)
        return AV1PF_ERR_INVALID_ARG;
    const int n = p->strength_count;
    int yfilt = 0, uvfilt = 0;
    for (int i = 0; i < n; i++) { if (p->y_strength[i]) yfilt = 1; if (p->uv_strength[i]) uvfilt = 1; }
    if (f->mono) uvfilt = 0;
    if (!yfilt && !uvfilt) return AV1PF_OK;
    for (int i = 0; i < cols * rows; i++)
        if (p->index[i] >= n) return AV1PF_ERR_INVALID_ARG;
    const int chroma_filtering = uvfilt;
    const int coeff_shift = f->bd - 8;

    uint8_t (*dirs)[8][8] = calloc((size_t)cols * rows, sizeof(*dirs));
    int32_t (*vars)[8][8] = calloc((size_t)cols * rows, sizeof(*vars));
    uint16_t *input = malloc(INPUT_SIZE * sizeof(uint16_t));
    uint16_t *unit_dst = malloc(INPUT_SIZE * sizeof(uint16_t));
    uint16_t *snap = NULL;
    if (!dirs || !vars || !input || !unit_dst) { free(dirs); free(vars); free(input); free(unit_dst); return AV1PF_ERR_INVALID_ARG; }
    const int aligned_lw = (f->w + 7) & ~7, aligned_lh = (f->h + 7) & ~7;

    for (int plane = 0; plane < 3; plane++) {
        int process = plane == 0 ? (yfilt || chroma_filtering) : uvfilt;
        if (plane && f->mono) process = 0;
        if (!process) continue;
        const int xd = plane ? f->ssx : 0, yd = plane ? f->ssy : 0;
        cdef_plane_src src;
        src.w = imin(aligned_lw >> xd, f->stride[plane]);
        src.h = aligned_lh >> yd;
        src.stride = src.w;
        free(snap);
        snap = malloc((size_t)src.w * src.h * sizeof(uint16_t));
        for (int y = 0; y < src.h; y++)
            memcpy(&snap[(size_t)y * src.w], &f->p[plane][(size_t)y * f->stride[plane]], (size_t)src.w * sizeof(uint16_t));
        src.pix = snap;
        const int usx = 64 >> xd, usy = 64 >> yd, bw = 8 >> xd, bh = 8 >> yd;
        const int bwl = 3 - xd, bhl = 3 - yd;
        for (int ur = 0; ur < rows; ur++) {
            for (int uc = 0; uc < cols; uc++) {
                const int ui = ur * cols + uc;
                const int idx = p->index[ui];
                if (idx < 0) continue;
                /* unitStrength (postfilter_cdef.go:936-958) */
                uint8_t packed = plane ? p->uv_strength[idx] : p->y_strength[idx];
                int need_dir = plane == 0 ? (strength_pri(p->y_strength[idx]) != 0 ||
                                             (chroma_filtering && strength_pri(p->uv_strength[idx]) != 0))
                                          : strength_pri(p->uv_strength[idx]) != 0;
                int dir_only = plane == 0 && chroma_filtering && packed == 0 && strength_pri(p->uv_strength[idx]) != 0;
                if (packed == 0 && !dir_only) continue;
                const int ux = (uc * 64) >> xd, uy = (ur * 64) >> yd;
                const int uw = imin(usx, src.w - ux), uh = imin(usy, src.h - uy);
                if (uw <= 0 || uh <= 0) continue;
                /* block list (postfilter_cdef.go:1571-1594) */
                const int bcols = (uw + bw - 1) / bw, brows = (uh + bh - 1) / bh;
                uint64_t m = p->noskip ? p->noskip[ui] : ~0ull;
                int nblk = 0;
                uint8_t blk[64][2];
                for (int by = 0; by < brows; by++)
                    for (int bx = 0; bx < bcols; bx++)
                        if ((m >> (by * 8 + bx)) & 1) { blk[nblk][0] = (uint8_t)by; blk[nblk][1] = (uint8_t)bx; nblk++; }
                if (nblk == 0) continue;
                fill_sentinels(input, uw, uh);
                copy_input(input, &src, ux, uy, uw, uh);
                if (!dir_only && nblk < bcols * brows)
                    for (int r = 0; r < uh; r++)
                        memcpy(&unit_dst[r * BSTRIDE], &snap[(size_t)(uy + r) * src.stride + ux], (size_t)uw * sizeof(uint16_t));
                /* FrameFilterParamsFromStrength + filterFrameBlocks (filter.go:373-478) */
                const int pri = strength_pri(packed) << coeff_shift;
                const int sec = strength_sec(packed) << coeff_shift;
                const int damping = p->damping + coeff_shift - (plane ? 1 : 0);
                const uint16_t *org = input + VBORDER * BSTRIDE + HBORDER;
                if (plane == 0 && need_dir) {
                    for (int b = 0; b < nblk; b++) {
                        int by = blk[b][0], bx = blk[b][1], d;
                        int32_t v;
                        orc_cdef_find_dir(org + ((by * BSTRIDE) << bhl) + (bx << bwl), BSTRIDE, coeff_shift, &d, &v);
                        dirs[ui][by][bx] = (uint8_t)d;
                        vars[ui][by][bx] = v;
                    }
                }
                if (need_dir && plane == 1 && xd != yd) {
                    static const uint8_t c422[8] = { 7, 0, 2, 4, 5, 6, 6, 6 };
                    static const uint8_t c440[8] = { 1, 2, 2, 2, 3, 4, 6, 0 };
                    for (int b = 0; b < nblk; b++) {
                        uint8_t *dd = &dirs[ui][blk[b][0]][blk[b][1]];
                        *dd = xd ? c422[*dd] : c440[*dd];
                    }
                }
                const int luma_adjust = plane == 0 && (need_dir || pri != 0);
                for (int b = 0; b < nblk; b++) {
                    int by = blk[b][0], bx = blk[b][1];
                    int st = luma_adjust ? orc_cdef_adjust_strength(pri, vars[ui][by][bx]) : pri;
                    int dir = pri ? dirs[ui][by][bx] : 0;
                    const uint16_t *so = org + ((by * BSTRIDE) << bhl) + (bx << bwl);
                    uint16_t *dd = unit_dst + ((by << bhl) * BSTRIDE) + (bx << bwl);
                    orc_cdef_filter_block(dd, BSTRIDE, so, st, sec, dir, damping, damping, coeff_shift, bw, bh);
                }
                if (dir_only) continue;
                for (int r = 0; r < uh; r++)
                    memcpy(&f->p[plane][(size_t)(uy + r) * f->stride[plane] + ux], &unit_dst[r * BSTRIDE], (size_t)uw * sizeof(uint16_t));
            }
        }
    }
    free(snap); free(dirs); free(vars); free(input); free(unit_dst);
    return AV1PF_OK;
}
