/*
 * CPU oracle: the encoder's CDEF strength search at SVT search level 1.
 * Test infrastructure only (see orc.h).
 *
 * Restates:
 *   encoder/cdef_search.go:48,72-110    level-1 candidate list: 16 primaries (secondary 0)
 *                                       then pf+1..pf+3 per primary; chroma tests the same list
 *   encoder/cdef_search.go:410-450      computeCDEFList (8x8 blocks with a non-skip MI, raster)
 *   encoder/cdef_search.go:453-477      computeCDEFDist8 (SSE, row step = subsampling)
 *   encoder/cdef_search.go:479-505      loadCDEFInput (VeryLarge outside the visible plane)
 *   encoder/cdef_search.go:580-634      damping = 3 + (q >> 6); plane sizes (W+1)>>1 for 4:2:0
 *   encoder/cdef_search.go:649-749      per filter block: luma directions from the recon, each
 *                                       trial filtered (unfiltered blocks keep the recon), chroma
 *                                       U+V summed into mse[1]
 *   cdef/filter_search.go:20-74         FilterBlocksU8SubsampledTrusted (luma adjust only when the
 *                                       trial primary is nonzero; chroma takes luma directions)
 *   encoder/cdef_search.go:780-945      searchOneDual / jointStrengthSearchDual / finish
 * 4:2:0 frames whose size is a multiple of 8 are searched.  The reference
 * encoder searches 8-bit frames only; 10-bit frames are an extension with no
 * reference counterpart: the trials are the decoder's high-bit-depth CDEF
 * (strengths and damping shifted by bd - 8, directions from samples shifted
 * down by bd - 8, the filter tap set chosen by the unshifted primary), which
 * orc_cdef.c restates from postfilter_cdef.go, so every table entry is the
 * SSE of the decoder CDEF at that strength.
 */
#include "orc.h"
#include <stdlib.h>
#include <string.h>

#define BSTRIDE 144
#define VERY_LARGE 0x4000

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* gi -> packed strength (pri * 4 + secondary code) for level 1. */
static int level1_packed(int gi) { return gi < 16 ? 4 * gi : 4 * ((gi - 16) / 3) + (gi - 16) % 3 + 1; }

static void load_input(uint16_t *in, const uint16_t *plane, int stride, int pw, int ph, int x0, int y0, int w, int h)
{
    for (int r = -2; r < h + 2; r++)
        for (int c = -8; c < w + 8; c++) {
            int y = y0 + r, x = x0 + c;
            in[(r + 2) * BSTRIDE + c + 8] = (y >= 0 && y < ph && x >= 0 && x < pw) ? plane[(size_t)y * stride + x] : VERY_LARGE;
        }
}

int orc_cdef_search_mse(const orc_frame *src, const orc_frame *rec, const uint64_t *noskip,
                        int damping, uint64_t *mse)
{
    if (!src || !rec || !mse || (rec->bd != 8 && rec->bd != 10) || src->bd != rec->bd || rec->ssx != 1 ||
        rec->ssy != 1 || rec->mono || (rec->w & 7) || (rec->h & 7))
        return AV1PF_ERR_INVALID_ARG;
    const int cs = rec->bd - 8;
    const int mi_cols = ((rec->w + 7) >> 3) << 1, mi_rows = ((rec->h + 7) >> 3) << 1;
    const int nhfb = (mi_cols + 15) / 16, nvfb = (mi_rows + 15) / 16, nfb = nhfb * nvfb;
    uint16_t *in = malloc(BSTRIDE * 132 * sizeof(uint16_t));
    uint16_t *tmp = malloc(64 * 64 * sizeof(uint16_t));
    const int pw[3] = { rec->w, (rec->w + 1) >> 1, (rec->w + 1) >> 1 };
    const int ph[3] = { rec->h, (rec->h + 1) >> 1, (rec->h + 1) >> 1 };
    for (int fbr = 0; fbr < nvfb; fbr++) {
        for (int fbc = 0; fbc < nhfb; fbc++) {
            const int fb = fbr * nhfb + fbc;
            const int lr = 16 * fbr, lc = 16 * fbc;
            const int nhb = imin(16, mi_cols - lc), nvb = imin(16, mi_rows - lr);
            uint64_t m = noskip ? noskip[fb] : ~0ull;
            uint8_t blk[64][2];
            int nblk = 0;
            for (int r = 0; r < nvb; r += 2)
                for (int c = 0; c < nhb; c += 2)
                    if ((m >> ((r >> 1) * 8 + (c >> 1))) & 1) { blk[nblk][0] = (uint8_t)(r >> 1); blk[nblk][1] = (uint8_t)(c >> 1); nblk++; }
            uint64_t *m0 = mse + (size_t)fb * 64, *m1 = mse + ((size_t)nfb + fb) * 64;
            memset(m0, 0, 64 * sizeof(uint64_t));
            memset(m1, 0, 64 * sizeof(uint64_t));
            if (nblk == 0) continue;
            uint8_t dirs[8][8];
            int32_t vars[8][8];
            for (int pli = 0; pli < 3; pli++) {
                const int dec = pli ? 1 : 0;
                const int x0 = (lc * 4) >> dec, y0 = (lr * 4) >> dec;
                const int w = (nhb * 4) >> dec, h = (nvb * 4) >> dec;
                const int bl = 3 - dec, bs = 1 << bl;
                const uint16_t *rp = rec->p[pli];
                const uint16_t *sp = src->p[pli];
                const int rs = rec->stride[pli], ss = src->stride[pli];
                load_input(in, rp, rs, pw[pli], ph[pli], x0, y0, w, h);
                if (pli == 0)
                    for (int b = 0; b < nblk; b++) {
                        int d;
                        int32_t v;
                        orc_cdef_find_dir(rp + (size_t)(y0 + (blk[b][0] << 3)) * rs + x0 + (blk[b][1] << 3), rs, cs, &d, &v);
                        dirs[blk[b][0]][blk[b][1]] = (uint8_t)d;
                        vars[blk[b][0]][blk[b][1]] = v;
                    }
                const int pd = damping + cs - (pli ? 1 : 0);
                for (int gi = 0; gi < 64; gi++) {
                    int packed = level1_packed(gi);
                    int pri = (packed >> 2) << cs, sec = packed & 3;
                    if (sec == 3) sec = 4;
                    sec <<= cs;
                    uint64_t sum = 0;
                    for (int b = 0; b < nblk; b++) {
                        int by = blk[b][0], bx = blk[b][1];
                        int st = (pli == 0 && pri) ? orc_cdef_adjust_strength(pri, vars[by][bx]) : pri;
                        const uint16_t *org = in + (2 + (by << bl)) * BSTRIDE + 8 + (bx << bl);
                        if (st == 0 && sec == 0) {
                            for (int r = 0; r < bs; r++)
                                for (int c = 0; c < bs; c++) tmp[r * bs + c] = org[r * BSTRIDE + c];
                        } else {
                            orc_cdef_filter_block(tmp, bs, org, st, sec, pri ? dirs[by][bx] : 0, pd, pd, cs, bs, bs);
                        }
                        for (int r = 0; r < bs; r++)
                            for (int c = 0; c < bs; c++) {
                                int e = (int)sp[(size_t)(y0 + (by << bl) + r) * ss + x0 + (bx << bl) + c] - tmp[r * bs + c];
                                sum += (uint64_t)(e * e);
                            }
                    }
                    if (pli == 0) m0[gi] = sum;
                    else m1[gi] += sum;
                }
            }
        }
    }
    free(in); free(tmp);
    (void)imax;
    return AV1PF_OK;
}

/* ---- finish_cdef_search (cdef_search.go:780-945) ------------------------ */
static uint64_t search_one_dual(const uint64_t *mse, int nfb, const int *fbs, int nsb,
                                int *lev0, int *lev1, int nb, int endGI, uint64_t (*tot)[64])
{
    if (endGI <= 0) { lev0[nb] = lev1[nb] = 0; return 1ull << 63; }
    if (nsb == 0) { lev0[nb] = lev1[nb] = 0; return 0; }
    if (nb == 0) {
        uint64_t s0[64] = { 0 }, s1[64] = { 0 };
        for (int i = 0; i < nsb; i++) {
            const uint64_t *a = mse + (size_t)fbs[i] * 64, *b = mse + ((size_t)nfb + fbs[i]) * 64;
            for (int j = 0; j < endGI; j++) { s0[j] += a[j]; s1[j] += b[j]; }
        }
        int b0 = 0, b1 = 0;
        for (int j = 1; j < endGI; j++) { if (s0[j] < s0[b0]) b0 = j; if (s1[j] < s1[b1]) b1 = j; }
        lev0[0] = b0; lev1[0] = b1;
        return s0[b0] + s1[b1];
    }
    for (int j = 0; j < endGI; j++) memset(tot[j], 0, sizeof(tot[j]));
    uint64_t sum_best = 0;
    for (int i = 0; i < nsb; i++) {
        const uint64_t *a = mse + (size_t)fbs[i] * 64, *b = mse + ((size_t)nfb + fbs[i]) * 64;
        uint64_t best = a[lev0[0]] + b[lev1[0]];
        for (int g = 1; g < nb; g++) { uint64_t c = a[lev0[g]] + b[lev1[g]]; if (c < best) best = c; }
        sum_best += best;
        for (int j = 0; j < endGI; j++) {
            if (a[j] >= best) continue;
            uint64_t rem = best - a[j];
            for (int k = 0; k < endGI; k++) if (b[k] < rem) tot[j][k] += rem - b[k];
        }
    }
    uint64_t bs = tot[0][0];
    int b0 = 0, b1 = 0;
    for (int j = 0; j < endGI; j++)
        for (int k = 0; k < endGI; k++)
            if (tot[j][k] > bs) { bs = tot[j][k]; b0 = j; b1 = k; }
    lev0[nb] = b0; lev1[nb] = b1;
    return sum_best - bs;
}

/* The finish with an optional trace of every searchOneDual step: per step
 * (75 of them: levels 0..3, nb greedy steps then 4 nb refinement steps
 * each) the level, the slot filled, the chosen pair (search-level-1
 * positions) and the returned total; per level the RD cost.  Test
 * infrastructure (the steps are checked against brute-force optima). */
typedef struct orc_finish_trace {
    int32_t step[75][4];   /* level, slot, lev0, lev1 */
    uint64_t total[75];
    uint64_t cost[4];
    int32_t nsteps;
} orc_finish_trace;

static int finish_impl(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda, uint8_t y_str[8],
                       uint8_t uv_str[8], uint8_t *fb_idx, orc_finish_trace *tr);

int orc_cdef_search_finish_trace(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda,
                                 uint8_t y_str[8], uint8_t uv_str[8], uint8_t *fb_idx, orc_finish_trace *tr)
{
    memset(tr, 0, sizeof(*tr));
    return finish_impl(mse, nfb, skip_fb, lambda, y_str, uv_str, fb_idx, tr);
}

/* Returns bits; writes y/uv packed strengths (8 each) and per-FB index. */
int orc_cdef_search_finish(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda,
                           uint8_t y_str[8], uint8_t uv_str[8], uint8_t *fb_idx)
{
    return finish_impl(mse, nfb, skip_fb, lambda, y_str, uv_str, fb_idx, NULL);
}

static void trace_step(orc_finish_trace *tr, int level, int slot, const int *lev0, const int *lev1, uint64_t t)
{
    if (!tr || tr->nsteps >= 75) return;
    int32_t *e = tr->step[tr->nsteps];
    e[0] = level; e[1] = slot; e[2] = lev0[slot]; e[3] = lev1[slot];
    tr->total[tr->nsteps++] = t;
}

static int finish_impl(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda, uint8_t y_str[8],
                       uint8_t uv_str[8], uint8_t *fb_idx, orc_finish_trace *tr)
{
    int *fbs = malloc(sizeof(int) * (nfb > 0 ? nfb : 1));
    uint64_t (*tot)[64] = malloc(sizeof(uint64_t) * 64 * 64);
    int nsb = 0;
    for (int fb = 0; fb < nfb; fb++) { fb_idx[fb] = 0; if (!skip_fb[fb]) fbs[nsb++] = fb; }
    const int endGI = 64;
    uint64_t best_tot = 1ull << 63;
    int bits = 0;
    int ys[8] = { 0 }, us[8] = { 0 };
    for (int i = 0; i <= 3; i++) {
        int lev0[8] = { 0 }, lev1[8] = { 0 };
        int nb = 1 << i;
        uint64_t t = 1ull << 63;
        for (int k = 0; k < nb; k++) {
            t = search_one_dual(mse, nfb, fbs, nsb, lev0, lev1, k, endGI, tot);
            trace_step(tr, i, k, lev0, lev1, t);
        }
        for (int k = 0; k < 4 * nb; k++) {
            for (int j = 0; j < nb - 1; j++) { lev0[j] = lev0[j + 1]; lev1[j] = lev1[j + 1]; }
            t = search_one_dual(mse, nfb, fbs, nsb, lev0, lev1, nb - 1, endGI, tot);
            trace_step(tr, i, nb - 1, lev0, lev1, t);
        }
        int total_bits = nsb * i + nb * 6 * 2;
        int64_t rate = (int64_t)total_bits << 9;
        uint64_t cost = (uint64_t)((rate * (int64_t)lambda + (1 << 8)) >> 9) + ((t * 16) << 7);
        if (tr) tr->cost[i] = cost;
        if (cost < best_tot) {
            best_tot = cost; bits = i;
            for (int j = 0; j < nb; j++) { ys[j] = lev0[j]; us[j] = lev1[j]; }
        }
    }
    int nb = 1 << bits;
    for (int i = nb; i < 8; i++) ys[i] = us[i] = 0;
    for (int i = 0; i < nsb; i++) {
        int bg = 0;
        uint64_t bm = 1ull << 63;
        const uint64_t *a = mse + (size_t)fbs[i] * 64, *b = mse + ((size_t)nfb + fbs[i]) * 64;
        for (int g = 0; g < nb; g++) { uint64_t c = a[ys[g]] + b[us[g]]; if (c < bm) { bg = g; bm = c; } }
        fb_idx[fbs[i]] = (uint8_t)bg;
    }
    for (int i = 0; i < 8; i++) {
        y_str[i] = i < nb ? (uint8_t)level1_packed(ys[i]) : 0;
        uv_str[i] = i < nb ? (uint8_t)level1_packed(us[i]) : 0;
    }
    free(fbs); free(tot);
    return bits;
}
