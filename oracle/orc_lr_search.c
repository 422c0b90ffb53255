/*
 * CPU oracle: the encoder's curated loop-restoration search (luma, 8-bit).
 * Test infrastructure only (see orc.h).
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
#include <stdlib.h>
#include <string.h>

static const int8_t kWienerCand[6][3] = { { 0, -3, 12 }, { 0, -5, 18 }, { 0, -2, 8 },
                                          { 0, 2, 10 },  { 1, 4, 14 },  { 0, -1, 4 } };
static const int kSgrPreset[4] = { 0, 2, 10, 14 };
static const int8_t kSgrXqd[4][2] = { { 8, 24 }, { 16, 40 }, { 0, 48 }, { 24, 72 } };

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
    if (!src || !rec || rec->bd != 8 || src->w != rec->w || src->h != rec->h || us < 32 || us > 256)
        return AV1PF_ERR_INVALID_ARG;
    const int W = rec->w, H = rec->h, pad = 3;
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
            uint64_t best = base;
            int wbest = -1;
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
                        pred = pred < 0 ? 0 : (pred > 255 ? 255 : pred);
                        const int d = (int)S[(size_t)y * ss + x] - pred;
                        s += (uint64_t)(d * d);
                    }
                e[1 + c] = s;
                if (s + (uint64_t)(w * h) < best) { best = s; wbest = c; }
            }
            const int valid = w > 0 && h > 0 && w <= 64 && h <= 64;
            const uint16_t *org = bord + (size_t)(y0 + pad) * bs + x0 + pad;
            uint64_t best_sse = UINT64_MAX;
            if (wbest >= 0 && valid) {
                orc_wiener_unit(org, bs, tmp, 64, w, h, kWienerCand[wbest], kWienerCand[wbest], 8);
                uint64_t s = 0;
                for (int y = 0; y < h; y++)
                    for (int x = 0; x < w; x++) {
                        const int d = (int)S[(size_t)(y0 + y) * ss + x0 + x] - (int)tmp[y * 64 + x];
                        s += (uint64_t)(d * d);
                    }
                e[7] = s;
                best_sse = s;
                out[u].type = AV1PF_LR_WIENER;
                memcpy(out[u].wiener_v, kWienerCand[wbest], 3);
                memcpy(out[u].wiener_h, kWienerCand[wbest], 3);
            }
            if (!valid) continue;
            uint64_t sbest = base;
            int pbest = -1;
            for (int p = 0; p < 4; p++) {
                orc_sgr_unit(org, bs, tmp, 64, w, h, kSgrPreset[p], kSgrXqd[p], 8);
                uint64_t s = 0;
                for (int y = 0; y < h; y++)
                    for (int x = 0; x < w; x++) {
                        const int d = (int)S[(size_t)(y0 + y) * ss + x0 + x] - (int)tmp[y * 64 + x];
                        s += (uint64_t)(d * d);
                    }
                e[8 + p] = s;
                if (s + (uint64_t)(w * h) < sbest) { sbest = s; pbest = p; }
            }
            if (pbest >= 0 && sbest < best_sse) {
                memset(&out[u], 0, sizeof(out[u]));
                out[u].type = AV1PF_LR_SGRPROJ;
                out[u].sgr_set = (uint8_t)kSgrPreset[pbest];
                out[u].sgr_xqd[0] = kSgrXqd[pbest][0];
                out[u].sgr_xqd[1] = kSgrXqd[pbest][1];
            }
        }
    }
    free(bord);
    free(tmp);
    return AV1PF_OK;
}
