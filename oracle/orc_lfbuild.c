/*
 * CPU oracle: deblocking side data from the loop-filter record map.  Test
 * infrastructure only.
 *
 * Restates goav1's order-independent mask build,
 * FrameWorkLoopFilterMasks.BuildFromMap (threading/loop_filter_masks.go:452-546):
 * blocks are visited at their origin cells in raster order; each block's
 * left / above edge context is derived from the neighbouring records
 * (fillLeftLumaContext / fillAboveLumaContext / fillLeftChromaContext /
 * fillAboveChromaContext, :582-695), then the lfmask builders add its edges
 * (CreateIntra / CreateInter, lfmask/build.go:59-89, 197-230, with
 * maskEdgesIntra / maskEdgesInter / maskEdgesChroma and decompTx,
 * lfmask/mask.go:76-528).  The level cache is filled as the decode-time build
 * fills it (LevelCache.fill, build.go:232-280) with the block levels of
 * blockLevels (loop_filter_masks.go:88-123) from CalcLFValues
 * (lfmask/values.go:57-119).
 *
 * A block's luma transform layout is a grid of leaves: intra blocks tile the
 * block with tx_y; inter blocks tile it with tx_y (the largest transform) and
 * split leaves by the two split words (decompTx: depth 0 and 1, bit
 * yOff * 4 + xOff of the tree position).  The reference's unsplit shortcut of
 * maskEdgesInter (mask.go:111-257) is the same tiling without splits.
 */
#include <string.h>
#include "orc.h"

/* tile.BlockSize in dav1d order (partition.go:43-66): size in 4x4 units */
static const uint8_t kBlkW4[22] = { 32, 32, 16, 16, 16, 16, 8, 8, 8, 8, 4, 4, 4, 4, 4, 2, 2, 2, 2, 1, 1, 1 };
static const uint8_t kBlkH4[22] = { 32, 16, 32, 16, 8, 4, 16, 8, 4, 2, 16, 8, 4, 2, 1, 8, 4, 2, 1, 4, 2, 1 };
/* tile.TransformSize (tx_size.go:20-40): log2 size in 4x4 units and the
 * size one split makes (the AV1 Split_Tx_Size table) */
static const uint8_t kTxL2W[19] = { 0, 1, 2, 3, 4, 0, 1, 1, 2, 2, 3, 3, 4, 0, 2, 1, 3, 2, 4 };
static const uint8_t kTxL2H[19] = { 0, 1, 2, 3, 4, 1, 0, 2, 1, 3, 2, 4, 3, 2, 0, 3, 1, 4, 2 };
static const uint8_t kTxSub[19] = { 0, 0, 1, 2, 3, 0, 0, 1, 1, 2, 2, 3, 3, 5, 6, 7, 8, 9, 10 };

const uint8_t *orc_lf_tables(int which)
{
    switch (which) {
    case 0: return kBlkW4;
    case 1: return kBlkH4;
    case 2: return kTxL2W;
    case 3: return kTxL2H;
    case 4: return kTxSub;
    default: return NULL;
    }
}

static int imin(int a, int b) { return a < b ? a : b; }
static int iclip(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

static int rec_ok(const av1pf_lf_record *r)
{
    return (r->flags & AV1PF_LF_VALID) && r->size < 22 && r->tx_y < 19 && r->tx_uv < 19;
}

/* ---- levels: calcLFValue / calcLFValueChroma / CalcLFValues (values.go:57-119)
 * and the table lookup of blockLevels (loop_filter_masks.go:88-123) */
static int lf_value(int base_lvl, int lf_delta, int seg_delta, int ref, int mode, const av1pf_lf_build_params *p)
{
    const int base = iclip(iclip(base_lvl + lf_delta, 0, 63) + seg_delta, 0, 63);
    if (!p->mode_ref_deltas) return base;
    const int sh = base >= 32;
    if (ref == 0) return iclip(base + p->ref_delta[0] * (1 << sh), 0, 63);
    return iclip(base + (p->mode_delta[mode] + p->ref_delta[ref]) * (1 << sh), 0, 63);
}

void orc_lf_block_levels(const av1pf_lf_build_params *p, const av1pf_lf_record *r, uint8_t out[4])
{
    if (p->level_y[0] == 0 && p->level_y[1] == 0) { memset(out, 0, 4); return; }
    const int seg = p->seg_enabled ? (r->segment & 7) : 0;
    const int ref = r->ref & 7, mode = r->ref ? (r->mode & 1) : 0;
    const int d0 = r->delta_lf[0];
    const int d1 = p->delta_lf_multi ? r->delta_lf[1] : d0;
    const int d2 = p->delta_lf_multi ? r->delta_lf[2] : d0;
    const int d3 = p->delta_lf_multi ? r->delta_lf[3] : d0;
    const int16_t *sd = p->seg_enabled ? p->seg_delta[seg] : NULL;
    out[0] = (uint8_t)lf_value(p->level_y[0], d0, sd ? sd[0] : 0, ref, mode, p);
    out[1] = (uint8_t)lf_value(p->level_y[1], d1, sd ? sd[1] : 0, ref, mode, p);
    out[2] = p->level_u ? (uint8_t)lf_value(p->level_u, d2, sd ? sd[2] : 0, ref, mode, p) : 0;
    out[3] = p->level_v ? (uint8_t)lf_value(p->level_v, d3, sd ? sd[3] : 0, ref, mode, p) : 0;
}

/* ---- luma transform leaves of one block (decompTx, mask.go:76-101):
 * per visible cell the clamped width / height class and whether a leaf starts
 * at the cell's column / row */
typedef struct {
    uint8_t cw[32][32], ch[32][32];   /* min(log2 size, 2) */
    uint8_t sx[32][32], sy[32][32];   /* leaf starts at this column / row */
} leaves;

static void put_leaf(leaves *L, int y0, int x0, int tx, int bw4, int bh4)
{
    const int w = 1 << kTxL2W[tx], h = 1 << kTxL2H[tx];
    for (int y = y0; y < y0 + h && y < bh4; y++)
        for (int x = x0; x < x0 + w && x < bw4; x++) {
            L->cw[y][x] = (uint8_t)imin(kTxL2W[tx], 2);
            L->ch[y][x] = (uint8_t)imin(kTxL2H[tx], 2);
            L->sx[y][x] = x == x0;
            L->sy[y][x] = y == y0;
        }
}

static void decomp(leaves *L, int y0, int x0, int tx, int depth, int yoff, int xoff, const uint16_t split[2],
                   int bw4, int bh4)
{
    int is_split = 0;
    if (tx != 0 && depth <= 1) is_split = (split[depth & 1] >> (yoff * 4 + xoff)) & 1;
    if (!is_split) { put_leaf(L, y0, x0, tx, bw4, bh4); return; }
    const int sub = kTxSub[tx];
    const int w4 = 1 << kTxL2W[tx], h4 = 1 << kTxL2H[tx];
    decomp(L, y0, x0, sub, depth + 1, yoff * 2, xoff * 2, split, bw4, bh4);
    if (w4 >= h4) decomp(L, y0, x0 + w4 / 2, sub, depth + 1, yoff * 2, xoff * 2 + 1, split, bw4, bh4);
    if (h4 >= w4) {
        decomp(L, y0 + h4 / 2, x0, sub, depth + 1, yoff * 2 + 1, xoff * 2, split, bw4, bh4);
        if (w4 >= h4) decomp(L, y0 + h4 / 2, x0 + w4 / 2, sub, depth + 1, yoff * 2 + 1, xoff * 2 + 1, split, bw4, bh4);
    }
}

/* leaves of the visible part (bw4 x bh4) of a block */
static void block_leaves(leaves *L, const av1pf_lf_record *r, int bw4, int bh4)
{
    const int tx = r->tx_y;
    const int tw = 1 << kTxL2W[tx], th = 1 << kTxL2H[tx];
    const int inter = !(r->flags & AV1PF_LF_INTRA);
    int yoff = 0;
    for (int y = 0; y < bh4; y += th, yoff++) {
        int xoff = 0;
        for (int x = 0; x < bw4; x += tw, xoff++) {
            if (inter) decomp(L, y, x, tx, 0, yoff, xoff, r->split, bw4, bh4);
            else put_leaf(L, y, x, tx, bw4, bh4);
        }
    }
}

static const av1pf_lf_record *rec_at(const av1pf_lf_record *m, int cols, int x, int y)
{
    return &m[(size_t)y * cols + x];
}

/* right-edge / bottom-edge context of the block owning cell (x, y)
 * (neighborRightLuma, InterEdgeContext / IntraEdgeContext, build.go:97-191):
 * the class of the leaf that covers the cell; -1 when the cell has no block */
static int nb_class(const av1pf_lf_record *m, int cols, int rows, int x, int y, int vertical)
{
    const av1pf_lf_record *r = rec_at(m, cols, x, y);
    if (!rec_ok(r)) return -1;
    const int bw4 = imin(cols - r->mi_col, kBlkW4[r->size]), bh4 = imin(rows - r->mi_row, kBlkH4[r->size]);
    const int lx = x - r->mi_col, ly = y - r->mi_row;
    if (bw4 <= 0 || bh4 <= 0 || lx < 0 || ly < 0 || lx >= bw4 || ly >= bh4) return -1;
    leaves L;
    memset(&L, 0, sizeof(L));
    block_leaves(&L, r, bw4, bh4);
    return vertical ? L.cw[ly][lx] : L.ch[ly][lx];
}

static void ybit(av1pf_lf_mask *m, int dir, int pos, int cls, int along)
{
    m->y[dir][pos][cls][along >> 4] |= (uint16_t)(1u << (along & 15));
}

static void uvbit(av1pf_lf_mask *m, int dir, int pos, int cls, int along, int half)
{
    const int h = along >= half;
    m->uv[dir][pos][cls][h] |= (uint16_t)(1u << (along - h * half));
}

static int uv_class(int tx, int vertical) { return (vertical ? kTxL2W[tx] : kTxL2H[tx]) != 0; }

int orc_deblock_build(const av1pf_lf_record *map, const av1pf_lf_build_params *p, uint8_t *level_cache,
                      av1pf_lf_mask *masks)
{
    const int cols = p->mi_cols, rows = p->mi_rows;
    if (cols <= 0 || rows <= 0 || (cols & 1) || (rows & 1)) return AV1PF_ERR_INVALID_ARG;
    const int ssx = p->ss_x, ssy = p->ss_y, chroma = !p->mono;
    const int sbw = (cols + 31) >> 5, sbh = (rows + 31) >> 5;
    memset(masks, 0, sizeof(av1pf_lf_mask) * (size_t)sbw * sbh);
    leaves L;
    for (int by = 0; by < rows; by++)
        for (int bx = 0; bx < cols; bx++) {
            const av1pf_lf_record *r = rec_at(map, cols, bx, by);
            if (!rec_ok(r) || r->mi_col != bx || r->mi_row != by) continue;
            const int W4 = kBlkW4[r->size], H4 = kBlkH4[r->size];
            const int bw4 = imin(cols - bx, W4), bh4 = imin(rows - by, H4);
            if (bw4 <= 0 || bh4 <= 0) continue;
            const int intra = (r->flags & AV1PF_LF_INTRA) != 0;
            const int no_inner = !intra && (r->flags & AV1PF_LF_SKIP);
            av1pf_lf_mask *m = &masks[(by >> 5) * sbw + (bx >> 5)];
            const int bx4 = bx & 31, by4 = by & 31;
            uint8_t lv[4];
            orc_lf_block_levels(p, r, lv);
            /* luma levels (LevelCache.fill, components 0 and 1) */
            for (int y = 0; y < bh4; y++)
                for (int x = 0; x < bw4; x++) {
                    uint8_t *c = level_cache + 4 * ((size_t)(by + y) * cols + bx + x);
                    c[0] = lv[0];
                    c[1] = lv[1];
                }
            /* luma contexts (fillLeftLumaContext / fillAboveLumaContext) */
            int ly[32], ay[32];
            for (int y = 0; y < bh4; y++) {
                ly[y] = 2;
                if (bx > 0) { int c = nb_class(map, cols, rows, bx - 1, by + y, 1); if (c >= 0) ly[y] = c; }
            }
            for (int x = 0; x < bw4; x++) {
                ay[x] = 2;
                if (by > 0) { int c = nb_class(map, cols, rows, bx + x, by - 1, 0); if (c >= 0) ay[x] = c; }
            }
            memset(&L, 0, sizeof(L));
            block_leaves(&L, r, bw4, bh4);
            /* block edges: class min(own leaf, neighbour context) */
            for (int y = 0; y < bh4; y++) ybit(m, 0, bx4, imin(L.cw[y][0], ly[y]), by4 + y);
            for (int x = 0; x < bw4; x++) ybit(m, 1, by4, imin(L.ch[0][x], ay[x]), bx4 + x);
            /* inner transform edges: class min of the two leaves */
            if (!no_inner) {
                for (int y = 0; y < bh4; y++)
                    for (int x = 1; x < bw4; x++)
                        if (L.sx[y][x]) ybit(m, 0, bx4 + x, imin(L.cw[y][x], L.cw[y][x - 1]), by4 + y);
                for (int x = 0; x < bw4; x++)
                    for (int y = 1; y < bh4; y++)
                        if (L.sy[y][x]) ybit(m, 1, by4 + y, imin(L.ch[y][x], L.ch[y - 1][x]), bx4 + x);
            }
            /* chroma: only the block carrying the shared chroma (dav1d has_chroma) */
            if (!chroma || !(r->flags & AV1PF_LF_HAS_UV)) continue;
            if (!((W4 > ssx || (bx & 1)) && (H4 > ssy || (by & 1)))) continue;
            const int cbw4 = imin(((cols + ssx) >> ssx) - (bx >> ssx), (W4 + ssx) >> ssx);
            const int cbh4 = imin(((rows + ssy) >> ssy) - (by >> ssy), (H4 + ssy) >> ssy);
            if (cbw4 <= 0 || cbh4 <= 0) continue;
            const int cbx = bx >> ssx, cby = by >> ssy;
            for (int y = 0; y < cbh4; y++)
                for (int x = 0; x < cbw4; x++) {
                    uint8_t *c = level_cache + 4 * ((size_t)(cby + y) * cols + cbx + x);
                    c[2] = lv[2];
                    c[3] = lv[3];
                }
            int luv[32], auv[32];
            for (int y = 0; y < cbh4; y++) {
                luv[y] = 1;
                if (bx > 0) {
                    const int lrow = imin(((cby + y) << ssy) + ssy, rows - 1);
                    const av1pf_lf_record *n = rec_at(map, cols, bx - 1, lrow);
                    if (rec_ok(n) && (n->flags & AV1PF_LF_HAS_UV)) luv[y] = uv_class(n->tx_uv, 1);
                }
            }
            for (int x = 0; x < cbw4; x++) {
                auv[x] = 1;
                if (by > 0) {
                    const int lcol = imin(((cbx + x) << ssx) + ssx, cols - 1);
                    const av1pf_lf_record *n = rec_at(map, cols, lcol, by - 1);
                    if (rec_ok(n) && (n->flags & AV1PF_LF_HAS_UV)) auv[x] = uv_class(n->tx_uv, 0);
                }
            }
            const int tw = uv_class(r->tx_uv, 1), th = uv_class(r->tx_uv, 0);
            const int cbx4 = bx4 >> ssx, cby4 = by4 >> ssy;
            const int vhalf = 16 >> ssy, hhalf = 16 >> ssx;
            for (int y = 0; y < cbh4; y++) uvbit(m, 0, cbx4, imin(tw, luv[y]), cby4 + y, vhalf);
            for (int x = 0; x < cbw4; x++) uvbit(m, 1, cby4, imin(th, auv[x]), cbx4 + x, hhalf);
            if (!no_inner) {
                const int hstep = 1 << kTxL2W[r->tx_uv], vstep = 1 << kTxL2H[r->tx_uv];
                for (int x = hstep; x < cbw4; x += hstep)
                    for (int y = 0; y < cbh4; y++) uvbit(m, 0, cbx4 + x, tw, cby4 + y, vhalf);
                for (int y = vstep; y < cbh4; y += vstep)
                    for (int x = 0; x < cbw4; x++) uvbit(m, 1, cby4 + y, th, cbx4 + x, hhalf);
            }
        }
    return AV1PF_OK;
}
