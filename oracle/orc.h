/*
 * orc.h — CPU oracle for the AV1 post-filter path: a plain-C restatement of
 * goav1's Go implementation, used ONLY as the parity checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  It is test
 * infrastructure, never linked into the product library.
 *
 * Every function cites the goav1 file:line it restates.  Samples are held as
 * uint16 for every bit depth (the reference's 8-bit byte paths are documented
 * bit-identical to its uint16 paths, SIMD_PORT.md:29-33).
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>
#include <stddef.h>
#include "av1pf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_frame {
    int w, h;              /* visible luma size */
    int bd, ssx, ssy, mono;
    int matrix_identity;
    uint16_t *p[3];        /* planes, each at least aligned8 extent */
    int stride[3];         /* samples */
    int rows[3];           /* allocated rows */
} orc_frame;

/* LR stripe boundary lines of one plane (tile/restoration_boundary.go:19-30). */
typedef struct orc_lr_bounds {
    uint16_t *above, *below;
    int stride, rows;
} orc_lr_bounds;

/* ---------------- deblocking (loopfilter/, decoder/postfilter_loopfilter_masks*.go) */
void orc_lf_thresholds(int level, int sharpness, int *limit, int *blimit, int *hev);
void orc_lpf_unit(uint16_t *pix, ptrdiff_t off, ptrdiff_t stridea, ptrdiff_t strideb,
                  int bd, int E, int I, int H, int wd);
/* libaom per-edge formulation (loopfilter/filter{4,6,8,14}.go), one line. */
void orc_lf_edge_libaom(uint16_t *pix, ptrdiff_t off, ptrdiff_t step, int bd,
                        int width, int limit, int blimit, int hev);
int  orc_deblock(orc_frame *f, const av1pf_deblock_params *p);

/* ---------------- CDEF (cdef/, decoder/postfilter_cdef.go) */
void orc_cdef_find_dir(const uint16_t *img, int stride, int coeff_shift, int *dir, int32_t *var);
void orc_cdef_filter_block(uint16_t *dst, int dst_stride, const uint16_t *in, /* in points at block origin, stride 144 */
                           int pri, int sec, int dir, int pri_damp, int sec_damp,
                           int coeff_shift, int w, int h);
int  orc_cdef_adjust_strength(int strength, int32_t var);
int  orc_cdef(orc_frame *f, const av1pf_cdef_params *p);

/* ---------------- loop restoration (restoration/, tile/restoration_*.go) */
void orc_wiener_unit(const uint16_t *src, int src_stride, uint16_t *dst, int dst_stride,
                     int w, int h, const int8_t vt[3], const int8_t ht[3], int bd);
void orc_sgr_unit(const uint16_t *src, int src_stride, uint16_t *dst, int dst_stride,
                  int w, int h, int set, const int8_t xqd[2], int bd);
void orc_sgr_decode_xq(int set, const int8_t xqd[2], int xq[2]);
int  orc_lr_plane_geometry(const orc_frame *f, const av1pf_lr_params *p, int plane,
                           int *plane_w, int *plane_h, int *horz, int *vert);
int  orc_lr_bounds_alloc(const orc_frame *f, const av1pf_lr_params *p, orc_lr_bounds b[3]);
void orc_lr_bounds_free(orc_lr_bounds b[3]);
int  orc_lr_save(const orc_frame *f, const av1pf_lr_params *p, orc_lr_bounds b[3], int after_cdef);
int  orc_lr_apply(orc_frame *f, const av1pf_lr_params *p, const orc_lr_bounds b[3]);

/* ---------------- film grain (filmgrain/, decoder/postfilter_filmgrain*.go) */
uint16_t orc_fg_random_step(uint16_t *reg, int bits);
void orc_fg_scaling_lut(uint8_t lut[256], const uint8_t (*pts)[2], int n);
int  orc_fg_luma_grain(int16_t *grain, const av1pf_grain_params *g, int bd);
int  orc_fg_chroma_grain(int16_t *grain, const int16_t *luma, const av1pf_grain_params *g,
                         int bd, int plane, int ssx, int ssy);
int  orc_filmgrain(const orc_frame *src, orc_frame *dst, const av1pf_grain_params *g);

/* ---------------- deblocking side data from the record map
 * (threading/loop_filter_masks.go BuildFromMap, lfmask/) */
int  orc_deblock_build(const av1pf_lf_record *map, const av1pf_lf_build_params *p, uint8_t *level_cache,
                       av1pf_lf_mask *masks);
void orc_lf_block_levels(const av1pf_lf_build_params *p, const av1pf_lf_record *r, uint8_t out[4]);
const uint8_t *orc_lf_tables(int which);

/* ---------------- super-resolution (superres/superres.go, decoder/postfilter_superres.go) */
const int16_t *orc_superres_filter(int phase);
int  orc_superres_row(const uint16_t *src, int srcw, int codedw, uint16_t *dst, int dstw, int bd);
int  orc_superres(const orc_frame *coded, orc_frame *up);
int  orc_lr_save_superres(const orc_frame *coded, const orc_frame *up, const av1pf_lr_params *p, orc_lr_bounds b[3]);

/* ---------------- border extension of a refreshing frame (frame/buffer.go:293-409) */
int  orc_extend_borders(orc_frame *f);

/* ---------------- chain (decoder/postfilter.go:589-845, 947-979) */
int  orc_chain(orc_frame *f, const av1pf_request *req, orc_frame *display);
int  orc_chain_sr(orc_frame *f, const av1pf_request *req, orc_frame *up, orc_frame *display);

/* ---------------- encoder CDEF search (encoder/cdef_search.go) */
int  orc_cdef_search_mse(const orc_frame *src, const orc_frame *rec, const uint64_t *noskip,
                         int damping, uint64_t *mse);
int  orc_cdef_search_finish(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda,
                            uint8_t y_str[8], uint8_t uv_str[8], uint8_t *fb_idx);

/* encoder curated LR search, luma 8-bit (orc_lr_search.c) */
int orc_lr_search_units(int pw, int ph, int us, int *hz, int *vt);
int orc_lr_search(const orc_frame *src, const orc_frame *rec, int us, av1pf_lr_unit *out, uint64_t *sse);
/* the unit decision of the curated search from its 12 scores */
void orc_lr_decide(const uint64_t e[12], int w, int h, av1pf_lr_unit *out);
/* exhaustive superset: curated candidates scored exactly, a least-squares
 * Wiener fit and all 16 SGR sets with a least-squares projection; sse[unit][28]
 * and the fitted parameters fit[unit][38] (Wiener v[3], h[3], then xqd[2] per set) */
int orc_lr_search_exhaustive(const orc_frame *src, const orc_frame *rec, int us, av1pf_lr_unit *out, uint64_t *sse,
                             int8_t *fit);
void orc_sgr_filters(const uint16_t *src, int ss, int w, int h, int set, int bd, int32_t *f0, int32_t *f1);

#ifdef __cplusplus
}
#endif
#endif
