/*
 * av1pf.h — C ABI of the MI355X (gfx950) AV1 post-filter library.
 *
 * This is the drop-in boundary that replaces goav1's frame post-filter runner
 * and the encoder's CDEF / loop-restoration candidate search.  Every entry
 * point takes plain pointers and sizes; no C++ or framework types cross it.
 *
 * Reference interfaces replaced (paths relative to the goav1 tree):
 *   - FrameWorkPostFilterRunner.Apply(FrameWorkPostFilterContext) error
 *       internal/av1/decoder/work.go:153-155, called once per completed frame
 *       from runFrameWorkPostFilterRunnerWithCompact (decoder/work.go:1912-1934)
 *       through the reusable runner in decoder_postfilter_bind.go:398-431.
 *   - The stage chain it runs, applySupportedPostFiltersInPlace
 *       internal/av1/decoder/postfilter.go:589-782 (deblock -> LR boundary save
 *       -> CDEF -> LR boundary save -> loop restoration -> film grain on a
 *       display copy, postfilter.go:810-845).
 *   - cdefSearcher.search  internal/av1/encoder/cdef_search.go:580-634
 *     (per-64x64 MSE table over the 64 strength pairs of search level 1).
 *
 * Memory model.  All pointers handed to av1pf_apply / av1pf_apply_batch /
 * av1pf_cdef_search are DEVICE pointers on the library's device (frames and
 * side data resident in HBM).  Calls are asynchronous on the given HIP stream
 * (NULL = the null stream); av1pf_sync() or a stream synchronisation fences
 * them, mirroring the reference contract that Apply returns only after Output
 * is written (decoder_postfilter_bind.go:414-431).
 *
 * Sample layout (frame/buffer.go:18-139): 8-bit planes are bytes; 10/12-bit
 * planes are little-endian uint16.  Each plane must hold the MI-aligned extent
 * ((W+7)&~7) >> ss_x samples by ((H+7)&~7) >> ss_y rows, because deblocking and
 * CDEF read and write that padding exactly as the reference does
 * (postfilter_cdef.go:1390-1412, postfilter_loopfilter.go:945-970).
 *
 * Errors are validated on the host before any device write
 * (postfilter.go:598-616); a nonzero status means nothing was launched.
 */
#ifndef AV1PF_H
#define AV1PF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define AV1PF_OK                 0
#define AV1PF_ERR_INVALID_ARG   -1  /* threading.ErrInvalidBatch / frame.ErrInvalidFormat */
#define AV1PF_ERR_UNSUPPORTED   -2  /* decoder.ErrUnsupportedPostFilter                   */
#define AV1PF_ERR_DEVICE        -3  /* HIP runtime failure                                */
#define AV1PF_ERR_SHORT_BUFFER  -4  /* frame.ErrShortBuffer                               */
#define AV1PF_ERR_NO_DEVICE     -5  /* no usable gfx950 device                            */

/* ---- stage bits (FrameWorkPostFilter* enum, postfilter.go:13-21) -------- */
#define AV1PF_STAGE_DEBLOCK    0x01u
#define AV1PF_STAGE_CDEF       0x02u
#define AV1PF_STAGE_SUPERRES   0x04u  /* needs av1pf_request.upscaled */
#define AV1PF_STAGE_LR         0x08u
#define AV1PF_STAGE_FILMGRAIN  0x10u

/* ---- frame descriptor (frame.Frame / frame.Layout, frame/buffer.go) ----- */
typedef struct av1pf_frame {
    int32_t width;            /* visible luma width  (FrameSize.CodedWidth == UpscaledWidth) */
    int32_t height;           /* visible luma height (FrameSize.Height) */
    int32_t bit_depth;        /* 8, 10 or 12 */
    int32_t ss_x, ss_y;       /* chroma subsampling shifts (0 or 1) */
    int32_t mono;             /* 1 = luma only */
    int32_t matrix_identity;  /* ColorConfig.MatrixCoefficients == identity (film grain clip) */
    /* Luma rows allocated (len(Y.Pix) / Y.Stride, frame.RequiredSize,
     * frame/buffer.go:141-236: the superblock-aligned height); chroma planes
     * hold alloc_height >> ss_y rows.  0 = the MI-aligned height (H+7)&~7.
     * Every plane's allocated width is stride / bytes per sample.  Only the
     * border extension of a refreshing frame (AV1PF_REQ_REFRESH) writes past
     * the MI-aligned extent.                                                */
    int32_t alloc_height;
    void   *plane[3];         /* Y, U, V sample pointers */
    int64_t stride[3];        /* row pitch in BYTES */
} av1pf_frame;

/* ---- deblocking side data ------------------------------------------------
 * One edge bitmask record per 128x128 region in lfmask.FilterMask layout
 * (internal/av1/lfmask/mask.go:38-41): y[dir][pos][class][half] with dir 0 =
 * vertical edges, pos = 4x4 column (dir 0) or row (dir 1) inside the region,
 * class 0/1/2 = filter width 4/8/16 (14-tap), half = low / high 16 bits;
 * uv[dir][pos][class][half] with class 0/1 = width 4/6 and the half split at
 * 16 >> ss (mask.go:406-470).  Masks must follow AV1's transform-spacing rule
 * (filter widths never reach the next edge of the same direction), which every
 * decoder-built mask does; edges of one direction are then independent.
 * level_cache holds four level bytes per 4x4 MI cell (Y vertical, Y
 * horizontal, U, V; lfmask/build.go:29-33); chroma levels sit at chroma MI
 * coordinates with the luma row stride mi_cols.                              */
typedef struct av1pf_lf_mask {
    uint16_t y[2][32][3][2];
    uint16_t uv[2][32][2][2];
} av1pf_lf_mask;

typedef struct av1pf_deblock_params {
    uint8_t  level_y[2];      /* loop_filter_level[0], [1]: both 0 disables chroma too */
    uint8_t  level_u, level_v;/* frame chroma levels: 0 skips that plane */
    uint8_t  sharpness;       /* 0..7 */
    uint8_t  pad[3];
    int32_t  mi_cols, mi_rows;/* ((W+7)>>3)<<1, ((H+7)>>3)<<1 */
    int32_t  sb128_w, sb128_h;/* (mi_cols+31)>>5, (mi_rows+31)>>5 */
    const uint8_t       *level_cache; /* 4*mi_cols*mi_rows bytes */
    const av1pf_lf_mask *masks;       /* sb128_w*sb128_h records, 4-byte aligned (read as 32-bit words) */
} av1pf_deblock_params;

/* ---- CDEF side data (parser/cdef.go:14-21, threading/cdef_index.go:13-18) */
typedef struct av1pf_cdef_params {
    uint8_t  damping;          /* cdef_damping (3..6) */
    uint8_t  strength_count;   /* 1 << cdef_bits (1..8) */
    uint8_t  y_strength[8];    /* packed primary<<2 | secondary code */
    uint8_t  uv_strength[8];
    uint8_t  pad[6];
    int32_t  unit_cols, unit_rows;  /* ceil(W/64), ceil(H/64) */
    const int8_t   *index;    /* per 64x64 unit: strength index, -1 = not read */
    const uint64_t *noskip;   /* per unit: bit by*8+bx set when the 8x8 block has a
                                 non-skip MI cell (dav1d noskip_mask); NULL = all */
} av1pf_cdef_params;

/* ---- loop restoration side data (parser/restoration.go:15-34,
 *      tile/restoration_schedule.go:40-48) --------------------------------- */
#define AV1PF_LR_NONE        0
#define AV1PF_LR_WIENER      1
#define AV1PF_LR_SGRPROJ     2
#define AV1PF_LR_SWITCHABLE  3   /* frame type only */

typedef struct av1pf_lr_unit {
    uint8_t type;           /* AV1PF_LR_NONE / WIENER / SGRPROJ */
    uint8_t sgr_set;        /* SGRProj parameter set 0..15 */
    int8_t  sgr_xqd[2];     /* coded xqd (restoration.DecodeSGRXQ input) */
    int8_t  wiener_v[3];    /* vertical taps 0..2 (tap 3 = -2*sum, mirrored) */
    int8_t  wiener_h[3];    /* horizontal taps 0..2 */
    uint8_t pad[6];
} av1pf_lr_unit;

typedef struct av1pf_lr_params {
    uint8_t  frame_type[3];   /* per plane AV1PF_LR_* */
    uint8_t  pad0;
    uint16_t unit_size[3];    /* per plane unit size (32..256) */
    uint16_t pad1;
    const av1pf_lr_unit *units[3]; /* per plane horz_units*vert_units records, row-major */
} av1pf_lr_params;

/* ---- film grain parameters (parser/film_grain.go:21-56) ----------------- */
typedef struct av1pf_grain_params {
    uint16_t seed;
    uint8_t  num_y_points, num_cb_points, num_cr_points;
    uint8_t  chroma_scaling_from_luma;
    uint8_t  scaling_shift;        /* 8..11 */
    uint8_t  ar_coeff_lag;         /* 0..3 */
    uint8_t  ar_coeff_shift;       /* 6..9 */
    uint8_t  grain_scale_shift;    /* 0..3 */
    uint8_t  overlap;
    uint8_t  clip_to_restricted_range;
    uint8_t  y_points[14][2];      /* (value, scaling) */
    uint8_t  cb_points[10][2];
    uint8_t  cr_points[10][2];
    int8_t   ar_coeffs_y[24];
    int8_t   ar_coeffs_cb[25];
    int8_t   ar_coeffs_cr[25];
    uint8_t  cb_mult, cb_luma_mult;
    uint16_t cb_offset;
    uint8_t  cr_mult, cr_luma_mult;
    uint16_t cr_offset;
} av1pf_grain_params;

/* ---- deblocking side data built on the device ------------------------------
 * One record per MI (4x4 luma) cell of the frame's loop-filter record map
 * (threading.FrameWorkLoopFilterBlockRecord, loop_filter_map.go:46-60): every
 * cell a block covers holds that block's record, cells covered by no block are
 * not valid.  Enumerations are goav1's (dav1d's): size is tile.BlockSize
 * (0 = 128x128 ... 21 = 4x4, partition.go:43-66), tx_y / tx_uv are
 * tile.TransformSize (0 = 4x4 ... 18 = 64x16, tx_size.go:20-40).           */
#define AV1PF_LF_VALID   0x01u
#define AV1PF_LF_SKIP    0x02u   /* SkipTransform */
#define AV1PF_LF_INTRA   0x04u
#define AV1PF_LF_HAS_UV  0x08u   /* TransformTree.HasUV */

typedef struct av1pf_lf_record {
    uint16_t mi_col, mi_row;  /* Block.MICol / MIRow: the block's top-left cell */
    uint8_t  size;            /* Block.Size */
    uint8_t  tx_y, tx_uv;     /* TransformTree.Y (largest luma transform) / .UV */
    uint8_t  flags;           /* AV1PF_LF_* */
    uint16_t split[2];        /* TransformTree.Split: inter transform split bits per depth */
    uint8_t  segment;         /* SegmentID */
    uint8_t  ref;             /* RefFrame: 0 intra, 1..7 LAST..ALTREF */
    uint8_t  mode;            /* loopfilter.ModeDeltaClass: 0 GLOBALMV, 1 any other */
    uint8_t  pad;
    int8_t   delta_lf[4];     /* the block's delta_lf state: DeltaLF[0..3] with
                                 delta_lf_multi, else {DeltaLFFromBase, ..}; zero
                                 when delta_lf_present is off */
} av1pf_lf_record;

/* Frame-header fields of the level computation (lfmask.LFValueParams,
 * values.go:15-26) and the MI grid.                                         */
typedef struct av1pf_lf_build_params {
    int32_t mi_cols, mi_rows;  /* ((W+7)>>3)<<1, ((H+7)>>3)<<1 */
    int32_t ss_x, ss_y, mono;
    uint8_t level_y[2], level_u, level_v;  /* loop_filter_level[0..3] */
    uint8_t mode_ref_deltas;   /* loop_filter_delta_enabled */
    uint8_t delta_lf_multi;
    uint8_t seg_enabled;
    uint8_t pad;
    int8_t  ref_delta[8];      /* loop_filter_ref_deltas */
    int8_t  mode_delta[2];     /* loop_filter_mode_deltas */
    int8_t  pad2[2];
    int16_t seg_delta[8][4];   /* per segment: delta_lf_y_v, _y_h, _u, _v */
} av1pf_lf_build_params;

/* ---- one frame's request (FrameWorkPostFilterRequest, postfilter.go:278-284) */
/* Request flags.
 * AV1PF_REQ_REFRESH: the frame refreshes a reference slot
 * (FrameSize.RefreshFrameFlags != 0).  After loop restoration (and the
 * upscale) every plane of the output frame -- the upscaled frame with
 * AV1PF_STAGE_SUPERRES -- is edge-replicated into its whole allocation before
 * film grain, as goav1 does before a frame becomes a reference
 * (postfilter.go:760-771, Frame.ExtendBordersFrom, frame/buffer.go:293-409):
 * columns [plane width, stride / bytes per sample) of each visible row take
 * the row's last visible sample, and rows [plane height, allocated rows)
 * repeat the extended last visible row.  Runs even when no stage is left.  */
#define AV1PF_REQ_REFRESH 0x01u

typedef struct av1pf_request {
    uint32_t stages;                        /* AV1PF_STAGE_* still to run */
    uint32_t flags;                         /* AV1PF_REQ_* */
    const av1pf_deblock_params *deblock;    /* host struct, device arrays */
    const av1pf_cdef_params    *cdef;
    const av1pf_lr_params      *lr;
    const av1pf_grain_params   *grain;
    const av1pf_frame          *display;    /* film-grain output frame (separate buffer) */
    /* AV1PF_STAGE_SUPERRES: the upscaled frame (FrameWorkSuperResPostFilterRequest
     * output, postfilter_superres.go:85-95).  Deblocking and CDEF run on the
     * coded frame, which keeps their result; the upscale writes this frame's
     * visible samples; loop restoration (with its deblocked context rows
     * upscaled from the coded frame) and film grain run on it, so lr and
     * display are validated against this frame.  Same bit depth, format and
     * height as the coded frame, width in (coded width, 2 * coded width],
     * no overlap with the coded frame's planes.                            */
    const av1pf_frame          *upscaled;
} av1pf_request;

typedef struct av1pf_ctx av1pf_ctx;

/* Library handle: one per decoder (the reference runner is not reentrant,
 * postfilter_parallel.go:143-146).  device < 0 picks the current device.
 * A handle's scratch arenas (second frame buffers, LR context rows, the
 * device selection state) are ordered by the stream the calls are issued on:
 * issue every call of one handle on one stream, or order the streams (an
 * event) between calls; a decoder thread drives its handle from one stream. */
int  av1pf_create(av1pf_ctx **out, int device);
void av1pf_destroy(av1pf_ctx *ctx);
const char *av1pf_version(void);
const char *av1pf_last_error(const av1pf_ctx *ctx);

/* Run the remaining post-filter stages of one frame in place on `frame`
 * (film grain goes to req->display).  Replaces FrameWorkPostFilterRunner.Apply
 * (decoder/work.go:153-155, decoder_postfilter_bind.go:414); with
 * AV1PF_STAGE_SUPERRES it is the caller chain ApplyCallerPostFilters
 * (postfilter.go:947-979), which moves LR and film grain to req->upscaled.  */
int  av1pf_apply(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_request *req, void *stream);

/* av1pf_apply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same chain, same results as av1pf_apply; every arena the call
 * needs (descriptor image, second frame buffers, LR context rows, film-grain
 * templates) is carved from one static device array of
 * AV1PF_NOALLOC_SCRATCH_BYTES per device, enough for one 3840x2160 10-bit
 * 4:2:0 frame without super-resolution; a larger request fails with
 * AV1PF_ERR_SHORT_BUFFER before anything is launched.  The array is shared by
 * every handle on a device and keeps the last call's contents: calls issued
 * on one stream are ordered, calls that overlap on one GPU (other streams,
 * other handles) race.  No graph replay or autotuning on this path.         */
#define AV1PF_NOALLOC_SCRATCH_BYTES (32u << 20)
int  av1pf_apply_noalloc(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_request *req, void *stream);

/* Run many frames of mixed size, bit depth and chroma format.  One launch per
 * stage covers every frame of the batch: each workgroup finds its frame,
 * plane and tile in a segment table (no counterpart in the reference, which
 * runs frames one at a time).  Equivalent to n av1pf_apply calls.           */
int  av1pf_apply_batch(av1pf_ctx *ctx, int n, const av1pf_frame *frames, const av1pf_request *reqs, void *stream);

/* ---- shape classes (no counterpart in the reference) -----------------------
 * A shape class is everything a batch's launches depend on besides data:
 * the frame count and, per stage, the number of tiles and of segments.
 * av1pf_set_graphs(ctx, 1): each class is captured once as a HIP graph
 * (descriptor upload, side-stream grain templates, every stage launch) and
 * replayed on later calls of the class with the new descriptors; two graph
 * instances per class alternate, so a call waits only for the replay two
 * calls back.  Off while per-stage profiling is on.
 * av1pf_autotune: av1pf_apply_batch (same results) that also times, for
 * CDEF, LR and film-grain apply, each work distribution of the stage's tiles
 * (1, 2 or 4 tiles per workgroup, or persistent workgroups taking tiles from
 * a device queue) on the stage's own inputs and keeps the fastest for the
 * batch's class; modes[3] (may be NULL) receives the choice (0 = queue).
 * Synchronous.  Later calls of the class, captured graphs included, use it.
 * av1pf_set_stage_mode forces one stage's mode for every class (-1 clears);
 * stage is AV1PF_STAGE_CDEF, _LR or _FILMGRAIN.                              */
int  av1pf_set_graphs(av1pf_ctx *ctx, int enable);
int  av1pf_autotune(av1pf_ctx *ctx, int n, const av1pf_frame *frames, const av1pf_request *reqs, int32_t *modes,
                    void *stream);
int  av1pf_set_stage_mode(av1pf_ctx *ctx, uint32_t stage, int mode);

/* ---- pipeline mode (no counterpart in the reference's GPU-less design; the
 * counterpart of its superblock-row pipelines, decoder/postfilter_sbrow.go:9-37
 * and postfilter_rowpipe.go:6-27) -------------------------------------------
 * AV1PF_PIPELINE_STAGED (default): one launch per stage over all frames of a
 * call, ordered by the stream.
 * AV1PF_PIPELINE_FUSED: deblocking, the LR context-row save, CDEF,
 * super-resolution and LR of all frames of a call run in ONE persistent
 * launch: the stages' tiles form a device work queue in dependency order and
 * each frame keeps a completion counter per stage and 64-row band, so a band
 * is CDEF-filtered as soon as the deblocked rows it reads are final, and
 * restored as soon as its CDEF output and context rows are.  Same results
 * bit for bit.  Applies to every later call of the handle (graphs included);
 * film grain, copies and border extension stay separate launches.          */
#define AV1PF_PIPELINE_STAGED 0
#define AV1PF_PIPELINE_FUSED  1
int  av1pf_set_pipeline(av1pf_ctx *ctx, int mode);

/* Wait for all work the handle has queued on `stream`.                      */
int  av1pf_sync(av1pf_ctx *ctx, void *stream);

/* Per-stage entry points (the stage functions the chain calls):
 *   deblock  -> ApplyLoopFilterEdgesFromMasks  (postfilter_loopfilter_masks.go:33)
 *   cdef     -> ApplyCDEFPostFilter            (postfilter_cdef.go:295)
 *   lr       -> ApplyLoopRestorationPostFilter (postfilter.go:1277); needs the two
 *               boundary saves, made by av1pf_lr_save_boundaries
 *   grain    -> ApplyFilmGrainPostFilter       (postfilter_filmgrain.go:516)
 *   superres -> ApplySuperResPostFilter        (postfilter_superres.go:198)
 * The chain av1pf_apply composes them; these exist for stage-level callers.
 * LR after a super-resolution upscale needs its deblocked context rows
 * upscaled from the coded frame: run it through the chain.                  */
int  av1pf_deblock(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_deblock_params *p, void *stream);
int  av1pf_cdef(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_cdef_params *p, void *stream);
int  av1pf_lr_save_boundaries(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_lr_params *p, int after_cdef, void *stream);
int  av1pf_lr(av1pf_ctx *ctx, const av1pf_frame *frame, const av1pf_lr_params *p, void *stream);
int  av1pf_filmgrain(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *dst, const av1pf_grain_params *p, void *stream);
/* Build the deblocking side data on the device from the record map (device
 * array of mi_rows * mi_cols records): the edge masks as
 * FrameWorkLoopFilterMasks.BuildFromMap derives them from neighbouring
 * records (threading/loop_filter_masks.go:452-546: per block the lfmask
 * builders CreateIntra / CreateInter, lfmask/build.go:59,197, mask.go:76-528)
 * and the level cache as the decode-time build fills it (build.go:232,
 * levels from loop_filter_masks.go:88-123 and values.go:57-119).  masks
 * receives sb128_w * sb128_h records, every word written; level_cache
 * receives 4 * mi_cols * mi_rows bytes, cells no valid block covers (and
 * chroma bytes of cells no chroma block covers) left as they were.       */
int  av1pf_deblock_build(av1pf_ctx *ctx, const av1pf_lf_record *records, const av1pf_lf_build_params *p,
                         uint8_t *level_cache, av1pf_lf_mask *masks, void *stream);
/* Super-resolution upscale of every plane of `coded` into `upscaled`
 * (ApplySuperResPostFilter, decoder/postfilter_superres.go:198-264, one
 * superres.UpscalePlaneCoded per plane, superres/superres.go:38-60): each
 * visible row is resampled with AV1's 8-tap normative filter from the coded
 * plane width to the upscaled one, reading the MI-aligned coded span.      */
int  av1pf_superres(av1pf_ctx *ctx, const av1pf_frame *coded, const av1pf_frame *upscaled, void *stream);

/* ---- encoder CDEF strength search (cdef_search.go:580-749) ---------------
 * For every 64x64 filter block that has a non-skip 8x8 block, evaluate the 64
 * (primary, secondary) strength pairs of search level 1 on luma and on chroma
 * (U+V summed) and write the SSE (uint64, times the subsampling factor 1) to
 * mse[2][nfb][64] (plane 0 = luma, 1 = chroma).  src and recon are 4:2:0
 * frames of identical geometry, sizes a multiple of 8, 8-bit as goav1's
 * encoder searches, or 10-bit (an extension with no reference counterpart:
 * every trial is the decoder's 10-bit CDEF, strengths and damping scaled as
 * the decoder scales them); damping = 3 + (qindex >> 6) (cdef_search.go:588). */
int  av1pf_cdef_search(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon,
                       const uint64_t *noskip, int damping, uint64_t *mse, void *stream);

/* Host finish of the search (finish / jointStrengthSearchDual / searchOneDual,
 * cdef_search.go:780-945): from mse[2][nfb][64] in HOST memory pick cdef_bits
 * (0..3), the 1 << bits packed luma / chroma strengths (primary * 4 +
 * secondary code; unused entries 0) and each filter block's index.  Filter
 * blocks with skip_fb[fb] != 0 (no non-skip 8x8 block) are left out and get
 * index 0.  Needs no device.                                                */
int  av1pf_cdef_search_finish(const uint64_t *mse, int nfb, const uint8_t *skip_fb, uint64_t lambda,
                              uint8_t y_strength[8], uint8_t uv_strength[8], uint8_t *fb_index, int *cdef_bits);

/* The same selection on the device, after av1pf_cdef_search on the same
 * stream: mse, noskip, decision and fb_index are DEVICE pointers; recon gives
 * the frame geometry (the active blocks are those with a non-skip 8x8 block
 * inside the frame, as the search counts them).  At most 4096 filter blocks
 * (4096x4096 luma).  Asynchronous.                                          */
typedef struct av1pf_cdef_decision {
    int32_t cdef_bits;
    uint8_t y_strength[8];
    uint8_t uv_strength[8];
    uint8_t pad[4];
} av1pf_cdef_decision;

int  av1pf_cdef_search_select(av1pf_ctx *ctx, const av1pf_frame *recon, const uint64_t *noskip, const uint64_t *mse,
                              uint64_t lambda, av1pf_cdef_decision *decision, uint8_t *fb_index, void *stream);

/* Apply the searched CDEF to the reconstruction on the device, in place, as
 * the encoder's cdefApplier.apply does (encoder/cdef_apply.go:326-371: the
 * decoder's CDEF with the frame's chosen parameters; the unit index map of
 * markFromLoopFilterMap's noskip path, :202-228, reads a unit when one of its
 * 8x8 blocks inside the frame has residual).  decision and fb_index are the
 * DEVICE outputs of av1pf_cdef_search_select (no host round trip); noskip
 * as for the search; damping = 3 + (qindex >> 6).  Asynchronous.           */
int  av1pf_cdef_search_apply(av1pf_ctx *ctx, const av1pf_frame *recon, const uint64_t *noskip, int damping,
                             const av1pf_cdef_decision *decision, const uint8_t *fb_index, void *stream);

/* Full-search lambda (svtCDEFFullLambda, cdef_search.go:274-299) from the
 * 8-bit luma DC quantizer of the frame's qindex.  Needs no device.          */
uint64_t av1pf_cdef_lambda(int dc_quant, int key_frame, int gf_update, int tl_factor);

/* ---- encoder loop-restoration search ------------------------------------
 * The keyframe encoder's curated search on the luma plane of 8-bit frames
 * (SearchWienerUnitU8 / SearchSGRProjUnitU8, encoder/restoration.go:149-262,
 * and, as an extension with no reference counterpart, of 10-bit frames: the
 * estimates clip to the 10-bit range, the exact scores use the decoder's
 * 10-bit filters and the acceptance margin w*h is scaled by 16)
 * and their driver, encoder/keyframe_lossy.go:2310-2386): per restoration
 * unit of size unit_size (32..256; units offset 8 rows as the decoder's),
 * six Wiener tap triples are screened with the reference's H-only estimate,
 * the winner is re-scored with the exact Wiener, four SGR presets are scored
 * exactly, and the unit record (NONE / WIENER / SGRPROJ, same layout as the
 * decoder input) is written to units[hz * vt] with hz = max((W + us/2) / us, 1)
 * and vt likewise.  sse (optional, [hz * vt][12] uint64): 0 baseline, 1..6
 * Wiener estimates, 7 exact Wiener, 8..11 SGR presets; UINT64_MAX where the
 * reference computes nothing.  Device pointers; asynchronous.               */
int  av1pf_lr_search(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                     av1pf_lr_unit *units, uint64_t *sse, void *stream);
/* The exhaustive superset mode (no counterpart in the reference): per unit
 * of at most 64x64, the six curated Wiener triples scored exactly, a
 * least-squares Wiener fit (symmetric Q7 taps, alternating H/V solves from
 * int64 window statistics, rounded and clamped to the tap ranges), the four
 * curated SGR presets and all 16 SGR sets with a least-squares projection
 * (coded xqd clamped to [-96,31] x [-32,95]); every candidate scored with the
 * decoder's filters and taken, in that order, when sse + w*h beats the
 * running best, so the decision is never worse than av1pf_lr_search's.
 * sse[unit][28]: 0 baseline, 1..6 curated Wiener, 7 fitted Wiener, 8..11
 * curated SGR presets, 12..27 SGR sets 0..15 fitted (UINT64_MAX when not
 * scored); fit[unit][38]: fitted Wiener v[3], h[3], then xqd[2] per set.   */
int  av1pf_lr_search_exhaustive(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                                av1pf_lr_unit *units, uint64_t *sse, int8_t *fit, void *stream);
/* The same with unit_mask[hz * vt] (device): units with mask 0 are written
 * NONE without being scored (their sse entries are UINT64_MAX).             */
int  av1pf_lr_search_masked(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon, int unit_size,
                            const uint8_t *unit_mask, av1pf_lr_unit *units, uint64_t *sse, void *stream);

/* ---- learned proposal network ---------------------------------------------
 * No counterpart in the reference: a 48,659-parameter CNN over 64x64 luma
 * (source, reconstruction) tiles, BF16 on the matrix cores with FP32
 * accumulation, whose 131 outputs are logits of the best luma CDEF strength
 * (64, search-level-1 list order), the best chroma strength (64) and the
 * loop-restoration unit type (3: NONE, WIENER, SGRPROJ).  Trained by the
 * library's own forward/backward kernels (softmax cross-entropy against
 * target distributions, Adam on FP32 master weights).
 * Tiles: BF16, [nfb][64][64][2] (channel 0 the reconstruction, channel 1
 * the coding error (source - reconstruction) times 16, both divided by
 * 2^bd - 1), one per 64x64 filter block in raster order, edge samples
 * repeated past the frame.  Parameters: FP32, per layer W [K][N]
 * row-major then b [N], layers conv1 (K 18: 3x3 taps x 2 channels, tap-major),
 * conv2 (144), conv3 (288), fc1 (64 -> 128), heads (128 -> 131).           */
typedef struct av1pf_nn av1pf_nn;
int64_t av1pf_nn_param_count(void);
int  av1pf_nn_create(av1pf_ctx *ctx, int max_batch, av1pf_nn **out);
void av1pf_nn_destroy(av1pf_nn *nn);
int  av1pf_nn_set_params(av1pf_nn *nn, const float *params);     /* host FP32; resets the optimizer */
int  av1pf_nn_get_params(av1pf_nn *nn, float *params);           /* host FP32 */
int  av1pf_nn_tiles(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon, uint16_t *tiles, void *stream);
int  av1pf_nn_forward(av1pf_nn *nn, const uint16_t *tiles, int n, float *logits, void *stream);
/* Inference straight from an 8-bit or 10-bit frame pair: logits [nfb][131] for every
 * 64x64 filter block, one fused launch (tiles built on chip, the same
 * rounding points as av1pf_nn_tiles + av1pf_nn_forward, FP32 sums in another
 * order).  What av1pf_search_learned runs.                                  */
int  av1pf_nn_infer(av1pf_nn *nn, const av1pf_frame *src, const av1pf_frame *recon, float *logits, void *stream);
/* forward + backward of n <= max_batch tiles against targets [n][131] (three
 * distributions); writes grads [param_count] and the mean loss [1]          */
int  av1pf_nn_backward(av1pf_nn *nn, const uint16_t *tiles, int n, const float *targets, float *grads, float *loss,
                       void *stream);
int  av1pf_nn_adam(av1pf_nn *nn, const float *grads, float lr, float weight_decay, void *stream);

/* ---- learned-proposal search ------------------------------------------------
 * No counterpart in the reference.  The CDEF search restricted to a frame
 * shortlist of AV1PF_SHORTLIST luma and AV1PF_SHORTLIST chroma strengths
 * (cand[2 * AV1PF_SHORTLIST], search-level-1 positions 0..63, ascending
 * within each half): their exact SSE for every filter block
 * (mse_sub[2][nfb][AV1PF_SHORTLIST], same rules as av1pf_cdef_search), then
 * the same greedy strength selection over the AV1PF_SHORTLIST^2 shortlist
 * pairs, in one launch.  8-bit or 10-bit 4:2:0 frames up to 2048 filter
 * blocks (4K).
 * Device pointers, asynchronous.                                            */
#define AV1PF_SHORTLIST 4
int  av1pf_search_shortlist_size(void);   /* AV1PF_SHORTLIST of the built library */
int  av1pf_cdef_search_shortlist(av1pf_ctx *ctx, const av1pf_frame *src, const av1pf_frame *recon,
                                 const uint64_t *noskip, int damping, uint64_t lambda, const int32_t *cand,
                                 uint64_t *mse_sub, av1pf_cdef_decision *decision, uint8_t *fb_index, void *stream);
/* Whole learned search of one frame: tiles -> network -> shortlist (the
 * AV1PF_SHORTLIST strengths per plane with the largest summed probability
 * over the active blocks; cand[2 * AV1PF_SHORTLIST] and the mean top luma
 * probability conf[1] are written)
 * -> av1pf_cdef_search_shortlist -> the curated luma LR search (unit size
 * 64) on the units whose predicted P(NONE) is at most lr_none_threshold,
 * the rest NONE.  A caller wanting the exhaustive search as a fallback
 * checks conf after av1pf_sync.                                             */
int  av1pf_search_learned(av1pf_ctx *ctx, av1pf_nn *nn, const av1pf_frame *src, const av1pf_frame *recon,
                          const uint64_t *noskip, int damping, uint64_t lambda, float lr_none_threshold,
                          av1pf_cdef_decision *decision, uint8_t *fb_index, av1pf_lr_unit *lr_units, int32_t *cand,
                          float *conf, void *stream);

/* ---- per-stage timing (no counterpart in the reference, whose closest aid is
 * the FrameWorkPostFilterResult counters, postfilter.go:309-316) ----------
 * While enabled, every stage launch is bracketed by HIP events on its stream
 * and the stage's algorithmic bytes (samples read + written, 2*S*b for a
 * whole-plane pass) are accumulated.  av1pf_profile_read synchronises the
 * device and returns, per stage id, total ms, launch count and bytes.
 * Stage ids: 0 deblock, 1 lr_save, 2 cdef, 3 lr, 4 copy, 5 grain_template,
 * 6 grain_apply, 7 cdef_search, 8 cdef_search_select, 9 lr_search,
 * 10 nn_forward, 11 nn_train, 12 cdef_search_learned, 13 lr_search_exhaustive,
 * 14 superres, 15 deblock_build, 16 border extension, 17 fused chain
 * (AV1PF_PIPELINE_FUSED: deblock + LR save + CDEF + superres + LR).        */
#define AV1PF_PROFILE_STAGES 18
int  av1pf_profile(av1pf_ctx *ctx, int enable);
int  av1pf_profile_read(av1pf_ctx *ctx, double ms[AV1PF_PROFILE_STAGES], int count[AV1PF_PROFILE_STAGES],
                        double bytes[AV1PF_PROFILE_STAGES]);

#ifdef __cplusplus
}
#endif
#endif /* AV1PF_H */
