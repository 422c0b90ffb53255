/*
 * CPU oracle: the post-filter stage chain.  Test infrastructure only.
 *
 * Restates decoder/postfilter.go:589-782 (applySupportedPostFiltersInPlace):
 *   deblock -> save LR boundary lines from the deblocked frame (afterCDEF=false,
 *   postfilter.go:683-692) -> CDEF -> save LR boundary lines from the CDEF output
 *   for the frame-edge stripes (afterCDEF=true, postfilter.go:733-738) -> loop
 *   restoration -> film grain into the display copy (postfilter.go:810-845).
 * With super-resolution it is the caller chain ApplyCallerPostFilters
 * (postfilter.go:947-979): deblock and CDEF on the coded frame, the deblocked
 * LR rows upscaled on save (postfilter.go:879-941), the upscale into the
 * upscaled frame, then the after-CDEF save, LR and film grain on it.
 * A refreshing frame (AV1PF_REQ_REFRESH: RefreshFrameFlags != 0) has its
 * borders edge-replicated after LR, before film grain (postfilter.go:760-771),
 * on the frame the tail ran on (the upscaled frame with super-resolution).
 */
#include "orc.h"

int orc_chain_sr(orc_frame *f, const av1pf_request *req, orc_frame *up, orc_frame *display)
{
    if (!f || !req) return AV1PF_ERR_INVALID_ARG;
    uint32_t st = req->stages;
    const int sr = (st & AV1PF_STAGE_SUPERRES) != 0;
    if (sr && !up) return AV1PF_ERR_INVALID_ARG;
    if ((st & AV1PF_STAGE_DEBLOCK) && !req->deblock) return AV1PF_ERR_INVALID_ARG;
    if ((st & AV1PF_STAGE_CDEF) && !req->cdef) return AV1PF_ERR_INVALID_ARG;
    if ((st & AV1PF_STAGE_LR) && !req->lr) return AV1PF_ERR_INVALID_ARG;
    if ((st & AV1PF_STAGE_FILMGRAIN) && (!req->grain || !display)) return AV1PF_ERR_INVALID_ARG;
    orc_frame *post = sr ? up : f;   /* the frame LR and film grain run on */
    int rc;
    if (st & AV1PF_STAGE_DEBLOCK)
        if ((rc = orc_deblock(f, req->deblock)) != AV1PF_OK) return rc;
    orc_lr_bounds b[3] = { { 0 } };
    if (st & AV1PF_STAGE_LR) {
        if ((rc = orc_lr_bounds_alloc(post, req->lr, b)) != AV1PF_OK) return rc;
        rc = sr ? orc_lr_save_superres(f, up, req->lr, b) : orc_lr_save(f, req->lr, b, 0);
        if (rc != AV1PF_OK) { orc_lr_bounds_free(b); return rc; }
    }
    if (st & AV1PF_STAGE_CDEF)
        if ((rc = orc_cdef(f, req->cdef)) != AV1PF_OK) { orc_lr_bounds_free(b); return rc; }
    if (sr)
        if ((rc = orc_superres(f, up)) != AV1PF_OK) { orc_lr_bounds_free(b); return rc; }
    if (st & AV1PF_STAGE_LR) {
        orc_lr_save(post, req->lr, b, 1);
        rc = orc_lr_apply(post, req->lr, b);
        orc_lr_bounds_free(b);
        if (rc != AV1PF_OK) return rc;
    }
    if (req->flags & AV1PF_REQ_REFRESH)
        if ((rc = orc_extend_borders(post)) != AV1PF_OK) return rc;
    if (st & AV1PF_STAGE_FILMGRAIN)
        if ((rc = orc_filmgrain(post, display, req->grain)) != AV1PF_OK) return rc;
    return AV1PF_OK;
}

int orc_chain(orc_frame *f, const av1pf_request *req, orc_frame *display)
{
    return orc_chain_sr(f, req, NULL, display);
}
