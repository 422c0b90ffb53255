/*
 * CPU oracle: reference-frame border extension.  Test infrastructure only.
 *
 * Restates Frame.ExtendBordersFrom(0) (frame/buffer.go:293-409), which goav1
 * runs after loop restoration and super-resolution, before film grain, when
 * the frame refreshes a reference slot (decoder/postfilter.go:760-771):
 *   extendRight  (buffer.go:367-409): for every visible row, the columns
 *                [width, stride / bytes per sample) take the last visible
 *                sample of the row;
 *   extendBorders (buffer.go:341-365): every allocated row past the visible
 *                height becomes a copy of the whole (already right-extended)
 *                last visible row.
 * Here a plane's stride is in samples and rows[] is the allocated row count.
 */
#include <string.h>
#include "orc.h"

static void extend_plane(uint16_t *pix, int stride, int alloc_rows, int width, int height)
{
    if (stride <= 0 || width <= 0 || height <= 0 || alloc_rows <= 0) return;
    const int w = width < stride ? width : stride;
    const int h = height < alloc_rows ? height : alloc_rows;
    /* extendRight over rows [0, h) */
    if (w < stride)
        for (int y = 0; y < h; y++) {
            uint16_t *row = pix + (size_t)y * stride;
            for (int x = w; x < stride; x++) row[x] = row[w - 1];
        }
    /* bottom rows: copies of the last valid row, corner included */
    const uint16_t *last = pix + (size_t)(h - 1) * stride;
    for (int y = h; y < alloc_rows; y++) memcpy(pix + (size_t)y * stride, last, sizeof(uint16_t) * (size_t)stride);
}

int orc_extend_borders(orc_frame *f)
{
    if (!f) return AV1PF_ERR_INVALID_ARG;
    const int np = f->mono ? 1 : 3;
    for (int i = 0; i < np; i++) {
        const int sx = i ? f->ssx : 0, sy = i ? f->ssy : 0;
        extend_plane(f->p[i], f->stride[i], f->rows[i], (f->w + sx) >> sx, (f->h + sy) >> sy);
    }
    return AV1PF_OK;
}
