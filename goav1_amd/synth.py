"""Seeded synthetic post-filter inputs.

goav1's encoder cannot emit every stage mix the benchmark workloads need (its
LR keyframe path turns CDEF off, the streaming encoder signals deblock level 0,
and the high-bit-depth encoders have no CDEF/LR search), so the inputs are
generated directly: a pre-filter reconstruction plus per-superblock filter
side data drawn from AV1's legal ranges.  The side data is built the way the
decoder builds it:

  - a random block partition of each 64x64 superblock with a transform size per
    block and a skip flag per block;
  - deblock edge masks in lfmask.FilterMask layout from the transform grid, the
    filter-width class of an edge being min(class of the two transforms it
    separates) exactly as dav1d/goav1 mask_edges_{intra,chroma} resolve it
    (internal/av1/lfmask/mask.go:330-470), so every mask obeys AV1's
    filter-width-versus-transform-spacing rule;
  - the 4-byte-per-MI level cache (lfmask/build.go:29-33), chroma levels at
    chroma MI coordinates (decoder/postfilter_loopfilter_masks.go:170-223);
  - CDEF strength index per 64x64 unit and the per-unit noskip word (bit by*8+bx
    set when any MI of the 8x8 block is not skip, postfilter_cdef.go:1276-1320);
  - loop-restoration unit records with legal Wiener taps and SGR sets/xqd
    (restoration/wiener.go:11-30, selfguided.go:24-30);
  - film-grain parameters.
"""
import numpy as np

from . import abi


def aligned8(v):
    return (v + 7) & ~7


class Plane:
    """One plane held as uint16 (all bit depths): visible w x h inside an
    allocation of alloc_w (the row pitch in samples) x alloc_h rows."""

    def __init__(self, w, h, alloc_w, alloc_h):
        self.w, self.h = w, h
        self.alloc_w, self.alloc_h = alloc_w, alloc_h
        self.pix = np.zeros((alloc_h, alloc_w), np.uint16)


class FrameData:
    """A frame: geometry plus planes and the per-frame post-filter side data.

    Planes are allocated at the MI-aligned extent, or with sb > 0 at the
    superblock-aligned extent goav1's frame.RequiredSize allocates
    (frame/buffer.go:141-236); pad_cols adds row-pitch columns past it (a
    stride alignment).  refresh marks a frame that refreshes a reference slot
    (its borders are edge-replicated after the filters)."""

    def __init__(self, w, h, bd, ssx, ssy, mono=False, sb=0, pad_cols=0):
        self.w, self.h, self.bd, self.ssx, self.ssy, self.mono = w, h, bd, ssx, ssy, int(mono)
        self.matrix_identity = 0
        self.refresh = False
        if sb:
            aw, ah = (w + sb - 1) // sb * sb, (h + sb - 1) // sb * sb
        else:
            aw, ah = aligned8(w), aligned8(h)
        assert pad_cols % 2 == 0
        aw += pad_cols
        self.planes = [Plane(w, h, aw, ah)]
        if not mono:
            cw, ch = (w + ssx) >> ssx, (h + ssy) >> ssy
            for _ in range(2):
                self.planes.append(Plane(cw, ch, aw >> ssx, ah >> ssy))
        self.deblock = None   # dict
        self.cdef = None
        self.lr = None
        self.grain = None
        self.upscaled = None   # FrameData of the upscaled geometry (super-resolution requests)
        self.stages = 0

    @property
    def samples(self):
        return sum(p.w * p.h for p in self.planes)

    @property
    def bytes_per_sample(self):
        return 1 if self.bd == 8 else 2

    def copy_pixels(self):
        out = FrameData.__new__(FrameData)
        out.__dict__.update(self.__dict__)
        out.planes = []
        for p in self.planes:
            q = Plane(p.w, p.h, p.alloc_w, p.alloc_h)
            q.pix = p.pix.copy()
            out.planes.append(q)
        return out


# ---------------------------------------------------------------- pixels
def fill_blocky(fr, rng, noise_sigma=4.0):
    """Per-8x8 random DC level in 16..235 + smooth bilinear gradient + Gaussian
    noise, clipped, scaled to the bit depth (the 720p keyframe recipe)."""
    scale = 1 << (fr.bd - 8)
    for pl in fr.planes:
        H, W = pl.alloc_h, pl.alloc_w
        dc = rng.integers(16, 236, size=((H + 7) // 8, (W + 7) // 8)).astype(np.float32)
        dc = np.repeat(np.repeat(dc, 8, 0), 8, 1)[:H, :W]
        g0, g1, g2, g3 = (float(v) for v in rng.uniform(-40, 40, size=4))
        u = np.linspace(0, 1, W, dtype=np.float32)[None, :]
        v = np.linspace(0, 1, H, dtype=np.float32)[:, None]
        grad = (g0 * (1 - v) + g2 * v) * (1 - u) + (g1 * (1 - v) + g3 * v) * u
        img = dc * 0.75 + 32 + grad
        img += rng.standard_normal((H, W), dtype=np.float32) * noise_sigma
        img *= scale
        pl.pix[:] = np.clip(np.rint(img), 0, (1 << fr.bd) - 1).astype(np.uint16)


def fill_rectangles(fr, rng, rects, t, sigma=6.0):
    """Background gradient plus textured rectangles translated by t*velocity
    (the moving-rectangle sequence recipe; sigma in 10-bit units)."""
    maxv = (1 << fr.bd) - 1
    s10 = sigma * (1 << fr.bd) / 1024.0
    for pi, pl in enumerate(fr.planes):
        sx = fr.ssx if pi else 0
        sy = fr.ssy if pi else 0
        H, W = pl.alloc_h, pl.alloc_w
        yy, xx = np.mgrid[0:H, 0:W]
        base = (0.25 + 0.5 * (xx / max(W - 1, 1)) * (0.6 + 0.4 * yy / max(H - 1, 1))) * maxv
        img = base.copy()
        for r in rects:
            x0 = int(round((r['x'] + r['vx'] * t))) >> sx
            y0 = int(round((r['y'] + r['vy'] * t))) >> sy
            rw, rh = max(r['w'] >> sx, 1), max(r['h'] >> sy, 1)
            xa, xb = max(x0, 0), min(x0 + rw, W)
            ya, yb = max(y0, 0), min(y0 + rh, H)
            if xa >= xb or ya >= yb:
                continue
            lvl = r['level'][pi] * maxv
            tex = r['tex'][pi]
            # texture is fixed to the rectangle, so it moves with it
            ty, tx = ya - y0, xa - x0
            img[ya:yb, xa:xb] = lvl + tex[ty:ty + (yb - ya), tx:tx + (xb - xa)] * s10
        pl.pix[:] = np.clip(np.rint(img), 0, maxv).astype(np.uint16)


def make_rects(rng, w, h, n=20, sigma=6.0):
    rects = []
    for _ in range(n):
        rw, rh = int(rng.integers(32, max(33, w // 4))), int(rng.integers(32, max(33, h // 4)))
        r = dict(x=float(rng.integers(0, w)), y=float(rng.integers(0, h)), w=rw, h=rh,
                 vx=float(rng.uniform(-8, 8)), vy=float(rng.uniform(-8, 8)),
                 level=rng.uniform(0.1, 0.9, size=3))
        r['vx'] = float(np.clip(r['vx'], -8, 8))
        r['tex'] = [rng.uniform(-1.7, 1.7, size=(rh + 1, rw + 1)) for _ in range(3)]
        rects.append(r)
    return rects


# ---------------------------------------------------------------- block structure
def _partition(rng, mi_rows, mi_cols, min_blk_log2=1):
    """Random quadtree partition of each 64x64 superblock.  Returns a list of
    blocks (y4, x4, log2 size in 4-units) clipped to the MI grid origin."""
    blocks = []
    stack = [(y, x, 4) for y in range(0, mi_rows, 16) for x in range(0, mi_cols, 16)]
    while stack:
        y, x, l2 = stack.pop()
        if y >= mi_rows or x >= mi_cols:
            continue
        n = 1 << l2
        must_split = (y + n > mi_rows or x + n > mi_cols) and l2 > min_blk_log2
        if l2 > min_blk_log2 and (must_split or rng.random() < 0.55):
            h = n >> 1
            stack += [(y, x, l2 - 1), (y, x + h, l2 - 1), (y + h, x, l2 - 1), (y + h, x + h, l2 - 1)]
        else:
            blocks.append((y, x, l2))
    return blocks


def _cls_luma(w4):
    return np.minimum(np.log2(np.maximum(w4, 1)).astype(np.int32), 2)


def _edge_classes(txid, txsz, luma):
    """Per-4x4 vertical-edge classes (-1 = no edge) from a transform id grid and
    transform widths in 4-units along the edge-normal axis."""
    cls = _cls_luma(txsz) if luma else (txsz > 1).astype(np.int32)
    e = np.full(txid.shape, -1, np.int32)
    diff = txid[:, 1:] != txid[:, :-1]
    c = np.minimum(cls[:, 1:], cls[:, :-1])
    e[:, 1:] = np.where(diff, c, -1)
    return e


def _pack(ev, eh, nclass, lane_v, lane_h, sbw, sbh, rcw, rch):
    """Pack per-4x4 edge class grids into [sb][dir][pos][class][half] uint16."""
    H, W = ev.shape
    Hp, Wp = sbh * rch, sbw * rcw
    evp = np.full((Hp, Wp), -1, np.int32)
    ehp = np.full((Hp, Wp), -1, np.int32)
    evp[:H, :W] = ev[:Hp, :Wp]
    ehp[:H, :W] = eh[:Hp, :Wp]
    out = np.zeros((sbh, sbw, 2, 32, nclass, 2), np.uint16)
    wv = (np.uint64(1) << np.arange(rch, dtype=np.uint64))
    wh = (np.uint64(1) << np.arange(rcw, dtype=np.uint64))
    for k in range(nclass):
        # vertical edges: pos = column, bits = rows
        bv = (evp == k).reshape(sbh, rch, sbw, rcw).astype(np.uint64)
        words = np.einsum('arbc,r->abc', bv, wv)          # (sbh, sbw, rcw)
        lo = words & np.uint64((1 << lane_v) - 1)
        hi = words >> np.uint64(lane_v)
        out[:, :, 0, :rcw, k, 0] = lo.astype(np.uint16)
        out[:, :, 0, :rcw, k, 1] = hi.astype(np.uint16)
        bh = (ehp == k).reshape(sbh, rch, sbw, rcw).astype(np.uint64)
        words = np.einsum('arbc,c->abr', bh, wh)          # (sbh, sbw, rch)
        lo = words & np.uint64((1 << lane_h) - 1)
        hi = words >> np.uint64(lane_h)
        out[:, :, 1, :rch, k, 0] = lo.astype(np.uint16)
        out[:, :, 1, :rch, k, 1] = hi.astype(np.uint16)
    return out


def build_side_data(fr, rng, *, lf_levels=None, lf_random=False, sharpness=0,
                    cdef=True, cdef_damping=None, lr_types=(abi.LR_SWITCHABLE,) * 3,
                    lr_unit=64, skip_prob=0.3, cdef_unread_prob=0.1):
    """Generate deblock / CDEF / LR side data for `fr` from a random partition."""
    mi_cols = ((fr.w + 7) >> 3) << 1
    mi_rows = ((fr.h + 7) >> 3) << 1
    blocks = _partition(rng, mi_rows, mi_cols)
    txid = np.zeros((mi_rows, mi_cols), np.int64)
    txw = np.ones((mi_rows, mi_cols), np.int32)
    skip = np.zeros((mi_rows, mi_cols), bool)
    lvl = np.zeros((mi_rows, mi_cols, 4), np.uint8)
    ssx, ssy = fr.ssx, fr.ssy
    cmi_rows, cmi_cols = (mi_rows + ssy) >> ssy, (mi_cols + ssx) >> ssx
    ctxid = np.zeros((cmi_rows, cmi_cols), np.int64)
    ctxw = np.ones((cmi_rows, cmi_cols), np.int32)
    tid = 1
    if lf_levels is None:
        lf_levels = (int(rng.integers(0, 64)),) * 4
    base = list(lf_levels)
    for (y, x, l2) in blocks:
        n = 1 << l2
        ye, xe = min(y + n, mi_rows), min(x + n, mi_cols)
        # luma transform: square, between n/4 and n, at most 16 units (64 px)
        tl2 = int(rng.integers(max(0, l2 - 2), min(l2, 4) + 1))
        t = 1 << tl2
        for ty in range(y, ye, t):
            for tx in range(x, xe, t):
                txid[ty:min(ty + t, ye), tx:min(tx + t, xe)] = tid
                tid += 1
        txw[y:ye, x:xe] = t
        skip[y:ye, x:xe] = rng.random() < skip_prob
        if lf_random:
            lv = [int(rng.integers(0, 64)) for _ in range(4)]
        else:
            lv = base
        lvl[y:ye, x:xe, 0] = lv[0]
        lvl[y:ye, x:xe, 1] = lv[1]
        if not fr.mono:
            cy, cx = y >> ssy, x >> ssx
            cn_y, cn_x = max(n >> ssy, 1), max(n >> ssx, 1)
            cye, cxe = min(cy + cn_y, cmi_rows), min(cx + cn_x, cmi_cols)
            # chroma transform: the largest square <= the chroma block, <= 32 px
            ct = min(cn_x, cn_y, 8)
            for ty in range(cy, cye, ct):
                for tx in range(cx, cxe, ct):
                    ctxid[ty:min(ty + ct, cye), tx:min(tx + ct, cxe)] = tid
                    tid += 1
            ctxw[cy:cye, cx:cxe] = ct
            lvl[cy:cye, cx:cxe, 2] = lv[2]
            lvl[cy:cye, cx:cxe, 3] = lv[3]
    sbw, sbh = (mi_cols + 31) >> 5, (mi_rows + 31) >> 5
    ev = _edge_classes(txid, txw, True)
    eh = _edge_classes(txid.T, txw.T, True).T
    ymask = _pack(ev, eh, 3, 16, 16, sbw, sbh, 32, 32)
    masks = np.zeros((sbh, sbw, 1280 // 2), np.uint16)
    masks[:, :, :384] = ymask.reshape(sbh, sbw, 384)
    if not fr.mono:
        cev = _edge_classes(ctxid, ctxw, False)
        ceh = _edge_classes(ctxid.T, ctxw.T, False).T
        cm = _pack(cev, ceh, 2, 16 >> ssy, 16 >> ssx, sbw, sbh, 32 >> ssx, 32 >> ssy)
        masks[:, :, 384:] = cm.reshape(sbh, sbw, 256)
    fr.deblock = dict(level_y=(base[0], base[1]), level_u=base[2], level_v=base[3],
                      sharpness=int(sharpness), mi_cols=mi_cols, mi_rows=mi_rows,
                      sb128_w=sbw, sb128_h=sbh,
                      level_cache=np.ascontiguousarray(lvl.reshape(-1)),
                      masks=np.ascontiguousarray(masks.reshape(-1)))
    fr.skip_mi = skip
    # transform grids the masks were built from (ids and square sizes in 4-sample
    # units, luma MI grid and chroma 4x4 grid), kept for the edge-list tests
    fr.tx = dict(luma=(txid, txw), chroma=(ctxid, ctxw))
    # ---- CDEF
    ucols, urows = (fr.w + 63) // 64, (fr.h + 63) // 64
    nsk = ~skip
    pad = np.zeros((urows * 16, ucols * 16), bool)
    pad[:mi_rows, :mi_cols] = nsk
    b8 = pad.reshape(urows * 8, 2, ucols * 8, 2).any(axis=(1, 3))
    b8 = b8.reshape(urows, 8, ucols, 8).transpose(0, 2, 1, 3).reshape(urows, ucols, 64)
    w64 = (b8.astype(np.uint64) << np.arange(64, dtype=np.uint64)).sum(axis=2, dtype=np.uint64)
    if cdef:
        bits = int(rng.integers(0, 4))
        cnt = 1 << bits
        ys = rng.integers(0, 64, size=8).astype(np.uint8)
        uvs = rng.integers(0, 64, size=8).astype(np.uint8)
        idx = rng.integers(0, cnt, size=(urows, ucols)).astype(np.int8)
        idx[w64 == 0] = -1
        idx[rng.random((urows, ucols)) < cdef_unread_prob] = -1
        fr.cdef = dict(damping=int(cdef_damping if cdef_damping is not None else rng.integers(3, 7)),
                       strength_count=cnt, y_strength=ys, uv_strength=uvs,
                       unit_cols=ucols, unit_rows=urows,
                       index=np.ascontiguousarray(idx.reshape(-1)),
                       noskip=np.ascontiguousarray(w64.reshape(-1)))
    # ---- loop restoration
    if lr_types is not None:
        types = list(lr_types)
        # chroma units halve only for 4:2:0 (parser/restoration.go:70-105)
        cu = max(lr_unit >> (ssx & ssy), 32)
        sizes = [lr_unit, cu, cu]
        units = []
        for pi in range(3):
            if pi >= len(fr.planes):
                types[pi] = abi.LR_NONE
            if types[pi] == abi.LR_NONE:
                units.append(None)
                continue
            pw, ph = fr.planes[pi].w, fr.planes[pi].h
            us = sizes[pi]
            hz, vt = max((pw + us // 2) // us, 1), max((ph + us // 2) // us, 1)
            arr = np.zeros(hz * vt, dtype=[('type', 'u1'), ('sgr_set', 'u1'), ('sgr_xqd', 'i1', 2),
                                           ('wiener_v', 'i1', 3), ('wiener_h', 'i1', 3), ('pad', 'u1', 6)])
            if types[pi] == abi.LR_SWITCHABLE:
                ut = rng.integers(0, 3, size=hz * vt)
            else:
                ut = np.where(rng.random(hz * vt) < 0.8, types[pi], abi.LR_NONE)
            arr['type'] = ut
            arr['sgr_set'] = rng.integers(0, 16, size=hz * vt)
            arr['sgr_xqd'][:, 0] = rng.integers(-96, 32, size=hz * vt)
            arr['sgr_xqd'][:, 1] = rng.integers(-32, 96, size=hz * vt)
            for key in ('wiener_v', 'wiener_h'):
                arr[key][:, 0] = rng.integers(-5, 11, size=hz * vt)
                arr[key][:, 1] = rng.integers(-23, 9, size=hz * vt)
                arr[key][:, 2] = rng.integers(-17, 47, size=hz * vt)
            units.append(arr)
        fr.lr = dict(frame_type=types, unit_size=sizes, units=units)


def random_grain(rng, bd, lag=3, n_y=8, mono=False, csfl=False):
    g = abi.GrainParams()
    g.seed = int(rng.integers(0, 65536))
    xs = np.sort(rng.choice(256, size=n_y, replace=False))
    g.num_y_points = n_y
    for i, x in enumerate(xs):
        g.y_points[2 * i] = int(x)
        g.y_points[2 * i + 1] = int(rng.integers(0, 256))
    if not mono and not csfl:
        for name, cnt_name in (('cb_points', 'num_cb_points'), ('cr_points', 'num_cr_points')):
            n = 4
            xs = np.sort(rng.choice(256, size=n, replace=False))
            setattr(g, cnt_name, n)
            arr = getattr(g, name)
            for i, x in enumerate(xs):
                arr[2 * i] = int(x)
                arr[2 * i + 1] = int(rng.integers(0, 256))
    g.chroma_scaling_from_luma = int(csfl)
    g.scaling_shift = 10
    g.ar_coeff_lag = lag
    g.ar_coeff_shift = 7
    g.grain_scale_shift = 0
    g.overlap = 1
    g.clip_to_restricted_range = 0
    ny = 2 * lag * (lag + 1)
    for i in range(ny):
        g.ar_coeffs_y[i] = int(rng.integers(-128, 128))
    for i in range(ny + 1):
        g.ar_coeffs_cb[i] = int(rng.integers(-128, 128))
        g.ar_coeffs_cr[i] = int(rng.integers(-128, 128))
    g.cb_mult, g.cb_luma_mult, g.cb_offset = int(rng.integers(0, 256)), int(rng.integers(0, 256)), int(rng.integers(0, 512))
    g.cr_mult, g.cr_luma_mult, g.cr_offset = int(rng.integers(0, 256)), int(rng.integers(0, 256)), int(rng.integers(0, 512))
    return g


# ---------------------------------------------------------------- workloads
def filter_level_from_qindex(q):
    """Deblock level used by the keyframe path: (ac*17563 - 421574 + 2^17) >> 18
    with ac the 8-bit AC quantizer (encoder/loopfilter_apply.go:43-60).  The AC
    quantizer values used by the workloads are pinned here from AV1's normative
    ac_qlookup table: q=120 -> 152, q=160 -> 305."""
    ac = {120: 152, 160: 305}[q]
    return max(0, min(63, (ac * 17563 - 421574 + (1 << 17)) >> 18))


def keyframe_720p(seed=0, w=1280, h=720, bd=8, ssx=1, ssy=1, stages=abi.STAGES_ALL & ~abi.STAGE_FILMGRAIN):
    """720p 8-bit 4:2:0 keyframe: blocky pixels, q120 deblock level, random
    CDEF strengths (damping 3 + (120 >> 6) = 4) and switchable 64-px LR units;
    allocated at the 64x64-superblock extent and refreshing a reference slot."""
    rng = np.random.default_rng(seed)
    fr = FrameData(w, h, bd, ssx, ssy, sb=64)
    fr.refresh = True
    fill_blocky(fr, rng)
    lv = filter_level_from_qindex(120)
    build_side_data(fr, rng, lf_levels=(lv, lv, lv, lv), cdef_damping=3 + (120 >> 6),
                    lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=64)
    fr.stages = stages
    return fr


def sequence_1080p10(seed=1, nframes=8, w=1920, h=1080, grain=True):
    """1080p 10-bit 4:2:0 moving-rectangle sequence with Wiener + SGR units and
    film grain (lag 3, 8 luma points, scaling shift 10); every frame allocated
    at the 64x64-superblock extent (1920x1088) and refreshing a reference slot."""
    rng = np.random.default_rng(seed)
    rects = make_rects(rng, w, h)
    frames = []
    for t in range(nframes):
        fr = FrameData(w, h, 10, 1, 1, sb=64)
        fr.refresh = True
        fill_rectangles(fr, rng, rects, t)
        build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                        lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=int(rng.choice([64, 128, 256])))
        fr.stages = abi.STAGES_ALL if grain else abi.STAGES_ALL & ~abi.STAGE_FILMGRAIN
        if grain:
            fr.grain = random_grain(rng, 10)
        frames.append(fr)
    return frames


def _dct_quantise(x, step):
    """8x8 orthonormal DCT, dead-zone quantisation (rounding offset 1/3 on AC,
    1/2 on DC) at `step`, inverse DCT: the blocking, ringing and detail loss of
    a transform codec."""
    import scipy.fft
    H, W = x.shape
    hb, wb = (H + 7) // 8, (W + 7) // 8
    pad = np.pad(x, ((0, hb * 8 - H), (0, wb * 8 - W)), mode="edge")
    blk = pad.reshape(hb, 8, wb, 8).transpose(0, 2, 1, 3)
    c = scipy.fft.dctn(blk, axes=(2, 3), norm="ortho")
    off = np.full((8, 8), 1.0 / 3.0)
    off[0, 0] = 0.5
    qs = np.full((8, 8), step)
    qs[0, 0] = step * 0.75
    c = np.sign(c) * np.floor(np.abs(c) / qs + off) * qs
    out = scipy.fft.idctn(c, axes=(2, 3), norm="ortho").transpose(0, 2, 1, 3).reshape(hb * 8, wb * 8)
    return out[:H, :W]


def search_pair(seed=3, w=1280, h=720, q=160, bd=8, skip_prob=0.15, skip_fb_prob=0.05, src=None, model="noise"):
    """Encoder-search input: a source frame (the 720p keyframe pixel recipe, or
    `src` when given), a reconstruction with coding error, the per-64x64
    non-skip bit masks (bit r*8+c = 8x8 block (r, c) has a non-skip MI; some
    filter blocks have none) and the damping 3 + (q >> 6) of the search.
    model "noise": per-8x8 smoothing toward the block mean plus noise growing
    with the quantiser; model "dct": 8x8 DCT dead-zone quantisation with a
    step that follows AV1's 8-bit AC quantiser (about 8 at q60, 20 at q120,
    39 at q160, 100 at q220, in orthonormal-DCT units)."""
    rng = np.random.default_rng(seed)
    if src is None:
        src = FrameData(w, h, bd, 1, 1)
        fill_blocky(src, rng)
    rec = src.copy_pixels()
    maxv = (1 << src.bd) - 1
    for ps, pr in zip(src.planes, rec.planes):
        x = ps.pix.astype(np.float64)
        if model == "dct":
            y = _dct_quantise(x, 8.0 * np.exp(0.0158 * (q - 60)) * (1 << (src.bd - 8)))
        else:
            step = 2 + q / 24.0
            H, W = x.shape
            hb, wb = (H + 7) // 8, (W + 7) // 8
            pad = np.pad(x, ((0, hb * 8 - H), (0, wb * 8 - W)), mode="edge")
            mean = pad.reshape(hb, 8, wb, 8).mean(axis=(1, 3))
            mean = np.repeat(np.repeat(mean, 8, 0), 8, 1)[:H, :W]
            y = x + 0.35 * (mean - x) + rng.normal(0, step * 0.6, size=x.shape)
        pr.pix[:] = np.clip(np.rint(y), 0, maxv).astype(np.uint16)
    mi_cols, mi_rows = aligned8(src.w) >> 2, aligned8(src.h) >> 2
    nhfb, nvfb = (mi_cols + 15) // 16, (mi_rows + 15) // 16
    bits = (rng.random((nvfb * nhfb, 64)) >= skip_prob).astype(np.uint64)
    bits[rng.random(nvfb * nhfb) < skip_fb_prob] = 0
    noskip = (bits << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    return src, rec, noskip, 3 + (q >> 6)


def search_dc_quant(q):
    """8-bit luma DC quantiser of the search workloads, pinned from AV1's
    normative dc_qlookup table: q=120 -> 123, q=160 -> 223."""
    return {120: 123, 160: 223}[q]


BATCH_SIZES = [(426, 240), (640, 360), (1280, 720), (1920, 1080), (3840, 2160)]


# ---------------------------------------------------------------- loop-filter record maps
LF_RECORD = np.dtype([('mi_col', '<u2'), ('mi_row', '<u2'), ('size', 'u1'), ('tx_y', 'u1'), ('tx_uv', 'u1'),
                      ('flags', 'u1'), ('split', '<u2', (2,)), ('segment', 'u1'), ('ref', 'u1'), ('mode', 'u1'),
                      ('pad', 'u1'), ('delta_lf', 'i1', (4,))])
assert LF_RECORD.itemsize == 20
# tile.BlockSize (dav1d order) as (width, height) in 4x4 units
BLOCK_W4H4 = [(32, 32), (32, 16), (16, 32), (16, 16), (16, 8), (16, 4), (8, 16), (8, 8), (8, 4), (8, 2), (4, 16),
              (4, 8), (4, 4), (4, 2), (4, 1), (2, 8), (2, 4), (2, 2), (2, 1), (1, 4), (1, 2), (1, 1)]
# tile.TransformSize as (log2 width, log2 height) in 4x4 units, and its one-split size
TX_L2 = [(0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2), (3, 4), (4, 3),
         (0, 2), (2, 0), (1, 3), (3, 1), (2, 4), (4, 2)]
TX_SUB = [0, 0, 1, 2, 3, 0, 0, 1, 1, 2, 2, 3, 3, 5, 6, 7, 8, 9, 10]
_BS_INDEX = {d: i for i, d in enumerate(BLOCK_W4H4)}
_TX_INDEX = {d: i for i, d in enumerate(TX_L2)}


def _largest_tx(w4, h4, cap_l2):
    lw, lh = min(int(w4).bit_length() - 1, cap_l2), min(int(h4).bit_length() - 1, cap_l2)
    lw, lh = min(lw, lh + 2), min(lh, lw + 2)   # transforms are at most 4:1
    return _TX_INDEX[(lw, lh)]


def record_map(rng, w, h, ssx=1, ssy=1, mono=False, intra_prob=0.4, skip_prob=0.3, delta_lf=False,
               split_prob=0.5):
    """A loop-filter record map for a w x h frame: a random AV1 partition of
    every 128x128 superblock (NONE / HORZ / VERT / SPLIT / HORZ_4 / VERT_4),
    intra blocks with one transform size (the largest or one or two splits
    smaller), inter blocks with the largest transform and random split words,
    random skip, segment, reference, mode class and delta_lf state.  Returns
    (records [mi_rows, mi_cols] of LF_RECORD, list of block origins)."""
    cols, rows = ((w + 7) >> 3) << 1, ((h + 7) >> 3) << 1
    recs = np.zeros((rows, cols), LF_RECORD)
    blocks = []

    def emit(x, y, w4, h4):
        if x >= cols or y >= rows:
            return
        r = np.zeros((), LF_RECORD)
        bs = _BS_INDEX[(w4, h4)]
        intra = rng.random() < intra_prob
        skip = rng.random() < skip_prob
        tmax = _largest_tx(w4, h4, 4)
        if intra:
            tx = tmax
            for _ in range(int(rng.integers(0, 3))):
                tx = TX_SUB[tx]
            r['tx_y'] = tx
        else:
            r['tx_y'] = tmax
            if rng.random() < split_prob:
                r['split'][0] = int(rng.integers(0, 1 << 16))
                r['split'][1] = int(rng.integers(0, 1 << 16))
        cw, ch = max(w4 >> ssx, 1), max(h4 >> ssy, 1)
        r['tx_uv'] = _largest_tx(cw, ch, 3)
        has_uv = (not mono) and (w4 > ssx or x & 1) and (h4 > ssy or y & 1)
        r['flags'] = 1 | (2 if skip else 0) | (4 if intra else 0) | (8 if has_uv else 0)
        r['mi_col'], r['mi_row'], r['size'] = x, y, bs
        r['segment'] = int(rng.integers(0, 8))
        r['ref'] = 0 if intra else int(rng.integers(1, 8))
        r['mode'] = 0 if intra else int(rng.integers(0, 2))
        if delta_lf:
            r['delta_lf'] = rng.integers(-63, 64, size=4)
        recs[y:min(y + h4, rows), x:min(x + w4, cols)] = r
        blocks.append((x, y))

    def part(x, y, n):
        if x >= cols or y >= rows:
            return
        opts = ['none', 'horz', 'vert', 'split']
        if n in (4, 8, 16):
            opts += ['horz4', 'vert4']
        weights = np.array([1.0, 1.0, 1.0, 3.0 if n > 4 else 1.5] + [0.5] * (len(opts) - 4))
        if n == 32:
            weights[0] = 0.2
        kind = opts[int(rng.choice(len(opts), p=weights / weights.sum()))]
        if kind == 'none':
            emit(x, y, n, n)
        elif kind == 'horz':
            emit(x, y, n, n // 2)
            emit(x, y + n // 2, n, n // 2)
        elif kind == 'vert':
            emit(x, y, n // 2, n)
            emit(x + n // 2, y, n // 2, n)
        elif kind == 'horz4':
            for k in range(4):
                emit(x, y + k * n // 4, n, n // 4)
        elif kind == 'vert4':
            for k in range(4):
                emit(x + k * n // 4, y, n // 4, n)
        elif n == 2:
            for k in range(4):
                emit(x + (k & 1), y + (k >> 1), 1, 1)
        else:
            for k in range(4):
                part(x + (k & 1) * n // 2, y + (k >> 1) * n // 2, n // 2)

    for sy in range(0, rows, 32):
        for sx in range(0, cols, 32):
            part(sx, sy, 32)
    return recs, blocks


def lf_build_params(rng, w, h, ssx=1, ssy=1, mono=False, seg=True, deltas=True, multi=False):
    """Random loop-filter header fields for av1pf_lf_build_params."""
    from . import abi
    p = abi.LFBuildParams()
    p.mi_cols, p.mi_rows = ((w + 7) >> 3) << 1, ((h + 7) >> 3) << 1
    p.ss_x, p.ss_y, p.mono = ssx, ssy, int(mono)
    lv = lambda: int(rng.choice([0, int(rng.integers(28, 36)), int(rng.integers(0, 64)), int(rng.integers(0, 64))]))
    p.level_y[0], p.level_y[1] = max(lv(), 1), lv()
    p.level_u, p.level_v = lv(), lv()
    p.mode_ref_deltas = int(deltas)
    p.delta_lf_multi = int(multi)
    p.seg_enabled = int(seg)
    for i in range(8):
        p.ref_delta[i] = int(rng.integers(-63, 64))
        for k in range(4):
            p.seg_delta[i][k] = int(rng.integers(-63, 64)) if rng.random() < 0.7 else 0
    for i in range(2):
        p.mode_delta[i] = int(rng.integers(-63, 64))
    return p


def superres_coded_width(up_w, denom):
    """Coded width of a super-resolution frame: (UpscaledWidth * 8 + denom / 2)
    / denom, at least min(16, UpscaledWidth) (AV1 superres_params)."""
    return max((up_w * 8 + denom // 2) // denom, min(16, up_w))


def superres_frame(seed=0, up_w=1920, h=1080, denom=12, bd=10, ssx=1, ssy=1, mono=False, grain=False,
                   lr_unit=64, stages=None):
    """A super-resolution frame: blocky coded picture with deblock and CDEF
    side data at the coded size, LR units on the upscaled grid, an empty
    upscaled frame, optional film grain (on the upscaled frame)."""
    rng = np.random.default_rng(seed)
    fr = FrameData(superres_coded_width(up_w, denom), h, bd, ssx, ssy, mono)
    fill_blocky(fr, rng)
    build_side_data(fr, rng, lr_types=None)
    up = FrameData(up_w, h, bd, ssx, ssy, mono, sb=64)
    fr.refresh = True
    grid = FrameData(up_w, h, bd, ssx, ssy, mono)
    build_side_data(grid, rng, cdef=False, lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=lr_unit)
    fr.lr = grid.lr
    fr.upscaled = up
    if grain:
        fr.grain = random_grain(rng, bd, mono=bool(mono))
    if stages is None:
        stages = abi.STAGES_ALL | abi.STAGE_SUPERRES
        if not grain:
            stages &= ~abi.STAGE_FILMGRAIN
    fr.stages = stages
    return fr


def batch_mixed(seed=2, n=64, sizes=BATCH_SIZES):
    """Mixed batch: sizes x {8,10}-bit x {4:2:0, 4:4:4}, random legal params."""
    rng = np.random.default_rng(seed)
    frames = []
    for _ in range(n):
        w, h = sizes[int(rng.integers(0, len(sizes)))]
        bd = int(rng.choice([8, 10]))
        sub = int(rng.integers(0, 2))
        fr = FrameData(w, h, bd, sub, sub, sb=64)
        fr.refresh = True
        fill_blocky(fr, rng)
        lrt = [int(rng.integers(0, 4)) for _ in range(3)]
        build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                        lr_types=lrt, lr_unit=int(rng.choice([64, 128, 256])))
        fr.stages = abi.STAGES_ALL & ~abi.STAGE_FILMGRAIN
        frames.append(fr)
    return frames
