"""Frame-level oracle checks against independent formulations.

- Deblocking: the oracle walks goav1's packed lfmask layout
  (postfilter_loopfilter_masks.go); here the same frames are deblocked from the
  transform grids the masks were built from, edge by edge, in the AV1 spec's
  order (per plane all vertical edges, then all horizontal ones, an edge
  filtered when it lies on screen, filter length from the smaller transform
  across it, level from the q side or, when zero, the p side), with libaom's
  per-edge filter, which the unit tests cross-check against the dav1d-style one.
- CDEF: the spec's per-8x8-block process over the whole deblocked frame with
  CdefLarge outside the MI-aligned frame (against goav1's per-unit staging,
  postfilter_cdef.go:1101-1250).
- Loop restoration: the spec's get_source_sample stripe rule, Wiener filter
  process and box filter process (against goav1's libaom-style striped flow
  with saved boundary rows, tile/restoration_*.go).
- Film grain: the spec's generate-grain and add-noise processes with noise
  stripes and the noise image (against goav1's per-32x32-block apply,
  filmgrain/apply.go).
"""
import ctypes as C

import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc


@pytest.fixture(scope="module")
def L():
    lib = orc.lib()
    lib.orc_lf_edge_libaom.argtypes = [C.c_void_p, C.c_ssize_t, C.c_ssize_t, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int]
    lib.orc_lf_thresholds.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return lib


def thresholds(L, level, sharp):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    L.orc_lf_thresholds(level, sharp, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def deblock_edge_list(L, fr):
    """Spec-order deblocking of a copy of fr from its transform grids."""
    out = fr.copy_pixels()
    d = fr.deblock
    lc = d["level_cache"].reshape(-1, 4)
    mi_cols = d["mi_cols"]
    luma_zero = d["level_y"][0] == 0 and d["level_y"][1] == 0
    nplanes = 1 if fr.mono else 3
    for plane in range(nplanes):
        if plane and (luma_zero or (d["level_u"] if plane == 1 else d["level_v"]) == 0):
            continue
        sx, sy = (fr.ssx, fr.ssy) if plane else (0, 0)
        txid, txw = fr.tx["luma" if plane == 0 else "chroma"]
        pix = out.planes[plane].pix
        buf = np.ascontiguousarray(pix, dtype=np.uint16)
        stride = buf.shape[1]
        rows4, cols4 = txid.shape
        for vertical in (True, False):
            comp = (0 if vertical else 1) if plane == 0 else plane + 1
            for r in range(rows4):
                for c in range(cols4):
                    # on screen: the edge's luma position inside the visible frame
                    if (c * 4) << sx >= fr.w or (r * 4) << sy >= fr.h:
                        continue
                    if vertical:
                        if c == 0 or txid[r, c] == txid[r, c - 1]:
                            continue
                        pr, pc = r, c - 1
                    else:
                        if r == 0 or txid[r, c] == txid[r - 1, c]:
                            continue
                        pr, pc = r - 1, c
                    t = min(txw[r, c], txw[pr, pc])
                    width = (4 if t == 1 else 8 if t == 2 else 14) if plane == 0 else (4 if t == 1 else 6)
                    lvl = int(lc[r * mi_cols + c, comp])
                    if lvl == 0:
                        lvl = int(lc[pr * mi_cols + pc, comp])
                    if lvl == 0:
                        continue
                    lim, blim, hev = thresholds(L, lvl, d["sharpness"])
                    for k in range(4):
                        y, x = (r * 4 + k, c * 4) if vertical else (r * 4, c * 4 + k)
                        off = y * stride + x
                        L.orc_lf_edge_libaom(buf.ctypes.data, off, 1 if vertical else stride, fr.bd, width,
                                             lim, blim, hev)
        pix[:] = buf
    return out


@pytest.mark.parametrize("geom", [(136, 72, 8, 1), (200, 136, 10, 1), (90, 50, 8, 1), (66, 66, 10, 0),
                                  (130, 98, 8, 0)])
def test_deblock_mask_walk_equals_edge_list(L, geom):
    w, h, bd, ss = geom
    rng = np.random.default_rng(600 + w + h)
    fr = synth.FrameData(w, h, bd, ss, ss)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)), lr_types=None)
    # zero a fifth of the per-MI levels so the p-side fallback is exercised
    lc = fr.deblock["level_cache"]
    lc[rng.random(lc.shape) < 0.2] = 0
    want = deblock_edge_list(L, fr)
    got, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK)
    for i, (a, b) in enumerate(zip(got.planes, want.planes)):
        va, vb = a.pix[:a.h, :a.w], b.pix[:b.h, :b.w]
        assert np.array_equal(va, vb), f"plane {i}: {np.argwhere(va != vb)[:3]}"


# ---------------------------------------------------------------- loop restoration
# AV1 Sgr_Params (r0, s0, r1, s1) per parameter set (normative table)
SGR_PARAMS = [(2, 140, 1, 3236), (2, 112, 1, 2158), (2, 93, 1, 1618), (2, 80, 1, 1438), (2, 70, 1, 1295),
              (2, 58, 1, 1177), (2, 47, 1, 1079), (2, 37, 1, 996), (2, 30, 1, 925), (2, 25, 1, 863),
              (0, -1, 1, 2589), (0, -1, 1, 1618), (0, -1, 1, 1177), (0, -1, 1, 925), (2, 56, 0, -1),
              (2, 22, 0, -1)]


def r2(v, n):
    """Round2 on int64 arrays or ints."""
    return v if n <= 0 else (v + (1 << (n - 1))) >> n


def box_filter(S, rows, pas, r, s, bd):
    """AV1 box filter process (spec 7.17.3) of one pass.  S: int64 source
    window with a 3-sample margin around the output region; rows: the output
    rows' plane y (for the parity rule of the r = 2 pass).  Returns F."""
    H, W = S.shape
    h, w = H - 6, W - 6
    n = (2 * r + 1) ** 2
    one_by_n = ((1 << 12) + n // 2) // n
    k = 2 * r + 1
    c1 = np.pad(S, ((1, 0), (1, 0))).cumsum(0).cumsum(1)
    c2 = np.pad(S * S, ((1, 0), (1, 0))).cumsum(0).cumsum(1)
    # window sums; entry [i, j] is centred on S[i + r, j + r]
    bs = c1[k:, k:] - c1[:-k, k:] - c1[k:, :-k] + c1[:-k, :-k]
    bq = c2[k:, k:] - c2[:-k, k:] - c2[k:, :-k] + c2[:-k, :-k]
    a = r2(bq, 2 * (bd - 8))
    d = r2(bs, bd - 8)
    p = np.maximum(a * n - d * d, 0)
    z = np.minimum(r2(p * s, 20), 255)
    A = np.where(z == 0, 1, np.where(z >= 255, 256, ((z << 8) + z // 2) // np.maximum(z + 1, 1)))
    B = r2((256 - A) * bs * one_by_n, 12)
    o = 3 - r                     # centre of output (0, 0) in A/B coordinates
    Aw = lambda dy, dx: A[o + dy:o + dy + h, o + dx:o + dx + w]
    Bw = lambda dy, dx: B[o + dy:o + dy + h, o + dx:o + dx + w]
    u = S[3:3 + h, 3:3 + w]
    if pas == 0:
        ea = 6 * (Aw(-1, 0) + Aw(1, 0)) + 5 * (Aw(-1, -1) + Aw(-1, 1) + Aw(1, -1) + Aw(1, 1))
        eb = 6 * (Bw(-1, 0) + Bw(1, 0)) + 5 * (Bw(-1, -1) + Bw(-1, 1) + Bw(1, -1) + Bw(1, 1))
        oa = 6 * Aw(0, 0) + 5 * (Aw(0, -1) + Aw(0, 1))
        ob = 6 * Bw(0, 0) + 5 * (Bw(0, -1) + Bw(0, 1))
        even = (np.asarray(rows) & 1) == 0
        return np.where(even[:, None], r2(ea * u + eb, 9), r2(oa * u + ob, 8))
    ca = 4 * (Aw(0, 0) + Aw(-1, 0) + Aw(1, 0) + Aw(0, -1) + Aw(0, 1)) + \
        3 * (Aw(-1, -1) + Aw(-1, 1) + Aw(1, -1) + Aw(1, 1))
    cb = 4 * (Bw(0, 0) + Bw(-1, 0) + Bw(1, 0) + Bw(0, -1) + Bw(0, 1)) + \
        3 * (Bw(-1, -1) + Bw(-1, 1) + Bw(1, -1) + Bw(1, 1))
    return r2(ca * u + cb, 9)


def wiener(S, vt, ht, bd):
    """Wiener filter process (spec 7.17.4) with the spec's rounding variables
    and offset form of the intermediate."""
    round0, round1 = (5, 9) if bd == 12 else (3, 11)
    hf = np.array([ht[0], ht[1], ht[2], 128 - 2 * (ht[0] + ht[1] + ht[2]), ht[2], ht[1], ht[0]], np.int64)
    vf = np.array([vt[0], vt[1], vt[2], 128 - 2 * (vt[0] + vt[1] + vt[2]), vt[2], vt[1], vt[0]], np.int64)
    H, W = S.shape
    h, w = H - 6, W - 6
    offset = 1 << (bd + 7 - round0 - 1)
    limit = (1 << (bd + 1 + 7 - round0)) - 1
    inter = sum(hf[t] * S[:, t:t + w] for t in range(7))
    inter = np.clip(r2(inter, round0), -offset, limit - offset)
    out = sum(vf[t] * inter[t:t + h, :] for t in range(7))
    return np.clip(r2(out, round1), 0, (1 << bd) - 1)


def lr_spec(D, Cf, fr):
    """Loop restoration of frame planes Cf (CDEF output) with stripe context
    from D (deblocked), following spec 7.17: get_source_sample clamps to the
    plane, reads rows above a stripe from D (at most 2 rows up) and rows below
    it from D (at most 2 rows down), and everything inside the stripe from Cf."""
    import ctypes as C
    lib = orc.lib()
    lib.orc_sgr_decode_xq.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    out = [p.pix.astype(np.int64).copy() for p in Cf.planes]
    lrp = fr.lr
    for pl in range(len(fr.planes)):
        t = lrp["frame_type"][pl]
        if t == abi.LR_NONE:
            continue
        sx, sy = (fr.ssx, fr.ssy) if pl else (0, 0)
        pw, ph = fr.planes[pl].w, fr.planes[pl].h
        us = lrp["unit_size"][pl]
        hz, vt = max((pw + us // 2) // us, 1), max((ph + us // 2) // us, 1)
        units = np.ascontiguousarray(lrp["units"][pl]).view(np.uint8).reshape(-1, 16)
        dpl = D.planes[pl].pix.astype(np.int64)
        cpl = Cf.planes[pl].pix.astype(np.int64)
        for k in range(0, (ph << sy) // 64 + 2):
            y0s = (-8 + 64 * k) >> sy
            y1s = y0s + (64 >> sy) - 1
            ylo, yhi = max(0, y0s), min(ph - 1, y1s)
            if ylo > yhi:
                continue
            ys = np.arange(ylo - 3, yhi + 4)
            yc = np.clip(ys, 0, ph - 1)
            above, below = yc < y0s, yc > y1s
            ysrc = np.where(above, np.maximum(y0s - 2, yc), np.where(below, np.minimum(y1s + 2, yc), yc))
            xs = np.clip(np.arange(-3, pw + 3), 0, pw - 1)
            src = np.where((above | below)[:, None], dpl[ysrc][:, xs], cpl[ysrc][:, xs])
            rows = np.arange(ylo, yhi + 1)
            # unit row of each output row (luma 4x4 row of the sample, offset 8)
            luma_y = ((rows << sy) >> 2) << 2
            urow = np.minimum(vt - 1, ((luma_y + 8) >> sy) // us)
            for ur in np.unique(urow):
                sel = np.nonzero(urow == ur)[0]
                r0, r1 = sel[0], sel[-1] + 1
                for uc in range(hz):
                    x0 = uc * us
                    x1 = pw if uc == hz - 1 else x0 + us
                    rec = units[ur * hz + uc]
                    win = src[r0:r1 + 6, x0:x1 + 6]
                    typ = rec[0]
                    if typ == abi.LR_WIENER:
                        vtap = rec[4:7].view(np.int8).astype(int)
                        htap = rec[7:10].view(np.int8).astype(int)
                        res = wiener(win, vtap, htap, fr.bd)
                    elif typ == abi.LR_SGRPROJ:
                        st = int(rec[1])
                        xq = (C.c_int * 2)()
                        xqd = np.ascontiguousarray(rec[2:4].view(np.int8))
                        lib.orc_sgr_decode_xq(st, xqd.ctypes.data, xq)
                        ra, sa, rb, sb = SGR_PARAMS[st]
                        u = win[3:-3, 3:-3] << 4
                        v = u << 7
                        if ra:
                            v = v + xq[0] * (box_filter(win, rows[r0:r1], 0, ra, sa, fr.bd) - u)
                        if rb:
                            v = v + xq[1] * (box_filter(win, rows[r0:r1], 1, rb, sb, fr.bd) - u)
                        res = np.clip(r2(v, 11), 0, (1 << fr.bd) - 1)
                    else:
                        continue
                    out[pl][ylo + r0:ylo + r1, x0:x1] = res
    return out


@pytest.mark.parametrize("geom", [(136, 72, 8, 1, 64), (200, 136, 10, 1, 64), (130, 98, 8, 0, 64),
                                  (264, 200, 10, 1, 128), (90, 50, 8, 1, 64)])
def test_lr_striped_equals_spec(geom):
    w, h, bd, ss, unit = geom
    rng = np.random.default_rng(700 + w + h)
    fr = synth.FrameData(w, h, bd, ss, ss)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)),
                          lr_types=(abi.LR_SWITCHABLE,) * 3, lr_unit=unit)
    # strong CDEF everywhere, so the deblocked context rows differ from the
    # CDEF output the stripe interior reads
    c = fr.cdef
    c["y_strength"][:] = 63
    c["uv_strength"][:] = 63
    c["index"][:] = np.where(c["noskip"] != 0, 0, -1)
    D, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK)
    Cf, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK | abi.STAGE_CDEF)
    got, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK | abi.STAGE_CDEF | abi.STAGE_LR)
    want = lr_spec(D, Cf, fr)
    for i, pl in enumerate(got.planes):
        a, b = pl.pix[:pl.h, :pl.w], want[i][:pl.h, :pl.w]
        assert np.array_equal(a, b), f"plane {i}: first diffs {np.argwhere(a != b)[:4].tolist()}"


# ---------------------------------------------------------------- CDEF
CDEF_LARGE = 0x4000


def cdef_spec(fr, pre):
    """CDEF of frame `pre` (the deblocked frame) per spec 7.15: every non-skip
    8x8 block of a 64x64 unit with an index is filtered from the unfiltered
    frame; samples outside the MI-aligned frame are CdefLarge; luma primary
    strength adjusted by the block variance, chroma damping one less, all
    strengths and damping scaled by bd - 8; blocks use the luma direction,
    or direction 0 when the (unadjusted) primary strength is zero."""
    import ctypes as C
    lib = orc.lib()
    lib.orc_cdef_find_dir.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int32)]
    lib.orc_cdef_filter_block.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.orc_cdef_adjust_strength.argtypes = [C.c_int, C.c_int32]
    cd = fr.cdef
    cs = fr.bd - 8
    mi_cols, mi_rows = synth.aligned8(fr.w) >> 2, synth.aligned8(fr.h) >> 2
    out = pre.copy_pixels()
    padded = []
    for pl in pre.planes:
        ah, aw = pl.alloc_h, pl.alloc_w
        P = np.full((ah + 4, aw + 16), CDEF_LARGE, np.uint16)
        P[2:2 + ah, 8:8 + aw] = pl.pix[:ah, :aw]
        padded.append(P)
    for fbr in range(cd["unit_rows"]):
        for fbc in range(cd["unit_cols"]):
            u = fbr * cd["unit_cols"] + fbc
            idx = int(cd["index"][u])
            if idx < 0:
                continue
            ns = int(cd["noskip"][u])
            for by in range(8):
                for bx in range(8):
                    r8, c8 = fbr * 8 + by, fbc * 8 + bx       # 8x8 luma block coordinates
                    if 2 * r8 >= mi_rows or 2 * c8 >= mi_cols or not (ns >> (by * 8 + bx)) & 1:
                        continue
                    L0 = pre.planes[0].pix
                    blk = np.ascontiguousarray(L0[r8 * 8:r8 * 8 + 8, c8 * 8:c8 * 8 + 8])
                    d, v = C.c_int(), C.c_int32()
                    lib.orc_cdef_find_dir(blk.ctypes.data, 8, cs, C.byref(d), C.byref(v))
                    for pi, pl in enumerate(out.planes):
                        sx, sy = (fr.ssx, fr.ssy) if pi else (0, 0)
                        strength = int(cd["y_strength" if pi == 0 else "uv_strength"][idx])
                        pri, sec = strength >> 2, strength & 3
                        sec = 4 if sec == 3 else sec
                        pri, sec = pri << cs, sec << cs
                        damping = cd["damping"] + cs - (1 if pi else 0)
                        direction = d.value if pri else 0       # chosen before the variance adjustment
                        if pi == 0 and pri:
                            pri = lib.orc_cdef_adjust_strength(pri, v.value)
                        if pri == 0 and sec == 0:
                            continue
                        bw, bh = 8 >> sx, 8 >> sy
                        y0, x0 = r8 * bh, c8 * bw
                        win = np.zeros((bh + 4, 144), np.uint16)
                        win[:, :bw + 16] = padded[pi][y0:y0 + bh + 4, x0:x0 + bw + 16]
                        dst = np.zeros((bh, bw), np.uint16)
                        lib.orc_cdef_filter_block(dst.ctypes.data, bw, win[2:].ctypes.data + 16, pri, sec,
                                                  direction, damping, damping, cs, bw, bh)
                        pl.pix[y0:y0 + bh, x0:x0 + bw] = dst
    return out


@pytest.mark.parametrize("geom", [(136, 72, 8, 1), (200, 136, 10, 1), (130, 98, 8, 0), (90, 50, 10, 1),
                                  (66, 66, 8, 1)])
def test_cdef_frame_equals_spec(geom):
    w, h, bd, ss = geom
    rng = np.random.default_rng(800 + w + h)
    fr = synth.FrameData(w, h, bd, ss, ss)
    synth.fill_blocky(fr, rng)
    synth.build_side_data(fr, rng, lf_random=True, sharpness=int(rng.integers(0, 8)), lr_types=None)
    pre, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK)
    got, _ = orc.run_chain(fr, abi.STAGE_DEBLOCK | abi.STAGE_CDEF)
    want = cdef_spec(fr, pre)
    for i, (a, b) in enumerate(zip(got.planes, want.planes)):
        va, vb = a.pix[:a.h, :a.w], b.pix[:b.h, :b.w]
        assert np.array_equal(va, vb), f"plane {i}: {np.argwhere(va != vb)[:4].tolist()}"


# ---------------------------------------------------------------- film grain
def _gaussian_table():
    import os
    import re
    txt = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "av1pf_gaussian_table.h")).read()
    body = txt[txt.index("#define AV1PF_GAUSSIAN_VALUES") + len("#define AV1PF_GAUSSIAN_VALUES"):]
    return [int(v) for v in re.findall(r"-?\d+", body)][:2048]


class SpecRng:
    def __init__(self, seed):
        self.r = seed & 0xFFFF

    def get(self, bits):
        r = self.r
        bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1
        self.r = ((r >> 1) | (bit << 15)) & 0xFFFF
        return (self.r >> (16 - bits)) & ((1 << bits) - 1)


def grain_templates_spec(g, bd, ssx, ssy, mono):
    """Generate grain process (spec 7.18.3.3)."""
    gauss = _gaussian_table()
    gmin, gmax = -(128 << (bd - 8)), (128 << (bd - 8)) - 1
    shift = 12 - bd + g.grain_scale_shift
    lag = g.ar_coeff_lag
    L = np.zeros((73, 82), np.int64)
    rng = SpecRng(g.seed)
    if g.num_y_points:
        for y in range(73):
            for x in range(82):
                L[y, x] = r2(gauss[rng.get(11)], shift)
    sh = g.ar_coeff_shift
    for y in range(3, 73):
        for x in range(3, 82 - 3):
            s, pos = 0, 0
            for dr in range(-lag, 1):
                for dc in range(-lag, lag + 1):
                    if dr == 0 and dc == 0:
                        break
                    s += L[y + dr, x + dc] * g.ar_coeffs_y[pos]
                    pos += 1
            L[y, x] = min(max(L[y, x] + r2(s, sh), gmin), gmax)
    if mono:
        return L, None, None
    cw, ch = (44 if ssx else 82), (38 if ssy else 73)
    out = []
    for pl, key, coefs in ((1, 0xb524, g.ar_coeffs_cb), (2, 0x49d8, g.ar_coeffs_cr)):
        Cg = np.zeros((73, 82), np.int64)
        rng = SpecRng(g.seed ^ key)
        npts = g.num_cb_points if pl == 1 else g.num_cr_points
        if npts or g.chroma_scaling_from_luma:
            for y in range(ch):
                for x in range(cw):
                    Cg[y, x] = r2(gauss[rng.get(11)], shift)
        for y in range(3, ch):
            for x in range(3, cw - 3):
                s, pos = 0, 0
                for dr in range(-lag, 1):
                    for dc in range(-lag, lag + 1):
                        if dr == 0 and dc == 0:
                            if g.num_y_points:
                                lx, ly = ((x - 3) << ssx) + 3, ((y - 3) << ssy) + 3
                                lum = sum(L[ly + i, lx + j] for i in range(ssy + 1) for j in range(ssx + 1))
                                s += r2(lum, ssx + ssy) * coefs[pos]
                            break
                        s += coefs[pos] * Cg[y + dr, x + dc]
                        pos += 1
                Cg[y, x] = min(max(Cg[y, x] + r2(s, sh), gmin), gmax)
        out.append(Cg)
    return L, out[0], out[1]


def grain_apply_spec(fr, img, g):
    """Add-noise synthesis (spec 7.18.3.5): noise stripes of 34 rows per 32
    luma rows with horizontal overlap blends, the noise image with vertical
    overlap blends, then scaling and clipping."""
    import ctypes as C
    lib = orc.lib()
    lib.orc_fg_scaling_lut.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    bd, ssx, ssy = fr.bd, fr.ssx, fr.ssy
    W, H = fr.w, fr.h
    mono = len(fr.planes) == 1
    tm = grain_templates_spec(g, bd, ssx, ssy, mono)
    gmin, gmax = -(128 << (bd - 8)), (128 << (bd - 8)) - 1
    nplanes = 1 if mono else 3
    stripes = {}
    for y in range(0, (H + 1) // 2, 16):
        num = y // 16
        rng = SpecRng(g.seed ^ (((num * 37 + 178) & 255) << 8) ^ ((num * 173 + 105) & 255))
        for pl in range(nplanes):
            sx, sy = (ssx, ssy) if pl else (0, 0)
            stripes.setdefault((num, pl), np.zeros((34, W + 34 + 2), np.int64))
        for x in range(0, (W + 1) // 2, 16):
            rnd = rng.get(8)
            ox, oy = rnd >> 4, rnd & 15
            for pl in range(nplanes):
                sx, sy = (ssx, ssy) if pl else (0, 0)
                px = 6 + ox if sx else 9 + ox * 2
                py = 6 + oy if sy else 9 + oy * 2
                S = stripes[(num, pl)]
                for i in range(34 >> sy):
                    for j in range(34 >> sx):
                        v = int(tm[pl][py + i, px + j])
                        if sx == 0:
                            if j < 2 and g.overlap and x > 0:
                                old = int(S[i, x * 2 + j])
                                v = old * 27 + v * 17 if j == 0 else old * 17 + v * 27
                                v = min(max(r2(v, 5), gmin), gmax)
                            S[i, x * 2 + j] = v
                        else:
                            if j == 0 and g.overlap and x > 0:
                                old = int(S[i, x + j])
                                v = min(max(r2(old * 23 + v * 22, 5), gmin), gmax)
                            S[i, x + j] = v
    out = [p.pix.astype(np.int64).copy() for p in img.planes]
    luts = []
    for pts, n in ((g.y_points, g.num_y_points), (g.cb_points, g.num_cb_points), (g.cr_points, g.num_cr_points)):
        lut = np.zeros(256, np.uint8)
        arr = np.array(list(pts), np.uint8)
        lib.orc_fg_scaling_lut(lut.ctypes.data, arr.ctypes.data, n)
        luts.append(lut.astype(np.int64))
    if g.chroma_scaling_from_luma:
        luts[1] = luts[2] = luts[0]

    def scale(pl, idx):
        x = idx >> (bd - 8)
        rem = idx - (x << (bd - 8))
        if bd == 8 or x == 255:
            return int(luts[pl][x])
        a, b = int(luts[pl][x]), int(luts[pl][x + 1])
        return a + r2((b - a) * rem, bd - 8)

    lo = (16 << (bd - 8)) if g.clip_to_restricted_range else 0
    hi_l = (235 << (bd - 8)) if g.clip_to_restricted_range else (256 << (bd - 8)) - 1
    hi_c = ((235 if fr.matrix_identity else 240) << (bd - 8)) if g.clip_to_restricted_range else (256 << (bd - 8)) - 1
    Lsrc = img.planes[0].pix.astype(np.int64)
    for pl in range(nplanes):
        sx, sy = (ssx, ssy) if pl else (0, 0)
        if pl == 0 and not g.num_y_points:
            continue
        if pl and not ((g.num_cb_points if pl == 1 else g.num_cr_points) or g.chroma_scaling_from_luma):
            continue
        pw, ph = (W + sx) >> sx, (H + sy) >> sy
        src = img.planes[pl].pix.astype(np.int64)
        for y in range(ph):
            num = y >> (5 - sy)
            i = y - (num << (5 - sy))
            for x in range(pw):
                v = int(stripes[(num, pl)][i, x])
                if sy == 0:
                    if i < 2 and num > 0 and g.overlap:
                        old = int(stripes[(num - 1, pl)][i + 32, x])
                        v = old * 27 + v * 17 if i == 0 else old * 17 + v * 27
                        v = min(max(r2(v, 5), gmin), gmax)
                else:
                    if i < 1 and num > 0 and g.overlap:
                        old = int(stripes[(num - 1, pl)][i + 16, x])
                        v = min(max(r2(old * 23 + v * 22, 5), gmin), gmax)
                orig = int(src[y, x])
                if pl == 0:
                    merged, top = orig, hi_l
                else:
                    lx, ly = x << sx, y << sy
                    lnx = min(lx + 1, W - 1)
                    avg = r2(int(Lsrc[ly, lx]) + int(Lsrc[ly, lnx]), 1) if sx else int(Lsrc[ly, lx])
                    if g.chroma_scaling_from_luma:
                        merged = avg
                    else:
                        mult, lmult, off = ((g.cb_mult, g.cb_luma_mult, g.cb_offset) if pl == 1 else
                                            (g.cr_mult, g.cr_luma_mult, g.cr_offset))
                        comb = avg * (lmult - 128) + orig * (mult - 128)
                        merged = min(max((comb >> 6) + ((off - 256) << (bd - 8)), 0), (1 << bd) - 1)
                    top = hi_c
                noise = r2(scale(pl, merged) * v, g.scaling_shift)
                out[pl][y, x] = min(max(orig + noise, lo), top)
    return out


@pytest.mark.parametrize("case", [
    dict(w=100, h=70, bd=8, ss=1, lag=3), dict(w=97, h=67, bd=10, ss=1, lag=2, restricted=1),
    dict(w=72, h=40, bd=10, ss=0, lag=1, overlap=0), dict(w=66, h=66, bd=8, ss=1, lag=0, csfl=1),
    dict(w=80, h=36, bd=10, ss=1, lag=3, gss=2, shift=8, arshift=9),
    dict(w=61, h=45, bd=8, ss=0, lag=3, restricted=1, identity=1)])
def test_film_grain_equals_spec(case):
    c = dict(restricted=0, overlap=1, csfl=0, gss=0, shift=10, arshift=7, identity=0)
    c.update(case)
    rng = np.random.default_rng(900 + c["w"] + c["h"])
    fr = synth.FrameData(c["w"], c["h"], c["bd"], c["ss"], c["ss"])
    synth.fill_blocky(fr, rng)
    fr.matrix_identity = c["identity"]
    g = synth.random_grain(rng, c["bd"], lag=c["lag"], csfl=bool(c["csfl"]))
    g.clip_to_restricted_range, g.overlap = c["restricted"], c["overlap"]
    g.grain_scale_shift, g.scaling_shift, g.ar_coeff_shift = c["gss"], c["shift"], c["arshift"]
    fr.grain = g
    fr.stages = abi.STAGE_FILMGRAIN
    _, disp = orc.run_chain(fr)
    want = grain_apply_spec(fr, fr, g)
    for i, pl in enumerate(disp.planes):
        a, b = pl.pix[:pl.h, :pl.w], want[i][:pl.h, :pl.w]
        assert np.array_equal(a, b), f"plane {i}: {np.argwhere(a != b)[:4].tolist()}"
