"""Deblocking side data from the loop-filter record map: the oracle pinned,
then av1pf_deblock_build against it.

Oracle pins (CPU):
  - goav1's known-answer cases of the lfmask builders (lfmask/mask_test.go:
    intra 4x4, inter 8x8, inter 16x16 with a split, chroma 4:2:0, the high
    half-lane, the level cache of CreateIntra), each as a one-block map;
  - the levels against AV1's filter-level derivation written from the
    specification (section 7.14.4: delta_lf, segment feature, reference and
    mode deltas scaled by 2 from level 32), over random headers;
  - whole maps against an independent per-edge formulation in numpy: every
    block's transform leaves on a cell grid, an edge wherever a block or a
    leaf starts (inter skip blocks: block edges only), class = the smaller of
    the clamped log2 sizes on the two sides, chroma from the block carrying
    each chroma cell.
GPU: masks and level cache bit-exact against the oracle on random maps
(4:2:0 / 4:2:2 / 4:4:4 / mono, several regions, up to 1080p), and the
deblocked frame from the device-built side data equals the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc


def one_block_map(cols, rows, x, y, size, tx_y, tx_uv=0, intra=True, skip=False, split=(0, 0), has_uv=True):
    recs = np.zeros((rows, cols), synth.LF_RECORD)
    w4, h4 = synth.BLOCK_W4H4[size]
    r = np.zeros((), synth.LF_RECORD)
    r['mi_col'], r['mi_row'], r['size'], r['tx_y'], r['tx_uv'] = x, y, size, tx_y, tx_uv
    r['flags'] = abi.LF_VALID | (abi.LF_INTRA if intra else 0) | (abi.LF_SKIP if skip else 0) | (abi.LF_HAS_UV if has_uv else 0)
    r['split'] = split
    r['ref'] = 0 if intra else 1
    recs[y:min(y + h4, rows), x:min(x + w4, cols)] = r
    return recs


def params(cols, rows, ssx=1, ssy=1, mono=False, levels=(32, 32, 32, 32)):
    p = abi.LFBuildParams()
    p.mi_cols, p.mi_rows, p.ss_x, p.ss_y, p.mono = cols, rows, ssx, ssy, int(mono)
    p.level_y[0], p.level_y[1], p.level_u, p.level_v = levels
    return p


def ymask(masks, region, dir_, pos, cls, half):
    return int(masks.reshape(-1, 640)[region][((dir_ * 32 + pos) * 3 + cls) * 2 + half])


def uvmask(masks, region, dir_, pos, cls, half):
    return int(masks.reshape(-1, 640)[region][384 + ((dir_ * 32 + pos) * 2 + cls) * 2 + half])


BS = {f"{4 * w}x{4 * h}": i for i, (w, h) in enumerate(synth.BLOCK_W4H4)}
TX = {f"{4 << a}x{4 << b}": i for i, (a, b) in enumerate(synth.TX_L2)}


# ---------------------------------------------------------------- oracle pins
def test_oracle_tables_match_the_dimension_lists():
    lib = orc.lib()
    lib.orc_lf_tables.restype = C.POINTER(C.c_uint8)
    w4, h4 = lib.orc_lf_tables(0), lib.orc_lf_tables(1)
    assert [(w4[i], h4[i]) for i in range(22)] == synth.BLOCK_W4H4
    l2w, l2h, sub = lib.orc_lf_tables(2), lib.orc_lf_tables(3), lib.orc_lf_tables(4)
    assert [(l2w[i], l2h[i]) for i in range(19)] == synth.TX_L2
    # a split halves the longer side, or both sides of a square
    for t in range(1, 19):
        a, b = synth.TX_L2[t]
        sa, sb = synth.TX_L2[sub[t]]
        assert (sa, sb) == ((a - 1, b - 1) if a == b else (a - 1, b) if a > b else (a, b - 1)), t


def test_kat_intra_4x4():          # mask_test.go:205-227
    m, _ = orc.deblock_build(one_block_map(2, 2, 0, 0, BS["4x4"], TX["4x4"]), params(2, 2))
    assert ymask(m, 0, 0, 0, 0, 0) == 1 and ymask(m, 0, 1, 0, 0, 0) == 1
    assert ymask(m, 0, 0, 0, 1, 0) == 0 and ymask(m, 0, 0, 0, 2, 0) == 0


def test_kat_inter_8x8():          # mask_test.go:229-259
    m, _ = orc.deblock_build(one_block_map(2, 2, 0, 0, BS["8x8"], TX["8x8"], intra=False), params(2, 2))
    assert ymask(m, 0, 0, 0, 1, 0) == 0b11 and ymask(m, 0, 1, 0, 1, 0) == 0b11
    assert ymask(m, 0, 0, 1, 0, 0) == 0 and ymask(m, 0, 0, 1, 1, 0) == 0


def test_kat_inter_16x16_split():  # mask_test.go:261-284
    m, _ = orc.deblock_build(one_block_map(4, 4, 0, 0, BS["16x16"], TX["16x16"], intra=False, split=(1, 0)),
                             params(4, 4))
    assert ymask(m, 0, 0, 2, 1, 0) == 0b1111 and ymask(m, 0, 1, 2, 1, 0) == 0b1111
    assert ymask(m, 0, 0, 0, 1, 0) == 0b1111


def test_kat_chroma_420():         # mask_test.go:286-304
    m, _ = orc.deblock_build(one_block_map(2, 2, 0, 0, BS["8x8"], TX["8x8"], TX["4x4"]), params(2, 2))
    assert uvmask(m, 0, 0, 0, 0, 0) == 1 and uvmask(m, 0, 1, 0, 0, 0) == 1


def test_kat_high_half_lane():     # mask_test.go:306-327
    m, _ = orc.deblock_build(one_block_map(2, 24, 0, 20, BS["4x16"], TX["4x16"], intra=False), params(2, 24))
    assert ymask(m, 0, 0, 0, 0, 1) == 0b11110000 and ymask(m, 0, 0, 0, 0, 0) == 0


def test_kat_level_cache():        # mask_test.go:329-357
    _, lc = orc.deblock_build(one_block_map(8, 8, 0, 0, BS["8x8"], TX["8x8"], TX["4x4"]),
                              params(8, 8, levels=(10, 11, 12, 13)))
    lc = lc.reshape(8, 8, 4)
    assert (lc[:2, :2, :2] == [10, 11]).all()
    assert list(lc[0, 0, 2:]) == [12, 13]


def spec_levels(p, r):
    """Filter levels of a block per AV1 section 7.14.4 (Y vertical, Y
    horizontal, U, V)."""
    if p.level_y[0] == 0 and p.level_y[1] == 0:
        return [0, 0, 0, 0]
    out = []
    bases = [p.level_y[0], p.level_y[1], p.level_u, p.level_v]
    for i in range(4):
        if i >= 2 and bases[i] == 0:
            out.append(0)
            continue
        delta = int(r['delta_lf'][i if p.delta_lf_multi else 0])
        lvl = min(max(delta + bases[i], 0), 63)
        if p.seg_enabled:
            lvl = min(max(lvl + p.seg_delta[int(r['segment'])][i], 0), 63)
        if p.mode_ref_deltas:
            n_shift = lvl >> 5
            ref = int(r['ref'])
            lvl += p.ref_delta[ref] << n_shift
            if ref > 0:
                lvl += p.mode_delta[int(r['mode'])] << n_shift
            lvl = min(max(lvl, 0), 63)
        out.append(lvl)
    return out


def test_oracle_levels_match_the_specification():
    rng = np.random.default_rng(7)
    lib = orc.lib()
    for it in range(3000):
        p = synth.lf_build_params(rng, 64, 64, seg=bool(rng.integers(0, 2)), deltas=bool(rng.integers(0, 2)),
                                  multi=bool(rng.integers(0, 2)))
        if rng.random() < 0.1:
            p.level_y[0] = p.level_y[1] = 0
        r = np.zeros((), synth.LF_RECORD)
        r['segment'], r['ref'], r['mode'] = rng.integers(0, 8), rng.integers(0, 8), rng.integers(0, 2)
        r['delta_lf'] = rng.integers(-63, 64, size=4)
        rr = np.ascontiguousarray(np.array([r]))
        out = (C.c_uint8 * 4)()
        lib.orc_lf_block_levels(C.byref(p), rr.ctypes.data, out)
        assert list(out) == spec_levels(p, rr[0]), (it, list(out))


def leaf_grid(r, cols, rows):
    """Per visible cell of block r: (leaf x0, y0, log2 w, log2 h), by
    enumerating the block's transform blocks."""
    x0, y0 = int(r['mi_col']), int(r['mi_row'])
    w4, h4 = synth.BLOCK_W4H4[int(r['size'])]
    g = {}

    def visit(tx, x, y, depth, pos):
        a, b = synth.TX_L2[tx]
        split = (not (r['flags'] & abi.LF_INTRA)) and tx != 0 and depth <= 1 and \
            (int(r['split'][depth]) >> (pos[1] * 4 + pos[0])) & 1
        if split:
            sub = synth.TX_SUB[tx]
            sa, sb = synth.TX_L2[sub]
            for dy in range(2 if b >= a else 1):
                for dx in range(2 if a >= b else 1):
                    visit(sub, x + (dx << sa), y + (dy << sb), depth + 1, (pos[0] * 2 + dx, pos[1] * 2 + dy))
            return
        for yy in range(y, y + (1 << b)):
            for xx in range(x, x + (1 << a)):
                if xx < w4 and yy < h4:
                    g[(xx, yy)] = (x, y, a, b)

    a, b = synth.TX_L2[int(r['tx_y'])]
    for uy, y in enumerate(range(0, h4, 1 << b)):
        for ux, x in enumerate(range(0, w4, 1 << a)):
            visit(int(r['tx_y']), x, y, 0, (ux, uy))
    return g


def per_edge_masks(recs, p):
    rows, cols = recs.shape
    sbw, sbh = (cols + 31) >> 5, (rows + 31) >> 5
    masks = np.zeros((sbh, sbw, 640), np.uint16)
    grids = {}

    def leaf(x, y):
        r = recs[y, x]
        if not r['flags'] & abi.LF_VALID:
            return None, None
        key = (int(r['mi_col']), int(r['mi_row']))
        if key not in grids:
            grids[key] = leaf_grid(r, cols, rows)
        return r, grids[key][(x - key[0], y - key[1])]

    def setbit(region, word, bit):
        masks.reshape(-1, 640)[region][word] |= np.uint16(1 << bit)

    for y in range(rows):
        for x in range(cols):
            r, lf = leaf(x, y)
            if r is None:
                continue
            region = (y >> 5) * sbw + (x >> 5)
            lx, ly = x - int(r['mi_col']), y - int(r['mi_row'])
            inner_ok = bool(r['flags'] & abi.LF_INTRA) or not (r['flags'] & abi.LF_SKIP)
            for d in range(2):
                first = (lx == 0) if d == 0 else (ly == 0)
                starts = (lf[0] == lx) if d == 0 else (lf[1] == ly)
                if not (first or (starts and inner_ok)):
                    continue
                own = min(lf[2 + d], 2)
                nx, ny = (x - 1, y) if d == 0 else (x, y - 1)
                other = 2
                if nx >= 0 and ny >= 0:
                    _, nl = leaf(nx, ny)
                    if nl is not None:
                        other = min(nl[2 + d], 2)
                cls = min(own, other)
                pos, along = ((x & 31), (y & 31)) if d == 0 else ((y & 31), (x & 31))
                setbit(region, ((d * 32 + pos) * 3 + cls) * 2 + (along >> 4), along & 15)
    if p.mono:
        return masks
    ssx, ssy = p.ss_x, p.ss_y

    def owner(cx, cy):
        r = recs[min((cy << ssy) + ssy, rows - 1), min((cx << ssx) + ssx, cols - 1)]
        w4, h4 = synth.BLOCK_W4H4[int(r['size'])]
        ok = (r['flags'] & abi.LF_VALID) and (r['flags'] & abi.LF_HAS_UV) and \
            (w4 > ssx or int(r['mi_col']) & 1) and (h4 > ssy or int(r['mi_row']) & 1)
        return r if ok else None

    for cy in range(rows >> ssy):
        for cx in range(cols >> ssx):
            r = owner(cx, cy)
            if r is None:
                continue
            region = ((cy << ssy) >> 5) * sbw + ((cx << ssx) >> 5)
            px, py = cx - (((cx << ssx) >> 5) * 32 >> ssx), cy - (((cy << ssy) >> 5) * 32 >> ssy)
            lx, ly = cx - (int(r['mi_col']) >> ssx), cy - (int(r['mi_row']) >> ssy)
            a, b = synth.TX_L2[int(r['tx_uv'])]
            inner_ok = bool(r['flags'] & abi.LF_INTRA) or not (r['flags'] & abi.LF_SKIP)
            for d in range(2):
                first = (lx == 0) if d == 0 else (ly == 0)
                step = a if d == 0 else b
                if not (first or (inner_ok and ((lx if d == 0 else ly) % (1 << step)) == 0)):
                    continue
                own = int(step != 0)
                cls = own
                if first:
                    other = 1
                    if d == 0 and int(r['mi_col']) > 0:
                        n = recs[min((cy << ssy) + ssy, rows - 1), int(r['mi_col']) - 1]
                    elif d == 1 and int(r['mi_row']) > 0:
                        n = recs[int(r['mi_row']) - 1, min((cx << ssx) + ssx, cols - 1)]
                    else:
                        n = None
                    if n is not None and (n['flags'] & abi.LF_VALID) and (n['flags'] & abi.LF_HAS_UV):
                        other = int(synth.TX_L2[int(n['tx_uv'])][d] != 0)
                    cls = min(own, other)
                pos, along = (px, py) if d == 0 else (py, px)
                half_n = 16 >> (ssy if d == 0 else ssx)
                setbit(region, 384 + ((d * 32 + pos) * 2 + cls) * 2 + (along >= half_n),
                       along - half_n * (along >= half_n))
    return masks


def drop_blocks(rng, recs, blocks, frac=0.03):
    """Take a few whole blocks out of the map (their cells become not valid):
    no edges inside them, frame-edge context beside them, old level bytes."""
    for (x, y) in blocks:
        if rng.random() < frac:
            cells = (recs['mi_col'] == x) & (recs['mi_row'] == y) & ((recs['flags'] & abi.LF_VALID) != 0)
            recs['flags'][cells] = 0


MAP_CASES = [(200, 136, 1, 1, False), (136, 72, 1, 0, False), (96, 80, 0, 0, False), (66, 66, 1, 1, True),
             (264, 136, 1, 1, False)]


@pytest.mark.parametrize("case", MAP_CASES, ids=lambda c: f"{c[0]}x{c[1]}_ss{c[2]}{c[3]}{'_mono' if c[4] else ''}")
def test_oracle_masks_match_per_edge_formulation(case):
    w, h, ssx, ssy, mono = case
    rng = np.random.default_rng(w * 7 + h)
    recs, blocks = synth.record_map(rng, w, h, ssx, ssy, mono)
    drop_blocks(rng, recs, blocks)
    p = synth.lf_build_params(rng, w, h, ssx, ssy, mono)
    got, _ = orc.deblock_build(recs, p)
    want = per_edge_masks(recs, p)
    assert np.array_equal(got, want)
    assert got.any()


# ---------------------------------------------------------------- GPU parity
def gpu_build(runner, recs, p, lc_init):
    import torch
    d_rec = torch.from_numpy(recs.view(np.uint8).reshape(-1).copy()).cuda()
    d_lc = torch.from_numpy(lc_init.copy()).cuda()
    sbw, sbh = (recs.shape[1] + 31) >> 5, (recs.shape[0] + 31) >> 5
    d_m = torch.full((sbh * sbw * 640,), 0x5a5a, dtype=torch.int16).cuda()
    runner.deblock_build(d_rec.data_ptr(), p, d_lc.data_ptr(), d_m.data_ptr())
    runner.sync()
    torch.cuda.synchronize()
    return d_m.cpu().numpy().view(np.uint16).reshape(sbh, sbw, 640), d_lc.cpu().numpy()


@pytest.fixture(scope="module", params=["staged", "fused"])
def runner(request):
    """A handle per pipeline mode: every test of the module runs once with one
    launch per stage and once with the fused chain launch."""
    from goav1_amd.lib import PIPELINE_FUSED, PostFilterRunner
    r = PostFilterRunner()
    if request.param == "fused":
        r.set_pipeline(PIPELINE_FUSED)
    yield r
    r.close()


GPU_CASES = MAP_CASES + [(1920, 1080, 1, 1, False), (1280, 720, 0, 0, False)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=lambda c: f"{c[0]}x{c[1]}_ss{c[2]}{c[3]}{'_mono' if c[4] else ''}")
def test_deblock_build_parity(runner, case):
    w, h, ssx, ssy, mono = case
    rng = np.random.default_rng(1000 + w + h)
    recs, blocks = synth.record_map(rng, w, h, ssx, ssy, mono, delta_lf=True)
    drop_blocks(rng, recs, blocks)
    p = synth.lf_build_params(rng, w, h, ssx, ssy, mono, multi=bool(w & 8))
    lc0 = rng.integers(0, 256, size=recs.size * 4).astype(np.uint8)
    print(f"deblock build {w}x{h} ss{ssx}{ssy} mono={mono}", flush=True)
    want_m, want_lc = orc.deblock_build(recs, p, lc0)
    got_m, got_lc = gpu_build(runner, recs, p, lc0)
    assert np.array_equal(got_m, want_m)
    assert np.array_equal(got_lc, want_lc)


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_deblock_from_device_built_side_data(runner, bd):
    """Deblocking with side data built on the device equals the oracle's
    deblocking with the oracle-built side data."""
    import torch
    from goav1_amd.lib import DeviceFrame, DeviceSideData
    w, h = 264, 200
    rng = np.random.default_rng(31 + bd)
    fr = synth.FrameData(w, h, bd, 1, 1)
    synth.fill_blocky(fr, rng)
    recs, _ = synth.record_map(rng, w, h, 1, 1)
    p = synth.lf_build_params(rng, w, h, 1, 1, seg=False)
    masks, lc = orc.deblock_build(recs, p)
    cols, rows = p.mi_cols, p.mi_rows
    fr.deblock = dict(level_y=(p.level_y[0], p.level_y[1]), level_u=p.level_u, level_v=p.level_v, sharpness=3,
                      mi_cols=cols, mi_rows=rows, sb128_w=(cols + 31) >> 5, sb128_h=(rows + 31) >> 5,
                      level_cache=lc, masks=np.ascontiguousarray(masks.reshape(-1)))
    fr.stages = abi.STAGE_DEBLOCK
    want, _ = orc.run_chain(fr)
    d = DeviceFrame(fr)
    side = DeviceSideData(fr)
    # replace the uploaded side data with the device-built one
    d_rec = torch.from_numpy(recs.view(np.uint8).reshape(-1).copy()).cuda()
    d_lc = torch.zeros(rows * cols * 4, dtype=torch.uint8).cuda()
    d_m = torch.zeros(masks.size, dtype=torch.int16).cuda()
    runner.deblock_build(d_rec.data_ptr(), p, d_lc.data_ptr(), d_m.data_ptr())
    side.req.deblock.contents.level_cache = d_lc.data_ptr()
    side.req.deblock.contents.masks = d_m.data_ptr()
    runner.apply(d, side)
    runner.sync()
    torch.cuda.synchronize()
    got = fr.copy_pixels()
    for pl, a in zip(got.planes, d.to_numpy()):
        pl.pix[:] = a
    assert orc.plane_md5(got) == orc.plane_md5(want)
    assert orc.plane_md5(want) != orc.plane_md5(fr)   # the edges were filtered
