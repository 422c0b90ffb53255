"""Encoder CDEF strength search (cdef_search.go:580-945).

CPU tests pin the oracle's MSE table two independent ways (the unfiltered
trial equals a plain block SSE; every trial equals the SSE of the decoder's
CDEF at that strength, which the decoder-side tests pin) and check the
library's host finish against the oracle's finish and against invariants of
the selection.  GPU tests require the device MSE table to equal the oracle's
bit for bit and the finish decisions to be identical.
"""
import numpy as np
import pytest

from goav1_amd import abi, synth
from goav1_amd.abi import SHORTLIST as K
from oracle import orc

# search-level-1 strength list: position -> packed (primary * 4 + secondary code)
LEVEL1 = [4 * g if g < 16 else 4 * ((g - 16) // 3) + (g - 16) % 3 + 1 for g in range(64)]


def test_level1_list_is_every_pair_once():
    assert sorted(LEVEL1) == list(range(64))
    assert LEVEL1[:16] == [4 * p for p in range(16)]


def block_list(fr, noskip, fb):
    mi_cols, mi_rows = synth.aligned8(fr.w) >> 2, synth.aligned8(fr.h) >> 2
    nhfb = (mi_cols + 15) // 16
    fr_, fc_ = divmod(fb, nhfb)
    out = []
    for r in range(8):
        for c in range(8):
            if (int(noskip[fb]) >> (r * 8 + c)) & 1 and fr_ * 16 + 2 * r < mi_rows and fc_ * 16 + 2 * c < mi_cols:
                out.append((fr_ * 8 + r, fc_ * 8 + c))
    return out


def block_sse(a, b, fr, blocks):
    """SSE over 8x8 luma blocks (8x8 block units) and their 4x4 chroma blocks (U + V)."""
    sy = sc = 0
    for by, bx in blocks:
        d = a.planes[0].pix[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].astype(np.int64) - \
            b.planes[0].pix[by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].astype(np.int64)
        sy += int((d * d).sum())
        for p in (1, 2):
            d = a.planes[p].pix[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4].astype(np.int64) - \
                b.planes[p].pix[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4].astype(np.int64)
            sc += int((d * d).sum())
    return sy, sc


def active_blocks(fr, noskip):
    """1 where the filter block has a non-skip 8x8 block inside the frame."""
    return np.array([1 if block_list(fr, noskip, fb) else 0 for fb in range(noskip.shape[0])], np.uint8)


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("size", [(128, 128), (200, 136)])
def test_unfiltered_trial_is_block_sse(size, bd):
    src, rec, noskip, damping = synth.search_pair(seed=11, w=size[0], h=size[1], bd=bd)
    mse = orc.cdef_search_mse(src, rec, noskip, damping)
    for fb in range(mse.shape[1]):
        sy, sc = block_sse(src, rec, src, block_list(src, noskip, fb))
        assert (int(mse[0, fb, 0]), int(mse[1, fb, 0])) == (sy, sc), f"fb {fb}"


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("size", [(128, 128), (200, 136)])
def test_every_trial_equals_decoder_cdef_sse(size, bd):
    """The search's trial gi is the decoder's CDEF with strength LEVEL1[gi] on
    every filter block, measured over the non-skip blocks (10-bit: the
    decoder's high-bit-depth CDEF, which the chain parity tests pin)."""
    src, rec, noskip, damping = synth.search_pair(seed=12, w=size[0], h=size[1], bd=bd)
    mse = orc.cdef_search_mse(src, rec, noskip, damping)
    ucols, urows = (rec.w + 63) // 64, (rec.h + 63) // 64
    assert mse.shape[1] == ucols * urows
    idx = np.where(noskip == 0, -1, 0).astype(np.int8)
    for gi in range(64):
        fr = rec.copy_pixels()
        fr.deblock = fr.lr = fr.grain = None
        ys = np.zeros(8, np.uint8)
        ys[0] = LEVEL1[gi]
        fr.cdef = dict(damping=damping, strength_count=1, y_strength=ys, uv_strength=ys.copy(),
                       unit_cols=ucols, unit_rows=urows, index=idx, noskip=np.ascontiguousarray(noskip))
        out, _ = orc.run_chain(fr, abi.STAGE_CDEF)
        for fb in range(mse.shape[1]):
            sy, sc = block_sse(src, out, src, block_list(src, noskip, fb))
            assert (int(mse[0, fb, gi]), int(mse[1, fb, gi])) == (sy, sc), f"gi {gi} fb {fb}"


def random_tables(rng, nfb, hi):
    mse = rng.integers(0, hi, size=(2, nfb, 64)).astype(np.uint64)
    skip = (rng.random(nfb) < 0.1).astype(np.uint8)
    mse[:, skip.astype(bool), :] = 0
    return mse, skip


def test_lambda_value():
    from goav1_amd.lib import cdef_lambda
    # q160 key frame: dc 223 -> (3.3 + 0.0015*223) * 223^2 = 180740.1 -> 180740;
    # * 150 >> 7 = 211804; * 100 >> 7 = 165471
    assert cdef_lambda(synth.search_dc_quant(160), key_frame=True) == 165471
    assert cdef_lambda(1, key_frame=False, tl_factor=1) == 1


@pytest.mark.parametrize("seed", range(6))
def test_finish_matches_oracle(seed):
    from goav1_amd.lib import cdef_lambda, cdef_search_finish
    rng = np.random.default_rng(seed)
    nfb = int(rng.integers(1, 300))
    hi = [8, 1000, 1 << 20, 1 << 30, 50, 1 << 12][seed]     # small ranges force ties
    mse, skip = random_tables(rng, nfb, hi)
    for lam in (0, cdef_lambda(223), 1 << 30):
        got = cdef_search_finish(mse, skip, lam)
        want = orc.cdef_search_finish(mse, skip, lam)
        assert got[0] == want[0]
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[2], want[2])
        np.testing.assert_array_equal(got[3], want[3])


def test_finish_invariants():
    from goav1_amd.lib import cdef_search_finish
    rng = np.random.default_rng(99)
    mse, skip = random_tables(rng, 120, 1 << 16)
    inv = {p: g for g, p in enumerate(LEVEL1)}
    # a large lambda (rate * lambda still inside int64) forces one pair: the
    # lowest luma and chroma column sums
    bits, ys, uvs, idx = cdef_search_finish(mse, skip, 1 << 30)
    act = ~skip.astype(bool)
    assert bits == 0
    assert inv[int(ys[0])] == int(np.argmin(mse[0, act].sum(axis=0)))
    assert inv[int(uvs[0])] == int(np.argmin(mse[1, act].sum(axis=0)))
    # with lambda 0 every active block takes its cheapest chosen pair
    bits, ys, uvs, idx = cdef_search_finish(mse, skip, 0)
    nb = 1 << bits
    assert all(ys[nb:] == 0) and all(uvs[nb:] == 0)
    for fb in range(mse.shape[1]):
        costs = [int(mse[0, fb, inv[int(ys[g])]]) + int(mse[1, fb, inv[int(uvs[g])]]) for g in range(nb)]
        assert idx[fb] == (0 if skip[fb] else int(np.argmin(costs)))


# ---------------------------------------------------------------- GPU parity
def gpu_search(src, rec, noskip, damping):
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        mse = torch.zeros((2, nfb, 64), dtype=torch.int64, device="cuda")
        r.cdef_search(ds, dr, ns.data_ptr(), damping, mse.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        return mse.cpu().numpy().view(np.uint64)
    finally:
        r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("size", [(1280, 720, 8), (200, 136, 8), (64, 64, 8), (1280, 720, 10), (200, 136, 10)],
                         ids=lambda s: f"{s[0]}x{s[1]}_{s[2]}b")
def test_gpu_search_parity(size):
    src, rec, noskip, damping = synth.search_pair(seed=21, w=size[0], h=size[1], bd=size[2])
    print(f"search {size} damping {damping}", flush=True)
    got = gpu_search(src, rec, noskip, damping)
    want = orc.cdef_search_mse(src, rec, noskip, damping)
    np.testing.assert_array_equal(got, want)
    from goav1_amd.lib import cdef_lambda, cdef_search_finish
    lam = cdef_lambda(synth.search_dc_quant(160))
    skip = 1 - active_blocks(src, noskip)
    g = cdef_search_finish(got, skip, lam)
    w = orc.cdef_search_finish(want, skip, lam)
    assert g[0] == w[0]
    for a, b in zip(g[1:], w[1:]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(1280, 720, 160, 0), (1920, 1080, 60, 1), (200, 136, 220, 2), (64, 64, 160, 3)])
def test_gpu_select_matches_host_finish(case):
    """Device strength selection = host finish, on tables from the device search."""
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner, cdef_lambda, cdef_search_finish
    w, h, q, seed = case
    src, rec, noskip, damping = synth.search_pair(seed=40 + seed, w=w, h=h, q=q)
    print(f"select {w}x{h} q{q}", flush=True)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        mse = torch.zeros((2, nfb, 64), dtype=torch.int64, device="cuda")
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.full((nfb,), 255, dtype=torch.uint8, device="cuda")
        r.cdef_search(ds, dr, ns.data_ptr(), damping, mse.data_ptr())
        for lam in (cdef_lambda(synth.search_dc_quant(160)), 0, 1 << 30):
            r.cdef_search_select(dr, ns.data_ptr(), mse.data_ptr(), lam, dec.data_ptr(), idx.data_ptr())
            r.sync()
            torch.cuda.synchronize()
            table = mse.cpu().numpy().view(np.uint64)
            bits, ys, uvs, fbi = cdef_search_finish(table, 1 - active_blocks(src, noskip), lam)
            d = dec.cpu().numpy()
            assert int(d[:4].view(np.int32)[0]) == bits, f"lambda {lam}"
            np.testing.assert_array_equal(d[4:12], ys)
            np.testing.assert_array_equal(d[12:20], uvs)
            np.testing.assert_array_equal(idx.cpu().numpy(), fbi)
    finally:
        r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(64, 64, 3, 0.0), (640, 360, 5, 0.2), (1920, 1080, 40, 0.1), (3840, 2160, 2, 0.0),
                                  (4096, 4096, 1000, 0.3)], ids=lambda c: f"{c[0]}x{c[1]}_v{c[2]}_skip{c[3]}")
def test_gpu_select_random_tables_with_ties(case):
    """Device selection = host finish on random tables over a small value
    range (many equal gains and sums: the first-maximum / first-minimum
    tie rules decide), up to the 4096-block limit, several calls in a row on
    one handle (the selection state carries over between launches)."""
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner, cdef_search_finish
    w, h, vmax, pskip = case
    rng = np.random.default_rng(w + vmax)
    rec = synth.FrameData(w, h, 8, 1, 1)
    nhfb, nvfb = (w + 63) // 64, (h + 63) // 64
    nfb = nhfb * nvfb
    noskip = np.where(rng.random(nfb) < pskip, 0, ~0).astype(np.int64).view(np.uint64)
    act = active_blocks(rec, noskip)
    print(f"select random tables {w}x{h} ({nfb} blocks)", flush=True)
    r = PostFilterRunner()
    try:
        dr = DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.full((nfb,), 255, dtype=torch.uint8, device="cuda")
        for rep, lam in enumerate((0, 3000, 1 << 24)):
            table = rng.integers(0, vmax + 1, size=(2, nfb, 64)).astype(np.uint64)
            mse = torch.from_numpy(table.view(np.int64)).cuda()
            r.cdef_search_select(dr, ns.data_ptr(), mse.data_ptr(), lam, dec.data_ptr(), idx.data_ptr())
            r.sync()
            torch.cuda.synchronize()
            bits, ys, uvs, fbi = cdef_search_finish(table, 1 - act, lam)
            d = dec.cpu().numpy()
            assert int(d[:4].view(np.int32)[0]) == bits, f"call {rep} lambda {lam}"
            np.testing.assert_array_equal(d[4:12], ys)
            np.testing.assert_array_equal(d[12:20], uvs)
            np.testing.assert_array_equal(idx.cpu().numpy(), fbi)
    finally:
        r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(1280, 720, 160, 0, 8), (200, 136, 100, 1, 8), (1920, 1080, 220, 2, 8),
                                  (2048, 1152, 120, 3, 8), (2304, 1856, 200, 4, 8), (1280, 720, 160, 5, 10),
                                  (200, 136, 60, 6, 10), (2304, 1856, 200, 7, 10), (640, 360, 0, 8, 10),
                                  (2304, 1856, 0, 9, 10)],
                         ids=lambda c: f"{c[0]}x{c[1]}_q{c[2]}_{c[4]}b")
def test_gpu_shortlist_equals_masked_host_finish(case):
    """The shortlist search's table equals the full table's shortlisted
    columns, and its one-launch selection equals the host finish over the full
    table with every other strength priced out (shortlists ascending, with 0).
    10-bit tables of realistic frames take the 32-bit selection, noise frames
    (q 0 here) the 64-bit one."""
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner, cdef_lambda, cdef_search_finish
    w, h, q, seed, bd = case
    src, rec, noskip, damping = synth.search_pair(seed=60 + seed, w=w, h=h, q=q, bd=bd)
    if q == 0:
        # independent uniform noise: block SSEs near 4096 * 1023^2 / 6, past the
        # 32-bit selection's range, so the 64-bit selection decides
        rng = np.random.default_rng(seed)
        for fr in (src, rec):
            for p in fr.planes:
                p.pix[:] = rng.integers(0, 1 << bd, size=p.pix.shape).astype(np.uint16)
    rng = np.random.default_rng(seed)
    cand = np.concatenate([np.sort(np.r_[0, rng.choice(np.arange(1, 64), K - 1, replace=False)]) for _ in range(2)])
    print(f"shortlist {w}x{h} q{q} cand {cand.tolist()}", flush=True)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        cd = torch.from_numpy(cand.astype(np.int32)).cuda()
        msub = torch.zeros((2, nfb, K), dtype=torch.int64, device="cuda")
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.zeros(nfb, dtype=torch.uint8, device="cuda")
        lam = cdef_lambda(synth.search_dc_quant(160))
        r.cdef_search_shortlist(ds, dr, ns.data_ptr(), damping, lam, cd.data_ptr(), msub.data_ptr(), dec.data_ptr(),
                                idx.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        full = orc.cdef_search_mse(src, rec, noskip, damping)
        sub = msub.cpu().numpy().view(np.uint64)
        np.testing.assert_array_equal(sub[0], full[0][:, cand[:K]])
        np.testing.assert_array_equal(sub[1], full[1][:, cand[K:]])
        masked = np.full_like(full, np.uint64(1 << 40))
        masked[0][:, cand[:K]] = full[0][:, cand[:K]]
        masked[1][:, cand[K:]] = full[1][:, cand[K:]]
        skip = 1 - active_blocks(src, noskip)
        masked[:, skip.astype(bool), :] = 0
        bits, ys, uvs, fbi = cdef_search_finish(masked, skip, lam)
        d = dec.cpu().numpy()
        assert int(d[:4].view(np.int32)[0]) == bits
        np.testing.assert_array_equal(d[4:12], ys)
        np.testing.assert_array_equal(d[12:20], uvs)
        np.testing.assert_array_equal(idx.cpu().numpy(), fbi)
    finally:
        r.close()


@pytest.mark.gpu
def test_gpu_learned_search_is_consistent():
    """The whole learned search (untrained weights): its shortlist run again
    through the shortlist entry point gives the same decision, every chosen
    strength is a legal level-1 strength from the shortlist, and LR units are
    NONE / WIENER / SGRPROJ."""
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner, ProposalNet, cdef_lambda
    from oracle import nn_ref
    src, rec, noskip, damping = synth.search_pair(seed=70, w=640, h=360, q=140)
    r = PostFilterRunner()
    net = ProposalNet(r, max_batch=64)
    try:
        net.set_params(nn_ref.init_params(np.random.default_rng(3)))
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        lam = cdef_lambda(synth.search_dc_quant(160))
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.zeros(nfb, dtype=torch.uint8, device="cuda")
        cand = torch.zeros(2 * K, dtype=torch.int32, device="cuda")
        conf = torch.zeros(1, dtype=torch.float32, device="cuda")
        hz, vt = (640 + 32) // 64, (360 + 32) // 64
        lru = torch.zeros((hz * vt, 16), dtype=torch.uint8, device="cuda")
        print("learned search 640x360", flush=True)
        r.search_learned(net, ds, dr, ns.data_ptr(), damping, lam, 0.9, dec.data_ptr(), idx.data_ptr(),
                         lru.data_ptr(), cand.data_ptr(), conf.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        c = cand.cpu().numpy()
        assert all(np.diff(c[:K]) > 0) and all(np.diff(c[K:]) > 0) and c.min() >= 0 and c.max() < 64
        assert 0 < conf.item() <= 1
        d1 = dec.cpu().numpy().copy()
        i1 = idx.cpu().numpy().copy()
        msub = torch.zeros((2, nfb, K), dtype=torch.int64, device="cuda")
        r.cdef_search_shortlist(ds, dr, ns.data_ptr(), damping, lam, cand.data_ptr(), msub.data_ptr(), dec.data_ptr(),
                                idx.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dec.cpu().numpy(), d1)
        np.testing.assert_array_equal(idx.cpu().numpy(), i1)
        nb = 1 << int(d1[:4].view(np.int32)[0])
        allowed_y = {LEVEL1[g] for g in c[:K]}
        allowed_uv = {LEVEL1[g] for g in c[K:]}
        assert set(d1[4:4 + nb].tolist()) <= allowed_y and set(d1[12:12 + nb].tolist()) <= allowed_uv
        assert set(lru.cpu().numpy()[:, 0].tolist()) <= {0, 1, 2}
    finally:
        net.close()
        r.close()


def in_frame_word(fr, fb):
    """Bits of the 8x8 blocks of filter block fb that lie inside the frame."""
    nhfb = (fr.w + 63) // 64
    fr_, fc_ = divmod(fb, nhfb)
    rows8, cols8 = (fr.h + 7) >> 3, (fr.w + 7) >> 3
    word = 0
    for by in range(min(8, rows8 - fr_ * 8)):
        for bx in range(min(8, cols8 - fc_ * 8)):
            word |= 1 << (by * 8 + bx)
    return word


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(1280, 720, 160, 0), (200, 136, 100, 1), (1920, 1080, 220, 2), (200, 136, 160, 3)])
def test_gpu_search_apply_equals_decoder_cdef(case):
    """The encoder's CDEF apply from the device decision (no host round trip)
    equals the decoder CDEF run with that decision's parameters and the index
    map of cdef_apply.go:202-228 (a unit is read when an in-frame 8x8 block has
    residual)."""
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner, cdef_lambda
    w, h, q, seed = case
    src, rec, noskip, damping = synth.search_pair(seed=70 + seed, w=w, h=h, q=q)
    lam = cdef_lambda(synth.search_dc_quant(160)) if seed != 3 else 1 << 40
    print(f"search apply {w}x{h} q{q} lambda {lam}", flush=True)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        mse = torch.zeros((2, nfb, 64), dtype=torch.int64, device="cuda")
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.zeros((nfb,), dtype=torch.uint8, device="cuda")
        r.cdef_search(ds, dr, ns.data_ptr(), damping, mse.data_ptr())
        r.cdef_search_select(dr, ns.data_ptr(), mse.data_ptr(), lam, dec.data_ptr(), idx.data_ptr())
        r.cdef_search_apply(dr, ns.data_ptr(), damping, dec.data_ptr(), idx.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        got = rec.copy_pixels()
        for pl, a in zip(got.planes, dr.to_numpy()):
            pl.pix[:] = a
        d = dec.cpu().numpy()
        fbi = idx.cpu().numpy()
    finally:
        r.close()
    bits = int(d[:4].view(np.int32)[0])
    index = np.array([int(fbi[u]) if int(noskip[u]) & in_frame_word(rec, u) else -1 for u in range(nfb)], np.int8)
    ref = rec.copy_pixels()
    ref.cdef = dict(damping=damping, strength_count=1 << bits, y_strength=d[4:12].copy(), uv_strength=d[12:20].copy(),
                    unit_cols=(rec.w + 63) // 64, unit_rows=(rec.h + 63) // 64, index=index,
                    noskip=np.ascontiguousarray(noskip))
    want, _ = orc.run_chain(ref, stages=abi.STAGE_CDEF)
    assert orc.plane_md5(got) == orc.plane_md5(want)
    if seed != 3:
        assert orc.plane_md5(want) != orc.plane_md5(rec)   # something was filtered
