"""Encoder loop-restoration search (curated: encoder/restoration.go:149-262,
driver encoder/keyframe_lossy.go:2310-2386).

CPU tests pin the oracle's exact scores against the decoder's loop
restoration (a unit scored with SGR preset p, or with the winning Wiener
taps, must have the SSE of the decoder LR output when every unit carries
those parameters and the frame is its own deblocked and CDEF output) and
check the H-only Wiener estimate with an independent numpy formulation.
GPU tests require identical unit records and all twelve SSE entries.
"""
import numpy as np
import pytest

from goav1_amd import abi, synth
from oracle import orc

WIENER = [(0, -3, 12), (0, -5, 18), (0, -2, 8), (0, 2, 10), (1, 4, 14), (0, -1, 4)]
SGR = [(0, (8, 24)), (2, (16, 40)), (10, (0, 48)), (14, (24, 72))]
NONE_SSE = np.uint64(2**64 - 1)


def unit_rects(w, h, us=64):
    hz, vt = orc.lr_search_units(w, h, us)
    out = []
    for row in range(vt):
        for col in range(hz):
            x0, x1 = col * us, (w if col == hz - 1 else col * us + us)
            y0, y1 = row * us, (h if row == vt - 1 else row * us + us)
            if y0 > 0:
                y0 -= 8
            if y1 < h:
                y1 -= 8
            out.append((x0, x1, y0, y1))
    return out


def forced_wiener_pair(w, h, seed=5, bd=8):
    """A source of zeros against a smooth reconstruction: the reference's
    H-only estimate (no identity tap) then beats the baseline, so the exact
    Wiener re-score runs."""
    rng = np.random.default_rng(seed)
    src = synth.FrameData(w, h, bd, 1, 1)
    rec = synth.FrameData(w, h, bd, 1, 1)
    k = 1 << (bd - 8)
    yy, xx = np.mgrid[0:rec.planes[0].alloc_h, 0:rec.planes[0].alloc_w]
    ramp = k * (60 + 0.2 * xx + 0.1 * yy + rng.normal(0, 1.0, xx.shape))
    ramp[:, : xx.shape[1] // 3] = 100 * k       # a flat third: SGR cannot beat the baseline there
    rec.planes[0].pix[:] = np.clip(ramp, 0, 255 * k).astype(np.uint16)
    for p in rec.planes[1:]:
        p.pix[:] = 128 * k
    return src, rec


def seam_safe(rec):
    """The decoder reads rows y0-2, y0-2, y0-1 above an interior stripe seam
    and y1, y1+1, y1+1 below it (two saved rows each side), while the search
    reads the frame itself; with row y0-3 equal to y0-2 and y1+2 equal to y1+1
    at every seam both see the same samples."""
    L = rec.planes[0].pix
    for s in range(56, rec.h, 64):
        L[s - 3] = L[s - 2]
        if s + 2 < rec.h:
            L[s + 2] = L[s + 1]
    return rec


def decoder_lr_luma(rec, unit):
    """Oracle decoder LR on the luma plane with `unit` (16-byte record) on every
    64-unit; the frame is its own deblocked and CDEF output."""
    fr = rec.copy_pixels()
    fr.deblock = fr.cdef = fr.grain = None
    hz, vt = orc.lr_search_units(fr.w, fr.h, 64)
    units = np.zeros(hz * vt, dtype=np.dtype((np.void, 16)))
    u = np.frombuffer(bytes(unit), np.uint8)
    arr = np.tile(u, hz * vt).reshape(hz * vt, 16)
    fr.lr = dict(frame_type=[abi.LR_SWITCHABLE, abi.LR_NONE, abi.LR_NONE], unit_size=[64, 32, 32],
                 units=[np.ascontiguousarray(arr).view(np.uint8).reshape(-1, 16), None, None])
    out, _ = orc.run_chain(fr, abi.STAGE_LR)
    return out.planes[0].pix


def unit_record(kind, p):
    b = bytearray(16)
    if kind == "sgr":
        b[0] = abi.LR_SGRPROJ
        b[1] = SGR[p][0]
        b[2:4] = np.array(SGR[p][1], np.int8).tobytes()
    else:
        b[0] = abi.LR_WIENER
        b[4:7] = np.array(WIENER[p], np.int8).tobytes()
        b[7:10] = np.array(WIENER[p], np.int8).tobytes()
    return b


def sse(a, b):
    d = a.astype(np.int64) - b.astype(np.int64)
    return int((d * d).sum())


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("size", [(256, 192), (200, 136)])
def test_sgr_scores_equal_decoder_lr(size, bd):
    src, rec, _, _ = synth.search_pair(seed=7, w=size[0], h=size[1], bd=bd)
    seam_safe(rec)
    units, e = orc.lr_search(src, rec, 64)
    S = src.planes[0].pix
    for p in range(4):
        out = decoder_lr_luma(rec, unit_record("sgr", p))
        for i, (x0, x1, y0, y1) in enumerate(unit_rects(rec.w, rec.h)):
            if x1 - x0 > 64 or y1 - y0 > 64:
                assert e[i, 8 + p] == NONE_SSE
                continue
            assert int(e[i, 8 + p]) == sse(S[y0:y1, x0:x1], out[y0:y1, x0:x1]), f"unit {i} preset {p}"


@pytest.mark.parametrize("bd", [8, 10])
def test_wiener_exact_score_equals_decoder_lr(bd):
    """10-bit: the margin a candidate must beat the running best by is w*h
    scaled by the squared sample range, 16 * w*h."""
    src, rec = forced_wiener_pair(192, 128, bd=bd)
    seam_safe(rec)
    units, e = orc.lr_search(src, rec, 64)
    assert (e[:, 7] != NONE_SSE).any(), "the forced case must re-score a Wiener winner"
    S = src.planes[0].pix
    for i, (x0, x1, y0, y1) in enumerate(unit_rects(rec.w, rec.h)):
        if e[i, 7] == NONE_SSE:
            continue
        best, c = int(e[i, 0]), None     # the estimate's winner: first strict improvement by w*h
        for k in range(6):
            if int(e[i, 1 + k]) + ((x1 - x0) * (y1 - y0) << (2 * (bd - 8))) < best:
                best, c = int(e[i, 1 + k]), k
        assert c is not None
        if units[i, 0] == abi.LR_WIENER:
            assert tuple(int(v) for v in units[i, 4:7].view(np.int8)) == WIENER[c]
        out = decoder_lr_luma(rec, unit_record("wiener", c))
        assert int(e[i, 7]) == sse(S[y0:y1, x0:x1], out[y0:y1, x0:x1]), f"unit {i}"


@pytest.mark.parametrize("bd", [8, 10])
@pytest.mark.parametrize("size", [(256, 192), (203, 141)])
def test_wiener_estimate_numpy(size, bd):
    src, rec, _, _ = synth.search_pair(seed=8, w=size[0], h=size[1], bd=bd)
    _, e = orc.lr_search(src, rec, 64)
    R = rec.planes[0].pix[:rec.h, :rec.w].astype(np.int64)
    S = src.planes[0].pix[:rec.h, :rec.w].astype(np.int64)
    padded = np.pad(R, ((0, 0), (3, 3)), mode="edge")
    for c, (t0, t1, t2) in enumerate(WIENER):
        f = np.array([t0, t1, t2, -2 * (t0 + t1 + t2), t2, t1, t0])
        hs = sum(f[k] * padded[:, k:k + rec.w] for k in range(7))
        pred = np.clip((hs + 64) >> 7, 0, (1 << bd) - 1)
        d2 = (S - pred) ** 2
        for i, (x0, x1, y0, y1) in enumerate(unit_rects(rec.w, rec.h)):
            if e[i, 0] == 0:
                continue
            assert int(e[i, 1 + c]) == int(d2[y0:y1, x0:x1].sum())


@pytest.mark.parametrize("bd", [8, 10])
def test_decisions_follow_scores(bd):
    src, rec, _, _ = synth.search_pair(seed=9, w=320, h=256, bd=bd)
    units, e = orc.lr_search(src, rec, 64)
    for i, (x0, x1, y0, y1) in enumerate(unit_rects(rec.w, rec.h)):
        wh = ((x1 - x0) * (y1 - y0)) << (2 * (bd - 8))
        t = units[i, 0]
        if t == abi.LR_SGRPROJ:
            p = [s for s, _ in SGR].index(units[i, 1])
            assert int(e[i, 8 + p]) + wh < int(e[i, 0]) and e[i, 8 + p] < e[i, 7]
            # presets are taken in order, each must beat the running winner by w*h
            run, want = int(e[i, 0]), None
            for q in range(4):
                if int(e[i, 8 + q]) + wh < run:
                    run, want = int(e[i, 8 + q]), q
            assert want == p
        elif t == abi.LR_NONE:
            ok = [int(e[i, 8 + q]) + wh < int(e[i, 0]) for q in range(4) if e[i, 8 + q] != NONE_SSE]
            assert not any(ok) or x1 - x0 > 64 or y1 - y0 > 64


# ---------------------------------------------------------------- GPU parity
def gpu_lr_search(src, rec, us):
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner
    hz, vt = orc.lr_search_units(rec.w, rec.h, us)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        units = torch.zeros((hz * vt, 16), dtype=torch.uint8, device="cuda")
        e = torch.zeros((hz * vt, 12), dtype=torch.int64, device="cuda")
        r.lr_search(ds, dr, us, units.data_ptr(), e.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        return units.cpu().numpy(), e.cpu().numpy().view(np.uint64)
    finally:
        r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", [("search", 1280, 720, 64, 8), ("search", 200, 136, 64, 8), ("search", 256, 192, 32, 8),
                                  ("search", 640, 360, 128, 8), ("wiener", 192, 128, 64, 8),
                                  ("search", 1280, 720, 64, 10), ("search", 200, 136, 64, 10),
                                  ("wiener", 192, 128, 64, 10)], ids=lambda c: f"{c[0]}_{c[1]}x{c[2]}_u{c[3]}_{c[4]}b")
def test_gpu_lr_search_parity(case):
    kind, w, h, us, bd = case
    if kind == "search":
        src, rec, _, _ = synth.search_pair(seed=50, w=w, h=h, bd=bd)
    else:
        src, rec = forced_wiener_pair(w, h, bd=bd)
    print(f"lr search {case}", flush=True)
    gu, ge = gpu_lr_search(src, rec, us)
    wu, we = orc.lr_search(src, rec, us)
    np.testing.assert_array_equal(ge, we)
    np.testing.assert_array_equal(gu, wu)


# ---------------------------------------------------------------- exhaustive superset mode
def decoder_lr_units(rec, recs):
    """Oracle decoder LR on the luma plane with per-unit 16-byte records
    (64-unit grid); the frame is its own deblocked and CDEF output."""
    fr = rec.copy_pixels()
    fr.deblock = fr.cdef = fr.grain = None
    fr.lr = dict(frame_type=[abi.LR_SWITCHABLE, abi.LR_NONE, abi.LR_NONE], unit_size=[64, 32, 32],
                 units=[np.ascontiguousarray(recs, np.uint8).reshape(-1, 16), None, None])
    out, _ = orc.run_chain(fr, abi.STAGE_LR)
    return out.planes[0].pix


def wiener_record(v, h):
    b = bytearray(16)
    b[0] = abi.LR_WIENER
    b[4:7] = np.array(v, np.int8).tobytes()
    b[7:10] = np.array(h, np.int8).tobytes()
    return np.frombuffer(bytes(b), np.uint8)


def sgr_record(s, xqd):
    b = bytearray(16)
    b[0] = abi.LR_SGRPROJ
    b[1] = s
    b[2:4] = np.array(xqd, np.int8).tobytes()
    return np.frombuffer(bytes(b), np.uint8)


def test_exhaustive_scores_equal_decoder_lr():
    """The fitted Wiener filter and the fitted projection of several SGR sets,
    put on every unit, give the decoder LR SSE the search reports."""
    src, rec, _, _ = synth.search_pair(seed=71, w=256, h=192, model="dct", q=150)
    seam_safe(rec)
    units, e, fit = orc.lr_search_exhaustive(src, rec, 64)
    S = src.planes[0].pix
    rects = unit_rects(rec.w, rec.h)
    cases = [("wiener", None)] + [("sgr", s) for s in (0, 5, 11, 14)]
    for kind, s in cases:
        recs = np.zeros((len(rects), 16), np.uint8)
        for i in range(len(rects)):
            recs[i] = wiener_record(fit[i, 0:3], fit[i, 3:6]) if kind == "wiener" else \
                sgr_record(s, fit[i, 6 + 2 * s:8 + 2 * s])
        out = decoder_lr_units(rec, recs)
        for i, (x0, x1, y0, y1) in enumerate(rects):
            if x1 - x0 > 64 or y1 - y0 > 64 or e[i, 0] == 0:
                continue
            k = 7 if kind == "wiener" else 12 + s
            assert int(e[i, k]) == sse(S[y0:y1, x0:x1], out[y0:y1, x0:x1]), f"{kind} {s} unit {i}"


def test_exhaustive_contains_curated():
    """The curated candidates keep their exact scores, and the exhaustive
    decision's SSE is never above the curated decision's."""
    for seed, model in ((72, "noise"), (73, "dct")):
        src, rec, _, _ = synth.search_pair(seed=seed, w=320, h=256, model=model, q=140)
        cu, ce = orc.lr_search(src, rec, 64)
        xu, xe, _ = orc.lr_search_exhaustive(src, rec, 64)
        np.testing.assert_array_equal(xe[:, 0], ce[:, 0])
        np.testing.assert_array_equal(xe[:, 8:12], ce[:, 8:12])
        for i in range(len(ce)):
            if ce[i, 7] != NONE_SSE:   # the curated exact Wiener re-score is one of the six
                assert ce[i, 7] in set(int(v) for v in xe[i, 1:7])

        def chosen(units, e, i, ex):
            t = units[i, 0]
            if t == abi.LR_NONE:
                return int(e[i, 0])
            if t == abi.LR_WIENER:
                return min(int(v) for v in (e[i, 1:8] if ex else e[i, 7:8]))
            return min(int(v) for v in (e[i, 8:28] if ex else e[i, 8:12]))
        for i in range(len(ce)):
            assert chosen(xu, xe, i, True) <= chosen(cu, ce, i, False), i


def bordered_unit(rec, x0, x1, y0, y1):
    R = rec.planes[0].pix[:rec.h, :rec.w].astype(np.int64)
    yy = np.clip(np.arange(y0 - 3, y1 + 3), 0, rec.h - 1)
    xx = np.clip(np.arange(x0 - 3, x1 + 3), 0, rec.w - 1)
    return R[yy][:, xx]


def test_exhaustive_fits_match_numpy_least_squares():
    """The fitted SGR projections equal numpy's least-squares solution of the
    same objective (rounded), and the fitted Wiener taps equal an independent
    numpy run of the alternating solves; both up to clamping and exact .5
    ties."""
    import ctypes as C
    src, rec, _, _ = synth.search_pair(seed=74, w=192, h=128, model="dct", q=120)
    units, e, fit = orc.lr_search_exhaustive(src, rec, 64)
    S = src.planes[0].pix[:rec.h, :rec.w].astype(np.int64)
    lib = orc.lib()
    lib.orc_sgr_filters.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    radius = [(2, 1)] * 10 + [(0, 1)] * 4 + [(2, 0)] * 2
    tap_lo, tap_hi = (-5, -23, -17), (10, 8, 46)
    checked = 0
    for i, (x0, x1, y0, y1) in enumerate(unit_rects(rec.w, rec.h)):
        w, h = x1 - x0, y1 - y0
        if w > 64 or h > 64 or e[i, 0] == 0:
            continue
        B = bordered_unit(rec, x0, x1, y0, y1)
        U = B[3:3 + h, 3:3 + w]
        Su = S[y0:y1, x0:x1]
        bu = np.ascontiguousarray(B, np.uint16)
        for s in range(16):
            f0 = np.zeros(w * h, np.int32)
            f1 = np.zeros(w * h, np.int32)
            lib.orc_sgr_filters(bu.ctypes.data + 2 * (3 * bu.shape[1] + 3), bu.shape[1], w, h, s, 8,
                                f0.ctypes.data, f1.ctypes.data)
            uu = (U << 4).ravel()
            cols = [c for c, r in zip((f0 - uu, f1 - uu), radius[s]) if r]
            A = np.stack(cols, axis=1).astype(np.float64)
            t = (((Su << 4).ravel() - uu) << 7).astype(np.float64)
            sol = np.linalg.lstsq(A, t, rcond=None)[0]
            r0, r1 = radius[s]
            xq = [0.0, 0.0]
            if r0 and r1:
                xq = list(sol)
            elif r0:
                xq[0] = sol[0]
            else:
                xq[1] = sol[0]
            if any(abs((v % 1.0) - 0.5) < 1e-6 for v in xq):
                continue
            q0, q1 = int(np.floor(xq[0] + 0.5)), int(np.floor(xq[1] + 0.5))
            if r0 and r1:
                want0 = min(max(q0, -96), 31)
                want = (want0, min(max(128 - want0 - q1, -32), 95))
            elif r0:
                want0 = min(max(q0, -96), 31)
                want = (want0, min(max(128 - want0, -32), 95))
            else:
                want = (0, min(max(128 - q1, -32), 95))
            assert tuple(int(v) for v in fit[i, 6 + 2 * s:8 + 2 * s]) == want, (i, s, xq)
            checked += 1
        # Wiener: the alternating solves in numpy
        def solve_dir(fixed, horiz):
            ft = np.array([fixed[0], fixed[1], fixed[2], 128 - 2 * sum(fixed), fixed[2], fixed[1], fixed[0]], np.int64)
            Y = np.zeros((7, h, w), np.int64)
            for k in range(7):
                for q in range(7):
                    if horiz:
                        Y[k] += ft[q] * B[q:q + h, k:k + w]
                    else:
                        Y[k] += ft[q] * B[k:k + h, q:q + w]
            z = np.stack([Y[0] + Y[6] - 2 * Y[3], Y[1] + Y[5] - 2 * Y[3], Y[2] + Y[4] - 2 * Y[3]]).reshape(3, -1)
            tt = ((Su << 14) - 128 * Y[3]).ravel()
            M = (z.astype(np.float64) @ z.T.astype(np.float64))
            r = z.astype(np.float64) @ tt.astype(np.float64)
            x = np.linalg.solve(M, r)
            return [int(min(max(np.floor(x[k] + 0.5), tap_lo[k]), tap_hi[k])) for k in range(3)]
        v, hh = [0, 0, 0], [0, 0, 0]
        for _ in range(2):
            hh = solve_dir(v, True)
            v = solve_dir(hh, False)
        assert list(fit[i, 0:3]) == v and list(fit[i, 3:6]) == hh, (i, v, hh, fit[i, :6])
        checked += 1
    assert checked > 50


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(1280, 720, 64, "noise"), (200, 136, 64, "dct"), (256, 192, 32, "dct"),
                                  (640, 360, 128, "dct")])
def test_gpu_lr_search_exhaustive_parity(case):
    import torch
    from goav1_amd.lib import DeviceFrame, PostFilterRunner
    w, h, us, model = case
    src, rec, _, _ = synth.search_pair(seed=51, w=w, h=h, model=model, q=130)
    hz, vt = orc.lr_search_units(rec.w, rec.h, us)
    print(f"exhaustive lr search {case}", flush=True)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        units = torch.zeros((hz * vt, 16), dtype=torch.uint8, device="cuda")
        e = torch.zeros((hz * vt, 28), dtype=torch.int64, device="cuda")
        f = torch.zeros((hz * vt, 38), dtype=torch.int8, device="cuda")
        r.lr_search_exhaustive(ds, dr, us, units.data_ptr(), e.data_ptr(), f.data_ptr())
        r.sync()
        torch.cuda.synchronize()
        gu, ge, gf = units.cpu().numpy(), e.cpu().numpy().view(np.uint64), f.cpu().numpy()
    finally:
        r.close()
    wu, we, wf = orc.lr_search_exhaustive(src, rec, us)
    np.testing.assert_array_equal(gf, wf)
    np.testing.assert_array_equal(ge, we)
    np.testing.assert_array_equal(gu, wu)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(1280, 720, 160, 0), (200, 136, 120, 1)])
def test_gpu_encoder_post_search_flow(case):
    """The encoder's filters after its searches, all on the device: save the
    deblocked LR context rows, apply the searched CDEF from the device decision
    (av1pf_cdef_search_apply), search LR on the CDEF output and apply the
    device unit records (av1pf_lr) -- equal to the decoder chain (CDEF then LR)
    run by the oracle with the same decisions."""
    import torch
    from goav1_amd.lib import DeviceFrame, DeviceSideData, PostFilterRunner, cdef_lambda
    from test_search import in_frame_word
    w, h, q, seed = case
    src, rec, noskip, damping = synth.search_pair(seed=90 + seed, w=w, h=h, q=q)
    hz, vt = orc.lr_search_units(w, h, 64)
    print(f"encoder flow {w}x{h} q{q}", flush=True)
    r = PostFilterRunner()
    try:
        ds, dr = DeviceFrame(src), DeviceFrame(rec)
        ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        nfb = noskip.shape[0]
        mse = torch.zeros((2, nfb, 64), dtype=torch.int64, device="cuda")
        dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        idx = torch.zeros((nfb,), dtype=torch.uint8, device="cuda")
        units = torch.zeros((hz * vt, 16), dtype=torch.uint8, device="cuda")
        lrf = rec.copy_pixels()
        lrf.deblock = lrf.cdef = lrf.grain = None
        lrf.lr = dict(frame_type=[abi.LR_SWITCHABLE, abi.LR_NONE, abi.LR_NONE], unit_size=[64, 32, 32],
                      units=[np.zeros((hz * vt, 16), np.uint8), None, None])
        side = DeviceSideData(lrf, stages=abi.STAGE_LR)
        side.req.lr.contents.units[0] = units.data_ptr()
        r.cdef_search(ds, dr, ns.data_ptr(), damping, mse.data_ptr())
        r.cdef_search_select(dr, ns.data_ptr(), mse.data_ptr(), cdef_lambda(synth.search_dc_quant(q)), dec.data_ptr(),
                             idx.data_ptr())
        r.lr_save_boundaries(dr, side, False)
        r.cdef_search_apply(dr, ns.data_ptr(), damping, dec.data_ptr(), idx.data_ptr())
        r.lr_search(ds, dr, 64, units.data_ptr())
        r.lr(dr, side)
        r.sync()
        torch.cuda.synchronize()
        got = rec.copy_pixels()
        for pl, a in zip(got.planes, dr.to_numpy()):
            pl.pix[:] = a
        d, fbi, u = dec.cpu().numpy(), idx.cpu().numpy(), units.cpu().numpy()
    finally:
        r.close()
    bits = int(d[:4].view(np.int32)[0])
    index = np.array([int(fbi[k]) if int(noskip[k]) & in_frame_word(rec, k) else -1 for k in range(nfb)], np.int8)
    ref = rec.copy_pixels()
    ref.deblock = ref.grain = None
    ref.cdef = dict(damping=damping, strength_count=1 << bits, y_strength=d[4:12].copy(), uv_strength=d[12:20].copy(),
                    unit_cols=(w + 63) // 64, unit_rows=(h + 63) // 64, index=index, noskip=np.ascontiguousarray(noskip))
    ref.lr = dict(frame_type=[abi.LR_SWITCHABLE, abi.LR_NONE, abi.LR_NONE], unit_size=[64, 32, 32],
                  units=[np.ascontiguousarray(u), None, None])
    want, _ = orc.run_chain(ref, stages=abi.STAGE_CDEF | abi.STAGE_LR)
    assert orc.plane_md5(got) == orc.plane_md5(want)
    assert (u[:, 0] != abi.LR_NONE).any()   # the LR stage had units to apply
