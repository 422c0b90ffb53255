"""Learned proposal for the encoder CDEF/LR search: data, training, evaluation.

The network (csrc/nn.hip) reads 64x64 luma tiles (the reconstruction and the
amplified coding error) and predicts, per filter block, a distribution over
the 64 luma strengths, one over the 64 chroma strengths and the
loop-restoration unit type.  Labels come from the exhaustive device search
itself:
  - strengths: a soft minimum of the block's exhaustive SSE row (probability
    proportional to exp(-(sse - min) / T), T = max(1, temp * minimum));
  - LR type: the curated search's decision for the co-located 64x64 unit.
Training runs the library's forward/backward/Adam kernels.  The learned
search (av1pf_search_learned) turns the predictions into a frame shortlist of
K + K strengths (K = AV1PF_SHORTLIST) evaluated exactly, and skips LR units
predicted NONE.

    python -m goav1_amd.learned train --frames 300 --out goav1_amd/models/proposal_net.npy

Frames: moving-rectangle 4:2:0 sources (the 1080p sequence recipe) at
1920x1080 and 3840x2160, 8- and 10-bit, with reconstructions from 8x8 DCT
dead-zone quantisation at a random qindex in 60..220.  One network serves
both depths (its inputs are divided by 2^bd - 1).  10-bit searches are an
extension of goav1's 8-bit encoder: the RD lambda is the 8-bit one times 16
(SSE grows with the squared sample range).

    python -m goav1_amd.learned eval --dc-table tests/golden/dc_qlookup_8bit.json --eval-frames 50
"""
import argparse
import json
import multiprocessing
import os
import time

import numpy as np

from . import abi, synth
from .abi import SHORTLIST as K

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "models", "proposal_net.npy")
LEVEL1 = [4 * g if g < 16 else 4 * ((g - 16) // 3) + (g - 16) % 3 + 1 for g in range(64)]
LEVEL1_INV = {p: g for g, p in enumerate(LEVEL1)}
TEMP = 0.0001   # label temperature relative to the block's smallest SSE


LAYERS = [(18, 16), (144, 32), (288, 64), (64, 128), (128, 131)]   # (K, N) per layer, as the library lays them out


def init_params(rng):
    """He-initialised weights, zero biases, in the library's flat layout."""
    P = []
    for k, n in LAYERS:
        P.append((rng.standard_normal((k, n)) * np.sqrt(2.0 / k)).astype(np.float32).ravel())
        P.append(np.zeros(n, np.float32))
    return np.concatenate(P)


# the search workloads: (width, height, bit depth)
CELLS = [(1920, 1080, 8), (1920, 1080, 10), (3840, 2160, 8), (3840, 2160, 10)]


def frame_pair(seed, w, h, q, bd=8):
    """(src, rec, noskip, damping) for a moving-rectangle source."""
    rng = np.random.default_rng(seed)
    src = synth.FrameData(w, h, bd, 1, 1)
    synth.fill_rectangles(src, rng, synth.make_rects(rng, w, h), int(rng.integers(0, 8)))
    return synth.search_pair(seed=seed + 1, w=w, h=h, q=q, src=src, model="dct")


def lambda_for(dc, q, bd):
    """RD lambda of the search at qindex q: goav1's full-search lambda from the
    8-bit DC quantiser, times 16 for 10-bit frames (whose SSEs are 16 times
    larger for the same picture)."""
    from .lib import cdef_lambda
    return cdef_lambda(dc[q] if dc else 223) << (2 * (bd - 8))


def _make(spec):
    seed, w, h, q, bd = spec
    src, rec, noskip, damping = frame_pair(seed, w, h, q, bd)
    return src, rec, noskip, damping, q


def pairs(specs, workers=None):
    """Frame pairs of (seed, w, h, q, bd) specs, in order.  With workers > 1 a
    process pool makes them; call it before this process touches the GPU (the
    pool forks) and consume the iterator afterwards."""
    if workers and workers > 1:
        pool = multiprocessing.get_context("fork").Pool(workers)
        it = pool.imap(_make, specs, chunksize=1)

        def gen():
            try:
                for x in it:
                    yield x
            finally:
                pool.terminate()
        return gen()
    return (_make(sp) for sp in specs)


def train_specs(n_frames, seed=1000):
    """Training frames over the four cells: 1080p and 4K in proportion 4 : 1
    (about as many filter blocks of each size), 8- and 10-bit alternately."""
    specs = []
    for f in range(n_frames):
        w, h = (3840, 2160) if f % 5 == 4 else (1920, 1080)
        bd = 10 if ((f // 5) if w == 3840 else f) % 2 else 8
        sd = seed + 7 * f
        specs.append((sd, w, h, int(np.random.default_rng(sd).integers(60, 221)), bd))
    return specs


def held_out_specs(n, w, h, bd, seed=900000):
    out = []
    for f in range(n):
        sd = seed + 11 * f + (bd << 20) + (w << 8)
        out.append((sd, w, h, int(np.random.default_rng(sd).integers(60, 221)), bd))
    return out


def geometry(w, h):
    mi_cols, mi_rows = synth.aligned8(w) >> 2, synth.aligned8(h) >> 2
    nhfb, nvfb = (mi_cols + 15) // 16, (mi_rows + 15) // 16
    hz, vt = max((w + 32) // 64, 1), max((h + 32) // 64, 1)
    return mi_cols, mi_rows, nhfb, nvfb, hz, vt


def active_blocks(w, h, noskip):
    """1 where the filter block has a non-skip 8x8 block inside the frame."""
    mi_cols, mi_rows, nhfb, nvfb, _, _ = geometry(w, h)
    out = np.zeros(nhfb * nvfb, bool)
    for fb in range(nhfb * nvfb):
        fr_, fc_ = divmod(fb, nhfb)
        nhb, nvb = min(16, mi_cols - 16 * fc_), min(16, mi_rows - 16 * fr_)
        valid = 0
        for r in range(0, nvb, 2):
            for c in range(0, nhb, 2):
                valid |= 1 << ((r >> 1) * 8 + (c >> 1))
        out[fb] = (int(noskip[fb]) & valid) != 0
    return out


class Device:
    """Device buffers of one frame for the exhaustive and learned searches."""

    def __init__(self, src, rec, noskip):
        import torch
        from .lib import DeviceFrame
        self.src, self.rec = src, rec
        self.ds, self.dr = DeviceFrame(src), DeviceFrame(rec)
        self.ns = torch.from_numpy(noskip.view(np.int64)).cuda()
        _, _, nhfb, nvfb, hz, vt = geometry(rec.w, rec.h)
        self.nfb, self.nhfb, self.nvfb, self.hz, self.vt = nhfb * nvfb, nhfb, nvfb, hz, vt
        self.mse = torch.zeros((2, self.nfb, 64), dtype=torch.int64, device="cuda")
        self.msub = torch.zeros((2, self.nfb, K), dtype=torch.int64, device="cuda")
        self.dec = torch.zeros(24, dtype=torch.uint8, device="cuda")
        self.idx = torch.zeros(self.nfb, dtype=torch.uint8, device="cuda")
        self.lru = torch.zeros((hz * vt, 16), dtype=torch.uint8, device="cuda")
        self.lse = torch.zeros((hz * vt, 12), dtype=torch.int64, device="cuda")
        self.cand = torch.zeros(2 * K, dtype=torch.int32, device="cuda")
        self.conf = torch.zeros(1, dtype=torch.float32, device="cuda")


_SIDE = []


def side_stream():
    import torch
    if not _SIDE:
        _SIDE.append(torch.cuda.Stream())
    return _SIDE[0]


def run_exhaustive(runner, d, damping, lam, with_sse=True):
    """The exhaustive search with the LR search on a second stream beside the
    CDEF search and selection (they are independent), as the learned search
    runs it."""
    import torch
    cur, side = torch.cuda.current_stream(), side_stream()
    side.wait_stream(cur)
    runner.lr_search(d.ds, d.dr, 64, d.lru.data_ptr(), d.lse.data_ptr() if with_sse else None, stream=side.cuda_stream)
    runner.cdef_search(d.ds, d.dr, d.ns.data_ptr(), damping, d.mse.data_ptr())
    runner.cdef_search_select(d.dr, d.ns.data_ptr(), d.mse.data_ptr(), lam, d.dec.data_ptr(), d.idx.data_ptr())
    cur.wait_stream(side)


def run_learned(runner, net, d, damping, lam, lr_thr):
    runner.search_learned(net, d.ds, d.dr, d.ns.data_ptr(), damping, lam, lr_thr, d.dec.data_ptr(), d.idx.data_ptr(),
                          d.lru.data_ptr(), d.cand.data_ptr(), d.conf.data_ptr())


def run_shortlist(runner, d, damping, lam, cand_dev):
    import torch
    cur, side = torch.cuda.current_stream(), side_stream()
    side.wait_stream(cur)
    runner.lr_search(d.ds, d.dr, 64, d.lru.data_ptr(), stream=side.cuda_stream)
    runner.cdef_search_shortlist(d.ds, d.dr, d.ns.data_ptr(), damping, lam, cand_dev.data_ptr(), d.msub.data_ptr(),
                                 d.dec.data_ptr(), d.idx.data_ptr())
    cur.wait_stream(side)


def targets_for(mse, lr_units, w, h, act, temp=TEMP):
    """Training targets [n_active][131] of one frame."""
    _, _, nhfb, nvfb, hz, vt = geometry(w, h)
    fbs = np.nonzero(act)[0]
    T = np.zeros((len(fbs), 131), np.float32)
    for p in range(2):
        m = mse[p][fbs].astype(np.float64)
        mn = m.min(axis=1, keepdims=True)
        z = -(m - mn) / np.maximum(1.0, temp * mn)
        e = np.exp(z - z.max(axis=1, keepdims=True))
        T[:, 64 * p:64 * p + 64] = e / e.sum(axis=1, keepdims=True)
    for i, fb in enumerate(fbs):
        r, c = divmod(int(fb), nhfb)
        u = min(r, vt - 1) * hz + min(c, hz - 1)
        T[i, 128 + int(lr_units[u][0])] = 1.0
    return fbs, T


def build_dataset(runner, net, frames, log=None, temp=TEMP):
    """Tiles and targets of frame pairs (src, rec, noskip, damping, q) from
    the exhaustive device search."""
    import torch
    tiles, tg, counts = [], [], np.zeros((2, 64), np.int64)
    for f, (src, rec, noskip, damping, q) in enumerate(frames):
        w, h = rec.w, rec.h
        d = Device(src, rec, noskip)
        run_exhaustive(runner, d, damping, 1000, with_sse=False)
        t = net.tiles(d.ds, d.dr)
        runner.sync()
        torch.cuda.synchronize()
        act = active_blocks(w, h, noskip)
        fbs, T = targets_for(d.mse.cpu().numpy().view(np.uint64), d.lru.cpu().numpy(), w, h, act, temp)
        idx = torch.from_numpy(fbs.astype(np.int64)).cuda()
        tiles.append(t.index_select(0, idx))
        tg.append(torch.from_numpy(T).cuda())
        for p in range(2):
            counts[p] += np.bincount(T[:, 64 * p:64 * p + 64].argmax(axis=1), minlength=64)
        if log and (f + 1) % 25 == 0:
            log(f"dataset: {f + 1} frames, {sum(x.shape[0] for x in tiles)} tiles")
    return torch.cat(tiles), torch.cat(tg), counts


def static_shortlist(counts):
    """The baseline without the network: the K strengths per plane that are
    most often the best in the training labels (ascending)."""
    return np.concatenate([np.sort(np.argsort(-counts[p], kind="stable")[:K]) for p in range(2)]).astype(np.int32)


def train(runner, net, tiles, targets, epochs=6, batch=256, lr=2e-3, seed=0, log=None):
    import torch
    n = tiles.shape[0]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    grads = torch.zeros(net.nparams, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    history = []
    steps = 0
    total = epochs * (n // batch)
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen).cuda()
        run = 0.0
        for b in range(n // batch):
            sel = perm[b * batch:(b + 1) * batch]
            bt = tiles.index_select(0, sel).contiguous()
            bg = targets.index_select(0, sel).contiguous()
            net.backward(bt, bg, grads, loss, n=batch)
            step_lr = lr * 0.5 * (1 + np.cos(np.pi * steps / max(total, 1)))   # cosine decay
            net.adam(grads, max(step_lr, 1e-5), weight_decay=1e-4)
            steps += 1
            if b % 50 == 0:
                run = loss.item()
        history.append(run)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {run:.4f}")
    return history


def psnr(sse, npx, bd=8):
    return 10 * np.log10(float((1 << bd) - 1) ** 2 * npx / max(sse, 1))


def cdef_luma_sse(full_mse, dec, idx, act, base_sse):
    """Frame luma SSE after CDEF with a decision, from the exhaustive table:
    the unfiltered SSE minus what each active block's chosen strength saves."""
    nb = 1 << int(np.frombuffer(bytes(dec[:4]), np.int32)[0])
    gy = [LEVEL1_INV[int(v)] for v in dec[4:4 + nb]]
    s = base_sse
    for fb in np.nonzero(act)[0]:
        s += int(full_mse[0][fb][gy[int(idx[fb])]]) - int(full_mse[0][fb][0])
    return s, nb


def lr_luma_sse(lse, units, base_sse):
    """Frame luma SSE after LR with unit decisions, from the exhaustive scores."""
    s = base_sse
    for u, rec in enumerate(units):
        t = int(rec[0])
        if t == abi.LR_WIENER:
            s += int(lse[u][7]) - int(lse[u][0])
        elif t == abi.LR_SGRPROJ:
            p = [0, 2, 10, 14].index(int(rec[1]))
            s += int(lse[u][8 + p]) - int(lse[u][0])
    return s


def topk_gain(lm, order, ks=(1, 2, 4, 8)):
    """Per-block luma CDEF gain kept by the first k strengths of `order`."""
    return {k: float((lm[:, 0] - lm[:, np.asarray(order[:k])].min(axis=1)).sum()) for k in ks}


def lr_exhaustive_mode_sse(runner, d, base):
    """Frame luma SSE after LR with the exhaustive superset mode's decisions
    (8-bit): per unit the running best of its 28 scores under the curated
    rule (a candidate replaces the best when sse + w*h is below it)."""
    import torch
    n = d.hz * d.vt
    units = torch.zeros((n, 16), dtype=torch.uint8, device="cuda")
    e = torch.zeros((n, 28), dtype=torch.int64, device="cuda")
    fit = torch.zeros((n, 38), dtype=torch.int8, device="cuda")
    runner.lr_search_exhaustive(d.ds, d.dr, 64, units.data_ptr(), e.data_ptr(), fit.data_ptr())
    runner.sync()
    torch.cuda.synchronize()
    E = e.cpu().numpy().view(np.uint64)
    W, H = d.rec.w, d.rec.h
    s = base
    for u in range(n):
        col, row = u % d.hz, u // d.hz
        x0, x1 = col * 64, (W if col == d.hz - 1 else col * 64 + 64)
        y0, y1 = row * 64, (H if row == d.vt - 1 else row * 64 + 64)
        y0 -= 8 if y0 > 0 else 0
        y1 -= 8 if y1 < H else 0
        wh = (x1 - x0) * (y1 - y0)
        best = int(E[u, 0])
        for c in range(1, 28):
            if E[u, c] != np.uint64(2 ** 64 - 1) and int(E[u, c]) + wh < best:
                best = int(E[u, c])
        s += best - int(E[u, 0])
    return s


def evaluate(runner, net, frames, lam_for, static_cand, lr_thr=0.9, reps=3, log=None, static_rank=None):
    """PSNR-Y loss, CDEF rate and time of the learned and static-shortlist
    searches against the exhaustive search on held-out frames (lam_for(q, bd)
    gives the RD lambda).  On 8-bit frames the learned LR decisions are also
    compared with the exhaustive superset LR mode."""
    import torch
    stream = torch.cuda.current_stream()
    cand_static = torch.from_numpy(static_cand).cuda()
    rows = []
    t_ex = t_le = t_st = 0.0
    for fi, (src, rec, noskip, damping, q) in enumerate(frames):
        bd = rec.bd
        lam = lam_for(q, bd)
        d = Device(src, rec, noskip)
        act = active_blocks(rec.w, rec.h, noskip)
        S, R = src.planes[0].pix[:rec.h, :rec.w].astype(np.int64), rec.planes[0].pix[:rec.h, :rec.w].astype(np.int64)
        base = int(((S - R) ** 2).sum())
        npx = rec.w * rec.h

        def timed(fn):
            for _ in range(1):
                fn()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            runner.sync()
            ev0.record(stream)
            for _ in range(reps):
                fn()
            ev1.record(stream)
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1) / reps

        t_ex += timed(lambda: run_exhaustive(runner, d, damping, lam))
        full = d.mse.cpu().numpy().view(np.uint64).copy()
        lse = d.lse.cpu().numpy().view(np.uint64).copy()
        ex_dec, ex_idx, ex_lru = d.dec.cpu().numpy().copy(), d.idx.cpu().numpy().copy(), d.lru.cpu().numpy().copy()
        t_le += timed(lambda: run_learned(runner, net, d, damping, lam, lr_thr))
        le_dec, le_idx, le_lru = d.dec.cpu().numpy().copy(), d.idx.cpu().numpy().copy(), d.lru.cpu().numpy().copy()
        le_cand = d.cand.cpu().numpy().copy()
        conf = d.conf.item()
        t_st += timed(lambda: run_shortlist(runner, d, damping, lam, cand_static))
        st_dec, st_idx, st_lru = d.dec.cpu().numpy().copy(), d.idx.cpu().numpy().copy(), d.lru.cpu().numpy().copy()
        row = {"q": q, "conf": conf}
        nact = int(act.sum())
        # how often each luma shortlist holds a block's exhaustive best strength;
        # "frame_counts" lists this frame's K most frequent best strengths
        best = full[0][act].argmin(axis=1)
        counts = np.bincount(best, minlength=64)
        row["hits"] = {"learned": int(np.isin(best, le_cand[:K]).sum()),
                       "static": int(np.isin(best, static_cand[:K]).sum()),
                       "frame_counts": int(np.sort(counts)[-K:].sum()), "blocks": int(best.size)}
        # the per-block CDEF luma gain (unfiltered SSE - best SSE) a shortlist
        # keeps when every block may take its best listed strength
        lm = full[0][act].astype(np.float64)
        room = (lm[:, 0] - lm.min(axis=1)).sum()
        keep = lambda c: (lm[:, 0] - lm[:, np.asarray(c[:K])].min(axis=1)).sum()
        row["gain"] = {"room": room, "learned": keep(le_cand), "static": keep(static_cand),
                       "frame_counts": keep(np.argsort(-counts, kind="stable")[:K])}
        # rankings for shorter lists: the network's summed luma probabilities
        # over the active blocks, the training-label counts, this frame's counts
        z = net.infer(d.ds, d.dr).cpu().numpy()[act][:, :64].astype(np.float64)
        pz = np.exp(z - z.max(axis=1, keepdims=True))
        pz /= pz.sum(axis=1, keepdims=True)
        row["topk"] = {"learned": topk_gain(lm, np.argsort(-pz.sum(axis=0), kind="stable")),
                       "frame_counts": topk_gain(lm, np.argsort(-counts, kind="stable"))}
        if static_rank is not None:
            row["topk"]["static"] = topk_gain(lm, static_rank)
        row["unfiltered_psnr"] = psnr(base, npx, bd)
        for name, dec, idx, lru in (("exhaustive", ex_dec, ex_idx, ex_lru), ("learned", le_dec, le_idx, le_lru),
                                    ("static", st_dec, st_idx, st_lru)):
            sse, nb = cdef_luma_sse(full, dec, idx, act, base)
            bits = int(np.frombuffer(bytes(dec[:4]), np.int32)[0])
            row[name] = {"cdef_psnr": psnr(sse, npx, bd), "cdef_rate_bits": nact * bits + nb * 12,
                         "lr_psnr": psnr(lr_luma_sse(lse, lru, base), npx, bd)}
        if bd == 8:
            row["lr_exhaustive_mode_psnr"] = psnr(lr_exhaustive_mode_sse(runner, d, base), npx, bd)
        rows.append(row)
        if log:
            log(f"eval frame {fi}: q {q} exh {row['exhaustive']['cdef_psnr']:.3f} "
                f"learned {row['learned']['cdef_psnr']:.3f} static {row['static']['cdef_psnr']:.3f} conf {conf:.2f}")
    n = len(frames)
    out = {"frames": n, "ms_exhaustive": t_ex / n, "ms_learned": t_le / n, "ms_static": t_st / n,
           "speedup_learned": t_ex / t_le, "rows": rows}
    tot = lambda k: sum(r["hits"][k] for r in rows)
    out["luma_best_in_shortlist"] = {k: round(tot(k) / max(tot("blocks"), 1), 4)
                                     for k in ("learned", "static", "frame_counts")}
    room = sum(r["gain"]["room"] for r in rows)
    out["luma_block_gain_kept"] = {k: round(sum(r["gain"][k] for r in rows) / max(room, 1.0), 4)
                                   for k in ("learned", "static", "frame_counts")}
    out["topk_gain_kept"] = {name: {k: round(sum(r["topk"][name][k] for r in rows) / max(room, 1.0), 4)
                                    for k in (1, 2, 4, 8)} for name in rows[0]["topk"]}
    out["cdef_gain_db_exhaustive"] = float(np.mean([r["exhaustive"]["cdef_psnr"] - r["unfiltered_psnr"] for r in rows]))
    if "lr_exhaustive_mode_psnr" in rows[0]:
        out["lr_psnr_loss_vs_exhaustive_mode"] = float(np.mean([r["lr_exhaustive_mode_psnr"] - r["learned"]["lr_psnr"]
                                                                for r in rows]))
        out["lr_psnr_loss_curated_vs_exhaustive_mode"] = float(np.mean(
            [r["lr_exhaustive_mode_psnr"] - r["exhaustive"]["lr_psnr"] for r in rows]))
    for name in ("learned", "static"):
        out[f"cdef_psnr_loss_{name}"] = float(np.mean([r["exhaustive"]["cdef_psnr"] - r[name]["cdef_psnr"] for r in rows]))
        out[f"lr_psnr_loss_{name}"] = float(np.mean([r["exhaustive"]["lr_psnr"] - r[name]["lr_psnr"] for r in rows]))
        out[f"cdef_rate_delta_{name}"] = float(np.mean([r[name]["cdef_rate_bits"] - r["exhaustive"]["cdef_rate_bits"]
                                                        for r in rows]))
    return out


def torch_log(t):
    return t.clamp_min(1e-30).log()


def held_out(n, w, h, bd=8, seed=900000):
    return list(pairs(held_out_specs(n, w, h, bd, seed)))


def cell_summary(res, w, h, bd):
    """The reported figures of one cell's evaluation."""
    px = w * h
    out = {"frames": res["frames"], "width": w, "height": h, "bit_depth": bd,
           "exhaustive_mpx_s": round(px / res["ms_exhaustive"] / 1e3, 1),
           "learned_mpx_s": round(px / res["ms_learned"] / 1e3, 1),
           "static_shortlist_mpx_s": round(px / res["ms_static"] / 1e3, 1),
           "ms_exhaustive": round(res["ms_exhaustive"], 4), "ms_learned": round(res["ms_learned"], 4),
           "speedup": round(res["speedup_learned"], 2),
           "cdef_psnr_loss_db": round(res["cdef_psnr_loss_learned"], 5),
           "lr_psnr_loss_db": round(res["lr_psnr_loss_learned"], 5),
           "cdef_rate_delta_bits": round(res["cdef_rate_delta_learned"], 2),
           "static_cdef_psnr_loss_db": round(res["cdef_psnr_loss_static"], 5),
           "static_lr_psnr_loss_db": round(res["lr_psnr_loss_static"], 5),
           "cdef_gain_db_exhaustive": round(res["cdef_gain_db_exhaustive"], 4),
           "luma_block_gain_kept": res["luma_block_gain_kept"], "topk_gain_kept": res["topk_gain_kept"]}
    if "lr_psnr_loss_vs_exhaustive_mode" in res:
        out["lr_psnr_loss_vs_exhaustive_lr_mode_db"] = round(res["lr_psnr_loss_vs_exhaustive_mode"], 5)
        out["curated_lr_psnr_loss_vs_exhaustive_lr_mode_db"] = round(res["lr_psnr_loss_curated_vs_exhaustive_mode"], 5)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--dc-table", default=None, help="JSON with the 8-bit DC quantizer per qindex ('dc')")
    ap.add_argument("--eval-frames", type=int, default=50, help="held-out frames per cell")
    ap.add_argument("--eval-out", default=None, help="JSON file for the per-cell evaluation")
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--temp", type=float, default=TEMP, help="label temperature (relative to the block's best SSE)")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="processes making the frames (before the GPU is used)")
    args = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    # the frames are made by a process pool that forks before this process
    # touches the GPU, and consumed as they arrive
    if args.cmd == "eval":
        cells = [(w, h, bd, held_out_specs(args.eval_frames, w, h, bd)) for (w, h, bd) in CELLS]
        stream = pairs([sp for c in cells for sp in c[3]], args.workers)
    else:
        stream = pairs(train_specs(args.frames, args.seed), args.workers)
    from .lib import PostFilterRunner, ProposalNet
    runner = PostFilterRunner()
    net = ProposalNet(runner, max_batch=1024 if args.cmd == "eval" else 256)
    dc = json.load(open(args.dc_table))["dc"] if args.dc_table else None
    if args.cmd == "eval":
        meta = json.load(open(os.path.splitext(args.weights)[0] + ".json"))
        net.set_params(np.load(args.weights))
        rank = None
        if "label_counts" in meta:
            rank = np.argsort(-np.asarray(meta["label_counts"][0]), kind="stable")
        report = {"weights": os.path.basename(args.weights), "frames_per_cell": args.eval_frames, "cells": []}
        # every frame is made before the first timed search, so the pool's
        # processes do not compete with the timed launches for the host
        made = list(stream)
        log(f"{len(made)} held-out frames made")
        for ci, (w, h, bd, specs) in enumerate(cells):
            frames = made[ci * args.eval_frames:(ci + 1) * args.eval_frames]
            res = evaluate(runner, net, frames, lambda q, b: lambda_for(dc, q, b),
                           np.array(meta["static_shortlist"], np.int32), log=log, static_rank=rank)
            summary = cell_summary(res, w, h, bd)
            report["cells"].append(summary)
            log(json.dumps(summary))
        if args.eval_out:
            json.dump(report, open(args.eval_out, "w"), indent=1)
        net.close()
        runner.close()
        return
    net.set_params(init_params(np.random.default_rng(args.seed)))
    t0 = time.time()
    tiles, targets, counts = build_dataset(runner, net, stream, log=log, temp=args.temp)
    ent = float(-(targets * torch_log(targets)).sum(dim=1).mean())
    log(f"dataset {tiles.shape[0]} tiles in {time.time() - t0:.1f} s; target entropy {ent:.4f} (the loss floor)")
    t0 = time.time()
    hist = train(runner, net, tiles, targets, epochs=args.epochs, lr=args.lr, log=log)
    log(f"trained in {time.time() - t0:.1f} s")
    P = net.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, P)
    meta = {"frames": args.frames, "cells": [list(c) for c in CELLS], "epochs": args.epochs, "loss": hist,
            "tiles": int(tiles.shape[0]), "target_entropy": ent, "temp": args.temp, "lr": args.lr,
            "static_shortlist": static_shortlist(counts).tolist(), "label_counts": counts.tolist(), "seed": args.seed}
    json.dump(meta, open(os.path.splitext(args.out)[0] + ".json", "w"))
    log(json.dumps({k: v for k, v in meta.items() if k not in ("loss", "label_counts")}))
    net.close()
    runner.close()


if __name__ == "__main__":
    main()
