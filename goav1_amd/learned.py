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

Frames: moving-rectangle 8-bit 4:2:0 sources (the 1080p sequence recipe at
8 bits) with reconstructions from 8x8 DCT dead-zone quantisation at a random
qindex in 60..220.
"""
import argparse
import json
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


def frame_pair(seed, w, h, q):
    """(src, rec, noskip, damping) for a moving-rectangle 8-bit source."""
    rng = np.random.default_rng(seed)
    src = synth.FrameData(w, h, 8, 1, 1)
    synth.fill_rectangles(src, rng, synth.make_rects(rng, w, h), int(rng.integers(0, 8)))
    return synth.search_pair(seed=seed + 1, w=w, h=h, q=q, src=src, model="dct")


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


def build_dataset(runner, net, n_frames, seed=1000, w=1280, h=720, log=None, temp=TEMP):
    """Tiles and targets of n_frames frames from the exhaustive device search."""
    import torch
    tiles, tg, counts = [], [], np.zeros((2, 64), np.int64)
    for f in range(n_frames):
        q = int(np.random.default_rng(seed + 7 * f).integers(60, 221))
        src, rec, noskip, damping = frame_pair(seed + 7 * f, w, h, q)
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
            log(f"dataset: {f + 1}/{n_frames} frames, {sum(x.shape[0] for x in tiles)} tiles")
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


def psnr(sse, npx):
    return 10 * np.log10(255.0 ** 2 * npx / max(sse, 1))


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


def evaluate(runner, net, frames, lam_for, static_cand, lr_thr=0.9, reps=3, log=None, static_rank=None):
    """PSNR-Y loss, CDEF rate and time of the learned and static-shortlist
    searches against the exhaustive search on held-out frames."""
    import torch
    stream = torch.cuda.current_stream()
    cand_static = torch.from_numpy(static_cand).cuda()
    rows = []
    t_ex = t_le = t_st = 0.0
    for fi, (src, rec, noskip, damping, q) in enumerate(frames):
        lam = lam_for(q)
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
        row["unfiltered_psnr"] = psnr(base, npx)
        for name, dec, idx, lru in (("exhaustive", ex_dec, ex_idx, ex_lru), ("learned", le_dec, le_idx, le_lru),
                                    ("static", st_dec, st_idx, st_lru)):
            sse, nb = cdef_luma_sse(full, dec, idx, act, base)
            bits = int(np.frombuffer(bytes(dec[:4]), np.int32)[0])
            row[name] = {"cdef_psnr": psnr(sse, npx), "cdef_rate_bits": nact * bits + nb * 12,
                         "lr_psnr": psnr(lr_luma_sse(lse, lru, base), npx)}
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
    for name in ("learned", "static"):
        out[f"cdef_psnr_loss_{name}"] = float(np.mean([r["exhaustive"]["cdef_psnr"] - r[name]["cdef_psnr"] for r in rows]))
        out[f"lr_psnr_loss_{name}"] = float(np.mean([r["exhaustive"]["lr_psnr"] - r[name]["lr_psnr"] for r in rows]))
        out[f"cdef_rate_delta_{name}"] = float(np.mean([r[name]["cdef_rate_bits"] - r["exhaustive"]["cdef_rate_bits"]
                                                        for r in rows]))
    return out


def torch_log(t):
    return t.clamp_min(1e-30).log()


def held_out(n, w, h, seed=900000):
    frames = []
    for f in range(n):
        q = int(np.random.default_rng(seed + 11 * f).integers(60, 221))
        src, rec, noskip, damping = frame_pair(seed + 11 * f, w, h, q)
        frames.append((src, rec, noskip, damping, q))
    return frames


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--dc-table", default=None, help="JSON with the 8-bit DC quantizer per qindex ('dc')")
    ap.add_argument("--eval-frames", type=int, default=20)
    ap.add_argument("--eval-width", type=int, default=1920)
    ap.add_argument("--eval-height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--temp", type=float, default=TEMP, help="label temperature (relative to the block's best SSE)")
    ap.add_argument("--lr", type=float, default=3e-3)
    args = ap.parse_args(argv)
    from .lib import PostFilterRunner, ProposalNet

    def log(m):
        print(m, flush=True)

    runner = PostFilterRunner()
    net = ProposalNet(runner, max_batch=1024 if args.cmd == "eval" else 256)
    if args.cmd == "eval":
        from .lib import cdef_lambda
        dc = json.load(open(args.dc_table))["dc"] if args.dc_table else None
        meta = json.load(open(os.path.splitext(args.weights)[0] + ".json"))
        net.set_params(np.load(args.weights))
        frames = held_out(args.eval_frames, args.eval_width, args.eval_height)
        rank = None
        if "label_counts" in meta:
            rank = np.argsort(-np.asarray(meta["label_counts"][0]), kind="stable")
        res = evaluate(runner, net, frames, lambda q: cdef_lambda(dc[q] if dc else 223),
                       np.array(meta["static_shortlist"], np.int32), log=log, static_rank=rank)
        log(json.dumps({k: v for k, v in res.items() if k != "rows"}))
        net.close()
        runner.close()
        return
    net.set_params(init_params(np.random.default_rng(args.seed)))
    t0 = time.time()
    tiles, targets, counts = build_dataset(runner, net, args.frames, seed=args.seed, w=args.width, h=args.height,
                                           log=log, temp=args.temp)
    ent = float(-(targets * torch_log(targets)).sum(dim=1).mean())
    log(f"dataset {tiles.shape[0]} tiles in {time.time() - t0:.1f} s; target entropy {ent:.4f} (the loss floor)")
    t0 = time.time()
    hist = train(runner, net, tiles, targets, epochs=args.epochs, lr=args.lr, log=log)
    log(f"trained in {time.time() - t0:.1f} s")
    P = net.get_params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, P)
    meta = {"frames": args.frames, "size": [args.width, args.height], "epochs": args.epochs, "loss": hist,
            "temp": args.temp, "lr": args.lr,
            "static_shortlist": static_shortlist(counts).tolist(), "label_counts": counts.tolist(), "seed": args.seed}
    json.dump(meta, open(os.path.splitext(args.out)[0] + ".json", "w"))
    log(json.dumps({k: v for k, v in meta.items() if k not in ("loss", "label_counts")}))
    net.close()
    runner.close()


if __name__ == "__main__":
    main()
