"""The learned-proposal network: CPU reference pinned by finite differences,
device forward/backward against the reference with the same BF16 rounding
points, and device training."""
import numpy as np
import pytest

from oracle import nn_ref


def rand_tiles(rng, n):
    return nn_ref.bf(rng.random((n, 64, 64, 2)).astype(np.float32))


def soft_targets(rng, n):
    t = np.zeros((n, 131), np.float32)
    for a, b in nn_ref.HEADS:
        z = rng.standard_normal((n, b - a)) * 2
        p = np.exp(z - z.max(axis=1, keepdims=True))
        t[:, a:b] = p / p.sum(axis=1, keepdims=True)
    return t


def test_param_count_matches_library():
    from goav1_amd import abi
    from goav1_amd.lib import load
    assert int(load().av1pf_nn_param_count()) == nn_ref.param_count() == 48659
    assert int(load().av1pf_search_shortlist_size()) == abi.SHORTLIST


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(5)
    P = nn_ref.init_params(rng).astype(np.float64)
    P += rng.standard_normal(P.shape) * 0.01          # nonzero biases
    tiles = rng.random((3, 64, 64, 2))
    tgt = soft_targets(rng, 3).astype(np.float64)
    nn_ref.EXACT[0] = True
    try:
        loss, g = nn_ref.backward(P, tiles, tgt)
        # probe parameters in every layer's weights and biases
        offs, o = [], 0
        for k, n in nn_ref.LAYERS:
            offs += [o + 3, o + k * n // 2, o + k * n + 1]
            o += k * n + n
        for i in offs:
            e = np.zeros_like(P)
            e[i] = 1e-5
            lp, _ = nn_ref.loss_and_dlogits(nn_ref.forward(P + e, tiles)[0], tgt)
            lm, _ = nn_ref.loss_and_dlogits(nn_ref.forward(P - e, tiles)[0], tgt)
            num = (lp - lm) / 2e-5
            assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])
    finally:
        nn_ref.EXACT[0] = False


# ---------------------------------------------------------------- device
@pytest.fixture(scope="module")
def net():
    from goav1_amd.lib import PostFilterRunner, ProposalNet
    r = PostFilterRunner()
    n = ProposalNet(r, max_batch=64)
    yield n
    n.close()
    r.close()


def to_dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def tiles_dev(tiles):
    import torch
    bits = (np.ascontiguousarray(tiles, np.float32).view(np.uint32) >> 16).astype(np.uint16)
    return torch.from_numpy(bits.view(np.int16)).cuda()


@pytest.mark.gpu
def test_forward_matches_reference(net):
    import torch
    rng = np.random.default_rng(11)
    P = nn_ref.init_params(rng)
    P[nn_ref.param_count() - 131:] = rng.standard_normal(131) * 0.1
    net.set_params(P)
    tiles = rand_tiles(rng, 70)                         # more than one batch of 64
    print("nn forward 70 tiles", flush=True)
    got = net.forward(tiles_dev(tiles)).cpu().numpy()
    torch.cuda.synchronize()
    want, _ = nn_ref.forward(P, tiles)
    err = np.abs(got - want).max() / (np.abs(want).max() + 1e-9)
    assert err < 2e-3, err


@pytest.mark.gpu
def test_backward_matches_reference(net):
    import torch
    rng = np.random.default_rng(12)
    P = nn_ref.init_params(rng)
    P += rng.standard_normal(P.shape).astype(np.float32) * 0.01
    net.set_params(P)
    tiles = rand_tiles(rng, 48)
    tgt = soft_targets(rng, 48)
    grads = torch.zeros(net.nparams, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    print("nn backward 48 tiles", flush=True)
    net.backward(tiles_dev(tiles), to_dev(tgt, torch.float32), grads, loss)
    torch.cuda.synchronize()
    wl, wg = nn_ref.backward(P, tiles, tgt)
    assert abs(loss.item() - wl) < 1e-4 * max(1.0, abs(wl))
    g = grads.cpu().numpy()
    o = 0
    for li, (k, n) in enumerate(nn_ref.LAYERS):
        for part, sz in (("W", k * n), ("b", n)):
            a, b = g[o:o + sz], wg[o:o + sz]
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
            assert rel < 1e-2, (li, part, rel)
            o += sz


@pytest.mark.gpu
def test_training_reduces_loss(net):
    """Adam steps on a learnable task (the targets depend on tile statistics)."""
    import torch
    rng = np.random.default_rng(13)
    net.set_params(nn_ref.init_params(rng))
    n = 64
    tiles = rand_tiles(rng, n) * np.linspace(0.2, 1.0, n, dtype=np.float32)[:, None, None, None]
    tiles = nn_ref.bf(tiles)
    tgt = np.zeros((n, 131), np.float32)
    lab = (np.arange(n) * 8 // n)                      # brightness class 0..7
    tgt[np.arange(n), lab] = 1
    tgt[np.arange(n), 64 + 7 - lab] = 1
    tgt[np.arange(n), 128 + lab % 3] = 1
    td, gd = tiles_dev(tiles), to_dev(tgt, torch.float32)
    grads = torch.zeros(net.nparams, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    losses = []
    print("nn training 200 steps", flush=True)
    for step in range(200):
        net.backward(td, gd, grads, loss)
        net.adam(grads, 3e-3)
        if step % 50 == 0 or step == 199:
            losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0], losses


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_tiles_match_reference(net, bd):
    """Tiles of a ragged frame pair: channel 0 the reconstruction, channel 1
    the amplified coding error, both over 2^bd - 1, edges repeated past the
    frame."""
    import torch
    from goav1_amd import synth
    from goav1_amd.lib import DeviceFrame
    src, rec, _, _ = synth.search_pair(seed=21, w=200, h=136, bd=bd)
    print(f"nn tiles 200x136 {bd}-bit", flush=True)
    got = net.tiles(DeviceFrame(src), DeviceFrame(rec)).cpu().numpy().view(np.uint16)
    torch.cuda.synchronize()
    want = np_tiles(src, rec)
    bits = (want.view(np.uint32) >> 16).astype(np.uint16)
    np.testing.assert_array_equal(got, bits)


def np_tiles(src, rec):
    """BF16-valued FP32 tiles [nfb][64][64][2] of a frame pair."""
    S = src.planes[0].pix[:rec.h, :rec.w].astype(np.float32)
    R = rec.planes[0].pix[:rec.h, :rec.w].astype(np.float32)
    nh, nv = (rec.w + 63) // 64, (rec.h + 63) // 64
    yy = np.minimum(np.arange(nv * 64), rec.h - 1)
    xx = np.minimum(np.arange(nh * 64), rec.w - 1)
    inv = np.float32(1.0 / ((1 << rec.bd) - 1))
    ch = [R[yy][:, xx] * inv, (S[yy][:, xx] - R[yy][:, xx]) * np.float32(16) * inv]
    want = np.stack([nn_ref.bf(c) for c in ch], axis=-1)            # [nv*64][nh*64][2]
    return np.ascontiguousarray(want.reshape(nv, 64, nh, 64, 2).transpose(0, 2, 1, 3, 4).reshape(nv * nh, 64, 64, 2))


@pytest.mark.gpu
@pytest.mark.parametrize("bd", [8, 10])
def test_fused_inference_matches_reference(net, bd):
    """The one-launch inference from frames against the reference forward of
    the same tiles, and against the layer-by-layer device pass, before and
    after an optimizer step (which repacks the convolution fragments)."""
    import torch
    from goav1_amd import synth
    from goav1_amd.lib import DeviceFrame
    rng = np.random.default_rng(14)
    P = nn_ref.init_params(rng)
    P[nn_ref.param_count() - 131:] = rng.standard_normal(131) * 0.1
    net.set_params(P)
    src, rec, _, _ = synth.search_pair(seed=22, w=328, h=200, bd=bd)
    ds, dr = DeviceFrame(src), DeviceFrame(rec)
    tiles = np_tiles(src, rec)
    for rnd in range(2):
        print(f"nn fused inference 328x200 {bd}-bit round {rnd}", flush=True)
        fused = net.infer(ds, dr).cpu().numpy()
        layered = net.forward(net.tiles(ds, dr)).cpu().numpy()
        torch.cuda.synchronize()
        want, _ = nn_ref.forward(net.get_params(), tiles)
        scale = np.abs(want).max() + 1e-9
        assert np.abs(fused - want).max() / scale < 2e-3
        assert np.abs(fused - layered).max() / scale < 2e-3
        if rnd == 0:
            td = tiles_dev(tiles[:32])
            tgt = soft_targets(rng, 32)
            grads = torch.zeros(net.nparams, dtype=torch.float32, device="cuda")
            loss = torch.zeros(1, dtype=torch.float32, device="cuda")
            net.backward(td, to_dev(tgt, torch.float32), grads, loss)
            net.adam(grads, 1e-2)
