"""CPU reference of the learned-proposal network (test infrastructure only).

A numpy restatement of the library's CNN (conv 3x3/2 x3 -> global average
pool -> fc -> three softmax heads) with the same parameter layout and the
same BF16 rounding points as the device passes (GEMM operands rounded to
BF16, FP32 accumulation, activations stored in BF16), so forward outputs and
gradients can be compared with tight tolerances.  Used by tests only.
"""
import numpy as np

LAYERS = [(18, 16), (144, 32), (288, 64), (64, 128), (128, 131)]   # (K, N)
HEADS = [(0, 64), (64, 128), (128, 131)]


def param_count():
    return sum(k * n + n for k, n in LAYERS)


def split(P):
    out, o = [], 0
    for k, n in LAYERS:
        W = P[o:o + k * n].reshape(k, n)
        o += k * n
        b = P[o:o + n]
        o += n
        out.append((W, b))
    return out


EXACT = [False]   # True: no BF16 rounding, float64 (finite-difference checks)


def bf(x):
    """Round float32 to the nearest BF16 (ties to even), returned as float32."""
    if EXACT[0]:
        return np.asarray(x, np.float64)
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def init_params(rng):
    P = []
    for k, n in LAYERS:
        P.append((rng.standard_normal((k, n)) * np.sqrt(2.0 / k)).astype(np.float32).ravel())
        P.append(np.zeros(n, np.float32))
    return np.concatenate(P)


def im2col(x):
    n, H, W, C = x.shape
    Ho, Wo = H // 2, W // 2
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    cols = np.empty((n, Ho, Wo, 9, C), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx, :] = xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :]
    return cols.reshape(n * Ho * Wo, 9 * C)


def col2im(dcols, shape):
    n, H, W, C = shape
    Ho, Wo = H // 2, W // 2
    d = dcols.reshape(n, Ho, Wo, 9, C)
    dx = np.zeros((n, H + 2, W + 2, C), dcols.dtype)
    for ky in range(3):
        for kx in range(3):
            dx[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :] += d[:, :, :, ky * 3 + kx, :]
    return dx[:, 1:H + 1, 1:W + 1, :]


def forward(P, tiles):
    """tiles: (n, 64, 64, 2) float32 already BF16-valued. Returns (logits, cache)."""
    L = split(P)
    x = tiles
    cache = {"cols": [], "act": []}
    for l in range(3):
        W, b = L[l]
        n, H = x.shape[0], x.shape[1]
        cols = im2col(x)
        y = np.maximum(cols @ bf(W) + b, 0)
        x = bf(y).reshape(n, H // 2, H // 2, W.shape[1])
        cache["cols"].append(cols)
        cache["act"].append(x)
    pool = bf(x.reshape(x.shape[0], -1, 64).mean(axis=1))
    W3, b3 = L[3]
    fc1 = bf(np.maximum(pool @ bf(W3) + b3, 0))
    W4, b4 = L[4]
    logits = fc1 @ bf(W4) + b4
    cache.update(pool=pool, fc1=fc1)
    return (logits if EXACT[0] else logits.astype(np.float32)), cache


def loss_and_dlogits(logits, target):
    n = logits.shape[0]
    d = np.zeros_like(logits)
    loss = 0.0
    for a, b in HEADS:
        z = logits[:, a:b]
        z = z - z.max(axis=1, keepdims=True)
        lse = np.log(np.exp(z).sum(axis=1, keepdims=True))
        p = np.exp(z - lse)
        d[:, a:b] = (p - target[:, a:b]) / n
        loss -= (target[:, a:b] * (z - lse)).sum() / n
    return loss, d


def backward(P, tiles, target):
    """Returns (loss, grads) with the device's rounding points."""
    L = split(P)
    logits, c = forward(P, tiles)
    n = tiles.shape[0]
    loss, dlog = loss_and_dlogits(logits, target)
    g = [None] * 5
    dlogb = bf(dlog)
    W4 = bf(L[4][0])
    g[4] = (c["fc1"].T @ dlogb, dlog.sum(axis=0))
    dfc1 = (dlogb @ W4.T) * (c["fc1"] > 0)
    dfc1b = bf(dfc1)
    W3 = bf(L[3][0])
    g[3] = (c["pool"].T @ dfc1b, dfc1.sum(axis=0))
    dpool = dfc1b @ W3.T
    act = c["act"]
    dact = bf(np.repeat(dpool[:, None, :], 64, axis=1).reshape(act[2].shape) / 64 * (act[2] > 0))
    shapes = [(n, 64, 64, 2), (n, 32, 32, 16), (n, 16, 16, 32)]
    for l in (2, 1, 0):
        cols = c["cols"][l]
        dY = dact.reshape(cols.shape[0], -1)
        g[l] = (cols.T @ dY, dY.sum(axis=0))
        if l == 0:
            break
        dcols = dY @ bf(L[l][0]).T
        dact = bf(col2im(dcols, shapes[l]) * (act[l - 1] > 0))
    grads = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in g])
    return loss, (grads if EXACT[0] else grads.astype(np.float32))
