"""CPU oracle of the warm-start network (test infrastructure: only tests/,
the smoke check and the benchmark's CPU leg use it; nothing in the product
imports it).

The network (six 3x3 convolutions, 3 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1
channels, ReLU between layers, zero "same" padding) has no counterpart in the
reference; its definition is include/tea_warmstart.h.  This module restates
the forward pass, the mean-squared-error loss and the backward pass in numpy
float64, with two modes:

* ``exact``: every value in float64 -- pinned by central finite differences of
  the loss (tests/test_warmstart_oracle.py) and by a direct per-pixel loop on
  a tiny image;
* ``bf16``: rounds to bfloat16 (round to nearest even) at the points where the
  GPU kernels store BF16 values -- the im2col input, the weights of every
  layer, every layer's activations and every backward activation gradient --
  so that the GPU result differs from it only by FP32 accumulation order.

Parameter layout (tea_warmstart.h): W1[27][32], b1[32], then for layers 2..5
W[9][32][32], b[32], then W6[9][32], b6.
"""
from __future__ import annotations

import numpy as np

CH = 32
TAPS = 9
K1 = 27
HID = TAPS * CH * CH + CH
OFF_W1 = 0
OFF_B1 = K1 * CH
OFF_HID = OFF_B1 + CH
OFF_W6 = OFF_HID + 4 * HID
OFF_B6 = OFF_W6 + TAPS * CH
NPARAM = OFF_B6 + 1
OFFSETS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]   # tap = (dy+1)*3 + (dx+1)


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); float64 result."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    w1 = p[OFF_W1:OFF_B1].reshape(K1, CH)
    b1 = p[OFF_B1:OFF_HID]
    hidden = []
    for l in range(4):
        o = OFF_HID + l * HID
        hidden.append((p[o:o + TAPS * CH * CH].reshape(TAPS, CH, CH), p[o + TAPS * CH * CH:o + HID]))
    w6 = p[OFF_W6:OFF_B6].reshape(TAPS, CH)
    return w1, b1, hidden, w6, p[OFF_B6]


def pack(w1, b1, hidden, w6, b6):
    parts = [w1.ravel(), b1]
    for w, b in hidden:
        parts += [w.ravel(), b]
    parts += [w6.ravel(), np.atleast_1d(b6)]
    return np.concatenate(parts)


def init_params(seed: int = 0, scale: float = 1.0):
    """He-style initial weights (the trainer's starting point); the output
    layer starts at zero so that the untrained guess is the extrapolation."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((K1, CH)) * np.sqrt(2.0 / K1) * scale
    hidden = [(rng.standard_normal((TAPS, CH, CH)) * np.sqrt(2.0 / (TAPS * CH)) * scale, np.zeros(CH))
              for _ in range(4)]
    return pack(w1, np.zeros(CH), hidden, np.zeros((TAPS, CH)), 0.0).astype(np.float32)


def _shift(a, dy, dx):
    """a[..., y+dy, x+dx, :] with zeros outside (a: B x H x W x C)."""
    B, H, W = a.shape[:3]
    out = np.zeros_like(a)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd] = a[:, ys, xs]
    return out


def im2col(planes):
    """planes: B x (>=3) x H x W -> B x H x W x 32 (k = tap*3 + feature)."""
    f = np.moveaxis(np.asarray(planes, dtype=np.float64)[:, :3], 1, -1)   # B H W 3
    cols = [_shift(f, dy, dx) for dy, dx in OFFSETS]
    x = np.concatenate(cols, axis=-1)
    return np.concatenate([x, np.zeros(x.shape[:3] + (CH - K1,))], axis=-1)


def forward(params, planes, mode="exact"):
    """Returns (out B x H x W, cache)."""
    rnd = bf16 if mode == "bf16" else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, hidden, w6, b6 = unpack(params)
    x1 = rnd(im2col(planes))
    w1p = np.concatenate([rnd(w1), np.zeros((CH - K1, CH))], axis=0)
    a = [x1, rnd(np.maximum(x1 @ w1p + b1, 0.0))]
    ws = [rnd(w) for w, _ in hidden]
    for (w, b), wr in zip(hidden, ws):
        z = sum(_shift(a[-1], dy, dx) @ wr[t] for t, (dy, dx) in enumerate(OFFSETS)) + b
        a.append(rnd(np.maximum(z, 0.0)))
    w6r = rnd(w6)
    out = sum(_shift(a[-1], dy, dx) @ w6r[t] for t, (dy, dx) in enumerate(OFFSETS)) + b6
    return out, dict(a=a, w1p=w1p, ws=ws, w6r=w6r, rnd=rnd)


def masked_operator(e, kx, ky, margin):
    """A_M e on B x H x W arrays: e + sum over the faces between two cells at
    least `margin` from the edge of k_face (e_i - e_j), zero outside; kx / ky
    hold each cell's left / bottom face coefficient."""
    B, H, W = e.shape
    m = np.zeros((H, W), dtype=bool)
    m[margin:H - margin, margin:W - margin] = True
    e = np.where(m, e, 0.0)
    out = e.copy()
    fx = m[:, 1:] & m[:, :-1]           # face between x-1 and x, both in M
    dx = (e[:, :, 1:] - e[:, :, :-1]) * kx[:, :, 1:] * fx
    out[:, :, 1:] += dx
    out[:, :, :-1] -= dx
    fy = m[1:, :] & m[:-1, :]
    dy = (e[:, 1:, :] - e[:, :-1, :]) * ky[:, 1:, :] * fy
    out[:, 1:, :] += dy
    out[:, :-1, :] -= dy
    return np.where(m, out, 0.0)


def loss_and_grad(params, samples, margin, mode="exact", loss="mse"):
    """Loss over the cells M at least `margin` from the crop edge, e = out -
    target: "mse" sum_M e^2 / |M|, or "energy" sum_M e (A_M e) / |M|; and its
    gradient with respect to the parameter vector.  samples: B x 6 x H x W
    (features, target, Kx, Ky)."""
    samples = np.asarray(samples, dtype=np.float64)
    out, c = forward(params, samples, mode)
    a, ws, w6r, rnd = c["a"], c["ws"], c["w6r"], c["rnd"]
    B, H, W = out.shape
    mask = np.zeros((H, W))
    mask[margin:H - margin, margin:W - margin] = 1.0
    n = B * mask.sum()
    d = (out - samples[:, 3]) * mask
    ad = masked_operator(d, samples[:, 4], samples[:, 5], margin) if loss == "energy" else d
    lossv = float((d * ad).sum() / n)
    g6 = 2.0 * ad / n
    if mode == "bf16":
        g6 = g6.astype(np.float32).astype(np.float64)
    grads = {}
    # layer 6
    a5 = a[5]
    grads["w6"] = np.stack([np.einsum("bhwc,bhw->c", _shift(a5, dy, dx), g6) for dy, dx in OFFSETS])
    grads["b6"] = g6.sum()
    g = sum(_shift(g6[..., None], -dy, -dx) * w6r[t] for t, (dy, dx) in enumerate(OFFSETS))
    g = rnd(g * (a5 > 0))
    hidden_grads = [None] * 4
    for l in range(3, -1, -1):       # layers 5..2
        a_in = a[l + 1]
        gw = np.stack([np.einsum("bhwi,bhwo->io", _shift(a_in, dy, dx), g) for dy, dx in OFFSETS])
        hidden_grads[l] = (gw, g.sum(axis=(0, 1, 2)))
        ga = sum(_shift(g, -dy, -dx) @ ws[l][t].T for t, (dy, dx) in enumerate(OFFSETS))
        g = rnd(ga * (a_in > 0))
    gw1 = np.einsum("bhwk,bhwo->ko", a[0], g)[:K1]
    return lossv, pack(gw1, g.sum(axis=(0, 1, 2)), hidden_grads, grads["w6"], grads["b6"])


def forward_loop(params, planes):
    """Direct per-pixel loops of the float64 forward pass (tiny images only):
    an independent restatement used to pin `forward`."""
    w1, b1, hidden, w6, b6 = unpack(params)
    planes = np.asarray(planes, dtype=np.float64)
    B, _, H, W = planes.shape

    def at(img, b, y, x):
        return img[b, y, x] if 0 <= y < H and 0 <= x < W else np.zeros(img.shape[-1])

    feats = np.moveaxis(planes[:, :3], 1, -1)
    a = np.zeros((B, H, W, CH))
    for b in range(B):
        for y in range(H):
            for x in range(W):
                v = b1.copy()
                for t, (dy, dx) in enumerate(OFFSETS):
                    f = at(feats, b, y + dy, x + dx)
                    for c in range(3):
                        v += f[c] * w1[t * 3 + c]
                a[b, y, x] = np.maximum(v, 0.0)
    for w, bias in hidden:
        nxt = np.zeros_like(a)
        for b in range(B):
            for y in range(H):
                for x in range(W):
                    v = bias.copy()
                    for t, (dy, dx) in enumerate(OFFSETS):
                        v += at(a, b, y + dy, x + dx) @ w[t]
                    nxt[b, y, x] = np.maximum(v, 0.0)
        a = nxt
    out = np.zeros((B, H, W))
    for b in range(B):
        for y in range(H):
            for x in range(W):
                out[b, y, x] = b6 + sum(at(a, b, y + dy, x + dx) @ w6[t] for t, (dy, dx) in enumerate(OFFSETS))
    return out
