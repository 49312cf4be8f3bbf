"""CPU oracle of the warm-start U-Net -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the network libaqflip trains (include/aqflip_unet.h):
channels-last tensors, 3x3x3 zero-padded convolutions with ReLU, 2x average
pooling, nearest 2x upsampling, skip concatenation and a 1x1x1 head, with the
parameter vector laid out exactly as aq_unet_get_params returns it.

`bf16=False` computes everything in FP64 (the reference for finite-difference
checks). `bf16=True` rounds to bfloat16 (nearest even) where the GPU stores
bf16 -- the input, the convolution weights, every convolution and pooling
output and the ReLU-masked output gradients -- and keeps the rest in FP64, so a
GPU result differs from it only by FP32 accumulation order.

The network has no counterpart in the reference repository; this oracle is
pinned by finite differences (tests/test_unet.py) and by the closed forms of its
pieces.
"""
from __future__ import annotations

import numpy as np

CONVS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
HEAD_C = 16


def layout():
    """Offsets of every parameter block: [(w_off, b_off)] per conv, head_w, head_b, total."""
    off = 0
    blocks = []
    for ci, co in CONVS:
        w = off
        off += 27 * ci * co
        b = off
        off += co
        blocks.append((w, b))
    hw = off
    off += HEAD_C
    hb = off
    off += 1
    return blocks, hw, hb, off


def unpack(params: np.ndarray):
    blocks, hw, hb, total = layout()
    assert params.size == total
    p = np.asarray(params, np.float64)
    Ws, bs = [], []
    for (w, b), (ci, co) in zip(blocks, CONVS):
        Ws.append(p[w:w + 27 * ci * co].reshape(co, 27, ci))
        bs.append(p[b:b + co])
    return Ws, bs, p[hw:hw + HEAD_C], p[hb:hb + 1]


def pack(Ws, bs, hw, hb) -> np.ndarray:
    parts = []
    for W, b in zip(Ws, bs):
        parts += [W.ravel(), b.ravel()]
    parts += [np.ravel(hw), np.ravel(hb)]
    return np.concatenate(parts)


def bf16(a):
    """Round to bfloat16, nearest even (values returned as float64)."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def conv3(x, W, b):
    B, D, H, Wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, D, H, Wd, W.shape[0]))
    for t, (dd, dh, dw) in enumerate(TAPS):
        y += xp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + Wd, :] @ W[:, t, :].T
    return y + b if b is not None else y


def conv3_dgrad(dy, W):
    B, D, H, Wd, _ = dy.shape
    dxp = np.zeros((B, D + 2, H + 2, Wd + 2, W.shape[2]))
    for t, (dd, dh, dw) in enumerate(TAPS):
        dxp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + Wd, :] += dy @ W[:, t, :]
    return dxp[:, 1:-1, 1:-1, 1:-1, :]


def conv3_wgrad(dy, x):
    B, D, H, Wd, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros((dy.shape[-1], 27, ci))
    for t, (dd, dh, dw) in enumerate(TAPS):
        xs = xp[:, 1 + dd:1 + dd + D, 1 + dh:1 + dh + H, 1 + dw:1 + dw + Wd, :]
        dW[:, t, :] = np.einsum("bdhwn,bdhwc->nc", dy, xs)
    return dW


def pool(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(dy):
    B, D, H, W, C = dy.shape
    return dy.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def forward(params, x2, bf=False, keep=False):
    """x2: (B, D, H, W, 2) -> y (B, D, H, W). keep=True also returns activations."""
    Ws, bs, hw, hb = unpack(params)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    if bf:
        Ws = [bf16(W) for W in Ws]
        bs = [np.asarray(b, np.float32).astype(np.float64) for b in bs]
    B, D, H, W, _ = x2.shape
    x0 = np.zeros((B, D, H, W, 8))
    x0[..., :2] = x2
    x0 = r(x0)

    def c(l, x):
        return r(np.maximum(conv3(x, Ws[l], bs[l]), 0.0))
    a1 = c(0, x0)
    s1 = c(1, a1)
    p1 = r(pool(s1))
    a2 = c(2, p1)
    s2 = c(3, a2)
    p2 = r(pool(s2))
    a3 = c(4, p2)
    b3 = c(5, a3)
    cat2 = np.concatenate([s2, up(b3)], -1)
    a4 = c(6, cat2)
    b4 = c(7, a4)
    cat1 = np.concatenate([s1, up(b4)], -1)
    a5 = c(8, cat1)
    b5 = c(9, a5)
    y = b5 @ hw + hb[0]
    if keep:
        return y, dict(x0=x0, a1=a1, s1=s1, p1=p1, a2=a2, s2=s2, p2=p2, a3=a3, b3=b3, cat2=cat2,
                       a4=a4, b4=b4, cat1=cat1, a5=a5, b5=b5, Ws=Ws, hw=hw)
    return y


def loss_and_grads(params, x2, target, mask, bf=False):
    """Masked MSE (mean over mask) and its gradient for every parameter."""
    y, A = forward(params, x2, bf=bf, keep=True)
    n = mask.sum()
    d = (y - target) * mask
    loss = (d * d).sum() / n
    gy = 2.0 * d / n
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws = A["Ws"]
    dW = [None] * 10
    db = [None] * 10
    g_b5 = gy[..., None] * A["hw"]
    dhw = np.einsum("bdhw,bdhwc->c", gy, A["b5"])
    dhb = np.array([gy.sum()])

    def layer(l, Y, G, X, need_dx=True):
        dyp = r(G * (Y > 0))
        db[l] = dyp.sum(axis=(0, 1, 2, 3))
        dW[l] = conv3_wgrad(dyp, X)
        return conv3_dgrad(dyp, Ws[l]) if need_dx else None

    g_a5 = layer(9, A["b5"], g_b5, A["a5"])
    g_cat1 = layer(8, A["a5"], g_a5, A["cat1"])
    g_b4 = up_bwd(g_cat1[..., 16:])
    g_a4 = layer(7, A["b4"], g_b4, A["a4"])
    g_cat2 = layer(6, A["a4"], g_a4, A["cat2"])
    g_b3 = up_bwd(g_cat2[..., 32:])
    g_a3 = layer(5, A["b3"], g_b3, A["a3"])
    g_p2 = layer(4, A["a3"], g_a3, A["p2"])
    g_s2 = g_cat2[..., :32] + pool_bwd(g_p2)
    g_a2 = layer(3, A["s2"], g_s2, A["a2"])
    g_p1 = layer(2, A["a2"], g_a2, A["p1"])
    g_s1 = g_cat1[..., :16] + pool_bwd(g_p1)
    g_a1 = layer(1, A["s1"], g_s1, A["a1"])
    layer(0, A["a1"], g_a1, A["x0"], need_dx=False)
    return loss, pack(dW, db, dhw, dhb)
