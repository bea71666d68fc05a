"""CPU oracle of the learned AO surrogate. Test infrastructure: only the tests, the smoke
check and the bench's CPU baseline use it; the product path never does.

The reference has no learned AO (its learned precedent is the roughness network of
scripts/rr_model.py:43-80 and src/rhi/vulkan/vk_rrnet.cpp, held there within 1e-2 of a
PyTorch forward pass). This module is the network's mathematical definition in float64
numpy, in the parameter layout of include/raw_ao_hip.h (per layer [cout][9*cin] weights with
k = (ky*3 + kx)*cin + ci and tap offset (ky-1, kx-1), then cout biases; head 32 weights +
bias), with two modes:

  exact=True   pure float64 (the definition; its backward is pinned by finite differences)
  exact=False  the GPU's storage points emulated: weights, activations and gradients
               between layers rounded to bf16 (round to nearest even) where the kernels
               store them, float32 where the kernels compute in float32 elementwise; the
               convolution sums themselves stay float64 (the GPU's float32 MFMA sums differ
               from them only by accumulation order).

Pins (tests/test_unet_oracle.py): the convolution against scipy.signal.correlate per
channel pair and against a delta/shift kernel, pooling and upsampling against reshape
means and np.kron, the adjoint identities <conv(x), y> = <x, conv^T(y)>, and the whole
backward against central finite differences of the forward along random directions.
"""
import numpy as np

# (c0, c1, cout, level): c0 channels at the layer's level, c1 channels from the level
# below, upsampled 2x and concatenated after them.
LAYERS = [(8, 0, 32, 0), (32, 0, 32, 0), (32, 0, 64, 1), (64, 0, 64, 1), (64, 0, 96, 2), (96, 0, 96, 2),
          (96, 0, 128, 3), (128, 0, 128, 3), (96, 128, 96, 2), (96, 0, 96, 2), (64, 96, 64, 1), (64, 0, 64, 1),
          (32, 64, 32, 0), (32, 0, 32, 0)]
FEATURES = 8
POOL_INTO = {2: 1, 4: 3, 6: 5}        # layer -> the encoder output it pools
SKIP_FROM = {8: 5, 10: 3, 12: 1}      # decoder layer -> its skip input


def param_count():
    return sum(co * 9 * (c0 + c1) + co for c0, c1, co, _ in LAYERS) + 33


def unpack(params):
    """[(W (cout, K), b (cout,))...], (head w (32,), head b)"""
    p = np.asarray(params, np.float64).ravel()
    assert p.size == param_count()
    out, off = [], 0
    for c0, c1, co, _ in LAYERS:
        k = 9 * (c0 + c1)
        w = p[off:off + co * k].reshape(co, k)
        off += co * k
        b = p[off:off + co]
        off += co
        out.append((w, b))
    return out, (p[off:off + 32], p[off + 32])


def pack(layers, head):
    parts = []
    for w, b in layers:
        parts += [np.ravel(w), np.ravel(b)]
    parts += [np.ravel(head[0]), np.atleast_1d(head[1])]
    return np.concatenate(parts)


# ---- bf16 ---------------------------------------------------------------------------
def bf16_bits(x):
    """float -> bf16 bits, round to nearest even through float32 (as the kernels' f2bf)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16(x):
    return from_bits(bf16_bits(x))


# ---- building blocks ------------------------------------------------------------------
def im2col(x):
    """(N, H, W, C) -> (N*H*W, 9*C), zero padding, column (ky*3 + kx)*C + ci = x[y+ky-1, x+kx-1, ci]."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:h + 1, 1:w + 1] = x
    cols = np.empty((n, h, w, 9, c), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx] = xp[:, ky:ky + h, kx:kx + w]
    return cols.reshape(n * h * w, 9 * c)


def col2im(dcols, shape):
    """The adjoint of im2col."""
    n, h, w, c = shape
    d = dcols.reshape(n, h, w, 9, c)
    xp = np.zeros((n, h + 2, w + 2, c), dcols.dtype)
    for ky in range(3):
        for kx in range(3):
            xp[:, ky:ky + h, kx:kx + w] += d[:, :, :, ky * 3 + kx]
    return xp[:, 1:h + 1, 1:w + 1]


def upsample2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def pool2(x, exact=True):
    """2x2 average; emulated: float32 (a + b) + (c + d), times 0.25, rounded to bf16."""
    if exact:
        n, h, w, c = x.shape
        return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))
    f = x.astype(np.float32)
    s = (f[:, 0::2, 0::2] + f[:, 0::2, 1::2]) + (f[:, 1::2, 0::2] + f[:, 1::2, 1::2])
    return bf16(s * np.float32(0.25))


def up_grad(d, exact=True):
    """The adjoint of upsample2: the sum over each 2x2 block."""
    if exact:
        n, h, w, c = d.shape
        return d.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))
    f = d.astype(np.float32)
    return bf16((f[:, 0::2, 0::2] + f[:, 0::2, 1::2]) + (f[:, 1::2, 0::2] + f[:, 1::2, 1::2]))


def conv3x3(in0, in1, w, b, relu, exact=True):
    """One layer: in0 (N,H,W,c0) ++ upsample2(in1) -> (N,H,W,cout)."""
    x = in0 if in1 is None else np.concatenate([in0, upsample2(in1)], axis=3)
    n, h, wd, _ = x.shape
    y = im2col(x) @ w.T
    if b is not None:
        y = y + b
    if relu:
        y = np.maximum(y, 0.0)
    y = y.reshape(n, h, wd, -1)
    return y if exact else bf16(y)


def conv3x3_wgrad(dy, in0, in1):
    """(gw (cout, 9*cin), gb (cout,)) of a layer from its output grad (float64 sums)."""
    x = in0 if in1 is None else np.concatenate([in0, upsample2(in1)], axis=3)
    d = dy.reshape(-1, dy.shape[-1])
    return d.T @ im2col(x), d.sum(axis=0)


def conv3x3_dgrad(dy, w, cin_shape):
    """The input grad of a layer (before splitting a concatenation)."""
    return col2im(dy.reshape(-1, dy.shape[-1]) @ w, cin_shape)


# ---- the network ----------------------------------------------------------------------
def forward(params, x, mask, exact=True):
    """x (N,H,W,8) features (bf16 values when emulating), mask (N,H,W). Returns
    (pred (N,H,W) before the clamp, cache)."""
    layers, (hw, hb) = unpack(params)
    if not exact:
        layers = [(bf16(w), b) for w, b in layers]
    acts, pooled = [], {}
    for i, (c0, c1, co, lv) in enumerate(LAYERS):
        if i == 0:
            in0, in1 = x, None
        elif i in POOL_INTO:
            pooled[i] = pool2(acts[POOL_INTO[i]], exact)
            in0, in1 = pooled[i], None
        elif i in SKIP_FROM:
            in0, in1 = acts[SKIP_FROM[i]], acts[i - 1]
        else:
            in0, in1 = acts[i - 1], None
        acts.append(conv3x3(in0, in1, layers[i][0], layers[i][1], True, exact))
    a = acts[-1]
    if exact:
        pred = a @ hw + hb
    else:   # float32, bias first, channels in order
        hw32 = hw.astype(np.float32)
        s = np.full(a.shape[:3], np.float32(hb), np.float32)
        a32 = a.astype(np.float32)
        for c in range(32):
            s = s + hw32[c] * a32[..., c]
        pred = s.astype(np.float64)
    return pred, dict(layers=layers, head=(hw, hb), acts=acts, pooled=pooled, x=x, mask=mask, exact=exact)


def layer_inputs(cache, i):
    if i == 0:
        return cache["x"], None
    if i in POOL_INTO:
        return cache["pooled"][i], None
    if i in SKIP_FROM:
        return cache["acts"][SKIP_FROM[i]], cache["acts"][i - 1]
    return cache["acts"][i - 1], None


def loss(pred, target, mask):
    m = np.asarray(mask) != 0
    cnt = max(int(m.sum()), 1)
    return float(((pred - target) ** 2)[m].sum() / cnt)


def backward(pred, target, cache):
    """Gradients of loss() in the parameter layout, and the loss."""
    exact = cache["exact"]
    layers, (hw, hb) = cache["layers"], cache["head"]
    acts = cache["acts"]
    m = np.asarray(cache["mask"]) != 0
    cnt = max(int(m.sum()), 1)
    if exact:
        g = np.where(m, 2.0 * (pred - target) / cnt, 0.0)
        dy = g[..., None] * hw
    else:
        inv = np.float32(1.0) / np.float32(cnt)
        e = pred.astype(np.float32) - np.asarray(target, np.float32)
        g = np.where(m, (np.float32(2.0) * e) * inv, np.float32(0.0)).astype(np.float32)
        dy = bf16(g[..., None] * hw.astype(np.float32))
        g = g.astype(np.float64)
    ghw = (g[..., None] * acts[-1]).reshape(-1, 32).sum(axis=0)
    ghb = g.sum()
    lval = float(((pred - target) ** 2)[m].sum() / cnt)
    grads = [None] * len(LAYERS)
    d_act = {len(LAYERS) - 1: dy}
    skip_pending = {}
    for i in range(len(LAYERS) - 1, -1, -1):
        c0, c1, co, lv = LAYERS[i]
        dpre = np.where(acts[i] > 0, d_act.pop(i), 0.0)
        in0, in1 = layer_inputs(cache, i)
        grads[i] = conv3x3_wgrad(dpre, in0, in1)
        if i == 0:
            break
        n, h, w, _ = dpre.shape
        dx = conv3x3_dgrad(dpre, layers[i][0], (n, h, w, c0 + c1))
        if not exact:
            dx = bf16(dx)
        if i in SKIP_FROM:
            skip_pending[SKIP_FROM[i]] = dx[..., :c0]
            d_act[i - 1] = up_grad(dx[..., c0:], exact)
        elif i in POOL_INTO:
            src = POOL_INTO[i]
            if exact:
                d_act[src] = skip_pending.pop(src) + 0.25 * upsample2(dx)
            else:
                d_act[src] = bf16(skip_pending.pop(src).astype(np.float32) +
                                  np.float32(0.25) * upsample2(dx).astype(np.float32))
        else:
            d_act[i - 1] = dx
    return pack(grads, (ghw, ghb)), lval


def features(pos, nrm, mask, ssao, noisy, eye):
    """The feature channels of one G-buffer (float64, unrounded)."""
    h, w = mask.shape
    f = np.zeros((h, w, FEATURES), np.float64)
    m = np.asarray(mask) != 0
    p = np.asarray(pos, np.float64).reshape(h, w, 3)
    nv = np.asarray(nrm, np.float64).reshape(h, w, 3)
    d = np.sqrt(((p - np.asarray(eye, np.float64)) ** 2).sum(-1))
    l1 = np.abs(nv).sum(-1)
    safe = np.where(l1 > 0, l1, 1.0)
    ox = np.where(l1 > 0, nv[..., 0] / safe, 0.0)
    oy = np.where(l1 > 0, nv[..., 2] / safe, 0.0)
    neg = nv[..., 1] < 0
    tx = (1.0 - np.abs(oy)) * np.where(ox >= 0, 1.0, -1.0)
    ty = (1.0 - np.abs(ox)) * np.where(oy >= 0, 1.0, -1.0)
    ox, oy = np.where(neg, tx, ox), np.where(neg, ty, oy)
    f[..., 0] = 1.0
    f[..., 1] = d * 0.05
    f[..., 2] = ox
    f[..., 3] = oy
    f[..., 4] = np.asarray(ssao, np.float64).reshape(h, w)
    if noisy is not None:
        f[..., 5] = np.asarray(noisy, np.float64).reshape(h, w)
        f[..., 7] = 1.0
    f[..., 6] = nv[..., 1]
    f[~m] = 0.0
    return f
