"""Pins of the learned-AO network oracle (oracle/unet_oracle.py) on the CPU: each building
block against an independent implementation or an identity that a plausible mistake
(a transposed kernel, a wrong tap order, a dropped term of the chain rule) breaks, and the
whole backward against finite differences of the forward."""
import numpy as np
import pytest
import scipy.signal
import torch

import unet_oracle as U


def rng(seed):
    return np.random.default_rng(seed)


def random_params(seed, scale=1.0):
    r = rng(seed)
    layers = []
    for c0, c1, co, _ in U.LAYERS:
        k = 9 * (c0 + c1)
        layers.append((r.normal(0, scale * np.sqrt(2.0 / k), (co, k)), r.normal(0, 0.05, co)))
    head = (r.normal(0, 0.2, 32), 0.3)
    return U.pack(layers, head)


def test_bf16_rounding_matches_torch():
    r = rng(1)
    x = np.concatenate([r.normal(0, 1, 4000), r.normal(0, 1e-30, 100), r.normal(0, 1e30, 100),
                        1.0 + np.arange(-64, 64) * 2.0 ** -9, [0.0, -0.0, 1.0, -2.5]]).astype(np.float32)
    mine = U.bf16_bits(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(mine, ref)
    # ties to even: 1 + 2^-8 -> 1.0, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert U.bf16_bits(np.float32(1 + 2.0 ** -8)) == 0x3F80
    assert U.bf16_bits(np.float32(1 + 3 * 2.0 ** -8)) == 0x3F82


def test_conv_matches_scipy_correlate():
    r = rng(2)
    x = r.normal(size=(2, 5, 7, 3))
    w = r.normal(size=(4, 27))
    b = r.normal(size=4)
    y = U.conv3x3(x, None, w, b, relu=False)
    kern = w.reshape(4, 3, 3, 3)   # [co][ky][kx][ci]
    for n in range(2):
        for co in range(4):
            ref = sum(scipy.signal.correlate(x[n, :, :, ci], kern[co, :, :, ci], mode="same") for ci in range(3)) + b[co]
            assert np.allclose(y[n, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_conv_tap_orientation():
    """Tap (ky, kx) = (0, 0) reads the pixel up-left: y[y, x] = x[y-1, x-1], zero at the border."""
    x = rng(3).normal(size=(1, 4, 5, 8))
    w = np.zeros((8, 72))
    w[:, 0:8] = np.eye(8)
    y = U.conv3x3(x, None, w, None, relu=False)
    assert np.array_equal(y[0, 1:, 1:], x[0, :-1, :-1])
    assert not y[0, 0].any() and not y[0, :, 0].any()
    w = np.zeros((8, 72))
    w[:, 32:40] = np.eye(8)   # centre tap: identity
    assert np.array_equal(U.conv3x3(x, None, w, None, relu=False), x)


def test_concat_upsample_input():
    """in1 is upsampled 2x (nearest) and its channels follow in0's."""
    r = rng(4)
    in0 = r.normal(size=(1, 4, 6, 8))
    in1 = r.normal(size=(1, 2, 3, 8))
    w = np.zeros((8, 9 * 16))
    w[:, 4 * 16 + 8:4 * 16 + 16] = np.eye(8)   # centre tap, second half
    y = U.conv3x3(in0, in1, w, None, relu=False)
    for yy in range(4):
        for xx in range(6):
            assert np.array_equal(y[0, yy, xx], in1[0, yy // 2, xx // 2])


def test_pool_and_upsample_blocks():
    r = rng(5)
    x = r.normal(size=(2, 4, 6, 3))
    p = U.pool2(x)
    for yy in range(2):
        for xx in range(3):
            assert np.allclose(p[:, yy, xx], x[:, 2 * yy:2 * yy + 2, 2 * xx:2 * xx + 2].mean(axis=(1, 2)))
    u = U.upsample2(p)
    assert np.array_equal(u, np.kron(p, np.ones((1, 2, 2, 1))))


def test_adjoint_identities():
    r = rng(6)
    x = r.normal(size=(2, 4, 6, 5))
    ycols = r.normal(size=(2 * 4 * 6, 45))
    assert np.isclose((U.im2col(x) * ycols).sum(), (x * U.col2im(ycols, x.shape)).sum(), rtol=1e-12)
    a = r.normal(size=(2, 2, 3, 5))
    b = r.normal(size=(2, 4, 6, 5))
    assert np.isclose((U.upsample2(a) * b).sum(), (a * U.up_grad(b)).sum(), rtol=1e-12)


def _net_case(seed, n=2, h=16, w=24):
    r = rng(seed)
    x = r.normal(size=(n, h, w, U.FEATURES))
    mask = r.random((n, h, w)) > 0.2
    x[~mask] = 0.0
    target = r.random((n, h, w))
    return x, mask.astype(np.uint8), target


def test_backward_matches_finite_differences():
    """Directional derivatives of the loss along random directions confined to each layer
    (and the head) equal <grad, direction> to 1e-6 relative."""
    params = random_params(7)
    x, mask, target = _net_case(8)
    pred, cache = U.forward(params, x, mask, exact=True)
    grads, lval = U.backward(pred, target, cache)
    assert np.isclose(lval, U.loss(pred, target, mask))
    r = rng(9)
    off = 0
    blocks = []
    for c0, c1, co, _ in U.LAYERS:
        size = co * 9 * (c0 + c1) + co
        blocks.append((off, size))
        off += size
    blocks.append((off, 33))
    # The loss is piecewise quadratic (ReLU kinks): a step that crosses a kink is off by
    # O(step); the best of a few small steps is exact to rounding unless the gradient is wrong.
    for off, size in blocks:
        v = np.zeros_like(params)
        v[off:off + size] = r.normal(size=size)
        an = float(grads @ v)
        errs = []
        for eps in (1e-6, 1e-7, 1e-8):
            lp = U.loss(U.forward(params + eps * v, x, mask)[0], target, mask)
            lm = U.loss(U.forward(params - eps * v, x, mask)[0], target, mask)
            errs.append(abs((lp - lm) / (2 * eps) - an))
        assert min(errs) <= 1e-6 * max(abs(an), 1e-3), (off, errs, an)


def test_emulated_mode_tracks_exact():
    """The bf16-emulating forward and backward stay close to the float64 definition (the size of
    the rounding the GPU carries)."""
    params = random_params(10)
    x, mask, target = _net_case(11)
    xb = U.bf16(x)
    pe, ce = U.forward(params, xb, mask, exact=True)
    pb, cb = U.forward(params, xb, mask, exact=False)
    assert np.abs(pe - pb).max() < 0.05 * np.abs(pe).max()
    ge, _ = U.backward(pe, target, ce)
    gb, _ = U.backward(pb, target, cb)
    cos = ge @ gb / np.linalg.norm(ge) / np.linalg.norm(gb)
    assert cos > 0.99


def test_features_octahedral_normal():
    """The octahedral map sends the 6 axis normals to the square's centre and corners/edges
    and keeps |ox| + |oy| <= 1."""
    nrm = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1]], np.float64)
    r = rng(12).normal(size=(58, 3))
    nrm = np.concatenate([nrm, r / np.linalg.norm(r, axis=1, keepdims=True)]).reshape(8, 8, 3)
    pos = np.zeros((8, 8, 3))
    mask = np.ones((8, 8), np.uint8)
    f = U.features(pos, nrm, mask, np.zeros((8, 8)), None, (0, 0, 1))
    o = f[..., 2:4].reshape(-1, 2)
    assert np.allclose(o[0], [0, 0])
    assert np.allclose(np.abs(o[1]), [1, 1])
    assert np.allclose(o[2], [1, 0]) and np.allclose(o[3], [-1, 0]) and np.allclose(o[4], [0, 1])
    # upper hemisphere inside the diamond |ox| + |oy| <= 1, lower hemisphere folded outside it
    ny = nrm.reshape(-1, 3)[:, 1]
    l1 = np.abs(o).sum(1)
    assert (np.abs(o) <= 1 + 1e-12).all()
    assert (l1[ny > 0] <= 1 + 1e-12).all() and (l1[ny < 0] >= 1 - 1e-12).all()
    # the standard octahedral decode recovers the normal
    ox, oy = o[:, 0], o[:, 1]
    y = 1 - np.abs(ox) - np.abs(oy)
    fx = np.where(y < 0, (1 - np.abs(oy)) * np.where(ox >= 0, 1, -1), ox)
    fz = np.where(y < 0, (1 - np.abs(ox)) * np.where(oy >= 0, 1, -1), oy)
    dec = np.stack([fx, y, fz], 1)
    dec /= np.linalg.norm(dec, axis=1, keepdims=True)
    assert np.allclose(dec, nrm.reshape(-1, 3), atol=1e-12)
    assert np.allclose(f[..., 1].reshape(-1)[0], 0.05)   # distance 1 from the eye, times 0.05


@pytest.mark.parametrize("seed", [13])
def test_param_layout_roundtrip(seed):
    p = random_params(seed)
    layers, head = U.unpack(p)
    assert np.array_equal(U.pack(layers, head), p)
    assert p.size == U.param_count() == 905_472 + sum(co for _, _, co, _ in U.LAYERS) + 33
