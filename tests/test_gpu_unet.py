"""The learned-AO network's HIP kernels against the float64 oracle (oracle/unet_oracle.py),
through the C ABI: the single-layer convolution and weight-gradient kernels, each network
layer's forward output from the GPU's own inputs, the head bit for bit, the gradients of
a training step, the Adam update bit for bit, and a loss that falls in training.

Tolerances: a convolution output may differ from the oracle's bf16 rounding of the exact
sum by one bf16 ulp (2^-7 relative), or by 1e-6 of the sum of |terms| where the sum is
near zero (the GPU sums in float32 in another order); a weight gradient by 1e-5 of the sum
of |terms|."""
import numpy as np
import pytest
import torch

import unet_oracle as U

pytestmark = pytest.mark.gpu


def dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    return torch.from_numpy(a).cuda()


def host16(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint16)


def round32(k):
    return (k + 31) // 32 * 32


def bf16_case(r, shape, scale=1.0):
    return U.bf16_bits(r.normal(0, scale, shape).astype(np.float32))


def check_conv(gpu, ref, absum):
    gap = np.abs(gpu - ref)
    bound = 2.0 ** -7 * np.abs(ref) + 1e-6 * absum
    assert (gap <= bound).all(), float((gap - bound).max())
    assert (gpu == ref).mean() >= 0.99


CONV_CASES = [  # n, h, w, c0, c1, cout, relu
    (1, 8, 8, 8, 0, 32, 1),
    (2, 10, 6, 32, 64, 32, 1),
    (1, 12, 20, 96, 128, 96, 1),
    (1, 6, 10, 96, 0, 224, 0),
    (3, 5, 7, 16, 0, 48, 0),
    (1, 16, 16, 128, 0, 128, 1),
    (2, 11, 37, 8, 0, 32, 1),
    (1, 24, 40, 8, 0, 64, 0),
    (3, 20, 72, 64, 0, 128, 1),
]


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv3x3_matches_oracle(ctx, case):
    n, h, w, c0, c1, cout, relu = case
    print("conv case", case, flush=True)
    r = np.random.default_rng(hash(case) % (1 << 32))
    k = 9 * (c0 + c1)
    kpad = round32(k)
    x0 = bf16_case(r, (n, h, w, c0))
    x1 = bf16_case(r, (n, h // 2, w // 2, c1)) if c1 else None
    wt = bf16_case(r, (cout, k), np.sqrt(2.0 / k))
    wpad = np.zeros((cout, kpad), np.uint16)
    wpad[:, :k] = wt
    bias = r.normal(0, 0.1, cout).astype(np.float32)
    out = torch.zeros(n * h * w * cout, dtype=torch.int16, device="cuda")
    keep = [dev(x0), dev(x1) if c1 else None, dev(wpad), dev(bias)]
    ctx.conv3x3_dev(keep[0], keep[1], c0, c1, keep[2], kpad, keep[3], cout, relu, n, w, h, out)
    gpu = U.from_bits(host16(out)).reshape(n, h, w, cout)
    f0 = U.from_bits(x0)
    f1 = U.from_bits(x1) if c1 else None
    ref = U.conv3x3(f0, f1, U.from_bits(wt), bias.astype(np.float64), relu, exact=False)
    xin = f0 if f1 is None else np.concatenate([f0, U.upsample2(f1)], 3)
    absum = (np.abs(U.im2col(xin)) @ np.abs(U.from_bits(wt)).T).reshape(ref.shape) + np.abs(bias)
    check_conv(gpu, ref, absum)


WGRAD_CASES = [  # n, h, w, c0, c1, cout
    (1, 8, 8, 8, 0, 32),
    (2, 10, 6, 32, 64, 32),
    (1, 12, 20, 96, 128, 96),
    (1, 16, 16, 128, 0, 128),
    (3, 9, 7, 64, 0, 64),
    (1, 96, 80, 32, 0, 64),
]


@pytest.mark.parametrize("case", WGRAD_CASES)
def test_conv3x3_wgrad_matches_oracle(ctx, case):
    n, h, w, c0, c1, cout = case
    print("wgrad case", case, flush=True)
    r = np.random.default_rng(hash(case) % (1 << 32) + 1)
    k = 9 * (c0 + c1)
    x0 = bf16_case(r, (n, h, w, c0))
    x1 = bf16_case(r, (n, h // 2, w // 2, c1)) if c1 else None
    dy = bf16_case(r, (n, h, w, cout), 0.01)
    gw = torch.zeros(cout * k, dtype=torch.float32, device="cuda")
    gb = torch.zeros(cout, dtype=torch.float32, device="cuda")
    keep = [dev(dy), dev(x0), dev(x1) if c1 else None]
    ctx.conv3x3_wgrad_dev(keep[0], keep[1], keep[2], c0, c1, cout, n, w, h, gw, gb)
    torch.cuda.synchronize()
    f0 = U.from_bits(x0)
    f1 = U.from_bits(x1) if c1 else None
    fdy = U.from_bits(dy)
    rw, rb = U.conv3x3_wgrad(fdy, f0, f1)
    xin = f0 if f1 is None else np.concatenate([f0, U.upsample2(f1)], 3)
    aw = np.abs(fdy.reshape(-1, cout)).T @ np.abs(U.im2col(xin))
    ab = np.abs(fdy.reshape(-1, cout)).sum(0)
    assert (np.abs(gw.cpu().numpy().reshape(cout, k) - rw) <= 1e-5 * aw + 1e-30).all()
    assert (np.abs(gb.cpu().numpy() - rb) <= 1e-5 * ab + 1e-30).all()
    # the chunks' partials are summed in a fixed order: a second call gives the same bits
    gw2, gb2 = torch.full_like(gw, 7.0), torch.full_like(gb, 7.0)
    ctx.conv3x3_wgrad_dev(keep[0], keep[1], keep[2], c0, c1, cout, n, w, h, gw2, gb2)
    torch.cuda.synchronize()
    assert torch.equal(gw.view(torch.int32), gw2.view(torch.int32)) and torch.equal(gb.view(torch.int32), gb2.view(torch.int32))


def random_params(seed):
    r = np.random.default_rng(seed)
    layers = []
    for c0, c1, co, _ in U.LAYERS:
        k = 9 * (c0 + c1)
        layers.append((r.normal(0, np.sqrt(2.0 / k), (co, k)), r.normal(0, 0.05, co)))
    return U.pack(layers, (r.normal(0, 0.2, 32), 0.3)).astype(np.float32)


def net_case(seed, n, h, w):
    r = np.random.default_rng(seed)
    x = r.normal(0, 1, (n, h, w, U.FEATURES))
    mask = (r.random((n, h, w)) > 0.2).astype(np.uint8)
    x[mask == 0] = 0.0
    target = r.random((n, h, w)).astype(np.float32)
    return U.bf16_bits(x.astype(np.float32)), mask, target


def gpu_cache(ctx, params, xb, mask, n, h, w):
    """The oracle's backward cache filled with the GPU's own activations."""
    layers, head = U.unpack(params.astype(np.float64))
    acts = []
    for i, (c0, c1, co, lv) in enumerate(U.LAYERS):
        cnt = n * (h >> lv) * (w >> lv) * co
        acts.append(U.from_bits(ctx.net_read_act(i, cnt)).reshape(n, h >> lv, w >> lv, co))
    pooled = {i: U.pool2(acts[s], exact=False) for i, s in U.POOL_INTO.items()}
    return dict(layers=[(U.bf16(wt), b) for wt, b in layers], head=head, acts=acts, pooled=pooled,
                x=U.from_bits(xb), mask=mask, exact=False)


def head_pred(cache):
    hw, hb = cache["head"]
    a = cache["acts"][-1].astype(np.float32)
    s = np.full(a.shape[:3], np.float32(hb), np.float32)
    for c in range(32):
        s = s + np.float32(hw[c]) * a[..., c]
    return s


@pytest.mark.parametrize("shape", [(2, 16, 24), (1, 64, 96)])
def test_net_forward_layers_match_oracle(ctx, shape):
    n, h, w = shape
    print("net forward", shape, flush=True)
    params = random_params(20)
    xb, mask, _ = net_case(21, n, h, w)
    ctx.net_set_params(params)
    ao = torch.zeros(n * h * w, dtype=torch.float32, device="cuda")
    keep = [dev(xb), dev(mask)]
    ctx.net_forward_dev(keep[0], keep[1], n, w, h, ao)
    cache = gpu_cache(ctx, params, xb, mask, n, h, w)
    acts = cache["acts"]
    for i, (c0, c1, co, lv) in enumerate(U.LAYERS):
        in0, in1 = U.layer_inputs(cache, i)
        wt, b = cache["layers"][i]
        ref = U.conv3x3(in0, in1, wt, b, True, exact=False)
        xin = in0 if in1 is None else np.concatenate([in0, U.upsample2(in1)], 3)
        absum = (np.abs(U.im2col(xin)) @ np.abs(wt).T).reshape(ref.shape) + np.abs(b)
        check_conv(acts[i], ref, absum)
    pred = head_pred(cache)
    want = np.where(mask != 0, np.clip(pred, 0, 1), 1).astype(np.float32)
    assert np.array_equal(ao.cpu().numpy().reshape(n, h, w).view(np.uint32), want.view(np.uint32))


def test_net_gradients_match_oracle(ctx):
    """GPU gradients of one step against the oracle's emulated backward over the GPU's own
    activations (per layer: relative L2 error <= 1e-2), and against the float64 definition
    (cosine >= 0.99: the size of bf16 training's rounding)."""
    n, h, w = 2, 32, 48
    params = random_params(30)
    xb, mask, target = net_case(31, n, h, w)
    ctx.net_set_params(params)
    keep = [dev(xb), dev(mask), dev(target)]
    ctx.net_train_step_dev(keep[0], keep[1], keep[2], n, w, h, 0.0)
    g = ctx.net_get_grads().astype(np.float64)
    cache = gpu_cache(ctx, params, xb, mask, n, h, w)
    pred = head_pred(cache).astype(np.float64)
    gref, lref = U.backward(pred, target.astype(np.float64), cache)
    assert abs(ctx.net_loss() - lref) <= 1e-5 * lref
    off = 0
    errs = []
    for c0, c1, co, _ in U.LAYERS + [(32, 0, 0, 0)]:
        size = co * 9 * (c0 + c1) + co if co else 33
        a, b = g[off:off + size], gref[off:off + size]
        errs.append(np.linalg.norm(a - b) / np.linalg.norm(b))
        off += size
    print("per-layer relative gradient error", ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-2
    pe, ce = U.forward(params.astype(np.float64), U.from_bits(xb), mask, exact=True)
    ge, _ = U.backward(pe, target.astype(np.float64), ce)
    assert g @ ge / np.linalg.norm(g) / np.linalg.norm(ge) >= 0.99


def test_adam_step_bit_exact(ctx):
    n, h, w = 1, 16, 16
    params = random_params(40)
    xb, mask, target = net_case(41, n, h, w)
    ctx.net_set_params(params)
    lr = np.float32(1e-3)
    keep = [dev(xb), dev(mask), dev(target)]
    ctx.net_train_step_dev(keep[0], keep[1], keep[2], n, w, h, float(lr))
    g = ctx.net_get_grads()
    p1 = ctx.net_get_params()
    b1, b2, one = np.float32(0.9), np.float32(0.999), np.float32(1)
    m = (one - b1) * g
    v = ((one - b2) * g) * g
    c1, c2 = one - b1, one - b2
    want = params - (lr * (m / c1)) / (np.sqrt(v / c2) + np.float32(1e-8))
    assert np.array_equal(p1.view(np.uint32), want.astype(np.float32).view(np.uint32))


def test_training_lowers_the_loss(ctx):
    """40 Adam steps on one batch fit a smooth function of the features."""
    n, h, w = 2, 32, 32
    r = np.random.default_rng(50)
    x = r.normal(0, 1, (n, h, w, U.FEATURES)).astype(np.float32)
    mask = np.ones((n, h, w), np.uint8)
    target = (0.5 + 0.4 * np.tanh(x[..., 1] - 0.5 * x[..., 4])).astype(np.float32)
    ctx.net_set_params(random_params(51))
    tx, tm, tt = dev(U.bf16_bits(x)), dev(mask), dev(target)
    losses = []
    for _ in range(40):
        ctx.net_train_step_dev(tx, tm, tt, n, w, h, 2e-3)
        losses.append(ctx.net_loss())
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.3 * losses[0]


def test_features_match_oracle(ctx):
    h, w = 24, 40
    r = np.random.default_rng(60)
    pos = r.normal(0, 5, (h, w, 3)).astype(np.float32)
    nrm = r.normal(size=(h, w, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=2, keepdims=True)).astype(np.float32)
    mask = (r.random((h, w)) > 0.3).astype(np.uint8)
    ssao = r.random((h, w)).astype(np.float32)
    noisy = r.random((h, w)).astype(np.float32)
    eye = np.array([1.0, 2.0, 3.0], np.float32)
    for nz in (None, noisy):
        out = torch.zeros(h * w * 8, dtype=torch.int16, device="cuda")
        keep = [dev(pos), dev(nrm), dev(mask), dev(ssao), None if nz is None else dev(nz)]
        ctx.net_features_dev(keep[0], keep[1], keep[2], keep[3], keep[4], eye, w, h, out)
        gpu = U.from_bits(host16(out)).reshape(h, w, 8)
        ref = U.features(pos, nrm, mask, ssao, nz, eye)
        assert (np.abs(gpu - U.bf16(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-7).all()


def _scene_frame(w=64, h=48):
    import raw_ao
    vpos, vnrm, idx = raw_ao.host.test_scene_mesh()
    eye = np.array([4, 4, 6], np.float32)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 1, 0), (0, 1, 0), 0.9,
                                              float(np.float32(w) / np.float32(h)), 0.1, 100.0, w, h)
    return raw_ao.host.mesh_tris(vpos, idx), pos, nrm, mask, eye


def test_learned_frame_composition(ctx):
    """learned_ao_dev without a probe == SSAO -> features -> forward, bit for bit; with a
    forced fallback == the ray-traced frame, bit for bit; the probe's RMSE == the network's
    error against ray tracing of the probe sub-image, computed here."""
    w, h = 64, 48
    tris, pos, nrm, mask, eye = _scene_frame(w, h)
    ctx.build_ploc(tris)
    ctx.net_set_params(random_params(70))
    P, N, M = dev(pos), dev(nrm), dev(mask)
    out = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    rep = ctx.learned_ao_dev(P, N, M, None, eye, w, h, out, ssao_samples=24, ssao_radius=2.0)
    torch.cuda.synchronize()
    learned = out.cpu().numpy().copy()
    ss = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    x = torch.zeros(h * w * 8, dtype=torch.int16, device="cuda")
    ao = torch.zeros((h, w), dtype=torch.float32, device="cuda")
    ctx.compute_ssao_dev(P, N, M, w, h, 24, 2.0, ss)
    ctx.net_features_dev(P, N, M, ss, None, eye, w, h, x)
    ctx.net_forward_dev(x, M, 1, w, h, ao)
    torch.cuda.synchronize()
    assert np.array_equal(learned.view(np.uint32), ao.cpu().numpy().view(np.uint32))
    assert rep["probe_pixels"] == 0 and not rep["fell_back"]
    # the probe: every 4th pixel from 2, traced as its own image
    st = 4
    rep = ctx.learned_ao_dev(P, N, M, None, eye, w, h, out, ssao_samples=24, ssao_radius=2.0, rt_spp=64, rt_radius=2.0,
                             probe_stride=st, fallback_rmse=1e9)
    sub = (slice(st // 2, h, st), slice(st // 2, w, st))
    prt = ctx.compute_rtao(np.ascontiguousarray(pos[sub]), np.ascontiguousarray(nrm[sub]), np.ascontiguousarray(mask[sub]),
                           64, 2.0)
    pm = mask[sub] != 0
    d = (learned[sub] - prt)[pm].astype(np.float64)
    assert rep["probe_pixels"] == pm.sum() > 0
    assert abs(rep["probe_rmse"] - np.sqrt((d * d).mean())) <= 1e-6 * max(1e-3, rep["probe_rmse"])
    assert rep["probe_max_error"] == np.float32(np.abs(d).max())
    assert not rep["fell_back"]
    # forced fallback: the full ray-traced frame
    rep = ctx.learned_ao_dev(P, N, M, None, eye, w, h, out, ssao_samples=24, ssao_radius=2.0, rt_spp=64, rt_radius=2.0,
                             probe_stride=st, fallback_rmse=-1.0)
    torch.cuda.synchronize()
    rt = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    assert rep["fell_back"]
    assert np.array_equal(out.cpu().numpy().view(np.uint32), rt.view(np.uint32))
    # the host-pointer form gives the same frame
    hao, hrep = ctx.learned_ao(pos, nrm, mask, eye, ssao_samples=24, ssao_radius=2.0)
    assert np.array_equal(hao.view(np.uint32), learned.view(np.uint32))


def test_weights_file_loads_through_the_abi(ctx, tmp_path):
    import raw_ao.learned as L
    p = random_params(80)
    f = str(tmp_path / "w.raow")
    L.save_weights(f, p, {"mode": "test"})
    ctx.net_load(f)
    assert np.array_equal(ctx.net_get_params(), p)
    bad = tmp_path / "bad.raow"
    bad.write_bytes(b"RAOX" + bytes(40))
    with pytest.raises(Exception):
        ctx.net_load(str(bad))


def test_persistent_conv_matches_lds_tiled(ctx):
    """The persistent forms (the LDS-tiled one for 32 inputs and 32 outputs, the direct-copy one
    for a concatenated input with 32 outputs by default and for every layer with option 4) and
    the one-tile-a-block form give the same bits: single layers at odd sizes, with several images
    and more tiles than resident blocks, with and without bias and ReLU, channel groups of 48
    with a remainder of 16 or 32 (and layers the persistent forms do not take, which every option
    routes alike); and every activation of a whole forward pass, the encoder's pooled outputs
    included."""
    import raw_ao
    r = np.random.default_rng(93)
    try:
        for (n1, h1, w1, c0, c1, cout, relu) in ((2, 14, 38, 32, 0, 32, 1), (3, 9, 33, 32, 0, 32, 0), (1, 1, 1, 32, 0, 32, 1),
                                                 (2, 600, 530, 32, 0, 32, 1), (1, 30, 70, 32, 64, 32, 1), (3, 9, 33, 64, 0, 64, 0),
                                                 (2, 34, 66, 96, 128, 96, 1), (1, 17, 45, 128, 0, 128, 0), (2, 40, 72, 32, 64, 32, 0),
                                                 (1, 2, 2, 64, 96, 64, 1)):
            print("persistent conv", n1, h1, w1, c0, c1, cout, flush=True)
            k = 9 * (c0 + c1)
            x0 = bf16_case(r, (n1, h1, w1, c0))
            x1 = bf16_case(r, (n1, h1 // 2, w1 // 2, c1)) if c1 else None
            wt = bf16_case(r, (cout, k), np.sqrt(2.0 / k))
            bias = torch.from_numpy(r.normal(0, 0.1, cout).astype(np.float32)).cuda()
            outs = {}
            for kern in (0, 3, 4):
                ctx.set_option(raw_ao.OPT_CONV_KERNEL, kern)
                out = torch.zeros(n1 * h1 * w1 * cout, dtype=torch.int16, device="cuda")
                kp = [dev(x0), dev(x1) if c1 else None, dev(wt)]
                ctx.conv3x3_dev(kp[0], kp[1], c0, c1, kp[2], k, bias, cout, relu, n1, w1, h1, out)
                outs[kern] = host16(out).copy()
            assert np.array_equal(outs[0], outs[3])
            assert np.array_equal(outs[4], outs[3])
        n, h, w = 2, 40, 72
        ctx.net_set_params(random_params(94))
        xb, mask, _ = net_case(95, n, h, w)
        keep = [dev(xb), dev(mask)]
        ao = torch.zeros(n * h * w, dtype=torch.float32, device="cuda")
        acts = {}
        for kern in (0, 3, 4):
            ctx.set_option(raw_ao.OPT_CONV_KERNEL, kern)
            ctx.net_forward_dev(keep[0], keep[1], n, w, h, ao)
            acts[kern] = [ctx.net_read_act(i, n * (h >> lv) * (w >> lv) * co) for i, (_, _, co, lv) in enumerate(U.LAYERS)]
        for a, b, c in zip(acts[0], acts[3], acts[4]):
            assert np.array_equal(a, b)
            assert np.array_equal(c, b)
    finally:
        ctx.set_option(raw_ao.OPT_CONV_KERNEL, 0)


def test_conv_kernels_agree(ctx):
    """The tap-major kernel sums in the general kernel's order: every layer's activations of
    a forward pass are identical (odd sizes: pixel tiles cross rows and images). The
    LDS-tiled kernel sums chunk-major: its single-layer outputs are within one bf16 rounding
    of the general kernel's."""
    import raw_ao
    n, h, w = 3, 24, 40
    ctx.net_set_params(random_params(90))
    xb, mask, _ = net_case(91, n, h, w)
    keep = [dev(xb), dev(mask)]
    ao = torch.zeros(n * h * w, dtype=torch.float32, device="cuda")
    acts = {}
    try:
        for kern in (1, 2):
            ctx.set_option(raw_ao.OPT_CONV_KERNEL, kern)
            ctx.net_forward_dev(keep[0], keep[1], n, w, h, ao)
            acts[kern] = [ctx.net_read_act(i, n * (h >> lv) * (w >> lv) * co) for i, (_, _, co, lv) in enumerate(U.LAYERS)]
        for a, b in zip(acts[1], acts[2]):
            assert np.array_equal(a, b)
        r = np.random.default_rng(92)
        for (n1, h1, w1, c0, c1, cout) in ((2, 14, 38, 32, 64, 32), (1, 9, 70, 96, 0, 96), (1, 16, 16, 128, 0, 224)):
            k = 9 * (c0 + c1)
            x0 = bf16_case(r, (n1, h1, w1, c0))
            x1 = bf16_case(r, (n1, h1 // 2, w1 // 2, c1)) if c1 else None
            wt = bf16_case(r, (cout, k), np.sqrt(2.0 / k))
            outs = {}
            for kern in (0, 1):
                ctx.set_option(raw_ao.OPT_CONV_KERNEL, kern)
                out = torch.zeros(n1 * h1 * w1 * cout, dtype=torch.int16, device="cuda")
                kp = [dev(x0), dev(x1) if c1 else None, dev(wt)]
                ctx.conv3x3_dev(kp[0], kp[1], c0, c1, kp[2], k, None, cout, 0, n1, w1, h1, out)
                outs[kern] = U.from_bits(host16(out))
            gap = np.abs(outs[0] - outs[1])
            assert (gap <= 2.0 ** -7 * np.abs(outs[1]) + 1e-4).all()
            assert (gap == 0).mean() > 0.9
        # the first layer's kernel (8 channels) sums in the general kernel's order
        x0 = bf16_case(r, (2, 13, 35, 8))
        wt = np.zeros((32, 96), np.uint16)
        wt[:, :72] = bf16_case(r, (32, 72), 0.2)
        outs = {}
        for kern in (0, 1):
            ctx.set_option(raw_ao.OPT_CONV_KERNEL, kern)
            out = torch.zeros(2 * 13 * 35 * 32, dtype=torch.int16, device="cuda")
            kp = [dev(x0), dev(wt)]
            ctx.conv3x3_dev(kp[0], None, 8, 0, kp[1], 96, None, 32, 1, 2, 35, 13, out)
            outs[kern] = host16(out).copy()
        assert np.array_equal(outs[0], outs[1])
    finally:
        ctx.set_option(raw_ao.OPT_CONV_KERNEL, 0)
