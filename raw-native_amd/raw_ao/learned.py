"""The learned AO surrogate and denoiser: training data rendered by the method itself,
training on the GPU through the network's own kernels, weights files, inference.

Data: seeded scenes from the same generator as the 1M-triangle workload (ground plane,
boxes and icospheres, sizes log-uniform in [0.05, 2.0], 40 x 40 x 10 volume) with varying
object counts, viewed from random orbit cameras; the G-buffer comes from the host
rasterizer (the reference's raw::rasterize), SSAO and the 64-spp ray-traced AO target from
this library's GPU kernels. For the denoiser the noisy input is the first k samples of the
same rays (--noisy-spp, default 4) and the target either the mean of the other 64 - k
((64 a64 - k ak) / (64 - k), exact: AO is a count over samples; input noise and target noise
independent) or the 64-spp AO itself (--target rt64: the network learns the ray-traced frame
it stands in for, the k input samples being part of it).

The network: include/raw_ao_hip.h raw_ao_net_* (a 4-level U-Net, bf16 MFMA kernels, fp32
master weights and Adam on the device). Nothing here computes the network on the CPU.
"""
import collections
import concurrent.futures as cf
import json
import os
import time

import numpy as np

from . import host

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(os.path.dirname(HERE), "weights")
SURROGATE = os.path.join(WEIGHTS_DIR, "ao_surrogate.raow")
DENOISER = os.path.join(WEIGHTS_DIR, "ao_denoiser.raow")
# trained at 1920 x 1080 on 8-spp (16-, 18-, 20-, 24-spp) inputs to give the 64-spp frame they are the first samples of
DENOISER_1080P = os.path.join(WEIGHTS_DIR, "ao_denoiser_1080p.raow")
DENOISER_1080P_16 = os.path.join(WEIGHTS_DIR, "ao_denoiser_1080p_16spp.raow")
DENOISER_1080P_18 = os.path.join(WEIGHTS_DIR, "ao_denoiser_1080p_18spp.raow")
DENOISER_1080P_20 = os.path.join(WEIGHTS_DIR, "ao_denoiser_1080p_20spp.raow")
DENOISER_1080P_24 = os.path.join(WEIGHTS_DIR, "ao_denoiser_1080p_24spp.raow")
FEATURES = 8

# (c0, c1, cout): layer shapes in parameter order (as raw_ao_net_param_count lays them out)
LAYERS = [(8, 0, 32), (32, 0, 32), (32, 0, 64), (64, 0, 64), (64, 0, 96), (96, 0, 96), (96, 0, 128), (128, 0, 128),
          (96, 128, 96), (96, 0, 96), (64, 96, 64), (64, 0, 64), (32, 64, 32), (32, 0, 32)]

Camera = collections.namedtuple("Camera", "eye target fovy")


def init_params(seed, head_bias=0.9):
    """He-normal convolution weights, zero biases, a small head around head_bias."""
    r = np.random.default_rng(seed)
    parts = []
    for c0, c1, co in LAYERS:
        k = 9 * (c0 + c1)
        parts += [r.normal(0, np.sqrt(2.0 / k), co * k), np.zeros(co)]
    parts += [r.normal(0, 0.01, 32), [head_bias]]
    return np.concatenate(parts).astype(np.float32)


def save_weights(path, params, meta):
    """A RAOW weights file (what raw_ao_net_load reads): "RAOW", uint32 version 1, uint64
    parameter count, uint32 metadata length, fp32 parameters, JSON metadata."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    p = np.ascontiguousarray(params, "<f4").ravel()
    blob = json.dumps(meta, sort_keys=True).encode()
    head = b"RAOW" + np.array([1], "<u4").tobytes() + np.array([p.size], "<u8").tobytes() + np.array([len(blob)], "<u4").tobytes()
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(head + p.tobytes() + blob)
    os.replace(tmp, path)


def load_weights(path):
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != b"RAOW" or np.frombuffer(data, "<u4", 1, 4)[0] != 1:
        raise ValueError("%s: not a version-1 RAOW weights file" % path)
    n = int(np.frombuffer(data, "<u8", 1, 8)[0])
    mlen = int(np.frombuffer(data, "<u4", 1, 16)[0])
    params = np.frombuffer(data, "<f4", n, 20).astype(np.float32)
    meta = json.loads(data[20 + 4 * n:20 + 4 * n + mlen].decode())
    return params, meta


def training_scene(seed, rng, objects=(600, 3000)):
    """A seeded scene of the 1M-triangle family with a random object count in `objects`."""
    s = host.SCENE_1M
    return host.SceneSpec(seed, int(rng.integers(objects[0], objects[1] + 1)), s.p_sphere, s.subdiv, s.smin, s.smax,
                          s.ground, s.height)


def random_camera(rng):
    a = rng.uniform(0, 2 * np.pi)
    rad = rng.uniform(14.0, 34.0)
    eye = np.array([rad * np.cos(a), rng.uniform(4.0, 18.0), rad * np.sin(a)], np.float32)
    tgt = np.array([rng.uniform(-6, 6), rng.uniform(0.0, 3.0), rng.uniform(-6, 6)], np.float32)
    return Camera(eye, tgt, 0.9)


def rasterize(mesh, cam, w, h):
    vpos, vnrm, idx = mesh
    return host.rasterize(vpos, vnrm, idx, cam.eye, cam.target, (0, 1, 0), cam.fovy, float(np.float32(w) / np.float32(h)),
                          0.1, 200.0, w, h)


class ViewSet:
    """Rendered views resident on the GPU: features (bf16 bits), mask, target AO and, for the
    denoiser, the noisy low-spp AO."""

    def __init__(self, n, w, h, noisy):
        import torch
        dev = torch.device("cuda:0")
        self.n, self.w, self.h, self.count = n, w, h, 0
        self.x = torch.zeros((n, h, w, FEATURES), dtype=torch.int16, device=dev)
        self.mask = torch.zeros((n, h, w), dtype=torch.uint8, device=dev)
        self.target = torch.zeros((n, h, w), dtype=torch.float32, device=dev)
        self.ssao = torch.zeros((n, h, w), dtype=torch.float32, device=dev)
        self.noisy = torch.zeros((n, h, w), dtype=torch.float32, device=dev) if noisy else None
        self.eyes = np.zeros((n, 3), np.float32)


def render_views(ctx, n_scenes, views_per_scene, w, h, seed=1000, spp=64, radius=1.0, ssao_samples=24, ssao_radius=1.0,
                 noisy_spp=0, threads=16, log=print, stream=None, objects=(600, 3000), target="independent"):
    """Renders n_scenes x views_per_scene views into a ViewSet. The scene seeds are
    seed + i; the 1M-triangle workload's scene (seed 0x5EED) is never among them. With a noisy
    input, target "independent" is the mean of the spp - noisy_spp samples the input does not
    hold (an estimate of the expectation), "rt64" the spp-spp ray-traced AO itself, whose first
    noisy_spp samples are the input's (the network then learns the ray-traced frame it stands
    in for)."""
    import torch
    assert not (seed <= host.SCENE_1M.seed < seed + n_scenes)
    rng = np.random.default_rng(seed)
    vs = ViewSet(n_scenes * views_per_scene, w, h, noisy_spp > 0)
    dev = torch.device("cuda:0")
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    pos_d = torch.empty((h, w, 3), dtype=torch.float32, device=dev)
    nrm_d = torch.empty_like(pos_d)
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=threads) as pool:
        for si in range(n_scenes):
            spec = training_scene(seed + si, rng, objects)
            vpos, vnrm, idx, _ = host.scene_seeded(spec)
            tris = torch.from_numpy(host.mesh_tris(vpos, idx)).to(dev)
            ctx.build_ploc_dev(tris, 16, st)
            cams = [random_camera(rng) for _ in range(views_per_scene)]
            for cam, (_, pos, nrm, mask) in zip(cams, pool.map(lambda c: rasterize((vpos, vnrm, idx), c, w, h), cams)):
                v = vs.count
                pos_d.copy_(torch.from_numpy(pos))
                nrm_d.copy_(torch.from_numpy(nrm))
                vs.mask[v].copy_(torch.from_numpy(mask))
                ctx.compute_ssao_dev(pos_d, nrm_d, vs.mask[v], w, h, ssao_samples, ssao_radius, vs.ssao[v], st)
                ctx.compute_rtao_dev(pos_d, nrm_d, vs.mask[v], w, h, spp, radius, vs.target[v], st)
                if noisy_spp:
                    ctx.compute_rtao_dev(pos_d, nrm_d, vs.mask[v], w, h, noisy_spp, radius, vs.noisy[v], st)
                    if target == "independent":   # the target from the other samples only
                        vs.target[v] = (vs.target[v] * spp - vs.noisy[v] * noisy_spp) / (spp - noisy_spp)
                ctx.net_features_dev(pos_d, nrm_d, vs.mask[v], vs.ssao[v], vs.noisy[v] if noisy_spp else None, cam.eye, w,
                                     h, vs.x[v], st)
                vs.eyes[v] = cam.eye
                vs.count += 1
            torch.cuda.synchronize()
            del tris
            log("scene %d/%d (%d objects): %d views, %.1f s" % (si + 1, n_scenes, spec.objects, vs.count, time.time() - t0))
    return vs


def lr_at(frac, lr_max, warm=0.02):
    if frac < warm:
        return lr_max * frac / warm
    return lr_max * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * min(1.0, (frac - warm) / (1 - warm)))))


def masked_rmse(a, b, mask):
    """(rmse, max) over covered pixels of two torch images (measurement only)."""
    m = mask.bool()
    d = (a - b)[m].double()
    if d.numel() == 0:
        return 0.0, 0.0
    return float(d.square().mean().sqrt()), float(d.abs().max())


def evaluate(ctx, vs, idx, stream=None):
    """Mean over views of the network's RMSE against the target, and of its input's (the
    noisy AO for the denoiser, SSAO otherwise)."""
    import torch
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    ao = torch.empty((vs.h, vs.w), dtype=torch.float32, device="cuda:0")
    net, ss = [], []
    for v in idx:
        ctx.net_forward_dev(vs.x[v], vs.mask[v], 1, vs.w, vs.h, ao, st)
        net.append(masked_rmse(ao, vs.target[v], vs.mask[v])[0])
        base = vs.noisy[v] if vs.noisy is not None else vs.ssao[v]
        ss.append(masked_rmse(base, vs.target[v], vs.mask[v])[0])
    return float(np.mean(net)), float(np.mean(ss))


def train(ctx, vs, minutes, batch=8, crop=256, lr_max=1e-3, seed=0, val_views=8, log=print, eval_every=500,
          params=None, stream=None):
    """Adam on random crops of the training views for a wall-clock budget; the last
    val_views views are held out. Returns (params, history)."""
    import torch
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(seed)
    ntrain = vs.count - val_views
    val = list(range(ntrain, vs.count))
    if params is None:
        num = den = 0.0
        for v in range(ntrain):   # per view: a mask over all views would index past 2^31 elements
            m = vs.mask[v].double()
            num += float((vs.target[v].double() * m).sum())
            den += float(m.sum())
        tm = num / max(den, 1.0)
        params = init_params(seed, head_bias=tm)
    ctx.net_set_params(params)
    dev = torch.device("cuda:0")
    xb = torch.empty((batch, crop, crop, FEATURES), dtype=torch.int16, device=dev)
    mb = torch.empty((batch, crop, crop), dtype=torch.uint8, device=dev)
    tb = torch.empty((batch, crop, crop), dtype=torch.float32, device=dev)
    hist = []
    t0 = time.time()
    budget = minutes * 60.0
    step, losses = 0, []
    while True:
        el = time.time() - t0
        if el >= budget:
            break
        vv = rng.integers(0, ntrain, batch)
        ys = rng.integers(0, (vs.h - crop) // 8 + 1, batch) * 8
        xs = rng.integers(0, (vs.w - crop) // 8 + 1, batch) * 8
        for b in range(batch):
            xb[b].copy_(vs.x[vv[b], ys[b]:ys[b] + crop, xs[b]:xs[b] + crop])
            mb[b].copy_(vs.mask[vv[b], ys[b]:ys[b] + crop, xs[b]:xs[b] + crop])
            tb[b].copy_(vs.target[vv[b], ys[b]:ys[b] + crop, xs[b]:xs[b] + crop])
        ctx.net_train_step_dev(xb, mb, tb, batch, crop, crop, lr_at(el / budget, lr_max), st)
        step += 1
        if step % 50 == 0:
            losses.append(ctx.net_loss())
        if step % eval_every == 0:
            vr, sr = evaluate(ctx, vs, val, st)
            rec = dict(step=step, seconds=round(time.time() - t0, 1), loss=float(np.mean(losses[-10:])), val_rmse=vr,
                       val_rmse_ssao=sr)
            hist.append(rec)
            log("step %d  %.0f s  loss %.5f  val RMSE %.4f (input %.4f)" % (step, rec["seconds"], rec["loss"], vr, sr))
    torch.cuda.synchronize()
    vr, sr = evaluate(ctx, vs, val, st)
    hist.append(dict(step=step, seconds=round(time.time() - t0, 1), val_rmse=vr, val_rmse_ssao=sr))
    log("done: %d steps, val RMSE %.4f (input %.4f)" % (step, vr, sr))
    return ctx.net_get_params(), hist


def time_network(w=1920, h=1080, reps=20, batch=8, crop=256, log=print):
    """Milliseconds of the network's forward pass on a w x h frame and of one training step
    (batch crops of crop x crop), random weights and inputs, on the current stream."""
    import raw_ao
    import torch
    ctx = raw_ao.Context(0)
    ctx.net_set_params(init_params(0))
    r = np.random.default_rng(0)
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    x = torch.from_numpy(r.integers(0, 1 << 14, (h, w, FEATURES)).astype(np.int16)).to(dev)
    m = torch.ones((h, w), dtype=torch.uint8, device=dev)
    ao = torch.empty((h, w), dtype=torch.float32, device=dev)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    fwd = timed(lambda: ctx.net_forward_dev(x, m, 1, w, h, ao, st))
    xb = torch.from_numpy(r.integers(0, 1 << 14, (batch, crop, crop, FEATURES)).astype(np.int16)).to(dev)
    mb = torch.ones((batch, crop, crop), dtype=torch.uint8, device=dev)
    tb = torch.rand((batch, crop, crop), device=dev)
    step = timed(lambda: ctx.net_train_step_dev(xb, mb, tb, batch, crop, crop, 1e-4, st))
    flops = 2.0 * sum(co * 9 * (c0 + c1) * (w >> lv) * (h >> lv) for (c0, c1, co), lv in
                      zip(LAYERS, (0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0)))
    res = dict(width=w, height=h, forward_ms=fwd, forward_tflops=flops / (fwd * 1e-3) / 1e12, forward_gflop=flops / 1e9,
               train_step_ms=step, train_batch="%dx%dx%d" % (batch, crop, crop))
    log(json.dumps(res))
    ctx.close()
    return res


def main(argv=None):
    import argparse
    import sys
    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "--time":
        rest = [int(v) for v in argv[1:4]]
        time_network(*rest)
        return 0
    ap = argparse.ArgumentParser(description="render training views with the method and train the learned AO network "
                                             "(or --time [W H REPS]: time the network's forward pass and a training step)")
    ap.add_argument("--mode", choices=["surrogate", "denoiser"], default="surrogate")
    ap.add_argument("--scenes", type=int, default=40)
    ap.add_argument("--views", type=int, default=24, help="views per scene")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--minutes", type=float, default=10.0)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--eval-every", type=int, default=500)
    ap.add_argument("--init", default=None, help="start from this weights file")
    ap.add_argument("--objects", type=int, nargs=2, default=(600, 3000), help="object count range of the scenes")
    ap.add_argument("--noisy-spp", type=int, default=4, help="denoiser: rays a pixel of the noisy input")
    ap.add_argument("--target", choices=["independent", "rt64"], default="independent",
                    help="denoiser: the mean of the samples the input does not hold, or the 64-spp AO itself")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    import raw_ao
    import torch
    torch.cuda.init()
    ctx = raw_ao.Context(0)
    noisy = a.noisy_spp if a.mode == "denoiser" else 0
    t0 = time.time()
    vs = render_views(ctx, a.scenes, a.views, a.width, a.height, seed=a.seed, noisy_spp=noisy, log=log,
                      objects=tuple(a.objects), target=a.target)
    gen_s = time.time() - t0
    init = load_weights(a.init)[0] if a.init else None
    params, hist = train(ctx, vs, a.minutes, batch=a.batch, crop=a.crop, lr_max=a.lr, seed=a.seed, log=log,
                         eval_every=a.eval_every, params=init)
    meta = dict(mode=a.mode, scenes=a.scenes, views_per_scene=a.views, width=a.width, height=a.height, rt_spp=64,
                rt_radius=1.0, ssao_samples=24, ssao_radius=1.0, noisy_spp=noisy, target=a.target if noisy else "rt64",
                scene_seeds=[a.seed, a.seed + a.scenes - 1],
                objects=list(a.objects),
                batch=a.batch, crop=a.crop, lr=a.lr, minutes=a.minutes, render_seconds=round(gen_s, 1),
                history=hist, init=os.path.basename(a.init) if a.init else None)
    save_weights(a.out, params, meta)
    log("wrote %s" % a.out)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
