"""Ablation of the learned part with its committed weights, on views no training run saw
(the 1M- and 2M-triangle workload scenes, seeds 0x5EED and 0x5EED2; training used seeds
1000-1059): what the network gains over the method without it.

  surrogate  RMSE against 64-spp ray tracing: the network's AO vs SSAO (its input), at
             1920 x 1080 on the ~1M-triangle scene's seeded orbit view.
  denoiser   RMSE against 256-spp ray tracing: the network's AO from 4-spp ray tracing vs
             the 4-spp ray tracing itself, at 3840 x 2160 on the ~2M-triangle scene.
  1080p      the denoisers trained on 8-spp and 16-spp inputs to give the 64-spp frame: against
             that frame and against 256-spp ray tracing, at 1920 x 1080 on the ~1M-triangle scene;
             from 16 spp the frame is within the 0.02 RMSE bound of the 64-spp frame.
  back-off   the learned frame's fall-back back-off (RAW_AO_OPT_LEARNED_BACKOFF).
"""
import os

import numpy as np
import pytest
import torch

import raw_ao
import raw_ao.learned as L

pytestmark = pytest.mark.gpu


def frame(spec, w, h):
    vpos, vnrm, idx, _ = raw_ao.host.scene_seeded(spec)
    eye = raw_ao.host.orbit_camera(1)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9,
                                              float(np.float32(w) / np.float32(h)), 0.1, 200.0, w, h)
    dev = torch.device("cuda:0")
    return (raw_ao.host.mesh_tris(vpos, idx), eye, mask, *(torch.from_numpy(x).to(dev) for x in (pos, nrm, mask)))


def rmse(ctx, a, b, mask_t, mask, w, h):
    err = torch.empty_like(a)
    mx = torch.zeros(4, device=a.device)
    ctx.error_map_dev(a, b, mask_t, w, h, err, mx)
    return float(raw_ao.host.rmse_of(err.cpu().numpy(), mask)[0])


@pytest.mark.skipif(not os.path.exists(L.SURROGATE), reason="no trained surrogate weights")
def test_surrogate_beats_ssao_on_a_held_out_view(ctx):
    w, h = 1920, 1080
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_1M, w, h)
    ctx.build_ploc(tris)
    img = lambda: torch.empty((h, w), dtype=torch.float32, device="cuda")
    rt, ss, lao = img(), img(), img()
    ctx.compute_rtao_dev(P, N, M, w, h, 64, 1.0, rt)
    ctx.compute_ssao_dev(P, N, M, w, h, 24, 1.0, ss)
    ctx.net_load(L.SURROGATE)
    rep = ctx.learned_ao_dev(P, N, M, None, eye, w, h, lao, rt_radius=1.0, probe_stride=8, fallback_rmse=1e9)
    torch.cuda.synchronize()
    r_net, r_ss = rmse(ctx, lao, rt, M, mask, w, h), rmse(ctx, ss, rt, M, mask, w, h)
    print("surrogate RMSE %.4f vs SSAO %.4f (probe estimate %.4f)" % (r_net, r_ss, rep["probe_rmse"]))
    assert r_net < 0.7 * r_ss
    # the probe's estimate of the frame's error is within a third of the truth
    assert abs(rep["probe_rmse"] - r_net) < r_net / 3


@pytest.mark.skipif(not os.path.exists(L.SURROGATE), reason="no trained surrogate weights")
@pytest.mark.parametrize("w,h", [(1920, 1080), (200, 136)])
def test_learned_frame_equals_its_parts(ctx, w, h):
    """The learned frame (raw_ao_learned_ao_dev, no probe) equals bit for bit its parts called one
    by one: SSAO -> features -> the network's forward pass; 200 x 136 leaves partial 8 x 32 tiles."""
    print("learned frame parts", w, h, flush=True)
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_1M, w, h)
    ctx.net_load(L.SURROGATE)
    lao = torch.empty((h, w), dtype=torch.float32, device="cuda")
    ctx.learned_ao_dev(P, N, M, None, eye, w, h, lao, probe_stride=0)
    ss = torch.empty((h, w), dtype=torch.float32, device="cuda")
    ctx.compute_ssao_dev(P, N, M, w, h, 24, 1.0, ss)
    x = torch.empty((h, w, 8), dtype=torch.int16, device="cuda")
    ctx.net_features_dev(P, N, M, ss, None, eye, w, h, x)
    ao = torch.empty(h * w, dtype=torch.float32, device="cuda")
    ctx.net_forward_dev(x, M, 1, w, h, ao)
    torch.cuda.synchronize()
    assert torch.equal(lao.reshape(-1).view(torch.int32), ao.view(torch.int32))
    assert float(lao[M != 0].std()) > 0.01


@pytest.mark.skipif(not os.path.exists(L.DENOISER), reason="no trained denoiser weights")
def test_denoiser_beats_low_spp_ray_tracing(ctx):
    w, h = 3840, 2160
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_2M, w, h)
    ctx.build_ploc(tris)
    img = lambda: torch.empty((h, w), dtype=torch.float32, device="cuda")
    ref, rt4, dn = img(), img(), img()
    ctx.compute_rtao_dev(P, N, M, w, h, 256, 1.0, ref)
    ctx.compute_rtao_dev(P, N, M, w, h, 4, 1.0, rt4)
    ctx.net_load(L.DENOISER)
    ctx.learned_ao_dev(P, N, M, rt4, eye, w, h, dn, rt_radius=1.0)
    torch.cuda.synchronize()
    r_dn, r_4 = rmse(ctx, dn, ref, M, mask, w, h), rmse(ctx, rt4, ref, M, mask, w, h)
    print("denoised RMSE %.4f vs 4-spp %.4f" % (r_dn, r_4))
    assert r_dn < 0.5 * r_4


@pytest.mark.skipif(not os.path.exists(L.DENOISER_1080P), reason="no trained 1080p denoiser weights")
def test_1080p_denoiser_from_8_spp(ctx):
    w, h = 1920, 1080
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_1M, w, h)
    ctx.build_ploc(tris)
    img = lambda: torch.empty((h, w), dtype=torch.float32, device="cuda")
    ref, rt64, rt8, dn = img(), img(), img(), img()
    ctx.compute_rtao_dev(P, N, M, w, h, 256, 1.0, ref)
    ctx.compute_rtao_dev(P, N, M, w, h, 64, 1.0, rt64)
    ctx.compute_rtao_dev(P, N, M, w, h, 8, 1.0, rt8)
    ctx.net_load(L.DENOISER_1080P)
    ctx.learned_ao_dev(P, N, M, rt8, eye, w, h, dn, rt_radius=1.0)
    torch.cuda.synchronize()
    r_dn, r_8 = rmse(ctx, dn, rt64, M, mask, w, h), rmse(ctx, rt8, rt64, M, mask, w, h)
    r_dn256, r_64_256 = rmse(ctx, dn, ref, M, mask, w, h), rmse(ctx, rt64, ref, M, mask, w, h)
    print("1080p denoised RMSE %.4f vs 64 spp (8-spp input %.4f); vs 256 spp %.4f (64 spp %.4f)"
          % (r_dn, r_8, r_dn256, r_64_256))
    assert r_dn < 0.5 * r_8
    # closer to the 256-spp frame than the 64-spp frame it was trained to give
    assert r_dn256 < r_64_256


@pytest.mark.parametrize("k", [16, 18, 20, 24])
def test_1080p_denoiser_from_more_samples(ctx, k):
    """The denoisers trained on 16-, 18-, 20- and 24-spp inputs: from 18 spp on the frame keeps the
    0.02 RMSE bound against the 64-spp frame (from 16 spp it comes within a few 1e-4 of it)."""
    path = {16: L.DENOISER_1080P_16, 18: L.DENOISER_1080P_18, 20: L.DENOISER_1080P_20, 24: L.DENOISER_1080P_24}[k]
    if not os.path.exists(path):
        pytest.skip("no trained %d-spp 1080p denoiser weights" % k)
    w, h = 1920, 1080
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_1M, w, h)
    ctx.build_ploc(tris)
    img = lambda: torch.empty((h, w), dtype=torch.float32, device="cuda")
    ref, rt64, rtk, dn = img(), img(), img(), img()
    ctx.compute_rtao_dev(P, N, M, w, h, 256, 1.0, ref)
    ctx.compute_rtao_dev(P, N, M, w, h, 64, 1.0, rt64)
    ctx.compute_rtao_dev(P, N, M, w, h, k, 1.0, rtk)
    ctx.net_load(path)
    ctx.learned_ao_dev(P, N, M, rtk, eye, w, h, dn, rt_radius=1.0)
    torch.cuda.synchronize()
    r_dn, r_k = rmse(ctx, dn, rt64, M, mask, w, h), rmse(ctx, rtk, rt64, M, mask, w, h)
    r_dn256, r_64_256 = rmse(ctx, dn, ref, M, mask, w, h), rmse(ctx, rt64, ref, M, mask, w, h)
    print("1080p denoised from %d spp: RMSE %.4f vs 64 spp (input %.4f); vs 256 spp %.4f (64 spp %.4f)"
          % (k, r_dn, r_k, r_dn256, r_64_256))
    if k >= 18:
        assert r_dn <= 0.02   # the learned frame's bound against 64-spp ray tracing
    assert r_dn < 0.7 * r_k
    assert r_dn256 < r_64_256


@pytest.mark.skipif(not os.path.exists(L.SURROGATE), reason="no trained surrogate weights")
def test_fallback_backoff_skips_the_network_then_retries(ctx):
    """RAW_AO_OPT_LEARNED_BACKOFF: after a fall-back the next calls of that size ray trace at once
    (fell_back 2, no probe), 1, 2, 4, ... of them, then the network is tried again; a passing
    probe clears it; another image size starts afresh. Skipped frames equal plain ray tracing."""
    w, h = 256, 192
    tris, eye, mask, P, N, M = frame(raw_ao.host.SCENE_1M, w, h)
    ctx.build_ploc(tris)
    ctx.net_load(L.SURROGATE)
    rt = torch.empty((h, w), dtype=torch.float32, device="cuda")
    out = torch.empty_like(rt)
    ctx.compute_rtao_dev(P, N, M, w, h, 64, 1.0, rt)
    ctx.set_option(raw_ao.OPT_LEARNED_BACKOFF, 4)

    def call(thr, ww=w, hh=h, PP=P, NN=N, MM=M, o=out):
        rep = ctx.learned_ao_dev(PP, NN, MM, None, eye, ww, hh, o, rt_radius=1.0, probe_stride=8, fallback_rmse=thr)
        torch.cuda.synchronize()
        return "skip" if rep["network_skipped"] else ("fell" if rep["fell_back"] else "net")

    seq = [call(0.0) for _ in range(12)]   # a bound no network meets: every tried frame falls back
    # tried, skip 1, tried, skip 2, tried, skip 4, tried, skip 4 (the cap), ...
    assert seq == ["fell", "skip", "fell", "skip", "skip", "fell", "skip", "skip", "skip", "skip", "fell", "skip"], seq
    assert torch.equal(out.view(torch.int32), rt.view(torch.int32))   # a skipped frame is the ray-traced frame
    # another image size starts afresh
    w2, h2 = 128, 96
    _, _, _, P2, N2, M2 = frame(raw_ao.host.SCENE_1M, w2, h2)
    o2 = torch.empty((h2, w2), dtype=torch.float32, device="cuda")
    assert call(0.0, w2, h2, P2, N2, M2, o2) == "fell"
    # back at the first size: a passing probe clears the back-off for the calls after it
    assert call(0.0) == "fell"
    r = call(1e9)
    while r == "skip":
        r = call(1e9)
    assert r == "net"
    assert [call(1e9) for _ in range(3)] == ["net", "net", "net"]
    ctx.set_option(raw_ao.OPT_LEARNED_BACKOFF, 0)
    assert call(0.0) == "fell" and call(0.0) == "fell"   # off: every frame tries the network
