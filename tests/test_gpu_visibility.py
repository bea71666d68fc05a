"""Lighting model 2's visibility (computeRTVisibility, src/renderer/ray_ao.cpp:18-63) on the
GPU against the oracle, which tests/test_oracle.py pins to the reference's own code: ao, sky
and sun bit for bit, with the default render's light (render.cpp:33: minus the test scene's
light direction, src/scene/scene.cpp:31), the no-sun control, spp other than 64, a seeded
scene and an empty one."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def f32_normalize(v):
    v = np.asarray(v, np.float32)
    length = np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), dtype=np.float32)
    return (v * (np.float32(1.0) / length)).astype(np.float32)


DEFAULT_TO_LIGHT = f32_normalize((-0.4, -1.0, -0.3)) * np.float32(-1.0)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("w,h,spp,forced", [(128, 128, 64, False), (96, 72, 33, False), (64, 64, 64, True),
                                            (80, 48, 100, False)])
def test_visibility_default_scene(ctx, w, h, spp, forced):
    print("visibility", w, h, spp, forced, flush=True)
    _, pos, nrm, mask = O.cli_gbuffer(w, h)
    vpos, _, idx = O.test_scene_mesh()
    tris = O.mesh_tris(vpos, idx)
    ctx.build_ploc(tris)
    got = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, spp, 2.0, forced)
    want = O.rtvis(tris, pos, nrm, mask, DEFAULT_TO_LIGHT, spp, 2.0, forced)
    for g, o in zip(got, want):
        assert np.array_equal(bits(g), bits(o))
    assert np.array_equal(bits(got[0]), bits(ctx.compute_rtao(pos, nrm, mask, spp, 2.0)))


def test_visibility_seeded_scene(ctx):
    tris = O.scene_tris("dense_block", 3000, 7)
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    c = (lo + hi) / 2
    vpos = tris.reshape(-1, 3)
    idx = np.arange(vpos.shape[0], dtype=np.int32).reshape(-1, 3)
    fn = np.cross(vpos[idx[:, 1]] - vpos[idx[:, 0]], vpos[idx[:, 2]] - vpos[idx[:, 0]])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-20)
    vnrm = np.repeat(fn, 3, axis=0).astype(np.float32)
    _, pos, nrm, mask = O.rasterize(vpos, vnrm, idx, c + (hi - lo) * np.float32(1.2), c, (0, 1, 0), 0.9, 1.0, 0.1, 1000.0,
                                    96, 96)
    ctx.build_ploc(tris)
    got = ctx.compute_rt_visibility(pos, nrm, mask, (0.4, 1.0, 0.3), 64, 1.0)
    want = O.rtvis(tris, pos, nrm, mask, (0.4, 1.0, 0.3), 64, 1.0)
    for g, o in zip(got, want):
        assert np.array_equal(bits(g), bits(o))


@pytest.mark.parametrize("spp", [64, 16])
def test_visibility_wave_cut(ctx, spp):
    """The 1M-triangle scene at radius 1: the AO rays from the wave's cut (the sky rays from the
    root), at 64 spp and below, against every ray from the root: ao, sky and sun bit for bit."""
    import raw_ao
    import raw_ao.host as H
    print("visibility wave cut", spp, flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded()
    _, pos, nrm, mask = H.rasterize(vpos, vnrm, idx, H.orbit_camera(3), (0, 2, 0), (0, 1, 0), 0.9, 160 / 90, 0.1, 200.0,
                                    160, 90)
    ctx.build_ploc(H.mesh_tris(vpos, idx))
    try:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 0)
        want = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, spp, 1.0)
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 2)
        got = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, spp, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 2)
    for g, o in zip(got, want):
        assert np.array_equal(bits(g), bits(o))
    assert 0.0 < float(np.mean(got[1][mask != 0])) < 1.0


def test_visibility_empty_scene(ctx):
    _, pos, nrm, mask = O.cli_gbuffer(32, 32)
    ctx.build_ploc(np.zeros((0, 9), np.float32))
    ao, sky, sun = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, 64, 2.0)
    assert (ao == 1).all() and (sky == 1).all() and (sun == 1).all()


@pytest.mark.parametrize("offset,radius", [(0.0, 1.0), (30.0, 1.0), (1000.0, 0.5), (5.0, 0.01)])
def test_visibility_blockers_just_past_radius(ctx, offset, radius):
    """Sky rays skip the boxes they leave before the AO radius (an open AO ray has no triangle
    nearer): a ceiling and a shell of small triangles just beyond the radius above a ground grid,
    at coordinates near the origin and far from it, must block the sky exactly as the oracle says."""
    print("visibility blockers past radius", offset, radius, flush=True)
    rng = np.random.default_rng(11)
    o = np.float32(offset)
    w = h = 24
    gx, gz = np.meshgrid(np.linspace(-1.5, 1.5, w, dtype=np.float32), np.linspace(-1.5, 1.5, h, dtype=np.float32))
    pos = np.stack([gx + o, np.zeros_like(gx), gz + o], -1).astype(np.float32)
    nrm = np.zeros_like(pos)
    nrm[..., 1] = 1.0
    mask = np.ones((h, w), np.uint8)
    mask[0, :5] = 0
    hc = np.float32(1e-3) + np.float32(radius) * np.float32(1.0005)   # ceiling: every hit at t > radius
    s = np.float32(4.0 * max(radius, 0.1))
    ceil = np.array([[[-s + o, hc, -s + o], [s + o, hc, -s + o], [s + o, hc, s + o]],
                     [[-s + o, hc, -s + o], [s + o, hc, s + o], [-s + o, hc, s + o]]], np.float32)
    ceil[:, :, 0] += np.float32(0.5 * s)   # an open side, so some sky rays pass
    # small triangles in a shell just past the radius around the grid's centre
    n = 400
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:, 1] = np.abs(d[:, 1])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = d * np.float32(radius * 1.003) + np.array([o, 1e-3, o], np.float32)
    e1 = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(0.05 * radius)
    e2 = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(0.05 * radius)
    shell = np.stack([c, c + e1, c + e2], 1).astype(np.float32)
    ground = np.array([[[-50 + o, -1e-2, -50 + o], [50 + o, -1e-2, 50 + o], [50 + o, -1e-2, -50 + o]]], np.float32)
    tris = np.ascontiguousarray(np.concatenate([ceil, shell, ground]).reshape(-1, 9))
    ctx.build_ploc(tris)
    got = ctx.compute_rt_visibility(pos, nrm, mask, (0.3, 1.0, 0.2), 64, radius)
    want = O.rtvis(tris, pos, nrm, mask, (0.3, 1.0, 0.2), 64, radius)
    for g, o_ in zip(got, want):
        assert np.array_equal(bits(g), bits(o_))
    sky, ao = want[1][mask != 0], want[0][mask != 0]
    assert float(np.mean(sky)) < float(np.mean(ao))   # the blockers past the radius matter


@pytest.mark.parametrize("refill", [0, 1, 40, 64])
def test_visibility_sky_refill(ctx, refill):
    """64 spp on the 1M-triangle scene: the sky rays traced apart from the AO rays with lanes
    refilled at `refill` idle lanes (0: each sky ray right after its AO ray, in the same lane)
    against the default threshold: ao, sky and sun bit for bit."""
    import raw_ao
    import raw_ao.host as H
    print("visibility sky refill", refill, flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded()
    _, pos, nrm, mask = H.rasterize(vpos, vnrm, idx, H.orbit_camera(2), (0, 2, 0), (0, 1, 0), 0.9, 320 / 180, 0.1, 200.0,
                                    320, 180)
    ctx.build_ploc(H.mesh_tris(vpos, idx))
    want = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, 64, 1.0)
    try:
        ctx.set_option(raw_ao.OPT_SKY_REFILL, refill)
        got = ctx.compute_rt_visibility(pos, nrm, mask, DEFAULT_TO_LIGHT, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_SKY_REFILL, 20)   # the default
    for g, o in zip(got, want):
        assert np.array_equal(bits(g), bits(o))
    assert 0.0 < float(np.mean(got[1][mask != 0])) < 1.0
