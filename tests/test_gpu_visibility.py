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
