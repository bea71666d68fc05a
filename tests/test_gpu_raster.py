"""The G-buffer rasterizer on the GPU (raw_ao_rasterize, replacing raw::rasterize,
src/renderer/raster.cpp:92-126) against the oracle's rasterizer, which tests/test_oracle.py pins
to the reference's code and golden masks: depth, position and normal bit for bit, and the mask.
Cases: the reference CLI's view of its test scene, the ~1M-triangle seeded scene at 1080p, a
camera inside a scene (near-plane clipping of the ground), a triangle soup that crosses the near
plane, coincident triangles (the earliest drawn wins), a vertex projected beyond the int range,
degenerate triangles, an empty mesh, the device-pointer form on a torch stream, and argument
errors."""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def check(got, want):
    gd, gp, gn, gm = got
    wd, wp, wn, wm = want
    assert np.array_equal(gm, wm)
    assert np.array_equal(bits(gd), bits(wd))
    assert np.array_equal(bits(gp), bits(wp))
    assert np.array_equal(bits(gn), bits(wn))


def both(ctx, vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h):
    got = ctx.rasterize(vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h)
    want = O.rasterize(vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h)
    return got, want


@pytest.mark.parametrize("w,h", [(256, 256), (97, 61), (1920, 1080)])
def test_cli_view_of_the_test_scene(ctx, w, h):
    print("raster test scene", w, h, flush=True)
    vpos, vnrm, idx = O.test_scene_mesh()
    aspect = float(np.float32(w) / np.float32(h))
    got = ctx.rasterize(vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, aspect, 0.1, 100.0, w, h)
    check(got, O.cli_gbuffer(w, h))
    assert got[3].sum() > 0.3 * w * h


def test_million_triangle_scene_at_1080p(ctx):
    import raw_ao.host as H
    print("raster 1M scene 1080p", flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded(H.SCENE_1M)
    eye = H.orbit_camera(1)
    got, want = both(ctx, vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9, 1920 / 1080, 0.1, 200.0, 1920, 1080)
    check(got, want)
    assert got[3].mean() > 0.5


def test_camera_inside_the_scene_clips_the_ground(ctx):
    import raw_ao.host as H
    print("raster inside camera", flush=True)
    spec = H.SceneSpec(0x51DE, 150, 0.4, 2, 0.05, 2.0, 40.0, 10.0)
    vpos, vnrm, idx, _ = H.scene_seeded(spec)
    for eye, center in [((0.5, 1.2, -0.3), (9.0, 0.4, 4.0)), ((-3.0, 0.3, 2.0), (-3.5, 0.0, -8.0))]:
        got, want = both(ctx, vpos, vnrm, idx, eye, center, (0, 1, 0), 1.2, 1.5, 0.1, 100.0, 240, 160)
        check(got, want)
        assert got[3].mean() > 0.5


def test_triangle_soup_across_the_near_plane(ctx):
    print("raster soup", flush=True)
    rng = np.random.default_rng(11)
    n = 6000
    centers = rng.uniform(-4, 4, (n, 1, 3)).astype(np.float32)
    vpos = (centers + rng.normal(0, 0.8, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    vnrm = rng.normal(0, 1, vpos.shape).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    for eye, near in [((0.0, 0.0, 0.0), 0.05), ((0.0, 0.0, 9.0), 0.1), ((1.0, -2.0, 0.5), 0.5)]:
        got, want = both(ctx, vpos, vnrm, idx, eye, (0.3, 0.2, -1.0), (0, 1, 0), 1.0, 1.25, near, 50.0, 160, 128)
        check(got, want)


def test_coincident_triangles_keep_the_earliest(ctx):
    print("raster coincident", flush=True)
    vpos, vnrm, idx = O.test_scene_mesh()
    nv = vpos.shape[0]
    vpos2 = np.concatenate([vpos, vpos])
    vnrm2 = np.concatenate([vnrm, -vnrm])
    idx2 = np.concatenate([idx, idx + nv])
    got, want = both(ctx, vpos2, vnrm2, idx2, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 128, 128)
    check(got, want)
    single = ctx.rasterize(vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 128, 128)
    check(got, single)


def test_vertex_projected_beyond_int_range(ctx):
    print("raster far vertex", flush=True)
    # camera at the origin looking down -z: view distance w = -z; a vertex at w = 1e-5 and
    # x = 1e4 lands ~1e11 pixels off screen, where the reference's (int) conversion gives INT_MIN
    vpos = np.array([[1e4, 0.0, -1e-5], [-1.0, -1.0, -2.0], [1.0, 1.0, -3.0],
                     [-0.5, -0.5, -1.5], [0.5, -0.5, -1.5], [0.0, 0.5, -1.5]], np.float32)
    vnrm = np.tile(np.array([[0, 0, 1]], np.float32), (6, 1))
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    got, want = both(ctx, vpos, vnrm, idx, (0, 0, 0), (0, 0, -1), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 64, 64)
    check(got, want)


def test_degenerate_and_empty_meshes(ctx):
    print("raster degenerate", flush=True)
    vpos = np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2], [0, 0, 0], [0, 0, 0], [0, 0, 0]], np.float32)
    vnrm = np.ones_like(vpos)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    got, want = both(ctx, vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 48, 32)
    check(got, want)
    assert got[3].sum() == 0
    depth, pos, nrm, mask = ctx.rasterize(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3), np.int32), (4, 4, 6),
                                          (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 40, 30)
    assert np.isinf(depth).all() and (mask == 0).all() and (pos == 0).all() and (nrm == 0).all()


def test_device_form_on_a_torch_stream(ctx):
    print("raster dev form", flush=True)
    vpos, vnrm, idx = O.test_scene_mesh()
    dev = torch.device("cuda:0")
    tv, tn, ti = (torch.from_numpy(a).to(dev) for a in (vpos, vnrm, idx))
    w, h = 200, 150
    aspect = float(np.float32(w) / np.float32(h))
    for _ in range(2):   # the second call reuses the context's buffers
        depth = torch.full((h, w), 7.0, device=dev)
        pos = torch.full((h, w, 3), 7.0, device=dev)
        nrm = torch.full((h, w, 3), 7.0, device=dev)
        mask = torch.full((h, w), 7, dtype=torch.uint8, device=dev)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            ctx.rasterize_dev(tv, tn, ti, idx.shape[0], (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, aspect, 0.1, 100.0, w, h,
                              depth, pos, nrm, mask, stream=stream.cuda_stream)
        stream.synchronize()
        check((depth.cpu().numpy(), pos.cpu().numpy(), nrm.cpu().numpy(), mask.cpu().numpy()), O.cli_gbuffer(w, h))


def test_argument_errors(ctx):
    import raw_ao
    vpos, vnrm, idx = O.test_scene_mesh()
    bad = idx.copy()
    bad[3, 1] = vpos.shape[0]
    with pytest.raises(raw_ao.RawAoError):
        ctx.rasterize(vpos, vnrm, bad, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 32, 32)
    with pytest.raises(raw_ao.RawAoError):
        ctx.rasterize(vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.0, 100.0, 32, 32)
    with pytest.raises(raw_ao.RawAoError):
        ctx.rasterize(vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 0, 32)
    got = ctx.rasterize(vpos, vnrm, idx, (4, 4, 6), (0, 1, 0), (0, 1, 0), 0.9, 1.0, 0.1, 100.0, 32, 32)
    check(got, O.cli_gbuffer(32, 32))
