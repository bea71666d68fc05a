"""The batched AO path (one persistent launch over many views of several scenes, the
hipGraph-captured build + trace + reconcile, the autotuner) against the single-view path,
which the golden and oracle tests pin: every view's AO and SSAO bit for bit, the error
map bit for bit, the per-view max error exactly and the RMSE to the last float bit."""
import numpy as np
import pytest
import torch

import raw_ao

pytestmark = pytest.mark.gpu

# (scene, w, h, spp, radius): uneven sizes, spp that are not multiples of 64, odd widths
VIEWS = [(0, 64, 48, 64, 2.0), (1, 100, 37, 16, 1.0), (2, 128, 96, 33, 1.0), (1, 33, 65, 100, 1.5), (0, 80, 80, 1, 2.0),
         (2, 96, 64, 128, 1.0), (3, 40, 24, 64, 1.0)]


def scenes():
    out = [raw_ao.host.test_scene_mesh()]
    for seed, objs in ((11, 60), (12, 150)):
        s = raw_ao.host.SCENE_1M
        spec = raw_ao.host.SceneSpec(seed, objs, s.p_sphere, 2, s.smin, s.smax, 12.0, 4.0)
        out.append(raw_ao.host.scene_seeded(spec)[:3])
    out.append((np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)))   # empty
    return out


def gbuffer(mesh, w, h, k):
    vpos, vnrm, idx = mesh
    if idx.shape[0] == 0:
        vpos, vnrm, idx = raw_ao.host.test_scene_mesh()   # the empty scene's view sees the test scene's G-buffer
    eye = (4.0 + k, 4.0, 6.0 - k) if k < 3 else (9.0, 5.0, 9.0)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 1, 0), (0, 1, 0), 0.9,
                                              float(np.float32(w) / np.float32(h)), 0.1, 100.0, w, h)
    return pos, nrm, mask


@pytest.fixture(scope="module")
def batch_case(ctx):
    dev = torch.device("cuda:0")
    meshes = scenes()
    b = raw_ao.Batch(0)
    tris = [raw_ao.host.mesh_tris(m[0], m[2]) if m[2].shape[0] else np.zeros((0, 9), np.float32) for m in meshes]
    for t in tris:
        if t.shape[0]:
            b.add_scene(torch.from_numpy(t).to(dev))
        else:
            b.add_scene(torch.zeros(9, device=dev), ntri=0)
    views = []
    for k, (sc, w, h, spp, rad) in enumerate(VIEWS):
        pos, nrm, mask = gbuffer(meshes[sc], w, h, k)
        P, N, M = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (pos, nrm, mask))
        ao, ss, err = (torch.full((h, w), -7.0, device=dev) for _ in range(3))
        b.add_view(sc, P, N, M, w, h, spp, rad, ao, ss, err, ssao_samples=24, ssao_radius=rad)
        views.append(dict(pos=pos, nrm=nrm, mask=mask, spp=spp, rad=rad, sc=sc, ao=ao, ss=ss, err=err))
    yield b, tris, views
    b.close()


def snapshot(views):
    torch.cuda.synchronize()
    return [(v["ao"].cpu().numpy().copy(), v["ss"].cpu().numpy().copy(), v["err"].cpu().numpy().copy()) for v in views]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batch_matches_single_view_path(ctx, batch_case):
    b, tris, views = batch_case
    b.run()
    got = snapshot(views)
    rmse, mx, px, rays = b.results()
    want_rays = 0
    built = None
    for k, (v, (ao, ss, err)) in enumerate(zip(views, got)):
        print("view", k, VIEWS[k], flush=True)
        if built != v["sc"]:
            ctx.build_ploc(tris[v["sc"]])
            built = v["sc"]
        ref_ao = ctx.compute_rtao(v["pos"], v["nrm"], v["mask"], v["spp"], v["rad"])
        ref_ss = ctx.compute_ssao(v["pos"], v["nrm"], v["mask"], 24, v["rad"])
        assert np.array_equal(bits(ao), bits(ref_ao))
        assert np.array_equal(bits(ss), bits(ref_ss))
        rec = ctx.reconcile(ref_ss, ref_ao, v["mask"], 0.12)
        assert np.array_equal(bits(err), bits(rec.error_map))
        assert mx[k] == np.float32(rec.max_error)
        assert abs(float(rmse[k]) - float(rec.rmse)) <= float(np.spacing(np.float32(rec.rmse)))
        assert px[k] == int(v["mask"].astype(bool).sum())
        want_rays += int(px[k]) * v["spp"]
    assert rays == want_rays


def test_batch_graph_replays_build_trace_reconcile(batch_case):
    b, _, views = batch_case
    b.run()
    eager = snapshot(views)
    b.capture(with_builds=True)
    for rep in range(3):
        for v in views:
            v["ao"].fill_(-1.0)
            v["ss"].fill_(-1.0)
            v["err"].fill_(-1.0)
        torch.cuda.synchronize()
        b.launch()
        b.check()
        got = snapshot(views)
        for (a0, s0, e0), (a1, s1, e1) in zip(eager, got):
            assert np.array_equal(bits(a0), bits(a1)) and np.array_equal(bits(s0), bits(s1))
            assert np.array_equal(bits(e0), bits(e1))


def test_batch_results_do_not_depend_on_tuning(batch_case):
    b, _, views = batch_case
    b.run()
    base = snapshot(views)
    for chunk, tq, bp in ((1, 64, 1), (32, 4, 16), (7, 16, 3)):
        b.set_tuning(chunk, tq, bp)
        b.run(rebuild=True)
        for (a0, _, _), (a1, _, _) in zip(base, snapshot(views)):
            assert np.array_equal(bits(a0), bits(a1))
    b.set_tuning(4, 16, 8)


def test_batch_wave_cut_does_not_change_results(batch_case):
    """Each 64-ray group walking the tree once for its pixels' union box (the default) or
    every ray from the root: the same AO bit for bit in every view."""
    b, _, views = batch_case
    b.run()
    base = snapshot(views)
    try:
        for k in (0, 8):
            b.set_option(raw_ao.OPT_WAVE_CUT, k)
            b.run(rebuild=True)
            for (a0, _, _), (a1, _, _) in zip(base, snapshot(views)):
                assert np.array_equal(bits(a0), bits(a1))
        with pytest.raises(raw_ao.RawAoError):
            b.set_option(raw_ao.OPT_TRI_PHASE, 8)
    finally:
        b.set_option(raw_ao.OPT_WAVE_CUT, 2)


def test_batch_tuner_and_cache(batch_case, tmp_path):
    b, _, views = batch_case
    cache = str(tmp_path / "tune.txt")
    ms = b.tune(cache)
    assert ms > 0
    t1 = b.tuning()
    b.set_tuning(1, 1, 1)
    assert b.tune(cache) == pytest.approx(ms)
    assert b.tuning() == t1
    assert len(open(cache).read().splitlines()) == 1


def test_batch_argument_errors(batch_case):
    b, _, views = batch_case
    dev = torch.device("cuda:0")
    z = torch.zeros(16, device=dev)
    with pytest.raises(raw_ao.RawAoError):
        b.add_view(99, z, z, z, 4, 4, 64, 1.0, z)
    with pytest.raises(raw_ao.RawAoError):
        b.set_tuning(0, 16, 8)
    fresh = raw_ao.Batch(0)
    try:
        with pytest.raises(raw_ao.RawAoError):
            fresh.launch()
        with pytest.raises(raw_ao.RawAoError):
            fresh.run()
    finally:
        fresh.close()
