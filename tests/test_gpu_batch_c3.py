"""The batched uneven workload at its full size against the oracle: 8 seeded scenes of the
1M-triangle scene's family (50k-2M triangles) and 64 square views of 256-2048 px at 16-128
rays a pixel, radius 1.0, SSAO 24 and the reconcile each, run as the captured graph of the 8
PLOC builds, the one persistent traversal launch and the reconciles.

What each view must equal is the reference's single-view path: computeRTAO
(src/renderer/ray_ao.cpp:20-58), computeSSAO (src/renderer/ssao.cpp:11-39) and the reconcile's
error map (src/cert/reconcile.cpp:4-21). The oracle restates those; it traces every 16th row
of each view (about 110 M of the 1.77 G rays), so the CPU time stays bounded, and the AO and
error map are compared there bit for bit. SSAO is compared on every pixel. Every view's whole
AO also equals the single-view HIP path's."""
import concurrent.futures as cf

import numpy as np
import pytest
import torch

import oracle_lib as O
import raw_ao
import raw_ao.host as H

pytestmark = pytest.mark.gpu

ROW_STEP = 16
RADIUS = 1.0


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.timeout(900)
def test_batched_workload_matches_oracle(ctx):
    specs, views = H.batched_workload()
    assert len(specs) == 8 and len(views) == 64
    meshes = [H.scene_seeded(s)[:3] for s in specs]
    tris = [H.mesh_tris(m[0], m[2]) for m in meshes]
    ntri = [t.shape[0] for t in tris]
    print("batched workload: scenes", ntri, flush=True)
    assert 50_000 <= min(ntri) and max(ntri) <= 2_000_000
    assert all(256 <= v[1] <= 2048 and 16 <= v[2] <= 128 for v in views)
    with cf.ThreadPoolExecutor(16) as pool:
        gbufs = list(pool.map(lambda v: H.batched_gbuffer(meshes[v[0]], v), views))
    dev = torch.device("cuda:0")
    b = raw_ao.Batch(0)
    try:
        for t in tris:
            b.add_scene(torch.from_numpy(t).to(dev))
        outs = []
        for (sc, side, spp, _), (_, pos, nrm, mask) in zip(views, gbufs):
            P, N, M = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (pos, nrm, mask))
            ao, ss, err = (torch.full((side, side), -7.0, device=dev) for _ in range(3))
            b.add_view(sc, P, N, M, side, side, spp, RADIUS, ao, ss, err, ssao_samples=24, ssao_radius=RADIUS)
            outs.append((P, N, M, ao, ss, err))
        b.capture(with_builds=True)
        b.launch()
        b.check()
        torch.cuda.synchronize()
        rmse, mx, px, rays = b.results()
        got = [(o[3].cpu().numpy(), o[4].cpu().numpy(), o[5].cpu().numpy()) for o in outs]
    finally:
        b.close()
    print("batched workload: %d rays traced" % rays, flush=True)
    assert rays == sum(int(px[k]) * v[2] for k, v in enumerate(views))
    checked_rays = 0
    for sc in range(8):
        mine = [k for k, v in enumerate(views) if v[0] == sc]
        if not mine:
            continue
        bvh = O.OracleBvh(tris[sc])
        ctx.build_ploc(tris[sc], 16)
        for k in mine:
            _, side, spp, _ = views[k]
            _, pos, nrm, mask = gbufs[k]
            ao, ss, err = got[k]
            print("view %d: scene %d, %d px square, %d spp" % (k, sc, side, spp), flush=True)
            rows = slice(0, side, ROW_STEP)
            ref_ao = bvh.rtao(pos, nrm, mask, spp, RADIUS, rows=(0, side), row_step=ROW_STEP)
            assert np.array_equal(bits(ao[rows]), bits(ref_ao[rows])), k
            checked_rays += int(mask[rows].astype(bool).sum()) * spp
            ref_ss = O.ssao(pos, nrm, mask, 24, RADIUS)
            assert np.array_equal(bits(ss), bits(ref_ss)), k
            em, _, emax, epx, _ = O.reconcile(ref_ss[rows], ref_ao[rows], mask[rows], 0.12)
            assert np.array_equal(bits(err[rows]), bits(em)), k
            assert px[k] == int(mask.astype(bool).sum())
            assert mx[k] >= emax
            single = ctx.compute_rtao(pos, nrm, mask, spp, RADIUS)
            assert np.array_equal(bits(ao), bits(single)), k
    print("batched workload: %d rays checked against the oracle" % checked_rays, flush=True)
    assert checked_rays * ROW_STEP >= rays // 2
