"""Closest hits with barycentrics (raw_ao_closest_hits, replacing rt::closest,
src/renderer/rt_bvh.cpp:125-152) and the RT stage R2 check (tests/r2_harness.py, the reference's
rtBvhParity) on the GPU: every hit bit for bit against the oracle's scan (t, u, v, triangle) on
the R2 rays, with finite tMax, in an empty scene and through the device form; the full R2 record
with 100,000 rays a scene against the reference's own record (tree hashes, rounds, hit counts,
the radius-8 control)."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import r2_harness
from cases import TREES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["retro_room", "clip_stress", "soup", "dense_block"])
def test_closest_hits_match_the_scan(ctx, name):
    print("closest hits", name, flush=True)
    tris = O.scene_tris(name, 20000, 7)
    ctx.build_ploc(tris)
    rays = O.r2_rays(name, tris.shape[0], 20000)
    rays[1::4, 6] = np.linspace(0.01, 20.0, len(rays[1::4]), dtype=np.float32)   # finite tMax on a quarter
    got = ctx.closest_hits(rays)
    want = O.tree_closest(*ctx.read_tree()[:3], tris, rays)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    full = rays[0::4]
    scan, _ = O.scan_hits(tris, full)
    assert np.array_equal(np.ascontiguousarray(got[0::4]).view(np.uint32), scan.view(np.uint32))
    assert (got["tri"] >= 0).mean() > 0.05


def test_closest_hits_edge_cases(ctx):
    print("closest hits edges", flush=True)
    ctx.build_ploc(np.zeros((0, 9), np.float32))
    rays = O.r2_rays("soup", 0, 100)
    h = ctx.closest_hits(rays)
    assert (h["tri"] == -1).all() and (h["t"] == -1).all()
    # a closed box seen from inside: every ray hits, none beyond tMax = 0 counts
    vpos, _, idx = O.test_scene_mesh()
    tris = O.mesh_tris(vpos, idx)[2:]
    ctx.build_ploc(tris)
    rng = np.random.default_rng(3)
    r = np.zeros((5000, 8), np.float32)
    r[:, 0:3] = rng.uniform(-0.9, 0.9, (5000, 3)) + np.array([0, 1, 0])
    d = rng.normal(size=(5000, 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 6] = 1e30
    h = ctx.closest_hits(r)
    assert (h["tri"] >= 0).all()
    assert np.array_equal(h.view(np.uint32), O.scan_hits(tris, r)[0].view(np.uint32))
    r[:, 6] = 0.0
    assert (ctx.closest_hits(r)["tri"] == -1).all()


def test_closest_hits_device_form(ctx):
    print("closest hits dev", flush=True)
    tris = O.scene_tris("iso_street")
    ctx.build_ploc(tris)
    rays = O.r2_rays("iso_street", tris.shape[0], 30000)
    want = ctx.closest_hits(rays)
    d_rays = torch.from_numpy(rays).cuda()
    d_hits = torch.zeros((30000, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx.closest_hits_dev(d_rays, 30000, d_hits, stream=stream.cuda_stream)
    stream.synchronize()
    assert np.array_equal(d_hits.cpu().numpy().view(np.uint32), want.view(np.uint32).reshape(-1, 4))


def test_r2_record(ctx):
    """The whole R2 check at the reference's 100,000 rays a scene."""
    print("r2 record", flush=True)
    rec = r2_harness.run(ctx)
    for s in rec["scenes"]:
        print(s, flush=True)
        g = TREES[s["scene"]]
        assert s["pass"], s
        assert s["cpu_tree_hash"] == s["gpu_tree_hash"] == "%016x" % int(g["tree_hash"], 16)
        assert s["rounds"] == g["rounds"] and s["triangles"] == g["triangles"]
        assert s["hits"] == g["r2_hits"] and s["rays"] == g["r2_rays"] and s["sah_ratio"] == g["r2_sah_ratio"]
        assert s["gpu_tie_exempt"] == 0 and s["gpu_hits_not_bit_identical"] == 0
    from cases import GOLDEN
    assert rec["control_radius_8_nodes_differing"] == GOLDEN["control_radius_8_nodes_differing_dense_block"]
    assert rec["control_short_far_hits_missed"] > 0
    assert rec["pass"]
