"""The RT stage R2 check (rtBvhParity, src/renderer/gpu/rt_bvh_parity.cpp:80-174) of this
library's BVH against the oracle: for the reference's four owned scenes and a 20,000-triangle
soup,
  B1  the GPU's PLOC tree is the CPU's node for node (the FNV-1a tree hashes and rounds);
  B2  the CPU's rt::closest over the GPU's tree finds the scan's nearest hit on every ray;
  B3  the GPU's closest hits (raw_ao_closest_hits) equal the scan's nearest hit, and its any
      hits (raw_ao_trace_rays) the scan's answer below a random tMax; no stack overflow; the
      tree's SAH within 1.25 of the binned-SAH build's;
and the two controls: PLOC radius 8 changes the tree, a slab exit shortened by 10% loses hits.
Besides the reference's counts, the GPU's closest hits are compared bit for bit (t, u, v,
triangle) with the scan, so no tie needs the reference's 1e-5 exemption. Test infrastructure:
the CPU side is the oracle. `python tests/r2_harness.py OUT.json [--quick]` writes the record in
the reference's raw-native.evidence/1 layout."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
for p in (HERE, os.path.join(os.path.dirname(HERE), "raw-native_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle_lib as O  # noqa: E402

SCENES = ["retro_room", "iso_street", "clip_stress", "dense_block", "soup"]


def any_tmax(n):
    """The any-hit rays' tMax: 8 x next() of the xorshift stream seeded 77 (rt_bvh_parity.cpp:104-106)."""
    s = 77
    out = np.zeros(n, np.float32)
    m = (1 << 64) - 1
    for k in range(n):
        s ^= (s << 13) & m
        s ^= s >> 7
        s ^= (s << 17) & m
        out[k] = np.float32(8.0) * np.float32((s >> 11) * (1.0 / 9007199254740992.0))
    return out


def run_scene(ctx, name, tris, n_rays):
    import time
    c = dict(scene=name, triangles=int(tris.shape[0]), rays=n_rays)
    cb, cl, croot, crounds = O.ploc(tris)
    t0 = time.perf_counter()
    ctx.build_ploc(tris)
    c["gpu_build_ms_with_transfers"] = round((time.perf_counter() - t0) * 1e3, 1)
    gb, gl, groot, grounds = ctx.read_tree()
    same = (cb.view(np.uint32) == gb.view(np.uint32)).all(axis=1) & (cl == gl).all(axis=1) if cb.shape == gb.shape else None
    c["nodes_differing"] = int((~same).sum()) if same is not None and croot == groot else max(len(cb), len(gb)) + 1
    c["rounds"] = int(grounds)
    c["cpu_tree_hash"] = "%016x" % O.tree_hash(cb, cl)
    c["gpu_tree_hash"] = "%016x" % O.tree_hash(gb, gl)
    c["leaf_order_differences"] = int((cl[:len(tris), 2] != gl[:len(tris), 2]).sum()) if cb.shape == gb.shape else -1
    c["sah_ratio"] = round(O.lib().orc_sah_cost(np.ascontiguousarray(gb), gb.shape[0], groot) /
                           O.lib().orc_binned_sah_cost(np.ascontiguousarray(tris), tris.shape[0], None), 4)
    rays = O.r2_rays(name, tris.shape[0], n_rays)
    scan, t2 = O.scan_hits(tris, rays)
    hit = scan["tri"] >= 0
    c["hits"] = int(hit.sum())
    # B2: the CPU's traversal of the GPU's tree
    cpu = O.tree_closest(gb, gl, groot, tris, rays)
    c["cpu_traversal_of_gpu_tree_differences"] = int(((cpu["tri"] != scan["tri"]) | (hit & (cpu["t"] != scan["t"]))).sum())
    # B3: the GPU's closest and any hits
    g = ctx.closest_hits(rays)
    tie = hit & (g["tri"] >= 0) & (t2 > 0) & (t2 - scan["t"] <= np.float32(1e-5) * scan["t"])
    diff = g["tri"] != scan["tri"]
    c["gpu_triangle_differences"] = int((diff & ~tie).sum())
    c["gpu_tie_exempt"] = int((diff & tie).sum())
    c["gpu_t_over_bound"] = int((~diff & hit & (np.abs(g["t"] - scan["t"]) > np.float32(1e-5) * scan["t"])).sum())
    c["gpu_hits_not_bit_identical"] = int((g.view(np.uint32).reshape(-1, 4) != scan.view(np.uint32).reshape(-1, 4))
                                          .any(axis=1).sum())
    amax = any_tmax(n_rays)
    arays = rays.copy()
    arays[:, 6] = amax
    a = ctx.trace_rays(arays, True)
    c["any_hit_differences"] = int(((a[:, 1] >= 0) != (hit & (scan["t"] < amax))).sum())
    c["stack_overflows"] = int((a[:, 1] == -2).sum())
    c["pass"] = (c["nodes_differing"] == 0 and c["cpu_traversal_of_gpu_tree_differences"] == 0 and
                 c["gpu_triangle_differences"] == 0 and c["gpu_t_over_bound"] == 0 and c["any_hit_differences"] == 0 and
                 c["stack_overflows"] == 0 and c["sah_ratio"] <= 1.25 and c["gpu_hits_not_bit_identical"] == 0)
    return c


def run(ctx, quick=False):
    n_rays = 20000 if quick else 100000
    adapter = ctx.adapter_info()
    rec = {"schema": "raw-native.evidence/1",
           "criterion": "RT stage R2, the compute BVH: B1 to B3 (evidence/rt-r2-bounds.json)",
           "backend": "hip", "adapter": "%s %s" % (adapter["device"], adapter["architecture"]), "error": "",
           "quick": quick, "scenes": []}
    for name in SCENES:
        tris = O.scene_tris(name, 20000, 7)
        rec["scenes"].append(run_scene(ctx, name, tris, n_rays))
    # controls: radius 8 on dense_block; the slab exit shortened by 10% on retro_room
    dense = O.scene_tris("dense_block")
    cb, cl, _, _ = O.ploc(dense)
    ctx.build_ploc(dense, 8)
    gb, gl, _, _ = ctx.read_tree()
    rec["control_radius_8_nodes_differing"] = int((~((cb.view(np.uint32) == gb.view(np.uint32)).all(axis=1) &
                                                      (cl == gl).all(axis=1))).sum())
    room = O.scene_tris("retro_room")
    ctx.build_ploc(room)
    rays = O.r2_rays("retro_room", room.shape[0], 100000)
    full = ctx.trace_rays(rays, False)
    short = ctx.trace_rays(rays, False, far_scale=0.9)
    rec["control_short_far_hits_missed"] = int(((full[:, 1] >= 0) & (short[:, 1] != full[:, 1])).sum())
    rec["pass"] = all(s["pass"] for s in rec["scenes"]) and rec["control_radius_8_nodes_differing"] > 0 and \
        rec["control_short_far_hits_missed"] > 0
    return rec


def main(argv):
    import raw_ao
    out = argv[1] if len(argv) > 1 else "r2_bvh.json"
    with raw_ao.Context(0) as ctx:
        rec = run(ctx, quick="--quick" in argv)
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({"pass": rec["pass"], "out": out}))
    return 0 if rec["pass"] else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv))
