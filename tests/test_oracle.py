"""The CPU oracle pinned to the reference's own golden vectors, closed forms and
invariants (no GPU). Everything the GPU path is later compared with is checked here."""
import numpy as np
import pytest

import oracle_lib as O
from cases import CASES, KNOWN, KNOWN_CASE, TREES, params, sha


def render(name, threads=None):
    p = params(name)
    _, pos, nrm, mask = O.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    vpos, _, idx = O.test_scene_mesh()
    tris = O.mesh_tris(vpos, idx)
    ao = O.rtao(tris, pos, nrm, mask, 64, 2.0, threads=threads or p["threads"])
    ss = O.ssao(pos, nrm, mask, 24, 2.0, threads=threads or p["threads"])
    return pos, nrm, mask, ao, ss


@pytest.mark.parametrize("name", sorted(CASES))
def test_identity_golden_hashes(name):
    """ao_rt.pfm, ao_ss.pfm, mask.pgm and the three previews byte for byte
    (evidence/identity-golden.json; src/core/image.cpp:34-69 formats)."""
    pos, nrm, mask, ao, ss = render(name)
    em, rmse, mx, px, within = O.reconcile(ss, ao, mask, 0.12)
    f = CASES[name]["files"]
    assert sha(O.mask_pgm_bytes(mask)) == f["mask.pgm"]
    assert sha(O.pfm1_bytes(ao)) == f["ao_rt.pfm"]
    assert sha(O.pfm1_bytes(ss)) == f["ao_ss.pfm"]
    assert sha(O.pgm_bytes(em)) == f["ao_error.pgm"]
    assert sha(O.pgm_bytes(ao)) == f["ao_rt.pgm"]
    assert sha(O.pgm_bytes(ss)) == f["ao_ss.pgm"]


@pytest.mark.parametrize("short", sorted(KNOWN_CASE))
def test_reconcile_known_answers(short):
    """pixels, rmse, maxError and verdict (evidence/wasm-vs-native-512.json)."""
    _, _, mask, ao, ss = render(KNOWN_CASE[short])
    _, rmse, mx, px, within = O.reconcile(ss, ao, mask, 0.12)
    k = KNOWN[short]
    assert px == k["pixels"]
    assert "%.9g" % rmse == "%.9g" % k["rmse"]
    assert "%.9g" % mx == "%.9g" % k["maxError"]
    assert ("verified" if within else "refuted") == k["verdict"]


def test_reconcile_default_256():
    """37,996 pixels, rmse 0.129436031 (docs/architecture/adr/0012-media-engine.md:96)."""
    _, _, mask, ao, ss = render("default")
    _, rmse, _, px, within = O.reconcile(ss, ao, mask, 0.12)
    assert (px, "%.9g" % rmse, within) == (37996, "0.129436031", False)


def test_reconcile_two_pixels():
    """tests/test_reconcile.cpp:4-21: max 0.4, rmse sqrt(0.16/2) > 0.1 refuted; self 0."""
    a = np.array([[0.5, 0.8]], np.float32)
    b = np.array([[0.5, 0.4]], np.float32)
    m = np.ones((1, 2), np.uint8)
    em, rmse, mx, px, within = O.reconcile(a, b, m, 0.1)
    assert px == 2 and not within
    assert mx == np.float32(0.8) - np.float32(0.4)
    e = np.float64(np.float32(0.8) - np.float32(0.4))
    assert rmse == np.float32(np.sqrt(e * e / 2))
    _, rmse2, _, _, within2 = O.reconcile(a, a, m, 0.1)
    assert rmse2 == 0.0 and within2


@pytest.mark.parametrize("scene", sorted(TREES))
def test_ploc_tree_hash(scene):
    """The whole PLOC tree by FNV-1a hash, and its rounds (evidence/rt-r2-bvh-*.json)."""
    t = O.scene_tris(scene, 20000, 7)
    g = TREES[scene]
    assert t.shape[0] == g["triangles"]
    boxes, links, root, rounds = O.ploc(t)
    assert "%016x" % O.tree_hash(boxes, links) == g["tree_hash"]
    assert rounds == g["rounds"]
    assert root == 2 * t.shape[0] - 2
    sah = O.lib().orc_sah_cost(boxes, boxes.shape[0], root)
    assert round(sah, 4) == g["sah"]
    nodes = O.SZ()
    binned = O.lib().orc_binned_sah_cost(t, t.shape[0], O.C.byref(nodes))
    assert round(binned, 4) == g["m3_binned_sah"] and nodes.value == g["binned_nodes"]


def test_ploc_radius_8_control():
    """Radius 8 must change the dense_block tree (rt_bvh_parity.cpp:161-166)."""
    t = O.scene_tris("dense_block")
    b16, l16, _, _ = O.ploc(t, 16)
    b8, l8, _, _ = O.ploc(t, 8)
    diff = np.count_nonzero(np.any(l16 != l8, axis=1) | np.any(b16.view(np.uint32) != b8.view(np.uint32), axis=1))
    assert diff > 1000


def test_ploc_sorted_order_is_stable_key_order():
    """Sorting (code, index) pairs == a stable sort by code (the radix sort's contract)."""
    t = O.scene_tris("soup", 20000, 7)
    keys, order, g4 = O.ploc_keys(t)
    assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.uint32))
    assert keys.max() < (1 << 30)


def test_intersect_known_answers():
    """tests/test_primitives.cpp:73-85."""
    tri = np.array([-1, -1, 0, 1, -1, 0, 0, 1, 0], np.float32)
    t, u, v = O.C.c_float(), O.C.c_float(), O.C.c_float()

    def hit(o, d):
        return O.lib().orc_intersect(np.array(o, np.float32), np.array(d, np.float32), tri, O.C.byref(t), O.C.byref(u),
                                     O.C.byref(v))
    assert hit((0, 0, 5), (0, 0, -1)) and abs(t.value - 5.0) < 1e-4
    assert not hit((5, 5, 5), (0, 0, -1))
    assert hit((0.5, -0.5, 5), (0, 0, -1)) and abs(u.value - 0.625) < 1e-5 and abs(v.value - 0.25) < 1e-5
    assert hit((0, 0, -2), (0, 0, 1))
    assert not hit((0, 0, 5e-5), (0, 0, -1))


def test_watertight_closed_mesh():
    """No ray from inside a closed mesh escapes through a shared edge or vertex
    (tests/test_primitives.cpp:86-97): an octahedron, rays at every vertex, every edge
    midpoint and random edge points."""
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32) * np.float32(1.7)
    F = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    tris = np.array([np.concatenate([V[a], V[b], V[c]]) for a, b, c in F], np.float32)
    rng = np.random.default_rng(5)
    targets = list(V)
    for a, b, c in F:
        for e0, e1 in ((a, b), (b, c), (c, a)):
            targets.append((V[e0] + V[e1]) * np.float32(0.5))
            for s in rng.random(6, dtype=np.float32):
                targets.append(V[e0] + (V[e1] - V[e0]) * s)
    rays = []
    for o in (rng.random((8, 3), dtype=np.float32) - np.float32(0.5)) * np.float32(0.3):
        for tg in targets:
            d = (np.asarray(tg, np.float32) - o).astype(np.float32)
            rays.append(np.concatenate([o, d / np.float32(np.sqrt(np.float32(d @ d)))]))
    rays = np.array(rays, np.float32)
    t, tri = O.scan_rays(tris, rays, np.full(len(rays), 1e30, np.float32), any_hit=False)
    assert np.all(tri >= 0), "%d rays leaked" % int(np.sum(tri < 0))


def test_bvh_equals_scan_on_random_rays():
    """Bvh::occluded equals the scan of every triangle (tests/test_bvh.cpp:28-64)."""
    t = O.scene_tris("soup", 3000, 11)
    bvh = O.OracleBvh(t)
    rng = np.random.default_rng(3)
    o = (rng.random((4000, 3), dtype=np.float32) * 24 - 12).astype(np.float32)
    d = rng.standard_normal((4000, 3)).astype(np.float32)
    d /= np.sqrt((d * d).sum(1, keepdims=True)).astype(np.float32)
    tmax = (rng.random(4000, dtype=np.float32) * 8).astype(np.float32)
    _, tri = O.scan_rays(t, np.hstack([o, d]), tmax, any_hit=True)
    got = [O.lib().orc_bvh_occluded(bvh.h, np.ascontiguousarray(o[k]), np.ascontiguousarray(d[k]), float(tmax[k]))
           for k in range(4000)]
    assert np.array_equal(np.array(got, bool), tri >= 0)


def test_ao_bvh_equals_scan_and_skew_breaks_it():
    """computeRTAO on the BVH == the scan; the skew control (x0.99) must differ
    (tests/test_ao_bvh.cpp:48-93, RAW_NATIVE_AO_ORACLE=scan|skew)."""
    _, pos, nrm, mask = O.cli_gbuffer(128, 128)
    vpos, _, idx = O.test_scene_mesh()
    tris = O.mesh_tris(vpos, idx)
    b = O.OracleBvh(tris)
    a_bvh = b.rtao(pos, nrm, mask, 64, 2.0, mode=0)
    a_scan = b.rtao(pos, nrm, mask, 64, 2.0, mode=1)
    a_skew = b.rtao(pos, nrm, mask, 64, 2.0, mode=2)
    assert np.array_equal(a_bvh, a_scan)
    assert not np.array_equal(a_skew, a_scan)


def test_open_plane_and_flat_ssao():
    """Closed forms: nothing above a plane -> every ray open (AO 1.0); SSAO of a flat plane
    is 1.0 (n . d = 0 for every tap)."""
    tris = np.array([[-50, 0, -50, 50, 0, -50, 50, 0, 50], [-50, 0, -50, 50, 0, 50, -50, 0, 50]], np.float32)
    h, w = 16, 16
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    pos = np.stack([xs * np.float32(0.1), np.zeros_like(xs), ys * np.float32(0.1)], -1).astype(np.float32)
    nrm = np.zeros_like(pos)
    nrm[..., 1] = 1
    mask = np.ones((h, w), np.uint8)
    assert np.all(O.rtao(tris, pos, nrm, mask, 64, 2.0) == 1.0)
    assert np.all(O.ssao(pos, nrm, mask, 24, 2.0) == 1.0)


def test_sampler_matches_cosine_closed_form():
    """A point under an infinite plane at height hgt < R is blocked iff cos(theta) > hgt/R;
    under cosine sampling P(cos > c) = 1 - c^2, so E[AO] = (hgt/R)^2."""
    hgt, R = 0.5, 1.0
    tris = np.array([[-1e3, hgt, -1e3, 1e3, hgt, 1e3, 1e3, hgt, -1e3], [-1e3, hgt, -1e3, -1e3, hgt, 1e3, 1e3, hgt, 1e3]],
                    np.float32)
    h = w = 96
    pos = np.zeros((h, w, 3), np.float32)
    pos[..., 1] = np.float32(-1e-3)   # the origin offset p + n * 1e-3 lands on y = 0
    nrm = np.zeros_like(pos)
    nrm[..., 1] = 1
    ao = O.rtao(tris, pos, nrm, np.ones((h, w), np.uint8), 64, R)
    n = h * w * 64
    expect = (hgt / R) ** 2
    assert abs(ao.mean() - expect) < 5 * np.sqrt(expect * (1 - expect) / n)


def test_threads_do_not_change_bytes():
    """Output identical for any thread count (raw/core/parallel.hpp:2-4)."""
    _, _, _, a1, s1 = render("default", threads=1)
    _, _, _, a5, s5 = render("default", threads=5)
    assert np.array_equal(a1, a5) and np.array_equal(s1, s5)


# ---- lighting model 2's visibility (computeRTVisibility, src/renderer/ray_ao.cpp:18-63) ----

def _default_frame(w=96, h=96):
    _, pos, nrm, mask = O.cli_gbuffer(w, h)
    vpos, _, idx = O.test_scene_mesh()
    return O.mesh_tris(vpos, idx), pos, nrm, mask


@pytest.mark.parametrize("to_light,forced", [((0.3, 1.0, 0.2), False), ((-1.0, 0.4, -0.5), False),
                                             ((0.0, 1.0, 0.0), True)])
def test_rtvis_ao_equals_rtao_and_scan(to_light, forced):
    """The visibility pass's AO is computeRTAO's (the same rays); BVH == scan for all three planes."""
    tris, pos, nrm, mask = _default_frame()
    ao, sky, sun = O.rtvis(tris, pos, nrm, mask, to_light, 64, 2.0, forced)
    assert np.array_equal(ao, O.rtao(tris, pos, nrm, mask, 64, 2.0))
    ao1, sky1, sun1 = O.rtvis(tris, pos, nrm, mask, to_light, 64, 2.0, forced, mode=1)
    assert np.array_equal(ao, ao1) and np.array_equal(sky, sky1) and np.array_equal(sun, sun1)
    assert (sky <= ao).all()                       # blocked within radius: blocked for the sky too
    assert (sun[mask == 0] == 1).all() and (sky[mask == 0] == 1).all()
    if forced:
        assert (sun == 1).all()
    else:
        assert set(np.unique(sun)) <= {0.0, 1.0} and (sun[mask != 0] == 0).any()


def test_rtvis_closed_forms():
    """Under a plane at height 0.5 every sky ray is blocked (sky 0) and a sun straight up is
    shadowed; with nothing above, sky = AO = 1 and the sun is lit."""
    h = w = 32
    pos = np.zeros((h, w, 3), np.float32)
    pos[..., 1] = np.float32(-1e-3)
    nrm = np.zeros_like(pos)
    nrm[..., 1] = 1
    mask = np.ones((h, w), np.uint8)
    roof = np.array([[-1e3, 0.5, -1e3, 1e3, 0.5, 1e3, 1e3, 0.5, -1e3], [-1e3, 0.5, -1e3, -1e3, 0.5, 1e3, 1e3, 0.5, 1e3]],
                    np.float32)
    ao, sky, sun = O.rtvis(roof, pos, nrm, mask, (0, 1, 0), 64, 1.0)
    assert (sky == 0).all() and (sun == 0).all() and 0 < ao.mean() < 1
    floor = np.array([[-1e3, -1.0, -1e3, 1e3, -1.0, -1e3, 1e3, -1.0, 1e3]], np.float32)
    ao, sky, sun = O.rtvis(floor, pos, nrm, mask, (0.2, 1, 0.1), 64, 1.0)
    assert (ao == 1).all() and (sky == 1).all() and (sun == 1).all()


# ---- the oracle against the reference's own code (oracle/_ref, built from its sources) ----

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")


@needs_ref
@pytest.mark.parametrize("to_light,forced", [((0.3, 1.0, 0.2), False), ((-1.0, 0.4, -0.5), False),
                                             ((0.0, 1.0, 0.0), True)])
def test_rtvis_matches_reference_code(to_light, forced):
    tris, pos, nrm, mask = _default_frame()
    mine = O.rtvis(tris, pos, nrm, mask, to_light, 64, 2.0, forced)
    ref = O.rtvis(tris, pos, nrm, mask, to_light, 64, 2.0, forced, impl=O.ref())
    for a, b in zip(mine, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@needs_ref
def test_rtvis_matches_reference_code_seeded_scene():
    tris = O.scene_tris("dense_block", 3000, 7)
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    c = (lo + hi) / 2
    eye = c + (hi - lo) * np.float32(1.2)
    vpos = tris.reshape(-1, 3)
    idx = np.arange(vpos.shape[0], dtype=np.int32).reshape(-1, 3)
    e0, e1 = vpos[idx[:, 1]] - vpos[idx[:, 0]], vpos[idx[:, 2]] - vpos[idx[:, 0]]
    fn = np.cross(e0, e1)
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-20)
    vnrm = np.repeat(fn, 3, axis=0).astype(np.float32)
    _, pos, nrm, mask = O.rasterize(vpos, vnrm, idx, eye, c, (0, 1, 0), 0.9, 1.0, 0.1, 1000.0, 64, 64)
    mine = O.rtvis(tris, pos, nrm, mask, (0.4, 1, 0.3), 32, 1.0)
    ref = O.rtvis(tris, pos, nrm, mask, (0.4, 1, 0.3), 32, 1.0, impl=O.ref())
    for a, b in zip(mine, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@needs_ref
def test_rtao_ssao_match_reference_code():
    tris, pos, nrm, mask = _default_frame(128, 96)
    R = O.ref()
    ref_ao = np.zeros(mask.shape, np.float32)
    R.ref_rtao(tris, tris.shape[0], 0, pos, nrm, mask, 128, 96, 64, 2.0, 4, ref_ao)
    assert np.array_equal(O.rtao(tris, pos, nrm, mask, 64, 2.0).view(np.uint32), ref_ao.view(np.uint32))
    ref_ss = np.zeros(mask.shape, np.float32)
    R.ref_ssao(pos, nrm, mask, 128, 96, 24, 2.0, 4, ref_ss)
    assert np.array_equal(O.ssao(pos, nrm, mask, 24, 2.0).view(np.uint32), ref_ss.view(np.uint32))


# ---- the default frame: shade, shade_hdr and motion on top of the passes above --------

def _frame_case(name, lighting):
    from cases import params
    p = params(name)
    sc = O.frame_test_scene(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    prev = None
    if p["prev_eye"]:
        ps = dict(sc)
        ps["eye"] = np.array(p["prev_eye"], np.float32)
        prev = O.view_proj(ps)
    return O.frame(sc, p["w"], p["h"], prev=prev, lighting=lighting, rtao_on=p["rt"])


@pytest.mark.parametrize("lighting,key", [(2, "files"), (1, "files_lighting_v1")])
@pytest.mark.parametrize("name", ["default", "motion-256", "nort-256", "close-512", "wide-512"])
def test_frame_golden_hashes(name, lighting, key):
    """frame.ppm and frame_hdr.pfm of the reference's identity renders (evidence/identity-golden.json
    for lighting model 2, identity-golden-v1.json for model 1) from the oracle's frame."""
    from cases import FRAMES
    f = _frame_case(name, lighting)
    g = FRAMES[name][key]
    assert sha(O.ppm_bytes(f["frame"])) == g["frame.ppm"]
    assert sha(O.pfm3_bytes(f["hdr"])) == g["frame_hdr.pfm"]
    if key == "files":
        assert sha(O.mask_pgm_bytes(f["mask"])) == g["mask.pgm"]
        assert sha(O.pfm1_bytes(f["ao_ss"])) == g["ao_ss.pfm"]


def test_motion_closed_forms():
    """A static camera gives zero motion on every covered pixel (all valid); a previous eye
    moves every covered pixel; projecting the current view-projection of a pixel's position
    lands inside that pixel (u*w, v*h within the pixel's square)."""
    w, h = 96, 64
    sc = O.frame_test_scene(w, h)
    f = O.frame(sc, w, h, rtao_on=False)
    assert (f["motion"] == 0).all() and f["motion_valid"] == f["motion_total"] == int(f["mask"].sum())
    ps = dict(sc)
    ps["eye"] = np.array([4.3, 4.0, 5.7], np.float32)
    g = O.frame(sc, w, h, prev=O.view_proj(ps), rtao_on=False)
    m = g["mask"] != 0
    assert g["motion_valid"] == int(m.sum())
    assert (np.abs(g["motion"][m]).sum(axis=1) > 0).mean() > 0.99
    vp = O.view_proj(sc).reshape(4, 4).astype(np.float64)
    ys, xs = np.nonzero(m)
    p = np.concatenate([g["position"][m].astype(np.float64), np.ones((len(xs), 1))], axis=1) @ vp.T
    u = (p[:, 0] / p[:, 3] * 0.5 + 0.5) * w
    v = (1 - (p[:, 1] / p[:, 3] * 0.5 + 0.5)) * h
    assert (np.abs(u - (xs + 0.5)) <= 0.5 + 1e-3).all() and (np.abs(v - (ys + 0.5)) <= 0.5 + 1e-3).all()


def test_shade_closed_forms():
    """Model 2 with sky = sun = 1 equals albedo x (0.2 + max(0, n.-l) li) computed in float64
    within float32 rounding; model 1 with ao = 0 is black; frame = clip(hdr, 0, 1)."""
    w, h = 80, 60
    sc = O.frame_test_scene(w, h)
    f = O.frame(sc, w, h)
    L = O.lib()
    ones = np.ones((h, w), np.float32)
    fr, hdr = np.zeros((h, w, 3), np.float32), np.zeros((h, w, 3), np.float32)
    L.orc_shade(f["normal"], f["albedo"], f["mask"], w, h, 2, ones, ones, sc["light_dir"], 1.5, fr, hdr)
    ndl = np.maximum(0, -(f["normal"].astype(np.float64) @ sc["light_dir"].astype(np.float64))) * 1.5
    want = f["albedo"] * (0.2 + ndl)[..., None] * (f["mask"] != 0)[..., None]
    assert np.allclose(hdr, want, rtol=1e-6, atol=1e-7)
    assert np.array_equal(fr, np.clip(hdr, 0, 1))
    L.orc_shade(f["normal"], f["albedo"], f["mask"], w, h, 1, np.zeros((h, w), np.float32), ones, sc["light_dir"], 1.0, fr, hdr)
    assert (hdr == 0).all() and (fr == 0).all()
    assert f["max_radiance"] == float(f["hdr"].max())


def _multi_mesh_scene(w, h):
    """A seeded scene split into meshes of uneven size, each with its own albedo, a light of
    intensity 1.7 and a camera low among the objects."""
    import raw_ao.host as H
    vpos, vnrm, idx, _ = H.scene_seeded(H.SceneSpec(0xF4A3, 60, 0.4, 2, 0.1, 1.5, 14.0, 4.0))
    rng = np.random.default_rng(3)
    ends = np.sort(rng.choice(np.arange(1, idx.shape[0]), 6, replace=False)).astype(np.uint32)
    ends = np.append(ends, np.uint32(idx.shape[0]))
    return dict(vpos=vpos, vnrm=vnrm, idx=idx, mesh_end=ends, albedo=rng.uniform(0.1, 1.0, (len(ends), 3)).astype(np.float32),
                eye=np.array([6.0, 2.5, 7.0], np.float32), center=np.array([0.0, 0.8, 0.0], np.float32),
                up=np.array([0, 1, 0], np.float32), fovy=1.0, aspect=float(np.float32(w) / np.float32(h)), near=0.1,
                far=100.0, has_light=True, light_dir=O.f32_normalize((0.3, -1.0, 0.5)), light_intensity=1.7)


def _same_frame(a, b):
    for k, v in b.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(np.ascontiguousarray(a[k]).view(np.uint8), np.ascontiguousarray(v).view(np.uint8)), k
        else:
            assert a[k] == v, k


@needs_ref
@pytest.mark.parametrize("opts", [dict(), dict(lighting=1), dict(no_sun=True), dict(rtao_on=False)])
def test_frame_matches_reference_code(opts):
    """Every plane and scalar of renderWithParams (the reference's code) from the oracle's frame:
    the test scene with a moving camera, and a seeded multi-mesh scene."""
    w, h = 96, 72
    sc = O.frame_test_scene(w, h)
    ps = dict(sc)
    ps["eye"] = np.array([4.3, 4.0, 5.7], np.float32)
    prev = O.view_proj(ps)
    _same_frame(O.frame(sc, w, h, prev=prev, **opts), O.ref_frame(sc, w, h, prev=prev, **opts))
    sc = _multi_mesh_scene(80, 60)
    _same_frame(O.frame(sc, 80, 60, **opts), O.ref_frame(sc, 80, 60, **opts))


# ---- the R2 check's rays, scan and CPU traversal ---------------------------------------

@pytest.mark.parametrize("name", ["retro_room", "iso_street", "clip_stress", "dense_block", "soup"])
def test_r2_rays_and_scan_match_the_evidence(name):
    """makeRays + bruteForce reproduce the hit counts of the reference's R2 record
    (evidence/rt-r2-bvh-gpu-d3d12-rtx4090.json, 100,000 rays a scene); the CPU rt::closest over
    the oracle's PLOC tree finds the scan's nearest hit (t, u, v, triangle) on every ray."""
    from cases import TREES
    tris = O.scene_tris(name, 20000, 7)
    g = TREES[name]
    rays = O.r2_rays(name, tris.shape[0], g["r2_rays"])
    scan, t2 = O.scan_hits(tris, rays)
    assert int((scan["tri"] >= 0).sum()) == g["r2_hits"]
    assert ((t2 >= scan["t"]) | (scan["tri"] < 0)).all()
    b, l, root, _ = O.ploc(tris)
    n = 20000
    cpu = O.tree_closest(b, l, root, tris, rays[:n])
    assert np.array_equal(cpu.view(np.uint32), scan[:n].view(np.uint32))


def test_r2_camera_draw_order_is_pinned():
    """The other order of the camera ray's two draws misses the evidence's count."""
    from cases import TREES
    tris = O.scene_tris("iso_street")
    rays = O.r2_rays("iso_street", tris.shape[0], 100000, cam_first_u=True)
    assert int((O.scan_hits(tris, rays)[0]["tri"] >= 0).sum()) != TREES["iso_street"]["r2_hits"]


@needs_ref
def test_tree_closest_matches_reference_code():
    """orc_tree_closest against the reference's own rt::closest on its own PLOC tree."""
    tris = O.scene_tris("retro_room")
    rays = O.r2_rays("retro_room", tris.shape[0], 20000)
    rays[::3, 6] = 2.5   # some finite tMax
    b, l, root, _ = O.ploc(tris)
    mine = O.tree_closest(b, l, root, tris, rays)
    ref = O.tree_closest(None, None, None, tris, rays, impl=O.ref())
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
