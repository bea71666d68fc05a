"""The HIP path against the oracle and the reference's golden vectors, through the C ABI.
Integer and byte results (Morton keys, sort order, trees, AO counts) must be bit-exact;
AO images are compared bit for bit, with the accuracy bound of the AO path (within 1/64 on >= 99.9%
of pixels and RMSE <= 1e-3) asserted as well where a different tree could matter."""
import numpy as np
import pytest

import oracle_lib as O
import raw_ao
from cases import CASES, KNOWN, KNOWN_CASE, TREES, params, sha

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_angle_table_matches_host_libm(ctx):
    """Every one of the 2^24 sample angles: the device's sin/cos == the host libm's."""
    s, c = ctx.angle_table()
    ls, lc = O.libm_angles()
    assert np.array_equal(bits(s), bits(ls))
    assert np.array_equal(bits(c), bits(lc))


def test_angle_table_correctly_rounded_mode(ctx):
    """With the host-libm table off, the angles are the correctly rounded values."""
    ctx.set_option(raw_ao.OPT_HOST_LIBM_TRIG, 0)
    try:
        s, c = ctx.angle_table()
    finally:
        ctx.set_option(raw_ao.OPT_HOST_LIBM_TRIG, 1)
    k = np.arange(1 << 24, dtype=np.float32) / np.float32(1 << 24)
    phi = (np.float32(6.2831853) * k).astype(np.float64)
    assert np.array_equal(bits(c), bits(np.cos(phi).astype(np.float32)))
    assert np.array_equal(bits(s), bits(np.sin(phi).astype(np.float32)))


@pytest.mark.parametrize("scene", sorted(TREES))
def test_ploc_tree_matches_oracle(ctx, scene):
    """Morton keys, sorted order and the whole tree node for node; the golden hash."""
    t = O.scene_tris(scene, 20000, 7)
    ctx.build_ploc(t, 16)
    keys, order = ctx.read_keys(t.shape[0])
    okeys, oorder, _ = O.ploc_keys(t)
    assert np.array_equal(keys, okeys)
    assert np.array_equal(order, oorder)
    boxes, links, root, rounds = ctx.read_tree()
    ob, ol, oroot, orounds = O.ploc(t, 16)
    assert (root, rounds) == (oroot, orounds)
    assert np.array_equal(links, ol)
    assert np.array_equal(bits(boxes), bits(ob))
    assert "%016x" % O.tree_hash(boxes, links) == TREES[scene]["tree_hash"]


def test_ploc_radius_8(ctx):
    t = O.scene_tris("dense_block")
    ctx.build_ploc(t, 8)
    boxes, links, root, rounds = ctx.read_tree()
    ob, ol, oroot, orounds = O.ploc(t, 8)
    assert (root, rounds) == (oroot, orounds)
    assert np.array_equal(links, ol) and np.array_equal(bits(boxes), bits(ob))


def test_ploc_1m_scene_matches_oracle(ctx):
    vpos, _, idx, _ = raw_ao.host.scene_seeded()
    t = raw_ao.host.mesh_tris(vpos, idx)
    ctx.build_ploc(t, 16)
    boxes, links, root, rounds = ctx.read_tree()
    ob, ol, oroot, orounds = O.ploc(t, 16)
    assert (root, rounds) == (oroot, orounds)
    assert np.array_equal(links, ol) and np.array_equal(bits(boxes), bits(ob))


def test_ploc_edge_sizes(ctx):
    """1, 2, 3 and 257 triangles (a partial block and a halo at both ends)."""
    for n in (1, 2, 3, 257):
        t = O.scene_tris("soup", n, 3)
        ctx.build_ploc(t, 16)
        boxes, links, root, rounds = ctx.read_tree()
        ob, ol, oroot, orounds = O.ploc(t, 16)
        assert (root, rounds) == (oroot, orounds), n
        assert np.array_equal(links, ol) and np.array_equal(bits(boxes), bits(ob)), n


def golden_frame(ctx, name):
    p = params(name)
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    ctx.build_ploc(raw_ao.host.mesh_tris(vpos, idx), 16)
    ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    ss = ctx.compute_ssao(pos, nrm, mask, 24, 2.0)
    rec = ctx.reconcile(ss, ao, mask, 0.12)
    return mask, ao, ss, rec


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_frame(ctx, name):
    """ao_rt.pfm, ao_ss.pfm and the error map's bytes from the GPU equal the reference's."""
    mask, ao, ss, rec = golden_frame(ctx, name)
    f = CASES[name]["files"]
    assert sha(O.pfm1_bytes(ao)) == f["ao_rt.pfm"]
    assert sha(O.pfm1_bytes(ss)) == f["ao_ss.pfm"]
    assert sha(O.pgm_bytes(rec.error_map)) == f["ao_error.pgm"]


@pytest.mark.parametrize("short", sorted(KNOWN_CASE))
def test_golden_reconcile_and_verdict(ctx, short):
    mask, ao, ss, rec = golden_frame(ctx, KNOWN_CASE[short])
    k = KNOWN[short]
    assert rec.pixels == k["pixels"]
    assert "%.9g" % rec.rmse == "%.9g" % k["rmse"] and "%.9g" % rec.max_error == "%.9g" % k["maxError"]
    assert ("verified" if rec.within_tolerance else "refuted") == k["verdict"]
    cert = raw_ao.host.certificate(rec.rmse, rec.max_error, rec.pixels, 0.12)
    assert '"verdict":"%s"' % k["verdict"] in cert and '"oracle":"raw-rt-ao-v1"' in cert


def big_frame(w, h, seed=1):
    vpos, vnrm, idx, _ = raw_ao.host.scene_seeded()
    eye = raw_ao.host.orbit_camera(seed)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9,
                                              float(np.float32(w) / np.float32(h)), 0.1, 200.0, w, h)
    return raw_ao.host.mesh_tris(vpos, idx), pos, nrm, mask


def test_1m_scene_ao_matches_oracle(ctx):
    """The ~1M-triangle scene at its full 1920 x 1080 frame, radius 1.0, 64 spp (101 M rays):
    every pixel against the CPU reference restatement (binned-SAH BVH), bit for bit. The
    AO accuracy bound (within 1/64 on >= 99.9% of pixels, RMSE <= 1e-3) is asserted too."""
    print("1M-triangle scene, 1920x1080, against the oracle", flush=True)
    tris, pos, nrm, mask = big_frame(1920, 1080)
    ctx.build_ploc(tris, 16)
    ao = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    ref = O.rtao(tris, pos, nrm, mask, 64, 1.0)
    d = np.abs(ao.astype(np.float64) - ref.astype(np.float64))
    assert mask.sum() > 1_500_000
    assert np.mean(d <= 1.0 / 64 + 1e-9) >= 0.999
    assert np.sqrt(np.mean(d ** 2)) <= 1e-3
    assert np.array_equal(bits(ao), bits(ref))


@pytest.mark.parametrize("spp", [8, 16, 18, 20, 24, 32])
def test_low_spp_refill_kernel_on_1m_scene(ctx, spp):
    """8, 16, 18, 20, 24 and 32 spp run the refill kernel with 256 / spp pixels a wave: on the ~1M-triangle
    scene at 1920 x 1080, radius 1.0, every pixel equals the one-ray-a-lane kernel's (refill off)
    and, on every 16th row, the oracle's, bit for bit."""
    print("low-spp refill kernel", spp, flush=True)
    tris, pos, nrm, mask = big_frame(1920, 1080)
    ctx.build_ploc(tris, 16)
    ao = ctx.compute_rtao(pos, nrm, mask, spp, 1.0)
    ctx.set_option(raw_ao.OPT_REFILL, 0)
    try:
        plain = ctx.compute_rtao(pos, nrm, mask, spp, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_REFILL, 48)
    assert np.array_equal(bits(ao), bits(plain))
    ref = O.OracleBvh(tris).rtao(pos, nrm, mask, spp, 1.0, row_step=16)
    rows = slice(0, 1080, 16)
    assert np.array_equal(bits(ao[rows]), bits(ref[rows]))


@pytest.mark.parametrize("name", ["default", "close-512"])
def test_skew_control_breaks_the_golden(ctx, name):
    """Negative control (src/renderer/accel.cpp:6-11,26-28; .github/workflows/ao-oracle.yml):
    with every AO query distance x 0.99 the GPU's ao_rt.pfm no longer matches the golden hash,
    and equals the oracle's own skew mode bit for bit; switched back, it matches again."""
    p = params(name)
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    tris = raw_ao.host.mesh_tris(vpos, idx)
    ctx.build_ploc(tris, 16)
    ctx.set_option(raw_ao.OPT_SKEW, 1)
    try:
        skew = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    finally:
        ctx.set_option(raw_ao.OPT_SKEW, 0)
    assert sha(O.pfm1_bytes(skew)) != CASES[name]["files"]["ao_rt.pfm"]
    assert np.array_equal(bits(skew), bits(O.rtao(tris, pos, nrm, mask, 64, 2.0, mode=2)))
    ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    assert sha(O.pfm1_bytes(ao)) == CASES[name]["files"]["ao_rt.pfm"]


def test_skew_control_on_1m_scene(ctx):
    """The skew control on the 1M-triangle scene at radius 1.0 moves pixels the oracle's skew
    mode moves, bit for bit, on the wide kernel's default 64-spp path."""
    tris, pos, nrm, mask = big_frame(320, 180, seed=4)
    ctx.build_ploc(tris, 16)
    ctx.set_option(raw_ao.OPT_SKEW, 1)
    try:
        skew = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_SKEW, 0)
    plain = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    assert np.sum(bits(skew) != bits(plain)) > 0
    assert np.array_equal(bits(skew), bits(O.rtao(tris, pos, nrm, mask, 64, 1.0, mode=2)))


@pytest.mark.parametrize("spp", [1, 8, 16, 24, 32, 33, 128])
def test_spp_other_than_64(ctx, spp):
    p = params("default")
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(96, 96)
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    tris = raw_ao.host.mesh_tris(vpos, idx)
    ctx.build_ploc(tris, 16)
    ao = ctx.compute_rtao(pos, nrm, mask, spp, 2.0)
    assert np.array_equal(bits(ao), bits(O.rtao(tris, pos, nrm, mask, spp, 2.0)))
    assert p is not None


def test_empty_scene_is_all_open(ctx):
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(64, 64)
    ctx.build_ploc(np.zeros((0, 9), np.float32), 16)
    ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    assert np.all(ao == 1.0)


def make_rays(n, seed):
    """Random rays in the [-12, 12]^3 cube (src/renderer/gpu/rt_bvh_parity.cpp:35-48)."""
    rng = np.random.default_rng(seed)
    o = (rng.random((n, 3), dtype=np.float32) * 24 - 12).astype(np.float32)
    d = (rng.random((n, 3), dtype=np.float32) - np.float32(0.5)).astype(np.float32)
    d = (d * (np.float32(1) / np.sqrt((d * d).sum(1, keepdims=True)).astype(np.float32))).astype(np.float32)
    return o, d


@pytest.mark.parametrize("scene", ["soup", "dense_block"])
def test_trace_rays_closest_and_any_equal_scan(ctx, scene):
    """Closest hit: the same triangle and t as the scan; any hit: the same hit/miss."""
    t = O.scene_tris(scene, 20000, 7)
    ctx.build_ploc(t, 16)
    n = 20000 if scene == "soup" else 4000
    o, d = make_rays(n, 0xB2B2 + t.shape[0])
    tmax = np.full(n, 1e30, np.float32)
    rays = np.hstack([o, d, tmax[:, None], np.zeros((n, 1), np.float32)]).astype(np.float32)
    out = ctx.trace_rays(rays, any_hit=False)
    st, stri = O.scan_rays(t, np.hstack([o, d]), tmax, any_hit=False)
    assert np.array_equal(out[:, 1].astype(np.int32), stri)
    hit = stri >= 0
    assert np.array_equal(bits(out[hit, 0]), bits(st[hit]))
    any_max = (np.random.default_rng(77).random(n, dtype=np.float32) * 8).astype(np.float32)
    rays[:, 6] = any_max
    out = ctx.trace_rays(rays, any_hit=True)
    _, atri = O.scan_rays(t, np.hstack([o, d]), any_max, any_hit=True)
    assert np.array_equal(out[:, 1] >= 0, atri >= 0)


def test_short_far_control_loses_hits(ctx):
    """A slab exit shortened to 0.9 must lose hits (rt_bvh_parity.cpp:167-173)."""
    t = O.scene_tris("retro_room")
    ctx.build_ploc(t, 16)
    o, d = make_rays(20000, 0xB2B2 + t.shape[0])
    rays = np.hstack([o, d, np.full((20000, 1), 1e30, np.float32), np.zeros((20000, 1), np.float32)]).astype(np.float32)
    full = ctx.trace_rays(rays, any_hit=False)
    short = ctx.trace_rays(rays, any_hit=False, far_scale=0.9)
    assert np.sum((full[:, 1] >= 0) & (short[:, 1] != full[:, 1])) > 0


def test_error_map_matches_oracle(ctx):
    rng = np.random.default_rng(9)
    a = rng.random((123, 77), dtype=np.float32)
    b = rng.random((123, 77), dtype=np.float32)
    m = (rng.random((123, 77)) < 0.7).astype(np.uint8)
    rec = ctx.reconcile(a, b, m, 0.3)
    em, rmse, mx, px, within = O.reconcile(a, b, m, 0.3)
    assert np.array_equal(bits(rec.error_map), bits(em))
    assert (float(rec.rmse), rec.max_error, rec.pixels, bool(rec.within_tolerance)) == (rmse, mx, px, within)


@pytest.mark.parametrize("name", ["default-512", "close-512"])
def test_binary_kernel_golden(ctx, name):
    """The binary-tree kernel gives the same golden bytes as the default wide one."""
    ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_BINARY)
    try:
        mask, ao, ss, rec = golden_frame(ctx, name)
    finally:
        ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_WIDE)
    assert sha(O.pfm1_bytes(ao)) == CASES[name]["files"]["ao_rt.pfm"]


def test_binary_and_wide_agree_on_1m_scene(ctx):
    tris, pos, nrm, mask = big_frame(320, 180, seed=5)
    ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_BINARY)
    try:
        ctx.build_ploc(tris, 16)
        a_bin = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_WIDE)
    ctx.build_ploc(tris, 16)
    a_wide = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    assert np.array_equal(bits(a_bin), bits(a_wide))


def test_option_change_without_rebuild_is_an_error(ctx):
    t = O.scene_tris("soup", 100, 1)
    ctx.build_ploc(t, 16)
    ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_BINARY)
    try:
        with pytest.raises(raw_ao.RawAoError):
            ctx.trace_rays(np.zeros((1, 8), np.float32), any_hit=True)
    finally:
        ctx.set_option(raw_ao.OPT_TRACE_KERNEL, raw_ao.TRACE_WIDE)


def test_cpp_host_interface_default_frame(tmp_path):
    """The C++ host interface (raw::computeRTAO / computeSSAO / reconcile, as render.cpp
    calls them) reproduces the default render's golden bytes and known reconcile."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(raw_ao.LIB_PATH), "host_interface_frame")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = CASES["default"]["files"]
    assert sha((tmp_path / "ao_rt.pfm").read_bytes()) == f["ao_rt.pfm"]
    assert sha((tmp_path / "ao_ss.pfm").read_bytes()) == f["ao_ss.pfm"]
    assert "pixels=37996 rmse=0.129436031" in r.stdout
    cert = (tmp_path / "certificate.json").read_text()
    assert '"verdict":"refuted"' in cert and '["rmse","0.1294"]' in cert


def test_repeatable(ctx):
    """Two runs give the same bytes (no scheduling-dependent results)."""
    tris, pos, nrm, mask = big_frame(128, 72, seed=3)
    ctx.build_ploc(tris, 16)
    b1 = ctx.read_tree()
    a1 = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    ctx.build_ploc(tris, 16)
    b2 = ctx.read_tree()
    a2 = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    assert np.array_equal(b1[1], b2[1]) and np.array_equal(bits(b1[0]), bits(b2[0]))
    assert np.array_equal(bits(a1), bits(a2))


@pytest.mark.parametrize("rounds", [1, 3, 7, 200])
def test_enqueued_build_rounds_do_not_change_the_tree(ctx, rounds):
    """The build enqueues its PLOC rounds and collapse levels without reading the cluster count
    back; too few (1, 3, 7: the rest continue with read-backs) or far too many (200: the extra
    ones do nothing) give the same tree node for node, the same root and rounds, and the same AO
    as the automatic count."""
    tris, pos, nrm, mask = big_frame(160, 90, seed=13)
    ctx.build_ploc(tris, 16)
    want_tree = ctx.read_tree()
    want_ao = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    ctx.set_option(raw_ao.OPT_BUILD_ROUNDS, rounds)
    try:
        ctx.build_ploc(tris, 16)
        got_tree = ctx.read_tree()
        got_ao = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_BUILD_ROUNDS, 0)
    assert got_tree[2:] == want_tree[2:]
    assert np.array_equal(got_tree[1], want_tree[1]) and np.array_equal(bits(got_tree[0]), bits(want_tree[0]))
    assert np.array_equal(bits(got_ao), bits(want_ao))
    ob, ol, oroot, orounds = O.ploc(tris, 16)
    assert (got_tree[2], got_tree[3]) == (oroot, orounds) and np.array_equal(got_tree[1], ol)


def test_recorded_grids_with_another_scene_of_as_many_triangles(ctx):
    """A build sizes each PLOC round's and collapse level's grid by the counts the last build of
    as many triangles recorded. Another scene with the same triangle count (its rounds keep other
    cluster counts, so some grids are too small, some too large) still gets the oracle's tree,
    since the kernels stride over their tiles, and the same 4-wide tree as a fresh context's: AO
    equal bit for bit on points of its own triangles."""
    import raw_ao
    n = 200000
    scenes = [O.scene_tris("soup", n, 3), O.scene_tris("soup", n, 4)]
    r = np.random.default_rng(5)
    for k in (0, 1, 0, 1):
        print("recorded grids, soup seed", 3 + k, flush=True)
        tris = scenes[k]
        ctx.build_ploc(tris, 16)
        got = ctx.read_tree()
        ob, ol, oroot, orounds = O.ploc(tris, 16)
        assert (got[2], got[3]) == (oroot, orounds) and np.array_equal(got[1], ol)
        # a 64 x 64 G-buffer of points on the scene's own triangles, their normals
        pick = r.choice(n, 64 * 64, replace=False)
        v = tris[pick].reshape(-1, 3, 3).astype(np.float64)
        pos = v.mean(axis=1).astype(np.float32)
        nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        nrm = (nrm / np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)).astype(np.float32)
        mask = np.ones(64 * 64, np.uint8)
        ao = ctx.compute_rtao(pos.reshape(64, 64, 3), nrm.reshape(64, 64, 3), mask.reshape(64, 64), 64, 0.05)
        fresh = raw_ao.Context(0)
        try:
            fresh.build_ploc(tris, 16)
            want = fresh.compute_rtao(pos.reshape(64, 64, 3), nrm.reshape(64, 64, 3), mask.reshape(64, 64), 64, 0.05)
        finally:
            fresh.close()
        assert np.array_equal(bits(ao), bits(want))


def test_ploc_tail_threshold_does_not_change_the_tree(ctx):
    """The single-workgroup tail and the grid-wide rounds run the same rounds: the tree is the
    same whichever takes over where."""
    tris = O.scene_tris("dense_block", 60000, 5)
    trees = []
    try:
        for tail in (4096, 1000, 64, 1):
            ctx.set_option(raw_ao.OPT_PLOC_TAIL, tail)
            ctx.build_ploc(tris)
            trees.append(ctx.read_tree())
    finally:
        ctx.set_option(raw_ao.OPT_PLOC_TAIL, 1024)
    for b, l, root, rounds in trees[1:]:
        assert np.array_equal(b.view(np.uint32), trees[0][0].view(np.uint32)) and np.array_equal(l, trees[0][1])
        assert (root, rounds) == trees[0][2:]


@pytest.mark.parametrize("refill,cut", [(1, 2), (24, 2), (64, 2), (24, 0), (64, 0)])
def test_refill_kernel_matches(ctx, refill, cut):
    """The 64-spp refill kernel (a wave owns 4 pixels' rays and refills idle lanes; from the
    wave's cut of their union box, or from the root) gives the one-pixel-a-wave kernel's AO bit
    for bit: on the 1M-triangle scene (odd width, so a wave's pixels cross rows) and on the
    golden default frame."""
    tris, pos, nrm, mask = big_frame(243, 131, seed=3)
    ctx.build_ploc(tris, 16)
    ctx.set_option(raw_ao.OPT_REFILL, 0)
    base = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    ctx.set_option(raw_ao.OPT_REFILL, refill)
    ctx.set_option(raw_ao.OPT_WAVE_CUT, cut)
    try:
        got = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
        assert np.array_equal(bits(got), bits(base))
        p = params("default")
        _, pos2, nrm2, mask2 = raw_ao.host.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
        vpos, _, idx = raw_ao.host.test_scene_mesh()
        ctx.build_ploc(raw_ao.host.mesh_tris(vpos, idx), 16)
        ao = ctx.compute_rtao(pos2, nrm2, mask2, 64, 2.0)
        assert sha(O.pfm1_bytes(ao)) == CASES["default"]["files"]["ao_rt.pfm"]
        ctx.build_ploc(np.zeros((0, 9), np.float32), 16)
        assert np.all(ctx.compute_rtao(pos2, nrm2, mask2, 64, 2.0) == 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_REFILL, REFILL_DEFAULT)
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)


def test_near_first_order_does_not_change_ao(ctx):
    """The 64-spp kernel's child order (nearest first, or slot order) leaves every pixel's AO
    bit for bit the same."""
    tris, pos, nrm, mask = big_frame(160, 90, seed=11)
    ctx.build_ploc(tris, 16)
    outs = []
    try:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 0)   # the child order applies to the one-pixel kernel from the root
        ctx.set_option(raw_ao.OPT_REFILL, 0)
        for nf in (1, 0):
            ctx.set_option(raw_ao.OPT_NEAR_FIRST, nf)
            outs.append(ctx.compute_rtao(pos, nrm, mask, 64, 1.0))
    finally:
        ctx.set_option(raw_ao.OPT_NEAR_FIRST, 1)
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)
        ctx.set_option(raw_ao.OPT_REFILL, REFILL_DEFAULT)
    assert np.array_equal(bits(outs[0]), bits(outs[1]))


WAVE_CUT_DEFAULT = 2
REFILL_DEFAULT = 48


@pytest.mark.parametrize("k,spp", [(1, 64), (2, 64), (8, 64), (64, 64), (2, 4), (2, 33), (8, 1)])
def test_wave_cut_does_not_change_ao(ctx, k, spp):
    """Starting each pixel's rays from the wave's shared cut of the tree (cut at k x radius;
    k = 1 fills the cut lists on many pixels, which then end the cut at the frontier of the
    level that would overflow; below 64 spp a wave spans several pixels and walks for their
    union) leaves every pixel's AO bit for bit the same as the traversal from the root,
    counting launch included."""
    print("wave cut", k, spp, flush=True)
    tris, pos, nrm, mask = big_frame(160, 90, seed=11)
    ctx.build_ploc(tris, 16)
    try:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 0)
        want = ctx.compute_rtao(pos, nrm, mask, spp, 1.0)
        ctx.set_option(raw_ao.OPT_WAVE_CUT, k)
        got = ctx.compute_rtao(pos, nrm, mask, spp, 1.0)
        ctx.count_next_trace()
        counted = ctx.compute_rtao(pos, nrm, mask, spp, 1.0)
        visits, _, rays = ctx.trace_counts()
    finally:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)
    assert np.array_equal(bits(got), bits(want))
    assert np.array_equal(bits(counted), bits(want))
    assert rays == spp * int(mask.sum()) and visits > 0


def test_wave_cut_full_frame(ctx):
    """At the bench's full size (the 1M-triangle scene, 1920 x 1080, 64 spp, radius 1.0) the
    cut kernel's AO equals the traversal from the root on every pixel, and so does the
    visibility pass's (ao, sky, sun)."""
    print("wave cut 1080p", flush=True)
    tris, pos, nrm, mask = big_frame(1920, 1080, seed=1)
    ctx.build_ploc(tris, 16)
    to_light = np.array([0.4, 1.0, 0.3], np.float32) / np.float32(np.sqrt(1.25))
    try:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 0)
        want = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
        want_vis = ctx.compute_rt_visibility(pos, nrm, mask, to_light, 64, 1.0)
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)
        got = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
        got_vis = ctx.compute_rt_visibility(pos, nrm, mask, to_light, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)
    assert int(mask.sum()) > 1_000_000
    assert np.array_equal(bits(got), bits(want))
    for g, w in zip(got_vis, want_vis):
        assert np.array_equal(bits(g), bits(w))


def test_wave_cut_small_scenes(ctx):
    """The cut on the 12-triangle test scene (radius 2, the tree's leaves reached in the wave's
    walk) and on an empty tree."""
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(96, 96)
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    tris = raw_ao.host.mesh_tris(vpos, idx)
    try:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 4)
        for t in (tris, tris[:1], np.zeros((0, 9), np.float32)):
            ctx.build_ploc(t, 16)
            ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
            assert np.array_equal(bits(ao), bits(O.rtao(t, pos, nrm, mask, 64, 2.0)))
    finally:
        ctx.set_option(raw_ao.OPT_WAVE_CUT, WAVE_CUT_DEFAULT)
    with pytest.raises(raw_ao.RawAoError):
        ctx.set_option(raw_ao.OPT_WAVE_CUT, 65)


def test_timings_before_every_kernel_ran_leave_no_error():
    """raw_ao_last_ms before the SSAO and error-map kernels have ever run reports 0 for them,
    and the next launches do not see that query's error."""
    import raw_ao
    tris, pos, nrm, mask = big_frame(64, 48, seed=2)
    with raw_ao.Context(0) as c:
        c.build_ploc(tris, 16)
        m = c.last_ms()
        assert m["ssao"] == 0.0 and m["error"] == 0.0
        ao = c.compute_rtao(pos, nrm, mask, 64, 1.0)
        assert ao.shape == mask.shape


def test_ssao_full_frame_matches_oracle(ctx):
    """SSAO of the 1M-triangle scene's 1920x1080 G-buffer (about 38M taps, a few thousand of
    them within the tap margin of a half-integer, where the exact sin/cos pair is taken)
    against the oracle, bit for bit."""
    import raw_ao.host as H
    print("ssao full frame", flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded()
    _, pos, nrm, mask = H.rasterize(vpos, vnrm, idx, H.orbit_camera(1), (0, 2, 0), (0, 1, 0), 0.9,
                                    float(np.float32(1920) / np.float32(1080)), 0.1, 200.0, 1920, 1080)
    got = ctx.compute_ssao(pos, nrm, mask, 24, 1.0)
    want = O.ssao(pos, nrm, mask, 24, 1.0)
    assert np.array_equal(bits(got), bits(want))
    assert 0.0 < float(got[mask != 0].mean()) < 1.0
