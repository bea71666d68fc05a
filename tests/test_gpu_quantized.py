"""The wide tree's 64 B quantized records (RAW_AO_OPT_QUANTIZED_NODES: 0 none, 1 the sky rays,
the default, 2 also the 64-spp AO rays), through the C ABI.

Each quantized slot box must contain the slot's 128 B box, with planes that are exact floats
(base + q * step, checked in float64, where those sums are exact), so that the kernels' fmaf
decode gives the planes themselves; the records' refs and masks must be the 128 B records'. The
AO (the golden frames' bytes, the oracle on the 1M-triangle scene's full 1080p frame) and the
visibility planes must then be the same bit for bit as with the 128 B records, and the counted
node visits may only grow (the boxes only grow)."""
import numpy as np
import pytest

import oracle_lib as O
import raw_ao
from cases import CASES, params, sha

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def big_frame(w, h, seed=1):
    vpos, vnrm, idx, _ = raw_ao.host.scene_seeded()
    eye = raw_ao.host.orbit_camera(seed)
    _, pos, nrm, mask = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9,
                                              float(np.float32(w) / np.float32(h)), 0.1, 200.0, w, h)
    return raw_ao.host.mesh_tris(vpos, idx), pos, nrm, mask


def check_records(r128, r64):
    """Containment, exact decode, refs and masks; returns the count of slots checked."""
    n = r128.shape[0]
    assert r64.shape == (n, 16)
    planes = r128[:, :24].view(np.float32).reshape(n, 6, 4).astype(np.float64)   # lo x, y, z, hi x, y, z; slot
    masks = r128[:, 28]
    assert np.array_equal(r64[:, 10], masks)
    assert np.array_equal(r64[:, 12:16], r128[:, 24:28])
    assert np.all(r64[:, 11] == 0)
    valid = (masks | (masks >> 4)) & 15
    exps = r64[:, 3]
    enter_all = exps >> 24
    assert np.all((enter_all == 0) | (enter_all == 15))
    base = r64[:, 0:3].view(np.float32).astype(np.float64)
    step = np.stack([np.ldexp(1.0, ((exps >> (8 * a)) & 255).astype(np.int64) - 127) for a in range(3)], 1)
    qlo = r64[:, 4:7]
    qhi = r64[:, 7:10]
    checked = 0
    for k in range(4):
        lo = base + ((qlo >> (8 * k)) & 255).astype(np.float64) * step
        hi = base + ((qhi >> (8 * k)) & 255).astype(np.float64) * step
        for p in (lo, hi):   # every decoded plane is a float: the kernels' fmaf gives it exactly
            assert np.array_equal(p.astype(np.float32).astype(np.float64), p)
        use = (((valid >> k) & 1) == 1) & (enter_all == 0)
        assert np.all(lo[use] <= planes[use, 0:3, k])
        assert np.all(hi[use] >= planes[use, 3:6, k])
        checked += int(use.sum())
    return checked


def test_records_contain_the_full_boxes(ctx):
    """On the 1M-triangle scene (513k records) and on seeded soups of other extents and offsets."""
    print("quantized records", flush=True)
    scenes = [big_frame(64, 48)[0]]
    rng = np.random.default_rng(5)
    for scale, offset, n in ((1.0, 0.0, 20000), (1e-3, 1e3, 5000), (1e4, -3e4, 5000)):
        c = rng.uniform(-1, 1, (n, 1, 3)) * scale * 50 + offset
        scenes.append((c + rng.normal(0, scale, (n, 3, 3))).astype(np.float32).reshape(n, 9))
    try:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 2)
        for tris in scenes:
            ctx.build_ploc(tris, 16)
            r128, r64 = ctx.read_wide_records()
            assert check_records(r128, r64) >= r128.shape[0] * 2
            assert int((r64[:, 3] >> 24).sum()) == 0   # finite scenes: every node quantized
    finally:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 1)


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_frames(ctx, name):
    """ao_rt.pfm's bytes from the quantized records equal the reference's."""
    p = params(name)
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    try:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 2)
        ctx.build_ploc(raw_ao.host.mesh_tris(vpos, idx), 16)
        ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
    finally:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 1)
    assert sha(O.pfm1_bytes(ao)) == CASES[name]["files"]["ao_rt.pfm"]


def test_full_frame_matches_oracle_and_full_records(ctx):
    """The 1M-triangle scene's 1920 x 1080 frame at 64 spp, radius 1.0: the AO from the
    quantized records equals the oracle's and the 128 B records' on every pixel, the visibility
    planes equal the 128 B records', and the counted visits are no fewer."""
    print("quantized 1080p", flush=True)
    tris, pos, nrm, mask = big_frame(1920, 1080)
    to_light = np.array([0.4, 1.0, 0.3], np.float32) / np.float32(np.sqrt(1.25))
    try:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 0)
        ctx.build_ploc(tris, 16)
        ctx.count_next_trace(True)
        want = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
        v_full, t_full, r_full = ctx.trace_counts()
        want_vis = ctx.compute_rt_visibility(pos, nrm, mask, to_light, 64, 1.0)
        # the default: the sky rays on the 64 B records, made by this visibility pass
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 1)
        sky_vis = ctx.compute_rt_visibility(pos, nrm, mask, to_light, 64, 1.0)
        # 2 set after the build: the AO rays' records are there already (the pass above made them)
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 2)
        ctx.count_next_trace(True)
        got = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
        v_q, t_q, r_q = ctx.trace_counts()
        got_vis = ctx.compute_rt_visibility(pos, nrm, mask, to_light, 64, 1.0)
        r128, r64 = ctx.read_wide_records()
        # 2 set after a build without them: made at the next trace call
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 0)
        ctx.build_ploc(tris, 16)
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 2)
        again = ctx.compute_rtao(pos, nrm, mask, 64, 1.0)
    finally:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 1)
    assert check_records(r128, r64) > 1_000_000
    assert np.array_equal(bits(got), bits(want))
    assert np.array_equal(bits(again), bits(want))
    for g, s, w in zip(got_vis, sky_vis, want_vis):
        assert np.array_equal(bits(g), bits(w))
        assert np.array_equal(bits(s), bits(w))
    assert r_q == r_full and t_q >= t_full and v_q >= v_full
    print("visits %d -> %d, tests %d -> %d" % (v_full, v_q, t_full, t_q), flush=True)
    ref = O.rtao(tris, pos, nrm, mask, 64, 1.0)
    assert np.array_equal(bits(got), bits(ref))


def test_small_and_empty_scenes(ctx):
    """One triangle, the 12-triangle test scene and no triangles, against the oracle; option
    values outside 0..2 are refused."""
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(96, 96)
    vpos, _, idx = raw_ao.host.test_scene_mesh()
    tris = raw_ao.host.mesh_tris(vpos, idx)
    try:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 2)
        for t in (tris, tris[:1], np.zeros((0, 9), np.float32)):
            ctx.build_ploc(t, 16)
            ao = ctx.compute_rtao(pos, nrm, mask, 64, 2.0)
            assert np.array_equal(bits(ao), bits(O.rtao(t, pos, nrm, mask, 64, 2.0)))
        ctx.build_ploc(tris[:1], 16)
        r128, r64 = ctx.read_wide_records()
        assert r128.shape[0] == 1 and check_records(r128, r64) == 1
    finally:
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 1)
    with pytest.raises(raw_ao.RawAoError):
        ctx.set_option(raw_ao.OPT_QUANTIZED_NODES, 3)
