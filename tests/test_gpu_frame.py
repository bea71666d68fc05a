"""The default frame on the GPU (raw_ao_render_frame, replacing renderGpu / renderWithParams,
src/renderer/render.cpp:71-105) against the oracle's frame, which tests/test_oracle.py pins to the
reference's code and golden hashes: every plane bit for bit (depth, position, normal, albedo,
mask, motion, ray-traced AO, SSAO, sky, sun, error map, frame, HDR), the reconcile, the motion
counts and maxRadiance; the golden frame.ppm / frame_hdr.pfm hashes of the reference's identity
renders straight from the GPU frame; the device-pointer form; and, at the ~1M-triangle 1080p
scale, the frame's composition (shade, motion, reconcile recomputed by the oracle from the GPU's
own planes)."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from cases import FRAMES, params, sha

pytestmark = pytest.mark.gpu

SCALARS = ("rmse", "max_error", "pixels", "within", "motion_valid", "motion_total", "max_radiance")


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def check_frame(got, want):
    for k, v in want.items():
        if k == "viewproj":
            assert same(got[k], v), k
        elif isinstance(v, np.ndarray):
            assert got[k].shape == v.shape and same(got[k], v), k
        elif k in SCALARS:
            assert got[k] == v, (k, got[k], v)


def case(name):
    p = params(name)
    sc = O.frame_test_scene(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    prev = None
    if p["prev_eye"]:
        ps = dict(sc)
        ps["eye"] = np.array(p["prev_eye"], np.float32)
        prev = O.view_proj(ps)
    return p, sc, prev


@pytest.mark.parametrize("lighting,key", [(2, "files"), (1, "files_lighting_v1")])
@pytest.mark.parametrize("name", ["default", "motion-256", "nort-256", "close-512", "wide-512"])
def test_golden_frames(ctx, name, lighting, key):
    print("frame", name, "lighting", lighting, flush=True)
    p, sc, prev = case(name)
    got = ctx.render_frame(sc, p["w"], p["h"], prev=prev, lighting=lighting, rtao=p["rt"])
    g, ao_files = FRAMES[name][key], FRAMES[name]["files"]
    assert sha(O.ppm_bytes(got["frame"])) == g["frame.ppm"]
    assert sha(O.pfm3_bytes(got["hdr"])) == g["frame_hdr.pfm"]
    assert sha(O.mask_pgm_bytes(got["mask"])) == ao_files["mask.pgm"]
    assert sha(O.pfm1_bytes(got["ao_ss"])) == ao_files["ao_ss.pfm"]
    if p["rt"]:
        assert sha(O.pfm1_bytes(got["ao_rt"])) == ao_files["ao_rt.pfm"]
    check_frame(got, O.frame(sc, p["w"], p["h"], prev=prev, lighting=lighting, rtao_on=p["rt"]))


@pytest.mark.parametrize("opts", [dict(no_sun=True), dict(lighting=1), dict()])
def test_multi_mesh_scene_moving_camera(ctx, opts):
    import raw_ao.host as H
    print("frame multi-mesh", opts, flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded(H.SceneSpec(0xF4A3, 60, 0.4, 2, 0.1, 1.5, 14.0, 4.0))
    rng = np.random.default_rng(3)
    ends = np.append(np.sort(rng.choice(np.arange(1, idx.shape[0]), 6, replace=False)), idx.shape[0]).astype(np.uint32)
    w, h = 160, 120
    sc = dict(vpos=vpos, vnrm=vnrm, idx=idx, mesh_end=ends, albedo=rng.uniform(0.1, 1.0, (len(ends), 3)).astype(np.float32),
              eye=np.array([6.0, 2.5, 7.0], np.float32), center=np.array([0.0, 0.8, 0.0], np.float32),
              up=np.array([0, 1, 0], np.float32), fovy=1.0, aspect=float(np.float32(w) / np.float32(h)), near=0.1, far=100.0,
              has_light=True, light_dir=O.f32_normalize((0.3, -1.0, 0.5)), light_intensity=1.7)
    ps = dict(sc)
    ps["eye"] = np.array([6.4, 2.4, 6.6], np.float32)
    prev = O.view_proj(ps)
    got = ctx.render_frame(sc, w, h, prev=prev, **opts)
    check_frame(got, O.frame(sc, w, h, prev=prev, lighting=opts.get("lighting", 2), no_sun=opts.get("no_sun", False)))
    assert 0 < got["motion_valid"] <= got["motion_total"] == int(got["mask"].sum())


def test_no_light_uses_the_default_light(ctx):
    print("frame default light", flush=True)
    w, h = 64, 64
    sc = dict(O.frame_test_scene(w, h), has_light=False)
    check_frame(ctx.render_frame(sc, w, h), O.frame(sc, w, h))


def test_device_form_on_a_torch_stream(ctx):
    print("frame dev form", flush=True)
    p, sc, prev = case("motion-256")
    w, h = p["w"], p["h"]
    want = ctx.render_frame(sc, w, h, prev=prev)
    dev = torch.device("cuda:0")
    dsc = dict(sc, vpos=torch.from_numpy(sc["vpos"]).to(dev), vnrm=torch.from_numpy(sc["vnrm"]).to(dev),
               idx=torch.from_numpy(sc["idx"]).to(dev))
    shapes = dict(depth=(h, w), position=(h, w, 3), normal=(h, w, 3), albedo=(h, w, 3), mask=(h, w), motion=(h, w, 2),
                  ao_rt=(h, w), ao_ss=(h, w), sky=(h, w), sun=(h, w), error=(h, w), frame=(h, w, 3), hdr=(h, w, 3))
    planes = {k: torch.full(s, 3, dtype=torch.uint8 if k == "mask" else torch.float32, device=dev) for k, s in shapes.items()}
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        st = ctx.render_frame_dev(dsc, w, h, planes, prev=prev, stream=stream.cuda_stream)
    stream.synchronize()
    for k, t in planes.items():
        assert same(t.cpu().numpy(), want[k]), k
    for k in SCALARS:
        assert st[k] == want[k], k


def test_argument_errors(ctx):
    import raw_ao
    w, h = 32, 32
    sc = O.frame_test_scene(w, h)
    for bad in (dict(mesh_end=np.array([2, 13], np.uint32)), dict(mesh_end=np.array([3, 2, 14], np.uint32),
                                                                  albedo=np.ones((3, 3), np.float32)),
                dict(near=0.0)):
        with pytest.raises(raw_ao.RawAoError):
            ctx.render_frame(dict(sc, **bad), w, h)
    with pytest.raises(raw_ao.RawAoError):
        ctx.render_frame(sc, w, h, lighting=3)
    check_frame(ctx.render_frame(sc, w, h), O.frame(sc, w, h))


def test_million_triangle_frame_composition(ctx):
    """At the bench's scale the oracle recomputes the frame's last passes from the GPU's own
    planes: shade and maxRadiance from normal, albedo, mask, sky and sun; motion from position;
    the reconcile from the two AO planes. The G-buffer equals the oracle rasterizer's."""
    import raw_ao.host as H
    print("frame 1M 1080p", flush=True)
    vpos, vnrm, idx, _ = H.scene_seeded(H.SCENE_1M)
    w, h = 1920, 1080
    rng = np.random.default_rng(5)
    nm = 40
    ends = np.append(np.sort(rng.choice(np.arange(1, idx.shape[0]), nm - 1, replace=False)), idx.shape[0]).astype(np.uint32)
    eye = H.orbit_camera(1)
    sc = dict(vpos=vpos, vnrm=vnrm, idx=idx, mesh_end=ends, albedo=rng.uniform(0.1, 1.0, (nm, 3)).astype(np.float32),
              eye=eye, center=np.array([0, 2, 0], np.float32), up=np.array([0, 1, 0], np.float32), fovy=0.9,
              aspect=float(np.float32(w) / np.float32(h)), near=0.1, far=200.0, has_light=True,
              light_dir=O.f32_normalize((-0.4, -1.0, -0.3)), light_intensity=1.0)
    ps = dict(sc)
    ps["eye"] = (eye + np.array([0.3, 0.0, -0.2], np.float32)).astype(np.float32)
    prev = O.view_proj(ps)
    got = ctx.render_frame(sc, w, h, prev=prev)
    d, pos, nrm, mask = O.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9, sc["aspect"], 0.1, 200.0, w, h)
    assert same(got["mask"], mask) and same(got["depth"], d) and same(got["position"], pos) and same(got["normal"], nrm)
    L = O.lib()
    fr, hdr = np.zeros((h, w, 3), np.float32), np.zeros((h, w, 3), np.float32)
    L.orc_shade(got["normal"], got["albedo"], got["mask"], w, h, 2, got["sky"], got["sun"], sc["light_dir"], 1.0, fr, hdr)
    assert same(got["frame"], fr) and same(got["hdr"], hdr)
    assert got["max_radiance"] == L.orc_max_radiance(hdr, w * h)
    motion = np.zeros((h, w, 2), np.float32)
    assert L.orc_motion(got["position"], got["mask"], w, h, O.view_proj(sc), prev, motion) == got["motion_valid"]
    assert same(got["motion"], motion)
    em, rmse, mx, px, within = O.reconcile(got["ao_ss"], got["ao_rt"], got["mask"], 0.12)
    assert same(got["error"], em) and got["rmse"] == rmse and got["max_error"] == mx and got["pixels"] == px
    assert got["within"] == within and px == got["motion_total"] == int(mask.sum())
    # albedo: the mesh of each covered pixel's nearest triangle has that albedo somewhere in the table
    a = got["albedo"][mask != 0]
    table = {tuple(r) for r in sc["albedo"].view(np.uint32)}
    assert all(tuple(r) in table for r in np.unique(a.view(np.uint32), axis=0))
    assert (got["albedo"][mask == 0] == 0).all()


@pytest.mark.parametrize("name", ["default", "motion-256", "nort-256", "wide-1440x900"])
def test_cli_files_from_the_gpu_frame(ctx, tmp_path, name):
    """The GPU frame written as the CLI writes it: every file of the reference's identity render
    (images, certificate.json, channels.json, receipt.json) hashes as the golden records, and
    verify passes."""
    import raw_ao.host as H
    print("frame files", name, flush=True)
    p, sc, prev = case(name)
    got = ctx.render_frame(sc, p["w"], p["h"], prev=prev, rtao=p["rt"])
    H.emit_outputs(tmp_path, got, H.cli_params(p["w"], p["h"], p["eye"], p["target"], fovy=p["fovy"], rt=p["rt"],
                                               prev_eye=p["prev_eye"]))
    g = FRAMES[name]["files"]
    assert {k: sha((tmp_path / k).read_bytes()) for k in g} == g
    rc, report = H.verify(tmp_path)
    assert rc == 0, report


@pytest.mark.parametrize("name,extra", [("default", []), ("motion-256", ["4.3", "4", "5.7"])])
def test_cpp_render_gpu_and_cli_files(tmp_path, name, extra):
    """The C++ host interface end to end: raw::renderGpu of the test scene, raw::emitOutputs,
    raw::verifyDir (tests/cpp/host_interface_frame.cpp, frame mode)."""
    import os
    import subprocess
    import raw_ao
    print("cpp frame", name, flush=True)
    exe = os.path.join(os.path.dirname(raw_ao.LIB_PATH), "host_interface_frame")
    r = subprocess.run([exe, str(tmp_path), "256", "256", "frame"] + extra, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.endswith("verify: all checks match\n")
    g = FRAMES[name]["files"]
    assert {k: sha((tmp_path / k).read_bytes()) for k in g} == g


def test_gpu_run_certificate(ctx, tmp_path):
    """The --gpu run (src/tools/gpu_run.cpp:70-93): the GPU frame against the CPU reference frame
    (here the oracle's) as raw-gpu-cert/1 with this device's adapter: identity MATCH, every
    channel at zero error, verified."""
    import json
    import raw_ao.host as H
    print("gpu certificate", flush=True)
    p, sc, prev = case("motion-256")
    g = ctx.render_frame(sc, p["w"], p["h"], prev=prev)
    c = O.frame(sc, p["w"], p["h"], prev=prev)
    H.gpu_certificate(tmp_path, H.cli_params(p["w"], p["h"], prev_eye=p["prev_eye"], gpu_backend="hip"), gpu=g, cpu=c,
                      gpu_ms=g["gpu_ms"], cpu_ms=0.0)
    cert = json.loads((tmp_path / "gpu_certificate.json").read_text())
    assert cert["verdict"] == "verified" and cert["identity"] == "MATCH"
    assert all(ch["rmse"] == 0 and ch["pass"] for ch in cert["channels"])
    assert cert["adapter"]["backend"] == "hip" and cert["adapter"]["architecture"].startswith("gfx950")
    rec = json.loads((tmp_path / "gpu_receipt.json").read_text())
    assert rec["reconcile"]["identity"] == "MATCH" and rec["reconcile"]["tolerance"]["verdict"] == "verified"
