"""Host-side pieces of the AO path (libraw_ao_host.so): the float64 row-major reconcile
sum, the raw-cert/2 writer, the reference's test scene, the seeded scene generator and
the G-buffer rasterizer that produce the path's inputs.
"""
import collections
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(os.path.dirname(HERE), "build", "libraw_ao_host.so")

# kSlabExit = 1 + 2 gamma(3), gamma(n) = n eps / (1 - n eps), eps = 2^-24 (raw/renderer/rt_bvh.hpp:34)
ROBUST_FAR_SCALE = float(np.float32(1.0) + np.float32(2.0) * (np.float32(3.0) * np.float32(5.96046448e-8))
                         / (np.float32(1.0) - np.float32(3.0) * np.float32(5.96046448e-8)))

ReconcileResult = collections.namedtuple("ReconcileResult", "error_map rmse max_error pixels within_tolerance")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("libraw_ao_host is not built (%s): run __graft_entry__.build()" % HOST_LIB)
        L = C.CDLL(HOST_LIB)
        L.rah_rmse.restype = C.c_float
        L.rah_rmse.argtypes = [_f32p, _u8p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.rah_certificate.restype = C.c_size_t
        L.rah_certificate.argtypes = [C.c_float, C.c_float, C.c_int32, C.c_float, C.c_int, C.c_int, C.c_char_p, C.c_char_p,
                                      C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_double, C.c_int,
                                      C.c_double, C.c_char_p, C.c_size_t]
        L.rah_test_scene.restype = C.c_size_t
        L.rah_test_scene.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rah_scene_seeded.restype = C.c_size_t
        L.rah_scene_seeded.argtypes = [C.c_uint64, C.c_int, C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int32)]
        L.rah_orbit_camera.restype = None
        L.rah_orbit_camera.argtypes = [C.c_uint64, C.c_float, C.c_float, _f32p, _f32p]
        L.rah_rasterize.restype = None
        L.rah_rasterize.argtypes = [_f32p, _f32p, _i32p, C.c_size_t, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_int, _f32p, _f32p, _f32p, _u8p]
        L.rah_emit_outputs.restype = C.c_int
        L.rah_emit_outputs.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]
        L.rah_verify.restype = C.c_int
        L.rah_verify.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.rah_gpu_certificate.restype = C.c_int
        L.rah_gpu_certificate.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p,
                                          C.c_double, C.c_double]
        L.rah_sha256.restype = None
        L.rah_sha256.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p]
        _lib = L
    return _lib


_F3 = C.c_float * 3


class CliParams(C.Structure):
    """rah_params: the CLI parameters a render's files record (raw/tools/cli_params.hpp)."""
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("eye", _F3), ("center", _F3), ("up", _F3),
                ("fovy", C.c_float), ("tolerance", C.c_float), ("rtao", C.c_int32), ("gpu", C.c_int32),
                ("gpu_backend", C.c_char_p), ("has_prev_eye", C.c_int32), ("has_prev_center", C.c_int32),
                ("has_prev_up", C.c_int32), ("prev_eye", _F3), ("prev_center", _F3), ("prev_up", _F3)]


class _Frame(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("depth", "position", "normal", "albedo", "mask", "motion", "ao_rt", "ao_ss",
                                          "error", "frame", "hdr")] + \
               [("rmse", C.c_float), ("max_error", C.c_float), ("pixels", C.c_int32), ("within", C.c_int32),
                ("motion_valid", C.c_int32), ("motion_total", C.c_int32)]


def cli_params(w=256, h=256, eye=(4, 4, 6), target=(0, 1, 0), up=(0, 1, 0), fovy=0.9, tolerance=0.12, rt=True,
               prev_eye=None, prev_target=None, prev_up=None, gpu_backend=None):
    p = CliParams(width=w, height=h, fovy=fovy, tolerance=tolerance, rtao=int(bool(rt)))
    p.eye, p.center, p.up = _F3(*eye), _F3(*target), _F3(*up)
    if gpu_backend is not None:
        p.gpu, p.gpu_backend = 1, gpu_backend.encode()
    for name, v in (("eye", prev_eye), ("center", prev_target), ("up", prev_up)):
        if v is not None:
            setattr(p, "has_prev_" + name, 1)
            setattr(p, "prev_" + name, _F3(*v))
    return p


def emit_outputs(out_dir, frame, params, backend="raw-native-cpu"):
    """Write a frame's files as the reference CLI does (frame.ppm, frame_hdr.pfm, the AO images,
    mask.pgm, certificate.json, channels.json, receipt.json). frame: a dict of planes and the
    reconcile / motion scalars (raw_ao.Context.render_frame's result)."""
    keep = []
    f = _frame_struct(frame, keep)
    err = C.create_string_buffer(512)
    rc = lib().rah_emit_outputs(str(out_dir).encode(), C.byref(params), C.byref(f), backend.encode(), err, 512)
    if rc != 0:
        raise RuntimeError("emit_outputs: %s" % err.value.decode())


def _frame_struct(frame, keep):
    def ptr(k, t):
        if frame.get(k) is None:
            return None
        a = np.ascontiguousarray(frame[k], t)
        keep.append(a)
        return a.ctypes.data

    f = _Frame(*[ptr(k, np.uint8 if k == "mask" else np.float32) for k in
                 ("depth", "position", "normal", "albedo", "mask", "motion", "ao_rt", "ao_ss", "error", "frame", "hdr")])
    f.rmse, f.max_error = frame["rmse"], frame["max_error"]
    f.pixels, f.within = frame["pixels"], int(bool(frame["within"]))
    f.motion_valid, f.motion_total = frame["motion_valid"], frame["motion_total"]
    return f


def gpu_certificate(out_dir, params, gpu=None, cpu=None, reason=None, adapter=None, gpu_ms=0.0, cpu_ms=0.0):
    """Write gpu_certificate.json (raw-gpu-cert/1) and gpu_receipt.json for a GPU frame against a
    CPU frame (dicts as emit_outputs takes; None: no frame). adapter: a dict of the raw_ao_adapter
    fields, None for this device's, {} for an empty one."""
    keep = []
    g = _frame_struct(gpu, keep) if gpu is not None else None
    c = _frame_struct(cpu, keep) if cpu is not None else None
    a = None
    if adapter is not None:
        import raw_ao
        a = raw_ao.Adapter(**{k: v.encode() for k, v in adapter.items()})
    rc = lib().rah_gpu_certificate(str(out_dir).encode(), C.byref(params), C.byref(g) if g is not None else None,
                                   C.byref(c) if c is not None else None, reason.encode() if reason else None,
                                   C.byref(a) if a is not None else None, gpu_ms, cpu_ms)
    if rc != 0:
        raise RuntimeError("gpu_certificate: could not write %s" % out_dir)


def verify(out_dir):
    """raw_native_cli verify DIR: (exit code, report)."""
    n = C.c_size_t()
    lib().rah_verify(str(out_dir).encode(), None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    rc = lib().rah_verify(str(out_dir).encode(), buf, n.value + 1, C.byref(n))
    return rc, buf.value.decode()


def rmse_of(err, mask):
    """float32 RMSE of the covered pixels: float64 row-major sum (src/cert/reconcile.cpp:9-18)."""
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = mask.shape
    px = C.c_int32()
    r = lib().rah_rmse(np.ascontiguousarray(err, np.float32), mask, w, h, None, C.byref(px))
    return np.float32(r), px.value


def certificate(rmse, max_error, pixels, tolerance=0.12, rt=64, ss=24, renderer="", params_json="{}", outputs=(),
                motion_coherence=None, hdr_headroom=None):
    """raw-cert/2 text (src/cert/certificate.cpp:38-85); absent channels are null."""
    names = (C.c_char_p * max(1, len(outputs)))(*[n.encode() for n, _ in outputs])
    hashes = (C.c_char_p * max(1, len(outputs)))(*[h.encode() for _, h in outputs])
    args = [C.c_float(rmse), C.c_float(max_error), int(pixels), C.c_float(tolerance), rt, ss, renderer.encode(),
            params_json.encode(), names, hashes, len(outputs), int(motion_coherence is not None),
            float(motion_coherence or 0.0), int(hdr_headroom is not None), float(hdr_headroom or 0.0)]
    n = lib().rah_certificate(*args, None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().rah_certificate(*args, buf, n + 1)
    return buf.value.decode()


def test_scene_mesh():
    """The reference's 14-triangle test scene (src/scene/scene.cpp:18-32): (vpos, vnrm, idx)."""
    n = lib().rah_test_scene(None, None, None)
    vpos = np.zeros((2 * n, 3), np.float32)
    vnrm = np.zeros((2 * n, 3), np.float32)
    idx = np.zeros((n, 3), np.int32)
    lib().rah_test_scene(vpos.ctypes.data, vnrm.ctypes.data, idx.ctypes.data)
    return vpos, vnrm, idx


SceneSpec = collections.namedtuple("SceneSpec", "seed objects p_sphere subdiv smin smax ground height")
# ~1M triangles: a 40 x 40 ground quad and 2,000 boxes and icospheres (subdivision 3,
# 1,280 triangles) in a 40 x 40 x 10 volume, sizes log-uniform in [0.05, 2.0].
SCENE_1M = SceneSpec(0x5EED, 2000, 0.385, 3, 0.05, 2.0, 40.0, 10.0)
# ~2M triangles: the same family with 4,000 objects (the 4K low-spp denoise workload).
SCENE_2M = SceneSpec(0x5EED2, 4000, 0.385, 3, 0.05, 2.0, 40.0, 10.0)


def scene_seeded(spec=SCENE_1M):
    """The seeded scene as (vpos, vnrm, idx, nspheres)."""
    L = lib()
    nv, ns = C.c_size_t(), C.c_int32()
    a = [spec.seed, spec.objects, spec.p_sphere, spec.subdiv, spec.smin, spec.smax, spec.ground, spec.height]
    n = L.rah_scene_seeded(*a, None, None, None, C.byref(nv), C.byref(ns))
    vpos = np.zeros((nv.value, 3), np.float32)
    vnrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((n, 3), np.int32)
    L.rah_scene_seeded(*a, vpos.ctypes.data, vnrm.ctypes.data, idx.ctypes.data, C.byref(nv), C.byref(ns))
    return vpos, vnrm, idx, ns.value


def mesh_tris(vpos, idx):
    return np.ascontiguousarray(vpos[idx].reshape(-1, 9), dtype=np.float32)


def orbit_camera(seed, radius=28.0, height=14.0, target=(0.0, 2.0, 0.0)):
    eye = np.zeros(3, np.float32)
    lib().rah_orbit_camera(seed, radius, height, np.asarray(target, np.float32), eye)
    return eye


def rasterize(vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h):
    """raw::rasterize with default options: (depth, position, normal, mask)."""
    depth = np.zeros((h, w), np.float32)
    pos = np.zeros((h, w, 3), np.float32)
    nrm = np.zeros((h, w, 3), np.float32)
    mask = np.zeros((h, w), np.uint8)
    lib().rah_rasterize(np.ascontiguousarray(vpos, np.float32), np.ascontiguousarray(vnrm, np.float32),
                        np.ascontiguousarray(idx, np.int32), idx.shape[0], np.asarray(eye, np.float32),
                        np.asarray(center, np.float32), np.asarray(up, np.float32), fovy, aspect, near, far, w, h,
                        depth, pos, nrm, mask)
    return depth, pos, nrm, mask


def cli_gbuffer(w=256, h=256, eye=(4, 4, 6), target=(0, 1, 0), fovy=0.9):
    """The G-buffer the reference CLI renders for its test scene (src/tools/run.cpp:12-26,
    src/tools/cli_params.cpp:12-18): aspect w/h, near 0.1, far 100, up (0,1,0)."""
    vpos, vnrm, idx = test_scene_mesh()
    aspect = float(np.float32(w) / np.float32(h))
    return rasterize(vpos, vnrm, idx, eye, target, (0, 1, 0), fovy, aspect, 0.1, 100.0, w, h)


def batched_workload(seed=0xBA7C):
    """The batched uneven workload: 8 seeded scenes of the 1M-triangle scene's family with
    50k-2M triangles (log-uniform), and 64 square views, sides log-uniform in [256, 2048] and
    16-128 rays a pixel (log-uniform), each from a seeded orbit camera of one of the scenes.
    Returns (specs, views) with views as (scene, side, spp, eye); build the meshes with
    scene_seeded(spec) and the G-buffers with batched_gbuffer."""
    rng = np.random.default_rng(seed)
    base = SCENE_1M
    tris_per_object = 0.385 * 1280 + 0.615 * 12
    specs = []
    for k in range(8):
        target = float(np.exp(rng.uniform(np.log(5e4), np.log(2e6))))
        specs.append(SceneSpec(0xBA7C00 + k, max(20, int(target / tris_per_object)), base.p_sphere, base.subdiv, base.smin,
                               base.smax, base.ground, base.height))
    views = []
    for v in range(64):
        side = int(round(np.exp(rng.uniform(np.log(256), np.log(2048)))))
        spp = int(round(np.exp(rng.uniform(np.log(16), np.log(128)))))
        views.append((int(rng.integers(0, 8)), side, spp, orbit_camera(1000 + v)))
    return specs, views


def batched_gbuffer(mesh, view):
    """One view's G-buffer (the host rasterizer; fovy 0.9, square, near 0.1, far 200, looking at
    (0, 2, 0)): (depth, position, normal, mask)."""
    vpos, vnrm, idx = mesh[:3]
    _, side, _, eye = view
    return rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9, 1.0, 0.1, 200.0, side, side)
