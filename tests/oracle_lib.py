"""ctypes bindings of the CPU oracle (oracle/build/liboracle_ao.so) and of the reference
built from its own sources (oracle/_ref/libref_ao.so). Test infrastructure only: the
checker the HIP path is compared with. The product never imports this module.
"""
import ctypes as C
import os
import numpy as np

# the host threads the oracle uses: the machine's cores, at most 16 (a GPU box's CPU share)
NTHREADS = max(1, min(16, os.cpu_count() or 1))

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "liboracle_ao.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_ao.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
SZ = C.c_size_t

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_hash01.argtypes = [C.c_uint32] * 3
        L.orc_hash01.restype = C.c_float
        L.orc_ao_ray.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, f32p, f32p, f32p, f32p]
        L.orc_intersect.argtypes = [f32p, f32p, f32p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_intersect.restype = C.c_int
        L.orc_rasterize.argtypes = [f32p, f32p, i32p, SZ, f32p, f32p, f32p, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_int, f32p, f32p, f32p, u8p]
        L.orc_test_scene.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_test_scene.restype = SZ
        L.orc_bvh_build.argtypes = [f32p, SZ]
        L.orc_bvh_build.restype = C.c_void_p
        L.orc_bvh_free.argtypes = [C.c_void_p]
        L.orc_bvh_node_count.argtypes = [C.c_void_p]
        L.orc_bvh_node_count.restype = SZ
        L.orc_bvh_occluded.argtypes = [C.c_void_p, f32p, f32p, C.c_float]
        L.orc_bvh_occluded.restype = C.c_int
        L.orc_rtao.argtypes = [C.c_void_p, f32p, SZ, C.c_int, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_float, C.c_int, f32p]
        L.orc_rtao_rows.argtypes = [C.c_void_p, f32p, SZ, C.c_int, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_float, C.c_int, f32p]
        L.orc_ssao.argtypes = [f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, f32p]
        L.orc_rtvis.argtypes = [C.c_void_p, f32p, SZ, C.c_int, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_float,
                                f32p, C.c_int, C.c_int, f32p, f32p, f32p]
        L.orc_reconcile.argtypes = [f32p, f32p, u8p, C.c_int, C.c_int, C.c_float, f32p, C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.orc_ploc_build.argtypes = [f32p, SZ, C.c_int, f32p, i32p, C.POINTER(C.c_uint32)]
        L.orc_ploc_build.restype = C.c_int32
        L.orc_ploc_keys.argtypes = [f32p, SZ, u32p, u32p, f32p]
        L.orc_tree_hash.argtypes = [f32p, i32p, SZ]
        L.orc_tree_hash.restype = C.c_uint64
        L.orc_sah_cost.argtypes = [f32p, SZ, C.c_int32]
        L.orc_sah_cost.restype = C.c_double
        L.orc_binned_sah_cost.argtypes = [f32p, SZ, C.POINTER(SZ)]
        L.orc_binned_sah_cost.restype = C.c_double
        L.orc_scan_rays.argtypes = [f32p, SZ, f32p, f32p, SZ, C.c_int, C.c_int, f32p, i32p]
        L.orc_scene_tris.argtypes = [C.c_char_p, SZ, C.c_uint64, C.c_void_p]
        L.orc_scene_tris.restype = SZ
        L.orc_libm_angles.argtypes = [f32p, f32p]
        L.orc_r2_rays.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, SZ, C.c_uint64, C.c_int, f32p]
        L.orc_scan_hits.argtypes = [f32p, SZ, f32p, SZ, C.c_int, f32p, C.c_void_p]
        L.orc_tree_closest.argtypes = [f32p, i32p, C.c_int32, f32p, f32p, SZ, C.c_int, f32p]
        L.orc_rasterize_scene.argtypes = [f32p, f32p, i32p, SZ, u32p, f32p, SZ, f32p, f32p, f32p, C.c_float, C.c_float,
                                          C.c_float, C.c_float, C.c_int, C.c_int, f32p, f32p, f32p, f32p, u8p]
        L.orc_view_proj.argtypes = [f32p, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float, f32p]
        L.orc_shade.argtypes = [f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p, C.c_float, f32p, f32p]
        L.orc_max_radiance.argtypes = [f32p, SZ]
        L.orc_max_radiance.restype = C.c_float
        L.orc_motion.argtypes = [f32p, u8p, C.c_int, C.c_int, f32p, f32p, f32p]
        L.orc_motion.restype = C.c_int32
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_rasterize.argtypes = lib().orc_rasterize.argtypes
        L.ref_render_test_scene.argtypes = [C.c_int, C.c_int, f32p, f32p, C.c_float, f32p, f32p, u8p]
        L.ref_rtao.argtypes = [f32p, SZ, C.c_int, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, f32p]
        L.ref_ssao.argtypes = lib().orc_ssao.argtypes
        L.ref_reconcile.argtypes = lib().orc_reconcile.argtypes
        L.ref_ploc_build.argtypes = lib().orc_ploc_build.argtypes
        L.ref_ploc_build.restype = C.c_int32
        L.ref_scene_tris.argtypes = lib().orc_scene_tris.argtypes
        L.ref_scene_tris.restype = SZ
        L.ref_binned_sah_cost.argtypes = lib().orc_binned_sah_cost.argtypes
        L.ref_binned_sah_cost.restype = C.c_double
        L.ref_accel_build.argtypes = [f32p, SZ]
        L.ref_accel_build.restype = C.c_void_p
        L.ref_accel_free.argtypes = [C.c_void_p]
        L.ref_rtao_accel.argtypes = [C.c_void_p, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, f32p]
        L.ref_rtvis.argtypes = [f32p, SZ, f32p, f32p, u8p, C.c_int, C.c_int, C.c_int, C.c_float, f32p, C.c_int, C.c_int,
                                f32p, f32p, f32p]
        L.ref_tree_closest.argtypes = [f32p, SZ, f32p, SZ, f32p]
        L.ref_render_frame.argtypes = [f32p, f32p, SZ, i32p, SZ, u32p, f32p, SZ, f32p, f32p, f32p, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.c_int, f32p, C.c_float, C.c_int, C.c_int, C.c_void_p,
                                       C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, f32p, f32p, f32p, f32p, u8p, f32p,
                                       f32p, f32p, f32p, f32p, f32p, f32p, f32p]
        _ref = L
    return _ref


# ---- convenience wrappers ---------------------------------------------------------

def test_scene_mesh():
    """The 14-triangle test scene (src/scene/scene.cpp:18-32): (vpos, vnrm, idx)."""
    L = lib()
    n = L.orc_test_scene(None, None, None)
    vpos = np.zeros((4 * n // 2, 3), np.float32)
    vnrm = np.zeros_like(vpos)
    idx = np.zeros((n, 3), np.int32)
    L.orc_test_scene(vpos.ctypes.data, vnrm.ctypes.data, idx.ctypes.data)
    return vpos, vnrm, idx


def mesh_tris(vpos, idx):
    return np.ascontiguousarray(vpos[idx].reshape(-1, 9), dtype=np.float32)


def rasterize(vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h, impl=None):
    L = impl or lib()
    fn = L.ref_rasterize if impl is not None else L.orc_rasterize
    depth = np.zeros((h, w), np.float32)
    pos = np.zeros((h, w, 3), np.float32)
    nrm = np.zeros((h, w, 3), np.float32)
    mask = np.zeros((h, w), np.uint8)
    fn(np.ascontiguousarray(vpos, np.float32), np.ascontiguousarray(vnrm, np.float32),
       np.ascontiguousarray(idx, np.int32), idx.shape[0], np.asarray(eye, np.float32),
       np.asarray(center, np.float32), np.asarray(up, np.float32), fovy, aspect, near, far, w, h,
       depth, pos, nrm, mask)
    return depth, pos, nrm, mask


def cli_gbuffer(w=256, h=256, eye=(4, 4, 6), target=(0, 1, 0), fovy=0.9):
    """The G-buffer the reference CLI renders for the test scene (src/tools/run.cpp:12-26,
    src/tools/cli_params.cpp:12-18): aspect w/h, near 0.1, far 100, up (0,1,0)."""
    vpos, vnrm, idx = test_scene_mesh()
    aspect = np.float32(w) / np.float32(h)
    return rasterize(vpos, vnrm, idx, eye, target, (0, 1, 0), fovy, float(aspect), 0.1, 100.0, w, h)


class OracleBvh:
    def __init__(self, tris9):
        self.tris = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
        self.h = lib().orc_bvh_build(self.tris, self.tris.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_free(self.h)
            self.h = None

    def rtao(self, pos, nrm, mask, samples, radius, mode=0, threads=None, rows=None, row_step=1):
        """computeRTAO on rows y0, y0 + row_step, ... below y1 (rows=(y0, y1); all rows by
        default); the other rows are left at 1.0."""
        h, w = mask.shape
        y0, y1 = rows if rows is not None else (0, h)
        ao = np.ones((h, w), np.float32)
        lib().orc_rtao_rows(self.h, self.tris, self.tris.shape[0], mode, np.ascontiguousarray(pos, np.float32),
                            np.ascontiguousarray(nrm, np.float32), np.ascontiguousarray(mask, np.uint8), w, h, y0, y1,
                            row_step, samples, radius, threads or NTHREADS, ao)
        return ao


def rtvis(tris9, pos, nrm, mask, to_light, samples=64, radius=2.0, sun_forced=False, mode=0, threads=None, impl=None):
    """computeRTVisibility: (ao, sky, sun); impl=ref() runs the reference's own code (BVH mode)."""
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    h, w = mask.shape
    out = [np.zeros((h, w), np.float32) for _ in range(3)]
    args = (np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(nrm, np.float32), np.ascontiguousarray(mask, np.uint8),
            w, h, samples, radius, np.asarray(to_light, np.float32), int(bool(sun_forced)), threads or NTHREADS)
    if impl is not None:
        impl.ref_rtvis(tris9, tris9.shape[0], *args, *out)
    else:
        b = OracleBvh(tris9)
        lib().orc_rtvis(b.h, b.tris, b.tris.shape[0], mode, *args, *out)
    return tuple(out)


def rtao(tris9, pos, nrm, mask, samples=64, radius=2.0, mode=0, threads=None):
    return OracleBvh(tris9).rtao(pos, nrm, mask, samples, radius, mode, threads)


def ssao(pos, nrm, mask, samples=24, radius=2.0, threads=None):
    h, w = mask.shape
    ao = np.zeros((h, w), np.float32)
    lib().orc_ssao(np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(nrm, np.float32),
                   np.ascontiguousarray(mask, np.uint8), w, h, samples, radius, threads or NTHREADS, ao)
    return ao


def reconcile(approx, truth, mask, tol=0.12):
    h, w = mask.shape
    em = np.zeros((h, w), np.float32)
    rmse, mx = C.c_float(), C.c_float()
    px, within = C.c_int32(), C.c_int32()
    lib().orc_reconcile(np.ascontiguousarray(approx, np.float32), np.ascontiguousarray(truth, np.float32),
                        np.ascontiguousarray(mask, np.uint8), w, h, tol, em, C.byref(rmse), C.byref(mx),
                        C.byref(px), C.byref(within))
    return em, rmse.value, mx.value, px.value, bool(within.value)


def scene_tris(name, count=0, seed=0, impl=None):
    if impl is not None:
        n = impl.ref_scene_tris(name.encode(), count, seed, None)
        out = np.zeros((n, 9), np.float32)
        impl.ref_scene_tris(name.encode(), count, seed, out.ctypes.data)
        return out
    n = lib().orc_scene_tris(name.encode(), count, seed, None)
    out = np.zeros((n, 9), np.float32)
    lib().orc_scene_tris(name.encode(), count, seed, out.ctypes.data)
    return out


def ploc(tris9, radius=16, impl=None):
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    n = tris9.shape[0]
    nn = max(2 * n - 1, 1)
    boxes = np.zeros((nn, 6), np.float32)
    links = np.zeros((nn, 3), np.int32)
    rounds = C.c_uint32()
    fn = impl.ref_ploc_build if impl is not None else lib().orc_ploc_build
    root = fn(tris9, n, radius, boxes, links, C.byref(rounds))
    return boxes, links, int(root), rounds.value


def tree_hash(boxes, links):
    return lib().orc_tree_hash(np.ascontiguousarray(boxes, np.float32), np.ascontiguousarray(links, np.int32),
                               boxes.shape[0])


def ploc_keys(tris9):
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    n = tris9.shape[0]
    keys = np.zeros(n, np.uint32)
    order = np.zeros(n, np.uint32)
    g4 = np.zeros(4, np.float32)
    lib().orc_ploc_keys(tris9, n, keys, order, g4)
    return keys, order, g4


def scan_rays(tris9, rays6, tmax, any_hit, threads=None):
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    rays6 = np.ascontiguousarray(rays6, np.float32).reshape(-1, 6)
    n = rays6.shape[0]
    t = np.zeros(n, np.float32)
    tri = np.zeros(n, np.int32)
    lib().orc_scan_rays(tris9, tris9.shape[0], rays6, np.ascontiguousarray(tmax, np.float32), n, int(any_hit),
                        threads or NTHREADS, t, tri)
    return t, tri


def libm_angles():
    s = np.zeros(1 << 24, np.float32)
    c = np.zeros(1 << 24, np.float32)
    lib().orc_libm_angles(s, c)
    return s, c


# ---- the reference's file formats (src/core/image.cpp:34-69), for hashing -----------

def pfm1_bytes(img):
    h, w = img.shape
    return b"Pf\n%d %d\n-1.0\n" % (w, h) + np.ascontiguousarray(img[::-1], dtype="<f4").tobytes()


def pgm_bytes(img):
    h, w = img.shape
    v = np.clip(img.astype(np.float32), np.float32(0), np.float32(1))
    b = np.floor(v * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)
    return b"P5\n%d %d\n255\n" % (w, h) + b.tobytes()


def rgb8(img):
    """to8 per channel (src/core/image.cpp:9-10): (unsigned char)(clamp(v, 0, 1) * 255 + 0.5)."""
    v = np.clip(np.asarray(img, np.float32), np.float32(0), np.float32(1))
    return np.floor(v * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def ppm_bytes(img):
    """writePPM (src/core/image.cpp:11-18)."""
    h, w, _ = img.shape
    return b"P6\n%d %d\n255\n" % (w, h) + rgb8(img).tobytes()


def pfm3_bytes(img):
    """writePFM (src/core/image.cpp:40-52): rows bottom to top, little-endian float32."""
    h, w, _ = img.shape
    return b"PF\n%d %d\n-1.0\n" % (w, h) + np.ascontiguousarray(img[::-1], dtype="<f4").tobytes()


def mask_pgm_bytes(mask):
    h, w = mask.shape
    return b"P5\n%d %d\n255\n" % (w, h) + np.where(mask != 0, 255, 0).astype(np.uint8).tobytes()


# ---- the default frame (renderWithParams, src/renderer/render.cpp:71-105) ------------

def f32_normalize(v):
    """normalize (raw/math/vec.hpp) in float32: v * (1 / sqrt(dot(v, v)))."""
    v = np.asarray(v, np.float32)
    length = np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), dtype=np.float32)
    return (v * (np.float32(1.0) / length)).astype(np.float32)


def frame_test_scene(w, h, eye=(4, 4, 6), target=(0, 1, 0), fovy=0.9):
    """buildTestScene (src/scene/scene.cpp:18-32) as a frame scene: the plane mesh (albedo 0.7)
    and the box mesh (0.8, 0.3, 0.2), one light; the camera as cameraFromParams makes it
    (src/tools/cli_params.cpp:12-18), aspect (float)w / (float)h."""
    vpos, vnrm, idx = test_scene_mesh()
    return dict(vpos=vpos, vnrm=vnrm, idx=idx, mesh_end=np.array([2, idx.shape[0]], np.uint32),
                albedo=np.array([[0.7, 0.7, 0.7], [0.8, 0.3, 0.2]], np.float32),
                eye=np.array(eye, np.float32), center=np.array(target, np.float32),
                up=np.array([0, 1, 0], np.float32), fovy=float(np.float32(fovy)), aspect=float(np.float32(w) / np.float32(h)),
                near=0.1, far=100.0, has_light=True, light_dir=f32_normalize((-0.4, -1.0, -0.3)), light_intensity=1.0)


def view_proj(sc):
    out = np.zeros(16, np.float32)
    lib().orc_view_proj(sc["eye"], sc["center"], sc["up"], sc["fovy"], sc["aspect"], sc["near"], sc["far"], out)
    return out


def frame(sc, w, h, prev=None, lighting=2, no_sun=False, rtao_on=True, tol=0.12, threads=None):
    """The oracle's frame: renderWithParams restated from the oracle's passes, in its order."""
    L = lib()
    vpos, vnrm, idx = (np.ascontiguousarray(sc[k], t) for k, t in (("vpos", np.float32), ("vnrm", np.float32), ("idx", np.int32)))
    ntri = idx.shape[0]
    depth = np.zeros((h, w), np.float32)
    pos, nrm, alb = (np.zeros((h, w, 3), np.float32) for _ in range(3))
    mask = np.zeros((h, w), np.uint8)
    L.orc_rasterize_scene(vpos, vnrm, idx, ntri, np.ascontiguousarray(sc["mesh_end"], np.uint32),
                          np.ascontiguousarray(sc["albedo"], np.float32), len(sc["mesh_end"]), sc["eye"], sc["center"],
                          sc["up"], sc["fovy"], sc["aspect"], sc["near"], sc["far"], w, h, depth, pos, nrm, alb, mask)
    vp = view_proj(sc)
    prev = vp if prev is None else np.ascontiguousarray(prev, np.float32)
    motion = np.zeros((h, w, 2), np.float32)
    valid = L.orc_motion(pos, mask, w, h, vp, prev, motion)
    ldir = np.asarray(sc["light_dir"], np.float32) if sc["has_light"] else np.array([0, -1, 0], np.float32)
    li = float(sc["light_intensity"]) if sc["has_light"] else 1.0
    model = lighting if rtao_on else 1
    out = dict(depth=depth, position=pos, normal=nrm, albedo=alb, mask=mask, motion=motion, viewproj=vp,
               motion_valid=int(valid), motion_total=int(mask.sum()))
    if rtao_on:
        tris = mesh_tris(vpos, idx)
        to_light = (ldir * np.float32(-1.0)).astype(np.float32)
        if model == 2:
            out["ao_rt"], out["sky"], out["sun"] = rtvis(tris, pos, nrm, mask, to_light, 64, 2.0, no_sun, threads=threads)
        else:
            out["ao_rt"] = rtao(tris, pos, nrm, mask, 64, 2.0, threads=threads)
    out["ao_ss"] = ssao(pos, nrm, mask, 24, 2.0, threads=threads)
    if rtao_on:
        out["error"], out["rmse"], out["max_error"], out["pixels"], out["within"] = reconcile(out["ao_ss"], out["ao_rt"], mask,
                                                                                            tol)
    else:
        out["error"], out["rmse"], out["max_error"], out["pixels"], out["within"] = (np.zeros((h, w), np.float32), 0.0,
                                                                                    0.0, 0, False)
    plane_a = out["sky"] if model == 2 else (out["ao_rt"] if rtao_on else out["ao_ss"])
    plane_sun = out["sun"] if model == 2 else plane_a
    fr, hdr = np.zeros((h, w, 3), np.float32), np.zeros((h, w, 3), np.float32)
    L.orc_shade(nrm, alb, mask, w, h, model, np.ascontiguousarray(plane_a), np.ascontiguousarray(plane_sun), ldir, li, fr, hdr)
    out["frame"], out["hdr"] = fr, hdr
    out["max_radiance"] = float(L.orc_max_radiance(hdr, w * h))
    return out


def ref_frame(sc, w, h, prev=None, lighting=2, no_sun=False, rtao_on=True, tol=0.12, threads=None):
    """renderWithParams itself (the reference's sources, oracle/_ref), the same planes."""
    vpos, vnrm, idx = (np.ascontiguousarray(sc[k], t) for k, t in (("vpos", np.float32), ("vnrm", np.float32), ("idx", np.int32)))
    depth = np.zeros((h, w), np.float32)
    pos, nrm, alb, fr, hdr = (np.zeros((h, w, 3), np.float32) for _ in range(5))
    mask = np.zeros((h, w), np.uint8)
    motion = np.zeros((h, w, 2), np.float32)
    ao_rt, ao_ss, em = (np.zeros((h, w), np.float32) for _ in range(3))
    st, vp = np.zeros(8, np.float32), np.zeros(16, np.float32)
    prev_p = None if prev is None else np.ascontiguousarray(prev, np.float32).ctypes.data
    ref().ref_render_frame(vpos, vnrm, vpos.shape[0], idx, idx.shape[0], np.ascontiguousarray(sc["mesh_end"], np.uint32),
                           np.ascontiguousarray(sc["albedo"], np.float32), len(sc["mesh_end"]), sc["eye"], sc["center"],
                           sc["up"], sc["fovy"], sc["aspect"], sc["near"], sc["far"], int(bool(sc["has_light"])),
                           np.asarray(sc["light_dir"], np.float32), float(sc["light_intensity"]), w, h, prev_p, lighting,
                           int(bool(no_sun)), int(bool(rtao_on)), tol, threads or NTHREADS, depth, pos, nrm,
                           alb, mask, motion, ao_rt, ao_ss, em, fr, hdr, st, vp)
    out = dict(depth=depth, position=pos, normal=nrm, albedo=alb, mask=mask, motion=motion, ao_ss=ao_ss, error=em,
               frame=fr, hdr=hdr, rmse=float(st[0]), max_error=float(st[1]), pixels=int(st[2]), within=bool(st[3]),
               motion_valid=int(st[4]), motion_total=int(st[5]), max_radiance=float(st[6]), viewproj=vp)
    if rtao_on:
        out["ao_rt"] = ao_rt
    return out


def sha256_file(path):
    import hashlib
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


# ---- the R2 check (src/renderer/gpu/rt_bvh_parity.cpp) -------------------------------

# The scenes' cameras (eye, target; up (0, 1, 0)): src/renderer/swr_scenes.cpp:41,65,80 and
# src/renderer/rt_scenes.cpp:35. The soup has none.
R2_CAMERAS = {"retro_room": ((0.0, 1.6, 1.0), (0.3, 1.2, -8.0)), "iso_street": ((24.0, 20.0, 24.0), (0.0, 0.5, 0.0)),
              "clip_stress": ((0.0, 0.25, 0.0), (0.2, 0.12, -3.0)), "dense_block": ((58.0, 40.0, 62.0), (0.0, 2.0, 0.0)),
              "soup": None}
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<i4")])


def r2_rays(name, ntri, n, cam_first_u=False):
    """makeRays (rt_bvh_parity.cpp:36-48) with seed 0xb2b2 + the triangle count; the camera ray's
    two draws in the order the reference's R2 evidence was made with (the up offset's first)."""
    rays = np.zeros((n, 8), np.float32)
    cam = R2_CAMERAS[name]
    if cam is None:
        lib().orc_r2_rays(None, None, None, n, 0xB2B2 + ntri, int(cam_first_u), rays)
    else:
        e, t, u = (np.array(v, np.float32) for v in (cam[0], cam[1], (0.0, 1.0, 0.0)))
        lib().orc_r2_rays(e.ctypes.data, t.ctypes.data, u.ctypes.data, n, 0xB2B2 + ntri, int(cam_first_u), rays)
    return rays


def scan_hits(tris9, rays8, threads=None):
    """bruteForce: (hits as HIT records, second-nearest t)."""
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    rays8 = np.ascontiguousarray(rays8, np.float32).reshape(-1, 8)
    n = rays8.shape[0]
    hits = np.zeros(n, HIT)
    t2 = np.zeros(n, np.float32)
    lib().orc_scan_hits(tris9, tris9.shape[0], rays8, n, threads or NTHREADS, hits.view(np.float32), t2.ctypes.data)
    return hits, t2


def tree_closest(boxes, links, root, tris9, rays8, threads=None, impl=None):
    """rt::closest over a given tree (impl=ref(): the reference's own, on its own PLOC tree)."""
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    rays8 = np.ascontiguousarray(rays8, np.float32).reshape(-1, 8)
    hits = np.zeros(rays8.shape[0], HIT)
    if impl is not None:
        impl.ref_tree_closest(tris9, tris9.shape[0], rays8, rays8.shape[0], hits.view(np.float32))
    else:
        lib().orc_tree_closest(np.ascontiguousarray(boxes, np.float32), np.ascontiguousarray(links, np.int32), root, tris9,
                               rays8, rays8.shape[0], threads or NTHREADS, hits.view(np.float32))
    return hits
