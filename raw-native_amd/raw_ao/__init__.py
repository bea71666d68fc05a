"""Python host side of raw_ao_hip: a ctypes binding of the C ABI (include/raw_ao_hip.h)
that mirrors the reference's AO/reconcile interface:

    computeRTAO  (raw/renderer/ray_ao.hpp:6-8)       -> Context.compute_rtao
    computeSSAO  (raw/renderer/ssao.hpp:5-6)         -> Context.compute_ssao
    reconcile    (raw/cert/reconcile.hpp:12-13)      -> Context.reconcile
    buildPlocGpu (raw/renderer/rt_gpu.hpp:19)        -> Context.build_ploc / read_tree
    traceGpu     (raw/renderer/rt_gpu.hpp:27-28)     -> Context.trace_rays

Arrays are numpy (host) or torch tensors on the GPU (the *_dev calls). Every call goes to
the HIP library; there is no CPU fallback: without the built library or a GPU this
module raises.
"""
import ctypes as C
import os

import numpy as np

from . import host  # noqa: F401  (float64 reconcile sum, certificate, scenes, rasterizer)

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "build", "libraw_ao_hip.so")

OK, ERR_ARG, ERR_DEVICE, ERR_BUILD, ERR_NO_DEVICE = 0, -1, -2, -3, -4
OPT_HOST_LIBM_TRIG, OPT_TRACE_KERNEL, OPT_TRI_PHASE, OPT_CONV_KERNEL, OPT_PLOC_TAIL, OPT_REFILL = 1, 2, 3, 4, 5, 6
OPT_NEAR_FIRST = 7
OPT_WAVE_CUT = 8
OPT_SKEW = 9
OPT_BUILD_ROUNDS = 10
OPT_SKY_REFILL = 11
OPT_LEARNED_BACKOFF = 12
OPT_QUANTIZED_NODES = 13
NOALLOC_MAX_PIXELS = 1 << 22
TRACE_WIDE, TRACE_BINARY = 0, 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_lib = None

# (name, restype, argtypes) of every symbol include/raw_ao_hip.h declares
SIGNATURES = [
    ("raw_ao_abi_version", C.c_uint32, []),
    ("raw_ao_hip_api", C.c_void_p, [C.c_uint32]),
    ("raw_ao_open", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("raw_ao_close", None, [_vp]),
    ("raw_ao_last_error", C.c_char_p, [_vp]),
    ("raw_ao_set_option", C.c_int, [_vp, C.c_int, C.c_int]),
    ("raw_ao_build_ploc", C.c_int, [_vp, _f32p, C.c_uint32, C.c_int]),
    ("raw_ao_build_ploc_dev", C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _vp]),
    ("raw_ao_tree_info", C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]),
    ("raw_ao_read_tree", C.c_int, [_vp, _f32p, _i32p]),
    ("raw_ao_read_keys", C.c_int, [_vp, _u32p, _u32p]),
    ("raw_ao_trace_ao", C.c_int, [_vp, _f32p, _f32p, _u8p, C.c_int, C.c_int, C.c_int, C.c_float, _f32p]),
    ("raw_ao_trace_ao_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp]),
    ("raw_ao_trace_ao_dev_noalloc", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp]),
    ("raw_ao_trace_visibility", C.c_int, [_vp, _f32p, _f32p, _u8p, C.c_int, C.c_int, C.c_int, C.c_float, _f32p, C.c_int,
                                          _f32p, _f32p, _f32p]),
    ("raw_ao_trace_visibility_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _f32p, C.c_int,
                                              _vp, _vp, _vp, _vp]),
    ("raw_ao_rasterize", C.c_int, [_vp, _f32p, _f32p, C.c_uint32, _i32p, C.c_uint32, _f32p, _f32p, _f32p, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _f32p, _f32p, _f32p, _u8p]),
    ("raw_ao_rasterize_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    ("raw_ao_render_frame", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    ("raw_ao_render_frame_dev", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, _vp]),
    ("raw_ao_render_opts_default", None, [C.c_void_p]),
    ("raw_ao_adapter_info", C.c_int, [_vp, C.c_void_p]),
    ("raw_ao_ssao", C.c_int, [_vp, _f32p, _f32p, _u8p, C.c_int, C.c_int, C.c_int, C.c_float, _f32p]),
    ("raw_ao_ssao_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp]),
    ("raw_ao_error_map", C.c_int, [_vp, _f32p, _f32p, _u8p, C.c_int, C.c_int, _f32p, C.POINTER(C.c_float)]),
    ("raw_ao_error_map_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp]),
    ("raw_ao_trace_rays", C.c_int, [_vp, _f32p, C.c_uint32, C.c_int, C.c_float, _f32p]),
    ("raw_ao_closest_hits", C.c_int, [_vp, _f32p, C.c_uint32, C.c_void_p]),
    ("raw_ao_closest_hits_dev", C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp]),
    ("raw_ao_count_next_trace", C.c_int, [_vp, C.c_int]),
    ("raw_ao_trace_counts", C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("raw_ao_last_ms", C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
    ("raw_ao_angle_table", C.c_int, [_vp, _f32p, _f32p]),
    ("raw_ao_bvh_bytes", C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    ("raw_ao_read_wide_records", C.c_int, [_vp, _vp, _vp, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("raw_ao_batch_open", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("raw_ao_batch_close", None, [_vp]),
    ("raw_ao_batch_last_error", C.c_char_p, [_vp]),
    ("raw_ao_batch_add_scene", C.c_int, [_vp, _vp, C.c_uint32, C.c_int, C.POINTER(C.c_int)]),
    ("raw_ao_batch_add_view", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                        C.c_float, _vp, _vp, _vp, C.POINTER(C.c_int)]),
    ("raw_ao_batch_set_tuning", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("raw_ao_batch_set_option", C.c_int, [_vp, C.c_int, C.c_int]),
    ("raw_ao_batch_get_tuning", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("raw_ao_batch_run", C.c_int, [_vp, C.c_int, _vp]),
    ("raw_ao_batch_capture", C.c_int, [_vp, C.c_int]),
    ("raw_ao_batch_launch", C.c_int, [_vp, _vp]),
    ("raw_ao_batch_check", C.c_int, [_vp]),
    ("raw_ao_batch_results", C.c_int, [_vp, _f32p, _f32p, _u32p, C.POINTER(C.c_uint64)]),
    ("raw_ao_batch_tune", C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_float)]),
    ("raw_ao_net_param_count", C.c_int, [_vp, C.POINTER(C.c_uint64)]),
    ("raw_ao_net_set_params", C.c_int, [_vp, _f32p, C.c_uint64]),
    ("raw_ao_net_get_params", C.c_int, [_vp, _f32p, C.c_uint64]),
    ("raw_ao_net_get_grads", C.c_int, [_vp, _f32p, C.c_uint64]),
    ("raw_ao_net_features_dev", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _f32p, C.c_int, C.c_int, _vp, _vp]),
    ("raw_ao_net_forward_dev", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    ("raw_ao_net_train_step_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp]),
    ("raw_ao_net_loss", C.c_int, [_vp, C.POINTER(C.c_float)]),
    ("raw_ao_net_load", C.c_int, [_vp, C.c_char_p]),
    ("raw_ao_learned_ao_dev", C.c_int, [_vp, _vp, _vp, _vp, _vp, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                        C.c_float, C.c_int, C.c_float, _vp, C.c_void_p, _vp]),
    ("raw_ao_learned_ao", C.c_int, [_vp, _f32p, _f32p, _u8p, C.c_void_p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                    C.c_int, C.c_float, C.c_int, C.c_float, _f32p, C.c_void_p]),
    ("raw_ao_net_read_act", C.c_int, [_vp, C.c_int, _u16p, C.c_uint64]),
    ("raw_ao_conv3x3_dev", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, _vp, _vp]),
    ("raw_ao_conv3x3_wgrad_dev", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           _vp, _vp, _vp]),
]


class LearnedReport(C.Structure):
    """raw_ao_learned_report"""
    _fields_ = [("probe_pixels", C.c_uint32), ("probe_rmse", C.c_float), ("probe_max_error", C.c_float),
                ("fell_back", C.c_int32)]

    def as_dict(self):
        return dict(probe_pixels=self.probe_pixels, probe_rmse=self.probe_rmse, probe_max_error=self.probe_max_error,
                    fell_back=bool(self.fell_back), network_skipped=self.fell_back == 2)


_F3 = C.c_float * 3


class SceneDesc(C.Structure):
    """raw_ao_scene"""
    _fields_ = [("vpos", C.c_void_p), ("vnrm", C.c_void_p), ("nverts", C.c_uint32), ("idx", C.c_void_p),
                ("ntri", C.c_uint32), ("mesh_end", C.c_void_p), ("albedo", C.c_void_p), ("nmesh", C.c_uint32),
                ("eye", _F3), ("center", _F3), ("up", _F3), ("fovy", C.c_float), ("aspect", C.c_float),
                ("near_z", C.c_float), ("far_z", C.c_float), ("has_light", C.c_int32), ("light_dir", _F3),
                ("light_intensity", C.c_float)]


class RenderOpts(C.Structure):
    """raw_ao_render_opts"""
    _fields_ = [("tolerance", C.c_float), ("rtao", C.c_int32), ("lighting", C.c_int32), ("no_sun", C.c_int32),
                ("rt_samples", C.c_int32), ("ss_samples", C.c_int32), ("radius", C.c_float), ("bvh_radius", C.c_int32),
                ("prev_viewproj", C.c_void_p)]


FRAME_PLANES = ("depth", "position", "normal", "albedo", "mask", "motion", "ao_rt", "ao_ss", "sky", "sun", "error",
                "frame", "hdr")


class FramePlanes(C.Structure):
    """raw_ao_frame"""
    _fields_ = [(k, C.c_void_p) for k in FRAME_PLANES]


class FrameStats(C.Structure):
    """raw_ao_frame_stats"""
    _fields_ = [("rmse", C.c_float), ("max_error", C.c_float), ("pixels", C.c_int32), ("within_tolerance", C.c_int32),
                ("motion_valid", C.c_int32), ("motion_total", C.c_int32), ("max_radiance", C.c_float),
                ("viewproj", C.c_float * 16), ("gpu_ms", C.c_float)]

    def as_dict(self):
        return dict(rmse=self.rmse, max_error=self.max_error, pixels=self.pixels, within=bool(self.within_tolerance),
                    motion_valid=self.motion_valid, motion_total=self.motion_total, max_radiance=self.max_radiance,
                    viewproj=np.array(self.viewproj[:], np.float32), gpu_ms=self.gpu_ms)


class Adapter(C.Structure):
    """raw_ao_adapter"""
    _fields_ = [("vendor", C.c_char * 64), ("architecture", C.c_char * 64), ("device", C.c_char * 128),
                ("description", C.c_char * 128), ("backend", C.c_char * 16), ("driver", C.c_char * 64)]


class RawAoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("raw_ao_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded HIP library. Raises when it is not built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("raw_ao_hip is not built (%s): run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(t):
    """Device pointer of a torch tensor (contiguous, on the GPU)."""
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return C.c_void_p(t.data_ptr())


def _optr(t):
    """_ptr, or NULL for None."""
    return _vp(None) if t is None else _ptr(t)


class Context:
    """One raw_ao_hip context (device memory, a HIP stream, the built BVH)."""

    def __init__(self, device=0):
        self._h = _vp()
        L = lib()
        rc = L.raw_ao_open(device, C.byref(self._h))
        if rc != OK:
            msg = L.raw_ao_last_error(self._h).decode() if self._h else "open failed"
            self.close()
            raise RawAoError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().raw_ao_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise RawAoError(rc, lib().raw_ao_last_error(self._h).decode())

    def set_option(self, opt, value):
        self._check(lib().raw_ao_set_option(self._h, opt, int(value)))

    # ---- scene ---------------------------------------------------------------------
    def build_ploc(self, tris9, radius=16):
        t = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
        self._check(lib().raw_ao_build_ploc(self._h, t, t.shape[0], radius))

    def build_ploc_dev(self, tris9_tensor, radius=16, stream=None):
        n = tris9_tensor.numel() // 9
        self._check(lib().raw_ao_build_ploc_dev(self._h, _ptr(tris9_tensor), n, radius, _vp(stream)))

    def tree_info(self):
        nn, root, rounds = C.c_uint32(), C.c_int32(), C.c_uint32()
        self._check(lib().raw_ao_tree_info(self._h, C.byref(nn), C.byref(root), C.byref(rounds)))
        return nn.value, root.value, rounds.value

    def read_tree(self):
        nn, root, rounds = self.tree_info()
        boxes = np.zeros((max(nn, 1), 6), np.float32)
        links = np.zeros((max(nn, 1), 3), np.int32)
        self._check(lib().raw_ao_read_tree(self._h, boxes, links))
        return boxes[:nn], links[:nn], root, rounds

    def read_keys(self, ntri):
        keys = np.zeros(max(ntri, 1), np.uint32)
        order = np.zeros(max(ntri, 1), np.uint32)
        self._check(lib().raw_ao_read_keys(self._h, keys, order))
        return keys[:ntri], order[:ntri]

    # ---- AO ------------------------------------------------------------------------
    def compute_rtao(self, position, normal, mask, samples=64, radius=2.0):
        mask = np.ascontiguousarray(mask, np.uint8)
        h, w = mask.shape
        ao = np.zeros((h, w), np.float32)
        self._check(lib().raw_ao_trace_ao(self._h, np.ascontiguousarray(position, np.float32),
                                          np.ascontiguousarray(normal, np.float32), mask, w, h, samples, radius, ao))
        return ao

    def compute_rtao_dev(self, position, normal, mask, w, h, samples, radius, out, stream=None):
        self._check(lib().raw_ao_trace_ao_dev(self._h, _ptr(position), _ptr(normal), _ptr(mask), w, h, samples, radius,
                                              _ptr(out), _vp(stream)))

    def compute_rtao_dev_noalloc(self, position, normal, mask, w, h, samples, radius, out, stream=None):
        """compute_rtao_dev through the entry point that allocates nothing (one scratch array per
        device, at most NOALLOC_MAX_PIXELS pixels)."""
        self._check(lib().raw_ao_trace_ao_dev_noalloc(self._h, _ptr(position), _ptr(normal), _ptr(mask), w, h, samples,
                                                      radius, _ptr(out), _vp(stream)))

    def compute_rt_visibility(self, position, normal, mask, to_light, samples=64, radius=2.0, sun_forced=False):
        """computeRTVisibility: (ao, sky, sun)."""
        mask = np.ascontiguousarray(mask, np.uint8)
        h, w = mask.shape
        out = [np.zeros((h, w), np.float32) for _ in range(3)]
        self._check(lib().raw_ao_trace_visibility(self._h, np.ascontiguousarray(position, np.float32),
                                                  np.ascontiguousarray(normal, np.float32), mask, w, h, samples, radius,
                                                  np.asarray(to_light, np.float32), int(bool(sun_forced)), *out))
        return tuple(out)

    def compute_rt_visibility_dev(self, position, normal, mask, w, h, samples, radius, to_light, ao, sky, sun,
                                  sun_forced=False, stream=None):
        self._check(lib().raw_ao_trace_visibility_dev(self._h, _ptr(position), _ptr(normal), _ptr(mask), w, h, samples,
                                                      radius, np.asarray(to_light, np.float32), int(bool(sun_forced)),
                                                      _ptr(ao), _ptr(sky), _ptr(sun), _vp(stream)))

    def adapter_info(self):
        a = Adapter()
        self._check(lib().raw_ao_adapter_info(self._h, C.byref(a)))
        return {k: getattr(a, k).decode() for k, _ in Adapter._fields_}

    # ---- the default frame ---------------------------------------------------------
    @staticmethod
    def _scene_desc(sc, keep, ptr):
        """raw_ao_scene from a dict (vpos, vnrm, idx, mesh_end, albedo, eye, center, up, fovy, aspect,
        near, far, has_light, light_dir, light_intensity); ptr maps the mesh arrays to addresses."""
        me = np.ascontiguousarray(sc["mesh_end"], np.uint32)
        al = np.ascontiguousarray(sc["albedo"], np.float32).reshape(-1, 3)
        keep += [me, al]
        d = SceneDesc()
        d.vpos, d.vnrm, d.idx = ptr(sc["vpos"]), ptr(sc["vnrm"]), ptr(sc["idx"])
        d.nverts = int(np.prod(sc["vpos"].shape)) // 3
        d.ntri = int(np.prod(sc["idx"].shape)) // 3
        d.mesh_end, d.albedo, d.nmesh = me.ctypes.data, al.ctypes.data, len(me)
        d.eye, d.center, d.up = (_F3(*[float(v) for v in sc[k]]) for k in ("eye", "center", "up"))
        d.fovy, d.aspect, d.near_z, d.far_z = sc["fovy"], sc["aspect"], sc["near"], sc["far"]
        d.has_light = int(bool(sc.get("has_light", False)))
        d.light_dir = _F3(*[float(v) for v in sc.get("light_dir", (0.0, -1.0, 0.0))])
        d.light_intensity = float(sc.get("light_intensity", 1.0))
        return d

    @staticmethod
    def _render_opts(keep, prev, lighting, no_sun, rtao, tolerance, rt_samples, ss_samples, radius):
        o = RenderOpts()
        lib().raw_ao_render_opts_default(C.byref(o))
        o.tolerance, o.rtao, o.lighting, o.no_sun = tolerance, int(bool(rtao)), lighting, int(bool(no_sun))
        o.rt_samples, o.ss_samples, o.radius = rt_samples, ss_samples, radius
        if prev is not None:
            pv = np.ascontiguousarray(prev, np.float32).reshape(16)
            keep.append(pv)
            o.prev_viewproj = pv.ctypes.data
        return o

    def render_frame(self, sc, w, h, prev=None, lighting=2, no_sun=False, rtao=True, tolerance=0.12, rt_samples=64,
                     ss_samples=24, radius=2.0):
        """renderWithParams on the GPU: a dict of host planes and the frame's scalars."""
        keep = []

        def host(a, t):
            a = np.ascontiguousarray(a, t)
            keep.append(a)
            return a.ctypes.data

        sc = dict(sc, vpos=np.ascontiguousarray(sc["vpos"], np.float32), vnrm=np.ascontiguousarray(sc["vnrm"], np.float32),
                  idx=np.ascontiguousarray(sc["idx"], np.int32))
        d = self._scene_desc(sc, keep, lambda a: host(a, a.dtype))
        o = self._render_opts(keep, prev, lighting, no_sun, rtao, tolerance, rt_samples, ss_samples, radius)
        shapes = dict(depth=(h, w), position=(h, w, 3), normal=(h, w, 3), albedo=(h, w, 3), mask=(h, w), motion=(h, w, 2),
                      ao_rt=(h, w), ao_ss=(h, w), sky=(h, w), sun=(h, w), error=(h, w), frame=(h, w, 3), hdr=(h, w, 3))
        out = {k: np.zeros(shapes[k], np.uint8 if k == "mask" else np.float32) for k in FRAME_PLANES}
        planes = FramePlanes(*[out[k].ctypes.data for k in FRAME_PLANES])
        st = FrameStats()
        self._check(lib().raw_ao_render_frame(self._h, C.byref(d), C.byref(o), w, h, C.byref(planes), C.byref(st)))
        if not rtao:
            del out["ao_rt"]
        if not (rtao and lighting == 2):
            del out["sky"], out["sun"]
        out.update(st.as_dict())
        return out

    def render_frame_dev(self, sc, w, h, planes, prev=None, lighting=2, no_sun=False, rtao=True, tolerance=0.12,
                         rt_samples=64, ss_samples=24, radius=2.0, stream=None):
        """The same with the mesh (vpos, vnrm, idx) and every plane in device memory (torch tensors);
        returns the frame's scalars."""
        keep = []
        d = self._scene_desc(sc, keep, lambda t: t.data_ptr())
        o = self._render_opts(keep, prev, lighting, no_sun, rtao, tolerance, rt_samples, ss_samples, radius)
        fp = FramePlanes(*[planes[k].data_ptr() if planes.get(k) is not None else None for k in FRAME_PLANES])
        st = FrameStats()
        self._check(lib().raw_ao_render_frame_dev(self._h, C.byref(d), C.byref(o), w, h, C.byref(fp), C.byref(st),
                                                  _vp(stream)))
        return st.as_dict()

    # ---- G-buffer ------------------------------------------------------------------
    def rasterize(self, vpos, vnrm, idx, eye, center, up, fovy, aspect, near, far, w, h):
        """raw::rasterize of one indexed mesh: (depth, position, normal, mask)."""
        vpos = np.ascontiguousarray(vpos, np.float32).reshape(-1, 3)
        vnrm = np.ascontiguousarray(vnrm, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        depth = np.zeros((h, w), np.float32)
        pos = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        mask = np.zeros((h, w), np.uint8)
        f3 = lambda v: np.ascontiguousarray(v, np.float32).reshape(3)
        self._check(lib().raw_ao_rasterize(self._h, vpos, vnrm, vpos.shape[0], idx, idx.shape[0], f3(eye), f3(center), f3(up),
                                           fovy, aspect, near, far, w, h, depth, pos, nrm, mask))
        return depth, pos, nrm, mask

    def rasterize_dev(self, vpos, vnrm, idx, ntri, eye, center, up, fovy, aspect, near, far, w, h, depth, pos, nrm, mask,
                      stream=None):
        f3 = lambda v: np.ascontiguousarray(v, np.float32).reshape(3)
        self._check(lib().raw_ao_rasterize_dev(self._h, _ptr(vpos), _ptr(vnrm), _ptr(idx), ntri, f3(eye), f3(center), f3(up),
                                               fovy, aspect, near, far, w, h, _ptr(depth), _ptr(pos), _ptr(nrm), _ptr(mask),
                                               _vp(stream)))

    def compute_ssao(self, position, normal, mask, samples=24, radius=2.0):
        mask = np.ascontiguousarray(mask, np.uint8)
        h, w = mask.shape
        ao = np.zeros((h, w), np.float32)
        self._check(lib().raw_ao_ssao(self._h, np.ascontiguousarray(position, np.float32),
                                      np.ascontiguousarray(normal, np.float32), mask, w, h, samples, radius, ao))
        return ao

    def compute_ssao_dev(self, position, normal, mask, w, h, samples, radius, out, stream=None):
        self._check(lib().raw_ao_ssao_dev(self._h, _ptr(position), _ptr(normal), _ptr(mask), w, h, samples, radius,
                                          _ptr(out), _vp(stream)))

    def error_map(self, approx, truth, mask):
        mask = np.ascontiguousarray(mask, np.uint8)
        h, w = mask.shape
        err = np.zeros((h, w), np.float32)
        mx = C.c_float()
        self._check(lib().raw_ao_error_map(self._h, np.ascontiguousarray(approx, np.float32),
                                           np.ascontiguousarray(truth, np.float32), mask, w, h, err, C.byref(mx)))
        return err, mx.value

    def error_map_dev(self, approx, truth, mask, w, h, err, maxbuf, stream=None):
        self._check(lib().raw_ao_error_map_dev(self._h, _ptr(approx), _ptr(truth), _ptr(mask), w, h, _ptr(err),
                                               _ptr(maxbuf), _vp(stream)))

    def reconcile(self, approx, truth, mask, tolerance=0.12):
        """reconcile (src/cert/reconcile.cpp:4-21): the error map and maximum on the GPU, the
        float64 row-major sum of squares on the host, as the reference orders it."""
        err, mx = self.error_map(approx, truth, mask)
        rmse, pixels = host.rmse_of(err, mask)
        return host.ReconcileResult(err, rmse, mx, pixels, rmse <= np.float32(tolerance))

    def trace_rays(self, rays8, any_hit, far_scale=None):
        r = np.ascontiguousarray(rays8, np.float32).reshape(-1, 8)
        out = np.zeros((r.shape[0], 4), np.float32)
        fs = host.ROBUST_FAR_SCALE if far_scale is None else far_scale
        self._check(lib().raw_ao_trace_rays(self._h, r, r.shape[0], int(bool(any_hit)), fs, out))
        return out

    HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("tri", "<i4")])

    def closest_hits(self, rays8):
        """rt::closest for every ray (rows of {o, d, tMax, pad}): a structured array of (t, u, v, tri)."""
        r = np.ascontiguousarray(rays8, np.float32).reshape(-1, 8)
        out = np.zeros(r.shape[0], self.HIT_DTYPE)
        self._check(lib().raw_ao_closest_hits(self._h, r, r.shape[0], out.ctypes.data))
        return out

    def closest_hits_dev(self, rays8, nrays, hits, stream=None):
        self._check(lib().raw_ao_closest_hits_dev(self._h, _ptr(rays8), nrays, _ptr(hits), _vp(stream)))

    # ---- measurement ---------------------------------------------------------------
    def count_next_trace(self, enable=True):
        self._check(lib().raw_ao_count_next_trace(self._h, int(bool(enable))))

    def trace_counts(self):
        v, t, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(lib().raw_ao_trace_counts(self._h, C.byref(v), C.byref(t), C.byref(r)))
        return v.value, t.value, r.value

    def last_ms(self):
        b, t, s, e = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        self._check(lib().raw_ao_last_ms(self._h, C.byref(b), C.byref(t), C.byref(s), C.byref(e)))
        return dict(build=b.value, trace=t.value, ssao=s.value, error=e.value)

    def angle_table(self):
        s = np.zeros(1 << 24, np.float32)
        c = np.zeros(1 << 24, np.float32)
        self._check(lib().raw_ao_angle_table(self._h, s, c))
        return s, c

    def bvh_bytes(self):
        nb, tb, rec = C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._check(lib().raw_ao_bvh_bytes(self._h, C.byref(nb), C.byref(tb), C.byref(rec)))
        return nb.value, tb.value, rec.value

    def read_wide_records(self, quantized=True):
        """The wide tree's 128 B records as (n, 32) uint32 words and, with quantized, their 64 B
        quantized forms as (n, 16) uint32 words (raw_ao_read_wide_records)."""
        n = C.c_uint32()
        self._check(lib().raw_ao_read_wide_records(self._h, None, None, 0, C.byref(n)))
        r128 = np.zeros((max(1, n.value), 32), np.uint32)
        r64 = np.zeros((max(1, n.value), 16), np.uint32) if quantized else None
        self._check(lib().raw_ao_read_wide_records(self._h, r128.ctypes.data, r64.ctypes.data if quantized else None, n.value,
                                                    C.byref(n)))
        return (r128[:n.value], r64[:n.value] if quantized else None)

    # ---- learned AO surrogate ------------------------------------------------------
    def net_param_count(self):
        n = C.c_uint64()
        self._check(lib().raw_ao_net_param_count(self._h, C.byref(n)))
        return n.value

    def net_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).ravel()
        self._check(lib().raw_ao_net_set_params(self._h, p, p.size))

    def net_get_params(self):
        p = np.zeros(self.net_param_count(), np.float32)
        self._check(lib().raw_ao_net_get_params(self._h, p, p.size))
        return p

    def net_get_grads(self):
        g = np.zeros(self.net_param_count(), np.float32)
        self._check(lib().raw_ao_net_get_grads(self._h, g, g.size))
        return g

    def net_features_dev(self, position, normal, mask, ssao, noisy, eye, w, h, out, stream=None):
        """out: (h, w, 8) bf16 bits as a torch int16 tensor."""
        self._check(lib().raw_ao_net_features_dev(self._h, _ptr(position), _ptr(normal), _ptr(mask), _ptr(ssao),
                                                  _optr(noisy), np.ascontiguousarray(eye, np.float32), w, h, _ptr(out),
                                                  _vp(stream)))

    def net_forward_dev(self, x, mask, n, w, h, ao, stream=None):
        self._check(lib().raw_ao_net_forward_dev(self._h, _ptr(x), _ptr(mask), n, w, h, _ptr(ao), _vp(stream)))

    def net_train_step_dev(self, x, mask, target, n, w, h, lr, stream=None):
        self._check(lib().raw_ao_net_train_step_dev(self._h, _ptr(x), _ptr(mask), _ptr(target), n, w, h, float(lr),
                                                    _vp(stream)))

    def net_loss(self):
        v = C.c_float()
        self._check(lib().raw_ao_net_loss(self._h, C.byref(v)))
        return v.value

    def net_read_act(self, layer, count):
        out = np.zeros(count, np.uint16)
        self._check(lib().raw_ao_net_read_act(self._h, layer, out, count))
        return out

    def conv3x3_dev(self, in0, in1, c0, c1, w, kpad, bias, cout, relu, n, width, height, out, stream=None):
        self._check(lib().raw_ao_conv3x3_dev(self._h, _ptr(in0), _optr(in1), c0, c1, _ptr(w), kpad, _optr(bias), cout,
                                             int(relu), n, width, height, _ptr(out), _vp(stream)))

    def conv3x3_wgrad_dev(self, dy, in0, in1, c0, c1, cout, n, width, height, gw, gb, stream=None):
        self._check(lib().raw_ao_conv3x3_wgrad_dev(self._h, _ptr(dy), _ptr(in0), _optr(in1), c0, c1, cout, n, width,
                                                   height, _ptr(gw), _ptr(gb), _vp(stream)))

    def net_load(self, path):
        self._check(lib().raw_ao_net_load(self._h, os.fsencode(path)))

    def learned_ao_dev(self, position, normal, mask, noisy, eye, w, h, ao, ssao_samples=24, ssao_radius=1.0, rt_spp=64,
                       rt_radius=1.0, probe_stride=0, fallback_rmse=0.05, stream=None):
        """The learned frame into ao (torch); returns the report (meaningful with a probe)."""
        rep = LearnedReport()
        self._check(lib().raw_ao_learned_ao_dev(self._h, _ptr(position), _ptr(normal), _ptr(mask), _optr(noisy),
                                                np.ascontiguousarray(eye, np.float32), w, h, ssao_samples, ssao_radius,
                                                rt_spp, rt_radius, probe_stride, fallback_rmse, _ptr(ao), C.byref(rep),
                                                _vp(stream)))
        return rep.as_dict()

    def learned_ao(self, position, normal, mask, eye, noisy=None, ssao_samples=24, ssao_radius=1.0, rt_spp=64,
                   rt_radius=1.0, probe_stride=0, fallback_rmse=0.05):
        mask = np.ascontiguousarray(mask, np.uint8)
        h, w = mask.shape
        ao = np.zeros((h, w), np.float32)
        rep = LearnedReport()
        nz = None if noisy is None else np.ascontiguousarray(noisy, np.float32)
        self._check(lib().raw_ao_learned_ao(self._h, np.ascontiguousarray(position, np.float32),
                                            np.ascontiguousarray(normal, np.float32), mask,
                                            None if nz is None else nz.ctypes.data_as(C.c_void_p),
                                            np.ascontiguousarray(eye, np.float32), w, h, ssao_samples, ssao_radius,
                                            rt_spp, rt_radius, probe_stride, fallback_rmse, ao, C.byref(rep)))
        return ao, rep.as_dict()


class Batch:
    """raw_ao_batch: many views of several scenes, one persistent traversal launch. The
    torch tensors handed in must stay alive while the batch uses them (it keeps them)."""

    def __init__(self, device=0):
        self._h = _vp()
        self._keep = []
        self.nviews = 0
        L = lib()
        rc = L.raw_ao_batch_open(device, C.byref(self._h))
        if rc != OK:
            msg = L.raw_ao_batch_last_error(self._h).decode() if self._h else "open failed"
            self.close()
            raise RawAoError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().raw_ao_batch_close(self._h)
            self._h = None
        self._keep = []

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise RawAoError(rc, lib().raw_ao_batch_last_error(self._h).decode())

    def add_scene(self, tris9_tensor, radius=16, ntri=None):
        sid = C.c_int()
        self._keep.append(tris9_tensor)
        n = tris9_tensor.numel() // 9 if ntri is None else ntri
        self._check(lib().raw_ao_batch_add_scene(self._h, _ptr(tris9_tensor), n, radius, C.byref(sid)))
        return sid.value

    def add_view(self, scene, position, normal, mask, w, h, spp, radius, ao, ssao=None, err=None, ssao_samples=24,
                 ssao_radius=None):
        vid = C.c_int()
        self._keep += [position, normal, mask, ao, ssao, err]
        self._check(lib().raw_ao_batch_add_view(self._h, scene, _ptr(position), _ptr(normal), _ptr(mask), w, h, spp, radius,
                                                ssao_samples, radius if ssao_radius is None else ssao_radius, _ptr(ao),
                                                _optr(ssao), _optr(err), C.byref(vid)))
        self.nviews = vid.value + 1
        return vid.value

    def set_tuning(self, chunk, tri_q, blocks_per_cu):
        self._check(lib().raw_ao_batch_set_tuning(self._h, chunk, tri_q, blocks_per_cu))

    def set_option(self, opt, value):
        self._check(lib().raw_ao_batch_set_option(self._h, opt, int(value)))

    def tuning(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().raw_ao_batch_get_tuning(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def run(self, rebuild=False, stream=None):
        self._check(lib().raw_ao_batch_run(self._h, int(bool(rebuild)), _vp(stream)))

    def capture(self, with_builds=True):
        self._check(lib().raw_ao_batch_capture(self._h, int(bool(with_builds))))

    def launch(self, stream=None):
        self._check(lib().raw_ao_batch_launch(self._h, _vp(stream)))

    def check(self):
        self._check(lib().raw_ao_batch_check(self._h))

    def results(self):
        n = max(self.nviews, 1)
        rmse, mx = np.zeros(n, np.float32), np.zeros(n, np.float32)
        px = np.zeros(n, np.uint32)
        rays = C.c_uint64()
        self._check(lib().raw_ao_batch_results(self._h, rmse, mx, px, C.byref(rays)))
        return rmse[:self.nviews], mx[:self.nviews], px[:self.nviews], rays.value

    def tune(self, cache_path=None):
        ms = C.c_float()
        self._check(lib().raw_ao_batch_tune(self._h, None if cache_path is None else os.fsencode(cache_path), C.byref(ms)))
        return ms.value
