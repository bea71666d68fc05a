"""raw-gpu-cert/1 (reconcileGpuCpu + gpuCertificateJson, src/renderer/gpu_reconcile.cpp) from
libraw_ao_host against the reference's own code (oracle/_ref), byte for byte, on the CPU: two
equal frames (identity MATCH, verified), a changed frame (coverage, AO and frame channels fail,
DRIFT, refuted), no frames with a reason (unverifiable), and a frame without the ray-traced pass.
The GPU receipt beside it is sealed and lists the certificate's digest."""
import ctypes as C
import json

import numpy as np
import pytest

import oracle_lib as O

H = pytest.importorskip("raw_ao.host")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")

PARAMS = ('{"backend":"hip","eye":[4,4,6],"fovy":0.899999976,"height":96,"prev_eye":null,"prev_target":null,'
          '"prev_up":null,"rt":%s,"target":[0,1,0],"tolerance":0.119999997,"up":[0,1,0],"width":128}')


class RefPlanes(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("position", C.c_void_p), ("normal", C.c_void_p), ("mask", C.c_void_p),
                ("motion", C.c_void_p), ("ao_rt", C.c_void_p), ("ao_ss", C.c_void_p), ("frame", C.c_void_p),
                ("rmse", C.c_float), ("within", C.c_int32)]


def ref_cert(gpu, cpu, w, h, rt, reason=None, gpu_ms=1.25, cpu_ms=37.5):
    R = O.ref()
    R.ref_gpu_certificate.restype = C.c_size_t
    R.ref_gpu_certificate.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p,
                                      C.c_double, C.c_double, C.c_char_p, C.c_size_t]
    keep = []

    def planes(f):
        if f is None:
            return None

        def p(k, t):
            if f.get(k) is None:
                return None
            a = np.ascontiguousarray(f[k], t)
            keep.append(a)
            return a.ctypes.data
        s = RefPlanes(*[p(k, np.uint8 if k == "mask" else np.float32) for k in
                        ("depth", "position", "normal", "mask", "motion", "ao_rt", "ao_ss", "frame")])
        s.rmse, s.within = f["rmse"], int(bool(f["within"]))
        keep.append(s)
        return C.byref(s)
    args = [planes(gpu), planes(cpu), w, h, int(rt), reason.encode() if reason else None,
            (PARAMS % ("true" if rt else "false")).encode(), gpu_ms, cpu_ms]
    n = R.ref_gpu_certificate(*args, None, 0)
    buf = C.create_string_buffer(n + 1)
    R.ref_gpu_certificate(*args, buf, n + 1)
    return buf.value.decode()


def mine(tmp_path, gpu, cpu, w, h, rt, reason=None):
    H.gpu_certificate(tmp_path, H.cli_params(w, h, rt=rt, gpu_backend="hip"), gpu=gpu, cpu=cpu, reason=reason,
                      adapter={}, gpu_ms=1.25, cpu_ms=37.5)
    return (tmp_path / "gpu_certificate.json").read_text()


def frames(rt=True):
    w, h = 128, 96
    sc = O.frame_test_scene(w, h)
    ps = dict(sc)
    ps["eye"] = np.array([4.3, 4.0, 5.7], np.float32)
    f = O.frame(sc, w, h, prev=O.view_proj(ps), rtao_on=rt)
    return f, w, h


@needs_ref
def test_equal_frames(tmp_path):
    f, w, h = frames()
    got = mine(tmp_path, f, f, w, h, True)
    assert got == ref_cert(f, f, w, h, True)
    c = json.loads(got)
    assert c["verdict"] == "verified" and c["identity"] == "MATCH" and all(ch["rmse"] == 0 for ch in c["channels"])
    rec = json.loads((tmp_path / "gpu_receipt.json").read_text())
    assert rec["schema"] == "superstack.receipt/1" and rec["backend"] == "raw-native-hip"
    assert rec["reconcile"]["tolerance"]["verdict"] == "verified" and rec["reconcile"]["identity"] == "MATCH"
    assert rec["outputs"]["gpu_certificate.json"] == O.sha256_file(tmp_path / "gpu_certificate.json")


@needs_ref
def test_changed_frame(tmp_path):
    f, w, h = frames()
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in f.items()}
    rng = np.random.default_rng(7)
    m = g["mask"] != 0
    g["ao_rt"] = np.where(m, np.clip(g["ao_rt"] + rng.normal(0, 0.05, g["ao_rt"].shape), 0, 1), 1).astype(np.float32)
    g["frame"] = (g["frame"] * np.float32(0.9)).astype(np.float32)
    g["depth"] = (g["depth"] * np.float32(1.0 + 3e-4)).astype(np.float32)
    ys, xs = np.nonzero(m)
    g["mask"][ys[:40], xs[:40]] = 0
    g["rmse"] = float(np.float32(f["rmse"] + 0.02))
    got = mine(tmp_path, g, f, w, h, True)
    assert got == ref_cert(g, f, w, h, True)
    c = json.loads(got)
    assert c["verdict"] == "refuted" and c["identity"] == "DRIFT"
    assert {ch["name"]: ch["pass"] for ch in c["channels"]} == {"depth_rel": False, "position": True, "normal": True,
                                                                   "motion": True, "ao_ss": True, "ao_rt": False,
                                                                   "frame": False}
    assert not c["ao_verdict"]["pass"]


@needs_ref
def test_no_frames(tmp_path):
    got = mine(tmp_path, None, None, 128, 96, True, reason="no GPU device")
    assert got == ref_cert(None, None, 128, 96, True, reason="no GPU device")
    c = json.loads(got)
    assert c["verdict"] == "unverifiable" and c["identity"] is None and c["reason"] == "no GPU device"
    rec = json.loads((tmp_path / "gpu_receipt.json").read_text())
    assert rec["reconcile"]["tolerance"]["reason"] == "no GPU device"
    assert rec["reconcile"]["reference"]["content_sha256"] == "none"


@needs_ref
def test_without_the_ray_traced_pass(tmp_path):
    f, w, h = frames(rt=False)
    got = mine(tmp_path, f, f, w, h, False)
    assert got == ref_cert(f, f, w, h, False)
    assert [ch["name"] for ch in json.loads(got)["channels"]] == ["depth_rel", "position", "normal", "motion", "ao_ss",
                                                                    "frame"]
