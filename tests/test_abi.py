"""The C ABI library and the host side without a GPU: every declared symbol is exported,
a missing device is reported as the reference reports it, and the host pieces (reconcile
sum, certificate, rasterizer, scenes) match the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import raw_ao
from cases import CASES, params, sha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "raw_ao_hip.h")).read()
    return sorted(set(re.findall(r"\b(raw_ao_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    L = raw_ao.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(n for n, _, _ in raw_ao.SIGNATURES) == names
    assert L.raw_ao_abi_version() == 1


def test_versioned_api_table():
    """raw_ao_hip_api(1): the version, its size and every entry the function of the same name;
    an unknown version gets NULL."""
    L = raw_ao.lib()
    names = ["open", "close", "last_error", "set_option", "build_ploc", "build_ploc_dev", "tree_info", "read_tree",
             "read_keys", "trace_ao", "trace_ao_dev", "trace_visibility", "ssao", "ssao_dev", "error_map", "error_map_dev",
             "trace_rays", "closest_hits", "render_frame"]

    class Api(C.Structure):
        _fields_ = [("abi_version", C.c_uint32), ("size", C.c_uint32)] + [(n, C.c_void_p) for n in names]

    p = L.raw_ao_hip_api(1)
    assert p
    api = Api.from_address(p)
    assert api.abi_version == 1 and api.size == C.sizeof(Api)
    for n in names:
        assert getattr(api, n) == C.cast(getattr(L, "raw_ao_" + n), C.c_void_p).value, n
    assert not L.raw_ao_hip_api(2) and not L.raw_ao_hip_api(0)


def gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(gpu_present(), reason="checks the no-device path")
def test_no_device_reports_exit_4_code():
    """No adapter -> the reference's 'no GPU' outcome (src/tools/gpu_run.cpp:52-59)."""
    with pytest.raises(raw_ao.RawAoError) as e:
        raw_ao.Context(0)
    assert e.value.code == raw_ao.ERR_NO_DEVICE


def test_null_context_is_an_argument_error():
    L = raw_ao.lib()
    assert L.raw_ao_trace_ao(None, np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.uint8), 1, 1, 64, 2.0,
                             np.zeros(1, np.float32)) == raw_ao.ERR_ARG
    assert L.raw_ao_last_error(None) == b"null context"


@pytest.mark.parametrize("name", ["default", "close-512", "wide-1440x900"])
def test_host_rasterizer_matches_oracle_and_golden_mask(name):
    p = params(name)
    _, pos, nrm, mask = raw_ao.host.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    d2, pos2, nrm2, mask2 = O.cli_gbuffer(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    assert sha(O.mask_pgm_bytes(mask)) == CASES[name]["files"]["mask.pgm"]
    assert np.array_equal(pos.view(np.uint32), pos2.view(np.uint32))
    assert np.array_equal(nrm.view(np.uint32), nrm2.view(np.uint32))


def test_host_rasterizer_on_large_scene_matches_oracle():
    vpos, vnrm, idx, _ = raw_ao.host.scene_seeded()
    eye = raw_ao.host.orbit_camera(1)
    a = raw_ao.host.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9, 16 / 9, 0.1, 200.0, 320, 180)
    b = O.rasterize(vpos, vnrm, idx, eye, (0, 2, 0), (0, 1, 0), 0.9, 16 / 9, 0.1, 200.0, 320, 180)
    for x, y in zip(a, b):
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))


def test_rmse_sum_is_float64_row_major():
    rng = np.random.default_rng(1)
    e = rng.random((300, 200), dtype=np.float32)
    m = (rng.random((300, 200)) < 0.6).astype(np.uint8)
    rmse, px = raw_ao.host.rmse_of(e, m)
    _, rmse_o, _, px_o, _ = O.reconcile(e, np.zeros_like(e), m, 0.1)
    assert (float(rmse), px) == (rmse_o, px_o)


def test_certificate_shape():
    """raw-cert/2 text: evidence strings, verdict and exact values
    (tests/test_certificate_from.cpp:9-14, src/cert/certificate.cpp:44-85)."""
    cert = raw_ao.host.certificate(np.float32(0.129436031), np.float32(0.609375), 37996, 0.12, renderer="raw-ao-mi355x",
                                   params_json='{"width":256}', outputs=[("ao_rt.pfm", "ab")])
    assert cert.startswith('{"claim":"screen-space AO matches ray-traced ground truth within tolerance","verdict":"refuted"')
    assert '["pixels","37996"]' in cert and '["rmse","0.1294"]' in cert and '["tolerance","0.1200"]' in cert
    assert '"exact":{"pixels":37996,"rmse":0.129436031,"maxError":0.609375,"tolerance":0.119999997}' in cert
    assert '"motion_coherence":null' in cert and '"schema":"raw-cert/2"' in cert
    assert cert.endswith('"outputs":{"ao_rt.pfm":"ab"}}')
    empty = raw_ao.host.certificate(0.0, 0.0, 0, 0.12)
    assert '"verdict":"unverifiable"' in empty and '"ao_fidelity":null' in empty


def test_scene_generator_is_seeded():
    a = raw_ao.host.scene_seeded()
    b = raw_ao.host.scene_seeded()
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert 900_000 < a[2].shape[0] < 1_200_000
    c = raw_ao.host.scene_seeded(raw_ao.host.SCENE_1M._replace(seed=7))
    assert not np.array_equal(a[0][:100], c[0][:100])
