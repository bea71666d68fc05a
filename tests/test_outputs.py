"""The files a render writes (raw_ao.host.emit_outputs over libraw_ao_host: src/tools/run.cpp,
channels_json.cpp, receipt.cpp, src/cert/certificate*.cpp, src/core/image.cpp) and verify
(src/tools/verify.cpp), on the CPU: from the oracle's frame, every file of the reference's
identity renders hashes as evidence/identity-golden.json records (certificate.json,
channels.json and receipt.json included), verify passes on them, and it catches a changed image,
a changed receipt and a missing file. SHA-256 against the FIPS 180-2 known answers."""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

import oracle_lib as O
from cases import FRAMES, params, sha

H = pytest.importorskip("raw_ao.host")


def oracle_case(name, lighting=2):
    p = params(name)
    sc = O.frame_test_scene(p["w"], p["h"], p["eye"], p["target"], p["fovy"])
    prev = None
    if p["prev_eye"]:
        ps = dict(sc)
        ps["eye"] = np.array(p["prev_eye"], np.float32)
        prev = O.view_proj(ps)
    f = O.frame(sc, p["w"], p["h"], prev=prev, lighting=lighting, rtao_on=p["rt"])
    cli = H.cli_params(p["w"], p["h"], p["eye"], p["target"], fovy=p["fovy"], rt=p["rt"], prev_eye=p["prev_eye"])
    return f, cli


def written(d, names):
    return {k: sha(open(os.path.join(d, k), "rb").read()) for k in names}


@pytest.mark.parametrize("name", ["default", "motion-256", "nort-256", "close-512", "threads4-256"])
def test_files_match_the_golden_renders(tmp_path, name):
    f, cli = oracle_case(name)
    H.emit_outputs(tmp_path, f, cli)
    g = FRAMES[name]["files"]
    assert written(tmp_path, g) == g
    expect = {"frame.ppm", "frame_hdr.pfm", "ao_ss.pgm", "ao_ss.pfm", "mask.pgm", "certificate.json", "channels.json",
              "receipt.json"} | ({"ao_rt.pgm", "ao_rt.pfm", "ao_error.pgm"} if params(name)["rt"] else set())
    assert set(os.listdir(tmp_path)) == expect
    rc, report = H.verify(tmp_path)
    assert rc == 0, report
    assert report.endswith("verify: all checks match\n")


def test_verify_catches_changes(tmp_path):
    f, cli = oracle_case("default")
    H.emit_outputs(tmp_path, f, cli)
    for name, offset in (("ao_rt.pfm", 200), ("receipt.json", 40)):
        d = tmp_path / ("bad_" + name)
        shutil.copytree(tmp_path, d, ignore=shutil.ignore_patterns("bad_*"))
        b = bytearray((d / name).read_bytes())
        b[offset] ^= 1
        (d / name).write_bytes(bytes(b))
        rc, report = H.verify(d)
        assert rc == 3 and "MISMATCH" in report, report
    d = tmp_path / "missing"
    shutil.copytree(tmp_path, d, ignore=shutil.ignore_patterns("bad_*", "missing"))
    os.remove(d / "mask.pgm")
    assert H.verify(d)[0] == 2


def test_gpu_backend_is_recorded(tmp_path):
    """A GPU render records its backend in params and the receipt (gpu_run.cpp:36-41)."""
    f, _ = oracle_case("default")
    H.emit_outputs(tmp_path, f, H.cli_params(gpu_backend="hip"), backend="raw-native-hip")
    cert = (tmp_path / "certificate.json").read_text()
    assert '"params":{"backend":"hip","eye":[4,4,6]' in cert
    assert '"backend":"raw-native-hip-ssao"' in (tmp_path / "receipt.json").read_text()
    assert H.verify(tmp_path)[0] == 0


@pytest.mark.parametrize("msg,want", [
    (b"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
    (b"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
    (b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
     "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1")])
def test_sha256_known_answers(msg, want):
    out = C.create_string_buffer(65)
    H.lib().rah_sha256(msg, len(msg), out)
    assert out.value.decode() == want


def test_sha256_long_input_matches_hashlib():
    import hashlib
    b = np.random.default_rng(1).integers(0, 256, 1_000_003, dtype=np.uint8).tobytes()
    out = C.create_string_buffer(65)
    H.lib().rah_sha256(b, len(b), out)
    assert out.value.decode() == hashlib.sha256(b).hexdigest()
