"""The SSAO taps' float32 sin/cos (sincos_fast, raw-native_amd/csrc/ao_math.hpp), checked on
the host over every one of the 2^24 sampler angles (6.2831853f * k / 2^24, the angle of
src/renderer/ssao.cpp:20-24 and ray_ao.cpp:32-38) against the exact pair the kernels otherwise
compute (sincos_base, and the host libm it is corrected to), and the tap margin that decides
when the exact pair is taken. The kernel's results stay bit-exact: the GPU SSAO parity tests
compare them with the oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "raw-native_amd", "csrc")


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fast_trig") / "fast_trig_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, os.path.join(HERE, "cpp", "fast_trig_check.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, check=True).stdout.split()
    e_base, e_libm, err, margin = (float(v) for v in out[:4])
    checked, bad, exact = (int(v) for v in out[4:])
    return dict(e_base=e_base, e_libm=e_libm, err=err, margin=margin, checked=checked, bad=bad, exact=exact)


def test_fast_pair_within_its_bound(check):
    print(check)
    assert check["e_base"] <= check["err"]
    assert check["e_libm"] <= check["err"]


def test_tap_margin_covers_the_bound(check):
    # an offset cos/sin x rad (rad <= 6) moves by at most 6 x err plus half an ulp of each
    # of the two products (2^-22 below 8)
    assert 6.0 * check["err"] + 2.0 * 2.0 ** -22 < check["margin"] / 3.0


def test_tap_decisions_match_exact(check):
    assert check["checked"] == 8000000
    assert check["bad"] == 0
    assert check["exact"] < check["checked"] // 1000   # the exact pair is rare
