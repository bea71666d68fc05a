"""Builds the native parts in-tree:

  build/libraw_ao_hip.so   the gfx950 kernels and the C ABI (hipcc)
  build/libraw_ao_host.so  the C++ host side (g++), linked to the former

Float32 exactness flags: -ffp-contract=off (no fused multiply-add), correctly rounded
division and square root, denormals kept. Packed-FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
HIP_SOURCES = ["csrc/capi.hip", "csrc/ploc.hip", "csrc/trace.hip", "csrc/bvh4.hip", "csrc/ssao.hip", "csrc/unet.hip", "csrc/batch.hip", "csrc/raster.hip", "csrc/frame.hip",
               "csrc/noalloc.hip", "csrc/trig_table.cpp"]
ARCH = os.environ.get("RAW_AO_ARCH", "gfx950")

HIPCC_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++23", "-fPIC", "-shared", "-ffp-contract=off",
               "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def run(cmd, cwd):
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hip_out = os.path.join(BUILD, "libraw_ao_hip.so")
    tmp = hip_out + ".tmp"
    out = run(["hipcc"] + HIPCC_FLAGS + HIP_SOURCES + ["-o", tmp], HERE)
    os.replace(tmp, hip_out)
    host_out = os.path.join(BUILD, "libraw_ao_host.so")
    tmp = host_out + ".tmp"
    out += run(["g++", "-O2", "-std=c++23", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "host/raw_ao_host.cpp", "host/raw_ao_outputs.cpp",
                "-L" + BUILD, "-lraw_ao_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp], HERE)
    os.replace(tmp, host_out)
    # the C++ host interface's frame test (tests/cpp), linked to the host library
    exe = os.path.join(BUILD, "host_interface_frame")
    out += run(["g++", "-O2", "-std=c++23", "-ffp-contract=off", "-Wall", "../tests/cpp/host_interface_frame.cpp",
                "-L" + BUILD, "-lraw_ao_host", "-lraw_ao_hip", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"], HERE)
    os.replace(exe + ".tmp", exe)
    if verbose and out.strip():
        print(out)
    return hip_out, host_out


if __name__ == "__main__":
    print(build(verbose=True))
