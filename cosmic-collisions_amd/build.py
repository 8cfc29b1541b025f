"""Builds the gfx950 library libcc_hip.so (and, with node headers present, the
N-API addon) in place.  Used by __graft_entry__.build(); runnable alone:
    python cosmic-collisions_amd/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the kernels keep the reference's FP32 operation order: no contraction,
    # correctly rounded division and square root, no packed FP32 forms
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["csrc/cc_kernels.hip", "csrc/cc_api.hip", "csrc/cc_ff.hip", "csrc/cc_fof.hip", "csrc/cc_fof_host.cpp",
           "csrc/cc_bodies.cpp", "csrc/cc_crater.cpp"]


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-8000:])
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_hip():
    out = os.path.join(HERE, "libcc_hip.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + SOURCES, HERE)
    os.replace(tmp, out)
    return out


def build_checks():
    """tests/recip_check: the GPU check of the correctly rounded helpers
    (csrc/cc_fpmath.h), run by tests/test_gpu_fpmath.py."""
    src = os.path.join(REPO, "tests", "recip_check.hip")
    out = os.path.join(REPO, "tests", "recip_check")
    run([HIPCC] + [f for f in HIP_FLAGS if f not in ("-fPIC", "-fhip-fp32-correctly-rounded-divide-sqrt")] +
        ["-o", out + ".tmp", src], HERE)
    os.replace(out + ".tmp", out)
    return out


def build_addon():
    """The N-API addon (js/cc_addon.cc) linked against libcc_hip.so; skipped
    where node's headers are absent."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    out = os.path.join(HERE, "js", "cc_addon.node")
    src = os.path.join(HERE, "js", "cc_addon.cc")
    if not os.path.exists(src):
        return None
    cxx = shutil.which("g++") or "g++"
    run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + inc, "-I" + os.path.join(REPO, "include"),
         "-o", out + ".tmp", src, "-L" + HERE, "-lcc_hip", "-Wl,-rpath,$ORIGIN/.."], HERE)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_hip())
    print(build_checks())
    print(build_addon())
