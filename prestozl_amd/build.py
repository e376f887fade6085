"""Build step for libaccelsearch_lib.so (gfx950) and the test oracle.

    python -m prestozl_amd.build            # library + oracle (+ oracle/_ref)

hipcc compiles the HIP kernels and the host C++ into one shared library in
prestozl_amd/lib/.  Packed FP32 instructions are disabled for the device code
(-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PZL_ARCH", "gfx950")

SOURCES = ["pzl_kernels.hip", "pzl_engine.cpp", "pzl_geometry.cpp", "pzl_stats.cpp", "accel_api.cpp"]
GLIB_INC = ["/opt/conda/include/glib-2.0", "/opt/conda/lib/glib-2.0/include"]
GLIB_LIB = "/opt/conda/lib"


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: %s" % cmd[0])
    return r


def build_library(out=None):
    os.makedirs(LIBDIR, exist_ok=True)
    out = out or os.path.join(LIBDIR, "libaccelsearch_lib.so")
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    have_glib = all(os.path.isdir(d) for d in GLIB_INC)
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-Wno-unused-value",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    if have_glib:
        cmd += ["-DPZL_HAVE_GLIB=1"] + ["-I" + d for d in GLIB_INC]
    cmd += srcs + ["-o", out + ".tmp"]
    if have_glib:
        cmd += ["-L" + GLIB_LIB, "-Wl,-rpath," + GLIB_LIB, "-lglib-2.0"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all():
    build_library()
    build_oracle()


if __name__ == "__main__":
    build_all()
