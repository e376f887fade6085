"""Build step for libaccelsearch_lib.so (gfx950) and the test oracle.

    python -m prestozl_amd.build            # library + driver + oracle (+ oracle/_ref)

hipcc compiles the HIP kernels and the host C++ into one shared library in
prestozl_amd/lib/.  Packed FP32 instructions are disabled for the device code
(-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PZL_ARCH", "gfx950")

SOURCES = ["pzl_kernels.hip", "pzl_sort.hip", "pzl_optimize.hip", "pzl_realfft.hip", "pzl_dedisp.hip", "pzl_engine.cpp", "pzl_optimize.cpp", "pzl_geometry.cpp",
           "pzl_stats.cpp", "pzl_server.cpp",
           "accel_api.cpp", "accel_output.cpp"]
GLIB_INC = ["/opt/conda/include/glib-2.0", "/opt/conda/lib/glib-2.0/include"]
# Headers from the conda tree; the library is the system's libglib-2.0.so.0
# (found by the loader without an rpath, so no other library of the conda
# tree, such as its older libstdc++, shadows the system's).
GLIB_LINK = ["-l:libglib-2.0.so.0"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: %s" % cmd[0])
    return r


def build_library(out=None, jobs=None):
    """Compiles every source to an object (in parallel), then links the
    shared library.  Objects go to _build/ at the repository root."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    out = out or os.path.join(LIBDIR, "libaccelsearch_lib.so")
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    have_glib = all(os.path.isdir(d) for d in GLIB_INC)
    flags = [HIPCC, "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    if have_glib:
        flags += ["-DPZL_HAVE_GLIB=1"] + ["-I" + d for d in GLIB_INC]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
    jobs = jobs or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda so: _run(flags + ["-c", so[0], "-o", so[1]]), zip(srcs, objs)))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-fPIC", "-shared"] + objs + ["-o", out + ".tmp"]
    if have_glib:
        cmd += GLIB_LINK
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_driver(lib=None):
    """The accelsearch command-line driver (accel.h API), linked to the library."""
    lib = lib or os.path.join(LIBDIR, "libaccelsearch_lib.so")
    out = os.path.join(LIBDIR, "accelsearch")
    if not all(os.path.isdir(d) for d in GLIB_INC):
        print("glib-2.0 headers not found: accelsearch driver not built", flush=True)
        return None
    cmd = ["gcc", "-O2", "-I" + os.path.join(ROOT, "include")] + ["-I" + d for d in GLIB_INC]
    cmd += [os.path.join(CSRC, "accelsearch_main.c"), "-o", out + ".tmp",
            "-L" + LIBDIR, "-l:" + os.path.basename(lib), "-Wl,-rpath,$ORIGIN",
            # the HIP runtime needs a newer libstdc++ than gcc 11 links against;
            # the system's, found at load time, has it
            "-Wl,--allow-shlib-undefined"] + GLIB_LINK
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_pipeline(lib=None):
    """The three-thread CPU1 -> GPU -> CPU2 pipeline driver over the accel.h API."""
    lib = lib or os.path.join(LIBDIR, "libaccelsearch_lib.so")
    out = os.path.join(LIBDIR, "accelsearch_pipeline")
    if not all(os.path.isdir(d) for d in GLIB_INC):
        print("glib-2.0 headers not found: accelsearch_pipeline not built", flush=True)
        return None
    cmd = ["gcc", "-O2", "-pthread", "-I" + os.path.join(ROOT, "include")] + ["-I" + d for d in GLIB_INC]
    cmd += [os.path.join(CSRC, "accelsearch_pipeline_main.c"), "-o", out + ".tmp",
            "-L" + LIBDIR, "-l:" + os.path.basename(lib), "-Wl,-rpath,$ORIGIN",
            "-Wl,--allow-shlib-undefined"] + GLIB_LINK
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_realfft(lib=None):
    """The realfft command (in-core path of realfft.c) on the GPU transform."""
    lib = lib or os.path.join(LIBDIR, "libaccelsearch_lib.so")
    out = os.path.join(LIBDIR, "realfft")
    cmd = ["gcc", "-O2", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "realfft_main.c"),
           "-o", out + ".tmp", "-L" + LIBDIR, "-l:" + os.path.basename(lib), "-Wl,-rpath,$ORIGIN",
           "-Wl,--allow-shlib-undefined"] + GLIB_LINK
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all():
    build_library()
    build_driver()
    build_pipeline()
    build_realfft()
    build_oracle()


if __name__ == "__main__":
    build_all()
