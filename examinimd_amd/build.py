"""Build recipe for the native parts of examinimd_amd.

  libemd_hip.so     HIP kernels + C ABI (include/emd_hip.h), gfx950
  libexaminimd.so   C++ host modules (Binning/Neighbor/Force/Integrator/Comm
                    subclasses, Input, driver) + the app C ABI (include/emd_app.h)
  ExaMiniMD         the command-line driver (same CLI as the reference)

Everything is compiled in-tree so the built files travel with the source tree.
Run `python -m examinimd_amd.build` (or __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# roctx ranges around the driver's module calls (host/trace.h)
ROCTX = ["-L", ROCM + "/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + ROCM + "/lib"]

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result", "-Wno-unused-value",
]

HIP_SOURCES = ["emd_core.hip", "emd_comm.hip", "emd_bin_neigh.hip", "emd_lj.hip",
               "emd_snap.hip", "emd_batch.hip"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_hip(jobs=8):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = HIP_SOURCES
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(o)
        cmd = [HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, s),
                                     "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("hipcc failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    out = os.path.join(LIBDIR, "libemd_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_host():
    srcs = sorted(f for f in os.listdir(HOST) if f.endswith(".cpp") and f != "main.cpp")
    if not srcs:
        return None
    cxx = os.environ.get("CXX", "g++")
    common = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable",
              "-I", INCLUDE, "-I", HOST, "-I", ROCM + "/include"]
    lib = os.path.join(LIBDIR, "libexaminimd.so")
    _run([cxx] + common + ["-shared", "-o", lib + ".tmp"] + [os.path.join(HOST, s) for s in srcs]
         + ["-L", LIBDIR, "-lemd_hip", "-Wl,-rpath,$ORIGIN"] + ROCTX)
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(LIBDIR, "ExaMiniMD")
    _run([cxx] + common + ["-o", exe + ".tmp", os.path.join(HOST, "main.cpp"),
                           "-L", LIBDIR, "-lexaminimd", "-lemd_hip", "-Wl,-rpath,$ORIGIN"] + ROCTX)
    os.replace(exe + ".tmp", exe)
    return lib


def build(jobs=8):
    build_hip(jobs)
    build_host()


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
