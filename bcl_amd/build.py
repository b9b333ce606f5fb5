"""Build recipe for the native libraries (run on the CPU host; hipcc cross-compiles gfx950).

  bcl_amd/lib/libbclx.so      HIP kernels + C ABI (include/bclx.h), gfx950 only
  bcl_amd/lib/libbclx_gen.so  host generators and tile cutting (g++ -fopenmp)
  oracle/liboracle.so         CPU oracle (test infrastructure), and oracle/_ref when the reference tree exists

python -m bcl_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bcl_amd", "csrc")
LIBDIR = os.path.join(ROOT, "bcl_amd", "lib")
OBJDIR = os.path.join(ROOT, "bcl_amd", "lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops off: packed fp32 math is avoided in every kernel of this library.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result",
]
HIP_SOURCES = ["core.hip", "spgemm.hip", "stage.hip", "mlp.hip", "spmm.hip", "gemm.hip", "accumulate.hip", "tiles.hip",
               "noalloc.hip"]
CPP_SOURCES = ["mmio.cpp"]  # host-only parts of the C ABI (g++)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_bclx(jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CPP_SOURCES:
        obj = os.path.join(OBJDIR, src + ".o")
        objs.append(obj)
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-fvisibility=hidden", "-Wall",
                     "-I" + os.path.join(ROOT, "include"),
                     "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIBDIR, "libbclx.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_gen() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libbclx_gen.so")
    _run(["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", os.path.join(CSRC, "gen.cpp"), "-o",
          out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_driver() -> str:
    """The C++ host-layer test driver (tests/cpp), linked against libbclx.so and the HIP runtime."""
    src = os.path.join(ROOT, "tests", "cpp", "host_layer_driver.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "host_layer_driver")
    _run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
          "-I/opt/rocm/include", src, "-o", out + ".tmp", "-L" + LIBDIR, "-lbclx", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j4"])


def build_all(jobs: int = 8) -> None:
    build_gen()
    build_oracle()
    build_bclx(jobs)
    build_cpp_driver()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--only", choices=["bclx", "gen", "oracle"], default=None)
    a = ap.parse_args()
    if a.only == "bclx":
        print(build_bclx(a.jobs))
    elif a.only == "gen":
        print(build_gen())
    elif a.only == "oracle":
        build_oracle()
    else:
        build_all(a.jobs)
        print("built", LIBDIR)
