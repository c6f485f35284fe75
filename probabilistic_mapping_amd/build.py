"""In-tree build of the native libraries (no JIT cache, so the .so files travel with the tree).

* ``libpm_hip.so``   -- HIP kernels for gfx950 + the C-ABI + the C++ StereoMapper host shell
* ``libpm_synth.so`` -- the seeded scene renderer (host only)

Run ``python -m probabilistic_mapping_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
HIP_LIB = os.path.join(HERE, "libpm_hip.so")
SYNTH_LIB = os.path.join(HERE, "libpm_synth.so")

ARCH = "gfx950"
HIP_SOURCES = ["cost.hip", "sgm.hip", "sweep.hip", "filter.hip", "fusion.hip", "fusion_cpu.hip", "batch.hip", "diff.hip", "capi.hip", "se3.cpp",
               "stereo_mapper.cpp", "parameters.cpp", "frame_driver.cpp", "frame_ring.cpp"]
HIP_HEADERS = ["common.hpp", "cost_tile.hpp", "sgm_scan.hpp"]

# Packed-FP32 VALU forms are disabled for every device compile.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_synth() -> str:
    _run(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c++17", "-march=x86-64-v2",
          "-o", SYNTH_LIB + ".tmp", os.path.join(CSRC, "synth.cpp")])
    os.replace(SYNTH_LIB + ".tmp", SYNTH_LIB)
    return SYNTH_LIB


def build_hip(jobs: int = 8) -> str:
    """Compile each translation unit to an object in parallel, then link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".", "_") + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [_hipcc()] + HIPCC_FLAGS + ["-I", INCLUDE, "-I", CSRC] + lang + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        while len([p for p in procs if p.poll() is None]) >= jobs:
            procs[0].wait()
    rc = [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError(f"hipcc failed: {rc}")
    _run([_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", HIP_LIB + ".tmp"] + objs)
    os.replace(HIP_LIB + ".tmp", HIP_LIB)
    return HIP_LIB


EVALUATE = os.path.join(HERE, "pm_evaluate")


def build_evaluate() -> str:
    """The file-based driver executable, linked against the in-tree library."""
    _run([_hipcc(), "-O2", "-std=c++17", "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "pm_evaluate.cpp"), "-L", HERE, "-lpm_hip",
          "-Wl,-rpath,$ORIGIN", "-o", EVALUATE + ".tmp"])
    os.replace(EVALUATE + ".tmp", EVALUATE)
    return EVALUATE


BENCH_MAPPER = os.path.join(HERE, "pm_bench_mapper")


def build_bench_mapper() -> str:
    """The C++ timing driver of the StereoMapper sequence (bench.py --form mapper)."""
    _run([_hipcc(), "-O2", "-std=c++17", "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "pm_bench_mapper.cpp"), "-L", HERE, "-lpm_hip",
          "-Wl,-rpath,$ORIGIN", "-o", BENCH_MAPPER + ".tmp"])
    os.replace(BENCH_MAPPER + ".tmp", BENCH_MAPPER)
    return BENCH_MAPPER


def build_all() -> None:
    build_synth()
    build_hip()
    build_evaluate()
    build_bench_mapper()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "hip": build_hip, "synth": build_synth,
     "evaluate": build_evaluate, "bench_mapper": build_bench_mapper}[what]()
