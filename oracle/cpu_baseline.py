"""CPU baseline timing -- TEST INFRASTRUCTURE ONLY (run by bench.py as a child process).

Times the float/OpenMP instantiation of the oracle (the C++/OpenMP transcription of the four
stages) on a bounded sample of the same seeded 752x480x64, 5-source, 8-path workload the GPU
bench runs, and prints one JSON line.

Settings follow the measurement plan for the CPU figure: the library is the -march=native build
of ``make native`` (compiled on the machine that times it), OMP_NUM_THREADS = the physical
cores available to this process (capped by ``--max-threads``) and OMP_PROC_BIND=close.  Both
OpenMP variables must be set before the OpenMP runtime starts, which is why this runs in a
fresh process: ``python -m oracle.cpu_baseline``.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)


def physical_cores() -> tuple[int, int]:
    """(physical cores, logical CPUs) among the CPUs this process may run on, from
    ``lscpu -p=CPU,CORE,SOCKET``; falls back to the logical count."""
    cpus = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             check=True).stdout
        cores = set()
        for line in out.splitlines():
            if not line or line.startswith("#"):
                continue
            cpu, core, sock = (int(v) if v else -1 for v in line.split(",")[:3])
            if cpu in cpus:
                cores.add((sock, core))
        if cores:
            return len(cores), len(cpus)
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return len(cpus), len(cpus)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build_native() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
    return os.path.join(_HERE, "_native", "libpm_oracle.so")


def child(seconds: float, seed: int, median: bool) -> dict:
    """Runs inside the timing process (environment already set)."""
    sys.path.insert(0, _ROOT)
    import oracle
    from oracle.sequence import OracleSequence
    from probabilistic_mapping_amd import synth
    s = synth.trajectory_scene(seed)
    seq = OracleSequence(s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"],
                         s["idep_min"], s["idep_max"], s["knobs"], s["paths"], median=median,
                         precision="f32")
    threads = oracle.max_threads()
    done, times = 0, []
    t0 = time.perf_counter()
    while done < len(s["images"]):
        t1 = time.perf_counter()
        seq.run_keyframe(done)
        times.append(time.perf_counter() - t1)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"keyframes": done, "seconds": dt, "ms_per_keyframe": [t * 1e3 for t in times],
            "threads": threads}


def run(seconds: float, seed: int, median: bool, max_threads: int) -> dict:
    """Builds the native library, then times it in a child process; returns the
    ``cpu_baseline`` object of the bench line."""
    lib = build_native()
    phys, logical = physical_cores()
    threads = max(1, min(phys, max_threads))
    env = dict(os.environ, PM_ORACLE_LIB=lib, OMP_NUM_THREADS=str(threads),
               OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--child", "--seconds", str(seconds),
           "--seed", str(seed), "--median", str(int(median))]
    out = subprocess.run(cmd, cwd=_ROOT, env=env, capture_output=True, text=True, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    # the first keyframe only resets the fusion state; every keyframe runs cost, SGM, filter
    # (and the median); the median of the per-keyframe times is reported beside the mean
    ms = sorted(r["ms_per_keyframe"])
    return {"value": r["keyframes"] / r["seconds"], "unit": "fused depth frames/s",
            "cores": r["threads"], "kind": "port",
            "sample": f"first {r['keyframes']} of 20 keyframes of the same 752x480x64, "
                      f"5-source, 8-path workload (float instantiation of the oracle, OpenMP)",
            "ms_per_keyframe_median": ms[len(ms) // 2],
            "build": "g++ -O3 -march=native -fopenmp",
            "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close",
                    "OMP_PLACES": "cores"},
            "host": {"cpu": cpu_model(), "physical_cores": phys, "logical_cpus": logical}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--seconds", type=float, default=12.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--median", type=int, default=1)
    ap.add_argument("--max-threads", type=int, default=16)
    a = ap.parse_args()
    if a.child:
        print(json.dumps(child(a.seconds, a.seed, bool(a.median))), flush=True)
    else:
        print(json.dumps(run(a.seconds, a.seed, bool(a.median), a.max_threads)), flush=True)


if __name__ == "__main__":
    main()
