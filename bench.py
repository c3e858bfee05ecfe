#!/usr/bin/env python3
"""Benchmark of the Wilson-clover operator on one MI355X.

Workload (one "step" = one application of Dw_dble): the 32^3x64 operator
throughput configuration -- open boundary conditions (cF=1), csw=1.77,
m0=-0.3, Gaussian spinor source, Haar-random SU(3) links smoothed by 100
forward-Euler steps (eps 0.01) of the Wilson flow.  The value is the algorithmic
HBM bandwidth of Dw_dble, 1536 bytes per lattice site (links 576, SW term
576, spinor in 192, out 192), over the timed steps.  Inputs are resident in
HBM before the timed region.

Also reported: the same figure for Dw, Dwhat_dble and Dwhat; a measured
device copy bandwidth; the roofline object of the Dw_dble kernel (HIP-event
timing on the library stream) and the reference CPU code timed on the host
(oracle/_ref, the reference's own AVX/FMA3 build, OpenMP).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "openqcd-2.4.2_amd"))

PEAK_HBM = 8000.0  # GB/s, MI355X published HBM3E bandwidth
L = (64, 32, 32, 32)
LTRD = (8, 8, 8, 8)
BYTES = {"Dw_dble": 1536, "Dw": 768, "Dwhat_dble": 1344, "Dwhat": 672}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_bandwidth(lib):
    """Device copy bandwidth (bytes read + written) of a plain 16-byte copy
    kernel over 4 GB, from the library (oqh_copy_bandwidth)."""
    import ctypes as C
    lib.oqh_copy_bandwidth.restype = C.c_double
    lib.oqh_copy_bandwidth.argtypes = [C.c_double, C.c_int]
    bw = lib.oqh_copy_bandwidth(4.0, 20)
    return bw if bw > 0 else None


def measured_traffic():
    """HBM bytes per Dw_dble launch from the committed counter summary
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated with the copy kernel)."""
    path = os.path.join(ROOT, "profiles", "Dw_dble_64x32x32x32_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == "Dw_dble":
                return int(f[2])
    except OSError:
        pass
    return None


def cpu_baseline(V):
    """Reference CPU Dw_dble (oracle/_ref/ref_64x32x32x32_avx: the
    reference's modules built with -DAVX -DFMA3, OpenMP, 16 threads)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_64x32x32x32_avx")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS="16")
    try:
        r = subprocess.run([exe, "bench", "-bc", "0", "-csw", "1.77", "-m0", "-0.3", "-seed", "2",
                            "-secs", "5"], capture_output=True, text=True, timeout=300, env=env,
                           cwd="/tmp")
        j = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # pragma: no cover
        log("cpu baseline failed:", e)
        return None
    sec = j["Dw_dble_sec"] / j["Dw_dble_calls"]
    return {"value": round(BYTES["Dw_dble"] * V / sec / 1e9, 2), "unit": "GB/s",
            "cores": j["nthread"], "kind": "reference",
            "sample": "Dw_dble of the reference (AVX/FMA3, OpenMP NTHREAD=%d) on its own "
                      "random_ud field, 32^3x64 open bc, %d calls in %.1f s; Dw %.1f GB/s, "
                      "Dwhat_dble %.1f GB/s" % (
                          j["nthread"], j["Dw_dble_calls"], j["Dw_dble_sec"],
                          BYTES["Dw"] * V / (j["Dw_sec"] / j["Dw_calls"]) / 1e9,
                          BYTES["Dwhat_dble"] * V / (j["Dwhat_dble_sec"] / j["Dwhat_dble_calls"]) / 1e9)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one process, one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyoqh

    lat = pyoqh.Lattice(L, LTRD, bc=0, cF=(1.0, 1.0), m0=-0.3, csw=1.77)
    V = lat.V
    rng = np.random.default_rng(2)
    log("bench: generating %d SU(3) links" % (4 * V))
    lat.import_ud(pyoqh.haar_su3(4 * V, rng))
    log("bench: Wilson flow, 100 Euler steps of 0.01 on the device")
    lat.lib.fwd_euler(100, 0.01)
    assert lat.sw_term(3) == 0
    s, r = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(s, rng.standard_normal((V, 24)))
    lib = lat.lib

    # warmup, then K timed steps bracketed by device synchronisation
    for _ in range(args.warmup):
        lib.Dw_dble(0.0, s, r)
    lat.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lib.Dw_dble(0.0, s, r)
    lat.sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = BYTES["Dw_dble"] * V / (ms_step * 1e-3) / 1e9

    # kernel time with HIP events on the library stream
    ms_ev = lib.oqh_time_dirac(0, 0.0, s, r, max(args.steps, 20))
    achieved = BYTES["Dw_dble"] * V / (ms_ev * 1e-3) / 1e9

    others = {}
    s32, r32 = lat.alloc_s(), lat.alloc_s()
    lat.put_s(s32, rng.standard_normal((V, 24)).astype(np.float32))
    ms = lib.oqh_time_dirac(1, 0.0, s32, r32, 50)
    others["Dw_GBps"] = round(BYTES["Dw"] * V / ms / 1e6, 1)
    assert lat.sw_term(2) == 0
    ms = lib.oqh_time_dirac(2, 0.0, s, r, 50)
    others["Dwhat_dble_GBps"] = round(BYTES["Dwhat_dble"] * V / ms / 1e6, 1)
    ms = lib.oqh_time_dirac(3, 0.0, s32, r32, 50)
    others["Dwhat_GBps"] = round(BYTES["Dwhat"] * V / ms / 1e6, 1)
    lat.close()

    copy_bw = copy_bandwidth(lib)
    cpu = None if args.no_cpu else cpu_baseline(V)

    out = {
        "metric": "Dw_dble HBM GB/s (% peak); time-to-1e-10 solve, mixed/learned vs plain SAP+GCR",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Haar-random SU(3) links (numpy, seed 2) smoothed by the Wilson flow "
                "(100 forward-Euler steps, eps 0.01, on the device), Gaussian FP64 source",
        "config": {"workload": "Dw_dble 32^3x64 open bc (type 0, cF=1), csw=1.77, m0=-0.3",
                   "lattice": "64x32x32x32", "bytes_per_site": BYTES["Dw_dble"],
                   "pct_of_8TBps": round(100 * value / PEAK_HBM, 1),
                   "pct_of_copy": round(100 * value / copy_bw, 1) if copy_bw else None,
                   "copy_GBps": round(copy_bw, 1) if copy_bw else None, **others},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4), "traffic": measured_traffic()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
