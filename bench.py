#!/usr/bin/env python3
"""Benchmark of the Wilson-clover operator on one MI355X.

Workload (one "step" = one application of Dw_dble): the 32^3x64 operator
throughput configuration -- open boundary conditions (cF=1), csw=1.77,
m0=-0.3, the reference's random field (start_ranlux(0, 2), random_ud: Haar
SU(3) links from the ranlxd streams, generated on the device) smoothed by 100
forward-Euler steps (eps 0.01) of the Wilson flow, Gaussian source
(random_sd).  The value is the algorithmic
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


NN_WIDTHS = [4, 8, 4]


def train_network():
    """Trains the learned preconditioner (nntrain.train_preconditioner) on a
    16^4 lattice with the physics of the solver workload (periodic bc,
    random_ud of seed 1 flowed 50 x 0.01, csw=1.77, m0=-0.35, SAP 4^4
    isolv=1 nmr=4 ncy=5): 32 pairs from Gaussian sources r, e = D^-1 r by
    SAP+GCR to 1e-12; the network input is the residual the 5 Schwarz cycles
    leave of r, its target the error they leave of e.  28 pairs train (1000
    Adam steps, batches of 8, learning rate 1e-2 then 2e-3), 4 are held out.
    Returns the weights and the training record."""
    import nntrain
    return nntrain.train_preconditioner(NN_WIDTHS, log=lambda m: log("bench: " + m))


def solve_times(args):
    """Time-to-solution of SAP+GCR on the 48^4 solver workload: periodic bc,
    random_ud links (start_ranlux(0, 3)) smoothed by 50 Euler steps (eps 0.01) of
    the Wilson flow, csw=1.77, m0=-0.35, Gaussian source; FGCR with nkv=24 to
    relative residual 1e-10, SAP on 4^4 blocks (isolv=1, nmr=4, ncy=5).  The
    plain solver (FP32 SAP, FP64 GCR), the FP16-storage variant (links, SW
    matrices and spinors in FP16 with per-site scales, FP32 arithmetic), the
    FP64 variant (FP64 Schwarz cycles) and the FP16 variant with the learned
    correction (trained at 16^4 by train_network) are timed on the same field
    and source; the true residual |eta - Dw_dble psi|/|eta| is recomputed and
    the error |psi - psi_ref|/|psi_ref| against a solve with FP64 Schwarz
    cycles to relative residual 1e-14 is reported."""
    wts, nn_rec = train_network()
    import ctypes as C
    import numpy as np
    import pyoqh
    P = C.c_void_p
    L4 = (48, 48, 48, 48)
    lat = pyoqh.Lattice(L4, (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    out = {"workload": "SAP+GCR 48^4 periodic, Wilson-flowed (50 x 0.01) random_ud links (seed 3), csw=1.77, "
                       "m0=-0.35, nkv=24, res=1e-10, SAP 4^4 blocks isolv=1 nmr=4 ncy=5"}
    try:
        V = lat.V
        log("bench: solver workload, random_ud of seed 3 (%d links)" % (4 * V))
        lib = lat.lib
        lib.start_ranlux(0, 3)
        lib.random_ud()
        lib.fwd_euler(50, 0.01)
        lib.set_ud_phase()
        lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.norm_square_dble.restype = pyoqh.Qflt
        lib.norm_square_dble.argtypes = [C.c_int, C.c_int, P]
        lib.mulr_spinor_add_dble.argtypes = [C.c_int, C.c_int, P, P, C.c_double]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        lib.random_sd(lat.VT, 2, e, 1.0)
        # the operators on this lattice (HIP events, 50 calls each)
        assert lat.sw_term(3) == 0
        ms = lib.oqh_time_dirac(0, 0.0, e, r, 50)
        out["Dw_dble_GBps"] = round(BYTES["Dw_dble"] * V / ms / 1e6, 1)
        out["Dw_dble_pct_of_8TBps"] = round(100 * BYTES["Dw_dble"] * V / ms / 1e6 / PEAK_HBM, 1)
        assert lat.sw_term(2) == 0
        ms = lib.oqh_time_dirac(2, 0.0, e, r, 50)
        out["Dwhat_dble_GBps"] = round(BYTES["Dwhat_dble"] * V / ms / 1e6, 1)
        log("bench: 48^4 Dw_dble %.1f GB/s, Dwhat_dble %.1f GB/s" % (out["Dw_dble_GBps"], out["Dwhat_dble_GBps"]))

        def nsq(f):
            q = lib.norm_square_dble(V, 0, f)
            return q.q[0] + q.q[1]

        w = np.array(NN_WIDTHS, np.int32)
        lib.oqh_nn_create(len(NN_WIDTHS) - 1, w.ctypes.data_as(P))
        lib.oqh_nn_set_weights(wts.ctypes.data_as(P))
        out["learned"] = nn_rec
        # reference solution: FP64 Schwarz cycles, FGCR to relative residual
        # 1e-14 (FP64 residual recomputation and solution update)
        pref = lat.alloc_sd()
        lib.oqh_set_sap_prec(2)
        ifail, status = C.c_int(), C.c_int()
        t0 = time.perf_counter()
        lib.sap_gcr(24, 2048, 0, 1e-14, 0.0, e, pref, C.byref(ifail), C.byref(status))
        lat.sync()
        dt = time.perf_counter() - t0
        lib.oqh_set_sap_prec(0)
        lib.Dw_dble(0.0, pref, r)
        lib.mulr_spinor_add_dble(V, 0, r, e, -1.0)
        rres = (nsq(r) / nsq(e)) ** 0.5
        out["reference_solve"] = {"sap": "fp64", "residual": 1e-14, "status": status.value,
                                  "ifail": ifail.value, "seconds": round(dt, 3),
                                  "true_residual": float("%.3e" % rres)}
        log("bench: reference solve (FP64 SAP) status %d ifail %d true residual %.2e" %
            (status.value, ifail.value, rres))
        nref = nsq(pref)
        for name, prec, nn in (("plain_fp32_sap", 0, 0), ("fp16_storage_sap", 1, 0),
                               ("fp64_sap", 2, 0), ("fp16_sap_learned", 1, 1)):
            lib.oqh_set_sap_prec(prec)
            lib.oqh_set_nn_prec(nn)
            ifail, status = C.c_int(), C.c_int()
            lat.sync()
            t0 = time.perf_counter()
            lib.sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
            lat.sync()
            dt = time.perf_counter() - t0
            lib.Dw_dble(0.0, p, r)
            lib.mulr_spinor_add_dble(V, 0, r, e, -1.0)
            res = (nsq(r) / nsq(e)) ** 0.5
            lib.mulr_spinor_add_dble(V, 0, p, pref, -1.0)
            err = (nsq(p) / nref) ** 0.5
            out[name] = {"seconds": round(dt, 3), "status": status.value, "ifail": ifail.value,
                         "true_residual": float("%.3e" % res), "error_vs_reference": float("%.3e" % err)}
            log("bench: %s %.3f s status %d residual %.2e error %.2e" % (name, dt, status.value, res, err))
        lib.oqh_set_sap_prec(0)
        lib.oqh_set_nn_prec(0)
        out["deflated"] = deflated_solves(lat, e, p, r, pref)
        dfl = out["deflated"]
        for name in ("fp16_storage_sap", "fp64_sap", "fp16_sap_learned"):
            if out["plain_fp32_sap"]["ifail"] == 0 and out[name]["ifail"] == 0:
                out["speedup_%s_vs_plain" % name] = round(out["plain_fp32_sap"]["seconds"] /
                                                          out[name]["seconds"], 3)
        for name in ("fp32_sap", "fp16_storage_sap"):
            if out["plain_fp32_sap"]["ifail"] == 0 and dfl.get(name, {}).get("ifail", 1) == 0:
                out["speedup_deflated_%s_vs_plain" % name] = round(out["plain_fp32_sap"]["seconds"] /
                                                                   dfl[name]["seconds"], 3)
    finally:
        lat.close()
    out["learned_held_out"] = learned_held_out(wts)
    return out


def deflated_solves(lat, e, p, r, pref=None):
    """Deflated SAP+GCR (dfl_sap_gcr) on the 48^4 solver workload (field,
    source and SAP parameters set by the caller) with the subspace
    parameters of the reference's ms4.in: 4^4 DFL blocks, Ns = 28, generated
    at the solve's mass (dfl_modes, then set_Awhat for the solve), then one
    solve with FP32 and one with FP16-storage Schwarz cycles; the true
    residual and, given a reference solution pref, the error against it."""
    import ctypes as C
    import numpy as np
    P = C.c_void_p
    lib, V = lat.lib, lat.V

    def nsq(f):
        q = lib.norm_square_dble(V, 0, f)
        return q.q[0] + q.q[1]

    dfl = {"subspace": "bs 4^4, Ns 28; projection nkv 16 nmx 64 res 1e-2 nmx_gcr 24 res_gcr 1e-2; "
                       "generation kappa 1/(2(4+m0)) mu 0 ninv 10 nmr 4 ncy 4, FP16 Schwarz cycles",
           "setup": "dfl_modes + set_Awhat(0): once per gauge field, shared by all solves on it"}
    bsd = np.array((4, 4, 4, 4), np.int32)
    lib.set_dfl_parms(bsd.ctypes.data_as(P), 28)
    lib.set_dfl_pro_parms(16, 64, 1e-2, 24, 1e-2)
    lib.set_dfl_gen_parms(1.0 / (2.0 * (4.0 - 0.35)), 0.0, 10, 4, 4)
    lib.set_Awhat.restype = C.c_int
    lib.set_Awhat.argtypes = [C.c_double]
    lib.start_ranlux(0, 7)
    dif = np.zeros(2, np.int32)
    dst = np.zeros(6, np.int32)
    lat.sync()
    t0 = time.perf_counter()
    # the generation's Schwarz cycles on FP16 fields (a smoother: the
    # subspace it spans serves both solves below)
    lib.oqh_set_sap_prec(1)
    lib.dfl_modes(dif.ctypes.data_as(P), dst.ctypes.data_as(P))
    lib.oqh_set_sap_prec(0)
    # the little operator of the solve (eoflg 0, mu 0) with its little modes
    if dif[0] == 0 and lib.set_Awhat(0.0) != 0:
        dif[0] = -2
    lat.sync()
    dfl["setup_seconds"] = round(time.perf_counter() - t0, 3)
    dfl["setup_ifail"] = int(dif[0])
    dfl["setup_status"] = [int(x) for x in dst[:2]]
    log("bench: dfl_modes + set_Awhat %.3f s ifail %d status %s" % (dfl["setup_seconds"], dif[0], dst[:2].tolist()))
    if dif[0] == 0:
        nref = nsq(pref) if pref is not None else None
        for name, prec in (("fp32_sap", 0), ("fp16_storage_sap", 1)):
            lib.oqh_set_sap_prec(prec)
            lat.sync()
            t0 = time.perf_counter()
            lib.dfl_sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, dif.ctypes.data_as(P), dst.ctypes.data_as(P))
            lat.sync()
            dt = time.perf_counter() - t0
            lib.Dw_dble(0.0, p, r)
            lib.mulr_spinor_add_dble(V, 0, r, e, -1.0)
            res = (nsq(r) / nsq(e)) ** 0.5
            rec = {"seconds": round(dt, 3), "status": [int(x) for x in dst[:3]], "ifail": int(dif[0]),
                   "true_residual": float("%.3e" % res)}
            if pref is not None:
                lib.mulr_spinor_add_dble(V, 0, p, pref, -1.0)
                rec["error_vs_reference"] = float("%.3e" % ((nsq(p) / nref) ** 0.5))
            dfl[name] = rec
            log("bench: deflated %s %.3f s status %s residual %.2e" % (name, dt, dst[:3].tolist(), res))
        lib.oqh_set_sap_prec(0)
    return dfl


def sap_cycles(lat, src, ncy=20):
    """Schwarz cycles alone on the 48^4 solver workload (4^4 blocks,
    isolv=1, nmr=4): ncy calls of sap() in each precision on the residual of
    a Gaussian source, timed with device synchronisation; algorithmic flops
    per cycle 2 x V/2 x ((nmr+1) x 2076 + 48) (block MinRes with the
    even-odd block operator, nmr+1 applications of it per colour, and the
    face update)."""
    import ctypes as C
    P = C.c_void_p
    lib, V = lat.lib, lat.V
    lib.sap.argtypes = [C.c_float, C.c_int, C.c_int, P, P]
    lib.set_s2zero.argtypes = [C.c_int, C.c_int, P]
    lib.assign_sd2s.argtypes = [C.c_int, C.c_int, P, P]
    psi, eta = lat.alloc_s(), lat.alloc_s()
    flops = 2 * (V // 2) * (5 * 2076 + 48)
    rec = {"workload": "48^4 solver field, SAP 4^4 blocks isolv=1 nmr=4, %d cycles per precision" % ncy,
           "gflop_per_cycle": round(flops / 1e9, 2)}
    for name, prec in (("fp32", 0), ("fp16_storage", 1), ("fp64", 2)):
        lib.oqh_set_sap_prec(prec)
        lib.set_s2zero(V, 0, psi)
        lib.assign_sd2s(V, 0, src, eta)
        lib.sap(0.0, 1, 4, psi, eta)     # warm-up (block fields, FP16 copies)
        lat.sync()
        t0 = time.perf_counter()
        for _ in range(ncy):
            lib.sap(0.0, 1, 4, psi, eta)
        lat.sync()
        ms = 1e3 * (time.perf_counter() - t0) / ncy
        rec[name] = {"ms_per_cycle": round(ms, 3), "algorithmic_TFLOPs": round(flops / ms / 1e9, 2)}
        log("bench: SAP cycle %s %.3f ms (%.1f TFLOP/s algorithmic)" % (name, ms, flops / ms / 1e9))
    lib.oqh_set_sap_prec(0)
    return rec


def solver_lattice():
    """The 48^4 solver workload's lattice, field, SAP parameters and a
    Gaussian source (see solve_times)."""
    import ctypes as C
    import numpy as np
    import pyoqh
    P = C.c_void_p
    lat = pyoqh.Lattice((48, 48, 48, 48), (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    lib = lat.lib
    lib.start_ranlux(0, 3)
    lib.random_ud()
    lib.fwd_euler(50, 0.01)
    lib.set_ud_phase()
    lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
    lib.sap_gcr.restype = C.c_double
    lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.norm_square_dble.restype = pyoqh.Qflt
    lib.norm_square_dble.argtypes = [C.c_int, C.c_int, P]
    lib.mulr_spinor_add_dble.argtypes = [C.c_int, C.c_int, P, P, C.c_double]
    bs = np.array((4, 4, 4, 4), np.int32)
    lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
    e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
    lib.random_sd(lat.VT, 2, e, 1.0)
    return lat, e, p, r


def learned_held_out(wts):
    """The trained network on fields it was not trained on: 32^3x64
    periodic, random_ud of seeds 5-8 flowed 50 x 0.01, quark phase,
    csw=1.77, m0=-0.35, a Gaussian source each; plain, FP16 and FP16 + network
    SAP+GCR (SAP 4^4 isolv=1 nmr=4 ncy=5, nkv=24) to relative residual 1e-10."""
    import ctypes as C
    import numpy as np
    import pyoqh
    P = C.c_void_p
    res = []
    lat = pyoqh.Lattice((64, 32, 32, 32), (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        w = np.array(NN_WIDTHS, np.int32)
        lib.oqh_nn_create(len(NN_WIDTHS) - 1, w.ctypes.data_as(P))
        lib.oqh_nn_set_weights(wts.ctypes.data_as(P))
        e, p = lat.alloc_sd(), lat.alloc_sd()
        for seed in (5, 6, 7, 8):
            lib.start_ranlux(0, seed)
            lib.random_ud()
            lib.fwd_euler(50, 0.01)
            lib.set_ud_phase()
            lib.random_sd(lat.VT, 2, e, 1.0)
            rec = {"seed": seed}
            for name, prec, nn in (("plain", 0, 0), ("fp16", 1, 0), ("fp16_learned", 1, 1)):
                lib.oqh_set_sap_prec(prec)
                lib.oqh_set_nn_prec(nn)
                ifail, status = C.c_int(), C.c_int()
                lat.sync()
                t0 = time.perf_counter()
                lib.sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
                lat.sync()
                rec[name] = {"seconds": round(time.perf_counter() - t0, 3), "status": status.value,
                             "ifail": ifail.value}
            lib.oqh_set_sap_prec(0)
            lib.oqh_set_nn_prec(0)
            log("bench: held-out field seed %d: plain %.3f s (%d), fp16 %.3f s (%d), learned %.3f s (%d)" %
                (seed, rec["plain"]["seconds"], rec["plain"]["status"], rec["fp16"]["seconds"],
                 rec["fp16"]["status"], rec["fp16_learned"]["seconds"], rec["fp16_learned"]["status"]))
            res.append(rec)
    finally:
        lat.close()
    return {"workload": "32^3x64 periodic, random_ud seeds 5-8 flowed 50 x 0.01, csw=1.77, m0=-0.35, "
                        "nkv=24, res=1e-10, SAP 4^4 isolv=1 nmr=4 ncy=5", "fields": res}


def multi_rhs():
    """Several right-hand sides on the 32^3x64 operator workload's field
    (open bc, cF=1, csw=1.77, random_ud links of seed 2 flowed by 100 steps
    of 0.01) at m0=0.1 (the reference build's SAP+GCR on this field fails to
    reach 1e-10 in 256 vectors at m0=0.0 and needs 20 at m0=0.1): the 12
    spin-colour point sources at x = (32, 0, 0, 0) (open bc
    make the time slices 0 and 63 inactive) and 4 Gaussian sources, solved
    together by sap_gcr_multi (FP32 SAP on 4^4 blocks, isolv=1, nmr=4,
    ncy=5; FGCR nkv=16 to 1e-10), against one sap_gcr call per source for
    two of them."""
    import ctypes as C
    import numpy as np
    import pyoqh
    P = C.c_void_p
    lat = pyoqh.Lattice((64, 32, 32, 32), (8, 8, 8, 8), bc=0, m0=0.1, csw=1.77)
    rec = {"workload": "32^3x64 open bc (cF=1), csw=1.77, m0=0.1, random_ud seed 2 flowed 100 x 0.01, "
                       "12 point sources at (32,0,0,0) + 4 Gaussian sources, FP32 SAP 4^4 isolv=1 nmr=4 ncy=5, "
                       "nkv=16, res 1e-10"}
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        V = lat.V
        lib.start_ranlux(0, 2)
        lib.random_ud()
        lib.fwd_euler(100, 0.01)
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        lib.bnd_sd2zero.argtypes = [C.c_int, P]
        n = 16
        es, ps = [lat.alloc_sd() for _ in range(n)], [lat.alloc_sd() for _ in range(n)]
        ix = int(lat.ipt()[32 * 32 * 32 * 32])
        for k in range(12):
            e = np.zeros((V, 24))
            e[ix, 2 * k] = 1.0
            lat.put_sd(es[k], e)
        for k in range(12, 16):
            lib.random_sd(lat.VT, 2, es[k], 1.0)
            lib.bnd_sd2zero(0, es[k])
        rho = np.zeros(n)
        ifail = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        lat.sync()
        t0 = time.perf_counter()
        lib.sap_gcr_multi(n, 16, 1024, 0, 1e-10, 0.0, (P * n)(*es), (P * n)(*ps), rho.ctypes.data_as(P),
                          ifail.ctypes.data_as(P), status.ctypes.data_as(P))
        lat.sync()
        t1 = time.perf_counter()
        rec["multi_s"] = round(t1 - t0, 3)
        rec["status"] = [int(x) for x in status]
        rec["ifail"] = [int(x) for x in ifail]
        single = []
        for k in (0, 12):
            fi, st = C.c_int(), C.c_int()
            lat.sync()
            t0 = time.perf_counter()
            lib.sap_gcr(16, 1024, 0, 1e-10, 0.0, es[k], ps[k], C.byref(fi), C.byref(st))
            lat.sync()
            single.append({"source": k, "seconds": round(time.perf_counter() - t0, 3), "status": st.value})
        rec["single"] = single
        per_vec_multi = rec["multi_s"] / max(1, int(status.sum()))
        per_vec_single = sum(x["seconds"] for x in single) / max(1, sum(x["status"] for x in single))
        rec["ms_per_krylov_vector_multi"] = round(1e3 * per_vec_multi, 3)
        rec["ms_per_krylov_vector_single"] = round(1e3 * per_vec_single, 3)
        log("bench: multi-RHS, 16 sources in %.2f s (%.2f ms per vector; single %.2f ms)" % (
            rec["multi_s"], 1e3 * per_vec_multi, 1e3 * per_vec_single))
    finally:
        lat.close()
    return rec


def uneven_batch():
    """The uneven batch: 16 independent solves on lattices 8^4, 12^4, 16^4 and
    24^4 (4 each; periodic bc with the quark phase, csw=1.77, m0=-0.35 as the
    solver workload), random_ud links of seeds 10-25 flowed by 50 Euler
    steps of 0.01 on the device, Gaussian sources; the SAP block size of each lattice is chosen by the autotuner
    (isolv=1, nmr=4, ncy=5; 12^4 admits only 6^4), FGCR nkv=24 to 1e-10.  The
    lattices run one after another (the library holds one geometry)."""
    import ctypes as C
    import numpy as np
    import pyoqh
    P = C.c_void_p
    rec = {"lattices": [], "workload": "16 solves, 4 each of 8^4, 12^4, 16^4, 24^4, periodic bc, "
                                       "csw=1.77, m0=-0.35, flow 50 x 0.01, autotuned SAP blocks, res 1e-10"}
    t_all = time.perf_counter()
    for k, n in enumerate([8] * 4 + [12] * 4 + [16] * 4 + [24] * 4):
        # thread layout of the reference run: one ranlxd stream per thread
        ltrd = {8: (4, 4, 4, 4), 12: None, 16: (8, 8, 8, 8), 24: (8, 8, 8, 8)}[n]
        lat = pyoqh.Lattice((n,) * 4, ltrd, bc=3, m0=-0.35, csw=1.77)
        try:
            lib = lat.lib
            lib.sap_gcr.restype = C.c_double
            lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
            lib.start_ranlux(0, 10 + k)
            lib.random_ud()
            lib.fwd_euler(50, 0.01)
            lib.set_ud_phase()
            e, p = lat.alloc_sd(), lat.alloc_sd()
            lib.random_sd(lat.VT, 2, e, 1.0)
            lat.sync()
            t0 = time.perf_counter()
            bs = np.zeros(4, np.int32)
            lib.oqh_sap_autotune(1, 4, 5, 0, bs.ctypes.data_as(P))
            t1 = time.perf_counter()
            ifail, status = C.c_int(), C.c_int()
            lib.sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
            lat.sync()
            t2 = time.perf_counter()
            rec["lattices"].append({"L": n, "seed": 10 + k, "block": [int(b) for b in bs],
                                    "autotune_s": round(t1 - t0, 3), "solve_s": round(t2 - t1, 3),
                                    "status": status.value, "ifail": ifail.value})
        finally:
            lat.close()
    rec["total_s"] = round(time.perf_counter() - t_all, 2)
    rec["solve_s"] = round(sum(x["solve_s"] for x in rec["lattices"]), 3)
    log("bench: uneven batch, 16 solves in %.2f s (solves %.2f s)" % (rec["total_s"], rec["solve_s"]))
    return rec


def host_info():
    """Cores of the host and of this job, and the CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        aff = None
    return {"nproc": os.cpu_count(), "affinity_cores": aff,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_solves():
    """SAP+GCR of the reference build on the host (FP32 SAP, FP64 GCR, SAP 4^4
    isolv=1 nmr=4 ncy=5): the 8^4 configuration (periodic, csw=1.0,
    m0=-0.5, random_ud of seed 1; NTHREAD=8, the most the block rules allow
    with 16 blocks) solved to 1e-10, and a bounded sample of the 48^4 solver
    workload (periodic, csw=1.77, m0=-0.35, random_ud of seed 3 without the
    flow, nkv=24; NTHREAD=16): 4 Krylov vectors, whose time per vector
    times the Krylov count of a solve gives its CPU time."""
    out = {}
    runs = (("8x8x8x8_avx", "config1_8^4", ["-bc", "3", "-csw", "1.0", "-m0", "-0.5", "-seed", "1",
                                              "-nkv", "16", "-nmx", "256"], "8"),
            ("48x48x48x48_avx", "solver_48^4_sample", ["-bc", "3", "-csw", "1.77", "-m0", "-0.35", "-seed",
                                                       "3", "-nkv", "24", "-nmx", "4"], "16"))
    for tag, name, extra, nt in runs:
        exe = os.path.join(ROOT, "oracle", "_ref", "ref_" + tag)
        if not os.path.exists(exe):
            continue
        env = dict(os.environ, OMP_NUM_THREADS=nt)
        try:
            r = subprocess.run([exe, "gcr", "-isolv", "1", "-bs", "4", "4", "4", "4", "-nmr", "4", "-ncy", "5",
                                "-res", "1e-10"] + extra, capture_output=True, text=True, timeout=600,
                               env=env, cwd="/tmp")
            j = json.loads(r.stdout.strip().splitlines()[-1])
        except Exception as e:  # pragma: no cover
            log("cpu solve %s failed: %s" % (tag, e))
            continue
        j["sec_per_krylov_vector"] = j["gcr_sec"] / max(j["gcr_status"], 1)
        out[name] = j
        log("bench: reference CPU SAP+GCR %s: status %d in %.3f s (%.3f s per Krylov vector)" %
            (name, j["gcr_status"], j["gcr_sec"], j["sec_per_krylov_vector"]))
    return out


def cpu_baseline(V):
    """Reference CPU Dw_dble (oracle/_ref/ref_64x32x32x32_avx: the
    reference's modules built with -DAVX -DFMA3, OpenMP, NTHREAD=16, the CPU
    share of the job) and the reference's SAP+GCR on the host."""
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
            "cores": j["nthread"], "kind": "reference", "host": host_info(),
            "sap_gcr": cpu_solves(),
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
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the solver section (training, reference solve, variants): "
                         "prints its JSON object")
    ap.add_argument("--sap-only", action="store_true",
                    help="only Schwarz cycles on the 48^4 solver workload (profiling): prints their JSON object")
    ap.add_argument("--deflated-only", action="store_true",
                    help="only the deflated solves of the 48^4 solver workload: prints their JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one process, one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pyoqh

    if args.sap_only:
        lat, e, p, r = solver_lattice()
        try:
            print(json.dumps({"sap_cycles": sap_cycles(lat, e)}))
        finally:
            lat.close()
        return

    if args.deflated_only:
        lat, e, p, r = solver_lattice()
        try:
            print(json.dumps({"deflated": deflated_solves(lat, e, p, r)}))
        finally:
            lat.close()
        return

    if args.solve_only:
        print(json.dumps({"solve": solve_times(args), "uneven_batch": uneven_batch(),
                          "multi_rhs": multi_rhs()}))
        return

    lat = pyoqh.Lattice(L, LTRD, bc=0, cF=(1.0, 1.0), m0=-0.3, csw=1.77)
    V = lat.V
    rng = np.random.default_rng(2)
    log("bench: random_ud of seed 2 (%d links)" % (4 * V))
    lat.lib.start_ranlux(0, 2)
    lat.lib.random_ud()
    log("bench: Wilson flow, 100 Euler steps of 0.01 on the device")
    lat.lib.fwd_euler(100, 0.01)
    assert lat.sw_term(3) == 0
    s, r = lat.alloc_sd(), lat.alloc_sd()
    lat.lib.random_sd(lat.VT, 2, s, 1.0)
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

    # batched operator: 12 fields (the spin-colour components of a point
    # source) per call, links and Pauli matrices loaded once for 8 of them
    import ctypes as C
    assert lat.sw_term(3) == 0
    nm = 12
    fs = [lat.alloc_sd() for _ in range(nm)]
    fr = [lat.alloc_sd() for _ in range(nm)]
    for f in fs:
        lat.put_sd(f, rng.standard_normal((V, 24)))
    arr_s, arr_r = (C.c_void_p * nm)(*fs), (C.c_void_p * nm)(*fr)
    ms = lib.oqh_time_dirac_multi(1, 0.0, nm, arr_s, arr_r, 20)
    others["Dw_dble_multi12_ms_per_field"] = round(ms / nm, 4)
    others["Dw_dble_multi12_effective_GBps"] = round(nm * BYTES["Dw_dble"] * V / ms / 1e6, 1)

    copy_bw = copy_bandwidth(lib)
    lat.close()
    solve = None if args.no_solve else solve_times(args)
    batch = None if args.no_solve else uneven_batch()
    mrhs = None if args.no_solve else multi_rhs()
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
        "data": "synthetic: the reference's random_ud field (start_ranlux(0, 2), ranlxd on the device) "
                "smoothed by the Wilson flow (100 forward-Euler steps, eps 0.01, on the device), "
                "Gaussian FP64 source (random_sd)",
        "config": {"workload": "Dw_dble 32^3x64 open bc (type 0, cF=1), csw=1.77, m0=-0.3",
                   "lattice": "64x32x32x32", "bytes_per_site": BYTES["Dw_dble"],
                   "pct_of_8TBps": round(100 * value / PEAK_HBM, 1),
                   "pct_of_copy": round(100 * value / copy_bw, 1) if copy_bw else None,
                   "copy_GBps": round(copy_bw, 1) if copy_bw else None, **others},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4), "traffic": measured_traffic()},
        "cpu_baseline": cpu,
        "solve": solve,
        "uneven_batch": batch,
        "multi_rhs": mrhs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
