#!/usr/bin/env python3
"""Writes the golden fixtures of the test suite.

ranlux_kat.json: the known-answer vectors xsn[96] (ranlxs) and xdn[48]
(ranlxd) of the reference's devel/nompi/random/check1.c:110-254, read as data
from that file, with the generator set-up of the same program
(rlxs_init(1,0,32767,1), rlxd_init(1,1,32767,1), 20 calls of 204 resp. 99
numbers, entries 60.. resp. 39.. of the last call).

ref_8x8x8x8_*.npz: per-point samples and norms of the reference build's
outputs (oracle/_ref, see oracle/build_ref.py) for the 8^4 configurations, so
that the oracle can be checked against the reference where the reference
build is absent.

ref_16x8x8x8_dfl.json: Krylov counts and true residuals of the reference
build's SAP+GCR and deflated SAP+GCR on one field and source (the GPU test
of the deflated solver compares its counts with them).
"""
import json
import os
import re
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def kat():
    src = open("/root/reference/devel/nompi/random/check1.c").read()
    xsn = {int(k): float(v) for k, v in re.findall(r"xsn\[(\d+)\]=([0-9.]+)f;", src)}
    xdn = {int(k): float(v) for k, v in re.findall(r"xdn\[(\d+)\]=([0-9.]+);", src)}
    assert sorted(xsn) == list(range(96)) and sorted(xdn) == list(range(48))
    out = {"source": "devel/nompi/random/check1.c:110-254",
           "ranlxs": {"level": 0, "seed": 32767, "calls": 20, "per_call": 204, "offset": 60,
                      "values_times_2p24": [xsn[k] for k in range(96)]},
           "ranlxd": {"level": 1, "seed": 32767, "calls": 20, "per_call": 99, "offset": 39,
                      "values_times_2p48": [xdn[k] for k in range(48)]}}
    json.dump(out, open(os.path.join(HERE, "ranlux_kat.json"), "w"), indent=1)


def ref_samples():
    import pyoracle as po
    cases = {"periodic_cfg1": dict(bc=3, csw=1.0, m0=-0.5, seed=1, mu=0.0, isolv=0),
             "open": dict(bc=0, csw=1.77, m0=-0.3, seed=2, mu=0.0, isolv=1)}
    rng = np.random.default_rng(0)
    for name, kw in cases.items():
        d = tempfile.mkdtemp()
        ld = po.ref_dump("8x8x8x8_t8", d, bs=(4, 4, 4, 4), nmr=4, ncy=5, nkv=16, nmx=256,
                         res=1e-10, cF=(1.0, 1.0), **kw)
        if ld is None:
            print("reference build not present")
            return
        V = 4096
        pts = np.sort(rng.choice(V, 96, replace=False)).astype(np.int32)
        info = json.load(open(os.path.join(d, "info.json")))
        arrs = {"points": pts, "info": json.dumps({**info, **kw})}
        for nm, dt, w in [("Dw_dble.f64", np.float64, 24), ("Dwhat_dble.f64", np.float64, 24),
                          ("Dw.f32", np.float32, 24), ("swd.f64", np.float64, 72),
                          ("eta.f64", np.float64, 24), ("sap_psi.f32", np.float32, 24),
                          ("gcr_psi.f64", np.float64, 24)]:
            a = ld(nm, dt, (V, w))
            key = nm.replace(".", "_")
            arrs[key] = a[pts]
            arrs[key + "_norm"] = np.linalg.norm(a.astype(np.float64))
        np.savez_compressed(os.path.join(HERE, "ref_8x8x8x8_%s.npz" % name), **arrs)


def flow_samples():
    """Plaquette force of random_ud (seed 2) and the field after 4 Euler
    steps of the Wilson flow (eps 0.02), sampled at 96 links, for periodic and
    open bc on 8^4."""
    import pyoracle as po
    rng = np.random.default_rng(1)
    for bc, name in ((3, "periodic"), (0, "open")):
        d = tempfile.mkdtemp()
        ld = po.ref_dump("8x8x8x8_t8", d, bc=bc, seed=2, flow=(4, 0.02))
        if ld is None:
            print("reference build not present")
            return
        nl = 4 * 4096
        links = np.sort(rng.choice(nl, 96, replace=False)).astype(np.int32)
        frc = ld("frc0.f64", np.float64, (nl, 8))
        udf = ld("ud_flow.f64", np.float64, (nl, 18))
        np.savez_compressed(os.path.join(HERE, "ref_8x8x8x8_flow_%s.npz" % name), links=links,
                            info=json.dumps({"bc": bc, "seed": 2, "nflow": 4, "eps": 0.02}),
                            frc0=frc[links], frc0_norm=np.linalg.norm(frc),
                            ud_flow=udf[links], ud_flow_norm=np.linalg.norm(udf))


DFL_CASE = dict(L=(16, 8, 8, 8), bc=3, m0=-0.5, csw=1.77, seed=4, flow=(50, 0.01), bs=(4, 4, 4, 4), isolv=1,
                nmr=4, ncy=5, nkv=16, nmx=256, res=1e-10, dbs=(4, 4, 4, 4), Ns=10,
                dpro=(16, 256, 1e-2, 24, 1e-2), dgen=(1.0 / (2.0 * (-0.5 + 4.0)), 0.0, 6, 4, 5))


def dfl_counts():
    """Krylov counts of the reference build's SAP+GCR and deflated SAP+GCR
    (dfl_modes, then dfl_sap_gcr) on its random field of seed 4 (16x8^3,
    50 flow steps, periodic bc with the quark phase) and random source."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_16x8x8x8_t8")
    if not os.path.exists(exe):
        print("reference build not present")
        return
    c = DFL_CASE
    common = ["-bc", str(c["bc"]), "-m0", repr(c["m0"]), "-csw", repr(c["csw"]), "-seed", str(c["seed"]),
              "-flow", str(c["flow"][0]), repr(c["flow"][1]), "-bs"] + [str(b) for b in c["bs"]] + \
             ["-isolv", str(c["isolv"]), "-nmr", str(c["nmr"]), "-ncy", str(c["ncy"]), "-nkv", str(c["nkv"]),
              "-nmx", str(c["nmx"]), "-res", repr(c["res"])]
    dfl = ["-dbs"] + [str(b) for b in c["dbs"]] + ["-Ns", str(c["Ns"]), "-dpro"] + \
          [repr(v) for v in c["dpro"]] + ["-dgen"] + [repr(v) for v in c["dgen"]]
    env = dict(os.environ, OMP_NUM_THREADS="8")
    out = {"case": c}
    for mode, extra in (("gcr", []), ("dfl", dfl)):
        r = subprocess.run([exe, mode] + common + extra, capture_output=True, text=True, env=env, timeout=900,
                           check=True)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
        out[mode] = json.loads(line)
    json.dump(out, open(os.path.join(HERE, "ref_16x8x8x8_dfl.json"), "w"), indent=1)


if __name__ == "__main__":
    kat()
    ref_samples()
    flow_samples()
    dfl_counts()
    print("fixtures written")
