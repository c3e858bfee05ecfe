"""Device random numbers and random fields (rng.hip) against the known-answer
vectors of devel/nompi/random/check1.c (the ranlxd stream, bit-exact) and
against the oracle's restatement of random_ud and random_sd on the same seed
and thread layout.  The streams are integer arithmetic and agree exactly;
the fields pass through the device's sin, cos and log, so they agree to the
last bits (1e-14)."""
import ctypes
import json
import os

import numpy as np
import pytest

import pyoqh
import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.gpu
def test_ranlxd_known_answers():
    k = json.load(open(os.path.join(GOLD, "ranlux_kat.json")))["ranlxd"]
    assert k["level"] == 1
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3)
    try:
        lat.lib.start_ranlux(0, k["seed"])  # ranlxd runs at level + 1
        n = k["calls"] * k["per_call"]
        r = np.empty(n)
        lat.lib.ranlxd(r.ctypes.data_as(ctypes.c_void_p), n)
        last = r[(k["calls"] - 1) * k["per_call"]:]
        got = last[k["offset"]:k["offset"] + 48] * 2.0 ** 48
        assert np.array_equal(got, np.array(k["values_times_2p48"]))
    finally:
        lat.close()


CFGS = [dict(name="8^4-periodic", L=(8, 8, 8, 8), T=None, bc=3, seed=1),
        dict(name="8^4-open-nthread16", L=(8, 8, 8, 8), T=(4, 4, 4, 4), bc=0, seed=2),
        dict(name="16x8^3-periodic", L=(16, 8, 8, 8), T=None, bc=3, seed=5),
        dict(name="16^4-open-nthread16", L=(16, 16, 16, 16), T=(8, 8, 8, 8), bc=0, seed=7)]


@pytest.mark.gpu
@pytest.mark.parametrize("c", CFGS, ids=lambda c: c["name"])
def test_random_fields(c):
    print("config", c, flush=True)
    o = po.Oracle(c["L"], c["T"], bc=c["bc"])
    o.start_ranlux(0, c["seed"])
    o.random_ud()
    ud_o = o.ud()
    sd_o = o.random_sd(1.0)
    lat = pyoqh.Lattice(c["L"], c["T"], bc=c["bc"])
    try:
        lat.lib.start_ranlux(0, c["seed"])
        lat.lib.random_ud()
        ud_g = lat.export_ud()
        dev = np.abs(ud_g - ud_o).max()
        print("random_ud max deviation", dev)
        assert dev < 1e-14
        s = lat.alloc_sd()
        lat.lib.random_sd(lat.VT, 2, s, 1.0)
        dev = np.abs(lat.get_sd(s) - sd_o).max()
        print("random_sd max deviation", dev)
        assert dev < 1e-13
        # the next draws: the random source of main/ms4.c on time slice 3,
        # then random_sd(vol, 0, ..) of the master thread on 100 points
        src_o = o.random_source(3)
        lat.lib.oqh_random_source(3, s)
        got = lat.get_sd(s)
        assert np.abs(got - src_o).max() < 1e-13
        lex = np.empty(lat.V, np.int64)
        lex[o.ipt()] = np.arange(lat.V)
        assert not got[lex // (c["L"][1] * c["L"][2] * c["L"][3]) != 3].any()
        m_o = o.random_sd_master(100, 0.5)
        lat.lib.random_sd(100, 0, s, 0.5)
        assert np.abs(lat.get_sd(s)[:100] - m_o).max() < 1e-13
        # the operators accept the generated field (links SU(3))
        z = (ud_g[:, 0::2] + 1j * ud_g[:, 1::2]).reshape(-1, 3, 3)
        nz = np.abs(z).max(axis=(1, 2)) > 0
        m = z[nz]
        assert np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - np.eye(3)).max() < 1e-14
    finally:
        lat.close()


@pytest.mark.gpu
def test_start_field_philox():
    """The start fields of dfl_modes: Philox4x32-10 keyed by two ranlxd
    numbers of stream 0, bit-exact against the numpy restatement
    (oracle/philox.py, pinned by the Random123 known answers)."""
    import philox as ph
    lat = pyoqh.Lattice((16, 8, 8, 8), (8, 8, 8, 8), bc=0)
    try:
        P = ctypes.c_void_p
        lat.lib.start_ranlux(0, 9)
        s = lat.alloc_s()
        lat.lib.oqh_random_start_field(s)
        lat.lib.oqh_random_start_field(s)      # the second field: key from the next two numbers
        got = lat.get_s(s)
        lat.lib.start_ranlux(0, 9)
        key = np.empty(4)
        lat.lib.ranlxd(key.ctypes.data_as(P), 4)
        ref = ph.start_field(lat.V, key[2:])
        assert np.array_equal(got, ref)
        assert not np.array_equal(ref, ph.start_field(lat.V, key[:2]))
        lat.free(s)
    finally:
        lat.close()


PAR = [dict(name="12^4-one-stream-level2", L=(12, 12, 12, 12), T=None, bc=3, level=1, seed=3),
       dict(name="16^4-open-nthread16", L=(16, 16, 16, 16), T=(8, 8, 8, 8), bc=0, level=0, seed=21),
       dict(name="12^4-open-nthread81", L=(12, 12, 12, 12), T=(4, 4, 4, 4), bc=0, level=0, seed=8)]


def _draws(lat, c):
    """random_ud, random_sd of every thread, the ms4 random source on time
    slice 1, random_sd of the master thread, random_sd again (FP32) and 50
    ranlxd numbers, all from one start_ranlux."""
    P = ctypes.c_void_p
    lat.lib.start_ranlux(c["level"], c["seed"])
    lat.lib.random_ud()
    out = [lat.export_ud()]
    s = lat.alloc_sd()
    lat.lib.random_sd(lat.VT, 2, s, 1.0)
    out.append(lat.get_sd(s))
    lat.lib.oqh_random_source(1, s)
    out.append(lat.get_sd(s))
    lat.lib.random_sd(77, 0, s, 0.5)
    out.append(lat.get_sd(s)[:77])
    lat.lib.random_ud()
    out.append(lat.export_ud())
    r = np.empty(50)
    lat.lib.ranlxd(r.ctypes.data_as(P), 50)
    out.append(r)
    lat.free(s)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("c", PAR, ids=lambda c: c["name"])
def test_parallel_streams_match_serial(c, monkeypatch):
    """The parallel generation (every refill of every stream reached by a
    jump, the links of random_ud assigned along their chain of draws) hands
    out the same numbers and leaves the streams in the same state as the
    walk of each stream by one lane (OQH_RNG_SERIAL=1), bit for bit."""
    print("config", c, flush=True)
    lat = pyoqh.Lattice(c["L"], c["T"], bc=c["bc"])
    try:
        monkeypatch.setenv("OQH_RNG_SERIAL", "1")
        ser = _draws(lat, c)
        monkeypatch.delenv("OQH_RNG_SERIAL")
        par = _draws(lat, c)
        names = ["random_ud", "random_sd", "source", "random_sd master", "random_ud again", "ranlxd"]
        for k, (a, b) in enumerate(zip(ser, par)):
            print(names[k], "max deviation", np.abs(a - b).max(), "equal", np.array_equal(a, b), flush=True)
        for k, (a, b) in enumerate(zip(ser, par)):
            assert np.array_equal(a, b), names[k]
    finally:
        lat.close()
