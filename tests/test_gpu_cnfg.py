"""Plaquette sums (plaq_sum_dble, plaq_wsum_dble) and the exchange of gauge
configurations (export_cnfg, import_cnfg, lat_sizes) and quark fields
(export_sfld, import_sfld) through the C-ABI, against the oracle: its
plaquette sums and its encoders of the file formats (bit-exact with the
reference build, tests/test_oracle_cnfg.py).

The sums are double-double reductions of per-point values on the device; the
per-plaquette products round as the reference's only up to the order of the
trace sum, so they agree to 1e-13 relative (exactly on the unit field).  The
file's link records are moved, not computed: they must be byte-identical."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoqh
import pyoracle as po

LAT = ((16, 8, 8, 8), (8, 8, 8, 8))


def last_error(lat):
    buf = C.create_string_buffer(512)
    n = lat.lib.oqh_last_error(buf, 512)
    return n, buf.value.decode()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [0, 3])
def test_plaq_sums(bc):
    print("plaquette sums, bc", bc, flush=True)
    L, T = LAT
    o = po.Oracle(L, T, bc=bc)
    o.start_ranlux(0, 5)
    o.random_ud()
    o.fwd_euler(2, 0.02)
    lat = pyoqh.Lattice(L, T, bc=bc)
    try:
        lat.import_ud(o.ud())
        for iw, fn in ((0, lat.lib.plaq_sum_dble), (1, lat.lib.plaq_wsum_dble)):
            ref = o.plaq_sum(iw)
            got = fn(1)
            print("iw", iw, got, ref, flush=True)
            assert abs(got - ref) <= 1e-13 * abs(ref)
        # unit field (links across the open boundary zero): exact counts
        V = lat.V
        ud = np.zeros((4 * V, 18))
        ud[:, [0, 8, 16]] = 1.0
        nt = V // L[0]
        if bc == 0:
            x0 = np.indices(L).reshape(4, -1)[0][np.argsort(lat.ipt())]
            h = V // 2
            t = x0[h:]
            ud[8 * np.flatnonzero(t == L[0] - 1)] = 0.0
            ud[8 * np.flatnonzero(t == 0) + 1] = 0.0
        lat.import_ud(ud)
        if bc == 3:
            assert lat.lib.plaq_sum_dble(1) == 18.0 * V == lat.lib.plaq_wsum_dble(1)
        else:
            assert lat.lib.plaq_sum_dble(1) == 9.0 * (V - nt) + 9.0 * V
            assert lat.lib.plaq_wsum_dble(1) == 18.0 * (V - nt)
    finally:
        lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [0, 3])
def test_export_import(tmp_path, bc):
    print("configuration files, bc", bc, flush=True)
    L, T = LAT
    o = po.Oracle(L, T, bc=bc)
    o.start_ranlux(0, 7)
    o.random_ud()
    ref = o.cnfg_bytes()
    path = str(tmp_path / "dev.cnfg")
    lat = pyoqh.Lattice(L, T, bc=bc)
    try:
        lat.import_ud(o.ud())
        lat.lib.export_cnfg(path.encode())
        data = open(path, "rb").read()
        assert len(data) == len(ref)
        assert data[:16] == ref[:16] and data[24:] == ref[24:]
        p_dev = np.frombuffer(data[16:24], "<f8")[0]
        p_ref = np.frombuffer(ref[16:24], "<f8")[0]
        assert abs(p_dev - p_ref) <= 1e-13 * abs(p_ref)
        nl = np.zeros(4, np.int32)
        lat.lib.lat_sizes(path.encode(), nl.ctypes.data_as(C.c_void_p))
        assert tuple(nl) == L
        # the oracle's file, read back into an overwritten field
        opath = str(tmp_path / "oracle.cnfg")
        open(opath, "wb").write(ref)
        lat.import_ud(np.roll(o.ud(), 1, axis=0))
        lat.lib.set_ud_phase()
        lat.lib.import_cnfg(opath.encode(), 0)
        assert np.array_equal(lat.export_ud(), o.ud())
        # the imported field carries no phase: exporting it again is allowed
        lat.lib.export_cnfg(path.encode())
        assert open(path, "rb").read()[24:] == ref[24:]

        lat.lib.oqh_set_error_mode(0)
        last_error(lat)
        # a phase-set field is refused (archive.c:812-813)
        lat.lib.set_ud_phase()
        lat.lib.export_cnfg(str(tmp_path / "x.cnfg").encode())
        n, msg = last_error(lat)
        assert n == 1 and "phase-modified" in msg
        lat.lib.unset_ud_phase()
        # a wrong average plaquette on the file (archive.c:966-971)
        bad = bytearray(ref)
        bad[16:24] = np.array([p_ref * (1.0 + 1e-9)], "<f8").tobytes()
        bpath = str(tmp_path / "bad.cnfg")
        open(bpath, "wb").write(bytes(bad))
        lat.lib.import_cnfg(bpath.encode(), 0)
        n, msg = last_error(lat)
        assert n == 1 and "Incorrect average plaquette" in msg
        # other lattice sizes (archive.c:335-357 with the whole lattice local)
        bad = bytearray(ref)
        bad[0:16] = np.array([8, 8, 8, 8], "<i4").tobytes()
        open(bpath, "wb").write(bytes(bad))
        lat.lib.import_cnfg(bpath.encode(), 0)
        n, msg = last_error(lat)
        assert n == 1 and "lattice sizes" in msg
        # a truncated file
        open(bpath, "wb").write(ref[:-144])
        lat.lib.import_cnfg(bpath.encode(), 0)
        n, msg = last_error(lat)
        assert n == 1 and "read count" in msg
        lat.lib.import_cnfg(str(tmp_path / "missing.cnfg").encode(), 0)
        n, msg = last_error(lat)
        assert n == 1 and "Unable to open" in msg
        # a field with nonzero links across the time boundary under open bc
        # (and the converse) does not match the boundary conditions
        other = po.Oracle(L, T, bc=3 - bc)
        other.start_ranlux(0, 9)
        other.random_ud()
        open(bpath, "wb").write(other.cnfg_bytes())
        lat.lib.import_cnfg(bpath.encode(), 0)
        n, msg = last_error(lat)
        assert n == 1 and "Incompatible boundary conditions" in msg
    finally:
        lat.lib.oqh_set_error_mode(1)
        lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("eo", [0, 1])
def test_sfld_export_import(tmp_path, eo):
    print("quark-field files, eo", eo, flush=True)
    L, T = LAT
    o = po.Oracle(L, T, bc=3)
    o.start_ranlux(0, 4)
    eta = o.random_sd(1.0)
    ipt = o.ipt()
    ref = po.sfld_encode(L, ipt, eta, eo)
    lat = pyoqh.Lattice(L, T, bc=3)
    try:
        V = lat.V
        n = V if eo == 0 else V // 2
        s, r = lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(s, eta)
        path = str(tmp_path / "dev.sfld")
        # device field and host array give the same file
        for arg in (s, eta.ctypes.data_as(C.c_void_p)):
            lat.lib.export_sfld(path.encode(), eo, arg)
            data = open(path, "rb").read()
            assert len(data) == len(ref) and data[:20] == ref[:20] and data[28:] == ref[28:]
            nd = np.frombuffer(data[20:28], "<f8")[0]
            nr = np.frombuffer(ref[20:28], "<f8")[0]
            assert abs(nd - nr) <= 1e-14 * nr
        # the oracle's file into a device field and into a host array; with
        # eo=1 the odd points keep their values
        opath = str(tmp_path / "oracle.sfld")
        open(opath, "wb").write(ref)
        fill = np.random.default_rng(1).standard_normal((V, 24))
        lat.put_sd(r, fill)
        lat.lib.import_sfld(opath.encode(), eo, r)
        got = lat.get_sd(r)
        assert np.array_equal(got[:n], eta[:n]) and np.array_equal(got[n:], fill[n:])
        host = fill.copy()
        lat.lib.import_sfld(opath.encode(), eo, host.ctypes.data_as(C.c_void_p))
        assert np.array_equal(host[:n], eta[:n]) and np.array_equal(host[n:], fill[n:])

        lat.lib.oqh_set_error_mode(0)
        last_error(lat)
        lat.lib.import_sfld(opath.encode(), 1 - eo, r)
        n_err, msg = last_error(lat)
        assert n_err == 1 and "Lattice sizes do not match" in msg
        bad = bytearray(ref)
        bad[20:28] = np.array([nr * (1.0 + 1e-10)], "<f8").tobytes()
        bpath = str(tmp_path / "bad.sfld")
        open(bpath, "wb").write(bytes(bad))
        lat.lib.import_sfld(bpath.encode(), eo, r)
        n_err, msg = last_error(lat)
        assert n_err == 1 and "Incorrect square norm" in msg
        lat.lib.export_sfld(path.encode(), 2, s)
        n_err, msg = last_error(lat)
        assert n_err == 1 and "out of range" in msg
    finally:
        lat.lib.oqh_set_error_mode(1)
        lat.close()
