"""The propagator loop of main/ms4.c (oqh_ms4_propagators) against the
oracle: on the reference's random 16x8^3 field of seed 11 (periodic bc,
quark phase)
three random sources on time slice 5 are drawn as ms4's random_source does,
solved by SAP+GCR and written as quark-field files <dir>/<nbase>n<icnfg>.s<k>.

* the file of each solution is byte-identical to the oracle's encoding of
  the solution read back from it (header, square norm, lexicographic order);
* the solution solves the oracle's own random source of ms4.c (drawn after
  the same RANLUX history) to the requested residual, recomputed with the
  oracle's Dw_dble (devel/sap/check3.c judges the reference's solver so);
* Krylov counts are reported per source, and the open-bc boundary slice is
  refused as a source slice."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoqh
import pyoracle as po

P = C.c_void_p


@pytest.mark.gpu
def test_ms4_propagators(tmp_path):
    L, T = (16, 8, 8, 8), (8, 8, 8, 8)
    lat = pyoqh.Lattice(L, T, bc=3, m0=-0.3, csw=1.77)
    try:
        lib = lat.lib
        lib.oqh_ms4_propagators.restype = C.c_int
        lib.oqh_ms4_propagators.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.c_char_p, C.c_char_p, C.c_int, P]
        # the reference's random field of seed 11 with the quark phase, then
        # one random_sd, as ms4 would find the streams after other work
        lib.start_ranlux(0, 11)
        lib.random_ud()
        lib.set_ud_phase()
        s = lat.alloc_sd()
        lib.random_sd(lat.VT, 2, s, 1.0)
        lat.free(s)
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        status = np.zeros(3, np.int32)
        nfail = lib.oqh_ms4_propagators(5, 3, 16, 256, 0, 1e-10, 0.0, str(tmp_path).encode(), b"run",
                                        7, status.ctypes.data_as(P))
        print("status", status.tolist(), "failed", nfail, flush=True)
        assert nfail == 0 and (status > 0).all()
        lib.oqh_set_error_mode(0)
        assert lib.oqh_ms4_propagators(L[0], 1, 16, 256, 0, 1e-10, 0.0, str(tmp_path).encode(), b"x", 0,
                                       status.ctypes.data_as(P)) == -1
        lib.oqh_last_error(None, 0)
        lib.oqh_set_error_mode(1)
    finally:
        lat.close()
    # the oracle with the same history draws the same sources
    o2 = po.Oracle(L, T, bc=3, m0=-0.3, csw=1.77)
    o2.start_ranlux(0, 11)
    o2.random_ud()
    o2.set_ud_phase()
    o2.random_sd()
    o2.sw_term(3)
    ipt = o2.ipt()
    for k in range(3):
        src = o2.random_source(5)
        path = os.path.join(str(tmp_path), "runn7.s%d" % k)
        data = open(path, "rb").read()
        Lf, eo, norm, psi = po.sfld_decode(data, ipt, o2.V)
        assert list(Lf) == list(L) and eo == 0 and not np.isnan(psi).any()
        assert data == po.sfld_encode(L, ipt, psi)
        res = np.linalg.norm(src - o2.Dw_dble(0.0, psi.copy())) / np.linalg.norm(src)
        print("source", k, "nonzero points", int((np.abs(src).max(axis=1) > 0).sum()), "true residual", res)
        assert res <= 1.01e-10
        assert int((np.abs(src).max(axis=1) > 0).sum()) == L[1] * L[2] * L[3]
