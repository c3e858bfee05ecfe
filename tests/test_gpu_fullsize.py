"""Parity at the benchmarks' full sizes.

The Dw_dble workload of bench.py (32^3x64, open bc with cF=1, csw=1.77,
m0=-0.3, 100 Wilson-flow steps of 0.01) on the reference's own random field
(random_ud after start_ranlux(0, 2), the oracle's restatement of the
reference): the flow runs on the device; the oracle then builds the SW term
and applies Dw_dble to the flowed field (about 20 s of CPU time), and every
point must agree to 1e-13 relative, as in the small-lattice tests.

The solver workload (48^4 periodic, the reference's random_ud field of seed
3 flowed 50 steps on the device, quark phase, csw=1.77, m0=-0.35, SAP 4^4
isolv=1 nmr=4 ncy=5, FGCR nkv=24 to 1e-10): the plain and the FP16-storage
solves must reach relative residual 1e-10 when the residual is recomputed by
the oracle's Dw_dble on the exported field and solution (devel/sap/check3.c
judges the reference's solver the same way)."""
import numpy as np
import pytest

import pyoqh
import pyoracle as po
from helpers import site_rel


@pytest.mark.gpu
def test_dw_dble_32x32x32x64_flowed():
    L, T = (64, 32, 32, 32), (8, 8, 8, 8)
    print("oracle: random_ud at 32^3x64", flush=True)
    o = po.Oracle(L, T, bc=0, m0=-0.3, csw=1.77)
    o.start_ranlux(0, 2)
    o.random_ud()
    lat = pyoqh.Lattice(L, T, bc=0, m0=-0.3, csw=1.77)
    try:
        lat.import_ud(o.ud())
        lat.lib.fwd_euler(100, 0.01)
        flowed = lat.export_ud()
        o.set_ud(flowed)
        print("oracle: SW term and Dw_dble", flush=True)
        o.sw_term(3)
        assert lat.sw_term(3) == 0
        eta = np.random.default_rng(2).standard_normal((lat.V, 24))
        o.bnd_sd2zero(eta)
        ref = o.Dw_dble(0.0, eta.copy())
        s, r = lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(s, eta)
        lat.lib.Dw_dble(0.0, s, r)
        rel, absz = site_rel(lat.get_sd(r), ref)
        print("Dw_dble 32^3x64 max per-point deviation", rel)
        assert rel < 1e-13 and absz == 0.0
    finally:
        lat.close()


@pytest.mark.gpu
def test_sap_gcr_48x48x48x48_true_residual():
    import ctypes as C
    P = C.c_void_p
    L, T = (48, 48, 48, 48), (8, 8, 8, 8)
    lat = pyoqh.Lattice(L, T, bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        V = lat.V
        print("device: random_ud of seed 3, 50 flow steps, phase", flush=True)
        lib.start_ranlux(0, 3)
        lib.random_ud()
        lib.fwd_euler(50, 0.01)
        lib.set_ud_phase()
        lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        e, p = lat.alloc_sd(), lat.alloc_sd()
        lib.random_sd(lat.VT, 2, e, 1.0)
        eta = lat.get_sd(e)
        sols = {}
        for prec in (0, 1):
            lib.oqh_set_sap_prec(prec)
            ifail, status = C.c_int(), C.c_int()
            rho = lib.sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
            print("prec", prec, "status", status.value, "ifail", ifail.value, "rho", rho, flush=True)
            assert ifail.value == 0 and rho <= 1e-10 * np.linalg.norm(eta)
            sols[prec] = lat.get_sd(p)
        lib.oqh_set_sap_prec(0)
        ud = lat.export_ud()
    finally:
        lat.close()
    print("oracle: SW term and Dw_dble at 48^4", flush=True)
    o = po.Oracle(L, T, bc=3, m0=-0.35, csw=1.77)
    o.set_ud(ud)
    del ud
    assert o.sw_term(3) == 0
    ne = np.linalg.norm(eta)
    for prec, psi in sols.items():
        res = np.linalg.norm(eta - o.Dw_dble(0.0, psi.copy())) / ne
        print("prec", prec, "true residual", res, flush=True)
        assert res <= 1.01e-10
