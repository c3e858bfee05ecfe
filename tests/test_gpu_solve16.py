"""SAP+GCR on a 16^4 lattice (periodic bc) for both block solvers, FP32 and
FP16 storage and FP64, 4^4 and 8^4 blocks, both forms of modified Gram-Schmidt
in fgcr, judged by
the double-precision true residual |eta - Dw_dble psi| / |eta| computed on
the device (devel/sap/check3.c:205-249).  The oracle is not run at this
size."""
import ctypes as C

import numpy as np
import pytest

import pyoqh

P = C.c_void_p


@pytest.mark.gpu
def test_sap_gcr_16_4():
    lat = pyoqh.Lattice((16, 16, 16, 16), (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    try:
        rng = np.random.default_rng(3)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        lib = lat.lib
        lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        V = lat.V
        eta = rng.standard_normal((V, 24))
        counts = {}
        for isolv, bs, prec, orth in ((0, (4, 4, 4, 4), 0, 1), (1, (8, 4, 4, 4), 0, 1),
                                      (1, (4, 4, 4, 4), 0, 1), (1, (4, 4, 4, 4), 0, 0),
                                      (1, (4, 4, 4, 4), 1, 1), (1, (4, 4, 4, 4), 1, 0),
                                      (0, (4, 4, 4, 4), 1, 1), (1, (4, 4, 4, 4), 2, 1),
                                      (1, (8, 8, 8, 8), 2, 1), (1, (8, 8, 8, 8), 0, 1)):
            b = np.array(bs, np.int32)
            lib.oqh_set_sap_parms(b.ctypes.data_as(P), isolv, 4, 5)
            lib.oqh_set_sap_prec(prec)
            lib.oqh_set_gcr_orth(orth)
            e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
            lat.put_sd(e, eta)
            ifail, status = C.c_int(), C.c_int()
            rho = lib.sap_gcr(16, 256, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
            lib.Dw_dble(0.0, p, r)
            res = np.linalg.norm(eta - lat.get_sd(r)) / np.linalg.norm(eta)
            print("16^4 isolv", isolv, "bs", bs, "prec", prec, "orth", orth, "status", status.value,
                  "rho", rho, "true residual", res)
            assert ifail.value == 0 and res <= 1.2e-10
            counts[(isolv, bs, prec, orth)] = status.value
            # the returned residual is the recomputed one of the last restart
            assert abs(rho / np.linalg.norm(eta) - res) <= 1e-3 * res
            for f in (e, p, r):
                lat.free(f)
        # both forms of modified Gram-Schmidt need nearly the same number of
        # Krylov vectors (equal coefficients in exact arithmetic)
        for prec in (0, 1):
            n0, n1 = counts[(1, (4, 4, 4, 4), prec, 0)], counts[(1, (4, 4, 4, 4), prec, 1)]
            assert abs(n0 - n1) <= max(2, 0.05 * n0), (prec, n0, n1)
    finally:
        lat.close()
