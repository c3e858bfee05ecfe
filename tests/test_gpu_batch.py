"""Block-size autotuner (oqh_sap_autotune) on lattices whose legal SAP block
sizes differ (README.block rules: even, >= 4, an even number of blocks per
axis; 12^4 admits only 6^4 among blocks up to 4096 points), and SAP+GCR with
the chosen block size, judged by the double-precision true residual."""
import ctypes as C

import numpy as np
import pytest

import pyoqh

P = C.c_void_p


def legal(l, b):
    return b >= 4 and b % 2 == 0 and l % b == 0 and (l // b) % 2 == 0


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(8, 10), (12, 14), (16, 18)])
def test_autotune_and_solve(n, seed):
    print("lattice", n, "seed", seed, flush=True)
    lat = pyoqh.Lattice((n,) * 4, None, bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        rng = np.random.default_rng(seed)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        lib.fwd_euler(50, 0.01)
        lib.set_ud_phase()
        bs = np.zeros(4, np.int32)
        t = lib.oqh_sap_autotune(1, 4, 5, 0, bs.ctypes.data_as(P))
        print("block size", bs, "seconds per e-fold", t)
        assert t > 0.0
        assert all(legal(n, int(b)) for b in bs) and bs[1] == bs[2] == bs[3]
        assert np.prod(bs) <= 4096
        if n == 12:
            assert list(bs) == [6, 6, 6, 6]
        eta = rng.standard_normal((lat.V, 24))
        e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(e, eta)
        ifail, status = C.c_int(), C.c_int()
        lib.sap_gcr(24, 512, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
        lib.Dw_dble(0.0, p, r)
        res = np.linalg.norm(eta - lat.get_sd(r)) / np.linalg.norm(eta)
        print("status", status.value, "true residual", res)
        assert ifail.value == 0 and res <= 1.2e-10
    finally:
        lat.close()
