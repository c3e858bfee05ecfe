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
        # SAP parameters set before: only the block size changes
        got = np.zeros(4, np.int32)
        iv, nm, nc = C.c_int(), C.c_int(), C.c_int()
        lib.oqh_get_sap_parms(got.ctypes.data_as(P), C.byref(iv), C.byref(nm), C.byref(nc))
        assert list(got) == list(bs) and (iv.value, nm.value, nc.value) == (1, 4, 5)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 0, 3, 4)
        bs2 = np.zeros(4, np.int32)
        assert lib.oqh_sap_autotune(1, 4, 5, 0, bs2.ctypes.data_as(P)) > 0.0
        lib.oqh_get_sap_parms(got.ctypes.data_as(P), C.byref(iv), C.byref(nm), C.byref(nc))
        assert list(got) == list(bs2) and (iv.value, nm.value, nc.value) == (0, 3, 4)
        lib.oqh_set_sap_parms(bs2.ctypes.data_as(P), 1, 4, 5)
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


@pytest.mark.gpu
def test_multi_rhs_solve():
    """Point sources (spin-colour unit vectors at the origin) and Gaussian
    sources solved together by sap_gcr_multi give bitwise the solutions and
    Krylov counts of one sap_gcr call per source, each with its true
    residual below the tolerance."""
    lat = pyoqh.Lattice((16, 16, 16, 16), (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        rng = np.random.default_rng(31)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        lib.fwd_euler(50, 0.01)
        lib.set_ud_phase()
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        V = lat.V
        srcs = []
        for sc in (0, 5, 11):
            e = np.zeros((V, 24))
            e[0, 2 * sc] = 1.0
            srcs.append(e)
        srcs += [rng.standard_normal((V, 24)) for _ in range(2)]
        n = len(srcs)
        es, ps = [lat.alloc_sd() for _ in range(n)], [lat.alloc_sd() for _ in range(n)]
        for f, e in zip(es, srcs):
            lat.put_sd(f, e)
        rho = np.zeros(n)
        ifail = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        lib.sap_gcr_multi(n, 24, 512, 0, 1e-10, 0.0, (P * n)(*es), (P * n)(*ps), rho.ctypes.data_as(P),
                          ifail.ctypes.data_as(P), status.ctypes.data_as(P))
        r = lat.alloc_sd()
        p1 = lat.alloc_sd()
        for k in range(n):
            lib.Dw_dble(0.0, ps[k], r)
            res = np.linalg.norm(srcs[k] - lat.get_sd(r)) / np.linalg.norm(srcs[k])
            fi, st = C.c_int(), C.c_int()
            rho1 = lib.sap_gcr(24, 512, 0, 1e-10, 0.0, es[k], p1, C.byref(fi), C.byref(st))
            print("source", k, "status", status[k], st.value, "true residual", res)
            assert ifail[k] == 0 and res <= 1.2e-10
            assert status[k] == st.value and rho[k] == rho1
            assert np.array_equal(lat.get_sd(ps[k]), lat.get_sd(p1))
    finally:
        lat.close()
