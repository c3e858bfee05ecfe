"""fgcr driven by host callbacks (linsolv/fgcr.c) and the unsafe-inversion
path of sap_gcr (sap/sap_gcr.c:181-251, sw_term/pauli_dble.c:1718-1758),
each on its own lattice."""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import make_oracle
from test_gpu_linalg_more import _sig

P = C.c_void_p


@pytest.mark.gpu
def test_fgcr_host_callbacks():
    """fgcr with Dop and Mop written on the host (ctypes callbacks): Mop runs
    ncy Schwarz cycles as sap_gcr's does (sap/sap_gcr.c:126-137); with the
    source normalised by its uniform norm this is sap_gcr itself, so the
    Krylov count and the solution agree with sap_gcr's, and the true
    residual recomputed by the oracle reaches the tolerance."""
    o, eta, _ = make_oracle((8, 8, 8, 8), None, bc=3, m0=-0.5, csw=1.0, seed=1)
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3, m0=-0.5, csw=1.0)
    try:
        lib = lat.lib
        _sig(lib)
        lat.import_ud(o.ud())
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        V, nkv = lat.V, 16
        DOP = C.CFUNCTYPE(None, P, P)
        MOP = C.CFUNCTYPE(None, C.c_int, P, P, P)
        calls = {"D": 0, "M": 0}

        def dop(s, r):
            calls["D"] += 1
            lib.Dw_dble(0.0, s, r)

        def mop(k, rho, phi, chi):
            calls["M"] += 1
            lib.set_s2zero(V, 0, phi)
            lib.assign_s2s(V, 0, rho, chi)
            for _ in range(5):
                lib.sap(0.0, 1, 4, phi, chi)
            lib.diff_s2s(V, 0, rho, chi)

        cd, cm = DOP(dop), MOP(mop)
        lib.fgcr.restype = C.c_double
        lib.fgcr.argtypes = [C.c_int, DOP, MOP, P, P, C.c_int, C.c_int, C.c_int, C.c_double, P, P,
                             C.POINTER(C.c_int)]
        e, p, p2 = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        rho0 = float(np.sqrt((eta ** 2).sum(axis=1)).max())
        lat.put_sd(e, eta / rho0)
        assert lat.sw_term(3) == 0
        lib.alloc_ws(2 * nkv + 1)
        lib.alloc_wsd(1)
        ws = lib.reserve_ws(2 * nkv + 1)
        wsd = lib.reserve_wsd(1)
        st = C.c_int()
        rn = lib.fgcr(V, cd, cm, ws, wsd, nkv, 256, 0, 1e-10, e, p, C.byref(st))
        lib.release_wsd()
        lib.release_ws()
        psi = lat.get_sd(p) * rho0
        o.sw_term(3)
        res = np.linalg.norm(eta - o.Dw_dble(0.0, psi.copy())) / np.linalg.norm(eta)
        print("fgcr status", st.value, "rn", rn, "calls", calls, "true residual", res, flush=True)
        assert st.value > 0 and calls["M"] == st.value and res <= 1.2e-10
        assert np.array_equal(lat.get_sd(e), eta / rho0)   # eta unchanged
        # the same solve through sap_gcr
        lat.put_sd(e, eta)
        ifail, st2 = C.c_int(), C.c_int()
        lib.sap_gcr(nkv, 256, 0, 1e-10, 0.0, e, p2, C.byref(ifail), C.byref(st2))
        assert ifail.value == 0 and st2.value == st.value
        assert np.abs(lat.get_sd(p2) - psi).max() <= 1e-12 * np.abs(psi).max()
    finally:
        lat.close()


@pytest.mark.gpu
def test_unsafe_inversion_sap_gcr():
    """With the diagonal of the SW term cancelled (m0 = -4, so c(x0) = 4 + m0
    = 0) the odd Pauli blocks are ill-conditioned: inv_pauli_dble reports an
    unsafe inversion exactly where the oracle does, sw_term(ODD_PTS) returns
    nonzero, and sap_gcr with the even-odd block solver (isolv=1, which
    inverts the odd blocks) returns ifail = -3, psi = 0 and the norm of eta
    (sap/sap_gcr.c:204-251)."""
    o, eta, _ = make_oracle((8, 8, 8, 8), None, bc=3, m0=-4.0, csw=1.0, seed=3)
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3, m0=-4.0, csw=1.0)
    try:
        lib = lat.lib
        _sig(lib)
        lat.import_ud(o.ud())
        ifail_o = o.sw_term(2)
        assert ifail_o != 0, "the oracle must find an unsafe inversion"
        assert lat.sw_term(2) != 0
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        e, p = lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(e, eta)
        lat.put_sd(p, eta)
        ifail, status = C.c_int(), C.c_int()
        rho = lib.sap_gcr(16, 64, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
        print("unsafe: ifail", ifail.value, "status", status.value, "rho", rho, flush=True)
        assert ifail.value == -3 and status.value == 0
        assert not lat.get_sd(p).any()
        assert abs(rho - np.linalg.norm(eta)) <= 1e-12 * np.linalg.norm(eta)
    finally:
        lat.close()
