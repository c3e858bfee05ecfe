"""GPU parity of the generic SAP block path (the block solver for block sizes
other than 4^4 and the whole FP64 Schwarz variant) against the oracle:
8^4 blocks on 16^4 and 6^4 blocks on 12^4 (and 4^4 blocks on 8^4 for the
FP64 variant), both block solvers (isolv 0 and 1), periodic and open bc.

* block operators Dw_blk and Dwhat_blk (dirac/Dw.c:1475-1907) on the first
  and the last block, with the odd half of Dwhat_blk's output left as it was;
* the block MinRes solvers blk_mres / blk_eo_mres (sap/blk_solv.c:1117-1186),
  blk_eo_mres leaving the odd points of the three block fields unchanged;
* Schwarz cycles (sap/sap.c:720-790): FP32 against the oracle's FP32 sweep,
  the FP64 variant against the oracle's FP64 sweep (FP64 links and block SW
  field), both rounded to FP32 between cycles as sap() does;
* SAP+GCR (sap/sap_gcr.c:120-254): Krylov count within 1 of the oracle's and
  the double-precision true residual recomputed by the oracle.

FP32 tolerances as in test_gpu_sap.py; the FP64 cycle agrees with the
oracle's to the FP32 rounding of the exported fields (2e-6 relative per
point after three cycles)."""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import make_oracle, site_rel

P = C.c_void_p

CFGS = [
    dict(name="16^4-bs8-periodic-isolv1", L=(16, 16, 16, 16), T=(8, 8, 8, 16), bs=(8, 8, 8, 8), bc=3,
         m0=-0.35, csw=1.77, seed=6, isolv=1, mu=0.0),
    dict(name="16^4-bs8-periodic-isolv0-mu", L=(16, 16, 16, 16), T=(8, 8, 8, 16), bs=(8, 8, 8, 8), bc=3,
         m0=-0.35, csw=1.77, seed=7, isolv=0, mu=0.05),
    dict(name="12^4-bs6-open-isolv1-mu", L=(12, 12, 12, 12), T=(6, 6, 12, 12), bs=(6, 6, 6, 6), bc=0,
         m0=-0.3, csw=1.77, seed=2, isolv=1, mu=0.1),
    dict(name="12^4-bs6-periodic-isolv0", L=(12, 12, 12, 12), T=(6, 6, 12, 12), bs=(6, 6, 6, 6), bc=3,
         m0=-0.35, csw=1.77, seed=4, isolv=0, mu=0.0),
    dict(name="8^4-bs4-open-isolv1", L=(8, 8, 8, 8), T=(8, 8, 8, 8), bs=(4, 4, 4, 4), bc=0,
         m0=-0.3, csw=1.77, seed=2, isolv=1, mu=0.0),
]


def _sig(lib):
    lib.sap.argtypes = [C.c_float, C.c_int, C.c_int, P, P]
    lib.sap_gcr.restype = C.c_double
    lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
    lib.assign_s2sblk.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int]
    lib.Dw_blk.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
    lib.Dwhat_blk.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
    lib.blk_mres.argtypes = [C.c_int, C.c_float, C.c_int]
    lib.blk_eo_mres.argtypes = [C.c_int, C.c_float, C.c_int]
    lib.oqh_export_sblk.argtypes = [C.c_int, C.c_int, P]
    lib.oqh_get_imb.argtypes = [C.c_int, P]
    lib.assign_sd2sdblk.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int]
    lib.assign_sdblk2sd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P]
    for nm in ("Dw_blk_dble", "Dwhat_blk_dble", "Dwee_blk_dble", "Dwoo_blk_dble"):
        getattr(lib, nm).argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
    for nm in ("Dwoe_blk_dble", "Dweo_blk_dble"):
        getattr(lib, nm).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.oqh_export_sdblk.argtypes = [C.c_int, C.c_int, P]


@pytest.fixture(scope="module", params=CFGS, ids=lambda c: c["name"])
def setup(request):
    c = request.param
    print("config", c, flush=True)
    o, eta, s32 = make_oracle(c["L"], c["T"], bc=c["bc"], m0=c["m0"], csw=c["csw"], seed=c["seed"])
    lat = pyoqh.Lattice(c["L"], c["T"], bc=c["bc"], m0=c["m0"], csw=c["csw"])
    _sig(lat.lib)
    lat.import_ud(o.ud())
    bs = np.array(c["bs"], np.int32)
    lat.lib.oqh_set_sap_parms(bs.ctypes.data_as(P), c["isolv"], 4, 5)
    o.set_sap(c["bs"], c["isolv"], 4, 5)
    o.sw_term(3)
    assert o.sap_prepare() == 0
    yield c, o, eta, s32, lat
    lat.lib.oqh_set_sap_prec(0)
    lat.close()


def _block(lat, n, k, bvol):
    a = np.empty((bvol, 24), np.float32)
    lat.lib.oqh_export_sblk(n, k, a.ctypes.data_as(P))
    return a


@pytest.mark.gpu
def test_generic_block_operators(setup):
    """Dw_blk, Dwhat_blk (isolv=1 grid: odd block SW term inverted) and the
    MinRes solver of the grid's kind on the first and the last block."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.oqh_set_sap_prec(0)
    mu = np.float32(c["mu"])
    other = np.random.default_rng(5).standard_normal(s32.shape).astype(np.float32)
    s, t = lat.alloc_s(), lat.alloc_s()
    lat.put_s(s, s32)
    lat.put_s(t, other)
    nb = o.nblocks()
    for n in (0, nb - 1):
        imb = o.block_imb(n)
        bvol, bh = len(imb), len(imb) // 2
        got_imb = np.empty(bvol, np.int32)
        lib.oqh_get_imb(n, got_imb.ctypes.data_as(P))
        assert np.array_equal(got_imb, imb)
        s1 = np.vstack([s32[imb], np.zeros((1, 24), np.float32)])
        # Dw_blk
        lib.assign_s2sblk(0, n, 0, s, 1)
        lib.Dw_blk(0, n, c["mu"], 1, 2)
        ref = o.Dw_blk(n, mu, s1.copy())[:-1]
        rel, absz = site_rel(_block(lat, n, 2, bvol), ref)
        print("block", n, "Dw_blk", rel, flush=True)
        assert rel < 1e-5 and absz == 0.0
        if c["isolv"]:
            # Dwhat_blk: even half against the oracle, odd half of the output
            # unchanged (Dw.c:1826-1905 keeps the intermediate in registers)
            lib.assign_s2sblk(0, n, 0, t, 2)
            before = _block(lat, n, 2, bvol)
            lib.Dwhat_blk(0, n, c["mu"], 1, 2)
            out = _block(lat, n, 2, bvol)
            ref = o.Dwhat_blk(n, mu, s1.copy())[:bh]
            rel, absz = site_rel(out[:bh], ref)
            print("block", n, "Dwhat_blk", rel, flush=True)
            assert rel < 1e-5 and absz == 0.0
            assert np.array_equal(out[bh:], before[bh:])
            # blk_eo_mres: s[0], s[1] even against the oracle; the odd points
            # of s[0], s[1], s[2] unchanged (blk_solv.c header)
            lib.assign_s2sblk(0, n, 0, t, 0)
            lib.assign_s2sblk(0, n, 0, s, 1)
            lib.assign_s2sblk(0, n, 0, t, 2)
            pre = [_block(lat, n, k, bvol) for k in range(3)]
            lib.blk_eo_mres(n, c["mu"], 4)
            post = [_block(lat, n, k, bvol) for k in range(3)]
            s0r, s1r = o.blk_eo_mres(n, mu, 4, s1.copy())
            rel0, _ = site_rel(post[0][:bh], s0r[:bh])
            rel1, _ = site_rel(post[1][:bh], s1r[:bh])
            print("block", n, "blk_eo_mres", rel0, rel1, flush=True)
            assert rel0 < 1e-4 and rel1 < 1e-3
            for k in range(3):
                assert np.array_equal(post[k][bh:], pre[k][bh:]), k
        else:
            lib.assign_s2sblk(0, n, 0, s, 1)
            lib.blk_mres(n, c["mu"], 4)
            s0r, s1r = o.blk_mres(n, mu, 4, s1.copy())
            rel0, _ = site_rel(_block(lat, n, 0, bvol), s0r[:-1])
            rel1, _ = site_rel(_block(lat, n, 1, bvol), s1r[:-1])
            print("block", n, "blk_mres", rel0, rel1, flush=True)
            assert rel0 < 1e-4 and rel1 < 1e-3
    lat.free(s)
    lat.free(t)


@pytest.mark.gpu
def test_block_operators_dble(setup):
    """The FP64 block operators (dirac/Dw_dble.c:1891-2290) on FP64 block
    fields against the oracle's FP64 restatement, on the first and the last
    block: Dw_blk_dble, Dwhat_blk_dble (odd half of the output unchanged),
    Dwee/Dwoo/Dwoe/Dweo_blk_dble with the input's boundary zeros; and the
    round trip assign_sd2sdblk / assign_sdblk2sd (bit-exact)."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.oqh_set_sap_prec(0)
    mu = c["mu"]
    other = np.random.default_rng(9).standard_normal(eta.shape)
    s, t, u = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(s, eta)
    lat.put_sd(t, other)
    nb = o.nblocks()

    def block(n, k, bvol):
        a = np.empty((bvol, 24))
        lib.oqh_export_sdblk(n, k, a.ctypes.data_as(P))
        return a

    for n in (0, nb - 1):
        imb = o.block_imb(n)
        bvol, bh = len(imb), len(imb) // 2
        s1 = np.vstack([eta[imb], np.zeros((1, 24))])
        # block fields in and out of the global field
        lib.assign_sd2sdblk(0, n, 0, s, 1)
        assert np.array_equal(block(n, 1, bvol), eta[imb])
        lat.put_sd(u, other)
        lib.assign_sdblk2sd(0, n, 0, 1, u)
        back = lat.get_sd(u)
        assert np.array_equal(back[imb], eta[imb])
        rest = np.ones(len(back), bool)
        rest[imb] = False
        assert np.array_equal(back[rest], other[rest])
        # Dw_blk_dble
        lib.Dw_blk_dble(0, n, mu, 1, 2)
        ref = o.Dw_blk_dble(n, mu, s1.copy())[:-1]
        rel, absz = site_rel(block(n, 2, bvol), ref)
        print("block", n, "Dw_blk_dble", rel, flush=True)
        assert rel < 1e-13 and absz == 0.0
        if c["isolv"]:
            lib.assign_sd2sdblk(0, n, 0, t, 2)
            before = block(n, 2, bvol)
            lib.Dwhat_blk_dble(0, n, mu, 1, 2)
            out = block(n, 2, bvol)
            ref = o.Dwhat_blk_dble(n, mu, s1.copy())[:bh]
            rel, absz = site_rel(out[:bh], ref)
            print("block", n, "Dwhat_blk_dble", rel, flush=True)
            assert rel < 1e-13 and absz == 0.0
            assert np.array_equal(out[bh:], before[bh:])
        calls = [lambda: lib.Dwee_blk_dble(0, n, mu, 1, 2), lambda: lib.Dwoo_blk_dble(0, n, mu, 1, 2),
                 lambda: lib.Dwoe_blk_dble(0, n, 1, 2), lambda: lib.Dweo_blk_dble(0, n, 1, 2)]
        for op, call in enumerate(calls):
            lib.assign_sd2sdblk(0, n, 0, t, 1)
            lib.assign_sd2sdblk(0, n, 0, s, 2)
            call()
            s_ref, r_ref = o.blk_eo_op_dble(n, op, mu, np.vstack([other[imb], np.zeros((1, 24))]), s1.copy())
            rel, absz = site_rel(block(n, 2, bvol), r_ref[:-1])
            rels, abss = site_rel(block(n, 1, bvol), s_ref[:-1])
            print("block", n, "eo op", op, rel, rels, flush=True)
            assert rel < 1e-13 and absz == 0.0 and rels == 0.0 and abss == 0.0
    lat.free(s)
    lat.free(t)
    lat.free(u)


@pytest.mark.gpu
def test_generic_sap_cycles_fp32(setup):
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.oqh_set_sap_prec(0)
    psi, rho = lat.alloc_s(), lat.alloc_s()
    lat.put_s(psi, np.zeros_like(s32))
    lat.put_s(rho, s32)
    psi_o = np.zeros_like(s32)
    rho_o = s32.copy()
    for _ in range(3):
        lib.sap(c["mu"], c["isolv"], 4, psi, rho)
        o.sap(np.float32(c["mu"]), c["isolv"], 4, psi_o, rho_o)
    rel_p, _ = site_rel(lat.get_s(psi), psi_o)
    g, r = lat.get_s(rho).astype(np.float64), rho_o.astype(np.float64)
    nr = np.linalg.norm(g - r) / np.linalg.norm(s32.astype(np.float64))
    print("fp32 sap psi", rel_p, "rho", nr, flush=True)
    assert rel_p < 1e-3 and nr <= 1e-4
    lat.free(psi)
    lat.free(rho)


@pytest.mark.gpu
def test_generic_sap_cycles_fp64(setup):
    """The FP64 variant (oqh_set_sap_prec(2)) against the oracle's FP64
    Schwarz sweep; sap() converts psi and rho to FP64 and back, so the
    oracle rounds to FP32 after every cycle as well."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.oqh_set_sap_prec(2)
    psi, rho = lat.alloc_s(), lat.alloc_s()
    lat.put_s(psi, np.zeros_like(s32))
    lat.put_s(rho, s32)
    psi_o = np.zeros_like(s32)
    rho_o = s32.copy()
    for cyc in range(3):
        lib.sap(c["mu"], c["isolv"], 4, psi, rho)
        p64, r64 = psi_o.astype(np.float64), rho_o.astype(np.float64)
        o.sap_dble(c["mu"], c["isolv"], 4, p64, r64)
        psi_o, rho_o = p64.astype(np.float32), r64.astype(np.float32)
        rel_p, _ = site_rel(lat.get_s(psi), psi_o)
        rel_r, _ = site_rel(lat.get_s(rho), rho_o)
        print("fp64 sap cycle", cyc, "psi", rel_p, "rho", rel_r, flush=True)
        assert rel_p < 2e-6 and rel_r < 2e-6
    lib.oqh_set_sap_prec(0)
    lat.free(psi)
    lat.free(rho)


@pytest.mark.gpu
def test_generic_sap_gcr(setup):
    """SAP+GCR with the generic block solver: the Krylov count of the FP32
    variant within 1 of the oracle's (modified Gram-Schmidt in the order of
    fgcr.c), and the FP64 true residual of the FP32 and FP64 variants
    recomputed by the oracle."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    e, p = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(e, eta)
    o.sw_term(3)
    rho_o, ifail_o, status_o, _ = o.sap_gcr(16, 256, 0, 1e-10, c["mu"], eta.copy())
    assert ifail_o == 0
    for prec in (0, 2):
        lib.oqh_set_sap_prec(prec)
        lib.oqh_set_gcr_orth(0)
        ifail, status = C.c_int(), C.c_int()
        lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
        lib.oqh_set_gcr_orth(1)
        o.sw_term(3)
        res = np.linalg.norm(eta - o.Dw_dble(c["mu"], lat.get_sd(p).copy())) / np.linalg.norm(eta)
        print("prec", prec, "status", status.value, "oracle", status_o, "true residual", res, flush=True)
        assert ifail.value == 0 and res <= 1.2e-10
        if prec == 0:
            assert abs(status.value - status_o) <= 1
        else:
            assert status.value <= status_o + 2
    lib.oqh_set_sap_prec(0)
    lat.free(e)
    lat.free(p)
