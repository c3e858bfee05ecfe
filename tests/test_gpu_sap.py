"""GPU parity of the SAP block solver, the Schwarz sweep and SAP+GCR.

FP32 block quantities are compared with the oracle to a relative 1e-5 after
one operator application and to 1e-3 after several Schwarz cycles (rounding
differences of the single-precision recursion grow with the number of steps).
Solutions of SAP+GCR are judged like the reference's devel/sap/check3.c: the
double-precision residual |eta - Dw_dble psi| / |eta|, recomputed by the
oracle on the exported solution, must reach the requested tolerance.
"""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import make_oracle, site_rel

P = C.c_void_p

CFGS = [
    dict(name="8^4-periodic-isolv0", L=(8, 8, 8, 8), bc=3, m0=-0.5, csw=1.0, seed=1, isolv=0, mu=0.0),
    dict(name="8^4-periodic-isolv1-mu", L=(8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77, seed=3, isolv=1, mu=0.05),
    dict(name="8^4-open-isolv0", L=(8, 8, 8, 8), bc=0, m0=-0.3, csw=1.77, seed=2, isolv=0, mu=0.0),
    dict(name="16x8^3-open-isolv1", L=(16, 8, 8, 8), bc=0, m0=-0.3, csw=1.77, seed=5, isolv=1, mu=0.1),
]


def _sig(lib):
    lib.sap.argtypes = [C.c_float, C.c_int, C.c_int, P, P]
    lib.sap_gcr.restype = C.c_double
    lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
    lib.assign_s2sblk.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int]
    lib.Dw_blk.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
    lib.blk_mres.argtypes = [C.c_int, C.c_float, C.c_int]
    lib.oqh_export_sblk.argtypes = [C.c_int, C.c_int, P]
    for nm in ("Dwee_blk", "Dwoo_blk"):
        getattr(lib, nm).argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
    for nm in ("Dwoe_blk", "Dweo_blk"):
        getattr(lib, nm).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.oqh_get_imb.argtypes = [C.c_int, P]


@pytest.fixture(params=CFGS, ids=lambda c: c["name"])
def setup(request):
    c = request.param
    print("config", c, flush=True)
    o, eta, s32 = make_oracle(c["L"], None, bc=c["bc"], m0=c["m0"], csw=c["csw"], seed=c["seed"])
    lat = pyoqh.Lattice(c["L"], None, bc=c["bc"], m0=c["m0"], csw=c["csw"])
    _sig(lat.lib)
    lat.import_ud(o.ud())
    bs = np.array([4, 4, 4, 4], np.int32)
    lat.lib.oqh_set_sap_parms(bs.ctypes.data_as(P), c["isolv"], 4, 5)
    o.set_sap((4, 4, 4, 4), c["isolv"], 4, 5)
    o.sw_term(3)
    yield c, o, eta, s32, lat
    lat.close()


@pytest.mark.gpu
def test_block_operator_and_mres(setup):
    """Dw_blk and blk_mres (Dw.c:1475-1582, blk_solv.c:1062-1186) against the
    oracle on the first and the last block; with isolv=1 on the grid that
    SAP prepares, whose odd block SW term is inverted, in both."""
    c, o, eta, s32, lat = setup
    assert o.sap_prepare() == 0
    s = lat.alloc_s()
    lat.put_s(s, s32)
    for n in (0, o.nblocks() - 1):
        imb = np.empty(256, np.int32)
        lat.lib.oqh_get_imb(n, imb.ctypes.data_as(P))
        assert np.array_equal(imb, o.block_imb(n))
        lat.lib.assign_s2sblk(0, n, 0, s, 1)
        lat.lib.Dw_blk(0, n, c["mu"], 1, 2)
        got = np.empty((256, 24), np.float32)
        lat.lib.oqh_export_sblk(n, 2, got.ctypes.data_as(P))
        s1 = np.vstack([s32[imb], np.zeros((1, 24), np.float32)])
        ref = o.Dw_blk(n, np.float32(c["mu"]), s1.copy())[:-1]
        rel, absz = site_rel(got, ref)
        assert rel < 1e-5 and absz == 0.0
        lat.lib.blk_mres(n, c["mu"], 4)
        s0g = np.empty((256, 24), np.float32)
        s1g = np.empty((256, 24), np.float32)
        lat.lib.oqh_export_sblk(n, 0, s0g.ctypes.data_as(P))
        lat.lib.oqh_export_sblk(n, 1, s1g.ctypes.data_as(P))
        s0r, s1r = o.blk_mres(n, np.float32(c["mu"]), 4, s1.copy())
        rel0, _ = site_rel(s0g, s0r[:-1])
        rel1, _ = site_rel(s1g, s1r[:-1])
        print("block", n, "Dw", rel, "mres", rel0, rel1)
        assert rel0 < 1e-4 and rel1 < 1e-3
    lat.free(s)


@pytest.mark.gpu
def test_block_eo_operators(setup):
    """Dwee_blk, Dwoo_blk, Dwoe_blk, Dweo_blk (Dw.c:1584-1826) against the
    oracle: output field and the input's boundary zeros, on the first and the
    last block; with isolv=1 the odd block SW term is inverted, as SAP
    leaves it.  Dwee_blk with k == l equals the out-of-place result."""
    c, o, eta, s32, lat = setup
    assert o.sap_prepare() == 0
    rng = np.random.default_rng(7)
    other = rng.standard_normal(s32.shape).astype(np.float32)
    s, t = lat.alloc_s(), lat.alloc_s()
    lat.put_s(s, s32)
    lat.put_s(t, other)
    mu = np.float32(c["mu"])

    def block(n, k):
        a = np.empty((256, 24), np.float32)
        lat.lib.oqh_export_sblk(n, k, a.ctypes.data_as(P))
        return a

    calls = [lambda n: lat.lib.Dwee_blk(0, n, c["mu"], 1, 2), lambda n: lat.lib.Dwoo_blk(0, n, c["mu"], 1, 2),
             lambda n: lat.lib.Dwoe_blk(0, n, 1, 2), lambda n: lat.lib.Dweo_blk(0, n, 1, 2)]
    for n in (0, o.nblocks() - 1):
        imb = o.block_imb(n)
        for op, call in enumerate(calls):
            # input with values on the open boundary (the operators zero them)
            lat.lib.assign_s2sblk(0, n, 0, t, 1)
            lat.lib.assign_s2sblk(0, n, 0, s, 2)
            call(n)
            s_ref, r_ref = o.blk_eo_op(n, op, mu, other[imb], s32[imb])
            rel, absz = site_rel(block(n, 2), r_ref)
            rels, abss = site_rel(block(n, 1), s_ref)
            print("block", n, "op", op, rel, rels)
            assert rel < 1e-5 and absz == 0.0 and rels == 0.0 and abss == 0.0
            if op == 0:
                out = block(n, 2)
                lat.lib.assign_s2sblk(0, n, 0, t, 1)
                lat.lib.Dwee_blk(0, n, c["mu"], 1, 1)
                inpl = block(n, 1)
                assert np.array_equal(inpl[:128], out[:128]) and np.array_equal(inpl[128:], s_ref[128:])
    lat.free(s)
    lat.free(t)


@pytest.mark.gpu
def test_dw_bnd(setup):
    """Dw_bnd (Dw_bnd.c:801-1011) against the oracle's global Dw: the Weyl
    field of face ifc, expanded to the Dirac spinor psi with
    (1 + gamma_mu) psi = 0 on the face -mu and (1 - gamma_mu) psi = 0 on
    +mu, is the value at the exterior neighbour of Dw applied to a field
    that equals the block's field on the block and vanishes elsewhere."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.Dw_bnd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.oqh_export_bnd.argtypes = [C.c_int, C.c_int, C.c_int, P]
    assert o.sap_prepare() == 0
    o.assign_ud2u()
    o.assign_swd2sw()
    iup, idn = o.iupdn()
    gam = np.zeros((4, 4, 4), complex)
    gam[0] = [[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]]
    gam[1] = [[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]]
    gam[2] = [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]]
    gam[3] = [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]]
    # a field that does not vanish on the open boundary: the time faces zero
    # it there before the other faces read it, as Dw zeroes its input
    src = np.random.default_rng(11).standard_normal(s32.shape).astype(np.float32)
    s = lat.alloc_s()
    lat.put_s(s, src)
    nb = o.nblocks()
    for n in (0, nb // 2 + 1, nb - 1):
        imb = o.block_imb(n)
        bvol, bh = len(imb), len(imb) // 2
        inblk = set(imb.tolist())
        lat.lib.assign_s2sblk(0, n, 0, s, 1)
        lib.Dw_bnd(0, n, 1, 0)
        F = np.zeros_like(s32)
        F[imb] = src[imb]
        D = o.Dw(np.float32(0.0), F.copy()).astype(np.float64)
        D = (D[:, 0::2] + 1j * D[:, 1::2]).reshape(-1, 4, 3)
        order = list(range(bh, bvol)) + list(range(bh))
        for ifc in range(8):
            mu, up = ifc // 2, ifc & 1
            nbr = iup if up else idn
            pts = [iy for iy in order if nbr[imb[iy], mu] not in inblk]
            assert len(pts) == bvol // 4
            w = np.empty((len(pts), 12), np.float32)
            lib.oqh_export_bnd(n, ifc, 0, w.ctypes.data_as(P))
            u = (w[:, 0::2] + 1j * w[:, 1::2]).astype(complex).reshape(-1, 2, 3)
            e = gam[mu][:2, 2:]
            low = np.einsum("ij,njc->nic", np.conj(e.T), u) * (1.0 if up else -1.0)
            psi = np.concatenate([u, low], axis=1)
            ref = D[[nbr[imb[iy], mu] for iy in pts]]
            err = np.abs(psi - ref).max(axis=(1, 2))
            scale = np.abs(ref).max(axis=(1, 2))
            print("block", n, "face", ifc, float((err / np.maximum(scale, 1e-30)).max()))
            assert (err <= 1e-5 * np.maximum(scale, 1.0)).all()
            assert np.array_equal(err[scale == 0], np.zeros((scale == 0).sum()))
    lat.free(s)


@pytest.mark.gpu
def test_sap_cycles(setup):
    c, o, eta, s32, lat = setup
    assert o.sap_prepare() == 0
    psi, rho = lat.alloc_s(), lat.alloc_s()
    lat.put_s(psi, np.zeros_like(s32))
    lat.put_s(rho, s32)
    psi_o = np.zeros_like(s32)
    rho_o = s32.copy()
    for cyc in range(3):
        lat.lib.sap(c["mu"], c["isolv"], 4, psi, rho)
        o.sap(np.float32(c["mu"]), c["isolv"], 4, psi_o, rho_o)
    rel_p, _ = site_rel(lat.get_s(psi), psi_o)
    rel_r, _ = site_rel(lat.get_s(rho), rho_o)
    print("sap psi", rel_p, "rho", rel_r)
    assert rel_p < 1e-3
    # the residual itself becomes small; compare in norm relative to |eta|
    g, r = lat.get_s(rho).astype(np.float64), rho_o.astype(np.float64)
    assert np.linalg.norm(g - r) <= 1e-4 * np.linalg.norm(s32.astype(np.float64))
    # SAP invariant (devel/sap/check2.c): rho = eta - D psi
    o.assign_ud2u()
    o.assign_swd2sw()
    d = o.Dw(np.float32(c["mu"]), lat.get_s(psi).copy())
    inv = np.linalg.norm(s32.astype(np.float64) - d - g) / np.linalg.norm(s32.astype(np.float64))
    print("invariant", inv)
    assert inv < 1e-5


@pytest.mark.gpu
def test_sap_gcr_solves(setup):
    c, o, eta, s32, lat = setup
    V = lat.V
    e, p = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(e, eta)
    ifail, status = C.c_int(), C.c_int()
    # modified Gram-Schmidt term by term, the order of fgcr.c, for the
    # comparison of the Krylov vector count with the oracle
    lat.lib.oqh_set_gcr_orth(0)
    rho = lat.lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
    lat.lib.oqh_set_gcr_orth(1)
    psi = lat.get_sd(p)
    o.sw_term(3)
    res = np.linalg.norm(eta - o.Dw_dble(c["mu"], psi.copy())) / np.linalg.norm(eta)
    rho_o, ifail_o, status_o, psi_o = o.sap_gcr(16, 256, 0, 1e-10, c["mu"], eta.copy())
    print("gcr", ifail.value, status.value, rho, "true residual", res, "oracle status", status_o)
    assert ifail.value == 0 and ifail_o == 0
    assert res <= 1.2e-10
    assert abs(status.value - status_o) <= 1
    assert np.array_equal(lat.get_sd(e), eta), "source must be preserved"
    # two-pass form of modified Gram-Schmidt (the default): same tolerance,
    # nearly the same count
    p2 = lat.alloc_sd()
    st2 = C.c_int()
    lat.lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p2, C.byref(ifail), C.byref(st2))
    res2 = np.linalg.norm(eta - o.Dw_dble(c["mu"], lat.get_sd(p2).copy())) / np.linalg.norm(eta)
    print("two-pass GS status", st2.value, "true residual", res2)
    assert ifail.value == 0 and res2 <= 1.2e-10 and abs(st2.value - status_o) <= 2
    lat.free(p2)
    # psi == eta (in place), as allowed by sap_gcr
    rho2 = lat.lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, e, C.byref(ifail), C.byref(status))
    assert ifail.value == 0
    rel, _ = site_rel(lat.get_sd(e), psi)
    assert rel < 1e-8


@pytest.mark.gpu
def test_sap_gcr_failure_modes(setup):
    c, o, eta, s32, lat = setup
    e, p = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(e, eta)
    ifail, status = C.c_int(), C.c_int()
    # too few Krylov vectors: ifail = -1, status = nmx (sap_gcr.c:223-228)
    lat.lib.sap_gcr(4, 2, 0, 1e-12, c["mu"], e, p, C.byref(ifail), C.byref(status))
    assert ifail.value == -1 and status.value == 2
    # zero source: psi = 0
    lat.put_sd(e, np.zeros_like(eta))
    lat.put_sd(p, eta)
    rho = lat.lib.sap_gcr(16, 64, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
    assert rho == 0.0 and ifail.value == 0 and not lat.get_sd(p).any()


@pytest.mark.gpu
def test_fp16_storage_variant(setup):
    """The FP16-storage Schwarz cycle (links, SW matrices and the spinor
    fields psi, rho, s0g in FP16, the spinors with per-site scales; FP32
    arithmetic) approximates the FP32 cycle to the FP16 rounding level, and
    SAP+GCR with it, where the whole preconditioner runs on FP16 fields,
    still reaches the requested double-precision residual (the outer FGCR
    corrects the preconditioner's rounding)."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    out = {}
    for prec in (0, 1):
        lib.oqh_set_sap_prec(prec)
        psi, rho = lat.alloc_s(), lat.alloc_s()
        lat.put_s(psi, np.zeros_like(s32))
        lat.put_s(rho, s32)
        lib.sap(c["mu"], c["isolv"], 4, psi, rho)
        out[prec] = lat.get_s(psi).astype(np.float64)
        lat.free(psi)
        lat.free(rho)
    rel = np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])
    print("fp16 vs fp32 SAP cycle", rel)
    assert rel < 5e-3
    e, p = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(e, eta)
    ifail, status = C.c_int(), C.c_int()
    lat.lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
    o.sw_term(3)
    res = np.linalg.norm(eta - o.Dw_dble(c["mu"], lat.get_sd(p).copy())) / np.linalg.norm(eta)
    print("fp16 SAP+GCR status", status.value, "true residual", res)
    assert ifail.value == 0 and res <= 1.2e-10
    lib.oqh_set_sap_prec(0)


@pytest.mark.gpu
def test_fp16_spinor_format(setup):
    """With nmr=0 a Schwarz cycle leaves psi as it is (with isolv=1 at the
    even points: the odd ones receive A_oo^-1 rho), and with isolv=0 also
    rho (apart from the open-bc boundary points), so in the FP16 variant it
    returns them after one trip through the FP16 format: per site, every real
    component is within 2^-12 of the site's largest one (FP16 rounding of
    values scaled into [-1, 1]), up to the FP32 rounding of the scaling
    (a few units of 2^-24 of the site maximum)."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    rng = np.random.default_rng(11)
    # components spread over six decades within a site
    psi0 = (s32 * np.exp(rng.uniform(-7, 0, s32.shape))).astype(np.float32)
    out = {}
    for prec in (0, 1):
        lib.oqh_set_sap_prec(prec)
        psi, rho = lat.alloc_s(), lat.alloc_s()
        lat.put_s(psi, psi0)
        lat.put_s(rho, s32)
        lib.sap(c["mu"], c["isolv"], 0, psi, rho)
        out[prec] = (lat.get_s(psi), lat.get_s(rho))
        lat.free(psi)
        lat.free(rho)
    lib.oqh_set_sap_prec(0)
    n = lat.V // 2 if c["isolv"] else lat.V
    for k in ((0,) if c["isolv"] else (0, 1)):
        ref = out[0][k][:n].astype(np.float64)
        m = np.abs(ref).max(axis=1, keepdims=True)
        dev = np.abs(out[1][k][:n] - ref)
        print("field", k, "max deviation / site max", (dev / np.maximum(m, 1e-300)).max())
        assert np.all(dev <= (2.0 ** -12 + 2.0 ** -22) * m)
    assert np.array_equal(out[0][0][:n], psi0[:n])


@pytest.mark.gpu
def test_fp16_preconditioner_graph(setup):
    """The FP16 Schwarz cycles of sap_gcr replayed as a HIP graph (once per
    Krylov vector, with the solver's own kernels and host synchronisations
    in between) give bitwise the solution of the kernel-by-kernel launches."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    lib.oqh_set_sap_prec(1)
    e = lat.alloc_sd()
    lat.put_sd(e, eta)
    out = {}
    for graph in (0, 1, 1):
        lib.oqh_set_sap_graph(graph)
        p = lat.alloc_sd()
        ifail, status = C.c_int(), C.c_int()
        rho = lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
        assert ifail.value == 0
        out.setdefault(graph, []).append((lat.get_sd(p), status.value, rho))
        lat.free(p)
    lib.oqh_set_sap_graph(1)
    lib.oqh_set_sap_prec(0)
    ref = out[0][0]
    for got in out[1]:
        print("graph status", got[1], "eager status", ref[1])
        assert got[1] == ref[1] and got[2] == ref[2]
        assert np.array_equal(got[0], ref[0])


@pytest.mark.gpu
def test_fp64_variant(setup):
    """The FP64 Schwarz cycle (generic block solver on double fields, FP64
    links and SW matrices) agrees with the FP32 cycle to single-precision
    rounding, and SAP+GCR with it reaches the requested residual."""
    c, o, eta, s32, lat = setup
    lib = lat.lib
    out = {}
    for prec in (0, 2):
        lib.oqh_set_sap_prec(prec)
        psi, rho = lat.alloc_s(), lat.alloc_s()
        lat.put_s(psi, np.zeros_like(s32))
        lat.put_s(rho, s32)
        lib.sap(c["mu"], c["isolv"], 4, psi, rho)
        out[prec] = (lat.get_s(psi).astype(np.float64), lat.get_s(rho).astype(np.float64))
        lat.free(psi)
        lat.free(rho)
    rel = np.linalg.norm(out[2][0] - out[0][0]) / np.linalg.norm(out[0][0])
    print("fp64 vs fp32 SAP cycle", rel)
    assert rel < 1e-4
    # SAP invariant rho = eta - D psi, up to the FP32 rounding of the exported fields
    o.assign_ud2u()
    o.assign_swd2sw()
    d = o.Dw(np.float32(c["mu"]), out[2][0].astype(np.float32))
    e64 = s32.astype(np.float64)
    inv = np.linalg.norm(e64 - d - out[2][1]) / np.linalg.norm(e64)
    print("fp64 invariant", inv)
    assert inv < 1e-5
    e, p = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(e, eta)
    ifail, status = C.c_int(), C.c_int()
    lib.sap_gcr(16, 256, 0, 1e-10, c["mu"], e, p, C.byref(ifail), C.byref(status))
    o.sw_term(3)
    res = np.linalg.norm(eta - o.Dw_dble(c["mu"], lat.get_sd(p).copy())) / np.linalg.norm(eta)
    print("fp64 SAP+GCR status", status.value, "true residual", res)
    assert ifail.value == 0 and res <= 1.2e-10
    lib.oqh_set_sap_prec(0)
