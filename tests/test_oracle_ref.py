"""Oracle versus the reference build (oracle/_ref, compiled from the
reference's own sources by oracle/build_ref.py).  Skipped where the
reference build is absent (it is made only where the reference sources are
present)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import make_oracle, site_rel

CASES = [
    ("8x8x8x8_t8", dict(bc=3, csw=1.0, m0=-0.5, seed=1, mu=0.0, isolv=0), {}),
    ("8x8x8x8_t8", dict(bc=3, csw=1.77, m0=-0.35, seed=3, mu=0.05, isolv=1), {}),
    ("8x8x8x8_t8", dict(bc=0, csw=1.77, m0=-0.3, seed=2, mu=0.0, isolv=0), {"cF": (1.0, 1.0)}),
    ("16x8x8x8_t8", dict(bc=0, csw=1.77, m0=-0.3, seed=5, mu=0.1, isolv=1, cF=(0.95, 0.9)), {}),
    ("8x8x8x8_t4", dict(bc=3, csw=1.2, m0=-0.35, seed=4, mu=0.0, isolv=0), {}),
    # generic block sizes: 8^4 blocks on 16^4, 6^4 blocks on 12^4
    ("16x16x16x16_t8", dict(bc=3, csw=1.77, m0=-0.35, seed=6, mu=0.0, isolv=1, bs=(8, 8, 8, 8)), {}),
    ("16x16x16x16_t8", dict(bc=3, csw=1.77, m0=-0.35, seed=7, mu=0.05, isolv=0, bs=(8, 8, 8, 8)), {}),
    ("12x12x12x12_t4", dict(bc=0, csw=1.77, m0=-0.3, seed=2, mu=0.1, isolv=1, bs=(6, 6, 6, 6)), {}),
    ("12x12x12x12_t4", dict(bc=3, csw=1.77, m0=-0.35, seed=4, mu=0.0, isolv=0, bs=(6, 6, 6, 6)), {}),
]
GEOM = {"8x8x8x8_t8": ((8, 8, 8, 8), (8, 8, 8, 8)), "16x8x8x8_t8": ((16, 8, 8, 8), (8, 8, 8, 8)),
        "8x8x8x8_t4": ((8, 8, 8, 8), (4, 4, 4, 4)),
        "16x16x16x16_t8": ((16, 16, 16, 16), (8, 8, 8, 16)),
        "12x12x12x12_t4": ((12, 12, 12, 12), (6, 6, 12, 12))}


@pytest.fixture(scope="module", params=range(len(CASES)))
def case(request, tmp_path_factory):
    tag, kw, _ = CASES[request.param]
    if po.ref_exe(tag) is None:
        pytest.skip("reference build not present")
    out = str(tmp_path_factory.mktemp("ref_%s_%d" % (tag, request.param)))
    args = dict(kw)
    cF = args.pop("cF", (1.0, 1.0))
    bs = args.pop("bs", (4, 4, 4, 4))
    ld = po.ref_dump(tag, out, bc=args["bc"], csw=args["csw"], m0=args["m0"], seed=args["seed"],
                     mu=args["mu"], isolv=args["isolv"], cF=cF, bs=bs, nmr=4, ncy=5,
                     nkv=16, nmx=256, res=1e-10)
    L, T = GEOM[tag]
    o, eta, s32 = make_oracle(L, T, bc=args["bc"], cF=cF, m0=args["m0"], csw=args["csw"],
                              seed=args["seed"])
    info = json.load(open(os.path.join(out, "info.json")))
    return dict(ld=ld, o=o, eta=eta, s32=s32, V=o.V, mu=args["mu"], isolv=args["isolv"], info=info, bs=bs)


def test_fields(case):
    ld, o, V = case["ld"], case["o"], case["V"]
    assert np.array_equal(o.ipt(), ld("ipt.i32", np.int32))
    assert np.abs(o.ud() - ld("ud.f64", np.float64, (4 * V, 18))).max() < 1e-15
    assert np.abs(case["eta"] - ld("eta.f64", np.float64, (V, 24))).max() < 1e-15
    assert np.array_equal(case["s32"], ld("s32.f32", np.float32, (V, 24)))
    ft = o.ftensor()
    for n in range(6):
        assert np.abs(ft[n] - ld("ft%d.f64" % n, np.float64, (V, 9))).max() < 1e-14


def test_sw_and_dirac(case):
    ld, o, V, mu, eta = case["ld"], case["o"], case["V"], case["mu"], case["eta"]
    o.sw_term(3)
    rel, _ = site_rel(o.swd(), ld("swd.f64", np.float64, (2 * V, 36)))
    assert rel < 1e-14
    for name in ("Dw_dble", "Dwee_dble", "Dwoo_dble", "Dwoe_dble", "Dweo_dble"):
        s = eta.copy()
        r = ld(name + "_r0.f64", np.float64, (V, 24)).copy()
        if name in ("Dwoe_dble", "Dweo_dble"):
            getattr(o, name)(s, r)
        else:
            getattr(o, name)(mu, s, r)
        rel, absz = site_rel(r, ld(name + ".f64", np.float64, (V, 24)))
        assert rel < 1e-14 and absz == 0.0, name
        sin = ld(name + "_in.f64", np.float64, (V, 24))
        assert np.abs(s - sin).max() < 1e-14, name
        assert np.array_equal(np.all(s == 0, axis=1), np.all(sin == 0, axis=1)), name
    o.assign_ud2u()
    o.assign_swd2sw()
    s = case["s32"].copy()
    r = o.Dw(np.float32(mu), s)
    rel, absz = site_rel(r, ld("Dw.f32", np.float32, (V, 24)))
    assert rel < 1e-6 and absz == 0.0
    assert o.sw_term(2) == case["info"]["ifail_inv"]
    rel, _ = site_rel(o.swd(), ld("swd_oinv.f64", np.float64, (2 * V, 36)))
    assert rel < 1e-14
    r = o.Dwhat_dble(mu, eta.copy())
    rel, _ = site_rel(r[:V // 2], ld("Dwhat_dble.f64", np.float64, (V, 24))[:V // 2])
    assert rel < 1e-14
    o.assign_swd2sw()
    r = o.Dwhat(np.float32(mu), case["s32"].copy())
    rel, _ = site_rel(r[:V // 2], ld("Dwhat.f32", np.float32, (V, 24))[:V // 2])
    assert rel < 1e-6


def test_qsum(case):
    ld, V = case["ld"], case["V"]
    q = ld("qsum.f64", np.float64)
    eta = ld("eta.f64", np.float64, (V, 24))
    # same summation order and double-double steps: bit-identical (hi, lo)
    n = po.norm_square_dble(eta)
    assert n[0] == q[4] and n[1] == q[5]
    qre, qim = po.spinor_prod_dble(eta, eta)
    assert qre[0] == q[0] and qre[1] == q[1]


def test_sap_and_solver(case):
    ld, o, V, mu = case["ld"], case["o"], case["V"], case["mu"]
    isolv = case["isolv"]
    if not case["info"]["sap"]:
        pytest.skip("block grid not allowed for this thread layout")
    o.sw_term(3)
    o.set_sap(case["bs"], isolv, 4, 5)
    assert o.sap_prepare() == case["info"]["ifail_bgr"]
    psi = np.zeros((V, 24), np.float32)
    rho = case["s32"].copy()
    for _ in range(5):
        o.sap(np.float32(mu), isolv, 4, psi, rho)
    rel, _ = site_rel(psi, ld("sap_psi.f32", np.float32, (V, 24)))
    assert rel < 1e-4
    rel, _ = site_rel(rho, ld("sap_rho.f32", np.float32, (V, 24)))
    assert rel < 1e-3
    if isolv == 0:
        imb = o.block_imb(0)
        assert np.array_equal(imb, ld("blk0_imb.i32", np.int32))
        s1 = np.vstack([case["s32"][imb], np.zeros((1, 24), np.float32)])
        r = o.Dw_blk(0, np.float32(mu), s1.copy())
        rel, _ = site_rel(r[:-1], ld("blk0_Dw.f32", np.float32, (len(imb), 24)))
        assert rel < 1e-6
        s0, s1b = o.blk_mres(0, np.float32(mu), 4, s1.copy())
        rel, _ = site_rel(s0[:-1], ld("blk0_s0.f32", np.float32, (len(imb), 24)))
        assert rel < 1e-5
    rho, ifail, status, psi = o.sap_gcr(16, 256, 0, 1e-10, mu, case["eta"].copy())
    info = case["info"]
    assert ifail == info["gcr_ifail"] == 0
    assert abs(status - info["gcr_status"]) <= 1
    rel, _ = site_rel(psi, ld("gcr_psi.f64", np.float64, (V, 24)))
    assert rel < 1e-8


def _unphase(o, ud, L):
    """Undo set_ud_phase of periodic bc exactly: the time-like links at
    x0 = N0-1 (slot 0 of odd points there, slot 1 of odd points at x0 = 0)."""
    V = o.V
    h = V // 2
    lex = np.empty(V, np.int64)
    lex[o.ipt()] = np.arange(V)
    t = lex[h:] // (L[1] * L[2] * L[3])
    ud = ud.reshape(h, 8, 18).copy()
    ud[t == L[0] - 1, 0] *= -1.0
    ud[t == 0, 1] *= -1.0
    return ud.reshape(4 * V, 18)


@pytest.mark.parametrize("tag,bc", [("8x8x8x8_t8", 3), ("8x8x8x8_t8", 0), ("8x8x8x8_t4", 3),
                                    ("16x8x8x8_t8", 0)])
def test_flow_matches_reference(tag, bc, tmp_path):
    """plaq_frc and fwd_euler of the oracle on the reference's own random
    field agree with the reference build to the last bit."""
    if po.ref_exe(tag) is None:
        pytest.skip("reference build not present")
    L, T = GEOM[tag]
    l0 = po.ref_dump(tag, str(tmp_path / "a"), bc=bc, seed=2)
    l1 = po.ref_dump(tag, str(tmp_path / "b"), bc=bc, seed=2, flow=(4, 0.02))
    o = po.Oracle(L, T, bc=bc)
    V = o.V
    ud = l0("ud.f64", np.float64, (4 * V, 18))
    if bc == 3:
        ud = _unphase(o, ud, L)
    o.set_ud(ud)
    assert np.array_equal(o.plaq_frc(), l1("frc0.f64", np.float64, (4 * V, 8)))
    o.fwd_euler(4, 0.02)
    assert np.array_equal(o.ud(), l1("ud_flow.f64", np.float64, (4 * V, 18)))


def test_reference_deflation_fixture(tmp_path):
    """The reference build's SAP+GCR and deflated SAP+GCR counts recorded in
    tests/golden/ref_16x8x8x8_dfl.json (used by the GPU deflation test) are
    those the current reference build gives, and both solves reach the
    requested residual."""
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(os.path.dirname(here), "oracle", "_ref", "ref_16x8x8x8_t8")
    if not os.path.exists(exe):
        pytest.skip("reference build not present")
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_golden
    ref = json.load(open(os.path.join(here, "golden", "ref_16x8x8x8_dfl.json")))
    assert ref["case"] == json.loads(json.dumps(make_golden.DFL_CASE))
    keep = make_golden.HERE
    try:
        make_golden.HERE = str(tmp_path)
        make_golden.dfl_counts()
        new = json.load(open(os.path.join(str(tmp_path), "ref_16x8x8x8_dfl.json")))
    finally:
        make_golden.HERE = keep
    for mode, keys in (("gcr", ("gcr_status", "gcr_ifail")), ("dfl", ("modes_status", "dfl_status", "dfl_ifail"))):
        for k in keys:
            assert new[mode][k] == ref[mode][k], (mode, k)
    assert ref["gcr"]["true_residual"] <= 1.01e-10 and ref["dfl"]["true_residual"] <= 1.01e-10
