"""Pins of the oracle's block operators, block MinRes and Schwarz sweep for
generic block sizes, in both precisions, by identities with the oracle's
global operators (which the reference build pins, test_oracle_ref.py):

* Dw_blk = Dw applied to the field that equals the block field on the block
  and vanishes elsewhere, read at the block points (Dirichlet exterior;
  devel/dirac/check7-8);
* Dwhat_blk = A_ee + i mu gamma5 - D_eo A_oo^-1 D_oe with the hopping terms
  restricted to the block, built from the global even/odd pieces
  (Dwoe_dble, Dwoo_dble on the inverted odd SW term, Dwee_dble, Dweo_dble);
* MinRes: the residual s1 = b - D s0 holds after every number of steps, one
  step is the closed form s0 = z b, z = (D b, b) / |D b|^2, and |s1| does
  not increase with the number of steps;
* the Schwarz sweep keeps rho = eta - D psi (devel/sap/check2.c), and the
  FP32 and FP64 sweeps agree to single-precision rounding.
A plausible slip (a wrong sign of a hopping term, a missing Dirichlet cut,
the wrong parity of the inverted SW term) fails one of these."""
import numpy as np
import pytest

import pyoracle as po
from helpers import make_oracle, site_rel

CFGS = [
    dict(L=(8, 8, 8, 8), T=(8, 8, 8, 8), bs=(4, 4, 4, 4), bc=3, m0=-0.35, csw=1.77, seed=3, mu=0.05),
    dict(L=(8, 8, 8, 8), T=(4, 4, 4, 4), bs=(4, 4, 4, 4), bc=0, m0=-0.3, csw=1.77, seed=2, mu=0.0),
    dict(L=(12, 12, 12, 12), T=(6, 6, 12, 12), bs=(6, 6, 6, 6), bc=0, m0=-0.3, csw=1.77, seed=2, mu=0.1),
    dict(L=(16, 8, 8, 8), T=(8, 8, 8, 8), bs=(8, 4, 4, 4), bc=3, m0=-0.35, csw=1.77, seed=5, mu=0.0),
]


def _ids(c):
    return "L%s-bs%s-bc%d" % ("x".join(map(str, c["L"])), c["bs"][0], c["bc"])


@pytest.fixture(params=CFGS, ids=_ids)
def cfg(request):
    c = request.param
    o, eta, s32 = make_oracle(c["L"], c["T"], bc=c["bc"], m0=c["m0"], csw=c["csw"], seed=c["seed"])
    return c, o, eta, s32


def _zero_ext(V, imb, blk):
    f = np.zeros((V, blk.shape[1]), blk.dtype)
    f[imb] = blk[:len(imb)]
    return f


@pytest.mark.parametrize("dble", [True, False])
def test_Dw_blk_is_Dirichlet_restriction(cfg, dble):
    c, o, eta, s32 = cfg
    o.sw_term(3)
    o.set_sap(c["bs"], 0, 4, 5)
    assert o.sap_prepare() == 0
    src = eta if dble else s32
    for n in (0, o.nblocks() // 2, o.nblocks() - 1):
        imb = o.block_imb(n)
        b = np.vstack([src[imb], np.zeros((1, 24), src.dtype)])
        if dble:
            got = o.Dw_blk_dble(n, c["mu"], b.copy())[:-1]
            full = o.Dw_dble(c["mu"], _zero_ext(o.V, imb, b))[imb]
            tol = 1e-14
        else:
            got = o.Dw_blk(n, np.float32(c["mu"]), b.copy())[:-1]
            o.assign_ud2u()
            o.assign_swd2sw()
            full = o.Dw(np.float32(c["mu"]), _zero_ext(o.V, imb, b))[imb]
            tol = 1e-6
        rel, absz = site_rel(got, full)
        assert rel < tol and absz == 0.0, (n, rel)
        # the hopping terms across the block faces are absent: the global
        # operator on the unrestricted field differs on the face points
        if n == 0:
            fu = (o.Dw_dble(c["mu"], src.copy()) if dble else o.Dw(np.float32(c["mu"]), src.copy()))[imb]
            assert site_rel(got, fu)[0] > 1e-3


def test_Dwhat_blk_dble_is_block_schur_complement(cfg):
    c, o, eta, s32 = cfg
    mu = c["mu"]
    o.sw_term(3)
    o.set_sap(c["bs"], 1, 4, 5)
    assert o.sap_prepare() == 0
    V, h = o.V, o.V // 2
    for n in (0, o.nblocks() - 1):
        imb = o.block_imb(n)
        bh = len(imb) // 2
        b = np.vstack([eta[imb], np.zeros((1, 24))])
        got = o.Dwhat_blk_dble(n, mu, b.copy())[:bh]
        gotf = o.Dwhat_blk(n, np.float32(mu), b.astype(np.float32))[:bh]
        # global pieces with the odd SW term inverted
        o.sw_term(2)
        u = np.zeros((V, 24))
        u[imb[:bh]] = eta[imb[:bh]]
        t = np.zeros((V, 24))
        o.Dwoe_dble(u.copy(), t)
        ti = np.zeros((V, 24))
        o.Dwoo_dble(0.0, t.copy(), ti)
        tr = np.zeros((V, 24))
        tr[imb[bh:]] = ti[imb[bh:]]       # D_oo^-1 D_oe restricted to the block
        r = np.zeros((V, 24))
        o.Dwee_dble(mu, u.copy(), r)
        o.Dweo_dble(tr, r)               # r_e -= D_eo (.)
        o.sw_term(3)
        ref = r[imb[:bh]]
        rel, absz = site_rel(got, ref)
        assert rel < 1e-13 and absz == 0.0, (n, rel)
        assert site_rel(gotf, ref)[0] < 1e-5


@pytest.mark.parametrize("dble", [True, False])
@pytest.mark.parametrize("eo", [0, 1])
def test_mres_identities(cfg, dble, eo):
    c, o, eta, s32 = cfg
    mu = c["mu"]
    o.sw_term(3)
    o.set_sap(c["bs"], eo, 4, 5)
    assert o.sap_prepare() == 0
    n = o.nblocks() - 1
    imb = o.block_imb(n)
    nv = len(imb) // 2 if eo else len(imb)
    src = eta if dble else s32
    b = np.vstack([src[imb], np.zeros((1, 24), src.dtype)])
    if dble:
        op = (lambda s: o.Dwhat_blk_dble(n, mu, s)) if eo else (lambda s: o.Dw_blk_dble(n, mu, s))
        mres = o.blk_eo_mres_dble if eo else o.blk_mres_dble
        m = mu
        tol = 1e-12
    else:
        m = np.float32(mu)
        op = (lambda s: o.Dwhat_blk(n, m, s)) if eo else (lambda s: o.Dw_blk(n, m, s))
        mres = o.blk_eo_mres if eo else o.blk_mres
        tol = 1e-5
    b64 = b.astype(np.float64)
    last = np.inf
    for nmr in (1, 2, 4, 8):
        s0, s1 = mres(n, m, nmr, b.copy())
        d = op(np.vstack([s0[:-1], np.zeros((1, 24), s0.dtype)])).astype(np.float64)
        res = np.linalg.norm(b64[:nv] - d[:nv] - s1[:nv]) / np.linalg.norm(b64[:nv])
        assert res < tol, (nmr, res)
        nr = np.linalg.norm(s1[:nv].astype(np.float64))
        assert nr <= last * (1 + 1e-6)
        last = nr
        if nmr == 1:
            db = op(b.copy()).astype(np.float64)[:nv]
            zb = b64[:nv, 0::2] + 1j * b64[:nv, 1::2]
            zd = db[:, 0::2] + 1j * db[:, 1::2]
            z = np.vdot(zd.ravel(), zb.ravel()) / np.vdot(zd.ravel(), zd.ravel()).real
            ref = z * zb
            s0z = s0[:nv, 0::2] + 1j * s0[:nv, 1::2]
            assert np.abs(s0z - ref).max() <= (1e-13 if dble else 1e-5) * np.abs(ref).max()


@pytest.mark.parametrize("isolv", [0, 1])
def test_sap_dble_invariant(cfg, isolv):
    c, o, eta, s32 = cfg
    mu = c["mu"]
    o.sw_term(3)
    o.set_sap(c["bs"], isolv, 4, 5)
    assert o.sap_prepare() == 0
    psi = np.zeros((o.V, 24))
    rho = eta.copy()
    psif = np.zeros((o.V, 24), np.float32)
    rhof = eta.astype(np.float32)
    for _ in range(2):
        o.sap_dble(mu, isolv, 4, psi, rho)
        o.sap(np.float32(mu), isolv, 4, psif, rhof)
    d = o.Dw_dble(mu, psi.copy())
    inv = np.linalg.norm(eta - d - rho) / np.linalg.norm(eta)
    assert inv < 1e-13, inv
    assert np.linalg.norm(rho) < 0.9 * np.linalg.norm(eta)
    rel = np.linalg.norm(psif - psi) / np.linalg.norm(psi)
    assert rel < 1e-4, rel
