"""Pins of the CPU oracle that do not need the reference build: the reference's
own known-answer vectors and closed forms of devel/, invariants of the
operator, textbook linear algebra, and the committed reference samples."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import pyoracle as po
from helpers import GAMMA5, cdot, gamma5, make_oracle, site_rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# gamma matrices of the reference (devel/dirac/check5.c:52-103)
G = np.zeros((4, 4, 4), complex)
G[0] = [[0, 0, -1, 0], [0, 0, 0, -1], [-1, 0, 0, 0], [0, -1, 0, 0]]
G[1] = [[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]]
G[2] = [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]]
G[3] = [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]]


def to_c(a):
    """[V,24] real -> [V,4,3] complex (spin, colour)."""
    return (a[:, 0::2] + 1j * a[:, 1::2]).reshape(-1, 4, 3)


def from_c(z):
    z = z.reshape(z.shape[0], 12)
    out = np.empty((z.shape[0], 24))
    out[:, 0::2] = z.real
    out[:, 1::2] = z.imag
    return out


def test_ranlux_known_answers():
    k = json.load(open(os.path.join(GOLD, "ranlux_kat.json")))
    s = k["ranlxs"]
    xs = po.kat_ranlxs(s["level"], s["seed"], s["calls"] * s["per_call"])
    last = xs[(s["calls"] - 1) * s["per_call"]:]
    got = last[s["offset"]:s["offset"] + 96].astype(np.float64) * 2.0 ** 24
    assert np.array_equal(got, np.array(s["values_times_2p24"]))
    d = k["ranlxd"]
    xd = po.kat_ranlxd(d["level"], d["seed"], d["calls"] * d["per_call"])
    last = xd[(d["calls"] - 1) * d["per_call"]:]
    got = last[d["offset"]:d["offset"] + 48] * 2.0 ** 48
    assert np.array_equal(got, np.array(d["values_times_2p48"]))


def test_geometry_is_a_bijection_with_even_first():
    o = po.Oracle((8, 8, 8, 16), (4, 4, 4, 8))
    ipt = o.ipt()
    assert np.array_equal(np.sort(ipt), np.arange(o.V))
    lx = np.arange(o.V)
    x = np.stack([lx // 1024, (lx // 128) % 8, (lx // 16) % 8, lx % 16], 1)
    par = x.sum(1) & 1
    assert (ipt[par == 0] < o.V // 2).all() and (ipt[par == 1] >= o.V // 2).all()
    iup, idn = o.iupdn()
    assert np.array_equal(idn[iup[:, 2], 2], np.arange(o.V))


@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_plane_wave_closed_form(mu):
    """Unit links, periodic bc with antiperiodic quarks: D acts on
    exp(ipx) u as m0 + sum(1-cos p) + i sum gamma_mu sin p_mu + i mu gamma5
    (devel/dirac/check5.c:184-330)."""
    L = (8, 4, 4, 6)
    o = po.Oracle(L, (4, 4, 4, 6), bc=3, m0=0.17, csw=1.3)
    ud = np.zeros((4 * o.V, 18))
    ud[:, [0, 8, 16]] = 1.0
    o.set_ud(ud)
    o.set_ud_phase()
    assert o.sw_term(3) == 0
    rng = np.random.default_rng(1)
    n = [3, 1, 2, 5]
    p = np.array([(2 * np.pi * n[0] + np.pi) / L[0]] + [2 * np.pi * n[k] / L[k] for k in (1, 2, 3)])
    u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    lex = np.empty(o.V, int)
    lex[o.ipt()] = np.arange(o.V)
    x = np.stack([lex // (L[1] * L[2] * L[3]), (lex // (L[2] * L[3])) % L[1], (lex // L[3]) % L[2], lex % L[3]], 1)
    ph = np.exp(1j * (x @ p))
    psi = ph[:, None, None] * u[None]
    M = (0.17 + np.sum(1 - np.cos(p))) * np.eye(4) + 1j * sum(np.sin(p[k]) * G[k] for k in range(4))
    M = M + 1j * mu * np.diag([1, 1, -1, -1])
    ref = ph[:, None, None] * (M @ u)[None]
    got = to_c(o.Dw_dble(mu, from_c(psi)))
    assert np.abs(got - ref).max() < 1e-13


def test_gamma5_hermiticity_and_eo_pieces():
    o, eta, _ = make_oracle((8, 8, 8, 8), (4, 4, 4, 4), bc=3, m0=-0.2, csw=1.5, seed=9)
    o.sw_term(3)
    rng = np.random.default_rng(3)
    a, b = eta, rng.standard_normal(eta.shape)
    Da, Db = o.Dw_dble(0.0, a.copy()), o.Dw_dble(0.0, b.copy())
    lhs, rhs = cdot(a, gamma5(Db)), cdot(gamma5(Da), b)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # Dw = Dee + Doo + Doe + Deo (devel/dirac/check6.c:245-267)
    mu = 0.2
    full = o.Dw_dble(mu, a.copy())
    r = o.Dwee_dble(mu, a.copy(), np.zeros_like(a))
    r = o.Dwoo_dble(mu, a.copy(), r)
    t = o.Dwoe_dble(a.copy(), np.zeros_like(a))
    h = o.V // 2
    r[h:] += t[h:]
    t2 = o.Dweo_dble(a.copy(), np.zeros_like(a))
    r[:h] -= t2[:h]
    rel, _ = site_rel(r, full)
    assert rel < 1e-14


def test_schur_complement_identity():
    """Dwhat = Dee - Deo Doo^-1 Doe (devel/dirac/check6.c:269-287)."""
    o, eta, _ = make_oracle((8, 8, 8, 8), None, bc=0, m0=-0.1, csw=1.7, seed=4)
    mu = 0.05
    h = o.V // 2
    s = eta.copy()
    s[h:] = 0.0
    o.sw_term(2)
    hat = o.Dwhat_dble(mu, s.copy())
    r = o.Dwee_dble(mu, s.copy(), np.zeros_like(s))
    t = o.Dwoe_dble(s.copy(), np.zeros_like(s))
    t = o.Dwoo_dble(0.0, t.copy(), t.copy())  # odd part holds the inverse
    r = o.Dweo_dble(t.copy(), r)
    rel, _ = site_rel(hat[:h], r[:h])
    assert rel < 1e-13


def _random_gauge(n, rng):
    z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, rr = np.linalg.qr(z)
    d = np.diagonal(rr, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    q[:, 2, :] /= np.linalg.det(q)[:, None]
    return q


def test_gauge_covariance():
    """D[U^g] g psi = g D[U] psi (devel/dirac/check4.c); also covers the SW term."""
    o, eta, _ = make_oracle((8, 8, 8, 8), None, bc=0, m0=-0.3, csw=1.77, seed=6)
    o.sw_term(3)
    r0 = o.Dw_dble(0.1, eta.copy())
    V, h = o.V, o.V // 2
    rng = np.random.default_rng(8)
    g = _random_gauge(V, rng)
    iup, idn = o.iupdn()
    udc =(o.ud()[:, 0::2] + 1j * o.ud()[:, 1::2]).reshape(h, 8, 3, 3)
    new = np.empty_like(udc)
    for mu in range(4):
        ox = np.arange(h, V)
        # U(x,mu) for odd x: g(x) U g(x+mu)^dag
        new[:, 2 * mu] = g[ox] @ udc[:, 2 * mu] @ np.conj(np.transpose(g[iup[ox, mu]], (0, 2, 1)))
        # U(x-mu,mu): g(x-mu) U g(x)^dag
        new[:, 2 * mu + 1] = g[idn[ox, mu]] @ udc[:, 2 * mu + 1] @ np.conj(np.transpose(g[ox], (0, 2, 1)))
    flat = new.reshape(4 * V, 9)
    udn = np.empty((4 * V, 18))
    udn[:, 0::2] = flat.real
    udn[:, 1::2] = flat.imag
    o.set_ud(udn)
    o.sw_term(3)
    geta = from_c(np.einsum("xij,xsj->xsi", g, to_c(eta)))
    r1 = o.Dw_dble(0.1, geta)
    gr0 = from_c(np.einsum("xij,xsj->xsi", g, to_c(r0)))
    rel, _ = site_rel(r1, gr0)
    assert rel < 1e-13


def test_pauli_inverse_and_product():
    rng = np.random.default_rng(2)
    for trial in range(20):
        A = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
        A = A + A.conj().T + 8 * np.eye(6)
        mu = 0.3 * rng.standard_normal()
        m = np.zeros(36)
        m[:6] = A.diagonal().real
        p = 6
        for i in range(6):
            for j in range(i + 1, 6):
                m[p], m[p + 1] = A[i, j].real, A[i, j].imag
                p += 2
        s = rng.standard_normal(12)
        r = po.mul_pauli_dble(mu, m, s)
        v = s[0::2] + 1j * s[1::2]
        w = (A + 1j * mu * np.eye(6)) @ v
        assert np.abs(r[0::2] + 1j * r[1::2] - w).max() < 1e-13
        ifail, im = po.inv_pauli_dble(mu, m)
        Mm = A + 1j * mu * np.eye(6)
        X = np.linalg.inv(Mm)
        H = 0.5 * (X + X.conj().T)
        # safe iff the Frobenius condition number is below 100
        condF = np.linalg.norm(Mm) * np.linalg.norm(X)
        if abs(condF - 100.0) > 1.0:
            assert ifail == int(condF > 100.0)
        assert abs(im[0] - H[0, 0].real) < 1e-14 and abs(im[6] - H[0, 1].real) < 1e-14
        assert abs(im[35] - H[4, 5].imag) < 1e-14
    # ill-conditioned: Frobenius condition number above 100 is unsafe
    A = np.diag([1.0, 1.0, 1.0, 1.0, 1.0, 1e-3]).astype(complex)
    m = np.zeros(36)
    m[:6] = A.diagonal().real
    assert po.inv_pauli_dble(0.0, m)[0] == 1


def test_qsum_exact():
    rng = np.random.default_rng(4)
    vals = list(rng.standard_normal(300) * 10.0 ** rng.integers(-12, 12, 300))
    q = np.zeros(2)
    for v in vals:
        q = po.acc_qflt(v, q)
    exact = sum(Fraction(v) for v in vals)
    assert abs(Fraction(q[0]) + Fraction(q[1]) - exact) <= Fraction(1, 10 ** 30) * sum(abs(Fraction(v)) for v in vals)
    qa = po.add_qflt([1e20, 1.0], [-1e20, 2.0 ** -60])
    assert Fraction(qa[0]) + Fraction(qa[1]) == Fraction(1) + Fraction(2) ** -60


@pytest.mark.parametrize("name", ["periodic_cfg1", "open"])
def test_reference_samples(name):
    """Oracle against the committed samples of the reference build."""
    z = np.load(os.path.join(GOLD, "ref_8x8x8x8_%s.npz" % name))
    info = json.loads(str(z["info"]))
    pts = z["points"]
    o, eta, s32 = make_oracle((8, 8, 8, 8), None, bc=info["bc"], m0=info["m0"], csw=info["csw"],
                              seed=info["seed"])
    mu = info["mu"]
    assert np.abs(eta[pts] - z["eta_f64"]).max() < 1e-15
    o.sw_term(3)
    sw = o.swd().reshape(4096, 72)
    assert site_rel(sw[pts], z["swd_f64"])[0] < 1e-14
    r = o.Dw_dble(mu, eta.copy())
    assert site_rel(r[pts], z["Dw_dble_f64"])[0] < 1e-14
    o.assign_ud2u()
    o.assign_swd2sw()
    r = o.Dw(np.float32(mu), s32.copy())
    assert site_rel(r[pts], z["Dw_f32"])[0] < 1e-6
    o.sw_term(2)
    r = o.Dwhat_dble(mu, eta.copy())
    ev = pts[pts < 2048]
    assert site_rel(r[ev], z["Dwhat_dble_f64"][pts < 2048])[0] < 1e-14
    rho, ifail, status, psi = o.sap_gcr(16, 256, 0, 1e-10, mu, eta.copy()) if o.set_sap((4, 4, 4, 4), info["isolv"], 4, 5) is None else (None,) * 4
    assert ifail == info["gcr_ifail"] and abs(status - info["gcr_status"]) <= 1
    assert site_rel(psi[pts], z["gcr_psi_f64"])[0] < 1e-8


def test_expXsu3_is_the_matrix_exponential():
    """expXsu3 (su3fcts/chexp.c:878) against the matrix exponential of scipy,
    including norms large enough for the halving-and-squaring branch."""
    from scipy.linalg import expm
    from helpers import alg_matrix
    rng = np.random.default_rng(11)
    for scale in (0.01, 0.3, 1.0, 5.0, 40.0):
        X = rng.standard_normal(8) * scale
        u = _random_gauge(1, rng)[0]
        ud = np.empty(18)
        ud[0::2] = u.ravel().real
        ud[1::2] = u.ravel().imag
        for eps in (0.02, -0.7):
            out = po.expXsu3(eps, X, ud)
            ref = expm(eps * alg_matrix(X)) @ u
            got = (out[0::2] + 1j * out[1::2]).reshape(3, 3)
            assert np.abs(got - ref).max() < 2e-14 * max(1.0, abs(eps) * scale), (scale, eps)


@pytest.mark.parametrize("name", ["periodic", "open"])
def test_flow_against_reference_samples(name):
    """The oracle's force and 4 flow steps from its own random field agree with
    the reference's at the sampled links (the two random fields differ in the
    last bit, see test_fields)."""
    g = np.load(os.path.join(GOLD, "ref_8x8x8x8_flow_%s.npz" % name))
    info = json.loads(str(g["info"]))
    o = po.Oracle((8, 8, 8, 8), None, bc=info["bc"])
    o.start_ranlux(0, info["seed"])
    o.random_ud()
    f = o.plaq_frc()
    assert np.abs(f[g["links"]] - g["frc0"]).max() < 1e-13 * float(g["frc0_norm"]) / math.sqrt(f.size)
    o.fwd_euler(info["nflow"], info["eps"])
    u = o.ud()
    assert np.abs(u[g["links"]] - g["ud_flow"]).max() < 1e-13


def test_flow_invariants():
    """The flow keeps the links in SU(3), increases the plaquette sum (it
    descends along the Wilson action) and its force is gauge covariant."""
    from helpers import alg_matrix, gauge_transform, link_base_points, plaquette_sum
    for bc in (3, 0):
        o = po.Oracle((8, 8, 8, 8), (4, 4, 4, 4), bc=bc)
        o.start_ranlux(0, 5)
        o.random_ud()
        iup, idn = o.iupdn()
        V, h = o.V, o.V // 2
        ud0 = o.ud()
        # gauge covariance of the force: F[U^g](x,mu) = g(x) F[U](x,mu) g(x)^+
        F0 = alg_matrix(o.plaq_frc()).reshape(h, 8, 3, 3)
        g = _random_gauge(V, np.random.default_rng(2))
        o.set_ud(gauge_transform(ud0, g, iup, idn))
        F1 = alg_matrix(o.plaq_frc()).reshape(h, 8, 3, 3)
        base = link_base_points(iup, idn, h)
        gb = g[base]
        rot = gb @ F0 @ np.conj(np.transpose(gb, (0, 1, 3, 2)))
        assert np.abs(F1 - rot).max() < 1e-13
        # monotone smoothing, unitarity
        o.set_ud(ud0)
        s = [plaquette_sum(o.ud(), iup)]
        for _ in range(3):
            o.fwd_euler(2, 0.02)
            s.append(plaquette_sum(o.ud(), iup))
        assert all(b > a for a, b in zip(s, s[1:])), s
        m = (o.ud()[:, 0::2] + 1j * o.ud()[:, 1::2]).reshape(-1, 3, 3)
        # open bc: the time-like links at x0 = N0-1 are zero and stay so
        live = np.abs(m).max(axis=(1, 2)) > 0
        assert live.sum() == (4 * V if bc == 3 else 4 * V - V // 8)
        m = m[live]
        eye = np.eye(3)
        assert np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - eye).max() < 1e-13
        assert np.abs(np.linalg.det(m) - 1).max() < 1e-13
