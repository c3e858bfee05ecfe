"""Pins of the numpy oracle of the gauge-equivariant network (oracle/ptc.py)
by properties the mathematics fixes: gauge covariance, the adjoint identity,
finite differences of the loss, and weights that reduce a layer to the
identity or to a covariant shift."""
import numpy as np
import pytest

import ptc
import pyoracle as po
from helpers import gauge_transform, link_field, ud_from_matrices


@pytest.fixture(scope="module")
def lat():
    o = po.Oracle((8, 8, 8, 8), None, bc=3)
    o.start_ranlux(0, 5)
    o.random_ud()
    iup, idn = o.iupdn()
    ud = o.ud()
    return o, ud, link_field(ud, iup), iup, idn


def rnd_field(rng, V, C):
    return rng.standard_normal((V, C, 3)) + 1j * rng.standard_normal((V, C, 3))


def su3_random(rng, n):
    z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1.0 / 3.0)


def test_gauge_covariance(lat):
    o, ud, U, iup, idn = lat
    rng = np.random.default_rng(1)
    V = U.shape[0]
    Ws = ptc.init_weights([4, 12, 16, 12, 4], rng)
    psi = rnd_field(rng, V, 4)
    g = su3_random(rng, V)
    U2 = link_field(gauge_transform(ud, g, iup, idn), iup)
    out = ptc.net_fwd(Ws, psi, U, iup, idn)[-1]
    out2 = ptc.net_fwd(Ws, np.einsum("xab,xib->xia", g, psi), U2, iup, idn)[-1]
    ref = np.einsum("xab,xib->xia", g, out)
    assert np.abs(out2 - ref).max() < 1e-12 * np.abs(ref).max()
    # a plain (non-covariant) shift would fail the check
    bad = np.roll(psi, 1, axis=0)
    assert np.abs(np.einsum("xab,xib->xia", g, bad) - np.roll(np.einsum("xab,xib->xia", g, psi), 1, axis=0)).max() > 1e-3


def test_adjoint_identity(lat):
    o, ud, U, iup, idn = lat
    rng = np.random.default_rng(2)
    V = U.shape[0]
    W = ptc.init_weights([12, 16], rng)[0]
    psi, g = rnd_field(rng, V, 12), rnd_field(rng, V, 16)
    lhs = np.vdot(g, ptc.layer_fwd(W, psi, U, iup, idn))
    rhs = np.vdot(ptc.layer_adj(W, g, U, iup, idn), psi)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_special_weights(lat):
    o, ud, U, iup, idn = lat
    rng = np.random.default_rng(3)
    V = U.shape[0]
    psi = rnd_field(rng, V, 4)
    W = np.zeros((9, 4, 4), complex)
    W[0] = np.eye(4)
    assert np.array_equal(ptc.layer_fwd(W, psi, U, iup, idn), psi)
    # W[2mu+1] = 1: out(x) = U(x,mu) psi(x+mu); with unit links a shift
    for mu in range(4):
        W = np.zeros((9, 4, 4), complex)
        W[2 * mu + 1] = np.eye(4)
        out = ptc.layer_fwd(W, psi, U, iup, idn)
        ref = np.einsum("xab,xib->xia", U[:, mu], psi[iup[:, mu]])
        assert np.allclose(out, ref, rtol=0, atol=1e-13)
        one = np.broadcast_to(np.eye(3), U.shape).copy()
        assert np.array_equal(ptc.layer_fwd(W, psi, one, iup, idn), psi[iup[:, mu]])
        W = np.zeros((9, 4, 4), complex)
        W[2 * mu + 2] = np.eye(4)
        assert np.array_equal(ptc.layer_fwd(W, psi, one, iup, idn), psi[idn[:, mu]])


def test_finite_differences(lat):
    o, ud, U, iup, idn = lat
    rng = np.random.default_rng(4)
    V = U.shape[0]
    widths = [4, 8, 4]
    Ws = ptc.init_weights(widths, rng)
    psi = rnd_field(rng, V, 4)
    e = rnd_field(rng, V, 4)

    def loss(Ws_, psi_):
        return ptc.loss_and_grad(ptc.net_fwd(Ws_, psi_, U, iup, idn)[-1], e)[0]

    acts = ptc.net_fwd(Ws, psi, U, iup, idn)
    L0, g = ptc.loss_and_grad(acts[-1], e)
    dWs, gin = ptc.net_bwd(Ws, acts, g, U, iup, idn)
    h = 1e-6
    for l in range(len(Ws)):
        shp = (9, widths[l + 1], widths[l])
        D = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        Wp = [w.copy() for w in Ws]
        Wm = [w.copy() for w in Ws]
        Wp[l] = Wp[l] + h * D
        Wm[l] = Wm[l] - h * D
        fd = (loss(Wp, psi) - loss(Wm, psi)) / (2 * h)
        an = np.sum(np.conj(dWs[l]) * D).real
        assert abs(fd - an) < 1e-6 * abs(an)
    D = rnd_field(rng, V, 4)
    fd = (loss(Ws, psi + h * D) - loss(Ws, psi - h * D)) / (2 * h)
    an = np.sum(np.conj(gin) * D).real
    assert abs(fd - an) < 1e-6 * abs(an)
