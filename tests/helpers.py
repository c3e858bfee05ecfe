"""Shared helpers of the test suite (parity metrics, field set-up)."""
import numpy as np

import pyoracle as po


def site_rel(a, b):
    """Maximum over points of |a(x)-b(x)| / |b(x)|; points where b vanishes
    must vanish in a as well (their absolute deviation is returned as the
    second value)."""
    a = np.asarray(a, np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, np.float64).reshape(b.shape[0], -1)
    d = np.linalg.norm(a - b, axis=1)
    n = np.linalg.norm(b, axis=1)
    m = n > 0
    rel = float((d[m] / n[m]).max()) if m.any() else 0.0
    absz = float(d[~m].max()) if (~m).any() else 0.0
    return rel, absz


def make_oracle(L, Ltrd=None, bc=3, cF=(1.0, 1.0), m0=-0.5, csw=1.0, seed=1, eoflg=0):
    """Oracle lattice with the reference's seeded fields:
    start_ranlux(0,seed); random_ud(); set_ud_phase(); random_sd(); random_s()."""
    o = po.Oracle(L, Ltrd, bc=bc, cF=cF, m0=m0, csw=csw, eoflg=eoflg)
    o.start_ranlux(0, seed)
    o.random_ud()
    o.set_ud_phase()
    eta = o.random_sd()
    o.bnd_sd2zero(eta)
    s32 = o.random_s()
    o.bnd_s2zero(s32)
    return o, eta, s32


GAMMA5 = np.array([1] * 12 + [-1] * 12, np.float64)


def gamma5(s):
    return s * GAMMA5[None, :].astype(s.dtype)


def cdot(a, b):
    """(a, b) = sum conj(a) b for [V,24] real-packed spinors."""
    za = a[:, 0::2] + 1j * a[:, 1::2]
    zb = b[:, 0::2] + 1j * b[:, 1::2]
    return np.vdot(za.ravel(), zb.ravel())


# ---- gauge-field helpers (reference link layout: 8 links per odd point) ----

def alg_matrix(X):
    """su3_alg_dble components c1..c8 (..., 8) -> antihermitian 3x3 matrices
    (the basis of su3fcts/chexp.c mapX2v)."""
    X = np.asarray(X)
    M = np.zeros(X.shape[:-1] + (3, 3), complex)
    c = [X[..., k] for k in range(8)]
    M[..., 0, 0] = 1j * (c[0] + c[1])
    M[..., 1, 1] = 1j * (c[1] - 2 * c[0])
    M[..., 2, 2] = 1j * (c[0] - 2 * c[1])
    M[..., 0, 1] = c[2] + 1j * c[3]
    M[..., 0, 2] = c[4] + 1j * c[5]
    M[..., 1, 2] = c[6] + 1j * c[7]
    M[..., 1, 0] = -c[2] + 1j * c[3]
    M[..., 2, 0] = -c[4] + 1j * c[5]
    M[..., 2, 1] = -c[6] + 1j * c[7]
    return M


def ud_matrices(ud):
    """(4V, 18) reference layout -> (V/2, 8, 3, 3) complex."""
    return (ud[:, 0::2] + 1j * ud[:, 1::2]).reshape(-1, 8, 3, 3)


def ud_from_matrices(m):
    flat = m.reshape(-1, 9)
    ud = np.empty((flat.shape[0], 18))
    ud[:, 0::2] = flat.real
    ud[:, 1::2] = flat.imag
    return ud


def link_field(ud, iup):
    """U(x,mu) for every point x: (V, 4, 3, 3)."""
    m = ud_matrices(ud)
    h = m.shape[0]
    V = 2 * h
    U = np.empty((V, 4, 3, 3), complex)
    for mu in range(4):
        U[h:, mu] = m[:, 2 * mu]
        U[:h, mu] = m[iup[:h, mu] - h, 2 * mu + 1]
    return U


def plaquette_sum(ud, iup):
    """sum_x sum_{mu<nu} Re tr U(x,mu) U(x+mu,nu) U(x+nu,mu)^+ U(x,nu)^+."""
    U = link_field(ud, iup)
    s = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            P = U[:, mu] @ U[iup[:, mu], nu] @ np.conj(np.transpose(U[iup[:, nu], mu], (0, 2, 1))) \
                @ np.conj(np.transpose(U[:, nu], (0, 2, 1)))
            s += np.trace(P, axis1=1, axis2=2).real.sum()
    return s


def gauge_transform(ud, g, iup, idn):
    """U(x,mu) -> g(x) U(x,mu) g(x+mu)^+ in the reference layout."""
    m = ud_matrices(ud)
    h = m.shape[0]
    V = 2 * h
    ox = np.arange(h, V)
    new = np.empty_like(m)
    for mu in range(4):
        new[:, 2 * mu] = g[ox] @ m[:, 2 * mu] @ np.conj(np.transpose(g[iup[ox, mu]], (0, 2, 1)))
        new[:, 2 * mu + 1] = g[idn[ox, mu]] @ m[:, 2 * mu + 1] @ np.conj(np.transpose(g[ox], (0, 2, 1)))
    return ud_from_matrices(new)


def link_base_points(iup, idn, h):
    """Point x of link (odd point, slot) with U(x,mu): x = o for even slots,
    o - mu for odd slots; (V/2, 8) array."""
    V = 2 * h
    ox = np.arange(h, V)
    base = np.empty((h, 8), np.int64)
    for mu in range(4):
        base[:, 2 * mu] = ox
        base[:, 2 * mu + 1] = idn[ox, mu]
    return base
