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
