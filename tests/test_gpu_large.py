"""Size-independent properties at the benchmark's full size (32^3x64, open bc).
The oracle is not run at these sizes; the checks are invariants of the operator:
gamma5-hermiticity of Dw_dble, Dw_dble = Dee + Doo + Doe + Deo, the Schur
complement identity, and FP32/FP64 agreement."""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import cdot, gamma5

P = C.c_void_p


@pytest.fixture(scope="module")
def big():
    L = (64, 32, 32, 32)
    lat = pyoqh.Lattice(L, (8, 8, 8, 8), bc=0, m0=-0.3, csw=1.77)
    rng = np.random.default_rng(21)
    print("large lattice: generating links", flush=True)
    lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
    yield lat, rng
    lat.close()


def _apply(lat, name, mu, s_host, r_host=None):
    s, r = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(s, s_host)
    if r_host is not None:
        lat.put_sd(r, r_host)
    fn = getattr(lat.lib, name)
    if mu is None:
        fn(s, r)
    else:
        fn(mu, s, r)
    out = lat.get_sd(r)
    lat.free(s)
    lat.free(r)
    return out


@pytest.mark.gpu
def test_gamma5_hermiticity_full_size(big):
    lat, rng = big
    assert lat.sw_term(3) == 0
    V = lat.V
    a = rng.standard_normal((V, 24))
    b = rng.standard_normal((V, 24))
    h0 = lat.lib.oqh_volume()
    assert h0 == V
    # zero the boundary slices of the inputs (the operator does it anyway)
    Da, Db = _apply(lat, "Dw_dble", 0.0, a), _apply(lat, "Dw_dble", 0.0, b)
    x0 = np.empty(V, int)
    ipt = lat.ipt()
    lex = np.empty(V, int)
    lex[ipt] = np.arange(V)
    x0 = lex // (32 ** 3)
    bnd = (x0 == 0) | (x0 == 63)
    a[bnd] = 0.0
    b[bnd] = 0.0
    lhs, rhs = cdot(a, gamma5(Db)), cdot(gamma5(Da), b)
    print("gamma5 hermiticity", abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.gpu
def test_eo_decomposition_full_size(big):
    lat, rng = big
    assert lat.sw_term(3) == 0
    V = lat.V
    h = V // 2
    a = rng.standard_normal((V, 24))
    mu = 0.07
    full = _apply(lat, "Dw_dble", mu, a)
    r = _apply(lat, "Dwee_dble", mu, a, np.zeros_like(a))
    r = _apply(lat, "Dwoo_dble", mu, a, r)
    t = _apply(lat, "Dwoe_dble", None, a, np.zeros_like(a))
    r[h:] += t[h:]
    t2 = _apply(lat, "Dweo_dble", None, a, np.zeros_like(a))
    r[:h] -= t2[:h]
    err = np.linalg.norm(r - full) / np.linalg.norm(full)
    print("eo decomposition", err)
    assert err < 1e-14
    # Schur complement identity with the odd part inverted
    assert lat.sw_term(2) == 0
    s = a.copy()
    s[h:] = 0.0
    hat = _apply(lat, "Dwhat_dble", mu, s)
    r = _apply(lat, "Dwee_dble", mu, s, np.zeros_like(s))
    t = _apply(lat, "Dwoe_dble", None, s, np.zeros_like(s))
    t = _apply(lat, "Dwoo_dble", 0.0, t, t)
    r = _apply(lat, "Dweo_dble", None, t, r)
    err = np.linalg.norm(hat[:h] - r[:h]) / np.linalg.norm(r[:h])
    print("Schur identity", err)
    assert err < 1e-13


@pytest.mark.gpu
def test_float_vs_double_full_size(big):
    lat, rng = big
    assert lat.sw_term(3) == 0
    lat.lib.assign_ud2u()
    lat.lib.assign_swd2sw()
    V = lat.V
    a = rng.standard_normal((V, 24)).astype(np.float32)
    rd = _apply(lat, "Dw_dble", 0.0, a.astype(np.float64))
    s, r = lat.alloc_s(), lat.alloc_s()
    lat.put_s(s, a)
    lat.lib.Dw(0.0, s, r)
    rf = lat.get_s(r).astype(np.float64)
    d = np.linalg.norm(rf - rd, axis=1)
    n = np.linalg.norm(rd, axis=1)
    m = n > 0
    print("fp32 vs fp64 max per-point", (d[m] / n[m]).max())
    assert (d[m] / n[m]).max() < 1e-5
