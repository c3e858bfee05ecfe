"""The remaining public entry points of the linear algebra (linalg/salg.c,
linalg/salg_dble.c, utils/qsum.c, utils/wspace.c); fgcr with host callbacks
and the unsafe-inversion path of sap_gcr are in test_gpu_fgcr_unsafe.py.

Exact references: double products against math.fsum of the exact double
products, double-double arithmetic against Python fractions, sign flips and
scalings bitwise; single-precision products within the rounding bound of
float colour-vector sums."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import pyoqh

P = C.c_void_p
G5 = np.array([1.0] * 12 + [-1.0] * 12)


def _sig(lib):
    cq, q = pyoqh.ComplexQflt, pyoqh.Qflt
    sigs = {
        "spinor_prod_re_dble": (q, [C.c_int, C.c_int, P, P]),
        "spinor_prod5_dble": (cq, [C.c_int, C.c_int, P, P]),
        "spinor_prod_dble": (cq, [C.c_int, C.c_int, P, P]),
        "norm_square_dble": (q, [C.c_int, C.c_int, P]),
        "spinor_prod_re": (C.c_float, [C.c_int, C.c_int, P, P]),
        "spinor_prod": (pyoqh.Complex, [C.c_int, C.c_int, P, P]),
        "norm_square": (C.c_float, [C.c_int, C.c_int, P]),
        "mulr_spinor_add": (None, [C.c_int, C.c_int, P, P, C.c_float]),
        "mulr_spinor_add_dble": (None, [C.c_int, C.c_int, P, P, C.c_double]),
        "mulc_spinor_add_dble": (None, [C.c_int, C.c_int, P, P, pyoqh.ComplexDble]),
        "scale": (None, [C.c_int, C.c_int, C.c_float, P]),
        "scale_dble": (None, [C.c_int, C.c_int, C.c_double, P]),
        "project": (None, [C.c_int, C.c_int, P, P]),
        "project_dble": (None, [C.c_int, C.c_int, P, P]),
        "normalize": (C.c_float, [C.c_int, C.c_int, P]),
        "normalize_dble": (C.c_double, [C.c_int, C.c_int, P]),
        "mulg5": (None, [C.c_int, C.c_int, P]), "mulmg5": (None, [C.c_int, C.c_int, P]),
        "mulg5_dble": (None, [C.c_int, C.c_int, P]), "mulmg5_dble": (None, [C.c_int, C.c_int, P]),
        "add_qflt": (None, [P, P, P]), "scl_qflt": (None, [C.c_double, P]),
        "mul_qflt": (None, [P, P, P]), "acc_qflt": (None, [C.c_double, P]),
        "global_qsum": (None, [C.c_int, P, P]),
        "alloc_ws": (None, [C.c_int]), "reserve_ws": (P, [C.c_int]), "release_ws": (C.c_int, []),
        "alloc_wsd": (None, [C.c_int]), "reserve_wsd": (P, [C.c_int]), "release_wsd": (C.c_int, []),
        "set_s2zero": (None, [C.c_int, C.c_int, P]), "assign_s2s": (None, [C.c_int, C.c_int, P, P]),
        "diff_s2s": (None, [C.c_int, C.c_int, P, P]),
        "sap": (None, [C.c_float, C.c_int, C.c_int, P, P]),
        "sap_gcr": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "unorm_dble": (C.c_double, [C.c_int, C.c_int, P]),
        "oqh_last_error": (C.c_int, [C.c_char_p, C.c_int]),
    }
    for n, (r, a) in sigs.items():
        getattr(lib, n).restype = r
        getattr(lib, n).argtypes = a


@pytest.fixture(scope="module")
def lat():
    lt = pyoqh.Lattice((8, 8, 8, 8), (4, 4, 4, 4), bc=3, m0=-0.5, csw=1.0)
    _sig(lt.lib)
    yield lt
    lt.close()


def _z(a):
    return a[:, 0::2] + 1j * a[:, 1::2]


def _fs(x):
    return math.fsum(np.asarray(x).ravel().tolist())


@pytest.mark.gpu
def test_double_products_and_updates(lat):
    lib, V = lat.lib, lat.V
    rng = np.random.default_rng(41)
    a = rng.standard_normal((V, 24)) * np.exp(rng.standard_normal((V, 1)))
    b = rng.standard_normal((V, 24))
    da, db = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(da, a)
    lat.put_sd(db, b)
    scale_ = _fs(np.abs(a * b))
    q = lib.spinor_prod_re_dble(V, 0, da, db)
    assert abs(q.q[0] + q.q[1] - _fs(a * b)) <= 1e-15 * scale_
    # (s, gamma5 r): spins 1,2 plus, 3,4 minus (salg_dble.c spinor_prod5_dble)
    q5 = lib.spinor_prod5_dble(V, 0, da, db)
    g = G5[None, :]
    re5 = _fs(g[:, 0::2] * (a[:, 0::2] * b[:, 0::2] + a[:, 1::2] * b[:, 1::2]))
    im5 = _fs(g[:, 0::2] * (a[:, 0::2] * b[:, 1::2] - a[:, 1::2] * b[:, 0::2]))
    assert abs(q5.re.q[0] + q5.re.q[1] - re5) <= 1e-15 * scale_
    assert abs(q5.im.q[0] + q5.im.q[1] - im5) <= 1e-15 * scale_
    # sign flips are exact
    lib.mulg5_dble(V, 0, da)
    assert np.array_equal(lat.get_sd(da), a * G5)
    lib.mulmg5_dble(V, 0, da)            # -gamma5 gamma5 a = -a
    assert np.array_equal(lat.get_sd(da), -a)
    # scale_dble, mulr/mulc_spinor_add_dble
    lat.put_sd(da, a)
    lib.scale_dble(V, 0, -0.375, da)
    assert np.array_equal(lat.get_sd(da), -0.375 * a)
    lat.put_sd(da, a)
    lib.mulr_spinor_add_dble(V, 0, da, db, 0.7)
    assert np.abs(lat.get_sd(da) - (a + 0.7 * b)).max() <= 1e-15 * np.abs(a + 0.7 * b).max()
    lat.put_sd(da, a)
    lib.mulc_spinor_add_dble(V, 0, da, db, pyoqh.ComplexDble(0.25, -1.5))
    ref = _z(a) + (0.25 - 1.5j) * _z(b)
    assert np.abs(_z(lat.get_sd(da)) - ref).max() <= 1e-15 * np.abs(ref).max() * 4
    # project_dble: s -= (r, s) r / ... with normalised r (salg_dble.c)
    lat.put_sd(da, a)
    nb = math.sqrt(_fs(b * b))
    lat.put_sd(db, b / nb)
    lib.project_dble(V, 0, da, db)
    got = _z(lat.get_sd(da))
    zb = _z(b / nb)
    assert abs(np.vdot(zb.ravel(), got.ravel())) <= 1e-13 * np.linalg.norm(got)
    ref = _z(a) - np.vdot(zb.ravel(), _z(a).ravel()) * zb
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # normalize_dble returns the norm and leaves a unit vector
    lat.put_sd(da, a)
    r = lib.normalize_dble(V, 0, da)
    assert abs(r - math.sqrt(_fs(a * a))) <= 1e-14 * r
    n = lib.norm_square_dble(V, 0, da)
    assert abs(n.q[0] + n.q[1] - 1.0) <= 1e-14
    lat.free(da)
    lat.free(db)


@pytest.mark.gpu
def test_float_products_and_updates(lat):
    lib, V = lat.lib, lat.V
    rng = np.random.default_rng(42)
    a = rng.standard_normal((V, 24)).astype(np.float32)
    b = rng.standard_normal((V, 24)).astype(np.float32)
    sa, sb = lat.alloc_s(), lat.alloc_s()
    lat.put_s(sa, a)
    lat.put_s(sb, b)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    bound = 1e-6 * float(np.abs(a64 * b64).sum())
    assert abs(lib.spinor_prod_re(V, 0, sa, sb) - _fs(a64 * b64)) <= bound
    lib.mulg5(V, 0, sa)
    assert np.array_equal(lat.get_s(sa), a * G5.astype(np.float32))
    lib.mulmg5(V, 0, sa)                 # -gamma5 gamma5 a = -a
    assert np.array_equal(lat.get_s(sa), -a)
    lat.put_s(sa, a)
    lib.scale(V, 0, np.float32(0.5), sa)
    assert np.array_equal(lat.get_s(sa), a * np.float32(0.5))
    lat.put_s(sa, a)
    lib.mulr_spinor_add(V, 0, sa, sb, np.float32(-0.3))
    ref = a64 + np.float64(np.float32(-0.3)) * b64
    assert np.abs(lat.get_s(sa) - ref).max() <= 2e-6 * np.abs(ref).max()
    lat.put_s(sa, a)
    r = lib.normalize(V, 0, sa)
    assert abs(r - math.sqrt(_fs(a64 * a64))) <= 1e-6 * r
    assert abs(lib.norm_square(V, 0, sa) - 1.0) <= 1e-5
    # project: s - (r, s) r with |r| = 1 leaves s orthogonal to r
    lat.put_s(sa, a)
    lib.project(V, 0, sa, sb)   # r = b is not normalised: s -= (b, s) b
    got = _z(lat.get_s(sa).astype(np.float64))
    zb = _z(b64)
    ref = _z(a64) - np.vdot(zb.ravel(), _z(a64).ravel()) * zb
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
    lat.free(sa)
    lat.free(sb)


@pytest.mark.gpu
def test_qsum_arithmetic():
    lib = pyoqh.load()
    _sig(lib)
    rng = np.random.default_rng(43)

    def F(q):
        return Fraction(q[0]) + Fraction(q[1])

    for _ in range(200):
        x = rng.standard_normal(4) * np.exp(rng.uniform(-30, 30, 4))
        qu = np.array([x[0], x[0] * 2.0 ** -60])
        qv = np.array([x[1], x[1] * 2.0 ** -58])
        # normalise the pairs (|lo| <= ulp(hi)/2)
        for q in (qu, qv):
            s = q[0] + q[1]
            q[1] = q[1] - (s - q[0])
            q[0] = s
        qr = np.zeros(2)
        lib.add_qflt(qu.ctypes.data_as(P), qv.ctypes.data_as(P), qr.ctypes.data_as(P))
        ex = F(qu) + F(qv)
        assert abs(F(qr) - ex) <= Fraction(2.0 ** -100) * (abs(F(qu)) + abs(F(qv)))
        qs = qu.copy()
        lib.scl_qflt(float(x[2]), qs.ctypes.data_as(P))
        ex = Fraction(x[2]) * F(qu)
        assert abs(F(qs) - ex) <= Fraction(2.0 ** -100) * abs(ex)
        qm = np.zeros(2)
        lib.mul_qflt(qu.ctypes.data_as(P), qv.ctypes.data_as(P), qm.ctypes.data_as(P))
        ex = F(qu) * F(qv)
        assert abs(F(qm) - ex) <= Fraction(2.0 ** -100) * abs(ex)
    # global_qsum with one process: the sums are the local values
    src = [np.array([1.5, 1e-20]), np.array([-2.0, 3e-19])]
    dst = [np.zeros(2), np.zeros(2)]
    pu = (P * 2)(*[s.ctypes.data_as(P) for s in src])
    pr = (P * 2)(*[d.ctypes.data_as(P) for d in dst])
    lib.global_qsum(2, pu, pr)
    assert all(np.array_equal(s, d) for s, d in zip(src, dst))


@pytest.mark.gpu
def test_workspace(lat):
    lib, V = lat.lib, lat.V
    lib.alloc_ws(3)
    lib.alloc_wsd(2)
    w1 = lib.reserve_ws(2)
    w2 = lib.reserve_ws(1)
    p1 = C.cast(w1, C.POINTER(P))
    p2 = C.cast(w2, C.POINTER(P))
    fields = [p1[0], p1[1], p2[0]]
    assert len(set(fields)) == 3
    for f in fields:
        assert not lat.get_s(f).any()   # reserved fields are zero
    lat.put_s(fields[0], np.ones((V, 24), np.float32))
    lib.oqh_set_error_mode(0)
    lib.oqh_last_error(None, 0)
    assert not lib.reserve_ws(1)       # all three are in use
    assert lib.oqh_last_error(None, 0) == 1
    lib.oqh_set_error_mode(1)
    assert lib.release_ws() == 1 and lib.release_ws() == 2
    wd = C.cast(lib.reserve_wsd(2), C.POINTER(P))
    assert wd[0] != wd[1] and not lat.get_sd(wd[0]).any()
    assert lib.release_wsd() == 2
    assert lib.release_wsd() == 0 and lib.release_ws() == 0
