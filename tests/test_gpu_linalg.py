"""GPU parity of the spinor linear algebra, field assignments and double-double
sums (linalg/salg*.c, sflds/sflds.c, sflds/unorm.c, utils/qsum.c)."""
import ctypes as C
import math

import numpy as np
import pytest

import pyoqh
import pyoracle as po

P = C.c_void_p


def _sig(lib):
    lib.spinor_prod_dble.restype = pyoqh.ComplexQflt
    lib.spinor_prod_dble.argtypes = [C.c_int, C.c_int, P, P]
    lib.norm_square_dble.restype = pyoqh.Qflt
    lib.norm_square_dble.argtypes = [C.c_int, C.c_int, P]
    lib.spinor_prod5_dble.restype = pyoqh.ComplexQflt
    lib.spinor_prod5_dble.argtypes = [C.c_int, C.c_int, P, P]
    lib.spinor_prod.restype = pyoqh.Complex
    lib.spinor_prod.argtypes = [C.c_int, C.c_int, P, P]
    lib.norm_square.restype = C.c_float
    lib.norm_square.argtypes = [C.c_int, C.c_int, P]
    lib.mulc_spinor_add_dble.argtypes = [C.c_int, C.c_int, P, P, pyoqh.ComplexDble]
    lib.mulc_spinor_add.argtypes = [C.c_int, C.c_int, P, P, pyoqh.Complex]
    lib.combine_spinor_dble.argtypes = [C.c_int, C.c_int, P, P, C.c_double, C.c_double]
    lib.unorm_dble.restype = C.c_double
    lib.unorm_dble.argtypes = [C.c_int, C.c_int, P]
    lib.unorm.restype = C.c_float
    lib.unorm.argtypes = [C.c_int, C.c_int, P]
    for n in ("assign_sd2s", "assign_s2sd", "add_s2sd", "diff_s2s", "assign_sd2sd"):
        getattr(lib, n).argtypes = [C.c_int, C.c_int, P, P]
    lib.diff_sd2s.argtypes = [C.c_int, C.c_int, P, P, P]


@pytest.fixture(scope="module")
def lat():
    lt = pyoqh.Lattice((8, 8, 8, 8), (4, 4, 4, 4), bc=3)
    _sig(lt.lib)
    yield lt
    lt.close()


def _cplx(a):
    return a[:, 0::2] + 1j * a[:, 1::2]


@pytest.mark.gpu
def test_dd_dot_products(lat):
    rng = np.random.default_rng(5)
    V = lat.V
    a = rng.standard_normal((V, 24)) * np.exp(rng.standard_normal((V, 1)) * 3)
    b = rng.standard_normal((V, 24))
    da, db = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(da, a)
    lat.put_sd(db, b)
    vt = V // 16  # VOLUME_TRD of the (4,4,4,4) thread layout, icom 3 = all threads
    q = lat.lib.spinor_prod_dble(vt, 3, da, db)
    # exact sum of the double products (correctly rounded by fsum)
    re_terms = (a[:, 0::2] * b[:, 0::2]).ravel().tolist() + (a[:, 1::2] * b[:, 1::2]).ravel().tolist()
    im_terms = (a[:, 0::2] * b[:, 1::2]).ravel().tolist() + (-(a[:, 1::2] * b[:, 0::2])).ravel().tolist()
    ex_re, ex_im = math.fsum(re_terms), math.fsum(im_terms)
    # plain double sums over a few points per work-item, double-double above
    # that (as the reference's 8-point chunks folded into a qflt)
    sabs = math.fsum(abs(t) for t in re_terms)
    assert abs((q.re.q[0] + q.re.q[1]) - ex_re) <= 1e-15 * sabs
    assert abs((q.im.q[0] + q.im.q[1]) - ex_im) <= 1e-15 * sabs
    n = lat.lib.norm_square_dble(vt, 3, da)
    ex_n = math.fsum((a * a).ravel().tolist())
    assert abs((n.q[0] + n.q[1]) - ex_n) <= 1e-15 * ex_n
    # agreement with the oracle's reference-order double-double sums
    qre, qim = po.spinor_prod_dble(a, b)
    assert abs(q.re.q[0] - qre[0]) <= 1e-15 * abs(qre[0])
    # half-volume slice through pointer offsets (psd + VOLUME/2)
    h = V // 2
    q2 = lat.lib.spinor_prod_dble(h, 0, da + h * 192, db + h * 192)
    ex2 = math.fsum((a[h:, 0::2] * b[h:, 0::2]).ravel().tolist() + (a[h:, 1::2] * b[h:, 1::2]).ravel().tolist())
    assert abs((q2.re.q[0] + q2.re.q[1]) - ex2) <= 1e-15 * sabs
    lat.free(da)
    lat.free(db)


@pytest.mark.gpu
def test_float_products_and_axpy(lat):
    rng = np.random.default_rng(6)
    V = lat.V
    a = rng.standard_normal((V, 24)).astype(np.float32)
    b = rng.standard_normal((V, 24)).astype(np.float32)
    sa, sb = lat.alloc_s(), lat.alloc_s()
    lat.put_s(sa, a)
    lat.put_s(sb, b)
    z = lat.lib.spinor_prod(V, 0, sa, sb)
    z0 = np.zeros(2, np.float32)
    po.lib().oc_spinor_prod(V, po.ptr(a), po.ptr(b), po.ptr(z0))
    # each colour-vector product is a float sum of 6 float products, so each
    # side carries a rounding error of at most ~7 ulp of sum_i |a_i b_i|
    # (contracted or not); the sums over points are in double
    bound = 1e-6 * float(np.abs(a.astype(np.float64) * b.astype(np.float64)).sum())
    assert abs(z.re - z0[0]) <= bound and abs(z.im - z0[1]) <= bound
    # exact double result of the same products, for scale
    ex = (_cplx(a.astype(np.float64)).conj() * _cplx(b.astype(np.float64))).sum()
    assert abs(z.re - ex.real) <= bound and abs(z.im - ex.imag) <= bound
    ns = lat.lib.norm_square(V, 0, sa)
    assert abs(ns - po.lib().oc_norm_square(V, po.ptr(a))) <= 1e-6 * ns
    lat.lib.mulc_spinor_add(V, 0, sa, sb, pyoqh.Complex(0.3, -1.1))
    ca, cb = _cplx(a.astype(np.float64)), _cplx(b.astype(np.float64))
    ref = ca + (0.3 - 1.1j) * cb
    got = _cplx(lat.get_s(sa).astype(np.float64))
    assert np.abs(got - ref).max() < 1e-5
    assert abs(lat.lib.unorm(V, 0, sb) - np.sqrt((b.astype(np.float64) ** 2).sum(1)).max()) < 1e-5
    lat.free(sa)
    lat.free(sb)


@pytest.mark.gpu
def test_assignments(lat):
    rng = np.random.default_rng(7)
    V = lat.V
    a = rng.standard_normal((V, 24))
    b = rng.standard_normal((V, 24))
    da, db, dr = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
    s1, s2 = lat.alloc_s(), lat.alloc_s()
    lat.put_sd(da, a)
    lat.put_sd(db, b)
    lat.lib.assign_sd2s(V, 0, da, s1)
    assert np.array_equal(lat.get_s(s1), a.astype(np.float32))
    lat.lib.diff_sd2s(V, 0, da, db, s2)
    assert np.array_equal(lat.get_s(s2), (a - b).astype(np.float32))
    lat.lib.assign_sd2sd(V, 0, da, dr)
    lat.lib.add_s2sd(V, 0, s2, dr)
    assert np.array_equal(lat.get_sd(dr), a + (a - b).astype(np.float32).astype(np.float64))
    lat.lib.diff_s2s(V, 0, s1, s2)  # s2 = s1 - s2
    assert np.array_equal(lat.get_s(s2), a.astype(np.float32) - (a - b).astype(np.float32))
    lat.lib.combine_spinor_dble(V, 0, da, db, 2.0, -0.5)
    assert np.allclose(lat.get_sd(da), 2.0 * a - 0.5 * b, rtol=0, atol=1e-14)
    assert abs(lat.lib.unorm_dble(V, 0, db) - np.sqrt((b ** 2).sum(1)).max()) < 1e-14
    for p in (da, db, dr, s1, s2):
        lat.free(p)


@pytest.mark.gpu
def test_host_qsum_arithmetic():
    lib = pyoqh.load()
    lib.acc_qflt.argtypes = [C.c_double, P]
    q = np.zeros(2)
    vals = [1e16, 1.0, -1e16, 3.5e-17, 2.0]
    for v in vals:
        lib.acc_qflt(v, q.ctypes.data_as(P))
    assert q[0] + q[1] == math.fsum(vals)
