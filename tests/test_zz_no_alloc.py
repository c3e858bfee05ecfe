"""The allocation-free variant of spinor_prod_dble
(oqh_spinor_prod_dble_noalloc, partial sums in a static device array)
against spinor_prod_dble (linalg/salg_dble.c) through the C ABI: the same
kernels and order, so the (hi, lo) results must be bitwise equal; two calls
in a row, the second on fewer points (a half field addressed with a pointer
offset), so stale partials of the first call must not leak into the second.
The refusals (host arrays, too many points, output not in device memory) are
recorded errors."""
import ctypes as C

import numpy as np
import pytest

import pyoqh

P = C.c_void_p


def _out(lat, f):
    """complex_qflt written at the base of device field f: the first two
    complex units of the field's site-block layout, i.e. component 0 of
    points 0 and 1."""
    g = lat.get_sd(f)
    return (g[0, 0], g[0, 1]), (g[1, 0], g[1, 1])


@pytest.mark.gpu
def test_noalloc_spinor_prod_dble():
    lat = pyoqh.Lattice((16, 8, 8, 8), (8, 8, 8, 8), bc=3, m0=-0.3, csw=1.77)
    try:
        lib = lat.lib
        lib.spinor_prod_dble.restype = pyoqh.ComplexQflt
        lib.spinor_prod_dble.argtypes = [C.c_int, C.c_int, P, P]
        lib.oqh_spinor_prod_dble_noalloc.restype = C.c_int
        lib.oqh_spinor_prod_dble_noalloc.argtypes = [C.c_int, C.c_int, P, P, P]
        V = lat.V
        rng = np.random.default_rng(21)
        a, b, out = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(a, rng.standard_normal((V, 24)))
        lat.put_sd(b, rng.standard_normal((V, 24)))
        lat.put_sd(out, np.zeros((V, 24)))
        half = (V // 2) * 192
        for vol, s, r in ((V, a, b), (V // 2 - 37, a + half, b + half)):
            print("vol", vol, flush=True)
            ref = lib.spinor_prod_dble(vol, 1, s, r)
            assert lib.oqh_spinor_prod_dble_noalloc(vol, 1, s, r, out) == 0
            lib.oqh_sync()
            re, im = _out(lat, out)
            print("noalloc", re, im, "spinor_prod_dble", tuple(ref.re.q), tuple(ref.im.q), flush=True)
            assert re == tuple(ref.re.q) and im == tuple(ref.im.q)
        # refusals, recorded without abort
        lib.oqh_set_error_mode(0)
        host = np.zeros((V, 24))
        assert lib.oqh_spinor_prod_dble_noalloc(V, 1, host.ctypes.data_as(P), b, out) == 1
        assert lib.oqh_spinor_prod_dble_noalloc(V // 2 + 1, 1, a + half, b, out) == 1
        z = np.zeros(4)
        assert lib.oqh_spinor_prod_dble_noalloc(V, 1, a, b, z.ctypes.data_as(P)) == 1
        lib.oqh_set_error_mode(1)
        for f in (a, b, out):
            lat.free(f)
    finally:
        lat.close()
