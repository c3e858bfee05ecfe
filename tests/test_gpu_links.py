"""The operators rebuild the third row of every link from the first two
(row 2 = +-conj(row0 x row1)) when the whole gauge field allows it, and read
all nine entries otherwise.  Both paths must give the oracle's Dw_dble and
Dwhat_dble to 1e-13 per point: on an SU(3) field, on the same field with the
quark phase (some links multiplied by -1), on a field whose links are scaled
(not SU(3)) and on an SU(3) field with a single perturbed link."""
import numpy as np
import pytest

import pyoqh
from helpers import make_oracle, site_rel

TOL64 = 1e-13


def check(o, lat, eta, mu):
    o.sw_term(3)
    assert lat.sw_term(3) == 0
    s, r = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(s, eta)
    lat.lib.Dw_dble(mu, s, r)
    rel, absz = site_rel(lat.get_sd(r), o.Dw_dble(mu, eta.copy()))
    assert rel < TOL64 and absz == 0.0, rel
    o.sw_term(2)
    assert lat.sw_term(2) == 0
    lat.lib.Dwhat_dble(mu, s, r)
    h = lat.V // 2
    ref = o.Dwhat_dble(mu, eta.copy())
    rel2, absz2 = site_rel(lat.get_sd(r)[:h], ref[:h])
    assert rel2 < TOL64 and absz2 == 0.0, rel2
    lat.free(s)
    lat.free(r)
    return rel, rel2


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [3, 0])
def test_rebuilt_and_full_links(bc):
    print("bc", bc, flush=True)
    o, eta, s32 = make_oracle((8, 8, 8, 8), None, bc=bc, m0=-0.3, csw=1.77, seed=4)
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=bc, m0=-0.3, csw=1.77)
    try:
        ud = o.ud()
        # with periodic bc the oracle's random field carries the quark phase
        # (time links at x0 = N0-1 times -1): links of both signs
        z = (ud[:, 0::2] + 1j * ud[:, 1::2]).reshape(-1, 3, 3)
        c = np.conj(np.cross(z[:, 0], z[:, 1]))
        zero = np.abs(z).max(axis=(1, 2)) == 0.0
        nneg = int(((np.abs(z[:, 2] + c).max(axis=1) < 1e-13) & ~zero).sum())
        # open bc: the 512 time links leaving the last time slice vanish
        assert nneg == (512 if bc == 3 else 0) and zero.sum() == (0 if bc == 3 else 512)
        lat.import_ud(ud)
        print("su3", nneg, "negated links", check(o, lat, eta, 0.1))
        # scaled links: not SU(3), all nine entries must be read
        o.set_ud(0.9 * ud)
        lat.import_ud(0.9 * ud)
        print("scaled", check(o, lat, eta, 0.0))
        # one perturbed link in an otherwise SU(3) field
        ud2 = ud.copy()
        ud2[123, 16] += 1e-6
        o.set_ud(ud2)
        lat.import_ud(ud2)
        print("one link", check(o, lat, eta, 0.0))
        # back to SU(3): the rebuilt path again
        o.set_ud(ud)
        lat.import_ud(ud)
        print("su3 again", check(o, lat, eta, 0.0))
    finally:
        lat.close()


@pytest.mark.gpu
def test_rebuild_threshold():
    """Row 2 is rebuilt only while every stored row 2 is within 64 ulp
    (1.4e-14) of +-conj(row0 x row1): one link perturbed by half of that
    keeps the rebuilt path, by twice that switches to nine-entry reads; in
    both cases Dw_dble and Dwhat_dble agree with the oracle on the stored
    field to 1e-14 per point."""
    import ctypes as C
    o, eta, s32 = make_oracle((8, 8, 8, 8), None, bc=3, m0=-0.3, csw=1.77, seed=5)
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3, m0=-0.3, csw=1.77)
    try:
        lib = lat.lib
        lib.oqh_link_rebuild.restype = C.c_int
        lib.oqh_link_rebuild.argtypes = [C.c_void_p]
        ud = o.ud()
        thr = 64.0 * np.finfo(np.float64).eps
        for f, rebuilt in ((0.0, 1), (0.5, 1), (2.0, 0)):
            ud2 = ud.copy()
            ud2[77, 12] += f * thr          # row 2, real part of column 0
            o.set_ud(ud2)
            lat.import_ud(ud2)
            dev = C.c_double()
            got = lib.oqh_link_rebuild(C.byref(dev))
            rels = check(o, lat, eta, 0.05)
            print("perturbation", f * thr, "rebuilt", got, "deviation", dev.value, "parity", rels, flush=True)
            assert got == rebuilt
            assert max(rels) < 1e-14
            if f > 0:
                assert abs(dev.value - f * thr) < 0.25 * thr
    finally:
        lat.close()
