"""GPU Wilson flow (wflow.hip) against the oracle: plaquette force
(forces/force0.c:281), forward-Euler steps (wflow/wflow.c:217), the quark
phase of periodic bc (uflds/uflds.c:338), and the Dirac operator on the
flowed field.  The oracle's force and flow agree with the reference build to
the last bit (tests/test_oracle_ref.py); the device sums the six plaquette
terms of a link in another order, so forces agree to rounding (1e-14 of the
force scale) and the flowed links to 1e-13."""
import numpy as np
import pytest

import pyoqh
import pyoracle as po
from helpers import plaquette_sum, site_rel

CFGS = [
    dict(name="8^4-periodic", L=(8, 8, 8, 8), T=None, bc=3, seed=1),
    dict(name="8^4-open", L=(8, 8, 8, 8), T=None, bc=0, seed=2),
    dict(name="8^4-nthread16", L=(8, 8, 8, 8), T=(4, 4, 4, 4), bc=3, seed=4),
    dict(name="16x8^3-open", L=(16, 8, 8, 8), T=None, bc=0, seed=5),
]


@pytest.fixture(params=CFGS, ids=lambda c: c["name"])
def cfg(request):
    c = request.param
    print("config", c, flush=True)
    o = po.Oracle(c["L"], c["T"], bc=c["bc"], m0=-0.3, csw=1.77)
    o.start_ranlux(0, c["seed"])
    o.random_ud()
    lat = pyoqh.Lattice(c["L"], c["T"], bc=c["bc"], m0=-0.3, csw=1.77)
    lat.import_ud(o.ud())
    yield c, o, lat
    lat.close()


@pytest.mark.gpu
def test_plaquette_force(cfg):
    c, o, lat = cfg
    lat.lib.plaq_frc()
    got = lat.export_frc()
    ref = o.plaq_frc()
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    print("force max deviation / scale", err / scale)
    assert err < 1e-14 * scale


@pytest.mark.gpu
def test_euler_steps_and_phase(cfg):
    c, o, lat = cfg
    lat.lib.fwd_euler(4, 0.02)
    o.fwd_euler(4, 0.02)
    err = np.abs(lat.export_ud() - o.ud()).max()
    print("flowed links max deviation", err)
    assert err < 1e-13
    # quark phase (periodic bc) and the operator on the flowed field
    flowed = o.ud()
    lat.import_ud(flowed)
    lat.lib.set_ud_phase()
    o.set_ud_phase()
    assert np.array_equal(lat.export_ud(), o.ud())
    lat.lib.set_ud_phase()  # no action when the phase is set
    assert np.array_equal(lat.export_ud(), o.ud())
    o.start_ranlux(0, 7)
    eta = o.random_sd()
    o.bnd_sd2zero(eta)
    o.sw_term(3)
    assert lat.sw_term(3) == 0
    s, r = lat.alloc_sd(), lat.alloc_sd()
    lat.put_sd(s, eta)
    lat.lib.Dw_dble(0.0, s, r)
    rel, absz = site_rel(lat.get_sd(r), o.Dw_dble(0.0, eta.copy()))
    print("Dw_dble on the flowed field", rel)
    assert rel < 1e-13 and absz == 0.0
    # unset_ud_phase restores the field (a second call does nothing)
    lat.lib.unset_ud_phase()
    lat.lib.unset_ud_phase()
    assert np.array_equal(lat.export_ud(), flowed)
    lat.lib.set_ud_phase()
    assert np.array_equal(lat.export_ud(), o.ud())


@pytest.mark.gpu
def test_flow_16_4_invariants():
    """At 16^4 (periodic) the device flow increases the plaquette sum step by
    step and keeps the links unitary with unit determinant."""
    lat = pyoqh.Lattice((16, 16, 16, 16), (8, 8, 8, 8), bc=3)
    try:
        rng = np.random.default_rng(9)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        o = po.Oracle((16, 16, 16, 16), (8, 8, 8, 8), bc=3)
        iup, _ = o.iupdn()
        s = [plaquette_sum(lat.export_ud(), iup)]
        for _ in range(4):
            lat.lib.fwd_euler(5, 0.01)
            s.append(plaquette_sum(lat.export_ud(), iup))
        print("plaquette sums", s)
        assert all(b > a for a, b in zip(s, s[1:]))
        ud = lat.export_ud()
        m = (ud[:, 0::2] + 1j * ud[:, 1::2]).reshape(-1, 3, 3)
        assert np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - np.eye(3)).max() < 1e-13
        assert np.abs(np.linalg.det(m) - 1).max() < 1e-13
    finally:
        lat.close()
