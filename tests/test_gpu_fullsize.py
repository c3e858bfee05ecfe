"""Parity at the benchmark's full size: the Dw_dble workload of bench.py
(32^3x64, open bc with cF=1, csw=1.77, m0=-0.3, 100 Wilson-flow steps of
0.01) on the reference's own random field (random_ud after
start_ranlux(0, 2), the oracle's restatement of the reference).  The flow
runs on the device; the oracle then builds the SW term and applies Dw_dble
to the flowed field (about 20 s of CPU time), and every point must agree to
1e-13 relative, as in the small-lattice tests."""
import numpy as np
import pytest

import pyoqh
import pyoracle as po
from helpers import site_rel


@pytest.mark.gpu
def test_dw_dble_32x32x32x64_flowed():
    L, T = (64, 32, 32, 32), (8, 8, 8, 8)
    print("oracle: random_ud at 32^3x64", flush=True)
    o = po.Oracle(L, T, bc=0, m0=-0.3, csw=1.77)
    o.start_ranlux(0, 2)
    o.random_ud()
    lat = pyoqh.Lattice(L, T, bc=0, m0=-0.3, csw=1.77)
    try:
        lat.import_ud(o.ud())
        lat.lib.fwd_euler(100, 0.01)
        flowed = lat.export_ud()
        o.set_ud(flowed)
        print("oracle: SW term and Dw_dble", flush=True)
        o.sw_term(3)
        assert lat.sw_term(3) == 0
        eta = np.random.default_rng(2).standard_normal((lat.V, 24))
        o.bnd_sd2zero(eta)
        ref = o.Dw_dble(0.0, eta.copy())
        s, r = lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(s, eta)
        lat.lib.Dw_dble(0.0, s, r)
        rel, absz = site_rel(lat.get_sd(r), ref)
        print("Dw_dble 32^3x64 max per-point deviation", rel)
        assert rel < 1e-13 and absz == 0.0
    finally:
        lat.close()
