"""Batched Dirac operator (oqh_Dw_dble_multi, oqh_Dw_multi): each of the n
fields gets what Dw_dble / Dw gives it alone (the per-field arithmetic
is the same; only the tile size differs, which changes the order in which the
hopping terms of a point are added), for every batch size and both boundary
conditions, on identical seeded fields; and the batch of 12 agrees with the
oracle's Dw_dble and Dw point by point (1e-13, 1e-6)."""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import make_oracle, site_rel

P = C.c_void_p


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [3, 0])
def test_multi_rhs_matches_single(bc):
    L = (8, 8, 8, 8)
    o, eta, _ = make_oracle(L, None, bc=bc, m0=-0.3, csw=1.77, seed=7)
    lat = pyoqh.Lattice(L, None, bc=bc, m0=-0.3, csw=1.77)
    try:
        lat.import_ud(o.ud())
        assert lat.sw_term(3) == 0
        V = lat.V
        rng = np.random.default_rng(5)
        src = [rng.standard_normal((V, 24)) for _ in range(12)]
        for s in src:
            o.bnd_sd2zero(s)
        for prec, tol in ((1, 1e-14), (0, 2e-6)):
            alloc = lat.alloc_sd if prec else lat.alloc_s
            put = lat.put_sd if prec else lat.put_s
            get = lat.get_sd if prec else lat.get_s
            single = lat.lib.Dw_dble if prec else lat.lib.Dw
            multi = lat.lib.oqh_Dw_dble_multi if prec else lat.lib.oqh_Dw_multi
            ins = [alloc() for _ in range(12)]
            outs = [alloc() for _ in range(12)]
            ref = alloc()
            for f, s in zip(ins, src):
                put(f, s if prec else s.astype(np.float32))
            for n in (1, 2, 3, 4, 5, 8, 12):
                arr_s = (P * n)(*ins[:n])
                arr_r = (P * n)(*outs[:n])
                multi(0.05, n, arr_s, arr_r)
                for k in range(n):
                    single(0.05, ins[k], ref)
                    rel, absz = site_rel(get(outs[k]), get(ref))
                    assert rel < tol and absz == 0.0, (prec, n, k, rel)
            # the batch of 12 against the oracle's Dw_dble / Dw per field
            # (1e-13 and 1e-6 per point, as the single-field operators)
            o.sw_term(3)
            if not prec:
                o.assign_ud2u()
                o.assign_swd2sw()
            for k in range(12):
                if prec:
                    oref = o.Dw_dble(0.05, src[k].copy())
                else:
                    oref = o.Dw(np.float32(0.05), src[k].astype(np.float32))
                rel, absz = site_rel(get(outs[k]), oref)
                assert rel < (1e-13 if prec else 1e-6) and absz == 0.0, (prec, k, rel)
            print("bc", bc, "multi-RHS agrees with single-field operator and the oracle")
    finally:
        lat.close()
