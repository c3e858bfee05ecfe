"""The block-grid set-up and map routines (block/blk_grid.c:139,
map_u2blk.c:308, map_sw2blk.c:102, map_s2blk.c:551-737) on the reference's
seeded 8^4 field with 4^4 SAP blocks, both boundary conditions:

* alloc_bgr and assign_ud2ubgr run and leave the block operators equal to
  the oracle's;
* assign_swd2swbgr(NO_PTS) then Dwoo_blk applies the uninverted odd SW term
  A_oo (against the oracle's apply_sw on the odd points); after
  assign_swd2swbgr(ODD_PTS) Dwoo_blk applies A_oo^-1, so that
  A_oo^-1 (A_oo x) = x on the odd points of the block;
* assign_sblk2s writes the block field back to exactly the points of the
  chosen set (ALL, EVEN, ODD) of the block and leaves every other point of
  the field as it was."""
import ctypes as C

import numpy as np
import pytest

import pyoqh
from helpers import make_oracle, site_rel

P = C.c_void_p
SAP_BLOCKS, ALL_PTS, EVEN_PTS, ODD_PTS, NO_PTS = 0, 0, 1, 2, 3


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [3, 0])
def test_block_maps(bc):
    L = (8, 8, 8, 8)
    o, eta, s32 = make_oracle(L, L, bc=bc, m0=-0.3, csw=1.77, seed=3)
    lat = pyoqh.Lattice(L, L, bc=bc, m0=-0.3, csw=1.77)
    try:
        lib = lat.lib
        lib.oqh_set_sap_parms.argtypes = [P, C.c_int, C.c_int, C.c_int]
        lib.alloc_bgr.argtypes = [C.c_int]
        lib.assign_ud2ubgr.argtypes = [C.c_int]
        lib.assign_swd2swbgr.restype = C.c_int
        lib.assign_swd2swbgr.argtypes = [C.c_int, C.c_int]
        lib.assign_s2sblk.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_int]
        lib.assign_sblk2s.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P]
        lib.Dwoo_blk.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
        lib.Dw_blk.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
        lib.oqh_export_sblk.argtypes = [C.c_int, C.c_int, P]
        lat.import_ud(o.ud())
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 0, 4, 5)
        o.set_sap((4, 4, 4, 4), 0, 4, 5)
        o.sw_term(3)
        assert o.sap_prepare() == 0
        lib.alloc_bgr(SAP_BLOCKS)
        lib.assign_ud2ubgr(SAP_BLOCKS)
        assert lib.assign_swd2swbgr(SAP_BLOCKS, NO_PTS) == 0

        def block(n, k):
            a = np.empty((256, 24), np.float32)
            lib.oqh_export_sblk(n, k, a.ctypes.data_as(P))
            return a

        s = lat.alloc_s()
        lat.put_s(s, s32)
        nb = o.nblocks()
        for n in (0, nb - 1):
            imb = o.block_imb(n)
            # the block operator on the uninverted grid equals the oracle's
            lib.assign_s2sblk(SAP_BLOCKS, n, ALL_PTS, s, 1)
            lib.Dw_blk(SAP_BLOCKS, n, 0.0, 1, 2)
            ref = o.Dw_blk(n, np.float32(0.0), np.vstack([s32[imb], np.zeros((1, 24), np.float32)]))[:-1]
            rel, absz = site_rel(block(n, 2), ref)
            assert rel < 1e-5 and absz == 0.0
            # A_oo then A_oo^-1 on the odd points
            x = block(n, 1)
            lib.Dwoo_blk(SAP_BLOCKS, n, 0.0, 1, 2)
            y = block(n, 2)
            _, ref = o.blk_eo_op(n, 1, np.float32(0.0), x, np.zeros_like(x))
            rel, absz = site_rel(y[128:], ref[128:])
            assert rel < 1e-5 and absz == 0.0
            assert lib.assign_swd2swbgr(SAP_BLOCKS, ODD_PTS) == 0
            lib.Dwoo_blk(SAP_BLOCKS, n, 0.0, 2, 0)
            back = block(n, 0)
            # open bc: the boundary time slices are zero in and out
            lex = np.empty(o.V, np.int64)
            lex[o.ipt()] = np.arange(o.V)
            x0 = lex // (L[1] * L[2] * L[3])
            keep = np.ones(128, bool) if bc == 3 else ~np.isin(x0[imb[128:]], (0, L[0] - 1))
            rel, _ = site_rel(back[128:][keep], x[128:][keep])
            print("block", n, "A_oo^-1 A_oo x", rel, flush=True)
            assert rel < 1e-5
            assert lib.assign_swd2swbgr(SAP_BLOCKS, NO_PTS) == 0
            # assign_sblk2s writes exactly the chosen points
            other = np.random.default_rng(n + 1).standard_normal(s32.shape).astype(np.float32)
            for st, sl in ((ALL_PTS, slice(0, 256)), (EVEN_PTS, slice(0, 128)), (ODD_PTS, slice(128, 256))):
                t = lat.alloc_s()
                lat.put_s(t, other)
                lib.assign_sblk2s(SAP_BLOCKS, n, st, 1, t)
                got = lat.get_s(t)
                pts = imb[sl]
                assert np.array_equal(got[pts], x[sl])
                rest = np.ones(len(got), bool)
                rest[pts] = False
                assert np.array_equal(got[rest], other[rest])
                lat.free(t)
        lat.free(s)
    finally:
        lat.close()
