"""The oracle's plaquette sums (uflds/plaq_sum.c) and its encoders of
configuration and quark-field files (archive/archive.c:804-870,
archive/sarchive.c:424-503, main/README.io) against the
reference build (bit-exact: file bytes and sums on the reference's own field)
and against what the mathematics fixes: the brute-force plaquette sum of the
test helpers, closed forms on the unit field, gauge invariance."""
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import gauge_transform, plaquette_sum

CASES = [("8x8x8x8_t4", (8, 8, 8, 8), (4, 4, 4, 4), 0),
         ("8x8x8x8_t4", (8, 8, 8, 8), (4, 4, 4, 4), 3),
         ("16x8x8x8_t8", (16, 8, 8, 8), (8, 8, 8, 8), 0),
         ("16x8x8x8_t8", (16, 8, 8, 8), (8, 8, 8, 8), 3)]


@pytest.mark.parametrize("tag,L,T,bc", CASES)
def test_against_reference_export(tmp_path, tag, L, T, bc):
    if po.ref_exe(tag) is None:
        pytest.skip("reference build not present")
    ld = po.ref_dump(tag, str(tmp_path), bc=bc, seed=5)
    raw = open(os.path.join(str(tmp_path), "cnfg.dat"), "rb").read()
    o = po.Oracle(L, T, bc=bc)
    ipt = ld("ipt.i32", np.int32)
    assert np.array_equal(ipt, o.ipt())
    nl, plaq, ud = po.cnfg_decode(raw, ipt)
    assert tuple(nl) == L
    # the dump's field carries the phase (time links at x0 = N0-1 times -1
    # for bc 3), the exported one does not
    ud_dump = ld("ud.f64", np.float64, (-1, 18))
    if bc == 0:
        assert np.array_equal(ud, ud_dump)
    else:
        assert np.array_equal(np.abs(ud), np.abs(ud_dump))
        assert 0 < np.count_nonzero(ud != ud_dump) <= 18 * 2 * (np.prod(L) // L[0])
    # plaquette sums and the whole file from the oracle, on the same field
    o.set_ud(ud)
    pq = ld("plaq.f64", np.float64)
    assert o.plaq_sum(0) == pq[0] and o.plaq_sum(1) == pq[1]
    assert plaq == pq[0] / (6.0 * np.prod(L))
    assert o.cnfg_bytes() == raw


@pytest.mark.parametrize("bc", [0, 3])
def test_plaq_sum_brute_force_and_invariance(bc):
    L = (8, 4, 4, 8)
    o = po.Oracle(L, (4, 4, 4, 4), bc=bc)
    o.start_ranlux(0, 11)
    o.random_ud()
    ud = o.ud()
    iup, idn = o.iupdn()
    ref = plaquette_sum(ud, iup)  # open bc: the links leaving x0 = N0-1 vanish
    assert abs(o.plaq_sum(0) - ref) < 1e-12 * np.prod(L)
    # weight 1/2 of the space-like plaquettes on the boundary slices (open bc)
    x0 = np.indices(L).reshape(4, -1)[0][np.argsort(o.ipt())]
    from helpers import link_field
    U = link_field(ud, iup)
    wsum = 0.0
    for mu, nu in ((1, 2), (1, 3), (2, 3)):
        P = U[:, mu] @ U[iup[:, mu], nu] @ np.conj(np.transpose(U[iup[:, nu], mu], (0, 2, 1))) \
            @ np.conj(np.transpose(U[:, nu], (0, 2, 1)))
        tr = np.trace(P, axis1=1, axis2=2).real
        w = np.where((x0 == 0) | (x0 == L[0] - 1), 0.5 if bc == 0 else 1.0, 1.0)
        wsum += (w * tr).sum()
    for mu, nu in ((0, 1), (0, 2), (0, 3)):
        P = U[:, mu] @ U[iup[:, mu], nu] @ np.conj(np.transpose(U[iup[:, nu], mu], (0, 2, 1))) \
            @ np.conj(np.transpose(U[:, nu], (0, 2, 1)))
        wsum += np.trace(P, axis1=1, axis2=2).real.sum()
    assert abs(o.plaq_sum(1) - wsum) < 1e-12 * np.prod(L)
    # a random gauge transformation leaves both sums unchanged
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.standard_normal((o.V, 3, 3)) + 1j * rng.standard_normal((o.V, 3, 3)))
    o.set_ud(gauge_transform(ud, q, iup, idn))
    assert abs(o.plaq_sum(0) - ref) < 1e-12 * np.prod(L)
    assert abs(o.plaq_sum(1) - wsum) < 1e-12 * np.prod(L)


@pytest.mark.parametrize("bc", [0, 3])
def test_plaq_sum_unit_field(bc):
    L = (8, 4, 4, 4)
    V = int(np.prod(L))
    o = po.Oracle(L, (4, 4, 4, 4), bc=bc)
    iup, idn = o.iupdn()
    ud = np.zeros((4 * V, 18))
    ud[:, [0, 8, 16]] = 1.0
    if bc == 0:
        # links across the open boundary vanish (lattice/bcnds.c:409-457)
        x0 = np.indices(L).reshape(4, -1)[0][np.argsort(o.ipt())]
        h = V // 2
        for j in range(h):
            if x0[h + j] == L[0] - 1:
                ud[8 * j] = 0.0
            if x0[h + j] == 0:
                ud[8 * j + 1] = 0.0
    o.set_ud(ud)
    if bc == 3:
        assert o.plaq_sum(0) == 18.0 * V and o.plaq_sum(1) == 18.0 * V
    else:
        nt = V // L[0]
        assert o.plaq_sum(0) == 9.0 * (V - nt) + 9.0 * V
        assert o.plaq_sum(1) == 9.0 * (V - nt) + 9.0 * (V - nt)


def test_codec_round_trip_and_order():
    L = (4, 4, 4, 8)
    o = po.Oracle(L, (4, 4, 4, 4), bc=3)
    o.start_ranlux(0, 2)
    o.random_ud()
    data = o.cnfg_bytes()
    assert len(data) == 24 + 4 * np.prod(L) * 144
    nl, plaq, ud = po.cnfg_decode(data, o.ipt())
    assert tuple(nl) == L and np.array_equal(ud, o.ud())
    assert plaq == o.plaq_sum(0) / (6.0 * np.prod(L))
    # the first record is the odd point (0,0,0,1), the second (0,0,0,3)
    ipt = o.ipt()
    h = np.prod(L) // 2
    first = np.frombuffer(data[24:24 + 8 * 144], "<f8").reshape(8, 18)
    assert np.array_equal(first, o.ud()[8 * (ipt[1] - h):8 * (ipt[1] - h) + 8])
    second = np.frombuffer(data[24 + 8 * 144:24 + 16 * 144], "<f8").reshape(8, 18)
    assert np.array_equal(second, o.ud()[8 * (ipt[3] - h):8 * (ipt[3] - h) + 8])


@pytest.mark.parametrize("tag,L,T,bc", CASES)
def test_sfld_against_reference_export(tmp_path, tag, L, T, bc):
    if po.ref_exe(tag) is None:
        pytest.skip("reference build not present")
    ld = po.ref_dump(tag, str(tmp_path), bc=bc, seed=6)
    eta = ld("eta.f64", np.float64, (-1, 24))
    ipt = ld("ipt.i32", np.int32)
    V = int(np.prod(L))
    for eo in (0, 1):
        raw = open(os.path.join(str(tmp_path), "eta_sfld%d.dat" % eo), "rb").read()
        assert po.sfld_encode(L, ipt, eta, eo) == raw
        nl, eo_f, norm, sd = po.sfld_decode(raw, ipt, V)
        assert tuple(nl) == L and eo_f == eo
        n = V if eo == 0 else V // 2
        assert np.array_equal(sd[:n], eta[:n]) and np.isnan(sd[n:]).all()
        assert norm == po.norm_square_dble(np.ascontiguousarray(eta[:n]))[0]
