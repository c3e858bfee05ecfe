"""GPU network of the learned preconditioner (ptc.hip) against the numpy
oracle (oracle/ptc.py, complex128): forward pass and gradients through the
C-ABI.  The device rounds the transported fields and the weights to BF16 and
accumulates in FP32, so outputs and gradients agree with the oracle to a
norm-relative 2e-2 (a layer alone to 1e-2); the oracle itself is pinned by
gauge covariance, the adjoint identity and finite differences
(tests/test_oracle_ptc.py).  Training and the preconditioned solve are
judged by the loss decrease and by the double-precision true residual."""
import ctypes as C

import numpy as np
import pytest

import ptc
import pyoqh
import pyoracle as po
from helpers import link_field

P = C.c_void_p


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def weights_flat(Ws):
    return np.ascontiguousarray(np.concatenate([W.ravel() for W in Ws]).view(np.float64))


def ptrs(fields):
    return (C.c_void_p * len(fields))(*fields)


CFGS = [dict(name="8^4-periodic", L=(8, 8, 8, 8), bc=3, seed=1),
        dict(name="8^4-open", L=(8, 8, 8, 8), bc=0, seed=2),
        dict(name="16x8^3-periodic", L=(16, 8, 8, 8), bc=3, seed=3)]


@pytest.fixture(params=CFGS, ids=lambda c: c["name"])
def cfg(request):
    c = request.param
    print("config", c, flush=True)
    o = po.Oracle(c["L"], None, bc=c["bc"], m0=-0.35, csw=1.77)
    o.start_ranlux(0, c["seed"])
    o.random_ud()
    iup, idn = o.iupdn()
    ud = o.ud()
    lat = pyoqh.Lattice(c["L"], None, bc=c["bc"], m0=-0.35, csw=1.77)
    lat.import_ud(ud)
    yield c, lat, link_field(ud, iup), iup, idn
    lat.lib.oqh_nn_free()
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("widths,tol", [([4, 4], 1e-2), ([4, 12, 16, 12, 4], 2e-2), ([4, 8, 16, 4], 2e-2)])
def test_forward(cfg, widths, tol):
    c, lat, U, iup, idn = cfg
    lib = lat.lib
    rng = np.random.default_rng(7)
    Ws = ptc.init_weights(widths, rng)
    w = np.array(widths, np.int32)
    assert lib.oqh_nn_create(len(widths) - 1, w.ctypes.data_as(P)) == 0
    assert lib.oqh_nn_nweights() == sum(W.size for W in Ws)
    lib.oqh_nn_set_weights(weights_flat(Ws).ctypes.data_as(P))
    back = np.empty(2 * lib.oqh_nn_nweights())
    lib.oqh_nn_get_weights(back.ctypes.data_as(P))
    assert np.abs(back - weights_flat(Ws)).max() < 1e-7
    s = rng.standard_normal((lat.V, 24)).astype(np.float32)
    a, b = lat.alloc_s(), lat.alloc_s()
    lat.put_s(a, s)
    lib.oqh_nn_apply(a, b)
    got = ptc.spinor_to_field(lat.get_s(b))
    ref = ptc.net_fwd(Ws, ptc.spinor_to_field(s), U, iup, idn)[-1]
    r = rel(got, ref)
    print("widths", widths, "forward deviation", r)
    assert r < tol
    # the preconditioner's form: phi += NN(r), BF16 activations between layers
    phi0 = rng.standard_normal((lat.V, 24)).astype(np.float32)
    lat.put_s(b, phi0)
    lib.oqh_nn_correct(a, b)
    got2 = ptc.spinor_to_field(lat.get_s(b) - phi0)
    r2 = rel(got2, ref)
    print("widths", widths, "preconditioner-form deviation", r2)
    assert r2 < 2 * tol
    lat.free(a)
    lat.free(b)


@pytest.mark.gpu
def test_gradients(cfg):
    c, lat, U, iup, idn = cfg
    lib = lat.lib
    rng = np.random.default_rng(8)
    widths = [4, 12, 16, 12, 4]
    Ws = ptc.init_weights(widths, rng)
    w = np.array(widths, np.int32)
    assert lib.oqh_nn_create(len(widths) - 1, w.ctypes.data_as(P)) == 0
    lib.oqh_nn_set_weights(weights_flat(Ws).ctypes.data_as(P))
    n = 2
    rs = [rng.standard_normal((lat.V, 24)).astype(np.float32) for _ in range(n)]
    es = [rng.standard_normal((lat.V, 24)).astype(np.float32) for _ in range(n)]
    fr, fe = [lat.alloc_s() for _ in range(n)], [lat.alloc_s() for _ in range(n)]
    for k in range(n):
        lat.put_s(fr[k], rs[k])
        lat.put_s(fe[k], es[k])
    dw = np.empty(2 * lib.oqh_nn_nweights())
    loss = lib.oqh_nn_grad(n, ptrs(fr), ptrs(fe), dw.ctypes.data_as(P))
    # oracle: mean loss and its gradient
    lo, dref = 0.0, None
    for k in range(n):
        acts = ptc.net_fwd(Ws, ptc.spinor_to_field(rs[k]), U, iup, idn)
        l, g = ptc.loss_and_grad(acts[-1], ptc.spinor_to_field(es[k]))
        dW, _ = ptc.net_bwd(Ws, acts, g, U, iup, idn)
        lo += l / n
        flat = np.concatenate([d.ravel() for d in dW]) / n
        dref = flat if dref is None else dref + flat
    got = dw.view(np.complex128)
    print("loss", loss, lo, "gradient deviation", rel(got, dref))
    assert abs(loss - lo) < 2e-2 * lo
    assert rel(got, dref) < 3e-2
    # per layer as well
    off = 0
    for W in Ws:
        sl = slice(off, off + W.size)
        assert rel(got[sl], dref[sl]) < 5e-2
        off += W.size
    for f in fr + fe:
        lat.free(f)


@pytest.mark.gpu
def test_training_and_solve():
    """Adam steps on pairs (r = Dw v, e = v) lower the loss on held-out
    pairs; sap_gcr with the trained network in the preconditioner (FP32 and
    FP16 Schwarz cycles) reaches the double-precision true residual."""
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        rng = np.random.default_rng(5)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        lib.fwd_euler(20, 0.02)
        lib.set_ud_phase()
        assert lat.sw_term(3) == 0
        widths = [4, 12, 16, 12, 4]
        w = np.array(widths, np.int32)
        assert lib.oqh_nn_create(4, w.ctypes.data_as(P)) == 0
        Ws = ptc.init_weights(widths, rng, scale=0.02)
        Ws[0][0, :4, :4] += np.eye(4)
        lib.oqh_nn_set_weights(weights_flat(Ws).ctypes.data_as(P))

        def pairs(n):
            r, e = [lat.alloc_s() for _ in range(n)], [lat.alloc_s() for _ in range(n)]
            for k in range(n):
                lat.put_s(e[k], rng.standard_normal((lat.V, 24)).astype(np.float32))
                lib.Dw(0.0, e[k], r[k])
            return r, e

        tr, te = pairs(4)
        hr, he = pairs(2)
        dw = np.empty(2 * lib.oqh_nn_nweights())
        held0 = lib.oqh_nn_grad(2, ptrs(hr), ptrs(he), dw.ctypes.data_as(P))
        losses = [lib.oqh_nn_train_step(4, ptrs(tr), ptrs(te), 2e-3) for _ in range(40)]
        held1 = lib.oqh_nn_grad(2, ptrs(hr), ptrs(he), dw.ctypes.data_as(P))
        print("training loss", losses[0], losses[-1], "held out", held0, held1)
        assert losses[-1] < 0.8 * losses[0] and held1 < 0.8 * held0
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        eta = rng.standard_normal((lat.V, 24))
        for prec in (0, 1):
            lib.oqh_set_sap_prec(prec)
            for nn in (0, 1):
                lib.oqh_set_nn_prec(nn)
                e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
                lat.put_sd(e, eta)
                ifail, status = C.c_int(), C.c_int()
                lib.sap_gcr(16, 256, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
                lib.Dw_dble(0.0, p, r)
                res = np.linalg.norm(eta - lat.get_sd(r)) / np.linalg.norm(eta)
                print("prec", prec, "network", nn, "status", status.value, "true residual", res)
                assert ifail.value == 0 and res <= 1.2e-10
                for f in (e, p, r):
                    lat.free(f)
        lib.oqh_set_nn_prec(0)
        lib.oqh_set_sap_prec(0)
    finally:
        lat.lib.oqh_nn_free()
        lat.close()


@pytest.mark.gpu
def test_learned_fallback_budget():
    """Plain SAP as the fallback of the learned preconditioner: with weights
    that spoil the preconditioner the learned solve exhausts its half of the
    Krylov budget, plain SAP+GCR continues from its solution, and the call
    uses at most nmx vectors in total (status counts both parts) and reaches
    the true residual; with nmx = 1 the call fails with ifail = -1 and
    status = nmx (sap/sap_gcr.c:223-228)."""
    lat = pyoqh.Lattice((8, 8, 8, 8), None, bc=3, m0=-0.35, csw=1.77)
    try:
        lib = lat.lib
        rng = np.random.default_rng(9)
        lat.import_ud(pyoqh.haar_su3(4 * lat.V, rng))
        lib.fwd_euler(20, 0.02)
        lib.set_ud_phase()
        assert lat.sw_term(3) == 0
        widths = [4, 8, 4]
        w = np.array(widths, np.int32)
        assert lib.oqh_nn_create(2, w.ctypes.data_as(P)) == 0
        lib.oqh_nn_set_weights(weights_flat(ptc.init_weights(widths, rng, scale=30.0)).ctypes.data_as(P))
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        eta = rng.standard_normal((lat.V, 24))
        e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        lat.put_sd(e, eta)
        ifail, status = C.c_int(), C.c_int()
        lib.oqh_set_nn_prec(0)
        lib.sap_gcr(16, 256, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
        plain = status.value
        assert ifail.value == 0
        nmx = 2 * plain + 10
        lib.oqh_set_nn_prec(1)
        lib.sap_gcr(16, nmx, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
        lib.Dw_dble(0.0, p, r)
        res = np.linalg.norm(eta - lat.get_sd(r)) / np.linalg.norm(eta)
        print("plain", plain, "nmx", nmx, "learned+fallback status", status.value, "ifail", ifail.value,
              "true residual", res, flush=True)
        assert ifail.value == 0 and res <= 1.2e-10
        assert nmx // 2 < status.value <= nmx, "the fallback must have run within the budget"
        lib.sap_gcr(16, 1, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
        assert ifail.value == -1 and status.value == 1
        assert np.array_equal(lat.get_sd(e), eta)
        for f in (e, p, r):
            lat.free(f)
        lib.oqh_set_nn_prec(0)
    finally:
        lat.lib.oqh_nn_free()
        lat.close()
