"""Training of the learned preconditioner of sap_gcr on the GPU (host-side
orchestration of the library's own kernels: pair generation by the solver,
oqh_nn_train_step for the forward/backward passes and Adam).

A pair is (r, e): r the residual that the Schwarz cycles of the
preconditioner leave of a Gaussian source eta, e the error they leave of
the solution D^-1 eta (solved by SAP+GCR to 1e-12), so that the network is
trained to correct what SAP leaves: phi = SAP(eta) + NN(r) approximates
D^-1 eta.  No counterpart in the reference.
"""
import ctypes as C
import time

import numpy as np

import pyoqh

P = C.c_void_p


def _ptrs(fields):
    return (C.c_void_p * len(fields))(*fields)


def train_preconditioner(widths=(4, 8, 4), L=(16, 16, 16, 16), Ltrd=(8, 8, 8, 8), seed=1, nflow=50,
                         m0=-0.35, csw=1.77, npairs=32, nheld=4, steps=1000, batch=8, lr=(1e-2, 2e-3),
                         init_scale=0.01, log=None):
    """Trains a network of the given channel widths on the lattice L
    (periodic bc, random_ud of the seed flowed nflow x 0.01, quark phase,
    SAP 4^4 isolv=1 nmr=4 ncy=5).  npairs pairs are generated, the last nheld
    are held out; steps Adam steps on random batches, learning rate lr[0]
    for the first half and lr[1] after.  Returns (weights, record): the
    weights as 2 * oqh_nn_nweights() doubles and the losses and timings."""
    widths = list(widths)
    rng = np.random.default_rng(seed)
    lat = pyoqh.Lattice(L, Ltrd, bc=3, m0=m0, csw=csw)
    rec = {"lattice": "x".join(str(x) for x in L), "pairs": npairs, "held_out": nheld, "adam_steps": steps,
           "batch": batch, "widths": widths}
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.sap.argtypes = [C.c_float, C.c_int, C.c_int, P, P]
        V = lat.V
        lib.start_ranlux(0, seed)
        lib.random_ud()
        lib.fwd_euler(nflow, 0.01)
        lib.set_ud_phase()
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        t0 = time.perf_counter()
        e_, p_ = lat.alloc_sd(), lat.alloc_sd()
        psi, rho = lat.alloc_s(), lat.alloc_s()
        fr, fe = [], []
        for _ in range(npairs):
            lib.random_sd(lat.VT, 2, e_, 1.0)
            r = lat.get_sd(e_)
            ifail, st = C.c_int(), C.c_int()
            lib.sap_gcr(24, 1024, 0, 1e-12, 0.0, e_, p_, C.byref(ifail), C.byref(st))
            e = lat.get_sd(p_)
            lat.put_s(psi, np.zeros((V, 24), np.float32))
            lat.put_s(rho, r.astype(np.float32))
            for _ in range(5):
                lib.sap(0.0, 1, 4, psi, rho)
            a, b = lat.alloc_s(), lat.alloc_s()
            lat.put_s(a, lat.get_s(rho))
            lat.put_s(b, (e - lat.get_s(psi)).astype(np.float32))
            fr.append(a)
            fe.append(b)
        rec["pair_generation_s"] = round(time.perf_counter() - t0, 2)
        w = np.array(widths, np.int32)
        if lib.oqh_nn_create(len(widths) - 1, w.ctypes.data_as(P)):
            raise RuntimeError("oqh_nn_create failed for widths %s" % widths)
        nw = lib.oqh_nn_nweights()
        w0 = init_scale * rng.standard_normal(2 * nw)
        lib.oqh_nn_set_weights(w0.ctypes.data_as(P))
        dw = np.empty(2 * nw)
        ntr = npairs - nheld
        hold = (_ptrs(fr[ntr:]), _ptrs(fe[ntr:]))
        rec["held_out_loss_before"] = round(lib.oqh_nn_grad(nheld, *hold, dw.ctypes.data_as(P)), 4)
        lat.sync()
        t0 = time.perf_counter()
        for it in range(steps):
            idx = rng.choice(ntr, batch, replace=False)
            lib.oqh_nn_train_step(batch, _ptrs([fr[i] for i in idx]), _ptrs([fe[i] for i in idx]),
                                  lr[0] if it < steps // 2 else lr[1])
        lat.sync()
        rec["training_s"] = round(time.perf_counter() - t0, 2)
        rec["held_out_loss_after"] = round(lib.oqh_nn_grad(nheld, *hold, dw.ctypes.data_as(P)), 4)
        wts = np.empty(2 * nw)
        lib.oqh_nn_get_weights(wts.ctypes.data_as(P))
        if log:
            log("network trained, held-out loss %.3f -> %.3f" % (rec["held_out_loss_before"],
                                                               rec["held_out_loss_after"]))
    finally:
        lat.lib.oqh_nn_free()
        lat.close()
    return wts, rec
