"""What the trained network gains: the learned preconditioner at the
configured 32^3x64 size on held-out fields.

The network (two layers 4 -> 8 -> 4, nntrain.train_preconditioner: trained
on the GPU by the library's forward/backward kernels on pairs the solver
generates at 16^4 from the reference's random field of seed 1) is applied
inside FGCR on 32^3x64 periodic lattices whose random fields (seeds 6 and
7, flowed 50 x 0.01, quark phase, csw=1.77, m0=-0.35) it has never seen.
FP16-storage SAP+GCR with the network must need fewer Krylov vectors than
FP16-storage SAP+GCR alone (together at most 95 %), and every solve must
reach the double-precision residual 1e-10 (recomputed on the device by
Dw_dble, which the oracle pins at this size)."""
import ctypes as C

import numpy as np
import pytest

import nntrain
import pyoqh

P = C.c_void_p


@pytest.mark.gpu
def test_learned_gain_held_out_32x32x32x64():
    print("training at 16^4", flush=True)
    wts, rec = nntrain.train_preconditioner((4, 8, 4))
    print("training record", rec, flush=True)
    assert rec["held_out_loss_after"] < rec["held_out_loss_before"]
    lat = pyoqh.Lattice((64, 32, 32, 32), (8, 8, 8, 8), bc=3, m0=-0.35, csw=1.77)
    counts = {0: 0, 1: 0}
    try:
        lib = lat.lib
        lib.sap_gcr.restype = C.c_double
        lib.sap_gcr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.norm_square_dble.restype = pyoqh.Qflt
        lib.norm_square_dble.argtypes = [C.c_int, C.c_int, P]
        lib.mulr_spinor_add_dble.argtypes = [C.c_int, C.c_int, P, P, C.c_double]
        bs = np.array((4, 4, 4, 4), np.int32)
        lib.oqh_set_sap_parms(bs.ctypes.data_as(P), 1, 4, 5)
        w = np.array((4, 8, 4), np.int32)
        assert lib.oqh_nn_create(2, w.ctypes.data_as(P)) == 0
        lib.oqh_nn_set_weights(wts.ctypes.data_as(P))
        V = lat.V
        e, p, r = lat.alloc_sd(), lat.alloc_sd(), lat.alloc_sd()
        lib.oqh_set_sap_prec(1)
        for seed in (6, 7):
            lib.start_ranlux(0, seed)
            lib.random_ud()
            lib.fwd_euler(50, 0.01)
            lib.set_ud_phase()
            lib.random_sd(lat.VT, 2, e, 1.0)
            for nn in (0, 1):
                lib.oqh_set_nn_prec(nn)
                ifail, status = C.c_int(), C.c_int()
                lib.sap_gcr(24, 1024, 0, 1e-10, 0.0, e, p, C.byref(ifail), C.byref(status))
                lib.Dw_dble(0.0, p, r)
                lib.mulr_spinor_add_dble(V, 0, r, e, -1.0)
                q1, q0 = lib.norm_square_dble(V, 0, r), lib.norm_square_dble(V, 0, e)
                res = ((q1.q[0] + q1.q[1]) / (q0.q[0] + q0.q[1])) ** 0.5
                print("seed", seed, "network", nn, "status", status.value, "true residual", res, flush=True)
                assert ifail.value == 0 and res <= 1.01e-10
                counts[nn] += status.value
        lib.oqh_set_nn_prec(0)
        lib.oqh_set_sap_prec(0)
    finally:
        lat.lib.oqh_nn_free()
        lat.close()
    print("Krylov vectors: FP16 SAP", counts[0], "FP16 SAP + network", counts[1])
    assert counts[1] <= 0.95 * counts[0]
