"""ctypes binding of libopenqcd_hip.so (the C ABI declared in
include/openqcd_hip.h).  Used by the tests, bench.py and __graft_entry__.

Host arrays use the reference layouts: spinor_dble float64 [V,24], spinor
float32 [V,24], gauge float64 [4V,18], pauli_dble float64 [2V,36]; the site
order is the reference's ipt order.  Device fields are opaque pointers
(ints) returned by alloc_sd/alloc_s; their contents are moved with
import_*/export_*.

The library refuses to run without its HIP build: load() raises if the
shared object is missing; there is no other implementation behind it.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBPATH = os.environ.get("OQH_LIBRARY") or os.path.join(HERE, "libopenqcd_hip.so")
HEADER = os.path.join(ROOT, "include", "openqcd_hip.h")

P = C.c_void_p
_I, _D, _F = C.c_int, C.c_double, C.c_float


class Complex(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class ComplexDble(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Qflt(C.Structure):
    _fields_ = [("q", C.c_double * 2)]


class ComplexQflt(C.Structure):
    _fields_ = [("re", Qflt), ("im", Qflt)]


class DflParms(C.Structure):
    _fields_ = [("bs", C.c_int * 4), ("Ns", C.c_int)]


class DflProParms(C.Structure):
    _fields_ = [("nkv", C.c_int), ("nmx", C.c_int), ("nmx_gcr", C.c_int), ("res", C.c_double),
                ("res_gcr", C.c_double)]


class DflGenParms(C.Structure):
    _fields_ = [("ninv", C.c_int), ("nmr", C.c_int), ("ncy", C.c_int), ("kappa", C.c_double),
                ("m0", C.c_double), ("mu", C.c_double)]


_SIG = {
    "oqh_init": (_I, [_I]), "oqh_finalize": (None, []),
    "oqh_set_error_mode": (None, [_I]), "oqh_last_error": (_I, [C.c_char_p, _I]),
    "oqh_geometry": (_I, [P, P]), "oqh_volume": (_I, []), "oqh_get_ipt": (None, [P]),
    "oqh_set_bc_parms": (None, [_I, _D, _D]), "oqh_set_sw_parms": (None, [_D, _D]),
    "oqh_set_tm_parms": (None, [_I]), "oqh_set_sap_parms": (None, [P, _I, _I, _I]),
    "oqh_import_ud": (None, [P]), "oqh_export_ud": (None, [P]),
    "oqh_export_swd": (None, [P]), "oqh_import_swd": (None, [P]),
    "oqh_alloc_s": (P, [_I]), "oqh_alloc_sd": (P, [_I]), "oqh_free": (None, [P]),
    "oqh_import_s": (None, [P, P, _I]), "oqh_export_s": (None, [P, P, _I]),
    "oqh_import_sd": (None, [P, P, _I]), "oqh_export_sd": (None, [P, P, _I]),
    "oqh_sync": (None, []), "oqh_time_dirac": (_D, [_I, _D, P, P, _I]),
    "sw_term": (_I, [_I]), "assign_swd2sw": (None, []), "assign_ud2u": (None, []),
    "oqh_copy_bandwidth": (_D, [_D, _I]), "oqh_export_frc": (None, [P]),
    "plaq_frc": (None, []), "fwd_euler": (None, [_I, _D]),
    "plaq_sum_dble": (_D, [_I]), "plaq_wsum_dble": (_D, [_I]),
    "export_cnfg": (None, [C.c_char_p]), "import_cnfg": (None, [C.c_char_p, _I]),
    "lat_sizes": (None, [C.c_char_p, P]),
    "export_sfld": (None, [C.c_char_p, _I, P]), "import_sfld": (None, [C.c_char_p, _I, P]),
    "set_ud_phase": (None, []), "unset_ud_phase": (None, []),
    "oqh_set_sap_prec": (None, [_I]),
    "oqh_set_gcr_orth": (None, [_I]),
    "oqh_set_sap_graph": (None, [_I]),
    "oqh_sap_autotune": (_D, [_I, _I, _I, _I, P]),
    "start_ranlux": (None, [_I, _I]), "random_ud": (None, []),
    "random_sd": (None, [_I, _I, P, _D]), "oqh_random_source": (None, [_I, P]), "oqh_random_start_field": (None, [P]), "ranlxd": (None, [P, _I]),
    "sap_gcr_multi": (None, [_I, _I, _I, _I, _D, _D, P, P, P, P, P]),
    "oqh_nn_create": (_I, [_I, P]), "oqh_nn_nweights": (_I, []),
    "oqh_nn_set_weights": (None, [P]), "oqh_nn_get_weights": (None, [P]),
    "oqh_nn_apply": (None, [P, P]), "oqh_nn_correct": (None, [P, P]), "oqh_nn_grad": (_D, [_I, P, P, P]),
    "oqh_nn_train_step": (_D, [_I, P, P, _D]), "oqh_set_nn_prec": (None, [_I]),
    "oqh_nn_free": (None, []),
    "oqh_Dw_dble_multi": (None, [_D, _I, P, P]), "oqh_Dw_multi": (None, [_F, _I, P, P]),
    "oqh_time_dirac_multi": (_D, [_I, _D, _I, P, P, _I]),
    "set_dfl_parms": (DflParms, [P, _I]), "dfl_parms": (DflParms, []),
    "set_dfl_pro_parms": (DflProParms, [_I, _I, _D, _I, _D]), "dfl_pro_parms": (DflProParms, []),
    "set_dfl_gen_parms": (DflGenParms, [_D, _D, _I, _I, _I]), "dfl_gen_parms": (DflGenParms, []),
    "dfl_modes": (None, [P, P]), "dfl_modes2": (None, [P, P]),
    "dfl_update": (None, [_I, P, P]), "dfl_update2": (None, [_I, P, P]), "dfl_restore_modes": (None, [P]),
    "dfl_sap_gcr": (_D, [_I, _I, _I, _D, _D, P, P, P, P]),
    "dfl_sap_gcr2": (_D, [_I, _I, _I, _D, _D, P, P, P, P]),
    "oqh_dfl_release": (None, []),
}
for _n in ("Dw_dble", "Dwee_dble", "Dwoo_dble", "Dwhat_dble"):
    _SIG[_n] = (None, [_D, P, P])
for _n in ("Dwoe_dble", "Dweo_dble", "Dwoe", "Dweo"):
    _SIG[_n] = (None, [P, P])
for _n in ("Dw", "Dwee", "Dwoo", "Dwhat"):
    _SIG[_n] = (None, [_F, P, P])

_LIB = None


def build():
    subprocess.run(["make", "-s", "-j8", "-C", HERE], check=True)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libopenqcd_hip.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def header_functions():
    """Names of all functions declared in include/openqcd_hip.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = []
    for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
        name = m.group(1)
        if name not in ("if", "typedef", "struct") and not txt[m.start():m.end()].lstrip().startswith("typedef"):
            names.append(name)
    return sorted(set(names))


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P)


class Lattice:
    """Library state for one geometry (the library holds one global state,
    so at most one Lattice is open at a time)."""

    _open = None

    def __init__(self, L, Ltrd=None, bc=3, cF=(1.0, 1.0), m0=0.0, csw=1.0, eoflg=0, device=0):
        if Lattice._open is not None:
            raise RuntimeError("another Lattice is open: close() it first (one global library state)")
        self.lib = load()
        self.L = tuple(int(x) for x in L)
        self.Ltrd = tuple(int(x) for x in (Ltrd or L))
        if self.lib.oqh_init(device):
            raise RuntimeError("oqh_init failed")
        a = np.array(self.L, np.int32)
        b = np.array(self.Ltrd, np.int32)
        if self.lib.oqh_geometry(_p(a), _p(b)):
            raise RuntimeError("oqh_geometry failed")
        self.V = self.lib.oqh_volume()
        self.VT = int(np.prod(self.Ltrd))   # VOLUME_TRD: points per reference thread
        self.lib.oqh_set_bc_parms(bc, cF[0], cF[1])
        self.lib.oqh_set_sw_parms(m0, csw)
        self.lib.oqh_set_tm_parms(eoflg)
        self._fields = []
        Lattice._open = self

    def close(self):
        self.lib.oqh_finalize()
        if Lattice._open is self:
            Lattice._open = None

    def ipt(self):
        a = np.empty(self.V, np.int32)
        self.lib.oqh_get_ipt(_p(a))
        return a

    def import_ud(self, ud):
        ud = np.ascontiguousarray(ud, np.float64)
        assert ud.size == 4 * self.V * 18
        self.lib.oqh_import_ud(_p(ud))

    def export_ud(self):
        a = np.empty((4 * self.V, 18), np.float64)
        self.lib.oqh_export_ud(_p(a))
        return a

    def export_frc(self):
        """Force field of the last plaq_frc(), (4V, 8) su3_alg_dble."""
        a = np.empty((4 * self.V, 8), np.float64)
        self.lib.oqh_export_frc(_p(a))
        return a

    def export_swd(self):
        a = np.empty((2 * self.V, 36), np.float64)
        self.lib.oqh_export_swd(_p(a))
        return a

    def sw_term(self, pset):
        return self.lib.sw_term(pset)

    def alloc_sd(self, n=0):
        p = self.lib.oqh_alloc_sd(n)
        if not p:
            raise MemoryError("oqh_alloc_sd")
        return p

    def alloc_s(self, n=0):
        p = self.lib.oqh_alloc_s(n)
        if not p:
            raise MemoryError("oqh_alloc_s")
        return p

    def free(self, p):
        self.lib.oqh_free(p)

    def put_sd(self, dev, host):
        host = np.ascontiguousarray(host, np.float64)
        self.lib.oqh_import_sd(dev, _p(host), host.shape[0])

    def get_sd(self, dev, n=None):
        n = n or self.V
        a = np.empty((n, 24), np.float64)
        self.lib.oqh_export_sd(dev, _p(a), n)
        return a

    def put_s(self, dev, host):
        host = np.ascontiguousarray(host, np.float32)
        self.lib.oqh_import_s(dev, _p(host), host.shape[0])

    def get_s(self, dev, n=None):
        n = n or self.V
        a = np.empty((n, 24), np.float32)
        self.lib.oqh_export_s(dev, _p(a), n)
        return a

    def sync(self):
        self.lib.oqh_sync()


def haar_su3(n, rng):
    """n Haar-distributed SU(3) matrices as float64 [n,18] (QR of complex
    Gaussian matrices with the phase convention that makes Q Haar; the
    determinant phase is divided out of the last row)."""
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q[:, 2, :] /= det[:, None]
    out = np.empty((n, 18))
    out[:, 0::2] = q.real.reshape(n, 9)
    out[:, 1::2] = q.imag.reshape(n, 9)
    return out
