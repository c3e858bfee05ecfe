"""ctypes binding of the CPU oracle (liboracle.so).  Test infrastructure only:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Fields are numpy arrays in the reference's formats and site order:
  spinor_dble -> float64 [V, 24], spinor -> float32 [V, 24]
  gauge (ud)  -> float64 [4V, 18] (8 links per odd site, README.global)
  pauli_dble  -> float64 [2V, 36]
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        i, d, f = C.c_int, C.c_double, C.c_float
        sig = {
            "oc_create": (P, [P, P]), "oc_destroy": (None, [P]),
            "oc_volume": (i, [P]), "oc_nthread": (i, [P]),
            "oc_get_ipt": (None, [P, P]), "oc_get_iupdn": (None, [P, P, P]),
            "oc_set_bc": (None, [P, i, d, d]), "oc_set_sw": (None, [P, d, d]),
            "oc_set_tm": (None, [P, i]), "oc_set_sap": (None, [P, P, i, i, i]),
            "oc_start_ranlux": (None, [P, i, i]),
            "oc_kat_ranlxs": (None, [i, i, P, i]), "oc_kat_ranlxd": (None, [i, i, P, i]),
            "oc_random_ud": (None, [P]), "oc_set_ud_phase": (None, [P]),
            "oc_plaq_sum": (d, [P, i]),
            "oc_udfld": (P, [P]), "oc_set_ud": (None, [P, P]),
            "oc_assign_ud2u": (None, [P]), "oc_ufld": (P, [P]),
            "oc_random_sd": (None, [P, P, d]), "oc_random_s": (None, [P, P, f]),
            "oc_random_sd_master": (None, [P, i, P, d]), "oc_random_source": (None, [P, i, P]),
            "oc_bnd_sd2zero": (None, [P, i, P]), "oc_bnd_s2zero": (None, [P, i, P]),
            "oc_unorm_dble": (d, [i, P]), "oc_unorm": (f, [i, P]),
            "oc_ftensor": (None, [P, P]), "oc_sw_term": (i, [P, i]),
            "oc_swdfld": (P, [P]), "oc_assign_swd2sw": (None, [P]), "oc_swfld": (P, [P]),
            "oc_mul_pauli_dble": (None, [d, P, P, P]), "oc_mul_pauli": (None, [f, P, P, P]),
            "oc_inv_pauli_dble": (i, [d, P, P]),
            "oc_apply_sw_dble": (None, [i, d, P, P, P]),
            "oc_acc_qflt": (None, [d, P]), "oc_add_qflt": (None, [P, P, P]),
            "oc_spinor_prod_dble": (None, [i, P, P, P, P]),
            "oc_norm_square_dble": (None, [i, P, P]),
            "oc_spinor_prod": (None, [i, P, P, P]), "oc_norm_square": (f, [i, P]),
            "oc_sap_nblocks": (i, [P]), "oc_sap_block_info": (None, [P, i, P, P]),
            "oc_sap_block_imb": (None, [P, i, P]), "oc_sap_prepare": (i, [P]),
            "oc_Dw_blk": (None, [P, i, f, P, P]), "oc_Dwhat_blk": (None, [P, i, f, P, P]),
            "oc_blk_eo_op": (None, [P, i, i, f, P, P]),
            "oc_blk_mres": (None, [P, i, f, i, P, P]),
            "oc_blk_eo_mres": (None, [P, i, f, i, P, P]),
            "oc_sap": (None, [P, f, i, i, P, P]),
            "oc_Dw_blk_dble": (None, [P, i, d, P, P]), "oc_Dwhat_blk_dble": (None, [P, i, d, P, P]),
            "oc_blk_eo_op_dble": (None, [P, i, i, d, P, P]),
            "oc_blk_mres_dble": (None, [P, i, d, i, P, P]),
            "oc_blk_eo_mres_dble": (None, [P, i, d, i, P, P]),
            "oc_sap_dble": (None, [P, d, i, i, P, P]),
            "oc_sap_gcr": (d, [P, i, i, i, d, d, P, P, P, P]),
            "oc_plaq_frc": (None, [P, P]), "oc_expXsu3": (None, [d, P, P]),
            "oc_fwd_euler": (None, [P, i, d]),
        }
        for op in ("Dw_dble", "Dwee_dble", "Dwoo_dble", "Dwhat_dble"):
            sig["oc_" + op] = (None, [P, d, P, P])
        for op in ("Dwoe_dble", "Dweo_dble"):
            sig["oc_" + op] = (None, [P, P, P])
        for op in ("Dw", "Dwee", "Dwoo", "Dwhat"):
            sig["oc_" + op] = (None, [P, f, P, P])
        for op in ("Dwoe", "Dweo"):
            sig["oc_" + op] = (None, [P, P, P])
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def iarr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


class Oracle:
    """One lattice + parameter set of the oracle."""

    def __init__(self, L, Ltrd=None, bc=3, cF=(1.0, 1.0), m0=0.0, csw=1.0, eoflg=0):
        self.L = tuple(int(x) for x in L)
        self.Ltrd = tuple(int(x) for x in (Ltrd or L))
        self._L = iarr(self.L)
        self._T = iarr(self.Ltrd)
        self.h = lib().oc_create(ptr(self._L), ptr(self._T))
        if not self.h:
            raise ValueError("illegal lattice geometry")
        self.V = lib().oc_volume(self.h)
        self.nthread = lib().oc_nthread(self.h)
        self.bc = bc
        lib().oc_set_bc(self.h, bc, cF[0], cF[1])
        lib().oc_set_sw(self.h, m0, csw)
        lib().oc_set_tm(self.h, eoflg)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    # geometry
    def ipt(self):
        a = np.empty(self.V, np.int32)
        lib().oc_get_ipt(self.h, ptr(a))
        return a

    def iupdn(self):
        a = np.empty((self.V, 4), np.int32)
        b = np.empty((self.V, 4), np.int32)
        lib().oc_get_iupdn(self.h, ptr(a), ptr(b))
        return a, b

    # fields
    def start_ranlux(self, level, seed):
        lib().oc_start_ranlux(self.h, level, seed)

    def random_ud(self):
        lib().oc_random_ud(self.h)

    def set_ud_phase(self):
        lib().oc_set_ud_phase(self.h)

    def ud(self):
        p = lib().oc_udfld(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), (4 * self.V, 18)).copy()

    def set_ud(self, ud):
        ud = np.ascontiguousarray(ud, np.float64)
        lib().oc_set_ud(self.h, ptr(ud))

    def assign_ud2u(self):
        lib().oc_assign_ud2u(self.h)

    def plaq_frc(self):
        """Plaquette force of the current FP64 field, (4V, 8) su3_alg_dble."""
        f = np.empty((4 * self.V, 8), np.float64)
        lib().oc_plaq_frc(self.h, ptr(f))
        return f

    def plaq_sum(self, iw=0):
        """plaq_sum_dble (iw=0) / plaq_wsum_dble (iw=1) of the FP64 field."""
        return lib().oc_plaq_sum(self.h, iw)

    def cnfg_bytes(self):
        """The field as export_cnfg writes it (archive/archive.c:804-870,
        main/README.io): int32 N0..N3, the average plaquette, then the 8
        links of every odd point in lexicographic order (x3 fastest)."""
        return cnfg_encode(self.L, self.ipt(), self.ud(),
                           self.plaq_sum(0) / (6.0 * np.prod(self.L, dtype=np.float64)))

    def fwd_euler(self, n, eps):
        lib().oc_fwd_euler(self.h, n, eps)

    def random_sd(self, sigma=1.0):
        a = np.empty((self.V, 24), np.float64)
        lib().oc_random_sd(self.h, ptr(a), sigma)
        return a

    def random_sd_master(self, vol, sigma=1.0):
        """random_sd(vol, 0, sd, sigma) called by the master thread."""
        a = np.empty((vol, 24), np.float64)
        lib().oc_random_sd_master(self.h, vol, ptr(a), sigma)
        return a

    def random_source(self, t0):
        """The random time-slice source of main/ms4.c (random_source)."""
        a = np.empty((self.V, 24), np.float64)
        lib().oc_random_source(self.h, t0, ptr(a))
        return a

    def random_s(self, sigma=1.0):
        a = np.empty((self.V, 24), np.float32)
        lib().oc_random_s(self.h, ptr(a), sigma)
        return a

    def bnd_sd2zero(self, sd, pset=0):
        lib().oc_bnd_sd2zero(self.h, pset, ptr(sd))

    def bnd_s2zero(self, s, pset=0):
        lib().oc_bnd_s2zero(self.h, pset, ptr(s))

    def ftensor(self):
        a = np.empty((6, self.V, 9), np.float64)
        lib().oc_ftensor(self.h, ptr(a))
        return a

    def sw_term(self, pset):
        return lib().oc_sw_term(self.h, pset)

    def swd(self):
        p = lib().oc_swdfld(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), (2 * self.V, 36)).copy()

    def assign_swd2sw(self):
        lib().oc_assign_swd2sw(self.h)

    def sw(self):
        """The single-precision SW field (after assign_swd2sw)."""
        p = lib().oc_swfld(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), (2 * self.V, 36)).copy()

    # operators (inputs may be modified, as in the reference)
    def _op_d(self, name, mu, s, r=None):
        if r is None:
            r = np.zeros((self.V, 24), np.float64)
        fn = getattr(lib(), "oc_" + name)
        if mu is None:
            fn(self.h, ptr(s), ptr(r))
        else:
            fn(self.h, mu, ptr(s), ptr(r))
        return r

    def _op_f(self, name, mu, s, r=None):
        if r is None:
            r = np.zeros((self.V, 24), np.float32)
        fn = getattr(lib(), "oc_" + name)
        if mu is None:
            fn(self.h, ptr(s), ptr(r))
        else:
            fn(self.h, mu, ptr(s), ptr(r))
        return r

    def Dw_dble(self, mu, s, r=None): return self._op_d("Dw_dble", mu, s, r)
    def Dwee_dble(self, mu, s, r=None): return self._op_d("Dwee_dble", mu, s, r)
    def Dwoo_dble(self, mu, s, r=None): return self._op_d("Dwoo_dble", mu, s, r)
    def Dwoe_dble(self, s, r=None): return self._op_d("Dwoe_dble", None, s, r)
    def Dweo_dble(self, s, r=None): return self._op_d("Dweo_dble", None, s, r)
    def Dwhat_dble(self, mu, s, r=None): return self._op_d("Dwhat_dble", mu, s, r)
    def Dw(self, mu, s, r=None): return self._op_f("Dw", mu, s, r)
    def Dwee(self, mu, s, r=None): return self._op_f("Dwee", mu, s, r)
    def Dwoo(self, mu, s, r=None): return self._op_f("Dwoo", mu, s, r)
    def Dwoe(self, s, r=None): return self._op_f("Dwoe", None, s, r)
    def Dweo(self, s, r=None): return self._op_f("Dweo", None, s, r)
    def Dwhat(self, mu, s, r=None): return self._op_f("Dwhat", mu, s, r)

    # SAP / solvers
    def set_sap(self, bs, isolv, nmr, ncy):
        b = iarr(bs)
        lib().oc_set_sap(self.h, ptr(b), isolv, nmr, ncy)

    def sap_prepare(self):
        return lib().oc_sap_prepare(self.h)

    def nblocks(self):
        return lib().oc_sap_nblocks(self.h)

    def block_imb(self, n):
        bo = np.zeros(4, np.int32)
        vol = C.c_int()
        lib().oc_sap_block_info(self.h, n, ptr(bo), C.byref(vol))
        a = np.empty(vol.value, np.int32)
        lib().oc_sap_block_imb(self.h, n, ptr(a))
        return a

    def Dw_blk(self, n, mu, s):
        r = np.zeros_like(s)
        lib().oc_Dw_blk(self.h, n, mu, ptr(s), ptr(r))
        return r

    def blk_eo_op(self, n, op, mu, s, r):
        """Dwee_blk (op 0), Dwoo_blk (1), Dwoe_blk (2), Dweo_blk (3) on
        block fields s (input, b.s[k]) and r (output, b.s[l]); returns the
        new (s, r)."""
        s = np.array(s, np.float32)
        r = np.array(r, np.float32)
        lib().oc_blk_eo_op(self.h, n, op, mu, ptr(s), ptr(r))
        return s, r

    def blk_mres(self, n, mu, nmr, s1):
        s0 = np.zeros_like(s1)
        lib().oc_blk_mres(self.h, n, mu, nmr, ptr(s0), ptr(s1))
        return s0, s1

    def Dwhat_blk(self, n, mu, s):
        """Dwhat_blk on block field s (vol+1 points); even half of the result
        (the odd half is zero)."""
        r = np.zeros_like(s)
        lib().oc_Dwhat_blk(self.h, n, mu, ptr(s), ptr(r))
        return r

    def blk_eo_mres(self, n, mu, nmr, s1):
        s0 = np.zeros_like(s1)
        lib().oc_blk_eo_mres(self.h, n, mu, nmr, ptr(s0), ptr(s1))
        return s0, s1

    def sap(self, mu, isolv, nmr, psi, rho):
        lib().oc_sap(self.h, mu, isolv, nmr, ptr(psi), ptr(rho))

    # FP64 block operators and the FP64 Schwarz sweep (FP64 links, FP64 block
    # SW field with the odd part inverted when isolv=1)
    def Dw_blk_dble(self, n, mu, s):
        r = np.zeros_like(s)
        lib().oc_Dw_blk_dble(self.h, n, mu, ptr(s), ptr(r))
        return r

    def Dwhat_blk_dble(self, n, mu, s):
        r = np.zeros_like(s)
        lib().oc_Dwhat_blk_dble(self.h, n, mu, ptr(s), ptr(r))
        return r

    def blk_eo_op_dble(self, n, op, mu, s, r):
        s = np.array(s, np.float64)
        r = np.array(r, np.float64)
        lib().oc_blk_eo_op_dble(self.h, n, op, mu, ptr(s), ptr(r))
        return s, r

    def blk_mres_dble(self, n, mu, nmr, s1):
        s0 = np.zeros_like(s1)
        lib().oc_blk_mres_dble(self.h, n, mu, nmr, ptr(s0), ptr(s1))
        return s0, s1

    def blk_eo_mres_dble(self, n, mu, nmr, s1):
        s0 = np.zeros_like(s1)
        lib().oc_blk_eo_mres_dble(self.h, n, mu, nmr, ptr(s0), ptr(s1))
        return s0, s1

    def sap_dble(self, mu, isolv, nmr, psi, rho):
        lib().oc_sap_dble(self.h, mu, isolv, nmr, ptr(psi), ptr(rho))

    def sap_gcr(self, nkv, nmx, istop, res, mu, eta, psi=None):
        if psi is None:
            psi = np.zeros_like(eta)
        ifail = C.c_int()
        status = C.c_int()
        rho = lib().oc_sap_gcr(self.h, nkv, nmx, istop, res, mu, ptr(eta), ptr(psi),
                               C.byref(ifail), C.byref(status))
        return rho, ifail.value, status.value, psi


def kat_ranlxs(level, seed, n):
    a = np.empty(n, np.float32)
    lib().oc_kat_ranlxs(level, seed, ptr(a), n)
    return a


def kat_ranlxd(level, seed, n):
    a = np.empty(n, np.float64)
    lib().oc_kat_ranlxd(level, seed, ptr(a), n)
    return a


def inv_pauli_dble(mu, m):
    m = np.ascontiguousarray(m, np.float64)
    im = np.empty(36, np.float64)
    ifail = lib().oc_inv_pauli_dble(mu, ptr(m), ptr(im))
    return ifail, im


def mul_pauli_dble(mu, m, s):
    m = np.ascontiguousarray(m, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    r = np.empty(12, np.float64)
    lib().oc_mul_pauli_dble(mu, ptr(m), ptr(s), ptr(r))
    return r


def apply_sw_dble(mu, m, s):
    """r = (A + i mu gamma5) s on len(s) points with the matrices m[2k],
    m[2k+1] (sw_term/pauli_dble.c:1840-1857)."""
    m = np.ascontiguousarray(m, np.float64)
    s = np.ascontiguousarray(s, np.float64)
    r = np.empty_like(s)
    lib().oc_apply_sw_dble(s.shape[0], mu, ptr(m), ptr(s), ptr(r))
    return r


def spinor_prod_dble(s, r):
    qre = np.zeros(2)
    qim = np.zeros(2)
    lib().oc_spinor_prod_dble(s.shape[0], ptr(s), ptr(r), ptr(qre), ptr(qim))
    return qre, qim


def norm_square_dble(s):
    q = np.zeros(2)
    lib().oc_norm_square_dble(s.shape[0], ptr(s), ptr(q))
    return q


def acc_qflt(u, q):
    q = np.ascontiguousarray(q, np.float64)
    lib().oc_acc_qflt(u, ptr(q))
    return q


def add_qflt(qu, qv):
    qu = np.ascontiguousarray(qu, np.float64)
    qv = np.ascontiguousarray(qv, np.float64)
    qr = np.zeros(2)
    lib().oc_add_qflt(ptr(qu), ptr(qv), ptr(qr))
    return qr


def odd_lex_order(L, ipt):
    """ipt indices of the odd points in lexicographic order (x3 fastest);
    ipt is indexed by x3 + L3*(x2 + L2*(x1 + L1*x0))."""
    x = np.indices(tuple(L)).reshape(4, -1)
    odd = (x.sum(axis=0) & 1) == 1
    return np.asarray(ipt)[np.flatnonzero(odd)]


def cnfg_encode(L, ipt, ud, plaq):
    """Bytes of a configuration file in the universal format."""
    L = [int(v) for v in L]
    h = int(np.prod(L)) // 2
    ix = odd_lex_order(L, ipt)
    links = np.asarray(ud, "<f8").reshape(h, 8, 18)[ix - h]
    return (np.array(L, "<i4").tobytes() + np.array([plaq], "<f8").tobytes() + links.tobytes())


def sfld_encode(L, ipt, sd, eo=0):
    """Bytes of a quark-field file (archive/sarchive.c:424-503): int32
    N0..N3 and eo, the square norm of the written points (norm_square_dble,
    double-double, high part), then their spinors in lexicographic order
    (eo=1: the even points only)."""
    L = [int(v) for v in L]
    x = np.indices(tuple(L)).reshape(4, -1)
    sel = np.ones(x.shape[1], bool) if eo == 0 else (x.sum(axis=0) & 1) == 0
    ix = np.asarray(ipt)[np.flatnonzero(sel)]
    n = len(ix)
    qn = norm_square_dble(np.ascontiguousarray(np.asarray(sd, np.float64)[:n]))
    return (np.array(L + [eo], "<i4").tobytes() + np.array([qn[0]], "<f8").tobytes()
            + np.asarray(sd, "<f8")[ix].tobytes())


def sfld_decode(data, ipt, V):
    """(sizes, eo, square norm, points written as a dict-free array): the
    spinors placed at their ipt positions in a (V, 24) array (NaN where the
    file holds no point)."""
    hdr = np.frombuffer(data[:20], "<i4").astype(int)
    L, eo = hdr[:4], int(hdr[4])
    norm = float(np.frombuffer(data[20:28], "<f8")[0])
    x = np.indices(tuple(L)).reshape(4, -1)
    sel = np.ones(x.shape[1], bool) if eo == 0 else (x.sum(axis=0) & 1) == 0
    ix = np.asarray(ipt)[np.flatnonzero(sel)]
    sd = np.full((V, 24), np.nan)
    sd[ix] = np.frombuffer(data[28:], "<f8").reshape(-1, 24)
    return L, eo, norm, sd


def cnfg_decode(data, ipt):
    """(sizes, average plaquette, ud in the layout of udfld()) of a file."""
    L = np.frombuffer(data[:16], "<i4").astype(int)
    plaq = float(np.frombuffer(data[16:24], "<f8")[0])
    h = int(np.prod(L)) // 2
    links = np.frombuffer(data[24:], "<f8").reshape(h, 8, 18)
    ud = np.empty((h, 8, 18))
    ud[odd_lex_order(L, ipt) - h] = links
    return L, plaq, ud.reshape(8 * h, 18)


def ref_exe(tag):
    """Path of a reference driver executable, or None when not built."""
    p = os.path.join(HERE, "_ref", "ref_" + tag)
    return p if os.path.exists(p) else None


def expXsu3(eps, X, u):
    """exp(eps X) u for one su3_alg_dble X (8) and su3_dble u (18)."""
    X = np.ascontiguousarray(X, np.float64)
    u = np.array(u, np.float64)
    lib().oc_expXsu3(eps, ptr(X), ptr(u))
    return u


def ref_dump(tag, outdir, **kw):
    """Run the reference driver in dump mode and return a loader."""
    exe = ref_exe(tag)
    if exe is None:
        return None
    args = [exe, "dump", outdir]
    for k, v in kw.items():
        args.append("-" + k)
        if isinstance(v, (list, tuple)):
            args += [str(x) for x in v]
        else:
            args.append(str(v))
    os.makedirs(outdir, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS="16")
    subprocess.run(args, check=True, capture_output=True, env=env, timeout=600)

    def load(name, dtype, shape=None):
        a = np.fromfile(os.path.join(outdir, name), dtype=dtype)
        return a if shape is None else a.reshape(shape)

    return load
