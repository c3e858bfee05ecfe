"""numpy restatement of the deflation subspace and the little Dirac operator
(TEST INFRASTRUCTURE ONLY: used by tests/ as the checker).

  block_gram_schmidt .... dfl/dfl_subspace.c:198-282 (dfl_subspace): the
                          modes restricted to each block, modified
                          Gram-Schmidt in double precision in mode order,
                          normalised with a real positive norm; the
                          coordinates of each mode in the basis
  s2vd / vd2s ........... dfl_subspace.c:108-196 (components along the
                          subspace, and back)
  little_operator ....... little/README.Aw: A[n,k; m,l] = (v_{n,k},
                          (Dw + i mu gamma5 I_eo) v_{m,l}), here by its
                          definition, applying the oracle's global Dw_dble
                          to each basis vector extended by zero
  ahat_dense ............ the even-odd form Ahat = 1 - Aee^-1 Aeo Aoo^-1 Aoe
                          (little/README.Aw)

Spinor fields are (V, 24) float arrays in the reference's layout; complex
spinors (V, 12).  Blocks are given by their point lists bpts[nb][bvol] (the
library's numbering, even blocks first).
"""
import numpy as np


def cplx(a):
    a = np.asarray(a, np.float64)
    return a[:, 0::2] + 1j * a[:, 1::2]


def real(z):
    out = np.empty((z.shape[0], 24))
    out[:, 0::2] = z.real
    out[:, 1::2] = z.imag
    return out


def block_gram_schmidt(mds, bpts):
    """Basis (Ns, V, 12) complex and coordinates vm[l][n][j] of mode l on
    block n (j <= l) from the modes mds[l] (V, 24)."""
    Ns = len(mds)
    V = mds[0].shape[0]
    basis = np.zeros((Ns, V, 12), complex)
    vm = np.zeros((Ns, len(bpts), Ns), complex)
    for n, pts in enumerate(bpts):
        for l in range(Ns):
            s = cplx(mds[l])[pts]
            for j in range(l):
                z = np.vdot(basis[j][pts].ravel(), s.ravel())
                vm[l, n, j] = z
                s = s - z * basis[j][pts]
            r = np.sqrt((np.abs(s) ** 2).sum())
            vm[l, n, l] = r
            basis[l][pts] = s / r
    return basis, vm


def s2vd(basis, bpts, s):
    """w[n, k] = (v_{n,k}, s) on block n."""
    zs = cplx(s)
    return np.array([[np.vdot(basis[k][pts].ravel(), zs[pts].ravel()) for k in range(len(basis))]
                     for pts in bpts])


def vd2s(basis, bpts, w):
    V = basis.shape[1]
    out = np.zeros((V, 12), complex)
    for n, pts in enumerate(bpts):
        out[pts] = np.tensordot(w[n], basis[:, pts], axes=(0, 0))
    return real(out)


def little_operator(o, basis, bpts, mu):
    """Dense little operator A[(n,k), (m,l)] by its definition: the
    oracle's Dw_dble (its SW term must be the uncompressed, non-inverted one)
    applied to v_{m,l} extended by zero."""
    Ns, V = basis.shape[0], basis.shape[1]
    nb = len(bpts)
    A = np.zeros((nb * Ns, nb * Ns), complex)
    for m, pm in enumerate(bpts):
        for l in range(Ns):
            f = np.zeros((V, 12), complex)
            f[pm] = basis[l][pm]
            g = cplx(o.Dw_dble(mu, real(f)))
            for n, pn in enumerate(bpts):
                A[n * Ns:(n + 1) * Ns, m * Ns + l] = basis[:, pn].reshape(Ns, -1).conj() @ g[pn].ravel()
    return A


def ahat_dense(A, nb, Ns):
    """Even-odd preconditioned little operator on the even blocks (the
    first nb/2 blocks)."""
    h = (nb // 2) * Ns
    Aee, Aeo, Aoe, Aoo = A[:h, :h], A[:h, h:], A[h:, :h], A[h:, h:]
    return np.eye(h) - np.linalg.solve(Aee, Aeo @ np.linalg.solve(Aoo, Aoe))
