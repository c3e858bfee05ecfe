"""Pins of the oracle's deflation restatement (oracle/dfl.py) on a 16x8^3
lattice with 4^4 deflation blocks (nb = 32, blocks numbered even first):

* the block Gram-Schmidt basis equals the Q factor of a QR decomposition of
  the block-restricted modes with positive real diagonal (the textbook form
  of Gram-Schmidt), is orthonormal on every block and reproduces every mode
  from its coordinates;
* s2vd and vd2s: P = vd2s s2vd is a projector, and s2vd inverts vd2s;
* the little operator is the Galerkin projection: s2vd(D vd2s(x)) = A x for
  a random little vector x (D applied to the whole prolongated field rather
  than to single basis vectors), and A vanishes between blocks that are not
  neighbours;
* the even-odd form: the even part of the solution of A x = eta solves
  Ahat x_e = Aee^-1 (eta_e - Aeo Aoo^-1 eta_o).
A transposed or conjugated index, a wrong block order or a dropped term
fails one of these."""
import numpy as np
import pytest

import dfl as od
from helpers import make_oracle

L, T, BS, NS = (16, 8, 8, 8), (8, 8, 8, 8), (4, 4, 4, 4), 4


def blocks(o):
    """Point lists of the 4^4 blocks, even blocks first, lexicographic."""
    lex = np.empty(o.V, np.int64)
    ipt = o.ipt()
    lex[ipt] = np.arange(o.V)
    x = np.stack(np.unravel_index(lex, L), axis=1)
    nbl = [L[mu] // BS[mu] for mu in range(4)]
    b = x // np.array(BS)
    order = []
    for col in (0, 1):
        for m in np.ndindex(*nbl):
            if sum(m) % 2 == col:
                order.append(m)
    bpts = []
    for m in order:
        sel = np.all(b == np.array(m), axis=1)
        bpts.append(np.flatnonzero(sel))
    return bpts, order


@pytest.fixture(scope="module")
def setup():
    o, eta, _ = make_oracle(L, T, bc=3, m0=-0.3, csw=1.77, seed=4)
    o.sw_term(3)
    rng = np.random.default_rng(12)
    mds = [rng.standard_normal((o.V, 24)) for _ in range(NS)]
    bpts, order = blocks(o)
    basis, vm = od.block_gram_schmidt(mds, bpts)
    return o, mds, bpts, order, basis, vm


def test_block_gram_schmidt(setup):
    o, mds, bpts, order, basis, vm = setup
    for n, pts in enumerate(bpts):
        B = basis[:, pts].reshape(NS, -1)
        assert np.abs(B.conj() @ B.T - np.eye(NS)).max() < 1e-13
        M = np.stack([od.cplx(m)[pts].ravel() for m in mds], axis=1)
        Q, Rm = np.linalg.qr(M)
        ph = np.diag(Rm) / np.abs(np.diag(Rm))
        Q = Q * ph[None, :]
        assert np.abs(Q.T - B).max() < 1e-12
        for l in range(NS):
            rec = (vm[l, n, :l + 1, None] * B[:l + 1]).sum(axis=0)
            assert np.abs(rec - M[:, l]).max() < 1e-12
            assert abs(vm[l, n, l].imag) == 0 and vm[l, n, l].real > 0


def test_projection(setup):
    o, mds, bpts, order, basis, vm = setup
    s = np.random.default_rng(3).standard_normal((o.V, 24))
    w = od.s2vd(basis, bpts, s)
    p1 = od.vd2s(basis, bpts, w)
    assert np.abs(od.vd2s(basis, bpts, od.s2vd(basis, bpts, p1)) - p1).max() < 1e-12
    assert np.abs(od.s2vd(basis, bpts, p1) - w).max() < 1e-12
    # the modes lie in the subspace
    for m in mds:
        assert np.abs(od.vd2s(basis, bpts, od.s2vd(basis, bpts, m)) - m).max() < 1e-12


def test_little_operator_galerkin(setup):
    o, mds, bpts, order, basis, vm = setup
    mu = 0.05
    A = od.little_operator(o, basis, bpts, mu)
    nb = len(bpts)
    x = np.random.default_rng(5).standard_normal((nb, NS)) + 1j * np.random.default_rng(6).standard_normal((nb, NS))
    ref = od.s2vd(basis, bpts, o.Dw_dble(mu, od.vd2s(basis, bpts, x)))
    assert np.abs((A @ x.ravel()).reshape(nb, NS) - ref).max() < 1e-12 * np.abs(ref).max()
    # locality: only the block itself and its neighbours couple
    nbl = [L[mu] // BS[mu] for mu in range(4)]
    for n, mn in enumerate(order):
        for m, mm in enumerate(order):
            d = [min((mn[k] - mm[k]) % nbl[k], (mm[k] - mn[k]) % nbl[k]) for k in range(4)]
            blk = A[n * NS:(n + 1) * NS, m * NS:(m + 1) * NS]
            if sum(d) > 1:
                assert np.abs(blk).max() == 0.0
            elif sum(d) == 1:
                assert np.abs(blk).max() > 0.0


def test_even_odd_form(setup):
    o, mds, bpts, order, basis, vm = setup
    A = od.little_operator(o, basis, bpts, 0.0)
    nb = len(bpts)
    h = (nb // 2) * NS
    rng = np.random.default_rng(8)
    eta = rng.standard_normal(nb * NS) + 1j * rng.standard_normal(nb * NS)
    x = np.linalg.solve(A, eta)
    Ah = od.ahat_dense(A, nb, NS)
    b = np.linalg.solve(A[:h, :h], eta[:h] - A[:h, h:] @ np.linalg.solve(A[h:, h:], eta[h:]))
    assert np.abs(Ah @ x[:h] - b).max() < 1e-10 * np.abs(b).max()
