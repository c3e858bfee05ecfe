"""CPU oracle of the gauge-equivariant network used as a learned
preconditioner (numpy, complex128).  Test infrastructure only: imported by
tests/ and bench.py, never by the library.

The network has no counterpart in the reference; its layers are
parallel-transport convolutions in the sense of Lehner and Wettig
(Phys. Rev. D 108 (2023) 034503).  A field with C channels holds C colour
vectors per point, psi[x, i, a] (i channel, a colour); a spinor is the field
with the 4 spin components as channels.  One layer maps C_in to C_out
channels through the 9-point stencil

    out(x)[o] = sum_{p=0..8} sum_i W[p, o, i] (T_p psi)(x)[i]

    T_0 psi(x)      = psi(x)
    T_{2mu+1} psi(x) = U(x, mu) psi(x + mu)
    T_{2mu+2} psi(x) = U(x - mu, mu)^+ psi(x - mu)

with complex weights W[p] acting on the channel index only, so the layer
commutes with gauge transformations psi(x) -> g(x) psi(x),
U(x, mu) -> g(x) U(x, mu) g(x + mu)^+.  The network is linear (no activation),
a learned approximation of D^-1.

Derivatives follow the convention g = dL/dRe(z) + i dL/dIm(z) for a real
loss L: a layer's input gradient is sum_p W[pbar]^+ (T_p g) (T_p^+ =
T_pbar, pbar the opposite direction) and dL/dW[p] = sum_{x,a} g(x)[:, a]
conj((T_p psi)(x)[:, a])^T.

Parity pins (tests/test_oracle_ptc.py): gauge covariance, the adjoint
identity <g, layer(psi)> = <layer^+(g), psi>, finite differences of the loss
against both gradients, and special weights that reduce a layer to the
identity or to a covariant shift.
"""
import numpy as np

NDIR = 9


def pbar(p):
    """Direction opposite to p (T_p^+ = T_pbar)."""
    if p == 0:
        return 0
    return p + 1 if p % 2 == 1 else p - 1


def transport(psi, U, iup, idn, p):
    """(T_p psi)(x) for a field psi[V, C, 3]; U[V, 4, 3, 3] = U(x, mu)."""
    if p == 0:
        return psi.copy()
    mu = (p - 1) // 2
    if p % 2 == 1:
        y = iup[:, mu]
        return np.einsum("xab,xib->xia", U[:, mu], psi[y])
    y = idn[:, mu]
    return np.einsum("xba,xib->xia", np.conj(U[y, mu]), psi[y])


def layer_fwd(W, psi, U, iup, idn):
    out = np.zeros((psi.shape[0], W.shape[1], 3), complex)
    for p in range(NDIR):
        out += np.einsum("oi,xia->xoa", W[p], transport(psi, U, iup, idn, p))
    return out


def layer_adj(W, g, U, iup, idn):
    """Adjoint of the layer applied to g[V, C_out, 3]: the input gradient."""
    out = np.zeros((g.shape[0], W.shape[2], 3), complex)
    for p in range(NDIR):
        out += np.einsum("oi,xoa->xia", np.conj(W[pbar(p)]), transport(g, U, iup, idn, p))
    return out


def layer_wgrad(psi, g, U, iup, idn):
    """dL/dW[p, o, i] = sum_{x,a} g[x, o, a] conj((T_p psi)[x, i, a])."""
    return np.stack([np.einsum("xoa,xia->oi", g, np.conj(transport(psi, U, iup, idn, p)))
                     for p in range(NDIR)])


def net_fwd(Ws, psi, U, iup, idn):
    """Activations of all layers: acts[0] = psi, acts[-1] = output."""
    acts = [psi]
    for W in Ws:
        acts.append(layer_fwd(W, acts[-1], U, iup, idn))
    return acts


def net_bwd(Ws, acts, g, U, iup, idn):
    """Weight gradients of all layers and the input gradient for the output
    gradient g."""
    dWs = [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dWs[l] = layer_wgrad(acts[l], g, U, iup, idn)
        g = layer_adj(Ws[l], g, U, iup, idn)
    return dWs, g


def loss_and_grad(out, e):
    """L = |out - e|^2 / |e|^2 and its output gradient 2 (out - e) / |e|^2."""
    d = out - e
    n = np.vdot(e, e).real
    return np.vdot(d, d).real / n, 2.0 * d / n


def spinor_to_field(s):
    """[V, 24] reals (4 spins x 3 colours x re/im) -> [V, 4, 3] complex."""
    s = np.asarray(s, np.float64)
    return (s[:, 0::2] + 1j * s[:, 1::2]).reshape(-1, 4, 3)


def field_to_spinor(f):
    f = f.reshape(f.shape[0], -1)
    out = np.empty((f.shape[0], 2 * f.shape[1]))
    out[:, 0::2] = f.real
    out[:, 1::2] = f.imag
    return out


def init_weights(widths, rng, scale=None):
    """Random complex weights [9, C_out, C_in] per layer, variance 1/(9 C_in)
    unless scale is given."""
    Ws = []
    for ci, co in zip(widths[:-1], widths[1:]):
        s = scale if scale is not None else 1.0 / np.sqrt(2.0 * NDIR * ci)
        Ws.append(s * (rng.standard_normal((NDIR, co, ci)) + 1j * rng.standard_normal((NDIR, co, ci))))
    return Ws
