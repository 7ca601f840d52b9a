"""Local observables, correlators and entanglement of finite and infinite MPS.

Mirrors (paths relative to the MPSKit.jl tree):
  expectation_value(psi, i => O)    src/algorithms/expval.jl:53-98 (one site: <AC|O|AC>;
                                    two neighbouring sites: <AC AR|O|AC AR>; / norm^2)
  correlator(psi, O1, O2, i, js)    src/algorithms/correlators.jl:11-38 (O1 on the center
                                    tensor, then right-orthonormal transfers to each j)
  entanglement_spectrum, entropy    src/algorithms/toolbox.jl:1-30, 80-101 (singular values
                                    of the bond matrix C; -sum s^2 log s^2 over s >= eps)

Sites are 0-based here.  A cut ``k`` of a finite chain lies between sites k-1 and
k (k = 0 .. L, the number of sites left of it), which is the reference's
``site = k`` in its 1-based numbering.  For an InfiniteMPS, cut k lies left of
site k of the unit cell (k = 0: the bond between cells, the reference's default).
Operators are real matrices O[b, h] = <b|O|h> (two-site O[b1, b2, h1, h2]).
The transfers of ``correlator`` are the HIP library's MPS transfer; the small
contractions at the ends and the SVD of C are torch float64 calls.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import torch

from . import tensors as T
from .krylov import cgs2
from .states import FiniteMPS, lq_right
from .transfer import transfer_left


def _op(O, ref):
    return torch.as_tensor(np.asarray(O, dtype=np.float64), device=ref.device)


def _finite(psi):
    return isinstance(psi, FiniteMPS)


def _center(psi, site):
    """(AC at ``site``, AR at ``site + 1`` or None) in the mixed gauge."""
    if _finite(psi):
        psi.move_center(site)
        return psi.AC, (psi.AR[site + 1] if site + 1 < psi.L else None)
    N = psi.N
    return psi.AC[site % N], psi.AR[(site + 1) % N]


def _ar(psi, j):
    return psi.AR[j] if _finite(psi) else psi.AR[j % psi.N]


def expectation_value_local(psi, site, O):
    """<psi| O |psi> / <psi|psi> for a one-site operator at ``site`` or a two-site
    operator (O of shape (d, d, d, d)) on ``site``, ``site + 1``."""
    AC, AR = _center(psi, site)
    O = _op(O, AC)
    nrm2 = float(torch.dot(AC.reshape(-1), AC.reshape(-1)))
    if O.dim() == 2:
        return float(torch.einsum("lbr,bh,lhr->", AC, O, AC)) / nrm2
    if AR is None:
        raise ValueError("expectation_value_local: two-site operator on the last site")
    X = torch.einsum("lbm,mcr->lbcr", AC, AR)
    return float(torch.einsum("lbcr,bchk,lhkr->", X, O, X)) / nrm2


def correlator(psi, O1, O2, i, js):
    """<O1_i O2_j> / <psi|psi> for every j in ``js`` (all > i), as the reference's
    correlator: O1 on the center tensor at i, right-orthonormal transfers up to j."""
    js = list(js)
    if not js or min(js) <= i:
        raise ValueError("correlator: every j must be larger than i")
    AC, _ = _center(psi, i)
    O1, O2 = _op(O1, AC), _op(O2, AC)
    nrm2 = float(torch.dot(AC.reshape(-1), AC.reshape(-1)))
    # V[r_bra, r_ket] = sum AC[l, b, r_bra] O1[b, h] AC[l, h, r_ket]
    V = T.as_colmajor(torch.einsum("lbr,bh,lhs->rs", AC, O1, AC))
    out, ctr = [], i + 1
    for j in sorted(js):
        while ctr < j:
            A = _ar(psi, ctr)
            V = transfer_left(V, A, A)
            ctr += 1
        A = _ar(psi, j)
        out.append(float(torch.einsum("rs,rbt,bh,sht->", V, A, O2, A)) / nrm2)
    order = {j: k for k, j in enumerate(sorted(js))}
    return [out[order[j]] for j in js]


def entanglement_spectrum(psi, cut=0):
    """Singular values of the bond matrix C at ``cut`` (descending)."""
    if _finite(psi):
        L = psi.L
        if not 0 <= cut <= L:
            raise ValueError(f"entanglement_spectrum: cut {cut} outside 0..{L}")
        if cut == 0 or cut == L:
            psi.move_center(0 if cut == 0 else L - 1)
            return torch.linalg.vector_norm(psi.AC).reshape(1)
        psi.move_center(cut)
        C, _ = lq_right(psi.AC, householder=True)
        return torch.linalg.svdvals(C)
    return torch.linalg.svdvals(psi.C[(cut - 1) % psi.N])


def entropy(psi, cut=0):
    """Von Neumann entropy -sum s^2 log s^2 of the spectrum at ``cut`` (values
    below machine epsilon contribute zero)."""
    s = entanglement_spectrum(psi, cut)
    s = s[s >= torch.finfo(torch.float64).eps]
    s2 = s * s
    return float(-(s2 * torch.log(s2)).sum())


def _mpo_product(W1, W2):
    """(W1 W2)[(g1 g2), b, h, (k1 k2)] = sum_c W1[g1, b, c, k1] W2[g2, c, h, k2]
    (H * H of the reference's MPO multiplication); the pair index is g1 + w1 g2."""
    wl1, d, _, wr1 = W1.shape
    wl2, _, _, wr2 = W2.shape
    P = np.einsum("abck,gchl->agbhkl", W1, W2)
    return np.asfortranarray(P.transpose(1, 0, 2, 3, 5, 4).reshape(wl1 * wl2, d, d, wr1 * wr2, order="F"))


def variance(psi: FiniteMPS, H):
    """<H^2> - <H>^2 for a finite MPS (variance(state::FiniteMPS, H), toolbox.jl:125-128):
    <psi|H H|psi> through the product MPO with the HIP MPO transfers, <H> as energy."""
    from .algorithms import expectation_value
    from .environments import FiniteEnvironments
    L = psi.L
    psi.move_center(L - 1)
    G = T.zeros((1, 1, 1))
    G.fill_(1.0)
    for j in range(L):
        W = np.asarray(H.host[j])
        A = psi.AL[j] if j < L - 1 else psi.AC
        G = transfer_left(G, T.from_numpy(_mpo_product(W, W)), A, A)
    nrm2 = float(torch.dot(psi.AC.reshape(-1), psi.AC.reshape(-1)))
    h2 = float(G.reshape(-1)[0]) / nrm2
    e = expectation_value(psi, H, FiniteEnvironments(psi, H))
    return h2 - e * e


def dot(psi1: FiniteMPS, psi2: FiniteMPS):
    """<psi1|psi2> of two finite MPS on the same chain (dot(psi1, psi2) of the
    reference, states/finitemps.jl): the left MPS transfer through all sites."""
    if psi1.L != psi2.L:
        raise ValueError("dot: chains of different length")
    V = T.zeros((1, 1))
    V.fill_(1.0)
    for j in range(psi1.L):
        V = transfer_left(V, psi2.tensor_at(j), psi1.tensor_at(j))
    return float(V.reshape(-1)[0])


def transfer_spectrum(psi, howmany=20, krylovdim=None, seed=0):
    """Leading (largest modulus) eigenvalues of the left transfer matrix of the
    unit cell of an InfiniteMPS (transfer_spectrum, transfer_spectrum.jl:27-60,
    one dense sector): one Arnoldi space of ``krylovdim`` (default max(30,
    2 howmany + 10)) vectors over the HIP MPS transfer, two Gram-Schmidt passes
    per step.  Returns (eigenvalues as a complex numpy array in descending
    modulus, their Ritz residual estimates)."""
    from .vumps import _flat, _unflat
    D = psi.AL[0].shape[0]
    n = D * D
    k = min(krylovdim or max(30, 2 * howmany + 10), n)
    rng = np.random.default_rng(seed)
    v = T.from_numpy(rng.standard_normal((D, D)))
    Vb = torch.empty((k + 1, n), dtype=torch.float64, device=v.device)
    x = _flat(v)
    Vb[0] = x / torch.linalg.vector_norm(x)
    H = np.zeros((k + 1, k))
    m = k
    for j in range(k):
        y = _unflat(Vb[j], (D, D))
        for s in range(psi.N):
            y = transfer_left(T.as_colmajor(y), psi.AL[s], psi.AL[s])
        w = _flat(y)
        hcol = cgs2(Vb, j + 1, w)
        H[: j + 1, j] = hcol[:-1]
        H[j + 1, j] = hcol[-1]
        if hcol[-1] <= 1e-14:
            m = j + 1
            break
        Vb[j + 1] = w / hcol[-1]
    ev, U = np.linalg.eig(H[:m, :m])
    order = np.argsort(-np.abs(ev))[: min(howmany, m)]
    res = np.abs(H[m, m - 1] * U[m - 1, order]) if m < H.shape[0] else np.zeros(len(order))
    return ev[order], res


def _simplest_rational(lo, hi):
    """The fraction of smallest denominator in the closed interval [lo, hi]
    (0 <= lo <= hi, exact Fractions), by the continued-fraction recursion."""
    fl = math.floor(lo)
    if fl == lo:
        return Fraction(fl)
    if fl + 1 <= hi:
        return Fraction(fl + 1)
    inner = _simplest_rational(1 / (hi - fl), 1 / (lo - fl))
    return fl + 1 / inner


def rationalize(x, tol):
    """The simplest fraction within ``tol`` of ``x`` (Julia's ``rationalize(x;
    tol)`` as used by approx_angles, transfer_spectrum.jl:107-116)."""
    x, tol = Fraction(float(x)), Fraction(float(tol))
    lo, hi = x - tol, x + tol
    if lo <= 0 <= hi:
        return Fraction(0)
    if hi < 0:
        return -_simplest_rational(-hi, -lo)
    return _simplest_rational(lo, hi)


def approx_angles(spectrum, tol_angle=0.1):
    """Angles of the eigenvalues snapped to the simplest fraction of pi within
    ``tol_angle`` (in units of pi), -pi mapped to pi (approx_angles,
    transfer_spectrum.jl:107-116).  Returned as exact fractions of pi."""
    out = []
    for lam in np.asarray(spectrum, dtype=np.complex128):
        a = rationalize(np.angle(lam) / np.pi, tol_angle)
        out.append(Fraction(1) if a == -1 else a)
    return out


def marek_gap(spectrum, tol_angle=0.1):
    """(epsilon, delta, theta) of a transfer spectrum (marek_gap, transfer_spectrum.jl:136-160):
    drop the eigenvalues of modulus 1, snap the angles to fractions of pi
    (``approx_angles``), take the angle theta of the largest remaining
    eigenvalue and the eigenvalues at exactly that snapped angle; epsilon =
    -log|lambda| of the first of them, delta the difference of the first two
    -log|lambda| when more than two share the angle (else inf), as the
    reference has it."""
    spec = np.asarray(spectrum, dtype=np.complex128)
    spec = spec[np.abs(spec) < 1 - 1e-12]
    if spec.size == 0:
        raise ValueError("marek_gap: no eigenvalue below modulus 1 (request more eigenvalues, or the MPS "
                         "is not injective)")
    spec = spec[np.argsort(-np.abs(spec), kind="stable")]
    angles = approx_angles(spec, tol_angle)
    theta = angles[0]
    at = spec[[a == theta for a in angles]]
    lambdas = -np.log(np.abs(at))
    delta = lambdas[1] - lambdas[0] if len(lambdas) > 2 else np.inf
    return float(lambdas[0]), float(delta), float(theta) * np.pi


def correlation_length(psi, **kw):
    """1 / epsilon of the transfer spectrum (correlation_length, transfer_spectrum.jl:170-180)."""
    ev, _ = transfer_spectrum(psi, **kw)
    return 1.0 / marek_gap(ev)[0]
