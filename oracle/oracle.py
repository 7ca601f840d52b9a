"""CPU oracle for the effective-Hamiltonian / transfer-matrix contraction path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU restatement of the reference's algorithm).  The
product path (``mpskit.jl_amd``) never imports it and has no CPU fallback.

This is a float64 numpy restatement of MPSKit.jl v0.13.13's contractions, one
function per reference routine, each citing the file:line it follows.  Array
conventions are those of a Trivial-symmetry ``TensorMap``: one array whose axes
are ``(codomain..., domain...)``; buffers exchanged with the HIP library are
Fortran (column-major) ordered so that fusing neighbouring legs is a reshape
with ``order="F"``.

    GL[a, g, f]     (D_bra  x w' <- D_ket)     abstract_envs.jl:36-37
    GR[i, l, c]     (D_ket  x w  <- D_bra)     abstract_envs.jl:50-51
    W[g, b, h, k]   (w_left x p_out <- p_in x w_right)
    A[l, p, r]      MPS tensor (D_l x p <- D_r)
    x2[f, h, i, j]  AC2 (D_l x p1 <- D_r x p2)  finitemps.jl:376-384

Jordan-block channel convention for a dense MPO bond of dimension ``w``
(jordanmpotensor.jl:44-69, :194-249): channel 0 is "not started", channel
``w-1`` is "finished", channels 1..w-2 carry the ``A``-block.  A left boundary
bond of dimension 1 holds only "not started"; a right boundary bond of
dimension 1 holds only "finished" (finite MPOHamiltonian first/last sites).

Pinning (what checks each function; see tests/test_oracle.py):
  * every derivative apply and every environment/transfer routine is checked
    against a brute-force dense Hamiltonian built from Kronecker products of
    Pauli/spin matrices (independent of all index strings here), projected on
    a random mixed-canonical MPS with *different* bra and ket states;
  * prepared == unprepared (mpo_derivatives.jl:233-288 vs :144-152);
  * Jordan == generic apply with identity corner environments;
  * closed-form energies (open TFIM free fermions, Heisenberg chain);
  * transfer fixed points and regularisation identities.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "F", "fmat", "ac2_apply_generic", "ac_apply_generic", "c_apply",
    "prepare_ac2_generic", "apply_prepared_ac2", "prepare_ac_generic",
    "apply_prepared_ac", "prepare_c", "apply_prepared_c",
    "jordan_split", "jordan_ac2_construct", "jordan_ac2_fold", "jordan_ac2_prepare", "jordan_ac2_apply",
    "jordan_ac_construct", "jordan_ac_fold", "jordan_ac_prepare", "jordan_ac_apply",
    "transfer_left_mps", "transfer_right_mps", "transfer_left_mpo",
    "transfer_right_mpo", "regularize",
]


def F(a):
    """Return a Fortran-ordered float64 copy (the library's memory layout)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def fmat(a, rows, cols):
    """Column-major reshape of ``a`` into a ``rows x cols`` matrix (free leg fusion)."""
    return np.reshape(a, (rows, cols), order="F")


def _es(spec, *ops):
    return np.einsum(spec, *ops, optimize=True)


# ---------------------------------------------------------------------------
# Generic (MPO) derivative applies -- mpo_derivatives.jl
# ---------------------------------------------------------------------------

def c_apply(GL, GR, x):
    """Zero-site derivative y[a,c] = GL[a,g,f] x[f,i] GR[i,g,c] (mpo_derivatives.jl:92-97)."""
    return _es("agi,igc->ac", _es("agf,fi->agi", GL, x), GR)


def ac_apply_generic(GL, W, GR, x):
    """Single-site derivative (mpo_derivatives.jl:99-106):
    y[-1 -2; -3] = GL[-1 5; 4] x[4 2; 1] W[5 -2; 2 3] GR[1 3; -3]."""
    t = _es("agf,fhi->aghi", GL, x)
    t = _es("aghi,gbhl->abil", t, W)
    return _es("abil,ilc->abc", t, GR)


def ac2_apply_generic(GL, W1, W2, GR, x):
    """Two-site derivative (mpo_derivatives.jl:144-152):
    y[-1 -2; -3 -4] = GL[-1 7; 6] x[6 5; 1 3] W1[7 -2; 5 4] W2[4 -4; 3 2] GR[1 2; -3]."""
    # pairwise, in the order L (x R) of the prepared form
    t = _es("fhij,ilc->fhjlc", x, GR)
    t = _es("fhjlc,kejl->fhkce", t, W2)
    t = _es("fhkce,gbhk->fgbce", t, W1)
    return _es("fgbce,agf->abce", t, GL)


# ---------------------------------------------------------------------------
# Prepared (precomputed) forms -- mpo_derivatives.jl:216-288
# ---------------------------------------------------------------------------

def prepare_ac2_generic(GL, W1, W2, GR):
    """``prepare_operator!!(::MPO_AC2_Hamiltonian)`` (mpo_derivatives.jl:233-247).

    Returns matrices (L, R) in the layout used by the prepared apply:
      L[(a,b), (f,h,k)] = sum_g GL[a,g,f] W1[g,b,h,k]          (dD x dD*w)
      R[(i,j), (k,c,e)] = sum_l W2[k,e,j,l] GR[i,l,c]          (dD x w*dD)
    """
    Dl, d1, Dr, d2 = GL.shape[0], W1.shape[1], GR.shape[0], W2.shape[1]
    wm = W1.shape[3]
    L = _es("agf,gbhk->abfhk", GL, W1)
    R = _es("kejl,ilc->ijkce", W2, GR)
    return (fmat(L, Dl * d1, Dl * d1 * wm).copy(order="F"),
            fmat(R, Dr * d2, wm * Dr * d2).copy(order="F"))


def apply_prepared_ac2(L, R, x):
    """``(H::PrecomputedAC2Derivative)(x)`` (mpo_derivatives.jl:276-288 -> :249-264):
    xR = x *_1 R (``mul_front!``, utility.jl:271-306) then y = L * xR."""
    Dl, d1, Dr, d2 = x.shape
    m, n = Dl * d1, Dr * d2
    w = R.shape[1] // n
    xR = fmat(x, m, n) @ R                      # (dD) x (w dD) == (dD w) x (dD)
    y = L @ fmat(xR, m * w, n)
    return np.reshape(y, (Dl, d1, Dr, d2), order="F")


def prepare_ac_generic(GL, W, GR):
    """``prepare_operator!!(::MPO_AC_Hamiltonian)`` (mpo_derivatives.jl:221-231).
    L[(a,b),(f,h,k)] = GL[a,g,f] W[g,b,h,k]; R = GR unchanged."""
    Dl, d, w = GL.shape[0], W.shape[1], W.shape[3]
    L = _es("agf,gbhk->abfhk", GL, W)
    return fmat(L, Dl * d, Dl * d * w).copy(order="F"), F(GR)


def apply_prepared_ac(L, GR, x):
    """``(H::PrecomputedACDerivative)(x)`` (mpo_derivatives.jl:266-274):
    xR[(f,h),(k,c)] = x[(f,h),i] GR[i,(k,c)], y = L xR."""
    Dl, d, Dr = x.shape
    w = GR.shape[1]
    xR = fmat(x, Dl * d, Dr) @ fmat(GR, Dr, w * GR.shape[2])
    y = L @ fmat(xR, Dl * d * w, GR.shape[2])
    return np.reshape(y, (Dl, d, GR.shape[2]), order="F")


def prepare_c(GL, GR):
    """``prepare_operator!!(::MPO_C_Hamiltonian)`` (mpo_derivatives.jl:216-220):
    L = _transpose_tail(GL), i.e. L[a,(f,g)] = GL[a,g,f]."""
    Da, w, Df = GL.shape
    L = np.transpose(GL, (0, 2, 1))
    return fmat(L, Da, Df * w).copy(order="F"), F(GR)


def apply_prepared_c(L, GR, x):
    """``(H::PrecomputedCDerivative)(x)`` (mpo_derivatives.jl:249-264)."""
    Dr, w, Dc = GR.shape
    xR = x @ fmat(GR, Dr, w * Dc)              # [f, (g, c)]
    return L @ fmat(xR, x.shape[0] * w, Dc)


# ---------------------------------------------------------------------------
# Jordan-block MPO structure -- jordanmpotensor.jl:194-249
# ---------------------------------------------------------------------------

def jordan_split(W, left_boundary=None, right_boundary=None):
    """Split a dense Jordan-form W into its blocks (jordanmpotensor.jl:194-249).

    Returns dict with keys C (p_out, p_in, mid_r), B (mid_l, p_out, p_in),
    D (p_out, p_in), A (mid_l, p_out, p_in, mid_r), and channel index lists.
    ``w_l == 1`` means the left bond holds only "not started"; ``w_r == 1``
    means the right bond holds only "finished".
    """
    wl, d, _, wr = W.shape
    s_l = 0
    e_l = wl - 1 if wl > 1 else None
    s_r = 0 if wr > 1 else None
    e_r = wr - 1
    mid_l = list(range(1, wl - 1)) if wl > 1 else []
    mid_r = list(range(1, wr - 1)) if wr > 1 else []
    blk = {
        "D": W[s_l, :, :, e_r],
        "C": W[s_l][:, :, mid_r] if mid_r else None,
        "B": W[mid_l][:, :, :, e_r] if mid_l else None,
        "A": W[np.ix_(mid_l, range(d), range(d), mid_r)] if (mid_l and mid_r) else None,
        "mid_l": mid_l, "mid_r": mid_r, "s_l": s_l, "e_l": e_l, "s_r": s_r, "e_r": e_r,
    }
    return blk


def _nonzero(t):
    return t is not None and t.size > 0 and np.any(t != 0)


def _connected_channels(A1, A2):
    """``_connected_channels`` (hamiltonian_derivatives.jl:243-270), per dense channel."""
    if not (_nonzero(A1) and _nonzero(A2)):
        return None
    fed = np.any(A1 != 0, axis=(0, 1, 2))        # middle channel reached from left
    feeds = np.any(A2 != 0, axis=(1, 2, 3))      # middle channel reaching right
    mids = np.nonzero(fed & feeds)[0]
    if len(mids) == 0:
        return None
    rows = np.nonzero(np.any(A1[..., mids] != 0, axis=(1, 2, 3)))[0]
    cols = np.nonzero(np.any(A2[mids] != 0, axis=(0, 1, 2)))[0]
    return rows, mids, cols


def jordan_ac2_construct(GL, W1, W2, GR):
    """``JordanMPO_AC2_Hamiltonian(GL, W1, W2, GR)`` constructor
    (hamiltonian_derivatives.jl:272-369): the unprepared terms (None ==
    ``missing``) in the reference's tensor layouts:
      II[c,i], IC[e,c,j,i], ID[e,j], CB[b,e,h,j], CA[b,e,c,h,j,i],
      AB[a,b,e,f,h,j], BE[a,b,f,h], DE[b,h], EE[a,f], AA=(GL2,A1,A2,GR2) or None.
    """
    b1 = jordan_split(W1)
    b2 = jordan_split(W2)
    d1, d2 = W1.shape[1], W2.shape[1]
    GL2 = GL[:, b1["mid_l"], :] if b1["mid_l"] else None
    GR2 = GR[:, b2["mid_r"], :] if b2["mid_r"] else None

    II = GR[:, 0, :].T.copy() if W2.shape[3] > 1 else None         # :291
    EE = GL[:, -1, :].copy() if W1.shape[0] > 1 else None          # :294
    C2, B1, C1, B2, A1, A2 = b2["C"], b1["B"], b1["C"], b2["B"], b1["A"], b2["A"]
    IC = _es("ejk,ikc->ecji", C2, GR2) if _nonzero(C2) else None    # :297-302
    DE = b1["D"].copy() if _nonzero(b1["D"]) else None             # :305
    ID = b2["D"].copy() if _nonzero(b2["D"]) else None             # :308
    CB = _es("bhk,kej->behj", C1, B2) if (_nonzero(C1) and _nonzero(B2)) else None
    CA = (_es("bhk,kejl,ilc->bechji", C1, A2, GR2)
          if (_nonzero(C1) and _nonzero(A2)) else None)             # :320-326
    AB = (_es("alf,lbhk,kej->abefhj", GL2, A1, B2)
          if (_nonzero(A1) and _nonzero(B2)) else None)             # :328-334
    BE = _es("agf,gbh->abfh", GL2, B1) if _nonzero(B1) else None    # :336-341
    ch = _connected_channels(A1, A2)                                # :352-362
    AA = None
    if ch is not None:
        rows, mids, cols = ch
        AA = (GL2[:, rows, :], A1[np.ix_(rows, range(d1), range(d1), mids)],
              A2[np.ix_(mids, range(d2), range(d2), cols)], GR2[:, cols, :])
    return dict(II=II, IC=IC, ID=ID, CB=CB, CA=CA, AB=AB, AA=AA, BE=BE, DE=DE, EE=EE)


def jordan_ac2_fold(T):
    """``prepare_operator!!(::JordanMPO_AC2_Hamiltonian)`` (hamiltonian_derivatives.jl:371-468):
    folds the small terms into the large ones with ``+= Id (x) X``, in the
    reference's order and with its fallbacks.  Returns a new dict."""
    II, IC, ID, CB, CA, AB, AA, BE, DE, EE = (T[k] for k in ("II", "IC", "ID", "CB", "CA", "AB", "AA", "BE",
                                                            "DE", "EE"))
    eye = np.eye
    if CA is not None and CB is not None:                           # :379-389
        CA = CA + _es("behj,ci->bechji", CB, eye(CA.shape[2])); CB = None
    elif AB is not None and CB is not None:
        AB = AB + _es("behj,af->abefhj", CB, eye(AB.shape[0])); CB = None
    if CA is not None and IC is not None:                           # :391-398
        CA = CA + _es("bh,ecji->bechji", eye(CA.shape[0]), IC); IC = None
    if AB is not None and BE is not None:                           # :400-407
        AB = AB + _es("abfh,ej->abefhj", BE, eye(AB.shape[2])); BE = None
    if BE is not None and DE is not None:                           # :409-421
        BE = BE + _es("af,bh->abfh", eye(BE.shape[0]), DE); DE = None
    elif AB is not None and DE is not None:
        AB = AB + _es("af,bh,ej->abefhj", eye(AB.shape[0]), DE, eye(AB.shape[2])); DE = None
    if IC is not None and ID is not None:                           # :423-434
        IC = IC + _es("ej,ci->ecji", ID, eye(IC.shape[1])); ID = None
    elif CA is not None and ID is not None:
        CA = CA + _es("bh,ej,ci->bechji", eye(CA.shape[0]), ID, eye(CA.shape[2])); ID = None
    if IC is not None and II is not None:                           # :436-446
        IC = IC + _es("ej,ci->ecji", eye(IC.shape[0]), II); II = None
    elif CA is not None and II is not None:
        CA = CA + _es("bh,ej,ci->bechji", eye(CA.shape[0]), eye(CA.shape[1]), II); II = None
    if BE is not None and EE is not None:                           # :448-458
        BE = BE + _es("af,bh->abfh", EE, eye(BE.shape[1])); EE = None
    elif AB is not None and EE is not None:
        AB = AB + _es("af,bh,ej->abefhj", EE, eye(AB.shape[1]), eye(AB.shape[2])); EE = None
    return dict(II=II, IC=IC, ID=ID, CB=CB, CA=CA, AB=AB, AA=AA, BE=BE, DE=DE, EE=EE)


def jordan_ac2_prepare(GL, W1, W2, GR):
    """``AC2_hamiltonian(...; prepare = true)`` (hamiltonian_derivatives.jl:206-217):
    the constructor followed by ``prepare_operator!!``."""
    return jordan_ac2_fold(jordan_ac2_construct(GL, W1, W2, GR))


def jordan_ac2_apply(T, x):
    """``(H::JordanMPO_AC2_Hamiltonian)(x)`` (hamiltonian_derivatives.jl:485-500)."""
    y = np.zeros_like(x) if T["AA"] is None else ac2_apply_generic(*T["AA"], x)
    if T["II"] is not None: y = y + _es("abie,ci->abce", x, T["II"])
    if T["IC"] is not None: y = y + _es("abij,ecji->abce", x, T["IC"])
    if T["ID"] is not None: y = y + _es("abcj,ej->abce", x, T["ID"])
    if T["CB"] is not None: y = y + _es("ahcj,behj->abce", x, T["CB"])
    if T["CA"] is not None: y = y + _es("ahij,bechji->abce", x, T["CA"])
    if T["AB"] is not None: y = y + _es("fhcj,abefhj->abce", x, T["AB"])
    if T["BE"] is not None: y = y + _es("fhce,abfh->abce", x, T["BE"])
    if T["DE"] is not None: y = y + _es("ahce,bh->abce", x, T["DE"])
    if T["EE"] is not None: y = y + _es("fbce,af->abce", x, T["EE"])
    return y


def jordan_ac_construct(GL, W, GR):
    """``JordanMPO_AC_Hamiltonian(GL, W, GR)`` constructor
    (hamiltonian_derivatives.jl:57-107): unprepared terms D[b,h], I[i,c],
    E[a,f], C[b,c,h,i], B[a,b,f,h], A=(GL2,WA,GR2) or None (None == missing)."""
    blk = jordan_split(W)
    GL2 = GL[:, blk["mid_l"], :] if blk["mid_l"] else None
    GR2 = GR[:, blk["mid_r"], :] if blk["mid_r"] else None
    Dm = blk["D"].copy() if _nonzero(blk["D"]) else None                # :68
    I = GR[:, 0, :].copy() if W.shape[3] > 1 else None                  # :71
    E = GL[:, -1, :].copy() if W.shape[0] > 1 else None                 # :74
    C = _es("bhk,ikc->bchi", blk["C"], GR2) if _nonzero(blk["C"]) else None    # :77-82
    B = _es("akf,kbh->abfh", GL2, blk["B"]) if _nonzero(blk["B"]) else None    # :85-90
    A = (GL2, blk["A"], GR2) if _nonzero(blk["A"]) else None                   # :93, :104
    return dict(D=Dm, I=I, E=E, C=C, B=B, A=A)


def jordan_ac_fold(T):
    """``prepare_operator!!(::JordanMPO_AC_Hamiltonian)`` (hamiltonian_derivatives.jl:109-157):
    D into C (else B), I into C, E into B."""
    Dm, I, E, C, B, A = (T[k] for k in ("D", "I", "E", "C", "B", "A"))
    if Dm is not None:
        if C is not None:
            C = C + _es("bh,ci->bchi", Dm, np.eye(C.shape[1])); Dm = None
        elif B is not None:
            B = B + _es("af,bh->abfh", np.eye(B.shape[0]), Dm); Dm = None
    if I is not None and C is not None:
        C = C + _es("bh,ic->bchi", np.eye(C.shape[0]), I); I = None
    if E is not None and B is not None:
        B = B + _es("af,bh->abfh", E, np.eye(B.shape[1])); E = None
    return dict(D=Dm, I=I, E=E, C=C, B=B, A=A)


def jordan_ac_prepare(GL, W, GR):
    """``AC_hamiltonian(...; prepare = true)`` (hamiltonian_derivatives.jl:44-55).  Layouts:
    D[b,h], I[i,c], E[a,f], C[b,c,h,i], B[a,b,f,h], A=(GL2,WA,GR2) or None."""
    return jordan_ac_fold(jordan_ac_construct(GL, W, GR))


def jordan_ac_apply(T, x):
    """``(H::JordanMPO_AC_Hamiltonian)(x)`` (hamiltonian_derivatives.jl:472-483)."""
    y = np.zeros_like(x) if T["A"] is None else ac_apply_generic(T["A"][0], T["A"][1], T["A"][2], x)
    if T["D"] is not None: y = y + _es("ahc,bh->abc", x, T["D"])
    if T["E"] is not None: y = y + _es("af,fbc->abc", T["E"], x)
    if T["I"] is not None: y = y + _es("abi,ic->abc", x, T["I"])
    if T["C"] is not None: y = y + _es("ahi,bchi->abc", x, T["C"])
    if T["B"] is not None: y = y + _es("abfh,fhc->abc", T["B"], x)
    return y


# ---------------------------------------------------------------------------
# Transfer matrices -- transfer.jl, transfermatrix.jl
# ---------------------------------------------------------------------------

def transfer_left_mps(v, A, Ab):
    """transfer_left(v, A, Abar) for a bond tensor v (transfer.jl:18-31):
    y[a,c] = v[m,n] A[n,h,c] conj(Ab[m,h,a])."""
    return _es("mhc,mha->ac", _es("mn,nhc->mhc", v, A), np.conj(Ab))


def transfer_right_mps(v, A, Ab):
    """transfer_right(v, A, Abar) (transfer.jl:43-55): y[a,b] = A[a,h,i] conj(Ab[b,h,j]) v[i,j]."""
    return _es("ahj,bhj->ab", _es("ahi,ij->ahj", A, v), np.conj(Ab))


def transfer_left_mpo(x, O, A, Ab):
    """MPO transfer_left (transfer.jl:125-127):
    y[-1 -2; -3] = x[1 2; 4] A[4 5; -3] O[2 3; 5 -2] conj(Ab[1 3; -1])."""
    t = _es("mgn,nhc->mghc", x, A)
    t = _es("mghc,gkhb->mkbc", t, O)
    return _es("mkbc,mka->abc", t, np.conj(Ab))


def transfer_right_mpo(v, O, A, Ab):
    """MPO transfer_right (transfer.jl:128-130):
    y[-1 -2; -3] = A[-1 2; 1] O[-2 4; 2 3] conj(Ab[-3 4; 5]) v[1 3; 5]."""
    t = _es("ahi,ilo->ahlo", A, v)
    t = _es("ahlo,bkhl->abko", t, O)
    return _es("abko,cko->abc", t, np.conj(Ab))


def regularize(v, lvec, rvec):
    """``regularize!`` (transfermatrix.jl:70-76): bond tensor v - tr(lvec v) rvec;
    for an environment-shaped v[a,g,c] the projection acts on each channel g."""
    if v.ndim == 2:
        return v - np.einsum("ij,ji->", lvec, v) * rvec
    lam = np.einsum("ij,jgi->g", lvec, v)
    return v - np.einsum("g,ac->agc", lam, rvec)


# ---------------------------------------------------------------------------
# Prepared nearest-neighbour Jordan apply in GEMM form (the CPU baseline path)
# ---------------------------------------------------------------------------

def jordan_ac2_nn_matrices(T, Dl, d1, Dr, d2):
    """GEMM-ready matrices of a prepared Jordan operator without A blocks:
    ICm[(i,j),(c,e)] = IC[e,c,j,i], BEm[(a,b),(f,h)] = BE[a,b,f,h]
    (terms of hamiltonian_derivatives.jl:490 and :495)."""
    assert T["CA"] is None and T["AB"] is None and T["AA"] is None
    ICm = np.transpose(T["IC"], (3, 2, 1, 0)).reshape(Dr * d2, Dr * d2, order="F")
    BEm = T["BE"].reshape(Dl * d1, Dl * d1, order="F")
    return np.asfortranarray(ICm), np.asfortranarray(BEm), T["CB"]


def jordan_ac2_apply_nn(ICm, BEm, CB, x):
    """The prepared nearest-neighbour Jordan two-site apply as two (dD)x(dD)
    GEMMs plus the d^2 x d^2 physical term (hamiltonian_derivatives.jl:485-500)."""
    Dl, d1, Dr, d2 = x.shape
    m = x.reshape(Dl * d1, Dr * d2, order="F")
    y = m @ ICm
    y += BEm @ m
    y = y.reshape(Dl, d1, Dr, d2, order="F")
    if CB is not None:
        # y[a,b,c,e] += sum_{h,j} x[a,h,c,j] CB[b,e,h,j], as one BLAS product
        # over the d^2 physical pairs
        y += np.tensordot(x, CB, axes=([1, 3], [2, 3])).transpose(0, 2, 1, 3)
    return y
