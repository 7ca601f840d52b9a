/*
 * mpskit_hip.h -- C ABI of the MI355X (gfx950) effective-Hamiltonian and
 * transfer-matrix contraction library.
 *
 * Drop-in boundary for MPSKit.jl v0.13.13's derivative operators, transfer
 * matrices and environment updates.  Every entry point below names the
 * reference routine it replaces (paths relative to the MPSKit.jl tree).
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer to float64 data in the memory
 *     layout of a Trivial-symmetry TensorKit.TensorMap: one column-major array
 *     with axes (codomain..., domain...):
 *        GL[a,g,f]   (D_bra x w <- D_ket)        src/environments/abstract_envs.jl:36-37
 *        GR[i,l,c]   (D_ket x w <- D_bra)        src/environments/abstract_envs.jl:50-51
 *        W[g,b,h,k]  (w_l x p_out <- p_in x w_r)  src/states/abstractmps.jl:8-11
 *        A[l,p,r]    (D_l x p <- D_r)
 *        AC2 x[f,h,i,j] (D_l x p1 <- D_r x p2)    src/states/finitemps.jl:376-384
 *     so a Julia caller passes `pointer(t.data)` of a device-adapted tensor.
 *   - MPO tensors are dense (materialised JordanMPOTensor / SparseBlockTensorMap).
 *     For the Jordan ("MPOHamiltonian") entry points, channel 0 of an MPO bond
 *     is "not started", the last channel "finished", the channels in between
 *     the A-block (src/operators/jordanmpotensor.jl:1-69); a left boundary
 *     bond of dimension 1 holds only "not started", a right boundary bond of
 *     dimension 1 only "finished".
 *   - Work is enqueued on the context's stream (mk_ctx_set_stream); calls
 *     return once enqueued unless documented otherwise.  Prepared operator
 *     handles are immutable; applies of operators that share a context must
 *     be issued from one host thread at a time (the context owns the scratch
 *     arena, which is stream-ordered).  Use one context per concurrent caller.
 *   - Status codes: MK_OK on success; otherwise mk_last_error() returns a
 *     thread-local message.  Mirrors the reference's SpaceMismatch /
 *     ArgumentError throws (src/utility/utility.jl:276-277, 359-360).
 */
#ifndef MPSKIT_HIP_H
#define MPSKIT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mk_status {
  MK_OK = 0,
  MK_BAD_SHAPE = 1,      /* SpaceMismatch analogue */
  MK_BAD_ARG = 2,        /* ArgumentError analogue */
  MK_OOM = 3,
  MK_HIP_ERROR = 4,
  MK_NOT_CONVERGED = 5,  /* KrylovKit non-convergence warning analogue */
  MK_NOT_SUPPORTED = 6,
  MK_NONFINITE = 7       /* NaN/Inf met in an input or iterate */
} mk_status;

/* Arithmetic used for the bulk contractions of an apply.  FP64 is the
 * reference arithmetic; FP32 runs the GEMMs on FP32 MFMA; BF16X3 splits each
 * float64 operand into bf16 hi+lo and accumulates hi*hi + hi*lo + lo*hi in
 * FP32 on BF16 MFMA.  Inputs and outputs stay float64 in every mode. */
typedef enum mk_precision { MK_FP64 = 0, MK_FP32 = 1, MK_BF16X3 = 2 } mk_precision;
/* bit masks for mk_prepare_* "precisions" arguments */
#define MK_PREC_FP64 (1 << MK_FP64)
#define MK_PREC_FP32 (1 << MK_FP32)
#define MK_PREC_BF16X3 (1 << MK_BF16X3)

typedef struct mk_ctx_s* mk_ctx;
typedef struct mk_op_s* mk_op;

/* ---- library / context -------------------------------------------------- */
int mk_version(void);
const char* mk_last_error(void);
mk_status mk_ctx_create(int device, mk_ctx* out);
mk_status mk_ctx_destroy(mk_ctx ctx);
/* stream is a hipStream_t (NULL = legacy default stream) */
mk_status mk_ctx_set_stream(mk_ctx ctx, void* stream);
mk_status mk_ctx_synchronize(mk_ctx ctx);

/* ---- derivative operators (prepare) -------------------------------------- */

/* Zero-site derivative, replaces MPO_C_Hamiltonian + prepare_operator!!
 * (src/algorithms/derivatives/mpo_derivatives.jl:37-44, :92-97, :216-220,
 *  PrecomputedCDerivative :249-264).
 * GL (Da x w x Df), GR (Di x w x Dc);  x (Df x Di) -> y (Da x Dc). */
mk_status mk_prepare_c(mk_ctx ctx, const double* GL, const double* GR, int64_t Da,
                       int64_t Df, int64_t w, int64_t Di, int64_t Dc, int precisions,
                       mk_op* out);

/* Single-site generic derivative, replaces MPO_AC_Hamiltonian + prepare
 * (mpo_derivatives.jl:46-54, :99-106, :221-231, PrecomputedACDerivative :266-274).
 * GL (Da x wl x Df), W (wl x d x d x wr), GR (Di x wr x Dc);
 * x (Df x d x Di) -> y (Da x d x Dc). */
mk_status mk_prepare_ac_generic(mk_ctx ctx, const double* GL, const double* W,
                                const double* GR, int64_t Da, int64_t Df, int64_t wl,
                                int64_t d, int64_t wr, int64_t Di, int64_t Dc,
                                int precisions, mk_op* out);

/* Two-site generic derivative, replaces MPO_AC2_Hamiltonian + prepare
 * (mpo_derivatives.jl:56-66, :144-152, :233-247, PrecomputedAC2Derivative :276-288).
 * GL (Da x wl x Df), W1 (wl x d1 x d1 x wm), W2 (wm x d2 x d2 x wr), GR (Di x wr x Dc);
 * x (Df x d1 x Di x d2) -> y (Da x d1 x Dc x d2). */
mk_status mk_prepare_ac2_generic(mk_ctx ctx, const double* GL, const double* W1,
                                 const double* W2, const double* GR, int64_t Da,
                                 int64_t Df, int64_t wl, int64_t d1, int64_t wm,
                                 int64_t d2, int64_t wr, int64_t Di, int64_t Dc,
                                 int precisions, mk_op* out);

/* Single-site Jordan-block (MPOHamiltonian) derivative, replaces
 * JordanMPO_AC_Hamiltonian ctor + prepare_operator!!
 * (src/algorithms/derivatives/hamiltonian_derivatives.jl:44-55, :57-107, :109-157,
 *  apply :472-483).  Bra == ket: GL (Dl x wl x Dl), W (wl x d x d x wr),
 * GR (Dr x wr x Dr); x (Dl x d x Dr) -> y (Dl x d x Dr).  As in the reference
 * the "not started" left level and the "finished" right level of the
 * environments are taken to be identities. */
mk_status mk_prepare_ac_jordan(mk_ctx ctx, const double* GL, const double* W,
                               const double* GR, int64_t Dl, int64_t wl, int64_t d,
                               int64_t wr, int64_t Dr, int precisions, mk_op* out);

/* Two-site Jordan-block derivative, replaces JordanMPO_AC2_Hamiltonian ctor +
 * prepare_operator!! (hamiltonian_derivatives.jl:206-217, :243-270, :272-369,
 * :371-468; apply :485-500).  GL (Dl x wl x Dl), W1 (wl x d1 x d1 x wm),
 * W2 (wm x d2 x d2 x wr), GR (Dr x wr x Dr); x (Dl x d1 x Dr x d2) -> y. */
mk_status mk_prepare_ac2_jordan(mk_ctx ctx, const double* GL, const double* W1,
                                const double* W2, const double* GR, int64_t Dl,
                                int64_t wl, int64_t d1, int64_t wm, int64_t d2,
                                int64_t wr, int64_t Dr, int precisions, mk_op* out);

/* ---- Jordan operators from their constructed terms --------------------------
 * The reference builds a JordanMPO_AC2_Hamiltonian / JordanMPO_AC_Hamiltonian
 * (the constructor contracts the environments with the MPO blocks) and then
 * calls the one-argument prepare_operator!!(H) on it (hamiltonian_derivatives.jl
 * :206-217 and :44-55).  These entry points replace that prepare_operator!!:
 * they take the constructed terms, device arrays in the layouts of the struct
 * fields (NULL = `missing`), fold them as prepare_operator!! does and return
 * a prepared operator for mk_apply / mk_eigsolve / batches.
 *
 * Two-site (hamiltonian_derivatives.jl:168-200 fields, :272-369 constructor,
 * :371-468 prepare):
 *   II[c,i] (Dr x Dr), IC[e,c,j,i] (d2 x Dr x d2 x Dr), ID[e,j] (d2 x d2),
 *   CB[b,e,h,j] (d1 x d2 x d1 x d2), CA[b,e,c,h,j,i] (d1 x d2 x Dr x d1 x d2 x Dr),
 *   AB[a,b,e,f,h,j] (Dl x d1 x d2 x Dl x d1 x d2), BE[a,b,f,h] (Dl x d1 x Dl x d1),
 *   DE[b,h] (d1 x d1), EE[a,f] (Dl x Dl);
 *   AA, the MPO_AC2_Hamiltonian over the connected channels (:352-362), as its
 *   four dense tensors: AA_GL (Dl x aa_wl x Dl), AA_W1 (aa_wl x d1 x d1 x aa_wm),
 *   AA_W2 (aa_wm x d2 x d2 x aa_wr), AA_GR (Dr x aa_wr x Dr).
 * One-site (hamiltonian_derivatives.jl:15-32 fields, :57-107 constructor,
 * :109-157 prepare):
 *   D[b,h] (d x d), I[i,c] (Dr x Dr), E[a,f] (Dl x Dl), C[b,c,h,i] (d x Dr x d x Dr),
 *   B[a,b,f,h] (Dl x d x Dl x d);  A, the MPO_AC_Hamiltonian over the middle
 *   channels: A_GL (Dl x a_wl x Dl), A_W (a_wl x d x d x a_wr), A_GR (Dr x a_wr x Dr). */
typedef struct mk_jordan_ac2_terms {
  const double *II, *IC, *ID, *CB, *CA, *AB, *BE, *DE, *EE;
  const double *AA_GL, *AA_W1, *AA_W2, *AA_GR;
  int64_t aa_wl, aa_wm, aa_wr;
} mk_jordan_ac2_terms;
typedef struct mk_jordan_ac_terms {
  const double *D, *I, *E, *C, *B;
  const double *A_GL, *A_W, *A_GR;
  int64_t a_wl, a_wr;
} mk_jordan_ac_terms;
mk_status mk_prepare_ac2_jordan_terms(mk_ctx ctx, const mk_jordan_ac2_terms* terms, int64_t Dl,
                                      int64_t d1, int64_t Dr, int64_t d2, int precisions, mk_op* out);
mk_status mk_prepare_ac_jordan_terms(mk_ctx ctx, const mk_jordan_ac_terms* terms, int64_t Dl, int64_t d,
                                     int64_t Dr, int precisions, mk_op* out);

/* ---- block-stored environments and MPO tensors -------------------------------
 * The layouts the reference's MPOHamiltonian path holds, passed without any
 * densification by the caller:
 *   - an environment is a BlockTensorMap over a SumSpace on its MPO leg, one
 *     TensorMap per Jordan level (src/environments/abstract_envs.jl:32-58):
 *     mk_env_levels lists the device pointers of the levels' `.data`, level L a
 *     column-major D1 x width[L] x D2 array; the levels in channel order;
 *   - a JordanMPOTensor stores its nonzero blocks, identity blocks as scalars
 *     (src/operators/jordanmpotensor.jl:44-69): mk_jordan_mpo lists, per block,
 *     its channel position (row, col) and either a device pointer to the d x d
 *     block (out x in, column-major) or NULL and a scalar (scalar x identity).
 * The library gathers the levels on the device (one strided copy per level,
 * stream-ordered) and assembles the MPO tensor; outputs are scattered back into
 * the caller's levels.  These calls synchronise the context's stream before
 * they return.  Otherwise each equals its dense counterpart, bit for bit. */
typedef struct mk_env_levels {
  int n;                      /* number of levels (1..64) */
  double* const* ptr;         /* host array of n device pointers */
  const int64_t* width;       /* host array of n channel counts */
} mk_env_levels;
typedef struct mk_jordan_mpo {
  int64_t wl, d, wr;          /* channels of the left / right MPO bond, physical dimension */
  int nblocks;
  const int* row;             /* host arrays of nblocks entries */
  const int* col;
  const double* const* block; /* device pointer of the d x d block, or NULL: scalar x identity */
  const double* scalar;
} mk_jordan_mpo;
/* mk_prepare_ac2_jordan / mk_prepare_ac_jordan / mk_prepare_c with block-stored
 * environments and MPO tensors (hamiltonian_derivatives.jl:206-217, :44-55;
 * mpo_derivatives.jl:37-44, where the reference densifies GL / GR itself). */
mk_status mk_prepare_ac2_jordan_levels(mk_ctx ctx, const mk_env_levels* GL, const mk_jordan_mpo* W1,
                                       const mk_jordan_mpo* W2, const mk_env_levels* GR, int64_t Dl,
                                       int64_t Dr, int precisions, mk_op* out);
mk_status mk_prepare_ac_jordan_levels(mk_ctx ctx, const mk_env_levels* GL, const mk_jordan_mpo* W,
                                      const mk_env_levels* GR, int64_t Dl, int64_t Dr, int precisions,
                                      mk_op* out);
mk_status mk_prepare_c_levels(mk_ctx ctx, const mk_env_levels* GL, const mk_env_levels* GR, int64_t Da,
                              int64_t Df, int64_t Di, int64_t Dc, int precisions, mk_op* out);
/* Environment updates of the MPOHamiltonian path (transfer.jl:125-130 on
 * BlockTensorMap environments, finite_envs.jl:110-150): x / v and the result y
 * are block-stored; A, Abar as in mk_transfer_left_mpo / mk_transfer_right_mpo. */
mk_status mk_transfer_left_mpo_levels(mk_ctx ctx, const mk_env_levels* x, const mk_jordan_mpo* O,
                                      const double* A, const double* Abar, int64_t Dm, int64_t Dn,
                                      int64_t Dc, int64_t Da, const mk_env_levels* y);
mk_status mk_transfer_right_mpo_levels(mk_ctx ctx, const mk_env_levels* v, const mk_jordan_mpo* O,
                                       const double* A, const double* Abar, int64_t Da, int64_t Di,
                                       int64_t Dc, int64_t Do, const mk_env_levels* y);

/* Two classical Gram-Schmidt passes of u against the k <= 32 orthonormal
 * columns of V (column-major, leading dim ldv; device pointers): u becomes
 * u - V (c1 + c2), h[0..k) = c1 + c2 and h[k] = |u| (host array of k + 1).
 * The orthogonalisation step of the Krylov solvers around the library
 * (KrylovKit's ClassicalGramSchmidt2 under the Arnoldi / GMRES / Lanczos calls
 * of infinite_envs.jl:207-346, ortho.jl:307-313, transfer_spectrum.jl:27-60,
 * tdvp.jl:72-133); no counterpart entry point in the reference. */
mk_status mk_krylov_orthogonalize(mk_ctx ctx, const double* V, int64_t ldv, int k, double* u, int64_t n,
                                  double* h);

/* Cholesky factorisation A = L L^T of an n x n symmetric positive definite
 * float64 matrix (column-major, leading dim lda, device pointer), in place:
 * the lower triangle becomes L and the strict upper triangle zero.  *info
 * (host) is 0 on success, else the 1-based column of the first non-positive
 * pivot (then A holds a partial result).  The factorisation inside the
 * Cholesky QR of the gauge steps' positive QR / LQ (qr_compact / lq_compact
 * with positive = true in ortho.jl:215-340 and the finite-MPS gauge moves of
 * finitemps.jl; the reference factors by Householder QR in LAPACK, which this
 * library replaces by shifted Cholesky QR3); no counterpart entry point in the
 * reference. */
mk_status mk_potrf(mk_ctx ctx, double* A, int64_t n, int64_t lda, int* info);

/* ---- U(1) charge blocks ------------------------------------------------------
 * Counterpart of TensorKit's block-sparse U(1) storage under the same
 * derivative operators (the reference applies one GEMM per fusion-tree block,
 * src/utility/utility.jl:285-302).  Tensors stay dense arrays in a
 * charge-sorted basis (each bond index space is a concatenation of charge
 * sectors); this call tells a prepared nearest-neighbour two-site Jordan
 * operator which blocks conserve charge, after which its applies (and batches
 * holding it) compute only those blocks, all of them in one grouped persistent
 * GEMM launch.  Left bond sectors are [loff[s], loff[s+1]) with charge lq[s]
 * (increasing), right bond sectors likewise, pq[h] the charge of physical
 * state h; x[f,h,i,j] may be nonzero only if lq(f) + pq(h) + pq(j) = rq(i),
 * and y is returned with the same structure (elements outside the blocks are
 * the physical term's).  Mixed precisions need every offset a multiple of 8. */
mk_status mk_op_set_charge_blocks(mk_op op, int nl, const int64_t* loff, const int* lq, int nr,
                                  const int64_t* roff, const int* rq, const int* pq);

/* Compact storage of the charge blocks (TensorKit's block-sparse storage, one
 * dense block per charge sector, src/utility/utility.jl:283-303): after
 * mk_op_set_charge_blocks, x and y of every apply hold only the blocks that
 * conserve charge, block b column-major (m_b x n_b) at element offset off_b
 * (mk_op_charge_blocks lists them), so the Krylov vectors, the physical term
 * and the mixed-precision conversions touch the blocks only.  mk_op_info then
 * reports the compact sizes.  No counterpart call in the reference: its tensors
 * are block-sparse from the start. */
mk_status mk_op_compact_charge_blocks(mk_op op);
/* The blocks of a charge-blocked operator: 5 entries per block (dense row
 * offset r0 and rows m of the fused (Dl d1) side, column offset c0 and columns
 * n of the fused (Dr d2) side, compact element offset), up to `cap` blocks;
 * *nblocks receives the count.  out may be NULL to query the count. */
mk_status mk_op_charge_blocks(mk_op op, int64_t* out, int cap, int* nblocks);

/* ---- derivative operators (apply) ---------------------------------------- */

/* y = alpha * H(x) + beta * y using arithmetic `precision` (which must have
 * been requested at prepare).  Replaces `(H::DerivativeOperator)(x)` / `H * x`
 * (src/algorithms/derivatives/derivatives.jl:8-11).  x and y must not alias. */
mk_status mk_apply(mk_op op, const double* x, double* y, double alpha, double beta,
                   int precision);

/* Allocation-free variant of mk_apply.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * operators and results as mk_apply, but the scratch of the apply (the
 * FP32 / BF16X3 operand views of x, the intermediate of a generic operator)
 * comes from a fixed device array of MK_NOALLOC_SCRATCH_MB MiB that the
 * library holds on every GPU it runs on until the process ends (its address
 * is resolved at mk_ctx_create); the call allocates nothing, creates no
 * events and does not synchronise.  An apply whose scratch exceeds the array
 * returns MK_BAD_SHAPE (160 MiB hold the BF16X3 views of a two-site Jordan
 * operator up to d*D = 4096, i.e. D = 2048 at d = 2; FP64 applies of Jordan
 * operators need none), and operators with charge blocks return
 * MK_NOT_SUPPORTED.  All calls on one GPU share the array: calls that overlap
 * on one GPU (different streams or threads) race; calls ordered on one stream
 * do not.  Nothing in the array needs to start at zero. */
#define MK_NOALLOC_SCRATCH_MB 160
mk_status mk_apply_noalloc(mk_op op, const double* x, double* y, double alpha, double beta,
                           int precision);

/* number of float64 elements of x and y, and the algorithmic flop count of one
 * apply (each contraction counted once, whatever the precision scheme) */
mk_status mk_op_info(mk_op op, int64_t* x_elems, int64_t* y_elems, double* flops);
mk_status mk_op_free(mk_op op);

/* ---- transfer matrices / environment updates ---------------------------- */

/* transfer_left(v, A, Abar), bond-tensor v (src/transfermatrix/transfer.jl:18-31):
 * y[a,c] = v[m,n] A[n,h,c] Abar[m,h,a];  v (Dm x Dn), A (Dn x d x Dc),
 * Abar (Dm x d x Da) -> y (Da x Dc). */
mk_status mk_transfer_left_mps(mk_ctx ctx, const double* v, const double* A,
                               const double* Abar, int64_t Dm, int64_t Dn, int64_t d,
                               int64_t Dc, int64_t Da, double* y);

/* transfer_right(v, A, Abar) (transfer.jl:43-55):
 * y[a,b] = A[a,h,i] Abar[b,h,j] v[i,j];  A (Da x d x Di), Abar (Db x d x Dj),
 * v (Di x Dj) -> y (Da x Db). */
mk_status mk_transfer_right_mps(mk_ctx ctx, const double* v, const double* A,
                                const double* Abar, int64_t Da, int64_t d, int64_t Di,
                                int64_t Db, int64_t Dj, double* y);

/* MPO transfer_left(x, O, A, Abar) (transfer.jl:125-127); the finite left
 * environment update GLs[j+1] = GLs[j] * TransferMatrix(AL, W, AL)
 * (src/environments/finite_envs.jl:137-150, transfermatrix.jl:48-54):
 * y[a,b,c] = x[m,g,n] A[n,h,c] O[g,k,h,b] Abar[m,k,a];
 * x (Dm x wl x Dn), O (wl x d x d x wr), A (Dn x d x Dc), Abar (Dm x d x Da)
 * -> y (Da x wr x Dc). */
mk_status mk_transfer_left_mpo(mk_ctx ctx, const double* x, const double* O,
                               const double* A, const double* Abar, int64_t Dm,
                               int64_t wl, int64_t Dn, int64_t d, int64_t wr,
                               int64_t Dc, int64_t Da, double* y);

/* MPO transfer_right(v, O, A, Abar) (transfer.jl:128-130); the finite right
 * environment update (finite_envs.jl:110-128):
 * y[a,b,c] = A[a,h,i] O[b,k,h,l] Abar[c,k,o] v[i,l,o];
 * A (Da x d x Di), O (wl x d x d x wr), Abar (Dc x d x Do), v (Di x wr x Do)
 * -> y (Da x wl x Dc). */
mk_status mk_transfer_right_mpo(mk_ctx ctx, const double* v, const double* O,
                                const double* A, const double* Abar, int64_t Da,
                                int64_t d, int64_t Di, int64_t wl, int64_t wr,
                                int64_t Dc, int64_t Do, double* y);

/* Quasiparticle transfers and transfers with legs passing through.  A
 * quasiparticle (excited) tensor A carries an excitation leg e between its
 * physical leg and its right bond, A (Dn x d x De x Dc) as the reference's
 * A[l p; e r].  O = NULL stands for the identity MPO on the passenger leg of
 * x / v (then wl == wr; wl == 1 is the MPS-only transfer).  All FP64, on the
 * context stream; braidings are trivial (bosonic spaces).
 *
 * transfer_left with an excited upper tensor: transfer.jl:58-60 (x a bond
 * tensor: wl = 1, O = NULL), :66-69 (x with a passing leg g: wl = Dg,
 * O = NULL), :133-135 (MPO transfer):
 *   y[a,b,e,c] = x[m,g,n] A[n,h,e,c] O[g,k,h,b] Abar[m,k,a];
 *   x (Dm x wl x Dn), O (wl x d x d x wr), Abar (Dm x d x Da) -> y (Da x wr x De x Dc). */
mk_status mk_transfer_left_excited(mk_ctx ctx, const double* x, const double* O, const double* A,
                                   const double* Abar, int64_t Dm, int64_t wl, int64_t Dn, int64_t d,
                                   int64_t De, int64_t wr, int64_t Dc, int64_t Da, double* y);

/* transfer_right with an excited upper tensor: transfer.jl:61-63, :70-73,
 * :136-138:
 *   y[a,b,e,c] = A[a,h,e,i] O[b,k,h,l] Abar[c,k,o] v[i,l,o];
 *   A (Da x d x De x Di), O (wl x d x d x wr), Abar (Db x d x Dj),
 *   v (Di x wr x Dj) -> y (Da x wl x De x Db). */
mk_status mk_transfer_right_excited(mk_ctx ctx, const double* v, const double* O, const double* A,
                                    const double* Abar, int64_t Da, int64_t d, int64_t De, int64_t Di,
                                    int64_t wl, int64_t wr, int64_t Db, int64_t Dj, double* y);

/* transfer_left of an environment with an excitation (or utility) leg e in
 * its domain: transfer.jl:101-104 (O = NULL), :141-143 (MPO transfer; the
 * left quasiparticle environments, src/environments/qp_envs.jl:149-153):
 *   y[a,b,e,c] = x[m,g,e,n] A[n,h,c] O[g,k,h,b] Abar[m,k,a];
 *   x (Dm x wl x De x Dn), A (Dn x d x Dc), Abar (Dm x d x Da) -> y (Da x wr x De x Dc).
 * With De = 1 and O = NULL this is also the transfer of a density matrix
 * with a utility leg (transfer.jl:76-79). */
mk_status mk_transfer_left_leg(mk_ctx ctx, const double* x, const double* O, const double* A,
                               const double* Abar, int64_t Dm, int64_t wl, int64_t De, int64_t Dn,
                               int64_t d, int64_t wr, int64_t Dc, int64_t Da, double* y);

/* transfer_right counterpart: transfer.jl:105-108, :144-148 (qp_envs.jl:156-160):
 *   y[a,b,e,c] = A[a,h,i] O[b,k,h,l] Abar[c,k,o] v[i,l,e,o];
 *   A (Da x d x Di), Abar (Db x d x Dj), v (Di x wr x De x Dj) -> y (Da x wl x De x Db).
 * De = 1, O = NULL: transfer.jl:80-83. */
mk_status mk_transfer_right_leg(mk_ctx ctx, const double* v, const double* O, const double* A,
                                const double* Abar, int64_t Da, int64_t d, int64_t Di, int64_t wl,
                                int64_t wr, int64_t De, int64_t Db, int64_t Dj, double* y);

/* An MPO tensor widened by identities on a virtual leg (ne) and a physical
 * leg (nq): out[(w,e),(k,q),(h,r),(v,f)] = O[w,k,h,v] delta(e,f) delta(q,r),
 * out ((wl ne) x (d nq) x (d nq) x (wr ne)), fused indices w + wl e and
 * k + d q; O = NULL is the identity MPO (wl == wr).  With nq > 1 it turns the
 * density-matrix transfers of transfer.jl:150-161 (an MPO acting on
 * the first of two physical legs) into mk_transfer_left/right_mpo over the
 * fused physical leg. */
mk_status mk_mpo_extend(mk_ctx ctx, const double* O, int64_t wl, int64_t d, int64_t wr, int64_t ne,
                        int64_t nq, double* out);

/* A transfer matrix prepared for repeated applies (a TransferMatrix of the
 * reference, src/transfermatrix/transfermatrix.jl:4-33, applied by the
 * fixed-point solvers of src/environments/infinite_envs.jl:207-346 and
 * src/states/ortho.jl:307-313): mk_apply(op, x, y, alpha, beta, precision)
 * computes y = alpha T(x) + beta y in any prepared precision, mk_eigsolve and
 * the batch solvers take it like a derivative operator when it is square.
 * The operator keeps its own copies of O, A and Abar.  Left:
 * T(x) = mk_transfer_left_mpo(x, O, A, Abar) with the same dimensions; right:
 * T(v) = mk_transfer_right_mpo(v, O, A, Abar).  O = NULL gives the MPS
 * transfer (mk_transfer_left_mps / mk_transfer_right_mps, wl = wr = 1).
 * FP32 / BF16x3 run both GEMMs on mixed copies of A and Abar made here, and
 * the MPO step writes the second GEMM's mixed operand directly. */
mk_status mk_prepare_transfer_left(mk_ctx ctx, const double* O, const double* A, const double* Abar,
                                   int64_t Dm, int64_t wl, int64_t Dn, int64_t d, int64_t wr,
                                   int64_t Dc, int64_t Da, int precisions, mk_op* out);
mk_status mk_prepare_transfer_right(mk_ctx ctx, const double* O, const double* A, const double* Abar,
                                    int64_t Da, int64_t d, int64_t Di, int64_t wl, int64_t wr,
                                    int64_t Dc, int64_t Do, int precisions, mk_op* out);

/* regularize!(v, lvec, rvec) (src/transfermatrix/transfermatrix.jl:70-80), in
 * place, for v (D1 x w x D2) (w = 1: a bond tensor):
 * v[a,g,c] -= (sum_ij lvec[i,j] v[j,g,i]) rvec[a,c];  lvec (D2 x D1), rvec (D1 x D2). */
mk_status mk_regularize(mk_ctx ctx, double* v, const double* lvec, const double* rvec,
                        int64_t D1, int64_t w, int64_t D2);

/* ---- Krylov eigensolver ---------------------------------------------------- */

/* Lowest (:SR) eigenpair of a prepared Hermitian derivative operator by
 * Lanczos with full FP64 reorthogonalisation on the device; replaces
 * fixedpoint(H, x0, :SR, Lanczos(...)) (src/algorithms/fixedpoint.jl:15-18,
 * KrylovKit.eigsolve).  x: in = start vector, out = normalised eigenvector.
 * Restarts after `krylovdim` vectors, at most `maxiter` restarts; converged
 * when the Ritz residual norm <= tol.  `precision` selects the apply
 * arithmetic; for MK_FP32/MK_BF16X3 every restart is closed by an FP64 apply
 * and the reported eigenvalue and residual are FP64 (iterative refinement).
 * Blocks the host thread (one scalar readback per Lanczos step). */
mk_status mk_eigsolve(mk_op op, double* x, int krylovdim, int maxiter, double tol,
                      int precision, double* eigval, double* resnorm, int* numops,
                      int* converged);

/* ---- plain GEMM ---------------------------------------------------------------
 * No single counterpart in the reference (BLAS gemm under TensorOperations'
 * tensorcontract!, src/utility/utility.jl:295-301, is what every contraction
 * above reduces to).  Column-major C = alpha op(A) op(B) + beta C through the
 * library's MFMA kernels in the requested arithmetic, for callers that hold
 * fused matrices of their own. */
mk_status mk_gemm(mk_ctx ctx, int transa, int transb, int64_t M, int64_t N, int64_t K,
                  double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                  double beta, double* C, int64_t ldc, int precision);

/* ---- batches -----------------------------------------------------------------
 * Counterpart of `mk_apply_batched(ops[], xs[], ys[], n)` and of running the
 * local eigensolves of many instances at once: the reference applies one
 * derivative operator per call (derivatives.jl:8-11) and solves one local
 * problem per fixedpoint call (fixedpoint.jl:15-18); callers that sweep many
 * instances in lockstep (parameter sweeps, symmetry sectors) bind these.
 *
 * A batch groups n prepared square operators of one context (they must stay
 * alive while the batch exists).  Vectors of a batch are segmented: instance
 * i occupies elements [offsets[i], offsets[i] + x_elems_i) of one device
 * buffer of `total` doubles; offsets are multiples of 4096 and the padding
 * between instances must be zero (it is never written).  When every operator
 * is a nearest-neighbour Jordan two-site operator (`grouped` = 1) an apply of
 * the whole batch is one grouped persistent GEMM launch plus two grouped
 * memory-bound passes; otherwise operators are applied one by one.  Tile
 * shapes of grouped launches are chosen by the library's autotuner. */
typedef struct mk_batch_s* mk_batch;
mk_status mk_batch_create(mk_ctx ctx, int n, const mk_op* ops, mk_batch* out);
mk_status mk_batch_layout(mk_batch b, int64_t* offsets, int64_t* total, int* grouped);
/* y = H x blockwise (padding of y not written). */
mk_status mk_batch_apply(mk_batch b, const double* x, double* y, int precision);
/* Lowest eigenpair of every instance (mk_eigsolve's algorithm in lockstep);
 * x holds the start vectors and receives the normalised eigenvectors.
 * use_graph = 1 captures each Krylov step's device work once into a hipGraph
 * and replays it (grouped batches only).  Outputs are arrays of length n. */
mk_status mk_batch_eigsolve(mk_batch b, double* x, int krylovdim, int maxiter, double tol, int precision,
                            int use_graph, double* eigvals, double* resnorms, int* numops, int* converged);
mk_status mk_batch_free(mk_batch b);
/* Autotuner decisions taken and cache hits so far in this process (no
 * counterpart in the reference). */
mk_status mk_autotune_stats(int* decisions, int* hits);

/* ---- learned Davidson preconditioner ------------------------------------------
 * No counterpart in the reference (its local solver is KrylovKit's Lanczos,
 * fixedpoint.jl:15-18): a Davidson solver for the same lowest-eigenpair
 * problem whose correction vectors come from a small network, for callers that
 * solve many similar local problems (sweeps, parameter scans).
 *
 * The network is a 3-layer FP32 MLP (512 -> 256 -> 256 -> 256, tanh) acting on
 * 256-element blocks of the residual: input = the residual block and the
 * diagonally preconditioned residual block q = r / (theta - diag H), output y =
 * a multiplicative refinement of the diagonal correction, t = q (1 + y); the
 * untrained network (zero last layer) is the diagonal correction.  Training
 * minimises the mean over blocks of |t*| (1 - cos(q (1 + y), t*)) with t* the
 * ideal correction.  Parameters are flattened [W0 b0 W1 b1 W2 b2] with
 * row-major (out x in) weights (mk_mlp_param_count floats).  Training pairs
 * (features, ideal correction of a Davidson iterate) are collected on the
 * device by mk_davidson(collect = 1); mk_mlp_train runs Adam on random
 * minibatches of them with the library's forward/backward kernels. */
typedef struct mk_mlp_s* mk_mlp;
mk_status mk_mlp_create(mk_ctx ctx, uint64_t seed, int64_t capacity, mk_mlp* out);
mk_status mk_mlp_free(mk_mlp m);
mk_status mk_mlp_param_count(int64_t* n);
mk_status mk_mlp_get_params(mk_mlp m, float* host);
mk_status mk_mlp_set_params(mk_mlp m, const float* host);
/* Y (rows x 256) = MLP(X (rows x 512)), device pointers. */
mk_status mk_mlp_forward(mk_mlp m, const float* X, int64_t rows, float* Y);
/* Loss to the targets T (rows x 256) and its parameter gradient (host, optional). */
mk_status mk_mlp_gradient(mk_mlp m, const float* X, const float* T, int64_t rows, float* grads_host, float* loss);
mk_status mk_mlp_dataset_size(mk_mlp m, int64_t* rows);
/* Adam steps on random minibatches of training rows [row_begin, row_end)
 * (row_end < 0: all rows); minibatch losses of the first and last step. */
mk_status mk_mlp_train(mk_mlp m, int steps, int batch, float lr, uint64_t seed, int64_t row_begin, int64_t row_end,
                       float* first_loss, float* last_loss);
/* Loss of the current parameters on training rows [row_begin, row_end). */
mk_status mk_mlp_eval(mk_mlp m, int64_t row_begin, int64_t row_end, float* loss);
/* Lowest eigenpair by Davidson iteration with float64 Rayleigh-Ritz and
 * residuals.  mode 0: residual correction, 1: diagonal correction, 2: MLP
 * correction (mlp required).  Falls back to mk_eigsolve (Lanczos) from the
 * current vector if not converged after maxiter steps.  collect = 1 appends
 * training pairs of the iterates to mlp's training set. */
mk_status mk_davidson(mk_op op, double* x, int maxdim, int maxiter, double tol, mk_mlp mlp, int mode, int collect,
                      double* eigval, double* resnorm, int* numops, int* converged);

/* ---- kernel timing ---------------------------------------------------------
 * No counterpart in the reference (its TimerOutputs sections time whole
 * algorithm steps, src/algorithms/groundstate/dmrg.jl:139-150).  When enabled,
 * every GEMM launch of the library is bracketed by HIP events recorded on its
 * own stream; mk_prof_query returns the summed device time and launch count of
 * one kernel family ("gemm_f64", "gemm_f32", "gemm_bf16x3"). */
mk_status mk_prof_enable(int on);
mk_status mk_prof_reset(void);
mk_status mk_prof_query(const char* kernel, double* total_ms, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif /* MPSKIT_HIP_H */
