// Memory-bound helper kernels (prepare, transfer middle step, regularise,
// Krylov vector algebra, precision conversion).  All arrays column-major.
#pragma once
#include "common.h"

namespace mk {

// L[(a,b),(f,h,kk)] = sum_g GLt(a,g,f) W[g,b,h,klist[kk]]     (Da*d x Df*d*nk)
// GLt = GL with channel `g_ident` replaced by the identity (g_ident < 0: none).
// W is a device copy, already masked on the host where terms are excluded.
mk_status k_prep_L(const double* GL, const double* W, const int* klist, int nk,
                   int64_t Da, int64_t Df, int wl, int d, int wr, int g_ident, double* L,
                   hipStream_t st);

// R[(i,j),(kk,c,e)] = sum_l W[klist[kk],e,j,l] GRt(i,l,c)      (Di*d x nk*Dc*d)
// GRt = GR with channel `l_ident` replaced by the identity (l_ident < 0: none).
mk_status k_prep_R(const double* GR, const double* W, const int* klist, int nk,
                   int64_t Di, int64_t Dc, int wl, int d, int wr, int l_ident, double* R,
                   hipStream_t st);

// C'[(h,i),(b,c)] = sum_l W[0,b,h,l] GRt(i,l,c)               (d*Di x d*Dc)
mk_status k_prep_Cp(const double* GR, const double* W, int64_t Di, int64_t Dc, int wl, int d,
                    int wr, int l_ident, double* Cp, hipStream_t st);

// CA'[(h,i,j),(b,c,e)] = sum_k W1[0,b,h,k] sum_l W2[k,e,j,l] GRt(i,l,c)
// (W1 already masked to the k that belong to this term)      (d1*Dr*d2 x d1*Dr*d2)
mk_status k_prep_CA(const double* GR, const double* W1, const double* W2, int64_t Dr,
                    int wl, int d1, int wm, int d2, int wr, int l_ident, double* CA,
                    hipStream_t st);

// AB'[(a,b),(f,h),e,j] = sum_{g,k} GLt(a,g,f) W1[g,b,h,k] W2[k,e,j,wr-1]
// (W1 masked on the host)                     blocks (Dl*d1 x Dl*d1), block (e,j) at e + d2*j
mk_status k_prep_AB(const double* GL, const double* W1, const double* W2, int64_t Dl,
                    int wl, int d1, int wm, int d2, int wr, int g_ident, double* AB,
                    hipStream_t st);

// out[a + D0*(f + D2*g)] = in[a + D0*(g + D1*f)]   (swap the last two legs of a 3-leg array)
mk_status k_swap23(const double* in, int64_t D0, int64_t D1, int64_t D2, double* out,
                   hipStream_t st);

// dst[i + Di*(k + nk*c)] = src[i + Di*(klist[k] + w*c)]     (channel subset of an env)
mk_status k_env_channels(const double* src, int64_t Di, int w, int64_t Dc, const int* klist,
                         int nk, double* dst, hipStream_t st);

// transfer-left middle step: t2[m,k,b,c] = sum_{g,h} t1[m,g,h,c] O[g,k,h,b]
mk_status k_tl_mid(const double* t1, const double* O, int64_t Dm, int wl, int d, int wr,
                   int64_t Dc, double* t2, hipStream_t st);
// transfer-right middle step: t2[a,b,k,o] = sum_{h,l} t1[a,h,l,o] O[b,k,h,l]
mk_status k_tr_mid(const double* t1, const double* O, int64_t Da, int wl, int d, int wr,
                   int64_t Do, double* t2, hipStream_t st);
// out[a,q,p,b] = in[a,p,q,b] for in of shape n0 x n1 x n2 x n3
mk_status k_swap_mid(const double* in, int64_t n0, int64_t n1, int64_t n2, int64_t n3, double* out,
                     hipStream_t st);
// out[(w,e),(k,q),(h,r),(v,f)] = O[w,k,h,v] delta(e,f) delta(q,r), shape
// (wl ne) x (d nq) x (d nq) x (wr ne); O == nullptr is the identity MPO (wl == wr)
mk_status k_mpo_extend(const double* O, int wl, int d, int wr, int ne, int nq, double* out, hipStream_t st);
// The MPO steps above written as the K-contiguous mixed operand (bf16 hi/lo if
// split, else fp32) of the transfer's second GEMM, row pitch ldo (zero padded).
mk_status k_tl_mid_mixed(const double* t1, const double* O, int64_t Dm, int wl, int d, int wr, int64_t Dc,
                         int64_t ldo, int split, void* hi, void* lo, hipStream_t st);
mk_status k_tr_mid_mixed(const double* t1, const double* O, int64_t Da, int wl, int d, int wr, int64_t Do,
                         int64_t ldo, int split, void* hi, void* lo, hipStream_t st);

// physical two-site term without a GEMM: y[a,b,c,e] = alpha*sum P[b,e,h,j] x[a,h,c,j] + beta*y
mk_status k_phys2(const double* P, const double* x, int64_t Dl, int d1, int64_t Dr, int d2,
                  double alpha, double beta, double* y, hipStream_t st);
// physical one-site term: y[a,b,c] = alpha*sum_h P[b,h] x[a,h,c] + beta*y
mk_status k_diag_nn(const double* IC, const double* BE, const double* P, int64_t Dl, int64_t Dr, int d1, int d2,
                    double* out, hipStream_t st);
mk_status k_phys1(const double* P, const double* x, int64_t Dl, int d, int64_t Dr,
                  double alpha, double beta, double* y, hipStream_t st);

// regularise in place: lam_g = sum_ij lvec[i,j] v[j,g,i]; v[a,g,c] -= lam_g rvec[a,c]
mk_status k_regularize(double* v, const double* lvec, const double* rvec, int64_t D1, int w,
                       int64_t D2, double* lam_scratch, hipStream_t st);

// ---- Krylov vector algebra (float64, device-side scalars) ----------------
// out[j] = <V_j, w> for j < k  (V column-major n x k, leading dim ldv)
mk_status k_multidot(const double* V, int64_t ldv, int k, const double* w, int64_t n,
                     double* out, double* partial, hipStream_t st);
// Fused classical Gram-Schmidt passes over the k <= 32 columns of V (ldv):
// u <- u - V cin, then out[0..k) = V^T u (update_dot) or out[0] = u^T u,
// out[1] = w^T u (update_norm), all in one read of V.  partial holds
// cgs_blocks(n) * k doubles.
int cgs_blocks(int64_t n);
mk_status k_cgs_update_dot(const double* V, int64_t ldv, int k, const double* cin, double* u, int64_t n,
                           double* out, double* partial, hipStream_t st);
mk_status k_cgs_update_norm(const double* V, int64_t ldv, int k, const double* cin, double* u, const double* w,
                            int64_t n, double* out, double* partial, hipStream_t st);
// w -= V c (c on device, length k)
mk_status k_multiaxpy(const double* V, int64_t ldv, int k, const double* c, double* w,
                      int64_t n, hipStream_t st);
// y = a*x + b*y with a, b host scalars
mk_status k_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t st);
// y *= s where s = *sdev (device scalar) or 1/s if inv
mk_status k_scale_dev(double* y, const double* sdev, int inv, int64_t n, hipStream_t st);
// x = V s  (s on device, length k)
mk_status k_gemv(const double* V, int64_t ldv, int k, const double* s, double* x, int64_t n,
                 hipStream_t st);
// count non-finite entries into *flag (device int)
mk_status k_nonfinite(const double* x, int64_t n, int* flag, hipStream_t st);

// ---- precision conversion --------------------------------------------------
// Convert a column-major float64 matrix M (rows x cols, leading dim ld) into a
// K-contiguous operand whose rows have length ldo (>= K, padding zeroed):
//   trans == 1: K = rows, out[c*ldo + r] = M(r,c)   (no data transposition)
//   trans == 0: K = cols, out[r*ldo + c] = M(r,c)   (LDS-tiled transposition)
// bf16: hi = bf16(M), lo = bf16(M - hi).
mk_status k_split_bf16(const double* M, int64_t rows, int64_t cols, int64_t ld, int trans,
                       int64_t ldo, void* hi, void* lo, hipStream_t st);
mk_status k_to_f32(const double* M, int64_t rows, int64_t cols, int64_t ld, int trans,
                   int64_t ldo, float* out, hipStream_t st);

// One pass over x (d1 = d2 = 2, M = 2 Dl and N = 2 Dr multiples of 8) for the
// mixed-precision nearest-neighbour apply: y = alpha CB(x) + beta y and the
// K-contiguous conversions of x as left factor (xa, rows of length ka >= N) and
// as right factor (xb, rows of length kb >= M); split = 1: bf16 hi/lo, 0: fp32
// (xa_hi / xb_hi only).  Row padding is not written.  packed = 1: both views
// in the packed panel layout of the 256 x 256 mixed kernel (ka = N, kb = M;
// 2 Dl and 2 Dr multiples of 256).
mk_status k_nn_prologue(const double* P, const double* x, int64_t Dl, int64_t Dr, double alpha, double beta,
                        double* y, int split, void* xa_hi, void* xa_lo, int64_t ka, void* xb_hi, void* xb_lo,
                        int64_t kb, int packed, hipStream_t st);
// Packed panel layout (packed_offset) of an operand held as MN K-contiguous
// rows of length kpad >= K, element size eb (2 or 4 bytes).
mk_status k_repack(const void* in, int64_t MN, int64_t K, int64_t kpad, int eb, void* out, hipStream_t st);


// ---- Jordan operator matrices from the constructed terms ------------------
// The terms are arrays in the layouts of the reference's JordanMPO_AC2 /
// JordanMPO_AC_Hamiltonian fields (null = missing); each routine writes one
// GEMM operand with the terms prepare_operator!! folds into it.
// ICm[(i,j),(c,e)] = IC[e,c,j,i] + ID[e,j] d_ci + d_ej II[c,i]          (Dr*d2 square)
mk_status k_asm_ic(const double* IC, const double* ID, const double* II, int64_t Dr, int d2, double* out,
                   hipStream_t st);
// BEm[(a,b),(f,h)] = BE[a,b,f,h] + d_af DE[b,h] + EE[a,f] d_bh           (Dl*d1 square)
mk_status k_asm_be(const double* BE, const double* DE, const double* EE, int64_t Dl, int d1, double* out,
                   hipStream_t st);
// CAm[(h,i,j),(b,c,e)] = CA[b,e,c,h,j,i] + CB[b,e,h,j] d_ci
//                        + d_bh (IC[e,c,j,i] + ID[e,j] d_ci + d_ej II[c,i])   (d1*Dr*d2 square)
mk_status k_asm_ca(const double* CA, const double* CB, const double* IC, const double* ID, const double* II,
                   int64_t Dr, int d1, int d2, double* out, hipStream_t st);
// ABm_(e,j)[(a,b),(f,h)] = AB[a,b,e,f,h,j] + CB[b,e,h,j] d_af
//                        + d_ej (BE[a,b,f,h] + d_af DE[b,h] + EE[a,f] d_bh)   blocks (e,j) at e + d2*j
mk_status k_asm_ab(const double* AB, const double* CB, const double* BE, const double* DE, const double* EE,
                   int64_t Dl, int d1, int d2, double* out, hipStream_t st);
// C'[(h,i),(b,c)] = C[b,c,h,i] + D[b,h] d_ic + d_bh I[i,c]                 (d*Dr square)
mk_status k_asm_cp(const double* C, const double* D, const double* I, int64_t Dr, int d, double* out,
                   hipStream_t st);

}  // namespace mk
