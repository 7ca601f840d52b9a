// C ABI: contexts, prepared derivative operators, transfers, regularisation.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "gemm.h"
#include "kernels.h"
#include "ops.h"

namespace mk {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------

mk_status DevBuf::alloc(size_t nbytes) {
  release();
  if (nbytes == 0) return MK_OK;
  hipError_t e = hipMalloc(&ptr, nbytes);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error("hipMalloc(%zu) failed: %s", nbytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? MK_OOM : MK_HIP_ERROR;
  }
  bytes = nbytes;
  return MK_OK;
}
void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

mk_status Ctx::scratch(size_t nbytes, void** out) {
  if (nbytes > arena.bytes) {
    // grow: outstanding work on the stream may still use the old arena
    MK_CHECK_HIP(hipStreamSynchronize(stream));
    MK_TRY(arena.alloc(nbytes + nbytes / 4));
  }
  *out = arena.ptr;
  return MK_OK;
}

mk_status Ctx::small(size_t nbytes, void** out) {
  MK_REQUIRE(fixed_scratch().base == nullptr, MK_NOT_SUPPORTED,
             "this call needs the context's own scratch and cannot run without allocating");
  if (nbytes > smallbuf.bytes) {
    MK_CHECK_HIP(hipStreamSynchronize(stream));
    MK_TRY(smallbuf.alloc(std::max<size_t>(nbytes, 1 << 20)));
  }
  *out = smallbuf.ptr;
  return MK_OK;
}


FixedScratch& fixed_scratch() {
  static thread_local FixedScratch f;
  return f;
}

// Carve several regions out of the context arena (or out of the calling
// thread's fixed scratch region, if one is installed).
mk_status carve(Ctx* ctx, std::initializer_list<std::pair<size_t, void**>> regions) {
  size_t total = 0;
  for (auto& r : regions) total += align256(r.first);
  void* base = nullptr;
  const FixedScratch& fx = fixed_scratch();
  if (fx.base) {
    MK_REQUIRE(total <= fx.bytes, MK_BAD_SHAPE,
               "this apply needs %zu bytes of scratch; the fixed scratch holds %zu", total, fx.bytes);
    base = fx.base;
  } else {
    MK_TRY(ctx->scratch(total, &base));
  }
  size_t off = 0;
  for (auto& r : regions) {
    *r.second = static_cast<char*>(base) + off;
    off += align256(r.first);
  }
  return MK_OK;
}

// ---------------------------------------------------------------------------
// Matrices with optional mixed-precision copies
// ---------------------------------------------------------------------------

// Row length (elements) of a K-contiguous mixed-precision operand: K rounded up
// to 8 (16-byte rows), plus 64 when that is a multiple of 128 and at least 512,
// so that long rows are not a multiple of a large power of two apart; the rows
// of one K slab then spread over the L2 and HBM channels instead of camping
// on one.  The padding is zero.
int64_t pad8(int64_t k) {
  int64_t p = (k + 7) / 8 * 8;
  if (p >= 512 && p % 128 == 0) p += 64;
  return p;
}

// A float64 column-major operator matrix (rows x cols) used as GEMM operand
// `role` ('A' = left factor, rows = M, cols = K;  'B' = right factor,
// rows = K, cols = N).  Mixed copies are K-contiguous with padded rows.
mk_status OpMat::make_mixed(int precisions, hipStream_t st) {
  const bool kfast = (role == 'B');   // col-major B: K is the fast index
  const int64_t K = kfast ? rows : cols;
  const int64_t MN = kfast ? cols : rows;
  kpad = pad8(K);
  if (precisions & MK_PREC_BF16X3) {
    MK_TRY(hi.alloc((size_t)MN * kpad * 2));
    MK_TRY(lo.alloc((size_t)MN * kpad * 2));
    MK_TRY(k_split_bf16(f64(), rows, cols, rows, kfast ? 1 : 0, kpad, hi.ptr, lo.ptr, st));
  }
  if (precisions & MK_PREC_FP32) {
    MK_TRY(f32.alloc((size_t)MN * kpad * 4));
    MK_TRY(k_to_f32(f64(), rows, cols, rows, kfast ? 1 : 0, kpad, static_cast<float*>(f32.ptr), st));
  }
  return MK_OK;
}

mk_status OpMat::make_packed(int precisions, hipStream_t st) {
  const bool kfast = (role == 'B');
  const int64_t K = kfast ? rows : cols;
  const int64_t MN = kfast ? cols : rows;
  if ((precisions & MK_PREC_BF16X3) && hi.ptr) {
    MK_TRY(phi.alloc((size_t)MN * K * 2));
    MK_TRY(plo.alloc((size_t)MN * K * 2));
    MK_TRY(k_repack(hi.ptr, MN, K, kpad, 2, phi.ptr, st));
    MK_TRY(k_repack(lo.ptr, MN, K, kpad, 2, plo.ptr, st));
  }
  if ((precisions & MK_PREC_FP32) && f32.ptr) {
    MK_TRY(pf32.alloc((size_t)MN * K * 4));
    MK_TRY(k_repack(f32.ptr, MN, K, kpad, 4, pf32.ptr, st));
  }
  packed = true;
  return MK_OK;
}

// ---------------------------------------------------------------------------
// GEMM call helpers over OpMat / raw x views
// ---------------------------------------------------------------------------

// A transient float64 view converted for a mixed apply (x, xR...).
MixedView view_layout(int64_t rows, int64_t cols, char role, int prec, void* mem) {
  const bool kfast = (role == 'B');
  const int64_t K = kfast ? rows : cols;
  const int64_t MN = kfast ? cols : rows;
  MixedView v;
  v.kpad = pad8(K);
  const size_t n = (size_t)MN * v.kpad;
  char* p = static_cast<char*>(mem);
  v.hi = p;
  v.lo = prec == MK_BF16X3 ? p + align256(n * 2) : nullptr;
  return v;
}

mk_status convert_view(Ctx* ctx, const double* M, int64_t rows, int64_t cols, int64_t ld, char role,
                       int prec, MixedView* out, void* mem) {
  const bool kfast = (role == 'B');
  *out = view_layout(rows, cols, role, prec, mem);
  if (prec == MK_BF16X3)
    MK_TRY(k_split_bf16(M, rows, cols, ld, kfast ? 1 : 0, out->kpad, out->hi, out->lo, ctx->stream));
  else
    MK_TRY(k_to_f32(M, rows, cols, ld, kfast ? 1 : 0, out->kpad, static_cast<float*>(out->hi), ctx->stream));
  return MK_OK;
}

size_t view_bytes(int64_t rows, int64_t cols, char role, int prec) {
  const bool kfast = (role == 'B');
  const int64_t K = kfast ? rows : cols;
  const int64_t MN = kfast ? cols : rows;
  const size_t n = (size_t)MN * pad8(K);
  return prec == MK_BF16X3 ? 2 * align256(n * 2) : align256(n * 4);
}

mk_status run_gemm(const ProbSpec* p, int np, int prec, GemmShape sh, hipStream_t st) {
  if (prec == MK_FP64) return gemm_f64(p, np, sh, st);
  if (prec == MK_BF16X3) return gemm_bf16x3(p, np, st);
  return gemm_f32(p, np, st);
}

}  // namespace mk

using namespace mk;

// ---------------------------------------------------------------------------
// C ABI: library / context
// ---------------------------------------------------------------------------

extern "C" int mk_version(void) { return 100; }
extern "C" const char* mk_last_error(void) { return last_error(); }

extern "C" mk_status mk_ctx_create(int device, mk_ctx* out) {
  MK_REQUIRE(out != nullptr, MK_BAD_ARG, "mk_ctx_create: null out");
  int n = 0;
  MK_CHECK_HIP(hipGetDeviceCount(&n));
  MK_REQUIRE(device >= 0 && device < n, MK_BAD_ARG, "mk_ctx_create: device %d of %d", device, n);
  MK_CHECK_HIP(hipSetDevice(device));
  // the allocation-free apply's scratch array: its address on this device
  MK_TRY(noalloc_setup(device));
  auto* c = new Ctx();
  c->device = device;
  c->stream = nullptr;
  *out = reinterpret_cast<mk_ctx>(c);
  return MK_OK;
}

extern "C" mk_status mk_ctx_destroy(mk_ctx h) {
  if (!h) return MK_OK;
  auto* c = reinterpret_cast<Ctx*>(h);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->release();
  return MK_OK;
}

extern "C" mk_status mk_ctx_set_stream(mk_ctx h, void* stream) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  auto* c = reinterpret_cast<Ctx*>(h);
  c->stream = static_cast<hipStream_t>(stream);
  return MK_OK;
}

extern "C" mk_status mk_ctx_synchronize(mk_ctx h) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  auto* c = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipStreamSynchronize(c->stream));
  return MK_OK;
}

// ---------------------------------------------------------------------------
// C ABI: transfers
// ---------------------------------------------------------------------------

#define CTX_OR_FAIL(h)                                                    \
  MK_REQUIRE((h) != nullptr, MK_BAD_ARG, "null context");                \
  Ctx* ctx = reinterpret_cast<Ctx*>(h);                                  \
  MK_CHECK_HIP(hipSetDevice(ctx->device))

static bool pos(int64_t v) { return v > 0 && v < (1ll << 31); }

extern "C" mk_status mk_transfer_left_mps(mk_ctx h, const double* v, const double* A,
                                          const double* Ab, int64_t Dm, int64_t Dn, int64_t d,
                                          int64_t Dc, int64_t Da, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Dm) && pos(Dn) && pos(d) && pos(Dc) && pos(Da), MK_BAD_SHAPE,
             "transfer_left: non-positive dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_left: null pointer");
  double* t = nullptr;
  MK_TRY(carve(ctx, {{(size_t)Dm * d * Dc * 8, (void**)&t}}));
  // t[m,(h,c)] = v[m,n] A[n,(h,c)]
  ProbSpec p1;
  p1.M = Dm; p1.N = d * Dc; p1.C = t; p1.ldc = Dm;
  p1.add(v, Dm, A, Dn, Dn);
  MK_TRY(gemm_f64(&p1, 1, GemmShape{false, false}, ctx->stream));
  // y[a,c] = Ab[(m,h),a]^T t[(m,h),c]
  ProbSpec p2;
  p2.M = Da; p2.N = Dc; p2.C = y; p2.ldc = Da;
  p2.add(Ab, Dm * d, t, Dm * d, Dm * d);
  return gemm_f64(&p2, 1, GemmShape{true, false}, ctx->stream);
}

extern "C" mk_status mk_transfer_right_mps(mk_ctx h, const double* v, const double* A,
                                           const double* Ab, int64_t Da, int64_t d, int64_t Di,
                                           int64_t Db, int64_t Dj, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Da) && pos(d) && pos(Di) && pos(Db) && pos(Dj), MK_BAD_SHAPE,
             "transfer_right: non-positive dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_right: null pointer");
  double* t = nullptr;
  MK_TRY(carve(ctx, {{(size_t)Da * d * Dj * 8, (void**)&t}}));
  // t[(a,h),j] = A[(a,h),i] v[i,j]
  ProbSpec p1;
  p1.M = Da * d; p1.N = Dj; p1.C = t; p1.ldc = Da * d;
  p1.add(A, Da * d, v, Di, Di);
  MK_TRY(gemm_f64(&p1, 1, GemmShape{false, false}, ctx->stream));
  // y[a,b] = t[a,(h,j)] Ab[b,(h,j)]^T
  ProbSpec p2;
  p2.M = Da; p2.N = Db; p2.C = y; p2.ldc = Da;
  p2.add(t, Da, Ab, Db, d * Dj);
  return gemm_f64(&p2, 1, GemmShape{false, true}, ctx->stream);
}

extern "C" mk_status mk_transfer_left_mpo(mk_ctx h, const double* x, const double* O,
                                          const double* A, const double* Ab, int64_t Dm, int64_t wl,
                                          int64_t Dn, int64_t d, int64_t wr, int64_t Dc, int64_t Da,
                                          double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Dm) && pos(wl) && pos(Dn) && pos(d) && pos(wr) && pos(Dc) && pos(Da), MK_BAD_SHAPE,
             "transfer_left_mpo: non-positive dimension");
  MK_REQUIRE(x && O && A && Ab && y, MK_BAD_ARG, "transfer_left_mpo: null pointer");
  double *t1 = nullptr, *t2 = nullptr;
  MK_TRY(carve(ctx, {{(size_t)Dm * wl * d * Dc * 8, (void**)&t1}, {(size_t)Dm * d * wr * Dc * 8, (void**)&t2}}));
  // t1[(m,g),(h,c)] = x[(m,g),n] A[n,(h,c)]
  ProbSpec p1;
  p1.M = Dm * wl; p1.N = d * Dc; p1.C = t1; p1.ldc = Dm * wl;
  p1.add(x, Dm * wl, A, Dn, Dn);
  MK_TRY(gemm_f64(&p1, 1, GemmShape{false, false}, ctx->stream));
  MK_TRY(k_tl_mid(t1, O, Dm, (int)wl, (int)d, (int)wr, Dc, t2, ctx->stream));
  // y[a,(b,c)] = Ab[(m,k),a]^T t2[(m,k),(b,c)]
  ProbSpec p2;
  p2.M = Da; p2.N = wr * Dc; p2.C = y; p2.ldc = Da;
  p2.add(Ab, Dm * d, t2, Dm * d, Dm * d);
  return gemm_f64(&p2, 1, GemmShape{true, false}, ctx->stream);
}

extern "C" mk_status mk_transfer_right_mpo(mk_ctx h, const double* v, const double* O,
                                           const double* A, const double* Ab, int64_t Da, int64_t d,
                                           int64_t Di, int64_t wl, int64_t wr, int64_t Dc, int64_t Do,
                                           double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Da) && pos(d) && pos(Di) && pos(wl) && pos(wr) && pos(Dc) && pos(Do), MK_BAD_SHAPE,
             "transfer_right_mpo: non-positive dimension");
  MK_REQUIRE(v && O && A && Ab && y, MK_BAD_ARG, "transfer_right_mpo: null pointer");
  double *t1 = nullptr, *t2 = nullptr;
  MK_TRY(carve(ctx, {{(size_t)Da * d * wr * Do * 8, (void**)&t1}, {(size_t)Da * wl * d * Do * 8, (void**)&t2}}));
  // t1[(a,h),(l,o)] = A[(a,h),i] v[i,(l,o)]
  ProbSpec p1;
  p1.M = Da * d; p1.N = wr * Do; p1.C = t1; p1.ldc = Da * d;
  p1.add(A, Da * d, v, Di, Di);
  MK_TRY(gemm_f64(&p1, 1, GemmShape{false, false}, ctx->stream));
  MK_TRY(k_tr_mid(t1, O, Da, (int)wl, (int)d, (int)wr, Do, t2, ctx->stream));
  // y[(a,b),c] = t2[(a,b),(k,o)] Ab[c,(k,o)]^T
  ProbSpec p2;
  p2.M = Da * wl; p2.N = Dc; p2.C = y; p2.ldc = Da * wl;
  p2.add(t2, Da * wl, Ab, Dc, d * Do);
  return gemm_f64(&p2, 1, GemmShape{false, true}, ctx->stream);
}

extern "C" mk_status mk_regularize(mk_ctx h, double* v, const double* lvec, const double* rvec,
                                   int64_t D1, int64_t w, int64_t D2) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(D1) && pos(w) && pos(D2), MK_BAD_SHAPE, "regularize: non-positive dimension");
  MK_REQUIRE(v && lvec && rvec, MK_BAD_ARG, "regularize: null pointer");
  void* lam = nullptr;
  MK_TRY(ctx->small((size_t)w * 8, &lam));
  return k_regularize(v, lvec, rvec, D1, (int)w, D2, static_cast<double*>(lam), ctx->stream);
}

// ---------------------------------------------------------------------------
// C ABI: plain GEMM through the library's kernels
// ---------------------------------------------------------------------------

extern "C" mk_status mk_gemm(mk_ctx h, int transa, int transb, int64_t M, int64_t N, int64_t K, double alpha,
                             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                             int64_t ldc, int precision) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(M >= 0 && N >= 0 && K >= 0 && M < (1ll << 31) && N < (1ll << 31) && K < (1ll << 31), MK_BAD_SHAPE,
             "gemm: bad dimensions");
  MK_REQUIRE(A && B && C, MK_BAD_ARG, "gemm: null pointer");
  MK_REQUIRE(lda >= (transa ? K : M) && ldb >= (transb ? N : K) && ldc >= M, MK_BAD_SHAPE, "gemm: leading dims");
  MK_REQUIRE(precision >= 0 && precision <= 2, MK_BAD_ARG, "gemm: precision %d", precision);
  if (M == 0 || N == 0) return MK_OK;
  ProbSpec p;
  p.M = M; p.N = N; p.C = C; p.ldc = ldc; p.alpha = alpha; p.beta = beta;
  if (precision == MK_FP64) {
    p.add(A, lda, B, ldb, K);
    return gemm_f64(&p, 1, GemmShape{transa != 0, transb != 0}, ctx->stream);
  }
  // K-contiguous copies: stored A is M x K (convert with transposition) or,
  // transposed, K x M (plain); stored B is K x N (plain) or N x K (transposed)
  void *ma = nullptr, *mb = nullptr;
  MK_TRY(carve(ctx, {{view_bytes(transa ? K : M, transa ? M : K, transa ? 'B' : 'A', precision), &ma},
                     {view_bytes(transb ? N : K, transb ? K : N, transb ? 'A' : 'B', precision), &mb}}));
  MixedView va, vb;
  MK_TRY(convert_view(ctx, A, transa ? K : M, transa ? M : K, lda, transa ? 'B' : 'A', precision, &va, ma));
  MK_TRY(convert_view(ctx, B, transb ? N : K, transb ? K : N, ldb, transb ? 'A' : 'B', precision, &vb, mb));
  p.add(va.hi, va.kpad, vb.hi, vb.kpad, K, va.lo, vb.lo);
  return run_gemm(&p, 1, precision, GemmShape{}, ctx->stream);
}
