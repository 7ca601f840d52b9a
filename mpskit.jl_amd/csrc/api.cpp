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

// The four FP64 transfers on caller-given scratch (the entry points below
// carve it from the context arena in one piece, so a composed transfer never
// hands an inner step a region that overlaps its own).

// y[a,c] = v[m,n] A[n,h,c] Ab[m,h,a]; t holds Dm d Dc doubles
static mk_status tl_mps_run(Ctx* ctx, const double* v, const double* A, const double* Ab, int64_t Dm, int64_t Dn,
                            int64_t d, int64_t Dc, int64_t Da, double* y, double* t) {
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

// y[a,b] = A[a,h,i] Ab[b,h,j] v[i,j]; t holds Da d Dj doubles
static mk_status tr_mps_run(Ctx* ctx, const double* v, const double* A, const double* Ab, int64_t Da, int64_t d,
                            int64_t Di, int64_t Db, int64_t Dj, double* y, double* t) {
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

// y[a,b,c] = x[m,g,n] A[n,h,c] O[g,k,h,b] Ab[m,k,a]; t1 holds Dm wl d Dc and
// t2 Dm d wr Dc doubles
static mk_status tl_mpo_run(Ctx* ctx, const double* x, const double* O, const double* A, const double* Ab, int64_t Dm,
                            int64_t wl, int64_t Dn, int64_t d, int64_t wr, int64_t Dc, int64_t Da, double* y,
                            double* t1, double* t2) {
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

// y[a,b,c] = A[a,h,i] O[b,k,h,l] Ab[c,k,o] v[i,l,o]; t1 holds Da d wr Do and
// t2 Da wl d Do doubles
static mk_status tr_mpo_run(Ctx* ctx, const double* v, const double* O, const double* A, const double* Ab, int64_t Da,
                            int64_t d, int64_t Di, int64_t wl, int64_t wr, int64_t Dc, int64_t Do, double* y,
                            double* t1, double* t2) {
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

static size_t f64b(int64_t n) { return (size_t)n * 8; }

extern "C" mk_status mk_transfer_left_mps(mk_ctx h, const double* v, const double* A,
                                          const double* Ab, int64_t Dm, int64_t Dn, int64_t d,
                                          int64_t Dc, int64_t Da, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Dm) && pos(Dn) && pos(d) && pos(Dc) && pos(Da), MK_BAD_SHAPE,
             "transfer_left: non-positive dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_left: null pointer");
  double* t = nullptr;
  MK_TRY(carve(ctx, {{f64b(Dm * d * Dc), (void**)&t}}));
  return tl_mps_run(ctx, v, A, Ab, Dm, Dn, d, Dc, Da, y, t);
}

extern "C" mk_status mk_transfer_right_mps(mk_ctx h, const double* v, const double* A,
                                           const double* Ab, int64_t Da, int64_t d, int64_t Di,
                                           int64_t Db, int64_t Dj, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Da) && pos(d) && pos(Di) && pos(Db) && pos(Dj), MK_BAD_SHAPE,
             "transfer_right: non-positive dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_right: null pointer");
  double* t = nullptr;
  MK_TRY(carve(ctx, {{f64b(Da * d * Dj), (void**)&t}}));
  return tr_mps_run(ctx, v, A, Ab, Da, d, Di, Db, Dj, y, t);
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
  MK_TRY(carve(ctx, {{f64b(Dm * wl * d * Dc), (void**)&t1}, {f64b(Dm * d * wr * Dc), (void**)&t2}}));
  return tl_mpo_run(ctx, x, O, A, Ab, Dm, wl, Dn, d, wr, Dc, Da, y, t1, t2);
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
  MK_TRY(carve(ctx, {{f64b(Da * d * wr * Do), (void**)&t1}, {f64b(Da * wl * d * Do), (void**)&t2}}));
  return tr_mpo_run(ctx, v, O, A, Ab, Da, d, Di, wl, wr, Dc, Do, y, t1, t2);
}

// ---- quasiparticle and passenger-leg transfers ----------------------------
// Every variant is one of the transfers above on reshaped operands: an
// excitation leg e of A fuses with its right bond (left transfers) or, after
// the two middle legs of A are exchanged, with its left bond (right
// transfers); a leg that passes through x / v widens the MPO by an identity
// on that leg (mpo_extend), and a missing MPO is the identity on the
// passenger legs.

static bool small_leg(int64_t a, int64_t b) { return a > 0 && b > 0 && a * b < (1 << 20); }

extern "C" mk_status mk_mpo_extend(mk_ctx h, const double* O, int64_t wl, int64_t d, int64_t wr, int64_t ne,
                                   int64_t nq, double* out) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(small_leg(wl, ne) && small_leg(wr, ne) && small_leg(d, nq), MK_BAD_SHAPE,
             "mpo_extend: bad dimension (wl %lld, d %lld, wr %lld, ne %lld, nq %lld)", (long long)wl, (long long)d,
             (long long)wr, (long long)ne, (long long)nq);
  MK_REQUIRE(O != nullptr || wl == wr, MK_BAD_SHAPE, "mpo_extend: the identity MPO needs wl == wr");
  MK_REQUIRE(out != nullptr, MK_BAD_ARG, "mpo_extend: null output");
  return k_mpo_extend(O, (int)wl, (int)d, (int)wr, (int)ne, (int)nq, out, ctx->stream);
}

extern "C" mk_status mk_transfer_left_excited(mk_ctx h, const double* x, const double* O, const double* A,
                                              const double* Ab, int64_t Dm, int64_t wl, int64_t Dn, int64_t d,
                                              int64_t De, int64_t wr, int64_t Dc, int64_t Da, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Dm) && small_leg(wl, 1) && pos(Dn) && small_leg(d, d) && small_leg(wr, 1) && pos(Dc) && pos(Da) &&
                 pos(De) && pos(De * Dc),
             MK_BAD_SHAPE, "transfer_left_excited: bad dimension");
  MK_REQUIRE(x && A && Ab && y, MK_BAD_ARG, "transfer_left_excited: null pointer");
  MK_REQUIRE(O != nullptr || wl == wr, MK_BAD_SHAPE, "transfer_left_excited: without an MPO wl must equal wr");
  const int64_t Dc2 = De * Dc;   // A[n,h,(e,c)]
  if (O == nullptr && wl == 1) {
    double* t = nullptr;
    MK_TRY(carve(ctx, {{f64b(Dm * d * Dc2), (void**)&t}}));
    return tl_mps_run(ctx, x, A, Ab, Dm, Dn, d, Dc2, Da, y, t);
  }
  double *t1 = nullptr, *t2 = nullptr, *Oi = nullptr;
  MK_TRY(carve(ctx, {{f64b(Dm * wl * d * Dc2), (void**)&t1},
                     {f64b(Dm * d * wr * Dc2), (void**)&t2},
                     {f64b(O ? 1 : wl * d * d * wl), (void**)&Oi}}));
  if (O == nullptr) {
    MK_TRY(k_mpo_extend(nullptr, (int)wl, (int)d, (int)wl, 1, 1, Oi, ctx->stream));
    O = Oi;
  }
  return tl_mpo_run(ctx, x, O, A, Ab, Dm, wl, Dn, d, wr, Dc2, Da, y, t1, t2);
}

extern "C" mk_status mk_transfer_right_excited(mk_ctx h, const double* v, const double* O, const double* A,
                                               const double* Ab, int64_t Da, int64_t d, int64_t De, int64_t Di,
                                               int64_t wl, int64_t wr, int64_t Db, int64_t Dj, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Da) && small_leg(d, d) && pos(De) && pos(Da * De) && pos(Di) && small_leg(wl, 1) &&
                 small_leg(wr, 1) && pos(Db) && pos(Dj),
             MK_BAD_SHAPE, "transfer_right_excited: bad dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_right_excited: null pointer");
  MK_REQUIRE(O != nullptr || wl == wr, MK_BAD_SHAPE, "transfer_right_excited: without an MPO wl must equal wr");
  const int64_t Da2 = Da * De;          // A'[(a,e),h,i]
  const bool permA = De > 1;
  const bool permY = De > 1 && wl > 1;  // the transfer yields y'[a,e,w,b]
  const bool mps = O == nullptr && wl == 1;
  double *Ap = nullptr, *yp = nullptr, *t1 = nullptr, *t2 = nullptr, *Oi = nullptr;
  MK_TRY(carve(ctx, {{f64b(permA ? Da2 * d * Di : 1), (void**)&Ap},
                     {f64b(permY ? Da2 * wl * Db : 1), (void**)&yp},
                     {f64b(mps ? Da2 * d * Dj : Da2 * d * wr * Dj), (void**)&t1},
                     {f64b(mps ? 1 : Da2 * wl * d * Dj), (void**)&t2},
                     {f64b(O || mps ? 1 : wl * d * d * wl), (void**)&Oi}}));
  const double* Au = A;
  if (permA) {
    MK_TRY(k_swap_mid(A, Da, d, De, Di, Ap, ctx->stream));
    Au = Ap;
  }
  double* yo = permY ? yp : y;
  if (mps) {
    MK_TRY(tr_mps_run(ctx, v, Au, Ab, Da2, d, Di, Db, Dj, yo, t1));
  } else {
    if (O == nullptr) {
      MK_TRY(k_mpo_extend(nullptr, (int)wl, (int)d, (int)wl, 1, 1, Oi, ctx->stream));
      O = Oi;
    }
    MK_TRY(tr_mpo_run(ctx, v, O, Au, Ab, Da2, d, Di, wl, wr, Db, Dj, yo, t1, t2));
  }
  if (permY) MK_TRY(k_swap_mid(yp, Da, De, wl, Db, y, ctx->stream));
  return MK_OK;
}

extern "C" mk_status mk_transfer_left_leg(mk_ctx h, const double* x, const double* O, const double* A,
                                          const double* Ab, int64_t Dm, int64_t wl, int64_t De, int64_t Dn,
                                          int64_t d, int64_t wr, int64_t Dc, int64_t Da, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Dm) && small_leg(wl, De) && small_leg(wr, De) && pos(Dn) && small_leg(d, d) && pos(Dc) && pos(Da),
             MK_BAD_SHAPE, "transfer_left_leg: bad dimension");
  MK_REQUIRE(x && A && Ab && y, MK_BAD_ARG, "transfer_left_leg: null pointer");
  MK_REQUIRE(O != nullptr || wl == wr, MK_BAD_SHAPE, "transfer_left_leg: without an MPO wl must equal wr");
  const int64_t WL = wl * De, WR = wr * De;   // x[m,(g,e),n]
  double *t1 = nullptr, *t2 = nullptr, *Oe = nullptr;
  MK_TRY(carve(ctx, {{f64b(Dm * WL * d * Dc), (void**)&t1},
                     {f64b(Dm * d * WR * Dc), (void**)&t2},
                     {f64b(WL * d * d * WR), (void**)&Oe}}));
  MK_TRY(k_mpo_extend(O, (int)wl, (int)d, (int)wr, (int)De, 1, Oe, ctx->stream));
  return tl_mpo_run(ctx, x, Oe, A, Ab, Dm, WL, Dn, d, WR, Dc, Da, y, t1, t2);
}

extern "C" mk_status mk_transfer_right_leg(mk_ctx h, const double* v, const double* O, const double* A,
                                           const double* Ab, int64_t Da, int64_t d, int64_t Di, int64_t wl,
                                           int64_t wr, int64_t De, int64_t Db, int64_t Dj, double* y) {
  CTX_OR_FAIL(h);
  MK_REQUIRE(pos(Da) && small_leg(d, d) && pos(Di) && small_leg(wl, De) && small_leg(wr, De) && pos(Db) && pos(Dj),
             MK_BAD_SHAPE, "transfer_right_leg: bad dimension");
  MK_REQUIRE(v && A && Ab && y, MK_BAD_ARG, "transfer_right_leg: null pointer");
  MK_REQUIRE(O != nullptr || wl == wr, MK_BAD_SHAPE, "transfer_right_leg: without an MPO wl must equal wr");
  const int64_t WL = wl * De, WR = wr * De;   // v[i,(l,e),j]
  double *t1 = nullptr, *t2 = nullptr, *Oe = nullptr;
  MK_TRY(carve(ctx, {{f64b(Da * d * WR * Dj), (void**)&t1},
                     {f64b(Da * WL * d * Dj), (void**)&t2},
                     {f64b(WL * d * d * WR), (void**)&Oe}}));
  MK_TRY(k_mpo_extend(O, (int)wl, (int)d, (int)wr, (int)De, 1, Oe, ctx->stream));
  return tr_mpo_run(ctx, v, Oe, A, Ab, Da, d, Di, WL, WR, Db, Dj, y, t1, t2);
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
