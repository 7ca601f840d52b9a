// Per-level (BlockTensorMap) environments and block-stored Jordan MPO tensors
// at the C ABI: the layouts MPSKit's MPOHamiltonian path holds, taken without
// any densification by the caller.
//
// An MPOHamiltonian environment is a BlockTensorMap over a SumSpace on its MPO
// leg, one TensorMap per Jordan level (src/environments/abstract_envs.jl:32-58):
// level L is a column-major D1 x w_L x D2 array of its own.  A JordanMPOTensor
// stores its nonzero (row level, column level) blocks, with identity blocks as
// scalars (src/operators/jordanmpotensor.jl:44-69).  These entry points gather
// the levels into the dense D1 x w x D2 layout on the device (one strided copy
// per level, stream-ordered), assemble the small MPO tensor from its blocks,
// run the dense routine and, for transfers, scatter the result back into the
// caller's levels.
#include <vector>

#include "ops.h"

namespace mk {
namespace {

bool okd(int64_t v) { return v > 0 && v < (1ll << 31); }

mk_status level_sum(const mk_env_levels* L, const char* what, int64_t* w) {
  MK_REQUIRE(L != nullptr && L->n >= 1 && L->n <= 64 && L->ptr && L->width, MK_BAD_ARG, "%s: bad level list", what);
  int64_t s = 0;
  for (int i = 0; i < L->n; ++i) {
    MK_REQUIRE(L->ptr[i] != nullptr && L->width[i] >= 1, MK_BAD_ARG, "%s: level %d is empty", what, i);
    s += L->width[i];
  }
  *w = s;
  return MK_OK;
}

// levels (D1 x w_L x D2 each) -> dense D1 x w x D2, or back
mk_status move_levels(const mk_env_levels* L, int64_t D1, int64_t w, int64_t D2, double* dense, bool to_dense,
                      hipStream_t st) {
  int64_t off = 0;
  for (int i = 0; i < L->n; ++i) {
    const int64_t wl = L->width[i];
    // for each column f: D1 * wl contiguous doubles in the level, at channel offset off of the dense array
    double* d = dense + D1 * off;
    const size_t row = (size_t)(D1 * wl) * 8;
    if (to_dense)
      MK_CHECK_HIP(hipMemcpy2DAsync(d, (size_t)(D1 * w) * 8, L->ptr[i], row, row, (size_t)D2, hipMemcpyDeviceToDevice,
                                    st));
    else
      MK_CHECK_HIP(hipMemcpy2DAsync(L->ptr[i], row, d, (size_t)(D1 * w) * 8, row, (size_t)D2, hipMemcpyDeviceToDevice,
                                    st));
    off += wl;
  }
  return MK_OK;
}

// dense W[g,b,h,k] (wl x d x d x wr) from its blocks, uploaded to `dev`
mk_status assemble_w(const mk_jordan_mpo* J, const char* what, DevBuf& dev, hipStream_t st) {
  MK_REQUIRE(J != nullptr && okd(J->wl) && okd(J->d) && okd(J->wr) && J->nblocks >= 0, MK_BAD_ARG,
             "%s: bad MPO block list", what);
  MK_REQUIRE(J->nblocks == 0 || (J->row && J->col && J->block && J->scalar), MK_BAD_ARG,
             "%s: null block arrays", what);
  const int64_t wl = J->wl, d = J->d, wr = J->wr;
  std::vector<double> W((size_t)(wl * d * d * wr), 0.0), blk((size_t)(d * d));
  std::vector<char> seen((size_t)(wl * wr), 0);
  for (int i = 0; i < J->nblocks; ++i) {
    const int g = J->row[i], k = J->col[i];
    MK_REQUIRE(g >= 0 && g < wl && k >= 0 && k < wr, MK_BAD_ARG, "%s: block %d at (%d, %d) outside %lld x %lld",
               what, i, g, k, (long long)wl, (long long)wr);
    MK_REQUIRE(!seen[g + wl * k], MK_BAD_ARG, "%s: block (%d, %d) given twice", what, g, k);
    seen[g + wl * k] = 1;
    if (J->block[i]) {
      MK_CHECK_HIP(hipMemcpy(blk.data(), J->block[i], blk.size() * 8, hipMemcpyDeviceToHost));
    } else {
      for (int64_t b = 0; b < d; ++b)
        for (int64_t h = 0; h < d; ++h) blk[b + d * h] = b == h ? J->scalar[i] : 0.0;
    }
    for (int64_t b = 0; b < d; ++b)
      for (int64_t h = 0; h < d; ++h) W[g + wl * (b + d * (h + d * k))] = blk[b + d * h];
  }
  MK_TRY(dev.alloc(W.size() * 8));
  MK_CHECK_HIP(hipMemcpyAsync(dev.ptr, W.data(), W.size() * 8, hipMemcpyHostToDevice, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}

Ctx* ctx_of(mk_ctx h) { return reinterpret_cast<Ctx*>(h); }

}  // namespace
}  // namespace mk

using namespace mk;

extern "C" mk_status mk_prepare_ac2_jordan_levels(mk_ctx h, const mk_env_levels* GL, const mk_jordan_mpo* W1,
                                                  const mk_jordan_mpo* W2, const mk_env_levels* GR, int64_t Dl,
                                                  int64_t Dr, int precisions, mk_op* out) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(okd(Dl) && okd(Dr), MK_BAD_SHAPE, "prepare_ac2_jordan_levels: non-positive dimension");
  Ctx* ctx = ctx_of(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t wl = 0, wr = 0;
  MK_TRY(level_sum(GL, "prepare_ac2_jordan_levels GL", &wl));
  MK_TRY(level_sum(GR, "prepare_ac2_jordan_levels GR", &wr));
  MK_REQUIRE(W1 && W2 && W1->wl == wl && W2->wr == wr && W1->wr == W2->wl, MK_BAD_SHAPE,
             "prepare_ac2_jordan_levels: MPO bonds do not match the environments' levels");
  DevBuf gl, gr, w1, w2;
  MK_TRY(gl.alloc((size_t)(Dl * wl * Dl) * 8));
  MK_TRY(gr.alloc((size_t)(Dr * wr * Dr) * 8));
  MK_TRY(move_levels(GL, Dl, wl, Dl, static_cast<double*>(gl.ptr), true, st));
  MK_TRY(move_levels(GR, Dr, wr, Dr, static_cast<double*>(gr.ptr), true, st));
  MK_TRY(assemble_w(W1, "prepare_ac2_jordan_levels W1", w1, st));
  MK_TRY(assemble_w(W2, "prepare_ac2_jordan_levels W2", w2, st));
  const mk_status s = mk_prepare_ac2_jordan(h, static_cast<double*>(gl.ptr), static_cast<double*>(w1.ptr),
                                            static_cast<double*>(w2.ptr), static_cast<double*>(gr.ptr), Dl, wl,
                                            W1->d, W1->wr, W2->d, wr, Dr, precisions, out);
  (void)hipStreamSynchronize(st);
  return s;
}

extern "C" mk_status mk_prepare_ac_jordan_levels(mk_ctx h, const mk_env_levels* GL, const mk_jordan_mpo* W,
                                                 const mk_env_levels* GR, int64_t Dl, int64_t Dr, int precisions,
                                                 mk_op* out) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(okd(Dl) && okd(Dr), MK_BAD_SHAPE, "prepare_ac_jordan_levels: non-positive dimension");
  Ctx* ctx = ctx_of(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t wl = 0, wr = 0;
  MK_TRY(level_sum(GL, "prepare_ac_jordan_levels GL", &wl));
  MK_TRY(level_sum(GR, "prepare_ac_jordan_levels GR", &wr));
  MK_REQUIRE(W && W->wl == wl && W->wr == wr, MK_BAD_SHAPE,
             "prepare_ac_jordan_levels: MPO bonds do not match the environments' levels");
  DevBuf gl, gr, w;
  MK_TRY(gl.alloc((size_t)(Dl * wl * Dl) * 8));
  MK_TRY(gr.alloc((size_t)(Dr * wr * Dr) * 8));
  MK_TRY(move_levels(GL, Dl, wl, Dl, static_cast<double*>(gl.ptr), true, st));
  MK_TRY(move_levels(GR, Dr, wr, Dr, static_cast<double*>(gr.ptr), true, st));
  MK_TRY(assemble_w(W, "prepare_ac_jordan_levels W", w, st));
  const mk_status s = mk_prepare_ac_jordan(h, static_cast<double*>(gl.ptr), static_cast<double*>(w.ptr),
                                           static_cast<double*>(gr.ptr), Dl, wl, W->d, wr, Dr, precisions, out);
  (void)hipStreamSynchronize(st);
  return s;
}

extern "C" mk_status mk_prepare_c_levels(mk_ctx h, const mk_env_levels* GL, const mk_env_levels* GR, int64_t Da,
                                         int64_t Df, int64_t Di, int64_t Dc, int precisions, mk_op* out) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(okd(Da) && okd(Df) && okd(Di) && okd(Dc), MK_BAD_SHAPE, "prepare_c_levels: non-positive dimension");
  Ctx* ctx = ctx_of(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t wl = 0, wr = 0;
  MK_TRY(level_sum(GL, "prepare_c_levels GL", &wl));
  MK_TRY(level_sum(GR, "prepare_c_levels GR", &wr));
  MK_REQUIRE(wl == wr, MK_BAD_SHAPE, "prepare_c_levels: GL and GR carry %lld and %lld channels", (long long)wl,
             (long long)wr);
  DevBuf gl, gr;
  MK_TRY(gl.alloc((size_t)(Da * wl * Df) * 8));
  MK_TRY(gr.alloc((size_t)(Di * wr * Dc) * 8));
  MK_TRY(move_levels(GL, Da, wl, Df, static_cast<double*>(gl.ptr), true, st));
  MK_TRY(move_levels(GR, Di, wr, Dc, static_cast<double*>(gr.ptr), true, st));
  const mk_status s = mk_prepare_c(h, static_cast<double*>(gl.ptr), static_cast<double*>(gr.ptr), Da, Df, wl, Di,
                                   Dc, precisions, out);
  (void)hipStreamSynchronize(st);
  return s;
}

extern "C" mk_status mk_transfer_left_mpo_levels(mk_ctx h, const mk_env_levels* x, const mk_jordan_mpo* O,
                                                 const double* A, const double* Abar, int64_t Dm, int64_t Dn,
                                                 int64_t Dc, int64_t Da, const mk_env_levels* y) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(okd(Dm) && okd(Dn) && okd(Dc) && okd(Da), MK_BAD_SHAPE, "transfer_left_mpo_levels: bad dimension");
  MK_REQUIRE(A && Abar, MK_BAD_ARG, "transfer_left_mpo_levels: null pointer");
  Ctx* ctx = ctx_of(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t wl = 0, wr = 0;
  MK_TRY(level_sum(x, "transfer_left_mpo_levels x", &wl));
  MK_TRY(level_sum(y, "transfer_left_mpo_levels y", &wr));
  MK_REQUIRE(O && O->wl == wl && O->wr == wr, MK_BAD_SHAPE,
             "transfer_left_mpo_levels: MPO bonds do not match the levels of x and y");
  DevBuf xd, yd, od;
  MK_TRY(xd.alloc((size_t)(Dm * wl * Dn) * 8));
  MK_TRY(yd.alloc((size_t)(Da * wr * Dc) * 8));
  MK_TRY(move_levels(x, Dm, wl, Dn, static_cast<double*>(xd.ptr), true, st));
  MK_TRY(assemble_w(O, "transfer_left_mpo_levels O", od, st));
  mk_status s = mk_transfer_left_mpo(h, static_cast<double*>(xd.ptr), static_cast<double*>(od.ptr), A, Abar, Dm, wl,
                                     Dn, O->d, wr, Dc, Da, static_cast<double*>(yd.ptr));
  if (s == MK_OK) s = move_levels(y, Da, wr, Dc, static_cast<double*>(yd.ptr), false, st);
  (void)hipStreamSynchronize(st);
  return s;
}

extern "C" mk_status mk_transfer_right_mpo_levels(mk_ctx h, const mk_env_levels* v, const mk_jordan_mpo* O,
                                                  const double* A, const double* Abar, int64_t Da, int64_t Di,
                                                  int64_t Dc, int64_t Do, const mk_env_levels* y) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(okd(Da) && okd(Di) && okd(Dc) && okd(Do), MK_BAD_SHAPE, "transfer_right_mpo_levels: bad dimension");
  MK_REQUIRE(A && Abar, MK_BAD_ARG, "transfer_right_mpo_levels: null pointer");
  Ctx* ctx = ctx_of(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int64_t wr = 0, wl = 0;
  MK_TRY(level_sum(v, "transfer_right_mpo_levels v", &wr));
  MK_TRY(level_sum(y, "transfer_right_mpo_levels y", &wl));
  MK_REQUIRE(O && O->wl == wl && O->wr == wr, MK_BAD_SHAPE,
             "transfer_right_mpo_levels: MPO bonds do not match the levels of v and y");
  DevBuf vd, yd, od;
  MK_TRY(vd.alloc((size_t)(Di * wr * Do) * 8));
  MK_TRY(yd.alloc((size_t)(Da * wl * Dc) * 8));
  MK_TRY(move_levels(v, Di, wr, Do, static_cast<double*>(vd.ptr), true, st));
  MK_TRY(assemble_w(O, "transfer_right_mpo_levels O", od, st));
  mk_status s = mk_transfer_right_mpo(h, static_cast<double*>(vd.ptr), static_cast<double*>(od.ptr), A, Abar, Da,
                                      O->d, Di, wl, wr, Dc, Do, static_cast<double*>(yd.ptr));
  if (s == MK_OK) s = move_levels(y, Da, wl, Dc, static_cast<double*>(yd.ptr), false, st);
  (void)hipStreamSynchronize(st);
  return s;
}
