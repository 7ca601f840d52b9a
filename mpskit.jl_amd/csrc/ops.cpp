// Prepared derivative operators: construction (prepare) and application.
//
// Generic operators follow the reference's precomputed form
// (mpo_derivatives.jl:216-288): L = fuse(GL * W1), R = fuse(W2 * GR) and
// y = L (x R).  Jordan-block operators (hamiltonian_derivatives.jl) use the
// identity structure of the MPO: with GL's "not started" level and GR's
// "finished" level equal to the identity, the two-site sum over MPO channels
// (g, k, l) splits exactly into
//     k = 0 family   (I x I) x R~_0       -> IC  (includes II, ID)
//     k = end family L~_end x (I x I)     -> BE  (includes DE, EE)
//     g = 0, k mid, l = end               -> CB  (physical d^2 x d^2 term)
//     g = 0, k mid, l mid                 -> CA
//     g mid, k mid, l = end               -> AB
//     g mid, k mid, l mid                 -> AA  (connected channels only)
// and, as prepare_operator!! does (hamiltonian_derivatives.jl:371-468), the
// k = 0 family and CB fold into CA when CA exists, the k = end family folds
// into AB when AB exists, and CB folds into AB when only AB exists.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "group.h"
#include "kernels.h"
#include "ops.h"
#include "seg.h"

namespace mk {

namespace {

bool okdim(int64_t v) { return v > 0 && v < (1ll << 31); }

// host copy of a small device array
mk_status fetch(const double* dev, size_t n, std::vector<double>& out) {
  out.resize(n);
  MK_CHECK_HIP(hipMemcpy(out.data(), dev, n * 8, hipMemcpyDeviceToHost));
  for (double v : out)
    MK_REQUIRE(std::isfinite(v), MK_NONFINITE, "MPO tensor holds a non-finite entry");
  return MK_OK;
}

mk_status upload(const void* host, size_t bytes, DevBuf& buf, hipStream_t st) {
  MK_TRY(buf.alloc(bytes));
  MK_CHECK_HIP(hipMemcpyAsync(buf.ptr, host, bytes, hipMemcpyHostToDevice, st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}

struct Wh {  // host view of a dense MPO tensor W[g,b,h,k]
  std::vector<double> v;
  int wl = 0, d = 0, wr = 0;
  double& at(int g, int b, int h, int k) { return v[g + wl * (b + d * (h + d * (size_t)k))]; }
  double at(int g, int b, int h, int k) const { return v[g + wl * (b + d * (h + d * (size_t)k))]; }
  bool block_nz(int g, int k) const {
    for (int b = 0; b < d; ++b)
      for (int h = 0; h < d; ++h)
        if (at(g, b, h, k) != 0.0) return true;
    return false;
  }
  bool is_identity(int g, int k) const {
    for (int b = 0; b < d; ++b)
      for (int h = 0; h < d; ++h)
        if (at(g, b, h, k) != (b == h ? 1.0 : 0.0)) return false;
    return true;
  }
  Wh masked(const std::vector<std::vector<bool>>& keep) const {
    Wh o = *this;
    for (int g = 0; g < wl; ++g)
      for (int k = 0; k < wr; ++k)
        if (!keep[g][k])
          for (int b = 0; b < d; ++b)
            for (int h = 0; h < d; ++h) o.at(g, b, h, k) = 0.0;
    return o;
  }
};

std::vector<std::vector<bool>> mask_none(int wl, int wr) {
  return std::vector<std::vector<bool>>(wl, std::vector<bool>(wr, false));
}

// Check the Jordan (upper triangular, identity corners) structure of W.
mk_status check_jordan(const Wh& W, const char* name) {
  const bool left_boundary = (W.wl == 1), right_boundary = (W.wr == 1);
  if (!right_boundary)
    MK_REQUIRE(W.is_identity(0, 0), MK_BAD_ARG, "%s: W[start,start] must be the identity", name);
  if (!left_boundary)
    MK_REQUIRE(W.is_identity(W.wl - 1, W.wr - 1), MK_BAD_ARG, "%s: W[end,end] must be the identity", name);
  // nothing flows back into "not started" or out of "finished"
  for (int g = 1; g < W.wl && !right_boundary; ++g)
    MK_REQUIRE(!W.block_nz(g, 0), MK_BAD_ARG, "%s: W[%d, start] must vanish (Jordan form)", name, g);
  if (!left_boundary)
    for (int k = 0; k + 1 < W.wr; ++k)
      MK_REQUIRE(!W.block_nz(W.wl - 1, k), MK_BAD_ARG, "%s: W[end, %d] must vanish (Jordan form)", name, k);
  return MK_OK;
}

// The mixed two-site apply of a nearest-neighbour Jordan operator (d = 2, x*IC +
// BE*x plus the physical term) converts x and computes the physical term in one
// pass over x.
// Packed operands for the fused apply on the 256 x 256 kernel; MK_PACKED=0
// keeps the K-contiguous rows (read once).
bool packing_enabled() {
  static const bool on = [] {
    const char* e = getenv("MK_PACKED");
    return !(e && e[0] == '0');
  }();
  return on;
}

bool nn_fused(const Op* op) {
  const int64_t M = op->Dl * op->d1, N = op->Dr * op->d2;
  return op->hasP && op->IC.present() && op->BE.present() && !op->CA.present() && op->AB.empty() &&
         op->d1 == 2 && op->d2 == 2 && M % 8 == 0 && N % 8 == 0;
}

mk_status make_all_mixed(Op* op) {
  hipStream_t st = op->ctx->stream;
  const int pm = op->precisions & (MK_PREC_BF16X3 | MK_PREC_FP32);
  if (!pm) return MK_OK;
  OpMat* mats[] = {&op->gen.L, &op->gen.R, &op->IC, &op->BE, &op->CA, &op->Cp, &op->Bp, &op->TA, &op->TB};
  for (OpMat* m : mats)
    if (m->present()) MK_TRY(m->make_mixed(pm, st));
  for (OpMat& m : op->AB) MK_TRY(m.make_mixed(pm, st));
  if (op->kind == OP_AC2_JOR && nn_fused(op) && packing_enabled()) {
    const int64_t M = op->Dl * op->d1, N = op->Dr * op->d2;
    const int64_t Ks[2] = {N, M};
    // BF16x3 only: the FP32 kernel measured 2% slower on packed operands
    if ((pm & MK_PREC_BF16X3) && mixed_packed_ok(M, N, Ks, 2)) {
      MK_TRY(op->IC.make_packed(MK_PREC_BF16X3, st));
      MK_TRY(op->BE.make_packed(MK_PREC_BF16X3, st));
      op->nn_packed = true;
    }
  }
  MK_CHECK_HIP(hipStreamSynchronize(st));
  return MK_OK;
}

// Build a generic "L (x R)" block: L from (GL, W1 masked) over channels klist,
// R from (W2 masked, GR) over the same channels.
mk_status build_lxr(Op* op, LxR& blk, const double* GL, const Wh& W1, int64_t Da, int64_t Df, int g_ident,
                    const double* GR, const Wh& W2, int64_t Di, int64_t Dc, int l_ident,
                    const std::vector<int>& klist) {
  hipStream_t st = op->ctx->stream;
  const int nk = (int)klist.size();
  const int64_t d1 = W1.d, d2 = W2.d;
  DevBuf w1d, w2d, kd;
  MK_TRY(upload(W1.v.data(), W1.v.size() * 8, w1d, st));
  MK_TRY(upload(W2.v.data(), W2.v.size() * 8, w2d, st));
  MK_TRY(upload(klist.data(), klist.size() * sizeof(int), kd, st));
  MK_TRY(blk.L.alloc('A', Da * d1, Df * d1 * nk));
  MK_TRY(blk.R.alloc('B', Di * d2, nk * Dc * d2));
  MK_TRY(k_prep_L(GL, static_cast<double*>(w1d.ptr), static_cast<int*>(kd.ptr), nk, Da, Df, W1.wl, (int)d1,
                  W1.wr, g_ident, blk.L.f64(), st));
  MK_TRY(k_prep_R(GR, static_cast<double*>(w2d.ptr), static_cast<int*>(kd.ptr), nk, Di, Dc, W2.wl, (int)d2,
                  W2.wr, l_ident, blk.R.f64(), st));
  MK_CHECK_HIP(hipStreamSynchronize(st));
  blk.Mx = Df * d1;
  blk.Ki = Di * d2;
  blk.nk = nk;
  blk.Ny = Dc * d2;
  blk.Mo = Da * d1;
  return MK_OK;
}

// ---------------------------------------------------------------------------
// applies
// ---------------------------------------------------------------------------

struct BetaSeq {
  double beta;
  bool first = true;
  double next() {
    double b = first ? beta : 1.0;
    first = false;
    return b;
  }
};

mk_status apply_lxr(Op* op, const LxR& blk, const double* x, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  const int64_t Mx = blk.Mx, Ki = blk.Ki, nk = blk.nk, Ny = blk.Ny, Mo = blk.Mo;
  const size_t xr_bytes = (size_t)Mx * nk * Ny * 8;
  if (prec == MK_FP64) {
    double* xR = nullptr;
    MK_TRY(carve(ctx, {{xr_bytes, (void**)&xR}}));
    ProbSpec p1;
    p1.M = Mx; p1.N = nk * Ny; p1.C = xR; p1.ldc = Mx;
    p1.add(x, Mx, blk.R.f64(), Ki, Ki);
    MK_TRY(gemm_f64(&p1, 1, GemmShape{}, ctx->stream));
    ProbSpec p2;
    p2.M = Mo; p2.N = Ny; p2.C = y; p2.ldc = Mo; p2.alpha = alpha; p2.beta = beta;
    p2.add(blk.L.f64(), Mo, xR, Mx * nk, Mx * nk);
    return gemm_f64(&p2, 1, GemmShape{}, ctx->stream);
  }
  const bool bf = prec == MK_BF16X3;
  double* xR = nullptr;
  void *mx = nullptr, *mxr = nullptr;
  MK_TRY(carve(ctx, {{xr_bytes, (void**)&xR},
                     {view_bytes(Mx, Ki, 'A', prec), &mx},
                     {view_bytes(Mx * nk, Ny, 'B', prec), &mxr}}));
  MixedView vx, vxr;
  MK_TRY(convert_view(ctx, x, Mx, Ki, Mx, 'A', prec, &vx, mx));
  ProbSpec p1;
  p1.M = Mx; p1.N = nk * Ny; p1.C = xR; p1.ldc = Mx;
  if (bf) p1.add(vx.hi, vx.kpad, blk.R.hi.ptr, blk.R.kpad, Ki, vx.lo, blk.R.lo.ptr);
  else p1.add(vx.hi, vx.kpad, blk.R.f32.ptr, blk.R.kpad, Ki);
  MK_TRY(run_gemm(&p1, 1, prec, GemmShape{}, ctx->stream));
  MK_TRY(convert_view(ctx, xR, Mx * nk, Ny, Mx * nk, 'B', prec, &vxr, mxr));
  ProbSpec p2;
  p2.M = Mo; p2.N = Ny; p2.C = y; p2.ldc = Mo; p2.alpha = alpha; p2.beta = beta;
  if (bf) p2.add(blk.L.hi.ptr, blk.L.kpad, vxr.hi, vxr.kpad, Mx * nk, blk.L.lo.ptr, vxr.lo);
  else p2.add(blk.L.f32.ptr, blk.L.kpad, vxr.hi, vxr.kpad, Mx * nk);
  return run_gemm(&p2, 1, prec, GemmShape{}, ctx->stream);
}

// Operand pointers of one GEMM side for precision `prec`.
struct Side {
  const void* p = nullptr;
  const void* lo = nullptr;
  int64_t ld = 0;
};
Side mat_side(const OpMat& m, int prec) {
  if (prec == MK_FP64) return {m.f64(), nullptr, m.rows};
  if (prec == MK_BF16X3) return {m.hi.ptr, m.lo.ptr, m.kpad};
  return {m.f32.ptr, nullptr, m.kpad};
}
Side view_side(const MixedView& v) { return {v.hi, v.lo, v.kpad}; }
// The packed copies (OpMat::make_packed); the leading dimension is the K length.
Side packed_side(const OpMat& m, int prec) {
  const int64_t K = m.role == 'B' ? m.rows : m.cols;
  if (prec == MK_BF16X3) return {m.phi.ptr, m.plo.ptr, K};
  return {m.pf32.ptr, nullptr, K};
}

}  // namespace

void block_probs(const Op* op, int prec, int64_t xoff, int64_t sa, int64_t sb, double beta,
                 std::vector<ProbSpec>& out) {
  const int64_t M = op->Dl * op->d1, N = op->Dr * op->d2;
  const int64_t kpA = pad8(N), kpB = pad8(M);
  const int64_t eb = prec == MK_BF16X3 ? 2 : 4;
  const int64_t loA = align256(M * kpA * eb), loB = align256(N * kpB * eb);
  for (const Op::BlockProb& bp : op->blocks) {
    ProbSpec p;
    p.M = bp.m;
    p.N = bp.n;
    p.C = const_cast<double*>(static_cast<const double*>(rel_ptr(2, xoff + (bp.r0 + bp.c0 * M) * 8)));
    p.ldc = M;
    p.alpha = 1.0;
    p.beta = beta;
    for (int q = 0; q < bp.nseg; ++q) {
      const Op::BlockSeg& sg = bp.seg[q];
      if (prec == MK_FP64) {
        if (sg.kind == 'I')
          p.add(rel_ptr(1, xoff + (bp.r0 + sg.k0 * M) * 8), M, op->IC.f64() + sg.k0 + bp.c0 * N, N, sg.k);
        else
          p.add(op->BE.f64() + bp.r0 + sg.k0 * M, M, rel_ptr(1, xoff + (sg.k0 + bp.c0 * M) * 8), M, sg.k);
        continue;
      }
      const bool bf = prec == MK_BF16X3;
      if (sg.kind == 'I') {
        const OpMat& m = op->IC;
        const char* hi = static_cast<const char*>(bf ? m.hi.ptr : m.f32.ptr) + (bp.c0 * m.kpad + sg.k0) * eb;
        const char* lo = bf ? static_cast<const char*>(m.lo.ptr) + (bp.c0 * m.kpad + sg.k0) * eb : nullptr;
        const int64_t a = sa + (bp.r0 * kpA + sg.k0) * eb;
        p.add(rel_ptr(3, a), kpA, hi, m.kpad, sg.k, bf ? rel_ptr(3, a + loA) : nullptr, lo);
      } else {
        const OpMat& m = op->BE;
        const char* hi = static_cast<const char*>(bf ? m.hi.ptr : m.f32.ptr) + (bp.r0 * m.kpad + sg.k0) * eb;
        const char* lo = bf ? static_cast<const char*>(m.lo.ptr) + (bp.r0 * m.kpad + sg.k0) * eb : nullptr;
        const int64_t b = sb + (bp.c0 * kpB + sg.k0) * eb;
        p.add(hi, m.kpad, rel_ptr(3, b), kpB, sg.k, lo, bf ? rel_ptr(3, b + loB) : nullptr);
      }
    }
    out.push_back(p);
  }
}

namespace {

// Charge-block apply: y = alpha H x + beta y with H restricted to the blocks.
mk_status apply_blocks(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  hipStream_t st = ctx->stream;
  const int64_t Dl = op->Dl, Dr = op->Dr, d1 = op->d1, d2 = op->d2;
  const int64_t M = Dl * d1, N = Dr * d2, n = op->x_elems;
  MK_REQUIRE(prec == MK_FP64 || op->blocks_aligned, MK_NOT_SUPPORTED,
             "charge-block apply: mixed precision needs sector ranges in multiples of 8");
  const size_t va = prec == MK_FP64 ? 0 : view_bytes(M, N, 'A', prec);
  const size_t vb = prec == MK_FP64 ? 0 : view_bytes(M, N, 'B', prec);
  void *mv = nullptr, *tmp = nullptr;
  MK_TRY(carve(ctx, {{va + vb, &mv}, {alpha != 1.0 ? (size_t)n * 8 : 0, &tmp}}));
  double* out = alpha != 1.0 ? static_cast<double*>(tmp) : y;
  const double b0 = alpha != 1.0 ? 0.0 : beta;
  // every element: physical term (or the beta scaling); then the blocks accumulate
  const bool fused = prec != MK_FP64 && op->hasP && d1 == 2 && d2 == 2 && M % 8 == 0 && N % 8 == 0;
  if (fused) {
    const MixedView xa = view_layout(M, N, 'A', prec, mv);
    const MixedView xb = view_layout(M, N, 'B', prec, static_cast<char*>(mv) + va);
    MK_TRY(k_nn_prologue(static_cast<const double*>(op->P.ptr), x, Dl, Dr, 1.0, b0, out, prec == MK_BF16X3, xa.hi,
                         xa.lo, xa.kpad, xb.hi, xb.lo, xb.kpad, 0, st));
  } else if (op->hasP) {
    MK_TRY(k_phys2(static_cast<const double*>(op->P.ptr), x, Dl, (int)d1, Dr, (int)d2, 1.0, b0, out, st));
  } else {
    MK_TRY(k_axpby(0.0, out, b0, out, n, st));
  }
  if (prec != MK_FP64 && !fused) {
    MixedView xa, xb;
    MK_TRY(convert_view(ctx, x, M, N, M, 'A', prec, &xa, mv));
    MK_TRY(convert_view(ctx, x, M, N, M, 'B', prec, &xb, static_cast<char*>(mv) + va));
  }
  auto* plan = static_cast<GroupPlan*>(op->bplan[prec].get());
  if (!plan) {
    std::vector<ProbSpec> probs;
    block_probs(op, prec, 0, 0, (int64_t)va, 1.0, probs);
    auto gp = std::make_shared<GroupPlan>();
    const int64_t bb[NBASE] = {0, n * 8, n * 8, (int64_t)(va + vb)};
    MK_TRY(group_plan(ctx, probs.data(), (int)probs.size(), prec, GemmShape{}, bb, -1, gp.get()));
    op->bplan[prec] = gp;
    plan = gp.get();
  }
  const void* base[NBASE] = {nullptr, x, out, mv};
  MK_TRY(plan->run(base, st));
  if (alpha != 1.0) MK_TRY(k_axpby(alpha, out, beta, y, n, st));
  return MK_OK;
}

// GEMM problems of the charge blocks on compact storage: x at slot 1, y at
// slot 2, the mixed views of x at slot 3 (per block: the left-factor view
// [m][pad8(n)], then the right-factor view [n][pad8(m)]; bf16 hi then lo).
struct CompactViews {
  std::vector<int64_t> a, b, alo, blo;   // byte offsets in slot 3
  int64_t bytes = 0;
};
CompactViews compact_views(const Op* op, int prec) {
  CompactViews v;
  if (prec == MK_FP64) return v;
  const int64_t eb = prec == MK_BF16X3 ? 2 : 4;
  const bool bf = prec == MK_BF16X3;
  int64_t o = 0;
  for (const Op::BlockProb& bp : op->blocks) {
    const int64_t na = bp.m * pad8(bp.n) * eb, nb = bp.n * pad8(bp.m) * eb;
    v.a.push_back(o); o += align256(na);
    v.alo.push_back(bf ? o : 0); if (bf) o += align256(na);
    v.b.push_back(o); o += align256(nb);
    v.blo.push_back(bf ? o : 0); if (bf) o += align256(nb);
  }
  v.bytes = o;
  return v;
}

int compact_block(const Op* op, int64_t r0, int64_t c0) {
  for (size_t b = 0; b < op->blocks.size(); ++b)
    if (op->blocks[b].r0 == r0 && op->blocks[b].c0 == c0) return (int)b;
  return -1;
}

void compact_probs(const Op* op, int prec, std::vector<ProbSpec>& out) {
  const int64_t M = op->Dl * op->d1, N = op->Dr * op->d2;
  const bool bf = prec == MK_BF16X3;
  const int64_t eb = bf ? 2 : 4;
  const CompactViews cv = compact_views(op, prec);
  for (size_t b = 0; b < op->blocks.size(); ++b) {
    const Op::BlockProb& bp = op->blocks[b];
    ProbSpec p;
    p.M = bp.m;
    p.N = bp.n;
    p.C = const_cast<double*>(static_cast<const double*>(rel_ptr(2, op->boff[b] * 8)));
    p.ldc = bp.m;
    p.alpha = 1.0;
    p.beta = 1.0;
    for (int q = 0; q < bp.nseg; ++q) {
      const Op::BlockSeg& sg = bp.seg[q];
      const int xb = sg.kind == 'I' ? compact_block(op, bp.r0, sg.k0) : compact_block(op, sg.k0, bp.c0);
      if (xb < 0) continue;   // x has no block there: the product is zero
      if (prec == MK_FP64) {
        if (sg.kind == 'I')
          p.add(rel_ptr(1, op->boff[xb] * 8), bp.m, op->IC.f64() + sg.k0 + bp.c0 * N, N, sg.k);
        else
          p.add(op->BE.f64() + bp.r0 + sg.k0 * M, M, rel_ptr(1, op->boff[xb] * 8), sg.k, sg.k);
        continue;
      }
      if (sg.kind == 'I') {
        const OpMat& m = op->IC;
        const char* hi = static_cast<const char*>(bf ? m.hi.ptr : m.f32.ptr) + (bp.c0 * m.kpad + sg.k0) * eb;
        const char* lo = bf ? static_cast<const char*>(m.lo.ptr) + (bp.c0 * m.kpad + sg.k0) * eb : nullptr;
        p.add(rel_ptr(3, cv.a[xb]), pad8(sg.k), hi, m.kpad, sg.k, bf ? rel_ptr(3, cv.alo[xb]) : nullptr, lo);
      } else {
        const OpMat& m = op->BE;
        const char* hi = static_cast<const char*>(bf ? m.hi.ptr : m.f32.ptr) + (bp.r0 * m.kpad + sg.k0) * eb;
        const char* lo = bf ? static_cast<const char*>(m.lo.ptr) + (bp.r0 * m.kpad + sg.k0) * eb : nullptr;
        p.add(hi, m.kpad, rel_ptr(3, cv.b[xb]), pad8(sg.k), sg.k, lo, bf ? rel_ptr(3, cv.blo[xb]) : nullptr);
      }
    }
    out.push_back(p);
  }
}

// Charge-block apply on compact storage: the physical term (every output
// block a combination of the blocks with its bond sectors), the mixed views
// of every block, then one grouped GEMM over the blocks' products.
mk_status apply_compact(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = op->x_elems;
  MK_REQUIRE(prec == MK_FP64 || op->blocks_aligned, MK_NOT_SUPPORTED,
             "charge-block apply: mixed precision needs sector ranges in multiples of 8");
  if (prec != MK_FP64 && !op->csplit[prec].ptr) {
    // the conversion items of the mixed views, made once per precision
    const CompactViews cv = compact_views(op, prec);
    std::vector<SplitItem> items;
    int mx = 0;
    for (size_t b = 0; b < op->blocks.size(); ++b) {
      const Op::BlockProb& bp = op->blocks[b];
      for (int role = 0; role < 2; ++role) {
        SplitItem it{};
        it.src = static_cast<const double*>(rel_ptr(1, op->boff[b] * 8));
        it.rows = bp.m;
        it.cols = bp.n;
        it.ld = bp.m;
        it.trans = role;   // 0: left-factor view (K = columns), 1: right-factor view (K = rows)
        it.ldo = role == 0 ? pad8(bp.n) : pad8(bp.m);
        it.nblocks = split_blocks(it.rows, it.cols, it.trans, it.ldo);
        it.hi = const_cast<void*>(rel_ptr(3, role == 0 ? cv.a[b] : cv.b[b]));
        it.lo = prec == MK_BF16X3 ? const_cast<void*>(rel_ptr(3, role == 0 ? cv.alo[b] : cv.blo[b])) : nullptr;
        mx = std::max(mx, (int)it.nblocks);
        items.push_back(it);
      }
    }
    MK_TRY(op->csplit[prec].alloc(items.size() * sizeof(SplitItem)));
    MK_CHECK_HIP(hipMemcpyAsync(op->csplit[prec].ptr, items.data(), items.size() * sizeof(SplitItem),
                                hipMemcpyHostToDevice, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    op->ncsplit[prec] = (int)items.size();
    op->csplit_max[prec] = mx;
    op->cview[prec] = cv.bytes;
  }
  const size_t vbytes = prec == MK_FP64 ? 0 : (size_t)op->cview[prec];
  void *mv = nullptr, *tmp = nullptr;
  MK_TRY(carve(ctx, {{vbytes, &mv}, {alpha != 1.0 ? (size_t)n * 8 : 0, &tmp}}));
  double* out = alpha != 1.0 ? static_cast<double*>(tmp) : y;
  const double b0 = alpha != 1.0 ? 0.0 : beta;
  const void* base[NBASE] = {nullptr, x, out, mv};
  if (op->hasP)
    MK_TRY(cphys_group(static_cast<const CPhysItem*>(op->cphys.ptr), op->ncphys, op->cphys_max, 1.0, b0, base, st));
  else
    MK_TRY(k_axpby(0.0, out, b0, out, n, st));
  if (prec != MK_FP64)
    MK_TRY(split_group(static_cast<const SplitItem*>(op->csplit[prec].ptr), op->ncsplit[prec], op->csplit_max[prec],
                       prec == MK_BF16X3, base, st));
  auto* plan = static_cast<GroupPlan*>(op->bplan[prec].get());
  if (!plan) {
    std::vector<ProbSpec> probs;
    compact_probs(op, prec, probs);
    auto gp = std::make_shared<GroupPlan>();
    const int64_t bb[NBASE] = {0, n * 8, n * 8, (int64_t)vbytes};
    MK_TRY(group_plan(ctx, probs.data(), (int)probs.size(), prec, GemmShape{}, bb, -1, gp.get()));
    op->bplan[prec] = gp;
    plan = gp.get();
  }
  MK_TRY(plan->run(base, st));
  if (alpha != 1.0) MK_TRY(k_axpby(alpha, out, beta, y, n, st));
  return MK_OK;
}

mk_status apply_ac2_jordan(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  if (op->compact) return apply_compact(op, x, y, alpha, beta, prec);
  if (!op->blocks.empty()) return apply_blocks(op, x, y, alpha, beta, prec);
  Ctx* ctx = op->ctx;
  const int64_t Dl = op->Dl, Dr = op->Dr, d1 = op->d1, d2 = op->d2;
  const int64_t M = Dl * d1, N = Dr * d2;
  BetaSeq bs{beta};
  // mixed-precision views of x: as left factor (M x N) and as right factor (K = M rows)
  MixedView xa, xb, xca;
  // nearest-neighbour chains with d = 2: the CB term and both conversions of x
  // come from one pass over x
  const bool fused = prec != MK_FP64 && nn_fused(op);
  // packed operands (views of x written packed by the prologue, IC / BE packed at prepare)
  const bool packed = fused && op->nn_packed && prec == MK_BF16X3;
  if (prec != MK_FP64) {
    void *ma = nullptr, *mb = nullptr, *mc = nullptr;
    const bool need_a = op->IC.present(), need_b = op->BE.present() || !op->AB.empty();
    const bool need_ca = op->CA.present();
    MK_TRY(carve(ctx, {{need_a ? view_bytes(M, N, 'A', prec) : 0, &ma},
                       {need_b ? view_bytes(M, N, 'B', prec) : 0, &mb},
                       {need_ca ? view_bytes(Dl, d1 * N, 'A', prec) : 0, &mc}}));
    if (fused) {
      xa = view_layout(M, N, 'A', prec, ma);
      xb = view_layout(M, N, 'B', prec, mb);
      if (packed) {
        // the same memory, packed: no row padding, so kpad is the K length
        // (N for xa, M for xb); the lo parts stay where the row layout puts them
        xa.kpad = N;
        xb.kpad = M;
      }
    } else {
      if (need_a) MK_TRY(convert_view(ctx, x, M, N, M, 'A', prec, &xa, ma));
      if (need_b) MK_TRY(convert_view(ctx, x, M, N, M, 'B', prec, &xb, mb));
      if (need_ca) MK_TRY(convert_view(ctx, x, Dl, d1 * N, Dl, 'A', prec, &xca, mc));
    }
  }
  const GemmShape nn{};
  // two-site CB term first: a memory-bound pass that reads x once
  if (fused)
    MK_TRY(k_nn_prologue(static_cast<const double*>(op->P.ptr), x, Dl, Dr, alpha, bs.next(), y,
                         prec == MK_BF16X3, xa.hi, xa.lo, xa.kpad, xb.hi, xb.lo, xb.kpad, packed ? 1 : 0,
                         ctx->stream));
  else if (op->hasP)
    MK_TRY(k_phys2(static_cast<const double*>(op->P.ptr), x, Dl, (int)d1, Dr, (int)d2, alpha, bs.next(), y,
                   ctx->stream));
  if (op->IC.present() || op->BE.present()) {
    ProbSpec p;
    p.M = M; p.N = N; p.C = y; p.ldc = M; p.alpha = alpha; p.beta = bs.next();
    if (op->IC.present()) {
      Side a = prec == MK_FP64 ? Side{x, nullptr, M} : view_side(xa);
      Side b = packed ? packed_side(op->IC, prec) : mat_side(op->IC, prec);
      p.add(a.p, a.ld, b.p, b.ld, N, a.lo, b.lo);
      p.seg[p.nseg - 1].packA = p.seg[p.nseg - 1].packB = packed;
    }
    if (op->BE.present()) {
      Side a = packed ? packed_side(op->BE, prec) : mat_side(op->BE, prec);
      Side b = prec == MK_FP64 ? Side{x, nullptr, M} : view_side(xb);
      p.add(a.p, a.ld, b.p, b.ld, M, a.lo, b.lo);
      p.seg[p.nseg - 1].packA = p.seg[p.nseg - 1].packB = packed;
    }
    MK_TRY(run_gemm(&p, 1, prec, nn, ctx->stream));
  }
  if (op->CA.present()) {
    // y[a,(b,c,e)] += x[a,(h,i,j)] CA[(h,i,j),(b,c,e)]
    ProbSpec p;
    p.M = Dl; p.N = d1 * N; p.C = y; p.ldc = Dl; p.alpha = alpha; p.beta = bs.next();
    Side a = prec == MK_FP64 ? Side{x, nullptr, Dl} : view_side(xca);
    Side b = mat_side(op->CA, prec);
    p.add(a.p, a.ld, b.p, b.ld, d1 * N, a.lo, b.lo);
    MK_TRY(run_gemm(&p, 1, prec, nn, ctx->stream));
  }
  if (!op->AB.empty()) {
    // y_e[(a,b),c] += sum_j AB_{e,j}[(a,b),(f,h)] x_j[(f,h),c]
    const double bcur = bs.next();
    ProbSpec ps[MAX_PROB];
    for (int e = 0; e < d2; ++e) {
      ProbSpec& p = ps[e];
      p.M = M; p.N = Dr; p.C = y + e * M * Dr; p.ldc = M; p.alpha = alpha; p.beta = bcur;
      for (int j = 0; j < d2; ++j) {
        Side a = mat_side(op->AB[e + d2 * j], prec);
        Side b;
        if (prec == MK_FP64) b = Side{x + j * M * Dr, nullptr, M};
        else b = Side{static_cast<const char*>(xb.hi) + (size_t)j * Dr * xb.kpad * (prec == MK_BF16X3 ? 2 : 4),
                      xb.lo ? static_cast<const char*>(xb.lo) + (size_t)j * Dr * xb.kpad * 2 : nullptr, xb.kpad};
        p.add(a.p, a.ld, b.p, b.ld, M, a.lo, b.lo);
      }
    }
    MK_TRY(run_gemm(ps, (int)d2, prec, nn, ctx->stream));
  }
  if (op->gen.present()) MK_TRY(apply_lxr(op, op->gen, x, y, alpha, bs.next(), prec));
  if (bs.first) MK_TRY(k_axpby(0.0, y, beta, y, op->y_elems, ctx->stream));
  return MK_OK;
}

mk_status apply_ac_jordan(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  const int64_t Dl = op->Dl, Dr = op->Dr, d = op->d1;
  BetaSeq bs{beta};
  MixedView xc, xb;
  if (prec != MK_FP64) {
    void *mc = nullptr, *mb = nullptr;
    MK_TRY(carve(ctx, {{op->Cp.present() ? view_bytes(Dl, d * Dr, 'A', prec) : 0, &mc},
                       {op->Bp.present() ? view_bytes(Dl * d, Dr, 'B', prec) : 0, &mb}}));
    if (op->Cp.present()) MK_TRY(convert_view(ctx, x, Dl, d * Dr, Dl, 'A', prec, &xc, mc));
    if (op->Bp.present()) MK_TRY(convert_view(ctx, x, Dl * d, Dr, Dl * d, 'B', prec, &xb, mb));
  }
  const GemmShape nn{};
  if (op->Cp.present()) {
    // y[a,(b,c)] = x[a,(h,i)] C'[(h,i),(b,c)]
    ProbSpec p;
    p.M = Dl; p.N = d * Dr; p.C = y; p.ldc = Dl; p.alpha = alpha; p.beta = bs.next();
    Side a = prec == MK_FP64 ? Side{x, nullptr, Dl} : view_side(xc);
    Side b = mat_side(op->Cp, prec);
    p.add(a.p, a.ld, b.p, b.ld, d * Dr, a.lo, b.lo);
    MK_TRY(run_gemm(&p, 1, prec, nn, ctx->stream));
  }
  if (op->Bp.present()) {
    // y[(a,b),c] += B'[(a,b),(f,h)] x[(f,h),c]
    ProbSpec p;
    p.M = Dl * d; p.N = Dr; p.C = y; p.ldc = Dl * d; p.alpha = alpha; p.beta = bs.next();
    Side a = mat_side(op->Bp, prec);
    Side b = prec == MK_FP64 ? Side{x, nullptr, Dl * d} : view_side(xb);
    p.add(a.p, a.ld, b.p, b.ld, Dl * d, a.lo, b.lo);
    MK_TRY(run_gemm(&p, 1, prec, nn, ctx->stream));
  }
  if (op->gen.present()) MK_TRY(apply_lxr(op, op->gen, x, y, alpha, bs.next(), prec));
  if (bs.first) MK_TRY(k_axpby(0.0, y, beta, y, op->y_elems, ctx->stream));
  return MK_OK;
}

// Left transfer y = alpha x T + beta y (transfer_left, transfer.jl:125-127 for
// an MPO, :18-20 for an MPS): t1 = x A over the ket's left bond, the MPO step,
// y = Abar^T t2.  In FP32 / BF16x3 the two GEMMs run on the prepared mixed
// copies of A and Abar with x converted on the way in, and the MPO step writes
// the second GEMM's mixed operand directly; the MPS transfer has no MPO step
// (t1 is already t2).
mk_status apply_transfer_left(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  hipStream_t st = ctx->stream;
  const int64_t Dm = op->tD[0], Dn = op->tD[1], Dc = op->tD[2], Da = op->tD[3];
  const int wl = op->tw[0], d = op->tw[1], wr = op->tw[2];
  const double* O = static_cast<const double*>(op->TO.ptr);
  const int64_t M1 = Dm * wl, N1 = (int64_t)d * Dc, K2 = Dm * d, N2 = (int64_t)wr * Dc;
  if (prec == MK_FP64) {
    double *t1 = nullptr, *t2 = nullptr;
    MK_TRY(carve(ctx, {{(size_t)M1 * N1 * 8, (void**)&t1}, {op->tmpo ? (size_t)K2 * N2 * 8 : 0, (void**)&t2}}));
    ProbSpec p1;
    p1.M = M1; p1.N = N1; p1.C = t1; p1.ldc = M1;
    p1.add(x, M1, op->TA.f64(), Dn, Dn);
    MK_TRY(gemm_f64(&p1, 1, GemmShape{}, st));
    if (op->tmpo) MK_TRY(k_tl_mid(t1, O, Dm, wl, d, wr, Dc, t2, st));
    else t2 = t1;
    ProbSpec p2;
    p2.M = Da; p2.N = N2; p2.C = y; p2.ldc = Da; p2.alpha = alpha; p2.beta = beta;
    p2.add(op->TB.f64(), K2, t2, K2, K2);
    return gemm_f64(&p2, 1, GemmShape{true, false}, st);
  }
  const bool bf = prec == MK_BF16X3;
  double* t1 = nullptr;
  void *mx = nullptr, *mt = nullptr;
  MK_TRY(carve(ctx, {{(size_t)M1 * N1 * 8, (void**)&t1},
                     {view_bytes(M1, Dn, 'A', prec), &mx},
                     {view_bytes(K2, N2, 'B', prec), &mt}}));
  MixedView vx, vt;
  MK_TRY(convert_view(ctx, x, M1, Dn, M1, 'A', prec, &vx, mx));
  ProbSpec p1;
  p1.M = M1; p1.N = N1; p1.C = t1; p1.ldc = M1;
  if (bf) p1.add(vx.hi, vx.kpad, op->TA.hi.ptr, op->TA.kpad, Dn, vx.lo, op->TA.lo.ptr);
  else p1.add(vx.hi, vx.kpad, op->TA.f32.ptr, op->TA.kpad, Dn);
  MK_TRY(run_gemm(&p1, 1, prec, GemmShape{}, st));
  if (op->tmpo) {
    vt = view_layout(K2, N2, 'B', prec, mt);
    MK_TRY(k_tl_mid_mixed(t1, O, Dm, wl, d, wr, Dc, vt.kpad, bf ? 1 : 0, vt.hi, vt.lo, st));
  } else {
    MK_TRY(convert_view(ctx, t1, K2, N2, K2, 'B', prec, &vt, mt));
  }
  ProbSpec p2;
  p2.M = Da; p2.N = N2; p2.C = y; p2.ldc = Da; p2.alpha = alpha; p2.beta = beta;
  if (bf) p2.add(op->TB.hi.ptr, op->TB.kpad, vt.hi, vt.kpad, K2, op->TB.lo.ptr, vt.lo);
  else p2.add(op->TB.f32.ptr, op->TB.kpad, vt.hi, vt.kpad, K2);
  return run_gemm(&p2, 1, prec, GemmShape{}, st);
}

// Right transfer y = alpha T v + beta y (transfer_right, transfer.jl:128-130
// for an MPO, :21-23 for an MPS): t1 = A v over the ket's right bond, the MPO
// step, y = t2 Abar^T; mixed precisions as for the left transfer.
mk_status apply_transfer_right(Op* op, const double* v, double* y, double alpha, double beta, int prec) {
  Ctx* ctx = op->ctx;
  hipStream_t st = ctx->stream;
  const int64_t Da = op->tD[0], Di = op->tD[1], Dc = op->tD[2], Do = op->tD[3];
  const int wl = op->tw[0], d = op->tw[1], wr = op->tw[2];
  const double* O = static_cast<const double*>(op->TO.ptr);
  const int64_t M1 = Da * d, N1 = (int64_t)wr * Do, M2 = Da * wl, K2 = (int64_t)d * Do;
  if (prec == MK_FP64) {
    double *t1 = nullptr, *t2 = nullptr;
    MK_TRY(carve(ctx, {{(size_t)M1 * N1 * 8, (void**)&t1}, {op->tmpo ? (size_t)M2 * K2 * 8 : 0, (void**)&t2}}));
    ProbSpec p1;
    p1.M = M1; p1.N = N1; p1.C = t1; p1.ldc = M1;
    p1.add(op->TA.f64(), M1, v, Di, Di);
    MK_TRY(gemm_f64(&p1, 1, GemmShape{}, st));
    if (op->tmpo) MK_TRY(k_tr_mid(t1, O, Da, wl, d, wr, Do, t2, st));
    else t2 = t1;
    ProbSpec p2;
    p2.M = M2; p2.N = Dc; p2.C = y; p2.ldc = M2; p2.alpha = alpha; p2.beta = beta;
    p2.add(t2, M2, op->TB.f64(), Dc, K2);
    return gemm_f64(&p2, 1, GemmShape{false, true}, st);
  }
  const bool bf = prec == MK_BF16X3;
  double* t1 = nullptr;
  void *mv = nullptr, *mt = nullptr;
  MK_TRY(carve(ctx, {{(size_t)M1 * N1 * 8, (void**)&t1},
                     {view_bytes(Di, N1, 'B', prec), &mv},
                     {view_bytes(M2, K2, 'A', prec), &mt}}));
  MixedView vv, vt;
  MK_TRY(convert_view(ctx, v, Di, N1, Di, 'B', prec, &vv, mv));
  ProbSpec p1;
  p1.M = M1; p1.N = N1; p1.C = t1; p1.ldc = M1;
  if (bf) p1.add(op->TA.hi.ptr, op->TA.kpad, vv.hi, vv.kpad, Di, op->TA.lo.ptr, vv.lo);
  else p1.add(op->TA.f32.ptr, op->TA.kpad, vv.hi, vv.kpad, Di);
  MK_TRY(run_gemm(&p1, 1, prec, GemmShape{}, st));
  if (op->tmpo) {
    vt = view_layout(M2, K2, 'A', prec, mt);
    MK_TRY(k_tr_mid_mixed(t1, O, Da, wl, d, wr, Do, vt.kpad, bf ? 1 : 0, vt.hi, vt.lo, st));
  } else {
    MK_TRY(convert_view(ctx, t1, M2, K2, M2, 'A', prec, &vt, mt));
  }
  ProbSpec p2;
  p2.M = M2; p2.N = Dc; p2.C = y; p2.ldc = M2; p2.alpha = alpha; p2.beta = beta;
  if (bf) p2.add(vt.hi, vt.kpad, op->TB.hi.ptr, op->TB.kpad, K2, vt.lo, op->TB.lo.ptr);
  else p2.add(vt.hi, vt.kpad, op->TB.f32.ptr, op->TB.kpad, K2);
  return run_gemm(&p2, 1, prec, GemmShape{}, st);
}

}  // namespace

mk_status op_apply(Op* op, const double* x, double* y, double alpha, double beta, int prec) {
  MK_REQUIRE(prec >= 0 && prec <= 2 && (op->precisions & (1 << prec)), MK_BAD_ARG,
             "apply: precision %d was not prepared (mask %d)", prec, op->precisions);
  MK_REQUIRE(x != nullptr && y != nullptr, MK_BAD_ARG, "apply: null pointer");
  MK_REQUIRE(x != y, MK_BAD_ARG, "apply: x and y must not alias");
  MK_CHECK_HIP(hipSetDevice(op->ctx->device));
  switch (op->kind) {
    case OP_C:
    case OP_AC_GEN:
    case OP_AC2_GEN:
      return apply_lxr(op, op->gen, x, y, alpha, beta, prec);
    case OP_AC2_JOR:
      return apply_ac2_jordan(op, x, y, alpha, beta, prec);
    case OP_AC_JOR:
      return apply_ac_jordan(op, x, y, alpha, beta, prec);
    case OP_TL:
      return apply_transfer_left(op, x, y, alpha, beta, prec);
    case OP_TR:
      return apply_transfer_right(op, x, y, alpha, beta, prec);
  }
  return MK_NOT_SUPPORTED;
}

// Diagonal of a square operator as far as the Davidson preconditioner needs it:
// the IC, BE and two-site physical terms of Jordan two-site operators (exact for
// nearest-neighbour chains), zero for every other term and operator kind.
mk_status op_diagonal(Op* op, double* diag, hipStream_t st) {
  const int64_t n = op->x_elems;
  if (op->kind != OP_AC2_JOR) {
    MK_CHECK_HIP(hipMemsetAsync(diag, 0, (size_t)n * 8, st));
    return MK_OK;
  }
  return k_diag_nn(op->IC.present() ? op->IC.f64() : nullptr, op->BE.present() ? op->BE.f64() : nullptr,
                   op->hasP ? static_cast<const double*>(op->P.ptr) : nullptr, op->Dl, op->Dr, (int)op->d1,
                   (int)op->d2, diag, st);
}

}  // namespace mk

using namespace mk;

// ---------------------------------------------------------------------------
// C ABI: charge blocks
// ---------------------------------------------------------------------------

extern "C" mk_status mk_op_set_charge_blocks(mk_op h, int nl, const int64_t* loff, const int* lq, int nr,
                                             const int64_t* roff, const int* rq, const int* pq) {
  MK_REQUIRE(h != nullptr && loff && lq && roff && rq && pq, MK_BAD_ARG, "charge blocks: null argument");
  Op* op = reinterpret_cast<Op*>(h);
  MK_REQUIRE(op->kind == OP_AC2_JOR && !op->CA.present() && op->AB.empty() && !op->gen.present(),
             MK_NOT_SUPPORTED, "charge blocks: only nearest-neighbour two-site Jordan operators");
  MK_REQUIRE(op->d1 == op->d2 && op->d1 + op->d2 <= MAX_SEG, MK_NOT_SUPPORTED,
             "charge blocks: equal physical spaces with d1 + d2 <= %d", MAX_SEG);
  const int64_t Dl = op->Dl, Dr = op->Dr, d = op->d1;
  MK_REQUIRE(nl >= 1 && nr >= 1 && loff[0] == 0 && loff[nl] == Dl && roff[0] == 0 && roff[nr] == Dr, MK_BAD_SHAPE,
             "charge blocks: sector offsets must cover [0, Dl) and [0, Dr)");
  for (int s = 0; s < nl; ++s)
    MK_REQUIRE(loff[s] < loff[s + 1] && (s == 0 || lq[s] > lq[s - 1]), MK_BAD_SHAPE,
               "charge blocks: left sectors must be non-empty with increasing charges");
  for (int s = 0; s < nr; ++s)
    MK_REQUIRE(roff[s] < roff[s + 1] && (s == 0 || rq[s] > rq[s - 1]), MK_BAD_SHAPE,
               "charge blocks: right sectors must be non-empty with increasing charges");
  auto find = [](int n, const int* q, int c) {
    for (int s = 0; s < n; ++s)
      if (q[s] == c) return s;
    return -1;
  };
  struct Part { int64_t o, len; };
  std::vector<Op::BlockProb> blocks;
  bool aligned = Dl % 8 == 0 && Dr % 8 == 0;
  for (int s = 0; s <= nl; ++s) aligned = aligned && loff[s] % 8 == 0;
  for (int s = 0; s <= nr; ++s) aligned = aligned && roff[s] % 8 == 0;
  // middle charges c = q_left + q_phys(h) = q_right - q_phys(j)
  std::vector<int> cs;
  for (int s = 0; s < nl; ++s)
    for (int64_t hh = 0; hh < d; ++hh) cs.push_back(lq[s] + pq[hh]);
  std::sort(cs.begin(), cs.end());
  cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
  double fl = 0.0;
  for (int c : cs) {
    std::vector<Part> rows, cols;
    for (int64_t hh = 0; hh < d; ++hh) {
      const int s = find(nl, lq, c - pq[hh]);
      if (s >= 0) rows.push_back({loff[s] + Dl * hh, loff[s + 1] - loff[s]});
    }
    for (int64_t j = 0; j < d; ++j) {
      const int t = find(nr, rq, c + pq[j]);
      if (t >= 0) cols.push_back({roff[t] + Dr * j, roff[t + 1] - roff[t]});
    }
    if (rows.empty() || cols.empty()) continue;
    for (const Part& rp : rows)
      for (const Part& cp : cols) {
        Op::BlockProb bp{};
        bp.r0 = rp.o; bp.m = rp.len; bp.c0 = cp.o; bp.n = cp.len;
        if (op->IC.present())
          for (const Part& kp : cols) bp.seg[bp.nseg++] = {'I', kp.o, kp.len};
        if (op->BE.present())
          for (const Part& kp : rows) bp.seg[bp.nseg++] = {'E', kp.o, kp.len};
        for (int q = 0; q < bp.nseg; ++q) fl += 2.0 * bp.m * bp.n * bp.seg[q].k;
        blocks.push_back(bp);
      }
  }
  MK_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  op->blocks = std::move(blocks);
  op->blocks_aligned = aligned;
  for (auto& b : op->bplan) b.reset();
  op->flops = fl + (op->hasP ? 2.0 * (double)op->x_elems * op->d1 * op->d2 : 0.0);
  return MK_OK;
}

extern "C" mk_status mk_op_compact_charge_blocks(mk_op h) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "compact charge blocks: null operator");
  Op* op = reinterpret_cast<Op*>(h);
  MK_REQUIRE(!op->blocks.empty(), MK_BAD_ARG, "compact charge blocks: set the charge blocks first");
  MK_REQUIRE(!op->compact, MK_BAD_ARG, "compact charge blocks: already compact");
  MK_CHECK_HIP(hipSetDevice(op->ctx->device));
  hipStream_t st = op->ctx->stream;
  const int64_t Dl = op->Dl, Dr = op->Dr, d1 = op->d1, d2 = op->d2;
  std::vector<int64_t> off;
  int64_t nnz = 0;
  for (const Op::BlockProb& bp : op->blocks) {
    off.push_back(nnz);
    nnz += bp.m * bp.n;
  }
  std::vector<CPhysItem> items;
  int mx = 0;
  if (op->hasP) {
    std::vector<double> P((size_t)d1 * d2 * d1 * d2);
    MK_CHECK_HIP(hipMemcpyAsync(P.data(), op->P.ptr, P.size() * 8, hipMemcpyDeviceToHost, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    std::map<std::pair<int64_t, int64_t>, int> at;
    for (size_t b = 0; b < op->blocks.size(); ++b) at[{op->blocks[b].r0, op->blocks[b].c0}] = (int)b;
    for (size_t b = 0; b < op->blocks.size(); ++b) {
      const Op::BlockProb& bp = op->blocks[b];
      // y[a,pb,c,pe] += sum P[pb,pe,h,j] x[a,h,c,j]: the same bond sectors at other physical indices
      const int64_t pb = bp.r0 / Dl, lo = bp.r0 % Dl, pe = bp.c0 / Dr, ro = bp.c0 % Dr;
      CPhysItem it{};
      it.y = static_cast<double*>(const_cast<void*>(rel_ptr(2, off[b] * 8)));
      it.len = bp.m * bp.n;
      it.nblocks = cphys_blocks(it.len);
      for (int64_t j = 0; j < d2; ++j)
        for (int64_t hh = 0; hh < d1; ++hh) {
          const double c = P[pb + d1 * (pe + d2 * (hh + d1 * j))];
          auto f = at.find({lo + Dl * hh, ro + Dr * j});
          if (c == 0.0 || f == at.end()) continue;
          const Op::BlockProb& sb = op->blocks[f->second];
          MK_REQUIRE(sb.m == bp.m && sb.n == bp.n, MK_BAD_SHAPE, "compact charge blocks: inconsistent sectors");
          MK_REQUIRE(it.nsrc < CPHYS_MAX, MK_NOT_SUPPORTED, "compact charge blocks: d1 d2 > %d", CPHYS_MAX);
          it.src[it.nsrc] = static_cast<const double*>(rel_ptr(1, off[f->second] * 8));
          it.coef[it.nsrc] = c;
          ++it.nsrc;
        }
      mx = std::max(mx, (int)it.nblocks);
      items.push_back(it);
    }
    MK_TRY(op->cphys.alloc(items.size() * sizeof(CPhysItem)));
    MK_CHECK_HIP(hipMemcpyAsync(op->cphys.ptr, items.data(), items.size() * sizeof(CPhysItem), hipMemcpyHostToDevice,
                                st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
  }
  op->ncphys = (int)items.size();
  op->cphys_max = mx;
  op->boff = std::move(off);
  op->compact = true;
  op->x_elems = op->y_elems = nnz;
  for (auto& b : op->bplan) b.reset();
  return MK_OK;
}

extern "C" mk_status mk_op_charge_blocks(mk_op h, int64_t* out, int cap, int* nblocks) {
  MK_REQUIRE(h != nullptr && nblocks != nullptr, MK_BAD_ARG, "charge blocks: null argument");
  Op* op = reinterpret_cast<Op*>(h);
  *nblocks = (int)op->blocks.size();
  int64_t o = 0;
  for (size_t b = 0; b < op->blocks.size() && (int)b < cap; ++b) {
    const Op::BlockProb& bp = op->blocks[b];
    if (out) {
      out[5 * b + 0] = bp.r0;
      out[5 * b + 1] = bp.m;
      out[5 * b + 2] = bp.c0;
      out[5 * b + 3] = bp.n;
      out[5 * b + 4] = o;
    }
    o += bp.m * bp.n;
  }
  return MK_OK;
}

// ---------------------------------------------------------------------------
// C ABI: prepares
// ---------------------------------------------------------------------------

namespace {

void drop(Op* op) {
  Ctx* ctx = op->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  delete op;
  ctx->release();
}

mk_status new_op(mk_ctx h, OpKind kind, int precisions, Op** out) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "null context");
  MK_REQUIRE(out != nullptr, MK_BAD_ARG, "null output handle");
  MK_REQUIRE(precisions > 0 && precisions < 8, MK_BAD_ARG, "precisions mask %d", precisions);
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  Op* op = new Op();
  op->ctx = ctx;
  ctx->retain();
  op->kind = kind;
  op->precisions = precisions | MK_PREC_FP64;  // FP64 is always available (refinement)
  *out = op;
  return MK_OK;
}

mk_status finish(Op* op, mk_op* out) {
  mk_status s = make_all_mixed(op);
  if (s == MK_OK && hipStreamSynchronize(op->ctx->stream) != hipSuccess) {
    set_error("prepare: device error");
    s = MK_HIP_ERROR;
  }
  if (s != MK_OK) {
    drop(op);
    return s;
  }
  *out = reinterpret_cast<mk_op>(op);
  return MK_OK;
}

#define OP_TRY(expr)            \
  do {                          \
    mk_status _s = (expr);      \
    if (_s != MK_OK) {          \
      drop(op);                 \
      return _s;                \
    }                           \
  } while (0)

}  // namespace

extern "C" mk_status mk_prepare_c(mk_ctx h, const double* GL, const double* GR, int64_t Da, int64_t Df,
                                  int64_t w, int64_t Di, int64_t Dc, int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Da) && okdim(Df) && okdim(w) && okdim(Di) && okdim(Dc), MK_BAD_SHAPE,
             "prepare_c: non-positive dimension");
  MK_REQUIRE(GL && GR, MK_BAD_ARG, "prepare_c: null pointer");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_C, precisions, &op));
  hipStream_t st = op->ctx->stream;
  LxR& g = op->gen;
  // L[a,(f,g)] = GL[a,g,f]  (_transpose_tail), R = GR
  OP_TRY(g.L.alloc('A', Da, Df * w));
  OP_TRY(k_swap23(GL, Da, w, Df, g.L.f64(), st));
  OP_TRY(g.R.alloc('B', Di, w * Dc));
  hipError_t e = hipMemcpyAsync(g.R.f64(), GR, (size_t)Di * w * Dc * 8, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) { drop(op); set_error("prepare_c: copy failed"); return MK_HIP_ERROR; }
  g.Mx = Df; g.Ki = Di; g.nk = w; g.Ny = Dc; g.Mo = Da;
  op->x_elems = Df * Di;
  op->y_elems = Da * Dc;
  op->flops = g.flops();
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac_generic(mk_ctx h, const double* GL, const double* W, const double* GR,
                                           int64_t Da, int64_t Df, int64_t wl, int64_t d, int64_t wr,
                                           int64_t Di, int64_t Dc, int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Da) && okdim(Df) && okdim(wl) && okdim(d) && okdim(wr) && okdim(Di) && okdim(Dc),
             MK_BAD_SHAPE, "prepare_ac: non-positive dimension");
  MK_REQUIRE(GL && W && GR, MK_BAD_ARG, "prepare_ac: null pointer");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC_GEN, precisions, &op));
  hipStream_t st = op->ctx->stream;
  Wh Wm;
  Wm.wl = (int)wl; Wm.d = (int)d; Wm.wr = (int)wr;
  OP_TRY(fetch(W, (size_t)wl * d * d * wr, Wm.v));
  LxR& g = op->gen;
  DevBuf wd, kd;
  std::vector<int> kl(wr);
  for (int k = 0; k < wr; ++k) kl[k] = k;
  OP_TRY(upload(Wm.v.data(), Wm.v.size() * 8, wd, st));
  OP_TRY(upload(kl.data(), kl.size() * sizeof(int), kd, st));
  OP_TRY(g.L.alloc('A', Da * d, Df * d * wr));
  OP_TRY(k_prep_L(GL, static_cast<double*>(wd.ptr), static_cast<int*>(kd.ptr), (int)wr, Da, Df, (int)wl, (int)d,
                  (int)wr, -1, g.L.f64(), st));
  OP_TRY(g.R.alloc('B', Di, wr * Dc));
  hipError_t e = hipMemcpyAsync(g.R.f64(), GR, (size_t)Di * wr * Dc * 8, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) { drop(op); set_error("prepare_ac: copy failed"); return MK_HIP_ERROR; }
  if (hipStreamSynchronize(st) != hipSuccess) { drop(op); set_error("prepare_ac: sync failed"); return MK_HIP_ERROR; }
  g.Mx = Df * d; g.Ki = Di; g.nk = wr; g.Ny = Dc; g.Mo = Da * d;
  op->x_elems = Df * d * Di;
  op->y_elems = Da * d * Dc;
  op->flops = g.flops();
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac2_generic(mk_ctx h, const double* GL, const double* W1, const double* W2,
                                            const double* GR, int64_t Da, int64_t Df, int64_t wl, int64_t d1,
                                            int64_t wm, int64_t d2, int64_t wr, int64_t Di, int64_t Dc,
                                            int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Da) && okdim(Df) && okdim(wl) && okdim(d1) && okdim(wm) && okdim(d2) && okdim(wr) &&
                 okdim(Di) && okdim(Dc),
             MK_BAD_SHAPE, "prepare_ac2: non-positive dimension");
  MK_REQUIRE(GL && W1 && W2 && GR, MK_BAD_ARG, "prepare_ac2: null pointer");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC2_GEN, precisions, &op));
  Wh A, B;
  A.wl = (int)wl; A.d = (int)d1; A.wr = (int)wm;
  B.wl = (int)wm; B.d = (int)d2; B.wr = (int)wr;
  OP_TRY(fetch(W1, (size_t)wl * d1 * d1 * wm, A.v));
  OP_TRY(fetch(W2, (size_t)wm * d2 * d2 * wr, B.v));
  std::vector<int> kl(wm);
  for (int k = 0; k < wm; ++k) kl[k] = k;
  OP_TRY(build_lxr(op, op->gen, GL, A, Da, Df, -1, GR, B, Di, Dc, -1, kl));
  op->x_elems = Df * d1 * Di * d2;
  op->y_elems = Da * d1 * Dc * d2;
  op->flops = op->gen.flops();
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac2_jordan(mk_ctx h, const double* GL, const double* W1, const double* W2,
                                           const double* GR, int64_t Dl, int64_t wl, int64_t d1, int64_t wm,
                                           int64_t d2, int64_t wr, int64_t Dr, int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Dl) && okdim(wl) && okdim(d1) && okdim(wm) && okdim(d2) && okdim(wr) && okdim(Dr),
             MK_BAD_SHAPE, "prepare_ac2_jordan: non-positive dimension");
  MK_REQUIRE(wm >= 2, MK_BAD_SHAPE, "prepare_ac2_jordan: the shared MPO bond needs >= 2 channels");
  MK_REQUIRE(d2 <= MAX_PROB && d2 <= MAX_SEG, MK_NOT_SUPPORTED, "prepare_ac2_jordan: d2 > %d", MAX_SEG);
  MK_REQUIRE(GL && W1 && W2 && GR, MK_BAD_ARG, "prepare_ac2_jordan: null pointer");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC2_JOR, precisions, &op));
  hipStream_t st = op->ctx->stream;
  op->Dl = Dl; op->Dr = Dr; op->d1 = d1; op->d2 = d2;
  Wh A, B;
  A.wl = (int)wl; A.d = (int)d1; A.wr = (int)wm;
  B.wl = (int)wm; B.d = (int)d2; B.wr = (int)wr;
  OP_TRY(fetch(W1, (size_t)wl * d1 * d1 * wm, A.v));
  OP_TRY(fetch(W2, (size_t)wm * d2 * d2 * wr, B.v));
  OP_TRY(check_jordan(A, "W1"));
  OP_TRY(check_jordan(B, "W2"));

  // channel sets: mid = A-block channels of each bond
  std::vector<int> midL, midM, midR;
  if (wl > 1) for (int g = 1; g + 1 < wl; ++g) midL.push_back(g);
  for (int k = 1; k + 1 < wm; ++k) midM.push_back(k);
  if (wr > 1) for (int l = 1; l + 1 < wr; ++l) midR.push_back(l);
  const int kEnd = (int)wm - 1, lEnd = (int)wr - 1;

  auto any_nz = [](const Wh& W, const std::vector<int>& gs, const std::vector<int>& ks) {
    for (int g : gs) for (int k : ks) if (W.block_nz(g, k)) return true;
    return false;
  };
  const bool C1nz = any_nz(A, {0}, midM);
  const bool A1nz = any_nz(A, midL, midM);
  const bool B2nz = any_nz(B, midM, {lEnd});
  const bool A2nz = any_nz(B, midM, midR);
  const bool hasCA = C1nz && A2nz;
  const bool hasAB = A1nz && B2nz;

  DevBuf w1d, w2d;
  const int64_t M = Dl * d1, N = Dr * d2;
  double fl = 0.0;

  if (hasCA) {
    // W1 restricted to row 0 and k != end: k = 0 family + CB + CA
    auto keep = mask_none((int)wl, (int)wm);
    for (int k = 0; k < kEnd; ++k) keep[0][k] = true;
    Wh A1m = A.masked(keep);
    OP_TRY(upload(A1m.v.data(), A1m.v.size() * 8, w1d, st));
    OP_TRY(upload(B.v.data(), B.v.size() * 8, w2d, st));
    OP_TRY(op->CA.alloc('B', d1 * N, d1 * N));
    OP_TRY(k_prep_CA(GR, static_cast<double*>(w1d.ptr), static_cast<double*>(w2d.ptr), Dr, (int)wl, (int)d1,
                     (int)wm, (int)d2, (int)wr, lEnd, op->CA.f64(), st));
    fl += 2.0 * Dl * (double)(d1 * N) * (d1 * N);
  } else {
    // IC: k = 0 family, R~ over channel 0 with GR's "finished" level = identity
    std::vector<int> kl = {0};
    DevBuf kd;
    OP_TRY(upload(B.v.data(), B.v.size() * 8, w2d, st));
    OP_TRY(upload(kl.data(), sizeof(int), kd, st));
    OP_TRY(op->IC.alloc('B', N, N));
    OP_TRY(k_prep_R(GR, static_cast<double*>(w2d.ptr), static_cast<int*>(kd.ptr), 1, Dr, Dr, (int)wm, (int)d2,
                    (int)wr, lEnd, op->IC.f64(), st));
    OP_TRY(hipStreamSynchronize(st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    fl += 2.0 * M * N * N;
  }
  if (hasAB) {
    auto keep = mask_none((int)wl, (int)wm);
    for (int g = 0; g < wl; ++g) keep[g][kEnd] = true;  // k = end family
    for (int k : midM) {
      for (int g : midL) keep[g][k] = true;              // AB
      if (!hasCA) keep[0][k] = true;                     // CB folds here when CA is absent
    }
    Wh A1m = A.masked(keep);
    OP_TRY(upload(A1m.v.data(), A1m.v.size() * 8, w1d, st));
    OP_TRY(upload(B.v.data(), B.v.size() * 8, w2d, st));
    OP_TRY(op->ABall.alloc((size_t)M * M * d2 * d2 * 8));
    OP_TRY(k_prep_AB(GL, static_cast<double*>(w1d.ptr), static_cast<double*>(w2d.ptr), Dl, (int)wl, (int)d1,
                     (int)wm, (int)d2, (int)wr, 0, static_cast<double*>(op->ABall.ptr), st));
    op->AB.resize(d2 * d2);
    for (int64_t b = 0; b < d2 * d2; ++b)
      op->AB[b].view('A', static_cast<double*>(op->ABall.ptr) + b * M * M, M, M);
    fl += 2.0 * d2 * d2 * (double)M * M * Dr;
  } else if (wl > 1 || A.block_nz(0, kEnd)) {
    // BE: k = end family, L~ over channel end with GL's "not started" level = identity
    std::vector<int> kl = {kEnd};
    DevBuf kd;
    OP_TRY(upload(A.v.data(), A.v.size() * 8, w1d, st));
    OP_TRY(upload(kl.data(), sizeof(int), kd, st));
    OP_TRY(op->BE.alloc('A', M, M));
    OP_TRY(k_prep_L(GL, static_cast<double*>(w1d.ptr), static_cast<int*>(kd.ptr), 1, Dl, Dl, (int)wl, (int)d1,
                    (int)wm, 0, op->BE.f64(), st));
    OP_TRY(hipStreamSynchronize(st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    fl += 2.0 * M * M * N;
  }
  if (!hasCA && !hasAB) {
    // CB[b,e,h,j] = sum_{k mid} W1[0,b,h,k] W2[k,e,j,end]
    std::vector<double> cb(d1 * d2 * d1 * d2, 0.0);
    bool nz = false;
    for (int b = 0; b < d1; ++b)
      for (int e = 0; e < d2; ++e)
        for (int hh = 0; hh < d1; ++hh)
          for (int j = 0; j < d2; ++j) {
            double s = 0.0;
            for (int k : midM) s += A.at(0, b, hh, k) * B.at(k, e, j, lEnd);
            cb[b + d1 * (e + d2 * (hh + d1 * j))] = s;
            nz = nz || s != 0.0;
          }
    if (nz) {
      OP_TRY(upload(cb.data(), cb.size() * 8, op->P, st));
      op->hasP = true;
      fl += 2.0 * d1 * d2 * (double)d1 * d2 * Dl * Dr;
    }
  }
  // AA over the connected channels (hamiltonian_derivatives.jl:243-270)
  if (A1nz && A2nz) {
    std::vector<int> mids;
    for (int k : midM) {
      bool fed = false, feeds = false;
      for (int g : midL) fed = fed || A.block_nz(g, k);
      for (int l : midR) feeds = feeds || B.block_nz(k, l);
      if (fed && feeds) mids.push_back(k);
    }
    if (!mids.empty()) {
      auto keepL = mask_none((int)wl, (int)wm);
      for (int g : midL) for (int k : mids) keepL[g][k] = true;
      auto keepR = mask_none((int)wm, (int)wr);
      for (int k : mids) for (int l : midR) keepR[k][l] = true;
      OP_TRY(build_lxr(op, op->gen, GL, A.masked(keepL), Dl, Dl, -1, GR, B.masked(keepR), Dr, Dr, -1, mids));
      fl += op->gen.flops();
    }
  }
  op->x_elems = M * N;
  op->y_elems = M * N;
  op->flops = fl;
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac2_jordan_terms(mk_ctx h, const mk_jordan_ac2_terms* t, int64_t Dl, int64_t d1,
                                                 int64_t Dr, int64_t d2, int precisions, mk_op* out) {
  MK_REQUIRE(t != nullptr, MK_BAD_ARG, "prepare_ac2_jordan_terms: null terms");
  MK_REQUIRE(okdim(Dl) && okdim(d1) && okdim(Dr) && okdim(d2), MK_BAD_SHAPE,
             "prepare_ac2_jordan_terms: non-positive dimension");
  MK_REQUIRE(d2 <= MAX_PROB && d2 <= MAX_SEG, MK_NOT_SUPPORTED, "prepare_ac2_jordan_terms: d2 > %d", MAX_SEG);
  const bool hasAA = t->AA_GL || t->AA_W1 || t->AA_W2 || t->AA_GR;
  MK_REQUIRE(!hasAA || (t->AA_GL && t->AA_W1 && t->AA_W2 && t->AA_GR && okdim(t->aa_wl) && okdim(t->aa_wm) &&
                        okdim(t->aa_wr)),
             MK_BAD_ARG, "prepare_ac2_jordan_terms: AA needs GL, W1, W2, GR and positive channel counts");
  MK_REQUIRE(t->II || t->IC || t->ID || t->CB || t->CA || t->AB || t->BE || t->DE || t->EE || hasAA, MK_BAD_ARG,
             "prepare_ac2_jordan_terms: every term is missing");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC2_JOR, precisions, &op));
  hipStream_t st = op->ctx->stream;
  op->Dl = Dl; op->Dr = Dr; op->d1 = d1; op->d2 = d2;
  const int64_t M = Dl * d1, N = Dr * d2;
  const bool hasCA = t->CA != nullptr, hasAB = t->AB != nullptr;
  double fl = 0.0;
  // prepare_operator!! folds (hamiltonian_derivatives.jl:379-462): CB into CA,
  // else into AB; IC, ID, II into CA when it exists; BE, DE, EE into AB when it
  // exists.  Terms the reference would keep apart (II / ID with neither IC nor
  // CA, DE / EE with neither BE nor AB) are added into an IC / BE of their own:
  // the same operator.
  if (hasCA) {
    OP_TRY(op->CA.alloc('B', d1 * N, d1 * N));
    OP_TRY(k_asm_ca(t->CA, t->CB, t->IC, t->ID, t->II, Dr, (int)d1, (int)d2, op->CA.f64(), st));
    fl += 2.0 * Dl * (double)(d1 * N) * (d1 * N);
  } else if (t->IC || t->ID || t->II) {
    OP_TRY(op->IC.alloc('B', N, N));
    OP_TRY(k_asm_ic(t->IC, t->ID, t->II, Dr, (int)d2, op->IC.f64(), st));
    fl += 2.0 * M * (double)N * N;
  }
  if (hasAB) {
    OP_TRY(op->ABall.alloc((size_t)M * M * d2 * d2 * 8));
    OP_TRY(k_asm_ab(t->AB, hasCA ? nullptr : t->CB, t->BE, t->DE, t->EE, Dl, (int)d1, (int)d2,
                    static_cast<double*>(op->ABall.ptr), st));
    op->AB.resize(d2 * d2);
    for (int64_t b = 0; b < d2 * d2; ++b)
      op->AB[b].view('A', static_cast<double*>(op->ABall.ptr) + b * M * M, M, M);
    fl += 2.0 * d2 * d2 * (double)M * M * Dr;
  } else if (t->BE || t->DE || t->EE) {
    OP_TRY(op->BE.alloc('A', M, M));
    OP_TRY(k_asm_be(t->BE, t->DE, t->EE, Dl, (int)d1, op->BE.f64(), st));
    fl += 2.0 * M * (double)M * N;
  }
  if (t->CB && !hasCA && !hasAB) {
    const size_t nb = (size_t)d1 * d2 * d1 * d2 * 8;
    OP_TRY(op->P.alloc(nb));
    OP_TRY(hipMemcpyAsync(op->P.ptr, t->CB, nb, hipMemcpyDeviceToDevice, st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    op->hasP = true;
    fl += 2.0 * d1 * d2 * (double)d1 * d2 * Dl * Dr;
  }
  if (hasAA) {
    // AA is the reference's MPO_AC2_Hamiltonian over the connected channels
    Wh A, B;
    A.wl = (int)t->aa_wl; A.d = (int)d1; A.wr = (int)t->aa_wm;
    B.wl = (int)t->aa_wm; B.d = (int)d2; B.wr = (int)t->aa_wr;
    OP_TRY(fetch(t->AA_W1, (size_t)A.wl * d1 * d1 * A.wr, A.v));
    OP_TRY(fetch(t->AA_W2, (size_t)B.wl * d2 * d2 * B.wr, B.v));
    std::vector<int> kl(A.wr);
    for (int k = 0; k < A.wr; ++k) kl[k] = k;
    OP_TRY(build_lxr(op, op->gen, t->AA_GL, A, Dl, Dl, -1, t->AA_GR, B, Dr, Dr, -1, kl));
    fl += op->gen.flops();
  }
  op->x_elems = M * N;
  op->y_elems = M * N;
  op->flops = fl;
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac_jordan_terms(mk_ctx h, const mk_jordan_ac_terms* t, int64_t Dl, int64_t d,
                                                int64_t Dr, int precisions, mk_op* out) {
  MK_REQUIRE(t != nullptr, MK_BAD_ARG, "prepare_ac_jordan_terms: null terms");
  MK_REQUIRE(okdim(Dl) && okdim(d) && okdim(Dr), MK_BAD_SHAPE, "prepare_ac_jordan_terms: non-positive dimension");
  const bool hasA = t->A_GL || t->A_W || t->A_GR;
  MK_REQUIRE(!hasA || (t->A_GL && t->A_W && t->A_GR && okdim(t->a_wl) && okdim(t->a_wr)), MK_BAD_ARG,
             "prepare_ac_jordan_terms: A needs GL, W, GR and positive channel counts");
  MK_REQUIRE(t->D || t->I || t->E || t->C || t->B || hasA, MK_BAD_ARG, "prepare_ac_jordan_terms: every term is missing");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC_JOR, precisions, &op));
  hipStream_t st = op->ctx->stream;
  op->Dl = Dl; op->Dr = Dr; op->d1 = d; op->d2 = 1;
  double fl = 0.0;
  // prepare_operator!! folds (hamiltonian_derivatives.jl:109-157): D into C,
  // else into B; I into C; E into B.  A D or I with no C goes into a C' of its
  // own, an E with no B into a B' of its own (the same operator).
  const bool d_to_c = t->C || !t->B;
  if (t->C || t->I || (t->D && d_to_c)) {
    OP_TRY(op->Cp.alloc('B', d * Dr, d * Dr));
    OP_TRY(k_asm_cp(t->C, d_to_c ? t->D : nullptr, t->I, Dr, (int)d, op->Cp.f64(), st));
    fl += 2.0 * Dl * (double)(d * Dr) * (d * Dr);
  }
  if (t->B || t->E || (t->D && !d_to_c)) {
    OP_TRY(op->Bp.alloc('A', Dl * d, Dl * d));
    OP_TRY(k_asm_be(t->B, d_to_c ? nullptr : t->D, t->E, Dl, (int)d, op->Bp.f64(), st));
    fl += 2.0 * (double)(Dl * d) * (Dl * d) * Dr;
  }
  if (hasA) {
    // A is the reference's MPO_AC_Hamiltonian over the middle channels
    LxR& gb = op->gen;
    Wh Wa;
    Wa.wl = (int)t->a_wl; Wa.d = (int)d; Wa.wr = (int)t->a_wr;
    OP_TRY(fetch(t->A_W, (size_t)Wa.wl * d * d * Wa.wr, Wa.v));
    std::vector<int> kl(Wa.wr);
    for (int k = 0; k < Wa.wr; ++k) kl[k] = k;
    DevBuf wad, kd;
    OP_TRY(upload(Wa.v.data(), Wa.v.size() * 8, wad, st));
    OP_TRY(upload(kl.data(), kl.size() * sizeof(int), kd, st));
    const int nA = Wa.wr;
    OP_TRY(gb.L.alloc('A', Dl * d, Dl * d * nA));
    OP_TRY(k_prep_L(t->A_GL, static_cast<double*>(wad.ptr), static_cast<int*>(kd.ptr), nA, Dl, Dl, Wa.wl, (int)d,
                    Wa.wr, -1, gb.L.f64(), st));
    OP_TRY(gb.R.alloc('B', Dr, nA * Dr));
    OP_TRY(hipMemcpyAsync(gb.R.f64(), t->A_GR, (size_t)Dr * nA * Dr * 8, hipMemcpyDeviceToDevice, st) == hipSuccess
               ? MK_OK : MK_HIP_ERROR);
    OP_TRY(hipStreamSynchronize(st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    gb.Mx = Dl * d; gb.Ki = Dr; gb.nk = nA; gb.Ny = Dr; gb.Mo = Dl * d;
    fl += gb.flops();
  }
  op->x_elems = Dl * d * Dr;
  op->y_elems = Dl * d * Dr;
  op->flops = fl;
  return finish(op, out);
}

extern "C" mk_status mk_prepare_ac_jordan(mk_ctx h, const double* GL, const double* W, const double* GR,
                                          int64_t Dl, int64_t wl, int64_t d, int64_t wr, int64_t Dr,
                                          int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Dl) && okdim(wl) && okdim(d) && okdim(wr) && okdim(Dr), MK_BAD_SHAPE,
             "prepare_ac_jordan: non-positive dimension");
  MK_REQUIRE(GL && W && GR, MK_BAD_ARG, "prepare_ac_jordan: null pointer");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_AC_JOR, precisions, &op));
  hipStream_t st = op->ctx->stream;
  op->Dl = Dl; op->Dr = Dr; op->d1 = d; op->d2 = 1;
  Wh Wm;
  Wm.wl = (int)wl; Wm.d = (int)d; Wm.wr = (int)wr;
  OP_TRY(fetch(W, (size_t)wl * d * d * wr, Wm.v));
  OP_TRY(check_jordan(Wm, "W"));
  const int lEnd = (int)wr - 1;
  std::vector<int> midL, midR;
  if (wl > 1) for (int g = 1; g + 1 < wl; ++g) midL.push_back(g);
  if (wr > 1) for (int l = 1; l + 1 < wr; ++l) midR.push_back(l);
  double fl = 0.0;
  DevBuf wd;
  OP_TRY(upload(Wm.v.data(), Wm.v.size() * 8, wd, st));
  // C'[(h,i),(b,c)] = sum_l W[0,b,h,l] GR~(i,l,c): row 0 (I, C and D)
  OP_TRY(op->Cp.alloc('B', d * Dr, d * Dr));
  OP_TRY(k_prep_Cp(GR, static_cast<double*>(wd.ptr), Dr, Dr, (int)wl, (int)d, (int)wr, lEnd, op->Cp.f64(), st));
  fl += 2.0 * Dl * (double)(d * Dr) * (d * Dr);
  if (wl > 1) {
    // B'[(a,b),(f,h)] = sum_{g>0} GL(a,g,f) W[g,b,h,end]: B and E
    auto keep = mask_none((int)wl, (int)wr);
    for (int g = 1; g < wl; ++g) keep[g][lEnd] = true;
    Wh Wb = Wm.masked(keep);
    DevBuf wbd, kd;
    std::vector<int> kl = {lEnd};
    OP_TRY(upload(Wb.v.data(), Wb.v.size() * 8, wbd, st));
    OP_TRY(upload(kl.data(), sizeof(int), kd, st));
    OP_TRY(op->Bp.alloc('A', Dl * d, Dl * d));
    OP_TRY(k_prep_L(GL, static_cast<double*>(wbd.ptr), static_cast<int*>(kd.ptr), 1, Dl, Dl, (int)wl, (int)d,
                    (int)wr, 0, op->Bp.f64(), st));
    OP_TRY(hipStreamSynchronize(st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    fl += 2.0 * (double)(Dl * d) * (Dl * d) * Dr;
  }
  bool anyA = false;
  for (int g : midL) for (int l : midR) anyA = anyA || Wm.block_nz(g, l);
  if (anyA) {
    // A: generic single-site operator over the middle channels
    auto keep = mask_none((int)wl, (int)wr);
    for (int g : midL) for (int l : midR) keep[g][l] = true;
    Wh Wa = Wm.masked(keep);
    LxR& gb = op->gen;
    DevBuf wad, kd;
    OP_TRY(upload(Wa.v.data(), Wa.v.size() * 8, wad, st));
    OP_TRY(upload(midR.data(), midR.size() * sizeof(int), kd, st));
    const int nA = (int)midR.size();
    OP_TRY(gb.L.alloc('A', Dl * d, Dl * d * nA));
    OP_TRY(k_prep_L(GL, static_cast<double*>(wad.ptr), static_cast<int*>(kd.ptr), nA, Dl, Dl, (int)wl, (int)d,
                    (int)wr, -1, gb.L.f64(), st));
    OP_TRY(gb.R.alloc('B', Dr, nA * Dr));
    OP_TRY(k_env_channels(GR, Dr, (int)wr, Dr, static_cast<int*>(kd.ptr), nA, gb.R.f64(), st));
    OP_TRY(hipStreamSynchronize(st) == hipSuccess ? MK_OK : MK_HIP_ERROR);
    gb.Mx = Dl * d; gb.Ki = Dr; gb.nk = nA; gb.Ny = Dr; gb.Mo = Dl * d;
    fl += gb.flops();
  }
  op->x_elems = Dl * d * Dr;
  op->y_elems = Dl * d * Dr;
  op->flops = fl;
  return finish(op, out);
}

// ---------------------------------------------------------------------------
// C ABI: prepared transfer matrices
// ---------------------------------------------------------------------------

namespace {

// Copies of A and Abar (and O) owned by the operator, so that later changes
// of the caller's tensors do not reach it.
mk_status own_copy(OpMat& m, char role, const double* src, int64_t rows, int64_t cols, hipStream_t st) {
  MK_TRY(m.alloc(role, rows, cols));
  MK_CHECK_HIP(hipMemcpyAsync(m.f64(), src, (size_t)rows * cols * 8, hipMemcpyDeviceToDevice, st));
  return MK_OK;
}

}  // namespace

extern "C" mk_status mk_prepare_transfer_left(mk_ctx h, const double* O, const double* A, const double* Abar,
                                              int64_t Dm, int64_t wl, int64_t Dn, int64_t d, int64_t wr, int64_t Dc,
                                              int64_t Da, int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Dm) && okdim(wl) && okdim(Dn) && okdim(d) && okdim(wr) && okdim(Dc) && okdim(Da), MK_BAD_SHAPE,
             "prepare_transfer_left: non-positive dimension");
  MK_REQUIRE(A && Abar, MK_BAD_ARG, "prepare_transfer_left: null pointer");
  MK_REQUIRE(O || (wl == 1 && wr == 1), MK_BAD_SHAPE, "prepare_transfer_left: an MPS transfer has wl = wr = 1");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_TL, precisions, &op));
  hipStream_t st = op->ctx->stream;
  // A (Dn x d Dc) is the right factor of t1 = x A; Abar (Dm d x Da), K = Dm d
  // fast, is stored as a right factor too, which makes its mixed copy the
  // K-contiguous left factor Abar^T of y = Abar^T t2
  OP_TRY(own_copy(op->TA, 'B', A, Dn, d * Dc, st));
  OP_TRY(own_copy(op->TB, 'B', Abar, Dm * d, Da, st));
  if (O) {
    op->tmpo = true;
    OP_TRY(op->TO.alloc((size_t)wl * d * d * wr * 8));
    hipError_t e = hipMemcpyAsync(op->TO.ptr, O, (size_t)wl * d * d * wr * 8, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) { drop(op); set_error("prepare_transfer_left: copy failed"); return MK_HIP_ERROR; }
  }
  op->tD[0] = Dm; op->tD[1] = Dn; op->tD[2] = Dc; op->tD[3] = Da;
  op->tw[0] = (int)wl; op->tw[1] = (int)d; op->tw[2] = (int)wr;
  op->x_elems = Dm * wl * Dn;
  op->y_elems = Da * wr * Dc;
  op->flops = 2.0 * Dm * wl * Dn * d * Dc + 2.0 * Da * Dm * d * wr * Dc;
  return finish(op, out);
}

extern "C" mk_status mk_prepare_transfer_right(mk_ctx h, const double* O, const double* A, const double* Abar,
                                               int64_t Da, int64_t d, int64_t Di, int64_t wl, int64_t wr, int64_t Dc,
                                               int64_t Do, int precisions, mk_op* out) {
  MK_REQUIRE(okdim(Da) && okdim(d) && okdim(Di) && okdim(wl) && okdim(wr) && okdim(Dc) && okdim(Do), MK_BAD_SHAPE,
             "prepare_transfer_right: non-positive dimension");
  MK_REQUIRE(A && Abar, MK_BAD_ARG, "prepare_transfer_right: null pointer");
  MK_REQUIRE(O || (wl == 1 && wr == 1), MK_BAD_SHAPE, "prepare_transfer_right: an MPS transfer has wl = wr = 1");
  Op* op = nullptr;
  MK_TRY(new_op(h, OP_TR, precisions, &op));
  hipStream_t st = op->ctx->stream;
  // A (Da d x Di) is the left factor of t1 = A v; Abar (Dc x d Do) is stored
  // as a left factor, which makes its mixed copy the K-contiguous right factor
  // Abar^T of y = t2 Abar^T
  OP_TRY(own_copy(op->TA, 'A', A, Da * d, Di, st));
  OP_TRY(own_copy(op->TB, 'A', Abar, Dc, d * Do, st));
  if (O) {
    op->tmpo = true;
    OP_TRY(op->TO.alloc((size_t)wl * d * d * wr * 8));
    hipError_t e = hipMemcpyAsync(op->TO.ptr, O, (size_t)wl * d * d * wr * 8, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) { drop(op); set_error("prepare_transfer_right: copy failed"); return MK_HIP_ERROR; }
  }
  op->tD[0] = Da; op->tD[1] = Di; op->tD[2] = Dc; op->tD[3] = Do;
  op->tw[0] = (int)wl; op->tw[1] = (int)d; op->tw[2] = (int)wr;
  op->x_elems = Di * wr * Do;
  op->y_elems = Da * wl * Dc;
  op->flops = 2.0 * Da * d * Di * wr * Do + 2.0 * Da * wl * d * Do * Dc;
  return finish(op, out);
}

extern "C" mk_status mk_apply(mk_op h, const double* x, double* y, double alpha, double beta, int precision) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "apply: null operator");
  return op_apply(reinterpret_cast<Op*>(h), x, y, alpha, beta, precision);
}

extern "C" mk_status mk_op_info(mk_op h, int64_t* x_elems, int64_t* y_elems, double* flops) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "op_info: null operator");
  Op* op = reinterpret_cast<Op*>(h);
  if (x_elems) *x_elems = op->x_elems;
  if (y_elems) *y_elems = op->y_elems;
  if (flops) *flops = op->flops;
  return MK_OK;
}

extern "C" mk_status mk_op_free(mk_op h) {
  if (!h) return MK_OK;
  Op* op = reinterpret_cast<Op*>(h);
  Ctx* ctx = op->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete op;
  ctx->release();
  return MK_OK;
}
