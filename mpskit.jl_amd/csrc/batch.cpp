// Batches of prepared derivative operators applied and solved together.
//
// A batch owns a segmented layout: instance i's vector occupies
// [off_i, off_i + n_i) of one buffer, off_i a multiple of SEG_CHUNK, padding
// zero.  When every operator is a nearest-neighbour Jordan two-site operator
// (terms IC, BE and the two-site physical term only, the MPSKit prepare result
// for TFIM/XXZ-type chains, hamiltonian_derivatives.jl:371-468) one apply of the
// whole batch is three launches: a grouped physical-term pass, (mixed modes) a
// grouped operand conversion, and one grouped persistent GEMM over all
// instances' x*IC + BE*x products.  Other operators are applied one by one.
//
// The batched eigensolver runs the thick-restart Lanczos / FP64-refinement
// scheme of mk_eigsolve for every instance in lockstep: one grouped apply per
// Krylov step, segmented float64 Gram-Schmidt, and one device->host readback per
// step.  Every instance keeps the same number of Ritz vectors at a restart (the
// fewest any active instance has), so the basis layout is common.
// Optionally the device work of Krylov step j is captured once into a hipGraph
// and replayed (all buffers and descriptor arrays it touches are owned by the
// batch and keep their addresses).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "group.h"
#include "kernels.h"
#include "ops.h"
#include "seg.h"

namespace mk {

void lanczos_lowest(const std::vector<double>& G, int ld, int k, double& lam, std::vector<double>& s);
void lanczos_lowest_p(const std::vector<double>& G, int ld, int k, int p, std::vector<double>& lam,
                      std::vector<double>& S);
bool lanczos_local_first_pass();

namespace {

// First basis column of the first Gram-Schmidt pass of krylov_step(j): the
// whole basis for the first Krylov vector of a cycle (it couples to x and the
// kept Ritz vectors), else the last three columns (the three-term recurrence;
// the full second pass removes the rest).
int first_pass_lo(int j, bool first) {
  return (first || !lanczos_local_first_pass()) ? 0 : std::max(0, j - 1);
}

struct PrecPlan {
  bool ready = false;
  GroupPlan gemm;
  DevBuf split;
  int nsplit = 0, split_max = 0;
  // nearest-neighbour d = 2 batches: CB term and both conversions in one pass
  DevBuf nn;
  int nnn = 0, nn_max = 0;
  int64_t scratch_bytes = 0;
};

}  // namespace

struct Batch {
  Ctx* ctx = nullptr;
  std::vector<Op*> ops;
  int n = 0;
  std::vector<int64_t> off, len;
  int64_t total = 0;
  DevBuf segtab;
  Segs sg;
  bool grouped = false;
  int d12 = 0;
  DevBuf phys;
  int nphys = 0, phys_max = 0;
  PrecPlan plan[3];
  DevBuf scratch;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  // eigensolver state (grown on demand; graphs refer to these addresses)
  int ks = 0, kq = 0;      // Krylov columns of Q, and the Ritz-vector columns after them
  DevBuf Q, W, U, coef, partial;
  double* host = nullptr;  // pinned readback / upload buffer
  size_t host_doubles = 0;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;

  ~Batch() {
    drop_graphs();
    if (host) (void)hipHostFree(host);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  int64_t xoff(int i) const { return off[i] * 8; }
  mk_status ensure_plan(int prec);
  // wait for the batch's work on its own stream and on the context's stream
  // (never device-wide: other contexts' streams keep running)
  mk_status wait_streams() {
    MK_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (stream) MK_CHECK_HIP(hipStreamSynchronize(stream));
    return MK_OK;
  }
  mk_status apply(const double* x, double* y, int prec, hipStream_t st);
  mk_status krylov_step(int j, int prec, bool first, hipStream_t st);
};

namespace {

bool nn_jordan(const Op* op) {
  // charge-blocked operators leave y outside their blocks to the physical pass
  return op->kind == OP_AC2_JOR && !op->CA.present() && op->AB.empty() && !op->gen.present() &&
         (op->IC.present() || op->BE.present()) && (op->blocks.empty() || op->hasP) && !op->compact;
}

}  // namespace

mk_status Batch::ensure_plan(int prec) {
  PrecPlan& pp = plan[prec];
  if (pp.ready) return MK_OK;
  for (Op* op : ops)
    MK_REQUIRE(op->precisions & (1 << prec), MK_BAD_ARG, "batch: precision %d not prepared for every operator", prec);
  if (!grouped) {
    pp.ready = true;
    return MK_OK;
  }
  const bool mixed = prec != MK_FP64;
  std::vector<ProbSpec> probs(n);
  std::vector<SplitItem> items;
  std::vector<NNItem> nn_items;
  bool all_nn = mixed;
  std::vector<ProbSpec> block_specs;
  int64_t sbytes = 0;
  for (int i = 0; i < n; ++i) {
    Op* op = ops[i];
    const int64_t M = op->Dl * op->d1, N = op->Dr * op->d2;
    ProbSpec& p = probs[i];
    p.M = M;
    p.N = N;
    p.C = const_cast<double*>(static_cast<const double*>(rel_ptr(2, xoff(i))));
    p.ldc = M;
    p.alpha = 1.0;
    p.beta = op->hasP ? 1.0 : 0.0;
    const void* x = rel_ptr(1, xoff(i));
    if (!mixed) {
      if (!op->blocks.empty()) {
        block_probs(op, prec, xoff(i), 0, 0, 1.0, block_specs);
        continue;
      }
      if (op->IC.present()) p.add(x, M, op->IC.f64(), N, N);
      if (op->BE.present()) p.add(op->BE.f64(), M, x, M, M);
      continue;
    }
    // x as left factor (K = N contiguous rows) and as right factor (K = M)
    const bool bf = prec == MK_BF16X3;
    const int64_t lda = pad8(N), ldb = pad8(M);
    const int64_t eb = bf ? 2 : 4;
    const int64_t a_hi = sbytes, a_lo = a_hi + align256(M * lda * eb);
    const int64_t b_hi = a_lo + (bf ? align256(M * lda * eb) : 0), b_lo = b_hi + align256(N * ldb * eb);
    sbytes = b_lo + (bf ? align256(N * ldb * eb) : 0);
    if (op->IC.present()) {
      SplitItem s{static_cast<const double*>(x), M, N, M, lda, 0, split_blocks(M, N, 0, lda),
                  const_cast<void*>(rel_ptr(3, a_hi)), bf ? const_cast<void*>(rel_ptr(3, a_lo)) : nullptr};
      items.push_back(s);
      const OpMat& m = op->IC;
      p.add(rel_ptr(3, a_hi), lda, bf ? m.hi.ptr : m.f32.ptr, m.kpad, N, bf ? rel_ptr(3, a_lo) : nullptr,
            bf ? m.lo.ptr : nullptr);
    }
    if (op->BE.present()) {
      SplitItem s{static_cast<const double*>(x), M, N, M, ldb, 1, split_blocks(M, N, 1, ldb),
                  const_cast<void*>(rel_ptr(3, b_hi)), bf ? const_cast<void*>(rel_ptr(3, b_lo)) : nullptr};
      items.push_back(s);
      const OpMat& m = op->BE;
      p.add(bf ? m.hi.ptr : m.f32.ptr, m.kpad, rel_ptr(3, b_hi), ldb, M, bf ? m.lo.ptr : nullptr,
            bf ? rel_ptr(3, b_lo) : nullptr);
    }
    if (op->hasP && op->IC.present() && op->BE.present() && op->blocks.empty() && op->d1 == 2 && op->d2 == 2 &&
        M % 8 == 0 && N % 8 == 0) {
      NNItem it{};
      it.P = static_cast<const double*>(op->P.ptr);
      it.x = static_cast<const double*>(x);
      it.y = const_cast<double*>(static_cast<const double*>(rel_ptr(2, xoff(i))));
      it.xa_hi = const_cast<void*>(rel_ptr(3, a_hi));
      it.xa_lo = bf ? const_cast<void*>(rel_ptr(3, a_lo)) : nullptr;
      it.xb_hi = const_cast<void*>(rel_ptr(3, b_hi));
      it.xb_lo = bf ? const_cast<void*>(rel_ptr(3, b_lo)) : nullptr;
      it.Dl = op->Dl;
      it.Dr = op->Dr;
      it.ka = lda;
      it.kb = ldb;
      nn_tiles(op->Dl, op->Dr, it.tiles_a, it.tiles);
      it.alpha = 1.0;
      it.beta = 0.0;
      nn_items.push_back(it);
    } else {
      all_nn = false;
    }
    if (!op->blocks.empty()) block_probs(op, prec, xoff(i), a_hi, b_hi, 1.0, block_specs);
  }
  // operators with charge blocks contribute their block problems instead of the dense one
  {
    std::vector<ProbSpec> kept;
    for (int i = 0; i < n; ++i)
      if (ops[i]->blocks.empty()) kept.push_back(probs[i]);
    kept.insert(kept.end(), block_specs.begin(), block_specs.end());
    probs.swap(kept);
  }
  if (all_nn && !nn_items.empty() && nn_items.size() == (size_t)n) {
    // the fused pass replaces the physical pass and both conversions
    MK_TRY(pp.nn.alloc(nn_items.size() * sizeof(NNItem)));
    MK_CHECK_HIP(hipMemcpy(pp.nn.ptr, nn_items.data(), nn_items.size() * sizeof(NNItem), hipMemcpyHostToDevice));
    pp.nnn = (int)nn_items.size();
    for (const NNItem& it : nn_items) pp.nn_max = std::max(pp.nn_max, it.tiles);
    items.clear();
  }
  if (!items.empty()) {
    MK_TRY(pp.split.alloc(items.size() * sizeof(SplitItem)));
    MK_CHECK_HIP(hipMemcpy(pp.split.ptr, items.data(), items.size() * sizeof(SplitItem), hipMemcpyHostToDevice));
    pp.nsplit = (int)items.size();
    for (const SplitItem& s : items) pp.split_max = std::max(pp.split_max, s.nblocks);
  }
  pp.scratch_bytes = sbytes;
  if (sbytes > 0 && (int64_t)scratch.bytes < sbytes) {
    MK_TRY(wait_streams());
    drop_graphs();
    scratch.release();
    MK_TRY(scratch.alloc((size_t)sbytes));
  }
  const int64_t base_bytes[NBASE] = {0, total * 8, total * 8, sbytes};
  for (int i = 0; i < n; ++i)
    MK_REQUIRE(prec == MK_FP64 || ops[i]->blocks.empty() || ops[i]->blocks_aligned, MK_NOT_SUPPORTED,
               "batch: charge blocks of operator %d are not multiples of 8 (mixed precision)", i);
  MK_TRY(group_plan(ctx, probs.data(), (int)probs.size(), prec, GemmShape{}, base_bytes, -1, &pp.gemm));
  pp.ready = true;
  return MK_OK;
}

mk_status Batch::apply(const double* x, double* y, int prec, hipStream_t st) {
  MK_TRY(ensure_plan(prec));
  if (!grouped) {
    for (int i = 0; i < n; ++i) MK_TRY(op_apply(ops[i], x + off[i], y + off[i], 1.0, 0.0, prec));
    return MK_OK;
  }
  const PrecPlan& pp = plan[prec];
  const void* base[NBASE] = {nullptr, x, y, scratch.ptr};
  if (pp.nnn) {
    MK_TRY(nn_group(static_cast<const NNItem*>(pp.nn.ptr), pp.nnn, pp.nn_max, prec == MK_BF16X3, base, st));
    return pp.gemm.run(base, st);
  }
  if (nphys) MK_TRY(phys2_group(static_cast<const PhysItem*>(phys.ptr), nphys, phys_max, d12, base, st));
  if (pp.nsplit)
    MK_TRY(split_group(static_cast<const SplitItem*>(pp.split.ptr), pp.nsplit, pp.split_max, prec == MK_BF16X3, base,
                       st));
  return pp.gemm.run(base, st);
}

// Device work of Krylov step j: U = H_p v_j, two Gram-Schmidt passes against
// Q[:, 0..j+1], |U|^2, v_{j+1} = U / |U| and <v_{j+1}, W>; coefficients are
// copied to the pinned host buffer.  Layout of `coef`: c1 | c2 (nseg x ks each),
// nn | xw (nseg each).  The first pass covers columns first_pass_lo(j)..j+1
// (entries of c1 below it are stale and read as zero).
mk_status Batch::krylov_step(int j, int prec, bool first, hipStream_t st) {
  double* Qp = static_cast<double*>(Q.ptr);
  double* Up = static_cast<double*>(U.ptr);
  double* Wp = static_cast<double*>(W.ptr);
  double* c = static_cast<double*>(coef.ptr);
  double* part = static_cast<double*>(partial.ptr);
  const int ns = sg.nseg;
  double* c1 = c;
  double* c2 = c + (int64_t)ns * ks;
  double* nn = c + 2 * (int64_t)ns * ks;
  double* xw = nn + ns;
  auto col = [&](int q) { return Qp + (int64_t)q * total; };
  MK_TRY(apply(col(j + 1), Up, prec, st));
  const int lo = first_pass_lo(j, first);
  MK_TRY(seg_dots(sg, col(lo), total, j + 2 - lo, Up, c1 + lo, ks, part, st));
  MK_TRY(seg_multiaxpy(sg, col(lo), total, j + 2 - lo, c1 + lo, ks, Up, st));
  MK_TRY(seg_dots(sg, Qp, total, j + 2, Up, c2, ks, part, st));
  MK_TRY(seg_multiaxpy(sg, Qp, total, j + 2, c2, ks, Up, st));
  MK_TRY(seg_dots(sg, Up, total, 1, Up, nn, 1, part, st));
  MK_TRY(seg_normalize(sg, Up, nn, 1, col(j + 2), st));
  MK_TRY(seg_dots(sg, col(j + 2), total, 1, Wp, xw, 1, part, st));
  MK_CHECK_HIP(hipMemcpyAsync(host, c, (size_t)(2 * ns * ks + 2 * ns) * 8, hipMemcpyDeviceToHost, st));
  return MK_OK;
}

}  // namespace mk

using namespace mk;

extern "C" {

mk_status mk_batch_create(mk_ctx h, int n, const mk_op* ops, mk_batch* out) {
  MK_REQUIRE(h != nullptr && out != nullptr, MK_BAD_ARG, "batch: null argument");
  MK_REQUIRE(n >= 1 && n < 65536 && ops != nullptr, MK_BAD_ARG, "batch: %d operators", n);
  Ctx* ctx = reinterpret_cast<Ctx*>(h);
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  Batch* b = new Batch();
  b->ctx = ctx;
  b->n = n;
  b->grouped = true;
  int64_t pos = 0;
  std::vector<int> chunk_seg, first, count;
  for (int i = 0; i < n; ++i) {
    Op* op = reinterpret_cast<Op*>(ops[i]);
    if (op == nullptr || op->ctx != ctx || op->x_elems != op->y_elems) {
      delete b;
      MK_REQUIRE(false, MK_BAD_ARG, "batch: operator %d is null, square-less or from another context", i);
    }
    b->ops.push_back(op);
    b->grouped = b->grouped && nn_jordan(op);
    b->off.push_back(pos);
    b->len.push_back(op->x_elems);
    const int64_t nch = std::max<int64_t>(1, cdiv(op->x_elems, SEG_CHUNK));
    first.push_back((int)(pos / SEG_CHUNK));
    count.push_back((int)nch);
    for (int64_t c = 0; c < nch; ++c) chunk_seg.push_back(i);
    pos += nch * SEG_CHUNK;
  }
  b->total = pos;
  // the grouped physical pass needs one d1*d2 in {4, 9, 16}
  std::vector<PhysItem> phys;
  for (int i = 0; i < n && b->grouped; ++i) {
    Op* op = b->ops[i];
    if (!op->hasP) continue;
    const int d12 = (int)(op->d1 * op->d2);
    if ((d12 != 4 && d12 != 9 && d12 != 16) || (b->d12 && b->d12 != d12)) {
      b->grouped = false;
      break;
    }
    b->d12 = d12;
    PhysItem it{static_cast<const double*>(op->P.ptr), static_cast<const double*>(rel_ptr(1, b->xoff(i))),
                const_cast<double*>(static_cast<const double*>(rel_ptr(2, b->xoff(i)))), op->Dl, op->Dr,
                (int32_t)op->d1, (int32_t)op->d2, phys2_blocks(op->Dl, op->Dr), 0, 1.0, 0.0};
    phys.push_back(it);
    b->phys_max = std::max(b->phys_max, it.nblocks);
  }
  mk_status s = MK_OK;
  const size_t tb = (chunk_seg.size() + 2 * (size_t)n) * sizeof(int);
  if ((s = b->segtab.alloc(tb)) == MK_OK) {
    std::vector<int> tab(chunk_seg);
    tab.insert(tab.end(), first.begin(), first.end());
    tab.insert(tab.end(), count.begin(), count.end());
    if (hipMemcpy(b->segtab.ptr, tab.data(), tb, hipMemcpyHostToDevice) != hipSuccess) s = MK_HIP_ERROR;
    const int* base = static_cast<const int*>(b->segtab.ptr);
    b->sg.nseg = n;
    b->sg.nchunks = (int)chunk_seg.size();
    b->sg.chunk_seg = base;
    b->sg.first = base + chunk_seg.size();
    b->sg.count = base + chunk_seg.size() + n;
  }
  if (s == MK_OK && b->grouped && !phys.empty()) {
    s = b->phys.alloc(phys.size() * sizeof(PhysItem));
    if (s == MK_OK &&
        hipMemcpy(b->phys.ptr, phys.data(), phys.size() * sizeof(PhysItem), hipMemcpyHostToDevice) != hipSuccess)
      s = MK_HIP_ERROR;
    b->nphys = (int)phys.size();
  }
  if (s == MK_OK && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) s = MK_HIP_ERROR;
  if (s == MK_OK && hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess) s = MK_HIP_ERROR;
  if (s != MK_OK) {
    delete b;
    return s;
  }
  ctx->retain();
  *out = reinterpret_cast<mk_batch>(b);
  return MK_OK;
}

mk_status mk_batch_layout(mk_batch h, int64_t* offsets, int64_t* total, int* grouped) {
  MK_REQUIRE(h != nullptr, MK_BAD_ARG, "batch: null handle");
  Batch* b = reinterpret_cast<Batch*>(h);
  if (offsets)
    for (int i = 0; i < b->n; ++i) offsets[i] = b->off[i];
  if (total) *total = b->total;
  if (grouped) *grouped = b->grouped ? 1 : 0;
  return MK_OK;
}

mk_status mk_batch_apply(mk_batch h, const double* x, double* y, int precision) {
  MK_REQUIRE(h != nullptr && x != nullptr && y != nullptr, MK_BAD_ARG, "batch apply: null argument");
  MK_REQUIRE(precision >= MK_FP64 && precision <= MK_BF16X3, MK_BAD_ARG, "batch apply: precision %d", precision);
  Batch* b = reinterpret_cast<Batch*>(h);
  MK_CHECK_HIP(hipSetDevice(b->ctx->device));
  return b->apply(x, y, precision, b->ctx->stream);
}

mk_status mk_batch_free(mk_batch h) {
  if (!h) return MK_OK;
  Batch* b = reinterpret_cast<Batch*>(h);
  Ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
  ctx->release();
  return MK_OK;
}

mk_status mk_batch_eigsolve(mk_batch h, double* x, int krylovdim, int maxiter, double tol, int precision,
                            int use_graph, double* eigvals, double* resnorms, int* numops, int* converged) {
  MK_REQUIRE(h != nullptr && x != nullptr, MK_BAD_ARG, "batch eigsolve: null argument");
  MK_REQUIRE(krylovdim >= 2 && krylovdim <= 128, MK_BAD_ARG, "batch eigsolve: krylovdim %d not in [2, 128]",
             krylovdim);
  MK_REQUIRE(maxiter >= 1 && tol >= 0.0, MK_BAD_ARG, "batch eigsolve: bad maxiter / tol");
  MK_REQUIRE(precision >= MK_FP64 && precision <= MK_BF16X3, MK_BAD_ARG, "batch eigsolve: precision %d", precision);
  Batch* b = reinterpret_cast<Batch*>(h);
  Ctx* ctx = b->ctx;
  MK_CHECK_HIP(hipSetDevice(ctx->device));
  MK_TRY(b->ensure_plan(MK_FP64));
  MK_TRY(b->ensure_plan(precision));
  const int n = b->n, ks = krylovdim + 1;
  const int pmax = std::max(1, krylovdim / 2);   // Ritz vectors kept at a restart (x included)
  const int64_t N = b->total;
  const bool graphs = use_graph && b->grouped;
  // buffers (addresses are part of captured graphs: reallocation drops them)
  if (b->ks < ks || b->kq < pmax || !b->Q.ptr) {
    MK_TRY(b->wait_streams());
    b->drop_graphs();
    b->Q.release();
    b->coef.release();
    b->partial.release();
    MK_TRY(b->Q.alloc((size_t)N * (ks + pmax) * 8));
    b->kq = pmax;
    MK_TRY(b->coef.alloc((size_t)n * (3 * ks + 8) * 8));
    MK_TRY(b->partial.alloc((size_t)ks * b->sg.nchunks * 8));
    if (!b->W.ptr) MK_TRY(b->W.alloc((size_t)N * 8));
    if (!b->U.ptr) MK_TRY(b->U.alloc((size_t)N * 8));
    if (b->host) (void)hipHostFree(b->host);
    b->host = nullptr;
    b->host_doubles = (size_t)n * (3 * ks + 8);
    MK_CHECK_HIP(hipHostMalloc((void**)&b->host, b->host_doubles * 8, hipHostMallocDefault));
    b->ks = ks;
  }
  const int kst = b->ks;  // coefficient stride of the allocated buffers (>= ks)
  // grouped batches run on their own stream (graph capture needs a capturable
  // stream), ordered after the caller's; per-operator applies use the context's
  hipStream_t st = b->grouped ? b->stream : ctx->stream;
  if (st != ctx->stream) {
    MK_CHECK_HIP(hipEventRecord(b->ev, ctx->stream));
    MK_CHECK_HIP(hipStreamWaitEvent(st, b->ev, 0));
  }
  double* Qp = static_cast<double*>(b->Q.ptr);
  double* Wp = static_cast<double*>(b->W.ptr);
  double* Up = static_cast<double*>(b->U.ptr);
  double* c = static_cast<double*>(b->coef.ptr);
  double* part = static_cast<double*>(b->partial.ptr);
  double* hb = b->host;
  const Segs& sg = b->sg;
  auto col = [&](int q) { return Qp + (int64_t)q * N; };
  // device scratch scalars after the step region: a | bb (nseg each), ritz coefficients (nseg x kst)
  double* dv_a = c + 2 * (int64_t)n * kst + 2 * n;
  double* dv_b = dv_a + n;
  double* dv_s = dv_b + n;
  double* dv_t = dv_s + (int64_t)n * kst;  // scalar results (nseg)
  MK_CHECK_HIP(hipMemsetAsync(Qp, 0, (size_t)N * (kst + b->kq) * 8, st));
  MK_CHECK_HIP(hipMemsetAsync(Wp, 0, (size_t)N * 8, st));
  MK_CHECK_HIP(hipMemsetAsync(Up, 0, (size_t)N * 8, st));
  for (int i = 0; i < n; ++i)
    MK_CHECK_HIP(hipMemcpyAsync(col(0) + b->off[i], x + b->off[i], (size_t)b->len[i] * 8, hipMemcpyDeviceToDevice, st));

  auto readback = [&](const double* d, int cnt) -> mk_status {
    MK_CHECK_HIP(hipMemcpyAsync(hb, d, (size_t)cnt * 8, hipMemcpyDeviceToHost, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    return MK_OK;
  };
  auto upload = [&](double* d, const std::vector<double>& v) -> mk_status {
    std::memcpy(hb, v.data(), v.size() * 8);
    MK_CHECK_HIP(hipMemcpyAsync(d, hb, v.size() * 8, hipMemcpyHostToDevice, st));
    MK_CHECK_HIP(hipStreamSynchronize(st));
    return MK_OK;
  };

  // normalise the start vectors
  MK_TRY(seg_dots(sg, col(0), N, 1, col(0), dv_t, 1, part, st));
  MK_TRY(readback(dv_t, n));
  for (int i = 0; i < n; ++i) {
    MK_REQUIRE(std::isfinite(hb[i]), MK_NONFINITE, "batch eigsolve: start vector %d is not finite", i);
    MK_REQUIRE(hb[i] > 0.0, MK_BAD_ARG, "batch eigsolve: zero start vector %d", i);
  }
  MK_TRY(seg_normalize(sg, col(0), dv_t, 1, col(0), st));

  std::vector<double> theta(n, 0.0), res(n, INFINITY), est_last(n, INFINITY), thresh(n, 0.0), rperp(n, 0.0);
  std::vector<int> conv(n, 0), ops_n(n, 0), active(n, 0), kk(n, 0);
  const int ldg = ks;
  std::vector<std::vector<double>> G(n, std::vector<double>((size_t)ldg * ldg));
  std::vector<std::vector<double>> lamk(n, std::vector<double>(1, 0.0)), cx(n);
  const double eta = precision == MK_FP64 ? 0.0 : 1e-3;
  std::vector<double> ones(n, 1.0), zeros(n, 0.0), neg(n), sv, lamv, Sv;
  int p = 1;   // kept vectors [x, y_2 .. y_p] in columns 0 .. p-1 of every segment
  for (int outer = 0;; ++outer) {
    if (precision == MK_FP64 && outer > 0)
      for (int i = 0; i < n; ++i)
        if (!conv[i] && est_last[i] <= tol) {
          conv[i] = 1;
          res[i] = est_last[i];
        }
    bool all = true;
    for (int i = 0; i < n; ++i) all = all && conv[i];
    if (all) break;
    // float64 Ritz values and residuals; <q_r, H x> of the kept vectors
    MK_TRY(b->apply(col(0), Wp, MK_FP64, st));
    for (int i = 0; i < n; ++i)
      if (!conv[i]) ++ops_n[i];
    MK_TRY(seg_dots(sg, Qp, N, p, Wp, c, kst, part, st));
    MK_TRY(readback(c, n * kst));
    for (int i = 0; i < n; ++i) {
      cx[i].assign(hb + (int64_t)i * kst, hb + (int64_t)i * kst + p);
      if (!conv[i]) theta[i] = cx[i][0];
      neg[i] = -cx[i][0];
    }
    // residual r = W - theta x in column p, re-orthogonalised against x
    MK_CHECK_HIP(hipMemcpyAsync(col(p), Wp, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
    MK_TRY(upload(dv_a, neg));
    MK_TRY(upload(dv_b, ones));
    MK_TRY(seg_axpby(sg, dv_a, col(0), dv_b, col(p), st));
    MK_TRY(seg_dots(sg, col(0), N, 1, col(p), dv_t, 1, part, st));
    MK_TRY(seg_multiaxpy(sg, col(0), N, 1, dv_t, 1, col(p), st));
    MK_TRY(seg_dots(sg, col(p), N, 1, col(p), dv_t, 1, part, st));
    MK_TRY(readback(dv_t, n));
    bool any = false;
    for (int i = 0; i < n; ++i) {
      if (conv[i]) continue;
      const double r = std::sqrt(std::max(hb[i], 0.0));
      MK_REQUIRE(std::isfinite(theta[i]) && std::isfinite(r), MK_NONFINITE, "batch eigsolve: non-finite Ritz pair %d", i);
      res[i] = r;
      if (r <= tol) conv[i] = 1;
      else any = true;
    }
    if (!any || outer == maxiter) break;
    // the Krylov space starts from the part of r outside the kept Ritz vectors
    for (int i = 0; i < n; ++i) rperp[i] = res[i];
    if (p > 1) {
      for (int pass = 0; pass < 2; ++pass) {
        MK_TRY(seg_dots(sg, Qp, N, p, col(p), c, kst, part, st));
        MK_TRY(seg_multiaxpy(sg, Qp, N, p, c, kst, col(p), st));
      }
      MK_TRY(seg_dots(sg, col(p), N, 1, col(p), dv_t, 1, part, st));
      MK_TRY(readback(dv_t, n));
      bool lost = false;
      for (int i = 0; i < n; ++i) {
        if (conv[i]) continue;
        rperp[i] = std::sqrt(std::max(hb[i], 0.0));
        lost = lost || !(rperp[i] > 1e-14 * std::max(std::fabs(theta[i]), res[i]));
      }
      if (lost) {
        // some residual lies in the kept space: restart every instance from x alone
        p = 1;
        for (int i = 0; i < n; ++i) {
          rperp[i] = res[i];
          cx[i].resize(1);
        }
        MK_CHECK_HIP(hipMemcpyAsync(col(1), Wp, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
        MK_TRY(upload(dv_a, neg));
        MK_TRY(seg_axpby(sg, dv_a, col(0), dv_b, col(1), st));
        MK_TRY(seg_dots(sg, col(0), N, 1, col(1), dv_t, 1, part, st));
        MK_TRY(seg_multiaxpy(sg, col(0), N, 1, dv_t, 1, col(1), st));
      }
    }
    {
      std::vector<double> sq(n);
      for (int i = 0; i < n; ++i) sq[i] = conv[i] ? 1.0 : rperp[i] * rperp[i];
      MK_TRY(upload(dv_t, sq));
    }
    MK_TRY(seg_normalize(sg, col(p), dv_t, 1, col(p), st));
    for (int i = 0; i < n; ++i) {
      std::vector<double>& Gi = G[i];
      std::fill(Gi.begin(), Gi.end(), 0.0);
      Gi[0] = theta[i];
      for (int r = 1; r < p; ++r) {
        Gi[(size_t)r * ldg + r] = lamk[i][r];
        Gi[r] = Gi[(size_t)r * ldg] = cx[i][r];
      }
      Gi[p] = Gi[(size_t)p * ldg] = rperp[i];
      kk[i] = 1;
      active[i] = conv[i] ? 0 : 1;
      thresh[i] = std::max(tol, eta * res[i]);
    }
    // Krylov step j applies column p + j and writes column p + j + 1 (at most ks - 1)
    const int mmax = ks - p - 1;
    for (int j = 0; j < mmax; ++j) {
      const int jj = p + j - 1;   // krylov_step(jj) applies column jj + 1
      if (graphs) {
        const auto key = std::make_pair(precision, 2 * jj + (j == 0 ? 1 : 0));
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
          hipGraph_t g = nullptr;
          hipGraphExec_t ex = nullptr;
          MK_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
          const mk_status s = b->krylov_step(jj, precision, j == 0, st);
          const hipError_t e = hipStreamEndCapture(st, &g);
          MK_TRY(s);
          MK_CHECK_HIP(e);
          const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          MK_CHECK_HIP(ie);
          it = b->graphs.emplace(key, ex).first;
        }
        MK_CHECK_HIP(hipGraphLaunch(it->second, st));
      } else {
        MK_TRY(b->krylov_step(jj, precision, j == 0, st));
      }
      MK_CHECK_HIP(hipStreamSynchronize(st));
      const double* c1 = hb;
      const double* c2 = hb + (int64_t)n * kst;
      const double* nn = hb + 2 * (int64_t)n * kst;
      const double* xw = nn + n;
      const int cj = p + j;
      bool still = false;
      const int lo1 = first_pass_lo(jj, j == 0);
      for (int i = 0; i < n; ++i) {
        if (!active[i]) continue;
        ++ops_n[i];
        std::vector<double>& Gi = G[i];
        for (int r = 1; r <= cj; ++r)
          Gi[(size_t)r * ldg + cj] = Gi[(size_t)cj * ldg + r] =
              (r >= lo1 ? c1[(int64_t)i * kst + r] : 0.0) + c2[(int64_t)i * kst + r];
        const double beta = std::sqrt(std::max(nn[i], 0.0));
        MK_REQUIRE(std::isfinite(beta), MK_NONFINITE, "batch eigsolve: non-finite Krylov vector %d", i);
        double lam;
        lanczos_lowest(Gi, ldg, cj + 1, lam, sv);
        est_last[i] = beta * std::fabs(sv[cj]);
        kk[i] = j + 1;
        if (est_last[i] <= thresh[i] || beta <= 1e-14 * std::fabs(theta[i]) || j + 1 == mmax) {
          active[i] = 0;
          continue;
        }
        Gi[(size_t)(cj + 1) * ldg + cj] = Gi[(size_t)cj * ldg + cj + 1] = beta;
        Gi[(size_t)(cj + 1)] = Gi[(size_t)(cj + 1) * ldg] = xw[i];
        still = true;
      }
      if (!still) break;
    }
    // the lowest Ritz vectors of every instance's space: x and the kept vectors
    // of the next cycle (converged instances keep their columns)
    int pn = pmax, kmax = 1;
    for (int i = 0; i < n; ++i)
      if (!conv[i]) pn = std::min(pn, p + kk[i]);
    std::vector<std::vector<double>> Si(n);
    for (int i = 0; i < n; ++i) {
      if (conv[i]) continue;
      lanczos_lowest_p(G[i], ldg, p + kk[i], pn, lamv, Si[i]);
      lamk[i] = lamv;
      theta[i] = lamv[0];
      kmax = std::max(kmax, p + kk[i]);
    }
    for (int q = 0; q < pn; ++q) {
      std::vector<double> S((size_t)n * kst, 0.0);
      for (int i = 0; i < n; ++i) {
        if (conv[i]) {
          S[(size_t)i * kst + q] = 1.0;
          continue;
        }
        const int k = p + kk[i];
        for (int r = 0; r < k; ++r) S[(size_t)i * kst + r] = Si[i][(size_t)q * k + r];
      }
      MK_TRY(upload(dv_s, S));
      MK_TRY(seg_gemv(sg, Qp, N, std::max(kmax, pn), dv_s, kst, col(ks + q), st));
    }
    MK_CHECK_HIP(hipMemcpyAsync(col(0), col(ks), (size_t)N * pn * 8, hipMemcpyDeviceToDevice, st));
    MK_TRY(seg_dots(sg, col(0), N, 1, col(0), dv_t, 1, part, st));
    MK_TRY(seg_normalize(sg, col(0), dv_t, 1, col(0), st));
    p = pn;
  }
  for (int i = 0; i < n; ++i)
    MK_CHECK_HIP(hipMemcpyAsync(x + b->off[i], col(0) + b->off[i], (size_t)b->len[i] * 8, hipMemcpyDeviceToDevice, st));
  if (st != ctx->stream) {
    MK_CHECK_HIP(hipEventRecord(b->ev, st));
    MK_CHECK_HIP(hipStreamWaitEvent(ctx->stream, b->ev, 0));
  }
  MK_CHECK_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < n; ++i) {
    if (eigvals) eigvals[i] = theta[i];
    if (resnorms) resnorms[i] = res[i];
    if (numops) numops[i] = ops_n[i];
    if (converged) converged[i] = conv[i];
  }
  return MK_OK;
}

mk_status mk_autotune_stats(int* decisions, int* hits) {
  autotune_stats(decisions, hits);
  return MK_OK;
}

}  // extern "C"
