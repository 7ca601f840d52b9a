// Grouped GEMM planning, descriptor upload and the tile-kind autotuner.
#include "group.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

namespace mk {

bool tile_valid(int prec, int kind) {
  if (prec == MK_FP64) return kind == TK_64 || kind == TK_128;
  return kind >= TK_64 && kind <= TK_256;
}

int group_grid(int prec, int kind, GemmShape sh, int total_tiles) {
  (void)sh;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
  }
  if (total_tiles <= 0) return 0;
  return prec == MK_FP64 ? f64_group_grid(kind, total_tiles, cus) : mixed_group_grid(prec, kind, total_tiles, cus);
}

namespace {

uint64_t low_bits(const void* p) { return reinterpret_cast<uint64_t>(p) & REL_MASK; }

int64_t ktot(const ProbSpec& p) {
  int64_t k = 0;
  for (int s = 0; s < p.nseg; ++s) k += std::max<int64_t>(p.seg[s].K, 0);
  return k;
}

mk_status fill(const std::vector<ProbSpec>& specs, int prec, int kind, std::vector<GemmProb>& out, int& total) {
  const int BM = tile_m(kind), BN = tile_n(kind);
  out.assign(specs.size(), GemmProb{});
  total = 0;
  for (size_t i = 0; i < specs.size(); ++i) {
    const ProbSpec& ps = specs[i];
    GemmProb& P = out[i];
    MK_REQUIRE(ps.M >= 0 && ps.N >= 0 && ps.M < (1ll << 31) && ps.N < (1ll << 31), MK_BAD_SHAPE,
               "grouped gemm: bad M/N");
    MK_REQUIRE(ps.nseg <= MAX_SEG, MK_BAD_ARG, "grouped gemm: too many segments");
    P.M = (int32_t)ps.M;
    P.N = (int32_t)ps.N;
    P.nseg = 0;
    for (int s = 0; s < ps.nseg; ++s) {
      const SegSpec& q = ps.seg[s];
      if (q.K <= 0) continue;
      MK_REQUIRE(q.K < (1ll << 31), MK_BAD_SHAPE, "grouped gemm: bad K");
      GemmSeg& g = P.seg[P.nseg++];
      g.A = q.A;
      g.B = q.B;
      g.lda = q.lda;
      g.ldb = q.ldb;
      if (prec == MK_FP64) {
        g.Alo = g.Blo = nullptr;
        g.K = (int32_t)q.K;
        g.vec = (low_bits(q.A) % 16 == 0 && low_bits(q.B) % 16 == 0 && q.lda % 2 == 0 && q.ldb % 2 == 0) ? 1 : 0;
      } else {
        MK_REQUIRE(q.lda % 8 == 0 && q.ldb % 8 == 0 && q.K <= q.lda && q.K <= q.ldb, MK_BAD_ARG,
                   "grouped mixed gemm: K-contiguous rows must be padded to 8");
        const void* ptrs[4] = {q.A, q.B, q.Alo, q.Blo};
        for (int z = 0; z < (prec == MK_BF16X3 ? 4 : 2); ++z)
          MK_REQUIRE(ptrs[z] != nullptr && low_bits(ptrs[z]) % 16 == 0, MK_BAD_ARG,
                     "grouped mixed gemm: operands must be non-null and 16-byte aligned");
        g.Alo = q.Alo;
        g.Blo = q.Blo;
        g.K = (int32_t)((q.K + 7) / 8 * 8);
        g.vec = 1;
      }
    }
    P.tiles_m = (int32_t)cdiv(ps.M, BM);
    P.tiles_n = (int32_t)cdiv(ps.N, BN);
    P.tile_start = total;
    total += P.tiles_m * P.tiles_n;
    P.C = ps.C;
    P.ldc = ps.ldc;
    P.alpha = ps.alpha;
    P.beta = ps.beta;
  }
  return MK_OK;
}

// ---- autotuner cache -------------------------------------------------------

using Key = std::tuple<int, int, int, int, int, int>;
std::mutex g_mu;
std::map<Key, int> g_cache;
int g_decisions = 0, g_hits = 0;

int ilog2(double v) { return v <= 1.0 ? 0 : (int)std::floor(std::log2(v)); }

Key signature(const std::vector<ProbSpec>& specs, int prec, GemmShape sh, double flops) {
  int64_t mn = 0;
  for (const ProbSpec& p : specs) mn = std::max<int64_t>(mn, p.M * p.N);
  return Key{prec, (sh.ta ? 1 : 0) | (sh.tb ? 2 : 0), ilog2((double)specs.size()), ilog2(flops), ilog2((double)mn), 0};
}

bool autotune_enabled() {
  static const bool on = [] {
    const char* e = getenv("MK_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  return on;
}

}  // namespace

int heuristic_kind(const ProbSpec* probs, int n, int prec) {
  int64_t t128 = 0, t256 = 0;
  for (int i = 0; i < n; ++i) {
    t128 += cdiv(probs[i].M, 128) * cdiv(probs[i].N, 128);
    t256 += cdiv(probs[i].M, 256) * cdiv(probs[i].N, 256);
  }
  if (prec != MK_FP64 && t256 >= 256) return TK_256;
  return t128 >= 256 ? TK_128 : TK_64;
}

void autotune_stats(int* decisions, int* hits) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (decisions) *decisions = g_decisions;
  if (hits) *hits = g_hits;
}

mk_status GroupPlan::build(int k, hipStream_t launched_on) {
  MK_REQUIRE(tile_valid(prec, k), MK_BAD_ARG, "grouped gemm: tile kind %d invalid for precision %d", k, prec);
  std::vector<GemmProb> host;
  int total = 0;
  MK_TRY(fill(specs, prec, k, host, total));
  const size_t bytes = std::max<size_t>(host.size(), 1) * sizeof(GemmProb);
  // launches of this plan (all on `launched_on`) may still be reading the
  // array: wait for that stream only before replacing or rewriting it
  if (dev.ptr && launched_on) MK_CHECK_HIP(hipStreamSynchronize(launched_on));
  if (dev.bytes < bytes) {
    dev.release();
    MK_TRY(dev.alloc(bytes));
  }
  if (!host.empty()) MK_CHECK_HIP(hipMemcpy(dev.ptr, host.data(), host.size() * sizeof(GemmProb), hipMemcpyHostToDevice));
  kind = k;
  nprob = (int)host.size();
  total_tiles = total;
  // relative pointers keep their offsets in the low bits and the base slots
  // are 16-byte aligned, so the alignment test reads the offsets
  f64_buf = prec == MK_FP64 && (k == TK_64 || k == TK_128) && f64_buf_ok(specs.data(), (int)specs.size(), sh, k);
  grid = group_grid(prec, k, sh, total);
  return MK_OK;
}

namespace {

const void* absolute(const void* p, const void* const base[NBASE]) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const int slot = (int)(v >> 62);
  if (slot == 0 || p == nullptr) return p;
  return static_cast<const char*>(base[slot]) + (v & REL_MASK);
}

// FP64 problems that each fill the GPU on their own (>= 256 tiles of 128 x 128)
// run as plain multi-problem launches: the per-tile kernel measured 20-25%
// faster than its persistent grouped instantiation at these sizes.
bool large_fp64(const std::vector<ProbSpec>& specs) {
  for (const ProbSpec& p : specs)
    if (cdiv(p.M, 128) * cdiv(p.N, 128) < 256) return false;
  return !specs.empty();
}

}  // namespace

mk_status GroupPlan::run(const void* const base[NBASE], hipStream_t st) const {
  if (total_tiles == 0) return MK_OK;
  if (prec == MK_FP64 && large_fp64(specs)) {
    for (size_t i0 = 0; i0 < specs.size(); i0 += MAX_PROB) {
      const int np = (int)std::min<size_t>(MAX_PROB, specs.size() - i0);
      ProbSpec ps[MAX_PROB];
      for (int i = 0; i < np; ++i) {
        ps[i] = specs[i0 + i];
        ps[i].C = static_cast<double*>(const_cast<void*>(absolute(ps[i].C, base)));
        for (int s = 0; s < ps[i].nseg; ++s) {
          ps[i].seg[s].A = absolute(ps[i].seg[s].A, base);
          ps[i].seg[s].B = absolute(ps[i].seg[s].B, base);
        }
      }
      MK_TRY(gemm_f64(ps, np, sh, st));
    }
    return MK_OK;
  }
  GemmGroup g{};
  g.probs = static_cast<const GemmProb*>(dev.ptr);
  g.nprob = nprob;
  g.total_tiles = total_tiles;
  const int64_t eb = prec == MK_FP32 ? 4 : 2;
  g.f64_buf = f64_buf ? 1 : 0;
  g.rows256_fit = 1;
  for (const ProbSpec& p : specs)
    for (int s = 0; s < p.nseg; ++s)
      if (256 * std::max(p.seg[s].lda, p.seg[s].ldb) * eb >= (1ll << 31)) g.rows256_fit = 0;
  for (int s = 0; s < NBASE; ++s) {
    g.base[s] = static_cast<const char*>(base ? base[s] : nullptr);
    MK_REQUIRE(reinterpret_cast<uintptr_t>(g.base[s]) % 16 == 0, MK_BAD_ARG,
               "grouped gemm: base slot %d is not 16-byte aligned", s);
  }
  if (prec == MK_FP64) return gemm_f64_group(g, kind, sh, grid, st);
  return gemm_mixed_group(g, prec, kind, grid, st);
}

mk_status group_plan(Ctx* ctx, const ProbSpec* probs, int n, int prec, GemmShape sh, const int64_t base_bytes[NBASE],
                     int kind, GroupPlan* out) {
  MK_REQUIRE(n >= 0, MK_BAD_ARG, "grouped gemm: negative problem count");
  MK_REQUIRE(prec == MK_FP64 || !(sh.ta || sh.tb), MK_BAD_ARG, "grouped mixed gemm: operands are pre-converted");
  out->prec = prec;
  out->sh = sh;
  out->specs.assign(probs, probs + n);
  // largest per-tile work first: persistent workgroups take tiles in order
  std::stable_sort(out->specs.begin(), out->specs.end(),
                   [](const ProbSpec& a, const ProbSpec& b) { return ktot(a) > ktot(b); });
  out->flops = 0.0;
  bool rel_out = true;
  for (const ProbSpec& p : out->specs) {
    out->flops += 2.0 * (double)p.M * (double)p.N * (double)ktot(p);
    rel_out = rel_out && (reinterpret_cast<uint64_t>(p.C) >> 62) != 0;
  }
  if (kind >= 0) return out->build(kind);
  const int heur = heuristic_kind(out->specs.data(), n, prec);
  if (!autotune_enabled() || !rel_out || n == 0) return out->build(heur);

  const Key key = signature(out->specs, prec, sh, out->flops);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      ++g_hits;
      return out->build(it->second);
    }
  }
  // time every valid kind on zeroed scratch copies of the relative slots
  DevBuf scratch[NBASE];
  const void* base[NBASE] = {nullptr, nullptr, nullptr, nullptr};
  hipStream_t st = ctx->stream;
  for (int s = 1; s < NBASE; ++s) {
    if (base_bytes && base_bytes[s] > 0) {
      MK_TRY(scratch[s].alloc((size_t)base_bytes[s]));
      MK_CHECK_HIP(hipMemsetAsync(scratch[s].ptr, 0, (size_t)base_bytes[s], st));
      base[s] = scratch[s].ptr;
    }
  }
  hipEvent_t e0, e1;
  MK_CHECK_HIP(hipEventCreate(&e0));
  MK_CHECK_HIP(hipEventCreate(&e1));
  int best = heur;
  float best_ms = 1e30f;
  for (int k = TK_64; k <= TK_256; ++k) {
    if (!tile_valid(prec, k)) continue;
    mk_status s = out->build(k, st);
    if (s == MK_OK) s = out->run(base, st);  // warm-up
    if (s == MK_OK && hipEventRecord(e0, st) != hipSuccess) s = MK_HIP_ERROR;
    for (int r = 0; r < 3 && s == MK_OK; ++r) s = out->run(base, st);
    if (s == MK_OK && hipEventRecord(e1, st) != hipSuccess) s = MK_HIP_ERROR;
    if (s == MK_OK && hipEventSynchronize(e1) != hipSuccess) s = MK_HIP_ERROR;
    float ms = 0.f;
    if (s == MK_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best_ms) {
      best_ms = ms;
      best = k;
    }
    if (s != MK_OK) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return s;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
    ++g_decisions;
  }
  return out->build(best, st);
}

}  // namespace mk
