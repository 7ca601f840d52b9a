// Grouped GEMM launches planned once and replayed many times.
//
// A GroupPlan holds the device descriptor array of one grouped launch (any
// number of problems, each a sum of up to MAX_SEG products), its tile kind and
// its persistent grid.  Descriptor pointers may be relative to three base
// slots (rel_ptr), so a plan built for one operator serves every apply whatever
// the addresses of x, y and the conversion scratch; a captured hipGraph that
// launches the plan stays valid as long as the plan's device array does.
//
// Tile kinds come from the autotuner: the first time a launch signature
// (precision, transposition, problem count, flop and size buckets) is seen,
// every valid kind is timed on scratch buffers and the fastest is cached for
// the process.  MK_AUTOTUNE=0 selects the size heuristic instead.
#pragma once
#include <vector>

#include "ops.h"

namespace mk {

int f64_group_grid(int kind, int total_tiles, int cus);
int mixed_group_grid(int prec, int kind, int total_tiles, int cus);

struct GroupPlan {
  int prec = MK_FP64;
  GemmShape sh{};
  int kind = TK_64;
  int nprob = 0;
  int total_tiles = 0;
  int grid = 0;
  double flops = 0.0;
  bool f64_buf = false;         // FP64 problems fit the branch-free kernel (set by build)
  std::vector<ProbSpec> specs;  // problems in launch order (largest per-tile work first)
  DevBuf dev;                   // GemmProb array
  // fill and upload the descriptors for one tile kind; a rebuild first waits for
  // the stream the plan's earlier launches went to
  mk_status build(int kind, hipStream_t launched_on = nullptr);
  mk_status run(const void* const base[NBASE], hipStream_t st) const;
};

// Plans a grouped launch of `n` problems.  `base_bytes[s]` is the extent of
// relative slot s (used to allocate autotuning scratch); outputs C of every
// problem must be relative for the autotuner to run, otherwise the heuristic
// decides.  `kind` < 0 asks the autotuner.
mk_status group_plan(Ctx* ctx, const ProbSpec* probs, int n, int prec, GemmShape sh, const int64_t base_bytes[NBASE],
                     int kind, GroupPlan* out);

// Heuristic tile kind for a set of problems.
int heuristic_kind(const ProbSpec* probs, int n, int prec);

// Number of autotuner decisions taken and cache hits (for tests and reports).
void autotune_stats(int* decisions, int* hits);

}  // namespace mk
