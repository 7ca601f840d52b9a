// Internal structures of contexts and prepared derivative operators.
#pragma once
#include <atomic>
#include <initializer_list>
#include <memory>
#include <utility>
#include <vector>

#include "common.h"
#include "gemm.h"

namespace mk {

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  ~DevBuf() { release(); }
  mk_status alloc(size_t nbytes);
  void release();
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // operators keep their context alive: mk_ctx_destroy drops the caller's
  // reference, the context is deleted when the last operator is freed
  std::atomic<int> refs{1};
  void retain() { refs.fetch_add(1); }
  void release() {
    if (refs.fetch_sub(1) == 1) delete this;
  }
  DevBuf arena;      // stream-ordered scratch for applies / transfers
  DevBuf smallbuf;   // small stream-ordered scratch (reduction partials, scalars)
  mk_status scratch(size_t nbytes, void** out);
  mk_status small(size_t nbytes, void** out);
};

mk_status carve(Ctx* ctx, std::initializer_list<std::pair<size_t, void**>> regions);

// A fixed scratch region for the calling host thread.  While one is installed
// (FixedScratchGuard), carve() takes its regions from it instead of the
// context arena and never allocates: a request larger than the region fails
// with MK_BAD_SHAPE.  Ctx::small() refuses to run while one is installed.
struct FixedScratch {
  char* base = nullptr;
  size_t bytes = 0;
};
FixedScratch& fixed_scratch();
struct FixedScratchGuard {
  explicit FixedScratchGuard(void* base, size_t bytes) {
    fixed_scratch() = FixedScratch{static_cast<char*>(base), bytes};
  }
  ~FixedScratchGuard() { fixed_scratch() = FixedScratch{}; }
  FixedScratchGuard(const FixedScratchGuard&) = delete;
  FixedScratchGuard& operator=(const FixedScratchGuard&) = delete;
};
// Resolves the device address of the allocation-free entry point's scratch
// array on `device` (which must be current) once per device; noalloc.hip.
mk_status noalloc_setup(int device);
inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }
int64_t pad8(int64_t k);

// An operator matrix used as GEMM operand: role 'A' (rows = M, cols = K) or
// 'B' (rows = K, cols = N).  float64 column-major, plus optional mixed copies.
struct OpMat {
  char role = 'A';
  int64_t rows = 0, cols = 0;
  DevBuf own;
  double* fp = nullptr;
  DevBuf hi, lo, f32;
  int64_t kpad = 0;
  double* f64() const { return fp; }
  bool present() const { return fp != nullptr; }
  mk_status alloc(char r, int64_t nr, int64_t nc) {
    role = r; rows = nr; cols = nc;
    MK_TRY(own.alloc((size_t)nr * nc * 8));
    fp = static_cast<double*>(own.ptr);
    return MK_OK;
  }
  void view(char r, double* p, int64_t nr, int64_t nc) { role = r; rows = nr; cols = nc; fp = p; }
  mk_status make_mixed(int precisions, hipStream_t st);
  // copies of the mixed operands in the packed panel layout of the 256 x 256
  // mixed kernel (packed_offset), besides the K-contiguous rows above
  DevBuf phi, plo, pf32;
  bool packed = false;
  mk_status make_packed(int precisions, hipStream_t st);
};

struct MixedView {
  void* hi = nullptr;
  void* lo = nullptr;
  int64_t kpad = 0;
};

mk_status convert_view(Ctx* ctx, const double* M, int64_t rows, int64_t cols, int64_t ld, char role,
                       int prec, MixedView* out, void* mem);
// Where convert_view puts the converted (rows x cols) view in `mem`, without converting.
MixedView view_layout(int64_t rows, int64_t cols, char role, int prec, void* mem);
size_t view_bytes(int64_t rows, int64_t cols, char role, int prec);
mk_status run_gemm(const ProbSpec* p, int np, int prec, GemmShape sh, hipStream_t st);

enum OpKind { OP_C = 0, OP_AC_GEN = 1, OP_AC2_GEN = 2, OP_AC_JOR = 3, OP_AC2_JOR = 4, OP_TL = 5, OP_TR = 6 };

// Two-step "L (x R)" block: xR = x (Mx x Ki) * R (Ki x nk*Ny);  y = L (Mo x Mx*nk) * xR.
struct LxR {
  OpMat L, R;
  int64_t Mx = 0, Ki = 0, nk = 0, Ny = 0, Mo = 0;
  bool present() const { return L.present(); }
  double flops() const { return 2.0 * Mx * Ki * nk * Ny + 2.0 * Mo * Mx * nk * Ny; }
};

struct Op {
  Ctx* ctx = nullptr;
  OpKind kind = OP_C;
  int precisions = MK_PREC_FP64;
  int64_t Dl = 0, Dr = 0, d1 = 0, d2 = 0;  // Jordan dims (bra == ket)
  int64_t x_elems = 0, y_elems = 0;
  double flops = 0.0;

  // generic operators and the A/AA parts of Jordan operators
  LxR gen;

  // Jordan two-site terms
  OpMat IC, BE, CA;
  DevBuf ABall;
  std::vector<OpMat> AB;  // (e, j) blocks at index e + d2*j
  DevBuf P;               // two-site (or one-site) physical term, if not folded
  bool hasP = false;
  // the fused nearest-neighbour BF16x3 apply runs on the 256 x 256 kernel with
  // packed operands (IC, BE packed at prepare; the views of x packed per apply)
  bool nn_packed = false;

  // Jordan one-site terms
  OpMat Cp, Bp;

  // transfer matrices (OP_TL, OP_TR): the ket tensor TA and the bra tensor TB
  // as GEMM operands, the MPO tensor (none for an MPS transfer) and the sizes
  //   left:  x (Dm, wl, Dn) -> y (Da, wr, Dc), A (Dn, d, Dc), Abar (Dm, d, Da)
  //   right: v (Di, wr, Do) -> y (Da, wl, Dc), A (Da, d, Di), Abar (Dc, d, Do)
  OpMat TA, TB;
  DevBuf TO;
  bool tmpo = false;
  int64_t tD[4] = {0, 0, 0, 0};     // left: Dm, Dn, Dc, Da;  right: Da, Di, Dc, Do
  int tw[3] = {1, 1, 1};            // wl, d, wr

  // U(1) charge blocks of a nearest-neighbour two-site operator: when set,
  // applies only compute the output blocks that conserve charge, each a sum of
  // x[r, k] IC[k, c] ('I') and BE[r, k] x[k, c] ('E') products over the
  // charge-compatible ranges k of the dense charge-sorted layout
  struct BlockSeg {
    char kind;
    int64_t k0, k;
  };
  struct BlockProb {
    int64_t r0, m, c0, n;
    int nseg = 0;
    BlockSeg seg[MAX_SEG];
  };
  std::vector<BlockProb> blocks;
  bool blocks_aligned = false;       // every range a multiple of 8 (mixed precisions allowed)
  std::shared_ptr<void> bplan[3];    // GroupPlan per precision, built on first use
  // compact storage of the charge blocks: x and y hold only the blocks, block b
  // column-major (m x n) at element offset boff[b]; the grouped items of the
  // physical term (device) and of the mixed views
  bool compact = false;
  std::vector<int64_t> boff;
  DevBuf cphys;
  int ncphys = 0, cphys_max = 0;
  DevBuf csplit[3];
  int ncsplit[3] = {0, 0, 0}, csplit_max[3] = {0, 0, 0};
  int64_t cview[3] = {0, 0, 0};      // bytes of the mixed views per precision
};

// Grouped-GEMM problems of an operator's charge blocks: x at slot 1 + xoff, y at
// slot 2 + xoff (bytes); mixed views of x at slot 3 + sa (left-factor view, row
// pitch pad8(N)) and slot 3 + sb (right-factor view, pitch pad8(M)).
void block_probs(const Op* op, int prec, int64_t xoff, int64_t sa, int64_t sb, double beta,
                 std::vector<ProbSpec>& out);

mk_status op_apply(Op* op, const double* x, double* y, double alpha, double beta, int prec);

}  // namespace mk
