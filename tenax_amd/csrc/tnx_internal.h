// Internal definitions shared by the libtnx translation units.
#pragma once
#include <cstring>
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tnx.h"

namespace tnx {

struct Error {
  int code;
  std::string msg;
};

void set_last_error(const std::string &m);

#define TNX_CHECK_HIP(x)                                                              \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      throw ::tnx::Error{TNX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define TNX_REQUIRE(cond, code, msg)                              \
  do {                                                            \
    if (!(cond)) throw ::tnx::Error{(code), std::string(msg)};    \
  } while (0)

// Runs f() and converts exceptions into status codes + last-error text.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return TNX_OK;
  } catch (const Error &e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return TNX_ERR_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return TNX_ERR_ARG;
  }
}

// Owning device buffer (plan-internal; never used on an execute path for allocation).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      hipError_t e = hipMalloc(&p, count * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        n = 0;
        throw Error{TNX_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e)};
      }
    }
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count, hipStream_t s) {
    if (count > n) alloc(count);
    if (count) TNX_CHECK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *h, size_t count, hipStream_t s) const {
    if (count) TNX_CHECK_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// ---------------------------------------------------------------- GEMM work lists
// One K-segment: the product of one A block and one B block feeding an output block.
struct GemmPair {
  int64_t a_off;
  int64_t b_off;
  int32_t K;
  int32_t lda;  // leading dimension of the A block in its layout
  int32_t ldb;  // leading dimension of the B block in its layout
  int32_t neg;  // 1: the product enters with a minus sign (fermionic exchange sign)
};

// One output block: C (M x N, row-major, ld = ldc) = sum over its pairs.
struct GemmOut {
  int64_t c_off;
  int32_t M, N;
  int32_t ldc;
  int32_t pair_begin, pair_end;
  int32_t pad;
};

// One workgroup's output tile, or a K-range of it (split K).  The K range
// [kt0, kt1) counts BK-wide K tiles over the concatenation of the output's
// pairs.  slot < 0: the tile is written to C; slot >= 0: the partial tile goes
// to workspace slot `slot` and a reduction adds the slots in order.
// out < 0 marks an empty entry (padding of the XCD-interleaved order).
struct GemmTask {
  int32_t out;
  int32_t m0, n0;
  int32_t kt0, kt1;
  int32_t slot;
  int32_t pad0, pad1;
};

__host__ __device__ inline GemmTask whole_tile(int32_t out, int32_t m0, int32_t n0) {
  return GemmTask{out, m0, n0, 0, 0x7fffffff, -1, 0, 0};
}

// A tile whose K range was split: slots slot0 .. slot0+nsplit-1 hold its partials.
struct SplitTile {
  int32_t out, m0, n0;
  int32_t slot0, nsplit;
  int32_t pad;
};

// Tiling of a grouped GEMM launch.
//   shape: 0 = 64x64, 1 = 128x64, 2 = 64x128, 3 = 128x128 (rows x cols per workgroup)
//   kchunk: K tiles per task; 0 = no split, < 0 = chosen from the work per workgroup slot
//   xcd: order the tasks so that tiles of one output run on one XCD (shared L2)
// kchunk value selecting the balanced two-wave split-K schedule
constexpr int32_t KCHUNK_BALANCED = -2;
struct TileCfg {
  int32_t shape = 0;
  int32_t kchunk = 0;
  int32_t xcd = 0;
};

enum : int { LAY_MK = 0, LAY_KM = 1 };  // A: MK = a[m*lda+k], KM = a[k*lda+m]
enum : int { LAY_KN = 0 };              // B: KN = b[k*ldb+n] (the only B layout)

constexpr int GEMM_BM = 64;
constexpr int GEMM_BN = 64;
constexpr int GEMM_BK = 16;
constexpr int GEMM_THREADS = 256;

// Where an output block's elements go when the output legs are stored in
// another order than [A free legs..., B free legs...] (a contraction whose
// requested output order the GEMM writes directly): element (row, col) of the
// block's M x N product lands at c_off + rowoff(row) + coloff(col), where
// rowoff decomposes row over nr row legs (last fastest) and sums digit x
// stride, and likewise for columns.  nr == 0: the plain row-major store
// c_off + row * ldc + col.  Legs contiguous in both orders are merged, so
// three per side cover the output permutations of rank-6 contractions.
constexpr int STORE_LEGS = 3;
struct GemmStore {
  int32_t nr, nc;
  int32_t rdim[STORE_LEGS], cdim[STORE_LEGS];
  int64_t rstr[STORE_LEGS], cstr[STORE_LEGS];
};

// Device-resident grouped-GEMM program: the work lists of one launch.
struct GemmProgram {
  int a_lay = LAY_MK, b_lay = LAY_KN;
  bool signs = false;  // some pair carries a fermionic sign (set when finalized)
  bool permuted = false;  // some output uses a GemmStore with nr > 0 (stores[] is then one per output)
  DevBuf<GemmPair> pairs;
  DevBuf<GemmOut> outs;
  DevBuf<GemmStore> stores;
  DevBuf<GemmTask> tasks;
  int32_t n_pairs = 0, n_outs = 0, n_tasks = 0;
  // host mirrors (used by the per-block path, tiling and stats)
  std::vector<GemmPair> h_pairs;
  std::vector<GemmOut> h_outs;
  std::vector<GemmStore> h_stores;
  double flops = 0.0;
  // tiled variant (set_tiling): used by launch_grouped_gemm when `tiled`
  bool tiled = false;
  TileCfg cfg;
  DevBuf<GemmTask> vtasks;
  DevBuf<SplitTile> red;
  DevBuf<double> ws;
  int32_t n_vtasks = 0, n_red = 0;
};

// Number of XCDs workgroups are dispatched across (round robin by workgroup id).
constexpr int N_XCD = 8;
// Cap on the split-K workspace of one program (doubles).
constexpr int64_t MAX_WS_DOUBLES = int64_t(32) << 20;

// Build the task list of `cfg` for a finalized program (host-side, from the
// host mirrors) and make it the launched one.  The default 64x64 unsplit
// device-built list stays in `tasks` (the per-block path uses it).
void set_tiling(GemmProgram &prog, const TileCfg &cfg, hipStream_t s);

// Launch the grouped GEMM: C[outs] (=|+=) sum_pairs A*B.
// reduce = false skips the split-K reduction (profiling of the GEMM kernel alone).
// stop: optional device flag; when it reads non-zero the launch does nothing
// (the matvecs of a Lanczos solve after its breakdown step).
void launch_grouped_gemm(const GemmProgram &prog, const double *A, const double *B, double *C,
                         int compute, bool accumulate, hipStream_t s, bool reduce = true,
                         const int32_t *stop = nullptr);
// The split-K reduction of a tiled program alone (nothing when it has no split tiles).
void launch_split_reduce(const GemmProgram &prog, double *C, bool accumulate, hipStream_t s,
                         const int32_t *stop = nullptr);
// Launch a prepared program slice (tasks/outs/pairs already on the device).
void launch_prepared(int a_lay, int b_lay, bool signs, int compute, int n_tasks, hipStream_t s,
                     const GemmTask *tasks, const GemmOut *outs, const GemmPair *pairs, const double *A,
                     const double *B, double *C, bool accumulate);

// Device-side tiling of a program whose pairs/outs are already on the device:
// builds the tile list (longest K first) and downloads host mirrors.
void finalize_program(GemmProgram &prog, hipStream_t s);

// ---------------------------------------------------------------- sector enumeration
// Sorted key set on the device (keys/shapes: n x rank, lexicographic order).
struct KeySet {
  int32_t rank = 0;
  int64_t n = 0;
  DevBuf<int32_t> keys, shapes;
};
// all_tuples: every charge tuple (reachability candidates) instead of the
// conservation-filtered valid set.
void enumerate_keys(const tnx_leg *legs, int rank, int sym_n, bool all_tuples, bool has_target, int target,
                    hipStream_t s, KeySet &out);

// ---------------------------------------------------------------- plans
enum class PlanKind { Pairwise, Transpose, Matvec, EnvLeft, EnvRight, BatchPairwise, BatchTranspose };

struct tnx_plan_base {
  virtual ~tnx_plan_base() = default;
  PlanKind kind;
  int32_t out_rank = 0;
  int32_t out_nblocks = 0;
  int64_t out_total = 0;
  std::vector<int32_t> out_keys, out_shapes;
  std::vector<int64_t> out_offsets;
  double flops = 0.0, bytes = 0.0;
  int32_t n_pairs = 0, n_tasks = 0;
  // bumped whenever the device work lists are rebuilt (tnx_plan_set_tiling);
  // captured graphs refuse to replay work lists of an older generation
  uint64_t generation = 0;
  // captured graphs recorded on this plan and not yet freed; the plan
  // refuses to be freed while any remain (they launch its work lists)
  mutable std::atomic<int32_t> live_graphs{0};
};

}  // namespace tnx

struct tnx_plan_s : public tnx::tnx_plan_base {};

namespace tnx {
// The grouped-GEMM program of a fused (matvec / environment) plan's stage 1 or 3.
GemmProgram *fused_program(tnx_plan_s *plan, int stage);
}  // namespace tnx

namespace tnx {
// ---------------------------------------------------------------- work-list validation
// Block layout of one operand for validation: half-open element ranges
// [off, off + size) of its blocks (any order; zero-size blocks ignored).
struct Extents {
  std::vector<int64_t> lo, hi;  // sorted by lo after finish()
  int64_t total = 0;            // buffer size in elements
  void add(int64_t off, int64_t size);
  void finish();  // sorts, and checks that blocks do not overlap
  // true when [a, b) lies inside one block
  bool covers(int64_t a, int64_t b) const;
};
// Extents of a packed block table (shapes n x rank, offsets n).
Extents extents_of(const int32_t *shapes, const int64_t *offsets, int32_t n, int32_t rank);

// Checks a grouped-GEMM work list against its operands' block layouts: every
// output's C range inside one C block, every K-segment's A and B range inside
// one block of its operand (A as M x K rows of pitch lda for the MK layout,
// K x M for KM; B as K x N), row pitches at least the row length, pair ranges
// inside the pair list, and every task on an existing tile of its output with
// a non-empty K range and a workspace slot inside the workspace.  bm x bn is
// the task tile; ws_tiles the number of workspace slots (0 without split K).
// Throws Error{TNX_ERR_BAD_STRUCTURE, what + detail} on the first violation.
void check_worklist(int a_lay, const GemmPair *pairs, int64_t n_pairs, const GemmOut *outs, int64_t n_outs,
                    const GemmTask *tasks, int64_t n_tasks, int bm, int bn, int64_t ws_tiles, const Extents &A,
                    const Extents &B, const Extents &C, const std::string &what);
// check_worklist over a finalized program: its device-built default task list
// and, when a tiling is set, the tiled list.  Downloads what it needs.
void check_program(const GemmProgram &prog, const Extents &A, const Extents &B, const Extents &C,
                   const std::string &what, hipStream_t s);
// TNX_VALIDATE set (and not "0") in the environment: every plan is checked
// when it is built and when its tiling changes.
bool validation_enabled();
// Both GEMM programs and the mix lists of a matvec / environment plan
// against the block layouts of its operands and intermediates.
void validate_fused(const tnx_plan_s *plan, hipStream_t s);
}  // namespace tnx
