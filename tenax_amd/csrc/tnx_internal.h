// Internal definitions shared by the libtnx translation units.
#pragma once
#include <cstring>
#include <hip/hip_runtime.h>

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
  int32_t pad;
};

// One output block: C (M x N, row-major, ld = ldc) = sum over its pairs.
struct GemmOut {
  int64_t c_off;
  int32_t M, N;
  int32_t ldc;
  int32_t pair_begin, pair_end;
  int32_t pad;
};

// One workgroup's output tile.
struct GemmTask {
  int32_t out;
  int32_t m0, n0;
  int32_t pad;
};

enum : int { LAY_MK = 0, LAY_KM = 1 };  // A: MK = a[m*lda+k], KM = a[k*lda+m]
enum : int { LAY_KN = 0 };              // B: KN = b[k*ldb+n] (the only B layout)

constexpr int GEMM_BM = 64;
constexpr int GEMM_BN = 64;
constexpr int GEMM_BK = 16;
constexpr int GEMM_THREADS = 256;

// Device-resident grouped-GEMM program: the work lists of one launch.
struct GemmProgram {
  int a_lay = LAY_MK, b_lay = LAY_KN;
  DevBuf<GemmPair> pairs;
  DevBuf<GemmOut> outs;
  DevBuf<GemmTask> tasks;
  int32_t n_pairs = 0, n_outs = 0, n_tasks = 0;
  // host mirrors (used by the per-block path and for stats)
  std::vector<GemmPair> h_pairs;
  std::vector<GemmOut> h_outs;
  double flops = 0.0;
};

// Launch the grouped GEMM: C[outs] (=|+=) sum_pairs A*B.
void launch_grouped_gemm(const GemmProgram &prog, const double *A, const double *B, double *C,
                         int compute, bool accumulate, hipStream_t s);
// Launch a prepared program slice (tasks/outs/pairs already on the device).
void launch_prepared(int a_lay, int b_lay, int compute, int n_tasks, hipStream_t s, const GemmTask *tasks,
                     const GemmOut *outs, const GemmPair *pairs, const double *A, const double *B, double *C,
                     bool accumulate);

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
enum class PlanKind { Pairwise, Transpose, Matvec, EnvLeft, EnvRight };

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
};

}  // namespace tnx

struct tnx_plan_s : public tnx::tnx_plan_base {};
