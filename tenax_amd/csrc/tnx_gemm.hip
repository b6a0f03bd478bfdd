// Grouped block GEMM for gfx950: uneven output blocks, each the sum of the
// K-segments of all block pairs that feed it, one 64x64 output tile per
// workgroup, LDS-staged operand tiles, MFMA 16x16x4 (FP64 or FP32).
//
// Replaces the per-combination einsum of the reference
// (src/tenax/contraction/contractor.py:444-476, src/tenax/algorithms/dmrg.py:1097-1109):
// instead of one dispatch per block pair plus one add per accumulation, every
// pair of a contraction runs in one launch and each output tile accumulates its
// pairs in a fixed order (deterministic, no atomics).
#include "tnx_internal.h"

namespace tnx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// LDS tile row pitch (elements): +1 breaks the power-of-two stride between k rows.
constexpr int APITCH = GEMM_BM + 1;
constexpr int BPITCH = GEMM_BN + 1;

// Per-thread staging of one A tile (64 x 16) and one B tile (16 x 64): 4 values each.
struct Stage {
  double a[4];
  double b[4];
};

template <int ALAY>
__device__ __forceinline__ void load_a(const double *__restrict__ A, const GemmPair &p, int M, int m0,
                                       int k0, int tid, double (&r)[4]) {
  if (ALAY == LAY_MK) {
    const int m = m0 + (tid >> 2);
    const int kq = k0 + ((tid & 3) << 2);
    const double *src = A + p.a_off + (int64_t)m * p.lda;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (m < M && kq + i < p.K) ? src[kq + i] : 0.0;
  } else {
    const int k = k0 + (tid >> 4);
    const int mq = m0 + ((tid & 15) << 2);
    const double *src = A + p.a_off + (int64_t)k * p.lda;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (k < p.K && mq + i < M) ? src[mq + i] : 0.0;
  }
}

// B operand is always K x N row-major (KN): b[k*ldb + n].
__device__ __forceinline__ void load_b(const double *__restrict__ B, const GemmPair &p, int N, int n0, int k0,
                                       int tid, double (&r)[4]) {
  const int k = k0 + (tid >> 4);
  const int nq = n0 + ((tid & 15) << 2);
  const double *src = B + p.b_off + (int64_t)k * p.ldb;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = (k < p.K && nq + i < N) ? src[nq + i] : 0.0;
}

// Store staged registers into LDS tiles laid out [k][m] and [k][n].
template <int ALAY, class T>
__device__ __forceinline__ void store_tiles(T *As, T *Bs, const Stage &st, int tid) {
  if (ALAY == LAY_MK) {
    const int m = tid >> 2, kq = (tid & 3) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) As[(kq + i) * APITCH + m] = (T)st.a[i];
  } else {
    const int k = tid >> 4, mq = (tid & 15) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) As[k * APITCH + mq + i] = (T)st.a[i];
  }
  {
    const int k = tid >> 4, nq = (tid & 15) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) Bs[k * BPITCH + nq + i] = (T)st.b[i];
  }
}

// Split-FP32 staging: hi = fp32(x), lo = fp32(x - hi).
template <int ALAY>
__device__ __forceinline__ void store_tiles_split(float *Ah, float *Al, float *Bh, float *Bl,
                                                  const Stage &st, int tid) {
  float ah[4], al[4], bh[4], bl[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ah[i] = (float)st.a[i];
    al[i] = (float)(st.a[i] - (double)ah[i]);
    bh[i] = (float)st.b[i];
    bl[i] = (float)(st.b[i] - (double)bh[i]);
  }
  if (ALAY == LAY_MK) {
    const int m = tid >> 2, kq = (tid & 3) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Ah[(kq + i) * APITCH + m] = ah[i];
      Al[(kq + i) * APITCH + m] = al[i];
    }
  } else {
    const int k = tid >> 4, mq = (tid & 15) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Ah[k * APITCH + mq + i] = ah[i];
      Al[k * APITCH + mq + i] = al[i];
    }
  }
  {
    const int k = tid >> 4, nq = (tid & 15) << 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Bh[k * BPITCH + nq + i] = bh[i];
      Bl[k * BPITCH + nq + i] = bl[i];
    }
  }
}

// Write one 16x16 MFMA accumulator.  Operand lanes are row/col + 16*k for both
// instructions; the accumulator layouts differ (measured on gfx950):
//   v_mfma_f32_16x16x4f32: lane l, reg r -> row 4*(l/16) + r, col l%16
//   v_mfma_f64_16x16x4f64: lane l, reg r -> row (l/16) + 4*r, col l%16
template <bool F64LAYOUT, class V>
__device__ __forceinline__ void store_acc(double *__restrict__ C, const GemmOut &o, int row0, int col0,
                                          int lane, const V &acc, bool accumulate) {
  const int col = col0 + (lane & 15);
  if (col >= o.N) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + (F64LAYOUT ? ((lane >> 4) + 4 * r) : (((lane >> 4) << 2) + r));
    if (row < o.M) {
      double *dst = C + o.c_off + (int64_t)row * o.ldc + col;
      const double v = (double)acc[r];
      *dst = accumulate ? *dst + v : v;
    }
  }
}

// COMPUTE: 0 = FP64 MFMA, 1 = FP32 MFMA, 2 = split FP32 (3 products, FP64 across K chunks)
template <int ALAY, int COMPUTE>
__global__ __launch_bounds__(GEMM_THREADS) void grouped_gemm_kernel(
    const GemmTask *__restrict__ tasks, const GemmOut *__restrict__ outs,
    const GemmPair *__restrict__ pairs, const double *__restrict__ A, const double *__restrict__ B,
    double *__restrict__ C, int accumulate) {
  constexpr int NT = (COMPUTE == 2) ? 2 : 1;  // hi/lo planes for split
  using T = typename std::conditional<COMPUTE == 0, double, float>::type;
  // Two LDS stages: tile t is read from stage t&1 while tile t+1 is written
  // to the other stage, so each K-tile needs a single barrier.
  __shared__ T As[2][NT][GEMM_BK * APITCH];
  __shared__ T Bs[2][NT][GEMM_BK * BPITCH];

  const GemmTask t = tasks[blockIdx.x];
  const GemmOut o = outs[t.out];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  const int lrow = lane & 15;
  const int lk = lane >> 4;

  d4 dacc[2][2];
  f4 facc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      dacc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      facc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }

  auto stage_store = [&](int buf, const Stage &st) {
    if (COMPUTE == 2)
      store_tiles_split<ALAY>((float *)As[buf][0], (float *)As[buf][NT - 1], (float *)Bs[buf][0],
                              (float *)Bs[buf][NT - 1], st, tid);
    else
      store_tiles<ALAY, T>(As[buf][0], Bs[buf][0], st, tid);
  };

  auto compute_tile = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < GEMM_BK; kk += 4) {
      const int kr = kk + lk;
      if (COMPUTE == 0) {
        const double *as = (const double *)As[buf][0];
        const double *bs = (const double *)Bs[buf][0];
        double a0 = as[kr * APITCH + wm + lrow];
        double a1 = as[kr * APITCH + wm + 16 + lrow];
        double b0 = bs[kr * BPITCH + wn + lrow];
        double b1 = bs[kr * BPITCH + wn + 16 + lrow];
        dacc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, dacc[0][0], 0, 0, 0);
        dacc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, dacc[0][1], 0, 0, 0);
        dacc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, dacc[1][0], 0, 0, 0);
        dacc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, dacc[1][1], 0, 0, 0);
      } else if (COMPUTE == 1) {
        const float *as = (const float *)As[buf][0];
        const float *bs = (const float *)Bs[buf][0];
        float a0 = as[kr * APITCH + wm + lrow];
        float a1 = as[kr * APITCH + wm + 16 + lrow];
        float b0 = bs[kr * BPITCH + wn + lrow];
        float b1 = bs[kr * BPITCH + wn + 16 + lrow];
        facc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, facc[0][0], 0, 0, 0);
        facc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, facc[0][1], 0, 0, 0);
        facc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, facc[1][0], 0, 0, 0);
        facc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, facc[1][1], 0, 0, 0);
      } else {
        const float *ah = (const float *)As[buf][0];
        const float *al = (const float *)As[buf][NT - 1];
        const float *bh = (const float *)Bs[buf][0];
        const float *bl = (const float *)Bs[buf][NT - 1];
        float a0h = ah[kr * APITCH + wm + lrow], a0l = al[kr * APITCH + wm + lrow];
        float a1h = ah[kr * APITCH + wm + 16 + lrow], a1l = al[kr * APITCH + wm + 16 + lrow];
        float b0h = bh[kr * BPITCH + wn + lrow], b0l = bl[kr * BPITCH + wn + lrow];
        float b1h = bh[kr * BPITCH + wn + 16 + lrow], b1l = bl[kr * BPITCH + wn + 16 + lrow];
#define TNX_SPLIT3(ACC, AH, AL, BH, BL)                             \
  ACC = __builtin_amdgcn_mfma_f32_16x16x4f32(AL, BH, ACC, 0, 0, 0); \
  ACC = __builtin_amdgcn_mfma_f32_16x16x4f32(AH, BL, ACC, 0, 0, 0); \
  ACC = __builtin_amdgcn_mfma_f32_16x16x4f32(AH, BH, ACC, 0, 0, 0);
        TNX_SPLIT3(facc[0][0], a0h, a0l, b0h, b0l)
        TNX_SPLIT3(facc[0][1], a0h, a0l, b1h, b1l)
        TNX_SPLIT3(facc[1][0], a1h, a1l, b0h, b0l)
        TNX_SPLIT3(facc[1][1], a1h, a1l, b1h, b1l)
#undef TNX_SPLIT3
      }
    }
    if (COMPUTE == 2) {
      // flush the FP32 chunk sums into FP64 accumulators
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) dacc[i][j][r] += (double)facc[i][j][r];
          facc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  };

  // Flattened iteration over (pair, k-tile) with a one-tile register prefetch.
  int pi = o.pair_begin;
  if (pi < o.pair_end) {
    int k0 = 0;
    GemmPair p = pairs[pi];
    Stage st;
    load_a<ALAY>(A, p, o.M, t.m0, k0, tid, st.a);
    load_b(B, p, o.N, t.n0, k0, tid, st.b);
    stage_store(0, st);
    __syncthreads();
    int buf = 0;
    while (true) {
      int nk0 = k0 + GEMM_BK;
      int npi = pi;
      GemmPair np = p;
      if (nk0 >= p.K) {
        npi = pi + 1;
        nk0 = 0;
        if (npi < o.pair_end) np = pairs[npi];
      }
      const bool more = npi < o.pair_end;
      if (more) {
        load_a<ALAY>(A, np, o.M, t.m0, nk0, tid, st.a);
        load_b(B, np, o.N, t.n0, nk0, tid, st.b);
      }
      compute_tile(buf);
      if (!more) break;
      stage_store(buf ^ 1, st);
      __syncthreads();
      buf ^= 1;
      pi = npi;
      k0 = nk0;
      p = np;
    }
  }
  const bool acc_flag = accumulate != 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int r0 = t.m0 + wm + i * 16;
      const int c0 = t.n0 + wn + j * 16;
      if (COMPUTE == 1)
        store_acc<false>(C, o, r0, c0, lane, facc[i][j], acc_flag);
      else
        store_acc<COMPUTE == 0>(C, o, r0, c0, lane, dacc[i][j], acc_flag);
    }
}

template <int ALAY>
static void launch_lay(int compute, dim3 grid, hipStream_t s, const GemmTask *tasks, const GemmOut *outs,
                       const GemmPair *pairs, const double *A, const double *B, double *C, int acc) {
  switch (compute) {
    case TNX_COMPUTE_F64:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 0>), grid, dim3(GEMM_THREADS), 0, s, tasks, outs, pairs, A, B,
                         C, acc);
      break;
    case TNX_COMPUTE_F32:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 1>), grid, dim3(GEMM_THREADS), 0, s, tasks, outs, pairs, A, B,
                         C, acc);
      break;
    case TNX_COMPUTE_SPLIT_F32:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 2>), grid, dim3(GEMM_THREADS), 0, s, tasks, outs, pairs, A, B,
                         C, acc);
      break;
    default:
      throw Error{TNX_ERR_DTYPE, "unknown compute mode " + std::to_string(compute)};
  }
}

static void launch_any(int a_lay, int b_lay, int compute, int n_tasks, hipStream_t s, const GemmTask *tasks,
                       const GemmOut *outs, const GemmPair *pairs, const double *A, const double *B, double *C,
                       int acc) {
  if (n_tasks <= 0) return;
  TNX_REQUIRE(b_lay == LAY_KN, TNX_ERR_ARG, "B operands must be K x N row-major");
  dim3 grid(n_tasks);
  if (a_lay == LAY_MK)
    launch_lay<LAY_MK>(compute, grid, s, tasks, outs, pairs, A, B, C, acc);
  else
    launch_lay<LAY_KM>(compute, grid, s, tasks, outs, pairs, A, B, C, acc);
  TNX_CHECK_HIP(hipGetLastError());
}

void launch_grouped_gemm(const GemmProgram &prog, const double *A, const double *B, double *C, int compute,
                         bool accumulate, hipStream_t s) {
  launch_any(prog.a_lay, prog.b_lay, compute, prog.n_tasks, s, prog.tasks.p, prog.outs.p, prog.pairs.p, A, B,
             C, accumulate ? 1 : 0);
}

// Per-block path helper: launch a prepared single-pair program.
void launch_prepared(int a_lay, int b_lay, int compute, int n_tasks, hipStream_t s, const GemmTask *tasks,
                     const GemmOut *outs, const GemmPair *pairs, const double *A, const double *B, double *C,
                     bool accumulate) {
  launch_any(a_lay, b_lay, compute, n_tasks, s, tasks, outs, pairs, A, B, C, accumulate ? 1 : 0);
}

}  // namespace tnx
