// Grouped block GEMM for gfx950: uneven output blocks, each the sum of the
// K-segments of all block pairs that feed it, in one launch; LDS-staged
// operand tiles, MFMA 16x16x4 (FP64 or FP32).
//
// Tiling (chosen per program): 64x64, 128x64, 64x128 or 128x128 output tiles
// per 256-thread workgroup; optionally split K (a task covers a range of the
// output's K tiles and writes a partial tile, a second launch adds the
// partials in slot order), and an XCD-aware task order (tiles sharing operand
// panels dispatched to one XCD's L2).  Built by set_tiling.
//
// Replaces the per-combination einsum of the reference
// (src/tenax/contraction/contractor.py:444-476, src/tenax/algorithms/dmrg.py:1097-1109):
// instead of one dispatch per block pair plus one add per accumulation, every
// pair of a contraction runs in one launch and each output tile accumulates its
// pairs in a fixed order (deterministic, no atomics).
#include "tnx_internal.h"

#include <type_traits>

namespace tnx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Per-thread staging of one A tile (BM x BK) and one B tile (BK x BN):
// 4*PA and 4*PB elements per thread.
template <int BM, int BN>
struct Stage {
  static constexpr int PA = BM / 64;
  static constexpr int PB = BN / 64;
  double a[PA][4];
  double b[PB][4];
  uint32_t neg_hi;  // 0x80000000 for a pair with a fermionic exchange sign (flips A's sign bit)
};

// A with the pair's sign: an XOR of the high word's sign bit (no FP64 multiply).
__device__ __forceinline__ double signed_a(double x, uint32_t neg_hi) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)neg_hi << 32));
}

// Operand loads go through a buffer resource spanning exactly one block, and
// the hardware returns 0 for any offset at or past its end.  That bounds K
// without per-element tests: B rows k >= K (ldb >= N) and KM-layout A rows
// k >= K (lda >= M) lie past the block's end.  Columns of B past N and rows
// of A past M that still fall inside the block read other elements of it;
// they only reach accumulator columns / rows outside the output block, which
// are never stored.  MK-layout A still masks k >= K within a row (those
// elements meet B's zero rows, but a non-finite value must not leak into the
// sum).  The loads carry no branches, so the compiler keeps them in flight
// across the MFMAs of the previous tile.  Offsets are 32-bit: finalize_program
// checks every pair's extent plus one tile of overhang against 2^32 bytes.
constexpr int32_t BUF_DWORD3 = 0x00020000;  // raw dword access, gfx9 family

__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int32_t)off, 0, 0));
}
// Two consecutive doubles (one 128-bit load).
__device__ __forceinline__ double2 buf_ld2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int32_t)off, 0, 0));
}

// Per-thread byte offsets of a thread's operand elements, relative to the
// first row (KM, KN) or column (MK) of the K tile.  They depend only on the
// block pair, so they are computed when the pair changes and stay in
// registers across its K tiles; each tile only moves the buffer descriptor's
// base (scalar arithmetic).
//   A MK: 8 threads per row of 16 k, 2 consecutive k each (one 128-bit
//   load per row), rows 32 apart;
//   A KM and B: one element per thread per k row, consecutive lanes on
//   consecutive m / n (coalesced 512-byte rows, conflict-free LDS stores),
//   rows RPP apart.
template <int ALAY, int BM>
__device__ __forceinline__ void a_offsets(const GemmPair &p, int m0, int tid, uint32_t (&o)[BM / 64][4]) {
#pragma unroll
  for (int q = 0; q < BM / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (ALAY == LAY_MK) {
        // element e = 4q + i is half h = e & 1 of 128-bit load j = e >> 1
        const int e = 4 * q + i;
        const uint32_t m = (uint32_t)(m0 + (tid >> 3) + 32 * (e >> 1));
        o[q][i] = 8u * (m * (uint32_t)p.lda + (uint32_t)(((tid & 7) << 1) + (e & 1)));
      } else {
        constexpr int RPP = 256 / BM;
        o[q][i] = 8u * ((uint32_t)(tid / BM + RPP * (4 * q + i)) * (uint32_t)p.lda + (uint32_t)(m0 + tid % BM));
      }
    }
}
template <int BN>
__device__ __forceinline__ void b_offsets(const GemmPair &p, int n0, int tid, uint32_t (&o)[BN / 64][4]) {
  constexpr int RPP = 256 / BN;
#pragma unroll
  for (int q = 0; q < BN / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      o[q][i] = 8u * ((uint32_t)(tid / BN + RPP * (4 * q + i)) * (uint32_t)p.ldb + (uint32_t)(n0 + tid % BN));
}

// A tile at K offset k0 of pair p (M rows of the output block): the
// descriptor starts at the tile's first column (MK) or row (KM) and ends at
// the block's end.  MK: rows past M lie past the end; k >= K within a row
// selects an out-of-range offset.  KM: rows k >= K lie past the end.
template <int ALAY, int BM>
__device__ __forceinline__ void load_a(const double *__restrict__ A, const GemmPair &p, int M, int k0, int tid,
                                       const uint32_t (&o)[BM / 64][4], double (&r)[BM / 64][4]) {
  const int32_t shift = ALAY == LAY_MK ? k0 : k0 * p.lda;  // elements; < 2^28 (finalize_program)
  const int32_t nb = (int32_t)(8 * (ALAY == LAY_MK ? (int64_t)(M - 1) * p.lda + p.K - k0
                                                    : (int64_t)(p.K - 1 - k0) * p.lda + M));
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)(A + p.a_off + shift), (short)0, nb, BUF_DWORD3);
  if (ALAY == LAY_MK) {
    if (k0 + GEMM_BK <= p.K) {  // wave-uniform: the whole K tile lies inside the pair
#pragma unroll
      for (int e = 0; e < BM / 32; ++e) {
        const double2 v = buf_ld2(rs, o[e >> 1][(e & 1) * 2]);
        r[(2 * e) >> 2][(2 * e) & 3] = v.x;
        r[(2 * e + 1) >> 2][(2 * e + 1) & 3] = v.y;
      }
    } else {
      // the pair's last K tile: element loads, k >= K selects an out-of-range
      // offset (a 128-bit load could straddle the block's end)
      const int lim = p.K - k0 - ((tid & 7) << 1);  // valid k of this thread's pair: <= 0, 1 or >= 2
#pragma unroll
      for (int e = 0; e < BM / 32; ++e) {
        const uint32_t base = o[e >> 1][(e & 1) * 2];
        r[(2 * e) >> 2][(2 * e) & 3] = buf_ld(rs, lim > 0 ? base : (uint32_t)nb);
        r[(2 * e + 1) >> 2][(2 * e + 1) & 3] = buf_ld(rs, lim > 1 ? base + 8u : (uint32_t)nb);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < BM / 64; ++q)
#pragma unroll
      for (int i = 0; i < 4; ++i) r[q][i] = buf_ld(rs, o[q][i]);
  }
}

// B operand is always K x N row-major (KN): b[k*ldb + n]; rows k >= K lie past
// the block's end, columns past N are never stored.
template <int BN>
__device__ __forceinline__ void load_b(const double *__restrict__ B, const GemmPair &p, int N, int k0,
                                       const uint32_t (&o)[BN / 64][4], double (&r)[BN / 64][4]) {
  const int32_t nb = (int32_t)(8 * ((int64_t)(p.K - 1 - k0) * p.ldb + N));
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)(B + p.b_off + (int64_t)k0 * p.ldb), (short)0, nb, BUF_DWORD3);
#pragma unroll
  for (int q = 0; q < BN / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) r[q][i] = buf_ld(rs, o[q][i]);
}

// LDS positions ([k][m] and [k][n], pitch BM+1 / BN+1) of a thread's staged values.
template <int ALAY, int BM>
__device__ __forceinline__ int a_lds(int tid, int q, int i) {
  constexpr int AP = BM + 1;
  if (ALAY == LAY_MK) {
    const int e = 4 * q + i;
    return (((tid & 7) << 1) + (e & 1)) * AP + (tid >> 3) + 32 * (e >> 1);
  }
  constexpr int RPP = 256 / BM;
  return (tid / BM + RPP * (4 * q + i)) * AP + tid % BM;
}
template <int BN>
__device__ __forceinline__ int b_lds(int tid, int q, int i) {
  constexpr int BP = BN + 1;
  constexpr int RPP = 256 / BN;
  return (tid / BN + RPP * (4 * q + i)) * BP + tid % BN;
}

// SIGNED: the program has pairs with a fermionic sign (otherwise no sign work).
template <int ALAY, int BM, int BN, bool SIGNED, class T>
__device__ __forceinline__ void store_tiles(T *As, T *Bs, const Stage<BM, BN> &st, int tid) {
#pragma unroll
  for (int q = 0; q < BM / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      As[a_lds<ALAY, BM>(tid, q, i)] = (T)(SIGNED ? signed_a(st.a[q][i], st.neg_hi) : st.a[q][i]);
#pragma unroll
  for (int q = 0; q < BN / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      Bs[b_lds<BN>(tid, q, i)] = (T)st.b[q][i];
}

// Split-FP32 staging: hi = fp32(x), lo = fp32(x - hi).
template <int ALAY, int BM, int BN, bool SIGNED>
__device__ __forceinline__ void store_tiles_split(float *Ah, float *Al, float *Bh, float *Bl,
                                                  const Stage<BM, BN> &st, int tid) {
#pragma unroll
  for (int q = 0; q < BM / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double x = SIGNED ? signed_a(st.a[q][i], st.neg_hi) : st.a[q][i];
      const float h = (float)x;
      const int at = a_lds<ALAY, BM>(tid, q, i);
      Ah[at] = h;
      Al[at] = (float)(x - (double)h);
    }
#pragma unroll
  for (int q = 0; q < BN / 64; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double x = st.b[q][i];
      const float h = (float)x;
      const int at = b_lds<BN>(tid, q, i);
      Bh[at] = h;
      Bl[at] = (float)(x - (double)h);
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

// Partial tile into a workspace slot (BM x BN row-major, every element written).
template <bool F64LAYOUT, int BN, class V>
__device__ __forceinline__ void store_partial(double *__restrict__ W, int row0, int col0, int lane, const V &acc) {
  const int col = col0 + (lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + (F64LAYOUT ? ((lane >> 4) + 4 * r) : (((lane >> 4) << 2) + r));
    W[row * BN + col] = (double)acc[r];
  }
}

// Offset of index x of a permuted output side: x decomposed over n legs (the
// last fastest), digit times stride summed.  x past the side's extent gives an
// offset that is never used (the element is not stored).
__device__ __forceinline__ int64_t store_offset(int32_t x, int n, const int32_t *dim, const int64_t *str) {
  int64_t off = 0;
  uint32_t v = (uint32_t)x;
#pragma unroll
  for (int l = STORE_LEGS - 1; l >= 0; --l) {
    if (l >= n) continue;
    const uint32_t d = (uint32_t)dim[l];
    off += (int64_t)(v % d) * str[l];
    v /= d;
  }
  return off;
}

// One 16x16 accumulator into a permuted output block: the element of tile row
// r and tile column c goes to c_off + roff[r] + coff[c] (tables of the tile).
template <bool F64LAYOUT, class V>
__device__ __forceinline__ void store_acc_perm(double *__restrict__ C, const GemmOut &o, const int64_t *roff,
                                               const int64_t *coff, int m0, int n0, int lr0, int lc0, int lane,
                                               const V &acc, bool accumulate) {
  const int lc = lc0 + (lane & 15);
  if (n0 + lc >= o.N) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int lr = lr0 + (F64LAYOUT ? ((lane >> 4) + 4 * r) : (((lane >> 4) << 2) + r));
    if (m0 + lr < o.M) {
      double *dst = C + o.c_off + roff[lr] + coff[lc];
      const double v = (double)acc[r];
      *dst = accumulate ? *dst + v : v;
    }
  }
}

// COMPUTE: 0 = FP64 MFMA, 1 = FP32 MFMA, 2 = split FP32 (3 products, FP64 across K chunks).
// BM x BN workgroup tile, 4 waves in 2 x 2, each wave (BM/2) x (BN/2) as
// (BM/32) x (BN/32) MFMA 16x16 blocks.
// At least two waves per SIMD (the 128-wide tiles held to 256 registers),
// except split FP32 at 128x128, whose hi/lo operands would then spill.
// PERM: outputs may carry a GemmStore (permuted output legs); the other
// instantiations skip that test.
template <int ALAY, int COMPUTE, int BM, int BN, bool SIGNED, bool PERM>
__global__ __launch_bounds__(GEMM_THREADS) __attribute__((amdgpu_waves_per_eu(
    (COMPUTE == 2 && BM * BN > 8192) ? 1 : 2))) void grouped_gemm_kernel(
    const GemmTask *__restrict__ tasks, const GemmOut *__restrict__ outs,
    const GemmPair *__restrict__ pairs, const double *__restrict__ A, const double *__restrict__ B,
    double *__restrict__ C, double *__restrict__ ws, int accumulate, const int32_t *__restrict__ stop,
    const GemmStore *__restrict__ stores) {
  constexpr int NT = (COMPUTE == 2) ? 2 : 1;  // hi/lo planes for split
  constexpr int AP = BM + 1, BP = BN + 1;
  constexpr int TI = BM / 32, TJ = BN / 32;
  constexpr bool ONE_STAGE = BM * BN > 8192;
  using T = typename std::conditional<COMPUTE == 0, double, float>::type;
  // Two LDS stages: tile t is read from stage t&1 while tile t+1 is written
  // to the other stage, so each K-tile needs a single barrier.
  __shared__ T As[2][NT][GEMM_BK * AP];
  __shared__ T Bs[2][NT][GEMM_BK * BP];
  __shared__ int64_t tbl[PERM ? BM + BN : 1];  // row / column offsets of a permuted output tile

  // a device-side stop flag (a Lanczos solve that has ended): no work
  if (stop && *stop) return;
  const GemmTask t = tasks[blockIdx.x];
  if (t.out < 0) return;  // padding entry (uniform over the workgroup)
  const GemmOut o = outs[t.out];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * (BM / 2);
  const int wn = (wave & 1) * (BN / 2);
  // 16x16 MFMA blocks of this wave that touch the output block (edge tiles
  // of uneven blocks have fewer): the others are skipped, which frees the
  // matrix pipe for the other waves on the SIMD.  Wave-uniform.
  const int na = __builtin_amdgcn_readfirstlane(min(TI, max(0, (o.M - t.m0 - wm + 15) / 16)));
  const int nb = __builtin_amdgcn_readfirstlane(min(BN / 32, max(0, (o.N - t.n0 - wn + 15) / 16)));
  const int lrow = lane & 15;
  const int lk = lane >> 4;

  d4 dacc[TI][TJ];
  f4 facc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      dacc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      facc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }

  auto stage_store = [&](int buf, const Stage<BM, BN> &st) {
    if (COMPUTE == 2)
      store_tiles_split<ALAY, BM, BN, SIGNED>((float *)As[buf][0], (float *)As[buf][NT - 1], (float *)Bs[buf][0],
                                      (float *)Bs[buf][NT - 1], st, tid);
    else
      store_tiles<ALAY, BM, BN, SIGNED, T>(As[buf][0], Bs[buf][0], st, tid);
  };

  // One BK-deep tile from LDS stage `buf`: the operands of k-step kk+1 are
  // read while the MFMAs of k-step kk run.
  // GUARD: skip the MFMA blocks outside the output block (edge tiles only).
  auto compute_tile = [&](int buf, auto guard) {
    constexpr bool GUARD = decltype(guard)::value;
    if (COMPUTE != 2) {
      const T *as = As[buf][0];
      const T *bs = Bs[buf][0];
      T a[2][TI], b[2][TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) a[0][i] = as[lk * AP + wm + 16 * i + lrow];
#pragma unroll
      for (int j = 0; j < TJ; ++j) b[0][j] = bs[lk * BP + wn + 16 * j + lrow];
#pragma unroll
      for (int kk = 0; kk < GEMM_BK / 4; ++kk) {
        const int cur = kk & 1;
        if (kk + 1 < GEMM_BK / 4) {
          const int kr = 4 * (kk + 1) + lk;
#pragma unroll
          for (int i = 0; i < TI; ++i) a[cur ^ 1][i] = as[kr * AP + wm + 16 * i + lrow];
#pragma unroll
          for (int j = 0; j < TJ; ++j) b[cur ^ 1][j] = bs[kr * BP + wn + 16 * j + lrow];
        }
        // keep the reads of step kk+1 ahead of the MFMAs of step kk
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j) {
            if (GUARD && (i >= na || j >= nb)) continue;
            if (COMPUTE == 0)
              dacc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[cur][i], b[cur][j], dacc[i][j], 0, 0, 0);
            else
              facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[cur][i], b[cur][j], facc[i][j], 0, 0, 0);
          }
      }
      return;
    }
#pragma unroll
    for (int kk = 0; kk < GEMM_BK; kk += 4) {
      const int kr = kk + lk;
      const float *ah = (const float *)As[buf][0];
      const float *al = (const float *)As[buf][NT - 1];
      const float *bh = (const float *)Bs[buf][0];
      const float *bl = (const float *)Bs[buf][NT - 1];
      float aH[TI], aL[TI], bH[TJ], bL[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        aH[i] = ah[kr * AP + wm + 16 * i + lrow];
        aL[i] = al[kr * AP + wm + 16 * i + lrow];
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        bH[j] = bh[kr * BP + wn + 16 * j + lrow];
        bL[j] = bl[kr * BP + wn + 16 * j + lrow];
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          if (GUARD && (i >= na || j >= nb)) continue;
          facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(aL[i], bH[j], facc[i][j], 0, 0, 0);
          facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(aH[i], bL[j], facc[i][j], 0, 0, 0);
          facc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(aH[i], bH[j], facc[i][j], 0, 0, 0);
        }
    }
    // flush the FP32 chunk sums into FP64 accumulators
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) dacc[i][j][r] += (double)facc[i][j][r];
        facc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      }
  };

  // Locate the task's first K tile (skip kt0 tiles over the output's pairs)
  // and count the tiles it covers.
  int pi = o.pair_begin;
  int skip = t.kt0;
  GemmPair p{};
  while (pi < o.pair_end) {
    p = pairs[pi];
    const int nkt = (p.K + GEMM_BK - 1) / GEMM_BK;
    if (skip < nkt) break;
    skip -= nkt;
    ++pi;
  }
  int n = 0;
  if (pi < o.pair_end) {
    int64_t avail = (p.K + GEMM_BK - 1) / GEMM_BK - skip;
    for (int q = pi + 1; q < o.pair_end && avail < (int64_t)(t.kt1 - t.kt0); ++q)
      avail += (pairs[q].K + GEMM_BK - 1) / GEMM_BK;
    n = (int)(avail < (int64_t)(t.kt1 - t.kt0) ? avail : (int64_t)(t.kt1 - t.kt0));
  }

  // Load cursor over the flattened (pair, k-tile) sequence.
  int lk0 = skip * GEMM_BK;
  uint32_t aoff[BM / 64][4], boff[BN / 64][4];
  a_offsets<ALAY, BM>(p, t.m0, tid, aoff);
  b_offsets<BN>(p, t.n0, tid, boff);
  auto advance = [&]() {
    lk0 += GEMM_BK;
    if (lk0 >= p.K) {
      lk0 = 0;
      if (pi + 1 < o.pair_end) {
        p = pairs[pi + 1];
        a_offsets<ALAY, BM>(p, t.m0, tid, aoff);
        b_offsets<BN>(p, t.n0, tid, boff);
      }
      ++pi;
    }
  };
  auto load = [&](Stage<BM, BN> &st) {
    load_a<ALAY, BM>(A, p, o.M, lk0, tid, aoff, st.a);
    load_b<BN>(B, p, o.N, lk0, boff, st.b);
    st.neg_hi = p.neg ? 0x80000000u : 0u;
    advance();
  };

  // Pipeline: tile i is computed from LDS stage i&1 while tile i+1 waits in
  // registers and tile i+2 is in flight from memory (two register stages).
  // GUARD as in compute_tile.  For the MK layout the loop is instantiated
  // twice, so that a wave whose MFMA blocks all lie inside the output block
  // runs its MFMAs as one straight-line sequence (a per-MFMA skip test splits
  // it into basic blocks, and the compiler then pads each MFMA with hazard
  // nops): 3% faster on an MK launch, but 3% slower on a KM one, which keeps
  // the single guarded loop.  Both versions pass the same barriers, so the
  // waves of a workgroup may take different ones.
  // step(i, X, Y): tile i is computed from LDS, X holds tile i+1 (stored
  // after the compute), Y (whose tile is already in LDS) receives the next
  // prefetch.  The prefetch is issued unconditionally (past the last tile it
  // re-reads the last pair's K tiles and is never stored), so the number of
  // loads in flight is the same on every path.
  auto step = [&](int i, Stage<BM, BN> &X, Stage<BM, BN> &Y, auto guard) {
    load(Y);
    compute_tile(i & 1, guard);
    if (i + 1 < n) {
      stage_store((i + 1) & 1, X);
      __syncthreads();
    }
  };
  auto main_loop = [&](auto guard) {
    if (ONE_STAGE) {
      // one register stage (128x128 tiles: a tile's 64 MFMAs per wave cover
      // the load latency of the next): tile i+1 is in flight from memory
      // while tile i is computed from LDS
      Stage<BM, BN> r0;
      load(r0);
      stage_store(0, r0);
      __syncthreads();
      for (int i = 0; i < n; ++i) {
        load(r0);
        compute_tile(i & 1, guard);
        if (i + 1 < n) {
          stage_store((i + 1) & 1, r0);
          __syncthreads();
        }
      }
      return;
    }
    // two register stages: tile i+1 waits in registers while tile i+2 is in flight
    Stage<BM, BN> r0, r1;
    load(r0);
    load(r1);
    stage_store(0, r0);
    __syncthreads();
    for (int i = 0; i < n; i += 2) {
      step(i, r1, r0, guard);
      if (i + 1 < n) step(i + 1, r0, r1, guard);
    }
  };
  if (n > 0) {
    if (ALAY == LAY_MK && na == TI && nb == TJ)
      main_loop(std::false_type{});
    else
      main_loop(std::true_type{});
  }
  if (t.slot >= 0) {
    double *W = ws + (int64_t)t.slot * (BM * BN);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        if (COMPUTE == 1)
          store_partial<false, BN>(W, wm + i * 16, wn + j * 16, lane, facc[i][j]);
        else
          store_partial<COMPUTE == 0, BN>(W, wm + i * 16, wn + j * 16, lane, dacc[i][j]);
      }
    return;
  }
  const bool acc_flag = accumulate != 0;
  if (PERM && stores[t.out].nr > 0) {  // uniform over the workgroup
    const GemmStore st = stores[t.out];
    if (tid < BM) tbl[tid] = store_offset(t.m0 + tid, st.nr, st.rdim, st.rstr);
    else if (tid < BM + BN) tbl[tid] = store_offset(t.n0 + tid - BM, st.nc, st.cdim, st.cstr);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        if (COMPUTE == 1)
          store_acc_perm<false>(C, o, tbl, tbl + BM, t.m0, t.n0, wm + i * 16, wn + j * 16, lane, facc[i][j],
                                acc_flag);
        else
          store_acc_perm<COMPUTE == 0>(C, o, tbl, tbl + BM, t.m0, t.n0, wm + i * 16, wn + j * 16, lane, dacc[i][j],
                                       acc_flag);
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int r0 = t.m0 + wm + i * 16;
      const int c0 = t.n0 + wn + j * 16;
      if (COMPUTE == 1)
        store_acc<false>(C, o, r0, c0, lane, facc[i][j], acc_flag);
      else
        store_acc<COMPUTE == 0>(C, o, r0, c0, lane, dacc[i][j], acc_flag);
    }
}

// Sum the partial tiles of each split tile in slot order (deterministic) into C.
// Each thread owns BM*BN/256 elements 256 apart (coalesced) and loads one
// partial of all of them before adding, so its loads are in flight together.
template <int BM, int BN, bool PERM>
__global__ __launch_bounds__(256) void split_reduce_kernel(const SplitTile *__restrict__ red,
                                                           const GemmOut *__restrict__ outs,
                                                           const double *__restrict__ ws, double *__restrict__ C,
                                                           int accumulate, const int32_t *__restrict__ stop,
                                                           const GemmStore *__restrict__ stores) {
  if (stop && *stop) return;
  constexpr int PER = BM * BN / 256;
  __shared__ int64_t tbl[PERM ? BM + BN : 1];
  const SplitTile r = red[blockIdx.x];
  const GemmOut o = outs[r.out];
  const bool perm = PERM && stores[r.out].nr > 0;  // uniform over the workgroup
  if (perm) {
    const GemmStore st = stores[r.out];
    if (threadIdx.x < BM) tbl[threadIdx.x] = store_offset(r.m0 + threadIdx.x, st.nr, st.rdim, st.rstr);
    else if (threadIdx.x < BM + BN) tbl[threadIdx.x] = store_offset(r.n0 + threadIdx.x - BM, st.nc, st.cdim, st.cstr);
    __syncthreads();
  }
  const double *w = ws + (int64_t)r.slot0 * (BM * BN) + threadIdx.x;
  double v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) v[i] = 0.0;
  for (int q = 0; q < r.nsplit; ++q) {
    double x[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) x[i] = w[(int64_t)q * (BM * BN) + i * 256];
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] += x[i];
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = threadIdx.x + i * 256;
    const int row = r.m0 + e / BN, col = r.n0 + e % BN;
    if (row >= o.M || col >= o.N) continue;
    double *dst = perm ? C + o.c_off + tbl[e / BN] + tbl[BM + e % BN] : C + o.c_off + (int64_t)row * o.ldc + col;
    *dst = accumulate ? *dst + v[i] : v[i];
  }
}

template <int ALAY, int BM, int BN, bool SIGNED, bool PERM>
static void launch_tile(int compute, dim3 grid, hipStream_t s, const GemmTask *tasks, const GemmOut *outs,
                        const GemmPair *pairs, const double *A, const double *B, double *C, double *ws, int acc,
                        const int32_t *stop, const GemmStore *stores) {
  switch (compute) {
    case TNX_COMPUTE_F64:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 0, BM, BN, SIGNED, PERM>), grid, dim3(GEMM_THREADS), 0, s, tasks,
                         outs, pairs, A, B, C, ws, acc, stop, stores);
      break;
    case TNX_COMPUTE_F32:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 1, BM, BN, SIGNED, PERM>), grid, dim3(GEMM_THREADS), 0, s, tasks,
                         outs, pairs, A, B, C, ws, acc, stop, stores);
      break;
    case TNX_COMPUTE_SPLIT_F32:
      hipLaunchKernelGGL((grouped_gemm_kernel<ALAY, 2, BM, BN, SIGNED, PERM>), grid, dim3(GEMM_THREADS), 0, s, tasks,
                         outs, pairs, A, B, C, ws, acc, stop, stores);
      break;
    default:
      throw Error{TNX_ERR_DTYPE, "unknown compute mode " + std::to_string(compute)};
  }
}

template <int ALAY, bool SIGNED, bool PERM>
static void launch_shape(int shape, int compute, dim3 grid, hipStream_t s, const GemmTask *tasks,
                         const GemmOut *outs, const GemmPair *pairs, const double *A, const double *B, double *C,
                         double *ws, int acc, const int32_t *stop, const GemmStore *stores) {
  switch (shape) {
    case 0: launch_tile<ALAY, 64, 64, SIGNED, PERM>(compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores); break;
    case 1: launch_tile<ALAY, 128, 64, SIGNED, PERM>(compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores); break;
    case 2: launch_tile<ALAY, 64, 128, SIGNED, PERM>(compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores); break;
    case 3: launch_tile<ALAY, 128, 128, SIGNED, PERM>(compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores); break;
    default: throw Error{TNX_ERR_ARG, "unknown tile shape " + std::to_string(shape)};
  }
}

// stores != nullptr: the program's outputs carry GemmStore descriptors
// (permuted output legs; bosonic programs only).
static void launch_any(int a_lay, int b_lay, bool signs, int shape, int compute, int n_tasks, hipStream_t s,
                       const GemmTask *tasks, const GemmOut *outs, const GemmPair *pairs, const double *A,
                       const double *B, double *C, double *ws, int acc, const int32_t *stop,
                       const GemmStore *stores = nullptr) {
  if (n_tasks <= 0) return;
  TNX_REQUIRE(b_lay == LAY_KN, TNX_ERR_ARG, "B operands must be K x N row-major");
  TNX_REQUIRE(!(signs && stores), TNX_ERR_ARG, "permuted output stores are not available with fermionic signs");
  dim3 grid(n_tasks);
  if (stores) {
    if (a_lay == LAY_MK)
      launch_shape<LAY_MK, false, true>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores);
    else
      launch_shape<LAY_KM, false, true>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, stores);
  } else if (a_lay == LAY_MK) {
    if (signs)
      launch_shape<LAY_MK, true, false>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, nullptr);
    else
      launch_shape<LAY_MK, false, false>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, nullptr);
  } else {
    if (signs)
      launch_shape<LAY_KM, true, false>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, nullptr);
    else
      launch_shape<LAY_KM, false, false>(shape, compute, grid, s, tasks, outs, pairs, A, B, C, ws, acc, stop, nullptr);
  }
  TNX_CHECK_HIP(hipGetLastError());
}

void launch_grouped_gemm(const GemmProgram &prog, const double *A, const double *B, double *C, int compute,
                         bool accumulate, hipStream_t s, bool reduce, const int32_t *stop) {
  const int acc = accumulate ? 1 : 0;
  const GemmStore *stores = prog.permuted ? prog.stores.p : nullptr;
  if (!prog.tiled) {
    launch_any(prog.a_lay, prog.b_lay, prog.signs, 0, compute, prog.n_tasks, s, prog.tasks.p, prog.outs.p, prog.pairs.p, A, B,
               C, nullptr, acc, stop, stores);
    return;
  }
  launch_any(prog.a_lay, prog.b_lay, prog.signs, prog.cfg.shape, compute, prog.n_vtasks, s, prog.vtasks.p, prog.outs.p,
             prog.pairs.p, A, B, C, prog.ws.p, acc, stop, stores);
  if (reduce) launch_split_reduce(prog, C, accumulate, s, stop);
}

template <bool PERM>
static void launch_reduce(const GemmProgram &prog, double *C, int acc, hipStream_t s, const int32_t *stop) {
  const dim3 g(prog.n_red), b(256);
  const GemmStore *st = PERM ? prog.stores.p : nullptr;
  switch (prog.cfg.shape) {
    case 0: hipLaunchKernelGGL((split_reduce_kernel<64, 64, PERM>), g, b, 0, s, prog.red.p, prog.outs.p, prog.ws.p, C, acc, stop, st); break;
    case 1: hipLaunchKernelGGL((split_reduce_kernel<128, 64, PERM>), g, b, 0, s, prog.red.p, prog.outs.p, prog.ws.p, C, acc, stop, st); break;
    case 2: hipLaunchKernelGGL((split_reduce_kernel<64, 128, PERM>), g, b, 0, s, prog.red.p, prog.outs.p, prog.ws.p, C, acc, stop, st); break;
    default: hipLaunchKernelGGL((split_reduce_kernel<128, 128, PERM>), g, b, 0, s, prog.red.p, prog.outs.p, prog.ws.p, C, acc, stop, st); break;
  }
  TNX_CHECK_HIP(hipGetLastError());
}

void launch_split_reduce(const GemmProgram &prog, double *C, bool accumulate, hipStream_t s, const int32_t *stop) {
  const int acc = accumulate ? 1 : 0;
  if (prog.tiled && prog.n_red > 0) {
    if (prog.permuted)
      launch_reduce<true>(prog, C, acc, s, stop);
    else
      launch_reduce<false>(prog, C, acc, s, stop);
  }
}

// Per-block path helper: launch a prepared single-pair program (64x64 whole tiles).
void launch_prepared(int a_lay, int b_lay, bool signs, int compute, int n_tasks, hipStream_t s,
                     const GemmTask *tasks, const GemmOut *outs, const GemmPair *pairs, const double *A,
                     const double *B, double *C, bool accumulate) {
  launch_any(a_lay, b_lay, signs, 0, compute, n_tasks, s, tasks, outs, pairs, A, B, C, nullptr, accumulate ? 1 : 0,
             nullptr);
}

}  // namespace tnx
