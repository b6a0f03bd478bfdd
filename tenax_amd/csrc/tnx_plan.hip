// Device-side integer planning for block-sparse contractions.
//
// Everything that decides the sector structure runs as GPU kernels:
//  * valid-sector enumeration (tenax core/tensor.py:53-165): one thread per
//    candidate charge tuple in mixed-radix order over each leg's sorted distinct
//    charges, U(1) last leg solved from the running sum, then an order-preserving
//    compaction -> keys come out lexicographically sorted;
//  * block matching (contractor.py:375-441, dmrg.py:1111-1136): one thread per
//    (A block, B block) candidate, equality of contracted-leg charges, output key
//    located by binary search in the sorted valid set (conservation filter) or in
//    the full candidate set (reachability semantics);
//  * output offsets, pair grouping by output block (stable radix sort, so pairs
//    accumulate in (A index, B index) order), GEMM tiles and their LPT order.
// Only the finished tables are copied back (the Python layer needs keys/offsets).
#include "tnx_internal.h"

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <numeric>

namespace tnx {

static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }

constexpr int MAX_RANK = 12;

// ------------------------------------------------------------------ primitives
template <class T>
static void dev_exclusive_scan(const T *in, T *out, size_t n, hipStream_t s, DevBuf<char> &tmp) {
  if (n == 0) return;
  size_t tb = 0;
  TNX_CHECK_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, T(0), n, rocprim::plus<T>(), s));
  tmp.alloc(tb ? tb : 1);
  TNX_CHECK_HIP(rocprim::exclusive_scan(tmp.p, tb, in, out, T(0), n, rocprim::plus<T>(), s));
}

template <class K, class V>
static void dev_sort_pairs(const K *kin, K *kout, const V *vin, V *vout, size_t n, int end_bit, hipStream_t s,
                           DevBuf<char> &tmp) {
  if (n == 0) return;
  size_t tb = 0;
  TNX_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, n, 0, end_bit, s));
  tmp.alloc(tb ? tb : 1);
  TNX_CHECK_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kin, kout, vin, vout, n, 0, end_bit, s));
}

static int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

static unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------------ legs on device
struct LegDev {
  int32_t off;   // into concatenated charge/mult arrays
  int32_t n;     // distinct charges
  int32_t flow;  // +1 / -1
  int32_t pad;
};

struct LegTable {
  std::vector<LegDev> h;
  std::vector<int32_t> ch, mu;
  DevBuf<LegDev> d_legs;
  DevBuf<int32_t> d_ch, d_mu;
  void build(const tnx_leg *legs, int rank, hipStream_t s) {
    h.resize(rank);
    ch.clear();
    mu.clear();
    for (int l = 0; l < rank; ++l) {
      TNX_REQUIRE(legs[l].n_charges > 0, TNX_ERR_BAD_STRUCTURE, "leg with no charges");
      TNX_REQUIRE(legs[l].flow == 1 || legs[l].flow == -1, TNX_ERR_ARG, "leg flow must be +1 or -1");
      h[l] = LegDev{(int32_t)ch.size(), legs[l].n_charges, legs[l].flow, 0};
      for (int c = 0; c < legs[l].n_charges; ++c) {
        if (c) TNX_REQUIRE(legs[l].charges[c] > legs[l].charges[c - 1], TNX_ERR_ARG, "leg charges must be strictly ascending");
        TNX_REQUIRE(legs[l].mult[c] > 0, TNX_ERR_ARG, "leg multiplicity must be positive");
        ch.push_back(legs[l].charges[c]);
        mu.push_back(legs[l].mult[c]);
      }
    }
    d_legs.upload(h.data(), h.size(), s);
    d_ch.upload(ch.data(), ch.size(), s);
    d_mu.upload(mu.data(), mu.size(), s);
  }
};

__device__ __forceinline__ int64_t pymod(int64_t a, int64_t n) {
  int64_t r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ int find_charge(const int32_t *ch, int n, int32_t q) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int32_t v = ch[mid];
    if (v == q) return mid;
    if (v < q) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

// mode: 0 = conservation filter (U(1): last leg solved), 1 = all tuples.
__global__ void enum_keys_kernel(int rank, const LegDev *__restrict__ legs, const int32_t *__restrict__ ch,
                                 const int32_t *__restrict__ mu, int sym_n, int all_tuples, int64_t target,
                                 int solve_last, int64_t ncand, int32_t *__restrict__ keys,
                                 int32_t *__restrict__ shapes, int32_t *__restrict__ flag) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= ncand) return;
  const int nenum = solve_last ? rank - 1 : rank;
  int32_t key[MAX_RANK], shp[MAX_RANK];
  int64_t r = idx;
  for (int l = nenum - 1; l >= 0; --l) {
    const int n = legs[l].n;
    const int d = (int)(r % n);
    r /= n;
    key[l] = ch[legs[l].off + d];
    shp[l] = mu[legs[l].off + d];
  }
  int ok = 1;
  if (!all_tuples) {
    if (rank == 1) {
      // tensor.py:89-95: single leg, raw flow*q == target
      ok = ((int64_t)legs[0].flow * key[0] == target);
    } else if (solve_last) {
      // tensor.py:134-149: U(1) running sum, last charge determined
      int64_t fused = 0;
      for (int l = 0; l < nenum; ++l) fused += (int64_t)legs[l].flow * key[l];
      const LegDev last = legs[rank - 1];
      const int64_t ql = (target - fused) * last.flow;
      int pos = -1;
      if (ql >= INT32_MIN && ql <= INT32_MAX) pos = find_charge(ch + last.off, last.n, (int32_t)ql);
      if (pos < 0) {
        ok = 0;
      } else {
        key[rank - 1] = (int32_t)ql;
        shp[rank - 1] = mu[last.off + pos];
      }
    } else {
      // tensor.py:99-128,150-163: Z_n, first leg raw, every fuse reduced mod n
      int64_t fused = (int64_t)legs[0].flow * key[0];
      for (int l = 1; l < rank; ++l) fused = pymod(fused + (int64_t)legs[l].flow * key[l], sym_n);
      ok = (fused == target);
    }
  }
  flag[idx] = ok;
  if (ok) {
    for (int l = 0; l < rank; ++l) {
      keys[idx * rank + l] = key[l];
      shapes[idx * rank + l] = shp[l];
    }
  }
}

__global__ void compact_keys_kernel(int rank, int64_t ncand, const int32_t *__restrict__ flag,
                                    const int64_t *__restrict__ pos, const int32_t *__restrict__ kin,
                                    const int32_t *__restrict__ sin, int32_t *__restrict__ kout,
                                    int32_t *__restrict__ sout) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= ncand || !flag[idx]) return;
  const int64_t o = pos[idx];
  for (int l = 0; l < rank; ++l) {
    kout[o * rank + l] = kin[idx * rank + l];
    sout[o * rank + l] = sin[idx * rank + l];
  }
}

__global__ void flag_to_i64_kernel(int64_t n, const int32_t *__restrict__ f, int64_t *__restrict__ o) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) o[i] = f[i];
}

static constexpr int64_t MAX_CANDIDATES = int64_t(1) << 28;

void enumerate_keys(const tnx_leg *legs, int rank, int sym_n, bool all_tuples, bool has_target,
                           int target, hipStream_t s, KeySet &out) {
  TNX_REQUIRE(rank >= 1 && rank <= MAX_RANK, TNX_ERR_ARG, "rank out of range");
  TNX_REQUIRE(sym_n >= 0, TNX_ERR_ARG, "sym_n must be >= 0");
  LegTable lt;
  lt.build(legs, rank, s);
  const bool u1 = sym_n == 0;
  const bool solve_last = !all_tuples && u1 && rank >= 2;
  const int nenum = solve_last ? rank - 1 : rank;
  int64_t ncand = 1;
  for (int l = 0; l < nenum; ++l) {
    ncand *= legs[l].n_charges;
    TNX_REQUIRE(ncand <= MAX_CANDIDATES, TNX_ERR_CAPACITY, "sector enumeration exceeds 2^28 candidates");
  }
  const int64_t tgt = has_target ? target : 0;
  DevBuf<int32_t> ckeys(ncand * rank), cshapes(ncand * rank), flag(ncand);
  DevBuf<int64_t> f64(ncand), pos(ncand);
  DevBuf<char> tmp;
  hipLaunchKernelGGL(enum_keys_kernel, dim3(nblk(ncand)), dim3(256), 0, s, rank, lt.d_legs.p, lt.d_ch.p,
                     lt.d_mu.p, sym_n, all_tuples ? 1 : 0, tgt, solve_last ? 1 : 0, ncand, ckeys.p, cshapes.p,
                     flag.p);
  TNX_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(flag_to_i64_kernel, dim3(nblk(ncand)), dim3(256), 0, s, ncand, flag.p, f64.p);
  dev_exclusive_scan(f64.p, pos.p, ncand, s, tmp);
  int64_t last_pos = 0, last_flag = 0;
  TNX_CHECK_HIP(hipMemcpyAsync(&last_pos, pos.p + ncand - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipMemcpyAsync(&last_flag, f64.p + ncand - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  out.rank = rank;
  out.n = last_pos + last_flag;
  out.keys.alloc(std::max<int64_t>(out.n, 1) * rank);
  out.shapes.alloc(std::max<int64_t>(out.n, 1) * rank);
  if (out.n)
    hipLaunchKernelGGL(compact_keys_kernel, dim3(nblk(ncand)), dim3(256), 0, s, rank, ncand, flag.p, pos.p,
                       ckeys.p, cshapes.p, out.keys.p, out.shapes.p);
  TNX_CHECK_HIP(hipGetLastError());
  TNX_CHECK_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ pairwise planning
__device__ __forceinline__ int key_cmp(const int32_t *a, const int32_t *b, int r) {
  for (int l = 0; l < r; ++l) {
    if (a[l] < b[l]) return -1;
    if (a[l] > b[l]) return 1;
  }
  return 0;
}

__device__ __forceinline__ int64_t find_key(const int32_t *V, int64_t nV, const int32_t *k, int r) {
  int64_t lo = 0, hi = nV - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int c = key_cmp(V + mid * r, k, r);
    if (c == 0) return mid;
    if (c < 0) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

struct PairSpec {
  int32_t nA, nB, rA, rB, nc, rOut;
  int32_t a_c0, b_c0;  // first contracted leg position in A / B
  int32_t a_f0, b_f0;  // first free leg position in A / B
};

__global__ void match_pairs_kernel(PairSpec sp, const int32_t *__restrict__ kA, const int32_t *__restrict__ kB,
                                   const int32_t *__restrict__ V, int64_t nV, int64_t *__restrict__ pair_out,
                                   int32_t *__restrict__ present) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)sp.nA * sp.nB) return;
  const int i = (int)(idx / sp.nB), j = (int)(idx % sp.nB);
  const int32_t *ka = kA + (int64_t)i * sp.rA;
  const int32_t *kb = kB + (int64_t)j * sp.rB;
  int64_t res = -1;
  bool ok = true;
  for (int c = 0; c < sp.nc; ++c)
    if (ka[sp.a_c0 + c] != kb[sp.b_c0 + c]) { ok = false; break; }
  if (ok) {
    int32_t ok_[MAX_RANK];
    int w = 0;
    for (int l = 0; l < sp.rA - sp.nc; ++l) ok_[w++] = ka[sp.a_f0 + l];
    for (int l = 0; l < sp.rB - sp.nc; ++l) ok_[w++] = kb[sp.b_f0 + l];
    res = find_key(V, nV, ok_, sp.rOut);
    if (res >= 0) present[res] = 1;
  }
  pair_out[idx] = res;
}

__global__ void pair_flags_kernel(int64_t n, const int64_t *__restrict__ pair_out, int64_t *__restrict__ f) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) f[i] = pair_out[i] >= 0 ? 1 : 0;
}

__global__ void compact_pairs_kernel(int64_t n, const int64_t *__restrict__ pair_out,
                                     const int64_t *__restrict__ pos, const int64_t *__restrict__ out_id,
                                     uint32_t *__restrict__ vp_oid, uint32_t *__restrict__ vp_idx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n || pair_out[i] < 0) return;
  vp_oid[pos[i]] = (uint32_t)out_id[pair_out[i]];
  vp_idx[pos[i]] = (uint32_t)i;
}

__global__ void present_to_i64_kernel(int64_t n, const int32_t *__restrict__ p, int64_t *__restrict__ o) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) o[i] = p[i] ? 1 : 0;
}

// Output block tables: keys, shapes, sizes and (M, N) split of each present key.
__global__ void out_tables_kernel(int64_t nV, int rOut, int nfreeA, const int32_t *__restrict__ present,
                                  const int64_t *__restrict__ out_id, const int32_t *__restrict__ V,
                                  const int32_t *__restrict__ Vs, int32_t *__restrict__ okeys,
                                  int32_t *__restrict__ oshapes, int64_t *__restrict__ osize,
                                  int32_t *__restrict__ oM, int32_t *__restrict__ oN) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= nV || !present[v]) return;
  const int64_t o = out_id[v];
  int64_t M = 1, N = 1;
  for (int l = 0; l < rOut; ++l) {
    okeys[o * rOut + l] = V[v * rOut + l];
    const int32_t d = Vs[v * rOut + l];
    oshapes[o * rOut + l] = d;
    if (l < nfreeA) M *= d; else N *= d;
  }
  osize[o] = M * N;
  oM[o] = (int32_t)M;
  oN[o] = (int32_t)N;
}

__global__ void segments_kernel(int64_t n, const uint32_t *__restrict__ soid, int32_t *__restrict__ pbeg,
                                int32_t *__restrict__ pend) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t o = soid[k];
  if (k == 0 || soid[k - 1] != o) pbeg[o] = (int32_t)k;
  if (k == n - 1 || soid[k + 1] != o) pend[o] = (int32_t)(k + 1);
}

// Fermionic exchange sign of a block pair (contractor.py:462-470): one factor
// -1 per listed leg pair (x, y) -- positions in the concatenated [A legs, B
// legs] -- whose two charges are both odd (parity = charge & 1).
__device__ __forceinline__ int pair_negative(const int32_t *ka, const int32_t *kb, int rA, const int32_t *sgn,
                                             int n_sgn) {
  int neg = 0;
  for (int t = 0; t < n_sgn; ++t) {
    const int x = sgn[2 * t], y = sgn[2 * t + 1];
    const int32_t qx = x < rA ? ka[x] : kb[x - rA];
    const int32_t qy = y < rA ? ka[y] : kb[y - rA];
    neg ^= (qx & 1) & (qy & 1);
  }
  return neg;
}

__global__ void build_pairs_kernel(int64_t n, PairSpec sp, int a_kfirst, int b_kfirst,
                                   const uint32_t *__restrict__ sidx, const int32_t *__restrict__ sA,
                                   const int32_t *__restrict__ sB, const int64_t *__restrict__ offA,
                                   const int64_t *__restrict__ offB, const int32_t *__restrict__ kA,
                                   const int32_t *__restrict__ kB, const int32_t *__restrict__ sgn, int n_sgn,
                                   GemmPair *__restrict__ pairs) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t p = sidx[k];
  const int i = (int)(p / sp.nB), j = (int)(p % sp.nB);
  int64_t K = 1, M = 1, N = 1;
  for (int c = 0; c < sp.nc; ++c) K *= sA[(int64_t)i * sp.rA + sp.a_c0 + c];
  for (int l = 0; l < sp.rA - sp.nc; ++l) M *= sA[(int64_t)i * sp.rA + sp.a_f0 + l];
  for (int l = 0; l < sp.rB - sp.nc; ++l) N *= sB[(int64_t)j * sp.rB + sp.b_f0 + l];
  GemmPair g;
  g.a_off = offA[i];
  g.b_off = offB[j];
  g.K = (int32_t)K;
  g.lda = a_kfirst ? (int32_t)M : (int32_t)K;
  g.ldb = (int32_t)N;
  (void)b_kfirst;
  g.neg = n_sgn ? pair_negative(kA + (int64_t)i * sp.rA, kB + (int64_t)j * sp.rB, sp.rA, sgn, n_sgn) : 0;
  pairs[k] = g;
}

__global__ void build_outs_kernel(int64_t nout, const int64_t *__restrict__ ooff, const int32_t *__restrict__ oM,
                                  const int32_t *__restrict__ oN, const int32_t *__restrict__ pbeg,
                                  const int32_t *__restrict__ pend, GemmOut *__restrict__ outs) {
  const int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (o >= nout) return;
  GemmOut g;
  g.c_off = ooff[o];
  g.M = oM[o];
  g.N = oN[o];
  g.ldc = oN[o];
  g.pair_begin = pbeg[o];
  g.pair_end = pend[o];
  g.pad = 0;
  outs[o] = g;
}

// ------------------------------------------------------------------ tiling (device)
__global__ void tile_count_kernel(int64_t nout, const GemmOut *__restrict__ outs, const GemmPair *__restrict__ pairs,
                                  int64_t *__restrict__ ntile, uint64_t *__restrict__ costkey, uint32_t *__restrict__ ids,
                                  uint64_t maxcost) {
  const int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (o >= nout) return;
  const GemmOut g = outs[o];
  ntile[o] = (int64_t)((g.M + GEMM_BM - 1) / GEMM_BM) * ((g.N + GEMM_BN - 1) / GEMM_BN);
  uint64_t cost = 0;
  for (int k = g.pair_begin; k < g.pair_end; ++k) cost += (uint64_t)((pairs[k].K + GEMM_BK - 1) / GEMM_BK);
  if (cost > maxcost) cost = maxcost;
  costkey[o] = maxcost - cost;  // ascending sort -> longest first
  ids[o] = (uint32_t)o;
}

__global__ void gather_i64_kernel(int64_t n, const uint32_t *__restrict__ idx, const int64_t *__restrict__ in,
                                  int64_t *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

__global__ void write_tasks_kernel(int64_t nout, const uint32_t *__restrict__ order, const int64_t *__restrict__ toff,
                                   const GemmOut *__restrict__ outs, GemmTask *__restrict__ tasks) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= nout) return;
  const int o = (int)order[r];
  const GemmOut g = outs[o];
  const int tn = (g.N + GEMM_BN - 1) / GEMM_BN;
  const int tm = (g.M + GEMM_BM - 1) / GEMM_BM;
  int64_t t = toff[r];
  for (int a = 0; a < tm; ++a)
    for (int b = 0; b < tn; ++b) tasks[t++] = whole_tile(o, a * GEMM_BM, b * GEMM_BN);
}

// The grouped GEMM addresses each operand block through a 32-bit buffer
// offset and relies on the block's end for its K bound: every pair needs
// lda >= K (MK) or lda >= M (KM), ldb >= N, and its extent plus one tile of
// overhang (the largest tile, 128 x 128 x 16) below 2^32 bytes.
static void check_operand_extents(const GemmProgram &prog) {
  constexpr int64_t LIM = int64_t(1) << 32;
  constexpr int64_t OVER = 128;
  for (const GemmOut &o : prog.h_outs)
    for (int k = o.pair_begin; k < o.pair_end; ++k) {
      const GemmPair &p = prog.h_pairs[k];
      if (p.K <= 0) continue;  // no K tiles: never loaded
      int64_t a_end, b_end;
      if (prog.a_lay == LAY_MK) {
        TNX_REQUIRE(p.lda >= p.K, TNX_ERR_ARG, "A block row pitch below K");
        a_end = 8 * ((int64_t)(o.M + OVER) * p.lda + p.K + GEMM_BK);
      } else {
        TNX_REQUIRE(p.lda >= o.M, TNX_ERR_ARG, "A block row pitch below M");
        a_end = 8 * ((int64_t)(p.K + GEMM_BK) * p.lda + o.M + OVER);
      }
      TNX_REQUIRE(p.ldb >= o.N, TNX_ERR_ARG, "B block row pitch below N");
      b_end = 8 * ((int64_t)(p.K + GEMM_BK) * p.ldb + o.N + OVER);
      TNX_REQUIRE(a_end < LIM && b_end < LIM, TNX_ERR_CAPACITY,
                  "block too large for the grouped GEMM's 32-bit operand offsets");
    }
}

void finalize_program(GemmProgram &prog, hipStream_t s) {
  const int64_t nout = prog.n_outs;
  prog.n_tasks = 0;
  if (nout == 0) {
    prog.h_outs.clear();
    prog.h_pairs.clear();
    prog.flops = 0;
    return;
  }
  DevBuf<int64_t> ntile(nout), ntile_sorted(nout), toff(nout);
  DevBuf<uint64_t> ck(nout), ck_sorted(nout);
  DevBuf<uint32_t> ids(nout), order(nout);
  DevBuf<char> tmp;
  const uint64_t maxcost = (uint64_t(1) << 40) - 1;
  hipLaunchKernelGGL(tile_count_kernel, dim3(nblk(nout)), dim3(256), 0, s, nout, prog.outs.p, prog.pairs.p, ntile.p,
                     ck.p, ids.p, maxcost);
  TNX_CHECK_HIP(hipGetLastError());
  dev_sort_pairs(ck.p, ck_sorted.p, ids.p, order.p, nout, 40, s, tmp);
  hipLaunchKernelGGL(gather_i64_kernel, dim3(nblk(nout)), dim3(256), 0, s, nout, order.p, ntile.p, ntile_sorted.p);
  dev_exclusive_scan(ntile_sorted.p, toff.p, nout, s, tmp);
  int64_t last_off = 0, last_n = 0;
  TNX_CHECK_HIP(hipMemcpyAsync(&last_off, toff.p + nout - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipMemcpyAsync(&last_n, ntile_sorted.p + nout - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t ntasks = last_off + last_n;
  TNX_REQUIRE(ntasks < (int64_t(1) << 31), TNX_ERR_CAPACITY, "too many GEMM tiles");
  prog.tasks.alloc(std::max<int64_t>(ntasks, 1));
  hipLaunchKernelGGL(write_tasks_kernel, dim3(nblk(nout)), dim3(256), 0, s, nout, order.p, toff.p, prog.outs.p,
                     prog.tasks.p);
  TNX_CHECK_HIP(hipGetLastError());
  prog.n_tasks = (int32_t)ntasks;
  prog.h_outs.resize(prog.n_outs);
  prog.h_pairs.resize(prog.n_pairs);
  prog.outs.download(prog.h_outs.data(), prog.n_outs, s);
  prog.pairs.download(prog.h_pairs.data(), prog.n_pairs, s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  check_operand_extents(prog);
  prog.signs = false;
  for (const GemmPair &p : prog.h_pairs)
    if (p.neg) prog.signs = true;
  double f = 0;
  for (const auto &o : prog.h_outs)
    for (int k = o.pair_begin; k < o.pair_end; ++k) f += 2.0 * o.M * (double)o.N * prog.h_pairs[k].K;
  prog.flops = f;
}

// ------------------------------------------------------------------ tiling variants (host)
static void tile_dims(int shape, int &bm, int &bn) {
  bm = (shape & 1) ? 128 : 64;
  bn = (shape & 2) ? 128 : 64;
}

// Resident workgroups per CU of each tile shape (LDS-bound, FP64 staging).
static int wg_per_cu(int shape) { return shape == 0 ? 4 : shape == 3 ? 2 : 3; }

void set_tiling(GemmProgram &prog, const TileCfg &cfg, hipStream_t s) {
  TNX_REQUIRE(cfg.shape >= 0 && cfg.shape <= 3, TNX_ERR_ARG, "tile shape must be 0..3");
  prog.cfg = cfg;
  prog.tiled = false;
  prog.n_vtasks = prog.n_red = 0;
  if (cfg.shape == 0 && cfg.kchunk == 0 && cfg.xcd == 0) return;  // the device-built default list
  int bm, bn;
  tile_dims(cfg.shape, bm, bn);
  const auto &ho = prog.h_outs;
  const auto &hp = prog.h_pairs;
  const int64_t nout = (int64_t)ho.size();
  std::vector<int64_t> kt(nout), ntile(nout);
  double total = 0.0;  // K tiles summed over all output tiles
  for (int64_t o = 0; o < nout; ++o) {
    int64_t c = 0;
    for (int k = ho[o].pair_begin; k < ho[o].pair_end; ++k) c += (hp[k].K + GEMM_BK - 1) / GEMM_BK;
    kt[o] = c;
    ntile[o] = (int64_t)((ho[o].M + bm - 1) / bm) * ((ho[o].N + bn - 1) / bn);
    total += (double)ntile[o] * (double)std::max<int64_t>(c, 1);
  }
  // K tiles per task: bounded by half the mean work of one resident workgroup
  // (kchunk < 0), or the balanced two-wave schedule (kchunk == KCHUNK_BALANCED):
  // a first wave of whole tiles and heads of the long ones, cut at the mean
  // work of one workgroup slot (all of it dispatched at once), and a second
  // wave of short tail pieces that the slots pick up as they free
  const bool balanced = cfg.kchunk == KCHUNK_BALANCED;
  const double slots = 256.0 * wg_per_cu(cfg.shape);
  const int64_t cap = std::max<int64_t>(8, (int64_t)std::ceil(total / slots));
  const int64_t piece = std::max<int64_t>(4, cap / 8);
  int64_t chunk = balanced ? 0 : cfg.kchunk;
  if (chunk < 0) chunk = std::max<int64_t>(8, (int64_t)std::ceil(total / (2.0 * slots)));
  // work units (an output, or a row of tiles of a large output) with their tasks
  struct Unit {
    int64_t cost;
    int32_t out, mt0, mt1;
  };
  std::vector<Unit> units;
  const double big = total / 64.0;
  for (int64_t o = 0; o < nout; ++o) {
    const int tm = (ho[o].M + bm - 1) / bm;
    const int64_t per_row = (int64_t)((ho[o].N + bn - 1) / bn) * std::max<int64_t>(kt[o], 1);
    if (cfg.xcd && tm > 1 && (double)per_row * tm > big) {
      for (int a = 0; a < tm; ++a) units.push_back(Unit{per_row, (int32_t)o, a, a + 1});
    } else {
      units.push_back(Unit{per_row * tm, (int32_t)o, 0, tm});
    }
  }
  std::stable_sort(units.begin(), units.end(), [](const Unit &x, const Unit &y) { return x.cost > y.cost; });
  const int nbin = cfg.xcd ? N_XCD : 1;
  std::vector<std::vector<GemmTask>> bins(nbin), tails(nbin);
  std::vector<int64_t> load(nbin, 0);
  std::vector<SplitTile> red;
  int64_t slot = 0;
  for (const Unit &u : units) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[b] += u.cost;
    const GemmOut &g = ho[u.out];
    const int tn = (g.N + bn - 1) / bn;
    const int64_t K = kt[u.out];
    if (balanced) {
      // head [0, cap) in the first wave, the rest in pieces of `piece` K tiles
      const int nsplit = K > cap ? 1 + (int)((K - cap + piece - 1) / piece) : 1;
      for (int a = u.mt0; a < u.mt1; ++a)
        for (int c = 0; c < tn; ++c) {
          if (nsplit == 1) {
            bins[b].push_back(whole_tile(u.out, a * bm, c * bn));
            continue;
          }
          red.push_back(SplitTile{u.out, a * bm, c * bn, (int32_t)slot, nsplit, 0});
          bins[b].push_back(GemmTask{u.out, a * bm, c * bn, 0, (int32_t)cap, (int32_t)slot, 0, 0});
          for (int q = 1; q < nsplit; ++q) {
            const int32_t k0 = (int32_t)(cap + (q - 1) * piece), k1 = (int32_t)std::min<int64_t>(K, cap + q * piece);
            tails[b].push_back(GemmTask{u.out, a * bm, c * bn, k0, k1, (int32_t)(slot + q), 0, 0});
          }
          slot += nsplit;
        }
      continue;
    }
    const int nsplit = (chunk > 0 && K > chunk) ? (int)((K + chunk - 1) / chunk) : 1;
    for (int a = u.mt0; a < u.mt1; ++a)
      for (int c = 0; c < tn; ++c) {
        if (nsplit == 1) {
          bins[b].push_back(whole_tile(u.out, a * bm, c * bn));
          continue;
        }
        red.push_back(SplitTile{u.out, a * bm, c * bn, (int32_t)slot, nsplit, 0});
        for (int q = 0; q < nsplit; ++q) {
          const int32_t k0 = (int32_t)(K * q / nsplit), k1 = (int32_t)(K * (q + 1) / nsplit);
          bins[b].push_back(GemmTask{u.out, a * bm, c * bn, k0, k1, (int32_t)(slot + q), 0, 0});
        }
        slot += nsplit;
      }
  }
  if (slot * bm * bn > MAX_WS_DOUBLES) {
    prog.cfg = TileCfg{};  // the program keeps the default tiling
    throw Error{TNX_ERR_CAPACITY, "split-K workspace exceeds its cap; use a larger kchunk"};
  }
  // each bin longest task first (LPT); ties keep unit order (tiles of one output together)
  auto cost = [](const GemmTask &t, const std::vector<int64_t> &kt) {
    return std::min<int64_t>(t.kt1, std::max<int64_t>(kt[t.out], 1)) - t.kt0;
  };
  size_t rounds = 0;
  for (int b = 0; b < nbin; ++b) {
    auto &v = bins[b];
    std::stable_sort(v.begin(), v.end(),
                     [&](const GemmTask &x, const GemmTask &y) { return cost(x, kt) > cost(y, kt); });
    // balanced schedule: the second wave (tail pieces, longest first) after all of the first
    auto &w = tails[b];
    std::stable_sort(w.begin(), w.end(),
                     [&](const GemmTask &x, const GemmTask &y) { return cost(x, kt) > cost(y, kt); });
    v.insert(v.end(), w.begin(), w.end());
    rounds = std::max(rounds, v.size());
  }
  std::vector<GemmTask> tasks;
  if (nbin == 1) {
    tasks = std::move(bins[0]);
  } else {
    // workgroup id w is dispatched to XCD w % N_XCD: interleave the bins
    tasks.assign(rounds * nbin, GemmTask{-1, 0, 0, 0, 0, -1, 0, 0});
    for (int b = 0; b < nbin; ++b)
      for (size_t r = 0; r < bins[b].size(); ++r) tasks[r * nbin + b] = bins[b][r];
  }
  TNX_REQUIRE(tasks.size() < (size_t(1) << 31), TNX_ERR_CAPACITY, "too many GEMM tiles");
  prog.vtasks.upload(tasks.data(), tasks.size(), s);
  prog.n_vtasks = (int32_t)tasks.size();
  prog.n_red = (int32_t)red.size();
  if (!red.empty()) {
    prog.red.upload(red.data(), red.size(), s);
    prog.ws.alloc((size_t)slot * bm * bn);
  } else {
    prog.ws.release();
  }
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  prog.tiled = true;
}

// ------------------------------------------------------------------ block transposes
// One block of a block-wise 2D transpose: in[r * cols + c] -> out[c * rows + r].
struct BlockT {
  int64_t in_off, out_off;
  int32_t rows, cols;
};

// One workgroup per (block, 32 x 32 tile): the tile goes through LDS so that
// both the reads and the writes are row-contiguous.
__global__ void __launch_bounds__(256) block_transpose_kernel(const BlockT *__restrict__ list, int32_t n_blocks,
                                                              const int64_t *__restrict__ tile_start,
                                                              const double *__restrict__ in, double *__restrict__ out) {
  __shared__ double t[32][33];
  // find this workgroup's block (tile_start: exclusive prefix of tiles per block)
  int lo = 0, hi = n_blocks - 1;
  const int64_t w = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const BlockT b = list[lo];
  const int64_t tile = w - tile_start[lo];
  const int tc = (b.cols + 31) / 32;
  const int r0 = (int)(tile / tc) * 32, c0 = (int)(tile % tc) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    if (r < b.rows && c < b.cols) t[y][tx] = in[b.in_off + (int64_t)r * b.cols + c];
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + tx;
    if (r < b.rows && c < b.cols) out[b.out_off + (int64_t)c * b.rows + r] = t[tx][y];
  }
}

__global__ void fill_zero_kernel(int64_t n, double *__restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

static void zero_fill(double *p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t g = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)g), dim3(256), 0, s, n, p);
}

// Block-wise transposes of a list (tile starts computed on the host).
struct TransposeList {
  DevBuf<BlockT> list;
  DevBuf<int64_t> start;
  int32_t n = 0;
  int64_t tiles = 0;
  void set(const std::vector<BlockT> &v, hipStream_t s) {
    n = (int32_t)v.size();
    std::vector<int64_t> st(v.size());
    tiles = 0;
    for (size_t i = 0; i < v.size(); ++i) {
      st[i] = tiles;
      tiles += (int64_t)((v[i].rows + 31) / 32) * ((v[i].cols + 31) / 32);
    }
    list.upload(v.data(), v.size(), s);
    start.upload(st.data(), st.size(), s);
  }
  void run(const double *in, double *out, hipStream_t s) const {
    if (!tiles) return;
    TNX_REQUIRE(tiles < (int64_t(1) << 31), TNX_ERR_CAPACITY, "too many transpose tiles");
    hipLaunchKernelGGL(block_transpose_kernel, dim3((unsigned)tiles), dim3(256), 0, s, list.p, n, start.p, in, out);
    TNX_CHECK_HIP(hipGetLastError());
  }
};

// ------------------------------------------------------------------ plan objects
struct PairwisePlan : public tnx_plan_s {
  GemmProgram prog;
  // per-block path: one single-pair program per pair (tasks contiguous per pair)
  DevBuf<GemmOut> pb_outs;
  DevBuf<GemmTask> pb_tasks;
  std::vector<int32_t> pb_task_off, pb_task_cnt;
  std::vector<uint8_t> pb_first;
  bool pb_ready = false;
  int64_t a_total = 0, b_total = 0;
  Extents ea, eb;  // operand block layouts (work-list validation)
  // adjoint (tnx_contract_vjp), derived from the forward pair list on first use
  std::mutex adj_mutex;
  bool adj_ready = false;
  GemmProgram adj_a, adj_b;   // dA = G . B^T, dB = A^T . G
  TransposeList bt_list;      // B blocks transposed into bt (K x N -> N x K)
  TransposeList at_list;      // KM forward: dA^T blocks (M x K) back to A's K x M
  DevBuf<double> bt, at;      // transposed B blocks; dA^T staging (KM forward only)
};

// A pairwise plan's program against its operands' and its output's block layouts.
static void validate_pairwise(const PairwisePlan &P, hipStream_t s) {
  const Extents ec = extents_of(P.out_shapes.data(), P.out_offsets.data(), P.out_nblocks, P.out_rank);
  check_program(P.prog, P.ea, P.eb, ec, "pairwise plan", s);
}

// One block of a leg permutation, with the output dimensions merged where
// they are also adjacent and in order in the input (size-1 dimensions
// dropped), so the index arithmetic runs over as few dimensions as possible,
// in 32 bits (a block holds fewer than 2^31 elements).
//   mode 0 (gather): each thread moves 4 output elements 256 apart; used when
//     the last output dimension is also the input's innermost (s[rank-1] ==
//     1) and at least PERM_RUN long, so both sides move long contiguous runs.
//   mode 1 (tile): one workgroup moves a tile of up to PERM_LDS elements
//     through LDS.  The tile spans the input's innermost dimensions (whole,
//     the outermost of them possibly in part) and the output's innermost
//     dimensions, chosen so that the tile is read in contiguous input runs
//     and written in contiguous output runs of at least PERM_RUN_MIN
//     elements where the block allows; at most PERM_MAXT dimensions are
//     tiled.  Short dimensions (e.g. the 24-wide sectors of a chi = 48 Z2
//     tensor) are combined until the runs are long.
constexpr int PERM_MAXT = 4;
constexpr int PERM_LDS = 4096;  // doubles per tile (32 KB)
struct PDesc {
  int64_t in_off, out_off;
  int64_t task0;       // first workgroup of this block
  int32_t size, rank;  // elements; merged rank
  int32_t neg;         // the block's Koszul sign is -1 (graded transpose)
  int32_t mode;
  int32_t ptile;       // tile mode: elements of a full tile
  int32_t pad;
  int32_t n[MAX_RANK];  // merged output dimensions
  int32_t s[MAX_RANK];  // input stride of each
  int32_t o[MAX_RANK];  // output stride of each
  int32_t t[MAX_RANK];  // tile mode: tile extent of each (1: not tiled)
  int32_t c[MAX_RANK];  // tile mode: tiles along each
  int32_t rd[PERM_MAXT], wd[PERM_MAXT];  // tile mode: tiled dimensions in input / output order, innermost first (-1: none)
  int32_t lt[PERM_MAXT];                 // tile mode: LDS stride of wd[j] (tile stored in output order)
};
constexpr int PERM_EPT = 4;       // gather mode: elements per thread
constexpr int PERM_RUN = 128;     // gather mode needs a contiguous last dimension at least this long
constexpr int PERM_RUN_MIN = 16;  // tile mode: shortest run target tried

struct TransposePlan : public tnx_plan_s {
  DevBuf<PDesc> desc;
  std::vector<PDesc> h_desc;  // host copy (batched execution merges these)
  int32_t n_desc = 0;
  int64_t n_wg = 0;
};

// Several instances' plans bound to their operand buffers and merged into
// one launch: the grouped GEMM of many pairwise plans (pair and output
// offsets rebased onto the lowest operand address of each kind), or the leg
// permutation of many transpose plans.
struct BatchPairwisePlan : public tnx_plan_s {
  GemmProgram prog;
  const double *A = nullptr, *B = nullptr;
  double *C = nullptr;
};
struct BatchTransposePlan : public tnx_plan_s {
  DevBuf<PDesc> desc;
  int32_t n_desc = 0;
  int64_t n_wg = 0;
  const double *in = nullptr;
  double *out = nullptr;
};

static void build_pairwise(PairwisePlan &P, const tnx_blocks *a, const tnx_blocks *b, int a_kfirst, int b_kfirst,
                           int nc, const tnx_leg *out_legs, int sym_n, int mode, bool has_target, int target,
                           hipStream_t s, int n_sign = 0, const int32_t *sign_pairs = nullptr) {
  TNX_REQUIRE(a && b && out_legs, TNX_ERR_ARG, "null argument");
  TNX_REQUIRE(n_sign >= 0 && (n_sign == 0 || sign_pairs), TNX_ERR_ARG, "bad sign-pair list");
  for (int t = 0; t < 2 * n_sign; ++t)
    TNX_REQUIRE(sign_pairs[t] >= 0 && sign_pairs[t] < a->rank + b->rank, TNX_ERR_ARG,
                "sign pair names a leg outside [A legs, B legs]");
  TNX_REQUIRE(nc >= 0 && nc <= a->rank && nc <= b->rank, TNX_ERR_ARG, "bad contracted leg count");
  const int rOut = a->rank + b->rank - 2 * nc;
  TNX_REQUIRE(rOut >= 1 && rOut <= MAX_RANK, TNX_ERR_ARG, "output rank out of range");
  TNX_REQUIRE(mode == TNX_MODE_PAIRWISE_FILTERED || mode == TNX_MODE_MULTI_REACHABLE, TNX_ERR_ARG, "bad mode");
  P.kind = PlanKind::Pairwise;
  P.out_rank = rOut;
  P.prog.a_lay = a_kfirst ? LAY_KM : LAY_MK;
  TNX_REQUIRE(b_kfirst == 1, TNX_ERR_ARG, "B operand legs must be [contracted..., free...] (b_kfirst = 1)");
  P.prog.b_lay = LAY_KN;
  P.ea = extents_of(a->shapes, a->offsets, a->n_blocks, a->rank);
  P.eb = extents_of(b->shapes, b->offsets, b->n_blocks, b->rank);
  for (int i = 0; i < a->n_blocks; ++i) {
    int64_t sz = 1;
    for (int l = 0; l < a->rank; ++l) sz *= a->shapes[i * a->rank + l];
    P.a_total = std::max(P.a_total, a->offsets[i] + sz);
  }
  for (int j = 0; j < b->n_blocks; ++j) {
    int64_t sz = 1;
    for (int l = 0; l < b->rank; ++l) sz *= b->shapes[j * b->rank + l];
    P.b_total = std::max(P.b_total, b->offsets[j] + sz);
  }

  KeySet V;
  enumerate_keys(out_legs, rOut, sym_n, mode == TNX_MODE_MULTI_REACHABLE, has_target, target, s, V);

  PairSpec sp;
  sp.nA = a->n_blocks;
  sp.nB = b->n_blocks;
  sp.rA = a->rank;
  sp.rB = b->rank;
  sp.nc = nc;
  sp.rOut = rOut;
  sp.a_c0 = a_kfirst ? 0 : a->rank - nc;
  sp.a_f0 = a_kfirst ? nc : 0;
  sp.b_c0 = b_kfirst ? 0 : b->rank - nc;
  sp.b_f0 = b_kfirst ? nc : 0;
  const int64_t npc = (int64_t)sp.nA * sp.nB;
  TNX_REQUIRE(npc < (int64_t(1) << 32), TNX_ERR_CAPACITY, "too many candidate block pairs");

  P.out_nblocks = 0;
  P.out_total = 0;
  P.out_keys.clear();
  P.out_shapes.clear();
  P.out_offsets.clear();
  if (npc == 0 || V.n == 0) {
    P.prog.n_outs = 0;
    P.prog.n_pairs = 0;
    finalize_program(P.prog, s);
    return;
  }
  DevBuf<int32_t> kA, kB, sA, sB;
  DevBuf<int64_t> oA, oB;
  kA.upload(a->keys, (size_t)sp.nA * sp.rA, s);
  kB.upload(b->keys, (size_t)sp.nB * sp.rB, s);
  sA.upload(a->shapes, (size_t)sp.nA * sp.rA, s);
  sB.upload(b->shapes, (size_t)sp.nB * sp.rB, s);
  oA.upload(a->offsets, sp.nA, s);
  oB.upload(b->offsets, sp.nB, s);

  DevBuf<int64_t> pair_out(npc), pflag(npc), ppos(npc);
  DevBuf<int32_t> present(V.n);
  DevBuf<int64_t> present64(V.n), out_id(V.n);
  DevBuf<char> tmp;
  TNX_CHECK_HIP(hipMemsetAsync(present.p, 0, V.n * sizeof(int32_t), s));
  hipLaunchKernelGGL(match_pairs_kernel, dim3(nblk(npc)), dim3(256), 0, s, sp, kA.p, kB.p, V.keys.p, V.n,
                     pair_out.p, present.p);
  TNX_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(pair_flags_kernel, dim3(nblk(npc)), dim3(256), 0, s, npc, pair_out.p, pflag.p);
  dev_exclusive_scan(pflag.p, ppos.p, npc, s, tmp);
  hipLaunchKernelGGL(present_to_i64_kernel, dim3(nblk(V.n)), dim3(256), 0, s, V.n, present.p, present64.p);
  dev_exclusive_scan(present64.p, out_id.p, V.n, s, tmp);
  int64_t h4[4];
  TNX_CHECK_HIP(hipMemcpyAsync(&h4[0], ppos.p + npc - 1, 8, hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipMemcpyAsync(&h4[1], pflag.p + npc - 1, 8, hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipMemcpyAsync(&h4[2], out_id.p + V.n - 1, 8, hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipMemcpyAsync(&h4[3], present64.p + V.n - 1, 8, hipMemcpyDeviceToHost, s));
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t npairs = h4[0] + h4[1];
  const int64_t nout = h4[2] + h4[3];
  TNX_REQUIRE(npairs < (int64_t(1) << 31) && nout < (int64_t(1) << 31), TNX_ERR_CAPACITY, "plan too large");
  P.prog.n_pairs = (int32_t)npairs;
  P.prog.n_outs = (int32_t)nout;
  P.out_nblocks = (int32_t)nout;
  if (nout == 0) {
    finalize_program(P.prog, s);
    return;
  }
  DevBuf<uint32_t> vp_oid(npairs), vp_idx(npairs), s_oid(npairs), s_idx(npairs);
  hipLaunchKernelGGL(compact_pairs_kernel, dim3(nblk(npc)), dim3(256), 0, s, npc, pair_out.p, ppos.p, out_id.p,
                     vp_oid.p, vp_idx.p);
  TNX_CHECK_HIP(hipGetLastError());
  dev_sort_pairs(vp_oid.p, s_oid.p, vp_idx.p, s_idx.p, npairs, bits_for((uint64_t)nout), s, tmp);

  DevBuf<int32_t> okeys(nout * rOut), oshapes(nout * rOut), oM(nout), oN(nout), pbeg(nout), pend(nout);
  DevBuf<int64_t> osize(nout), ooff(nout);
  hipLaunchKernelGGL(out_tables_kernel, dim3(nblk(V.n)), dim3(256), 0, s, V.n, rOut, a->rank - nc, present.p,
                     out_id.p, V.keys.p, V.shapes.p, okeys.p, oshapes.p, osize.p, oM.p, oN.p);
  dev_exclusive_scan(osize.p, ooff.p, nout, s, tmp);
  hipLaunchKernelGGL(segments_kernel, dim3(nblk(npairs)), dim3(256), 0, s, npairs, s_oid.p, pbeg.p, pend.p);
  P.prog.pairs.alloc(npairs);
  P.prog.outs.alloc(nout);
  DevBuf<int32_t> sgn;
  if (n_sign) sgn.upload(sign_pairs, 2 * (size_t)n_sign, s);
  hipLaunchKernelGGL(build_pairs_kernel, dim3(nblk(npairs)), dim3(256), 0, s, npairs, sp, a_kfirst, b_kfirst,
                     s_idx.p, sA.p, sB.p, oA.p, oB.p, kA.p, kB.p, sgn.p, n_sign, P.prog.pairs.p);
  hipLaunchKernelGGL(build_outs_kernel, dim3(nblk(nout)), dim3(256), 0, s, nout, ooff.p, oM.p, oN.p, pbeg.p,
                     pend.p, P.prog.outs.p);
  TNX_CHECK_HIP(hipGetLastError());
  finalize_program(P.prog, s);

  P.out_keys.resize(nout * rOut);
  P.out_shapes.resize(nout * rOut);
  P.out_offsets.resize(nout);
  okeys.download(P.out_keys.data(), nout * rOut, s);
  oshapes.download(P.out_shapes.data(), nout * rOut, s);
  ooff.download(P.out_offsets.data(), nout, s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  int64_t total = 0;
  for (int64_t o = 0; o < nout; ++o) {
    int64_t sz = 1;
    for (int l = 0; l < rOut; ++l) sz *= P.out_shapes[o * rOut + l];
    total = std::max(total, P.out_offsets[o] + sz);
  }
  P.out_total = total;
  P.flops = P.prog.flops;
  P.bytes = 8.0 * (double)(P.a_total + P.b_total + P.out_total);
  P.n_pairs = P.prog.n_pairs;
  P.n_tasks = P.prog.n_tasks;
  if (validation_enabled()) validate_pairwise(P, s);
}

// Store the output legs of a built pairwise plan in the order out_perm (output
// leg j = natural leg out_perm[j]; natural = [A free legs..., B free legs...]):
// the output blocks are re-keyed, re-sorted and re-laid out in that order, and
// every output of the program gets a GemmStore so the GEMM writes each element
// straight to its place in the permuted block (no separate transpose).  Sides
// (rows = A free legs, columns = B free legs) merge legs that stay adjacent;
// more than STORE_LEGS legs left on a side is refused (TNX_ERR_ARG).
static void apply_output_perm(PairwisePlan &P, int nfa, const int32_t *perm, hipStream_t s) {
  const int r = P.out_rank;
  TNX_REQUIRE(perm, TNX_ERR_ARG, "null output permutation");
  TNX_REQUIRE(nfa >= 0 && nfa <= r, TNX_ERR_ARG, "bad free-leg split");
  std::vector<int> inv(r, -1);
  bool identity = true;
  for (int j = 0; j < r; ++j) {
    TNX_REQUIRE(perm[j] >= 0 && perm[j] < r && inv[perm[j]] < 0, TNX_ERR_ARG, "out_perm is not a permutation");
    inv[perm[j]] = j;
    identity = identity && perm[j] == j;
  }
  if (identity || P.prog.n_outs == 0) return;
  TNX_REQUIRE(!P.prog.signs, TNX_ERR_ARG, "permuted output stores are not available with fermionic signs");
  const int64_t nout = P.out_nblocks;
  TNX_REQUIRE((int64_t)P.prog.h_outs.size() == nout, TNX_ERR_ARG, "plan has no host mirror of its outputs");
  // permuted keys and shapes, blocks sorted by permuted key
  std::vector<int32_t> pk(nout * r), ps(nout * r);
  for (int64_t o = 0; o < nout; ++o)
    for (int j = 0; j < r; ++j) {
      pk[o * r + j] = P.out_keys[o * r + perm[j]];
      ps[o * r + j] = P.out_shapes[o * r + perm[j]];
    }
  std::vector<int64_t> order(nout);
  for (int64_t o = 0; o < nout; ++o) order[o] = o;
  std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    return std::lexicographical_compare(&pk[x * r], &pk[x * r] + r, &pk[y * r], &pk[y * r] + r);
  });
  std::vector<int64_t> noff(nout);
  std::vector<int32_t> keys2(nout * r), shapes2(nout * r);
  std::vector<int64_t> offs2(nout);
  int64_t at = 0;
  for (int64_t q = 0; q < nout; ++q) {
    const int64_t o = order[q];
    int64_t sz = 1;
    for (int j = 0; j < r; ++j) sz *= ps[o * r + j];
    noff[o] = at;
    offs2[q] = at;
    std::copy(&pk[o * r], &pk[o * r] + r, &keys2[q * r]);
    std::copy(&ps[o * r], &ps[o * r] + r, &shapes2[q * r]);
    at += sz;
  }
  // the stores: natural leg i sits at permuted position inv[i]
  auto side = [&](int64_t o, int l0, int l1, int32_t *dim, int64_t *str) {
    std::vector<int32_t> d;
    std::vector<int64_t> st;
    for (int i = l0; i < l1; ++i) {
      int64_t stride = 1;
      for (int j = inv[i] + 1; j < r; ++j) stride *= ps[o * r + j];
      const int32_t di = P.out_shapes[o * r + i];
      if (di == 1) continue;
      if (!d.empty() && st.back() == stride * di) {  // contiguous with the leg before: merge
        d.back() *= di;
        st.back() = stride;
        continue;
      }
      d.push_back(di);
      st.push_back(stride);
    }
    if (d.empty()) {
      d.push_back(1);
      st.push_back(0);
    }
    TNX_REQUIRE((int)d.size() <= STORE_LEGS, TNX_ERR_ARG,
                "output permutation needs more than 3 row or column legs after merging");
    for (size_t k = 0; k < d.size(); ++k) {
      dim[k] = d[k];
      str[k] = st[k];
    }
    return (int32_t)d.size();
  };
  P.prog.h_stores.assign(nout, GemmStore{});
  for (int64_t o = 0; o < nout; ++o) {
    GemmStore g{};
    g.nr = side(o, 0, nfa, g.rdim, g.rstr);
    g.nc = side(o, nfa, r, g.cdim, g.cstr);
    GemmOut &h = P.prog.h_outs[o];
    int64_t last = 0;
    for (int k = 0; k < g.nr; ++k) last += (int64_t)(g.rdim[k] - 1) * g.rstr[k];
    for (int k = 0; k < g.nc; ++k) last += (int64_t)(g.cdim[k] - 1) * g.cstr[k];
    TNX_REQUIRE((int64_t)h.M * h.N == 0 || last == (int64_t)h.M * h.N - 1, TNX_ERR_ARG,
                "permuted store does not cover its block");
    h.c_off = noff[o];
    P.prog.h_stores[o] = g;
  }
  P.prog.outs.upload(P.prog.h_outs.data(), nout, s);
  P.prog.stores.upload(P.prog.h_stores.data(), nout, s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  P.prog.permuted = true;
  P.out_keys.swap(keys2);
  P.out_shapes.swap(shapes2);
  P.out_offsets.swap(offs2);
}

// Upload a host-built program and let the device build its tile list.
static void finalize_host_program(GemmProgram &g, int a_lay, const std::vector<GemmPair> &pairs,
                                  const std::vector<GemmOut> &outs, hipStream_t s) {
  g.a_lay = a_lay;
  g.b_lay = LAY_KN;
  g.n_pairs = (int32_t)pairs.size();
  g.n_outs = (int32_t)outs.size();
  if (!pairs.empty()) g.pairs.upload(pairs.data(), pairs.size(), s);
  if (!outs.empty()) g.outs.upload(outs.data(), outs.size(), s);
  finalize_program(g, s);
}

// Adjoint programs of a pairwise plan, from its forward pair list
// (C_o = sum over pairs (i, j) -> o of A_i . B_j):
//   dB_j = sum over its pairs A_i^T . G_o   (A read transposed: the other A layout)
//   dA_i = sum over its pairs G_o . B_j^T   (B blocks transposed into bt first)
// For a KM forward (A blocks stored K x M) dA_i^T is formed as M x K in a
// staging buffer and transposed into place.  Output blocks that no pair
// touches stay zero.
static void ensure_adjoint(PairwisePlan &P, hipStream_t s) {
  std::lock_guard<std::mutex> lk(P.adj_mutex);
  if (P.adj_ready) return;
  const auto &ho = P.prog.h_outs;
  const auto &hp = P.prog.h_pairs;
  struct Use {
    int32_t pair, out;
  };
  std::map<int64_t, std::vector<Use>> by_a, by_b;  // block offset -> its pairs, in forward order
  for (int32_t o = 0; o < (int32_t)ho.size(); ++o)
    for (int32_t k = ho[o].pair_begin; k < ho[o].pair_end; ++k)
      if (hp[k].K > 0 && ho[o].M > 0 && ho[o].N > 0) {
        by_a[hp[k].a_off].push_back(Use{k, o});
        by_b[hp[k].b_off].push_back(Use{k, o});
      }
  const bool mk = P.prog.a_lay == LAY_MK;
  // dB: outputs = B blocks (K x N, pitch ldb); A' = A^T, B' = G
  std::vector<GemmPair> pb;
  std::vector<GemmOut> ob;
  std::vector<BlockT> bt;
  for (const auto &kv : by_b) {
    const GemmPair &f = hp[kv.second[0].pair];
    const GemmOut &fo = ho[kv.second[0].out];
    GemmOut g{kv.first, f.K, fo.N, f.ldb, (int32_t)pb.size(), 0, 0};
    for (const Use &u : kv.second) {
      const GemmPair &q = hp[u.pair];
      const GemmOut &qo = ho[u.out];
      TNX_REQUIRE(q.K == f.K && qo.N == fo.N && q.ldb == f.ldb, TNX_ERR_BAD_STRUCTURE,
                  "pairs of one B block disagree on its shape");
      pb.push_back(GemmPair{q.a_off, qo.c_off, qo.M, q.lda, qo.ldc, q.neg});
    }
    g.pair_end = (int32_t)pb.size();
    ob.push_back(g);
    bt.push_back(BlockT{kv.first, kv.first, f.K, f.ldb});  // K x N (pitch = N) -> N x K
  }
  // dA: outputs = A blocks (MK: M x K, pitch lda) or their transposes (KM)
  std::vector<GemmPair> pa;
  std::vector<GemmOut> oa;
  std::vector<BlockT> at;
  for (const auto &kv : by_a) {
    const GemmPair &f = hp[kv.second[0].pair];
    const GemmOut &fo = ho[kv.second[0].out];
    TNX_REQUIRE(mk ? f.lda == f.K : f.lda == fo.M, TNX_ERR_BAD_STRUCTURE, "adjoint needs packed A blocks");
    GemmOut g{kv.first, fo.M, f.K, f.K, (int32_t)pa.size(), 0, 0};
    for (const Use &u : kv.second) {
      const GemmPair &q = hp[u.pair];
      const GemmOut &qo = ho[u.out];
      TNX_REQUIRE(q.K == f.K && qo.M == fo.M, TNX_ERR_BAD_STRUCTURE, "pairs of one A block disagree on its shape");
      // G_o (M x N_o, MK) . B_j^T (N_o x K, stored transposed in bt at the block's offset)
      pa.push_back(GemmPair{qo.c_off, q.b_off, qo.N, qo.ldc, f.K, q.neg});
    }
    g.pair_end = (int32_t)pa.size();
    oa.push_back(g);
    if (!mk) at.push_back(BlockT{kv.first, kv.first, fo.M, f.K});  // M x K -> K x M
  }
  finalize_host_program(P.adj_b, mk ? LAY_KM : LAY_MK, pb, ob, s);
  finalize_host_program(P.adj_a, LAY_MK, pa, oa, s);
  P.bt_list.set(bt, s);
  P.at_list.set(at, s);
  P.bt.alloc(std::max<int64_t>(P.b_total, 1));
  if (!mk) P.at.alloc(std::max<int64_t>(P.a_total, 1));
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  if (validation_enabled()) {
    const Extents ec = extents_of(P.out_shapes.data(), P.out_offsets.data(), P.out_nblocks, P.out_rank);
    check_program(P.adj_b, P.ea, ec, P.eb, "adjoint dB", s);
    check_program(P.adj_a, ec, P.eb, P.ea, "adjoint dA", s);
  }
  P.adj_ready = true;
}

static void ensure_perblock(PairwisePlan &P, hipStream_t s) {
  if (P.pb_ready) return;
  const auto &ho = P.prog.h_outs;
  const auto &hp = P.prog.h_pairs;
  std::vector<GemmOut> outs(hp.size());
  std::vector<GemmTask> tasks;
  P.pb_task_off.assign(hp.size(), 0);
  P.pb_task_cnt.assign(hp.size(), 0);
  P.pb_first.assign(hp.size(), 0);
  for (size_t o = 0; o < ho.size(); ++o) {
    for (int k = ho[o].pair_begin; k < ho[o].pair_end; ++k) {
      GemmOut g = ho[o];
      g.pair_begin = k;
      g.pair_end = k + 1;
      outs[k] = g;
      P.pb_first[k] = (k == ho[o].pair_begin);
      P.pb_task_off[k] = (int32_t)tasks.size();
      for (int m0 = 0; m0 < g.M; m0 += GEMM_BM)
        for (int n0 = 0; n0 < g.N; n0 += GEMM_BN) tasks.push_back(whole_tile(k, m0, n0));
      P.pb_task_cnt[k] = (int32_t)tasks.size() - P.pb_task_off[k];
    }
  }
  P.pb_outs.upload(outs.data(), outs.size(), s);
  P.pb_tasks.upload(tasks.data(), tasks.size(), s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  P.pb_ready = true;
}

// ------------------------------------------------------------------ transpose
__global__ void transpose_code_kernel(int nb, int rank, const int32_t *__restrict__ keys,
                                      const int32_t *__restrict__ perm, const LegDev *__restrict__ legs,
                                      const int32_t *__restrict__ ch, uint64_t *__restrict__ code,
                                      uint32_t *__restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  uint64_t c = 0;
  for (int d = 0; d < rank; ++d) {
    const int l = perm[d];
    const LegDev L = legs[l];
    const int pos = find_charge(ch + L.off, L.n, keys[(int64_t)i * rank + l]);
    c = c * (uint64_t)L.n + (uint64_t)(pos < 0 ? 0 : pos);
  }
  code[i] = c;
  ids[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void permute_kernel(const PDesc *__restrict__ desc, int32_t n_desc,
                                                      const double *__restrict__ in, double *__restrict__ out) {
  __shared__ double lds[PERM_LDS];
  // this workgroup's block: the last with task0 <= blockIdx.x (wave-uniform search)
  const int64_t w = blockIdx.x;
  int lo = 0, hi = n_desc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (desc[mid].task0 <= w) lo = mid; else hi = mid - 1;
  }
  const PDesc &d = desc[lo];
  const int32_t t = (int32_t)(w - d.task0);
  const double sg = d.neg ? -1.0 : 1.0;
  const double *src = in + d.in_off;
  double *dst = out + d.out_off;
  const int rank = d.rank;
  if (d.mode == 0) {
    const int32_t size = d.size;
    int32_t e[PERM_EPT], off[PERM_EPT];
#pragma unroll
    for (int j = 0; j < PERM_EPT; ++j) {
      e[j] = t * (256 * PERM_EPT) + (int32_t)threadIdx.x + 256 * j;
      int32_t r = e[j] < size ? e[j] : 0, x = 0;
      for (int k = rank - 1; k > 0; --k) {
        const int32_t q = r / d.n[k];
        x += (r - q * d.n[k]) * d.s[k];
        r = q;
      }
      off[j] = x + r * d.s[0];
    }
    double v[PERM_EPT];
#pragma unroll
    for (int j = 0; j < PERM_EPT; ++j) v[j] = src[off[j]];  // in bounds: off of element 0 for the idle lanes
#pragma unroll
    for (int j = 0; j < PERM_EPT; ++j)
      if (e[j] < size) dst[e[j]] = sg * v[j];
    return;
  }
  // tile mode: the tile's origin from the workgroup's tile coordinates, and
  // the extent of each tiled dimension (edge tiles are cut at the block's end)
  int32_t rT[PERM_MAXT], rS[PERM_MAXT], rL[PERM_MAXT], rE[PERM_MAXT];
  int32_t wT[PERM_MAXT], wO[PERM_MAXT], wL[PERM_MAXT], wE[PERM_MAXT];
  int32_t rdim[PERM_MAXT], wdim[PERM_MAXT];
#pragma unroll
  for (int j = 0; j < PERM_MAXT; ++j) {
    rdim[j] = d.rd[j];
    wdim[j] = d.wd[j];
    rT[j] = wT[j] = 1;
    rS[j] = wO[j] = rL[j] = wL[j] = 0;
    rE[j] = wE[j] = 1;
  }
  int32_t rem = t, ib = 0, ob = 0;
  for (int k = rank - 1; k >= 0; --k) {
    const int32_t ck = d.c[k], tk = d.t[k];
    const int32_t q = rem / ck;
    const int32_t x0 = (rem - q * ck) * tk;
    rem = q;
    ib += x0 * d.s[k];
    ob += x0 * d.o[k];
    const int32_t ext = min(tk, d.n[k] - x0);
#pragma unroll
    for (int j = 0; j < PERM_MAXT; ++j) {
      if (rdim[j] == k) {
        rT[j] = tk;
        rS[j] = d.s[k];
        rE[j] = ext;
      }
      if (wdim[j] == k) {
        wT[j] = tk;
        wO[j] = d.o[k];
        wE[j] = ext;
        wL[j] = d.lt[j];
      }
    }
  }
  // LDS stride of each read-order dimension: the tile is stored in output order
#pragma unroll
  for (int j = 0; j < PERM_MAXT; ++j)
#pragma unroll
    for (int i = 0; i < PERM_MAXT; ++i)
      if (rdim[j] >= 0 && wdim[i] == rdim[j]) rL[j] = d.lt[i];
  const int32_t P = d.ptile;
  // read in input order (consecutive threads along the input's innermost
  // dimensions), store in output order
#pragma unroll 4
  for (int32_t q = threadIdx.x; q < P; q += 256) {
    int32_t rr = q, off = ib, li = 0;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < PERM_MAXT; ++j) {
      const int32_t qq = rr / rT[j];
      const int32_t l = rr - qq * rT[j];
      rr = qq;
      ok = ok && l < rE[j];
      off += l * rS[j];
      li += l * rL[j];
    }
    if (ok) lds[li] = src[off];
  }
  __syncthreads();
  // write in output order
#pragma unroll 4
  for (int32_t q = threadIdx.x; q < P; q += 256) {
    int32_t rr = q, off = ob, li = 0;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < PERM_MAXT; ++j) {
      const int32_t qq = rr / wT[j];
      const int32_t l = rr - qq * wT[j];
      rr = qq;
      ok = ok && l < wE[j];
      off += l * wO[j];
      li += l * wL[j];
    }
    if (ok) dst[off] = sg * lds[li];
  }
}

// Batched 2-D strided copy (tnx_copy2d_execute): a workgroup copies one
// TNX_COPY2D_CHUNK-element range of one copy's row-major rows x cols
// rectangle, consecutive threads on consecutive elements of a row.
static_assert(sizeof(tnx_copy2d) == 56, "tnx_copy2d layout");
__global__ __launch_bounds__(256) void copy2d_kernel(const tnx_copy2d *__restrict__ desc, int32_t n) {
  const int64_t w = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (desc[mid].task0 <= w) lo = mid; else hi = mid - 1;
  }
  const tnx_copy2d &d = desc[lo];
  const int64_t size = (int64_t)d.rows * d.cols;
  const int64_t e0 = (w - d.task0) * TNX_COPY2D_CHUNK;
  const double sc = d.scale;
#pragma unroll 4
  for (int j = 0; j < TNX_COPY2D_CHUNK / 256; ++j) {
    const int64_t e = e0 + threadIdx.x + 256 * j;
    if (e < size) {
      const int64_t r = e / d.cols, c = e - r * d.cols;
      d.dst[r * d.dst_ld + c] = sc * d.src[r * d.src_ld + c];
    }
  }
}

// Koszul sign of a block under a leg permutation (tensor.py:32-50, 853-856):
// -1 per inversion (i < j, perm[i] > perm[j]) whose two legs carry odd
// charges (parity = charge & 1).
static int koszul_negative(const int32_t *key, const int32_t *perm, int r) {
  int neg = 0;
  for (int i = 0; i < r; ++i)
    for (int j = i + 1; j < r; ++j)
      if (perm[i] > perm[j]) neg ^= (key[perm[i]] & 1) & (key[perm[j]] & 1);
  return neg;
}

// Tile of a tile-mode block (the dimensions of d, merged, in output order,
// with their input strides): returns the number of tiles, or 0 when no tile
// of at most PERM_MAXT dimensions and PERM_LDS elements is found (the block
// then takes the gather mode).  The input's innermost dimensions are taken
// whole until a contiguous input run of `target` elements is reached (the
// last one in part), likewise the output's; smaller targets are tried until
// the tile fits; then the outermost partly taken dimension of either side is
// doubled while the tile stays within PERM_LDS (fewer, larger workgroups).
static int64_t choose_perm_tile(PDesc &d) {
  const int nr = d.rank;
  int in_ord[MAX_RANK], out_ord[MAX_RANK];
  for (int k = 0; k < nr; ++k) in_ord[k] = out_ord[k] = k;
  std::sort(in_ord, in_ord + nr, [&](int x, int y) { return d.s[x] < d.s[y]; });
  std::reverse(out_ord, out_ord + nr);
  int32_t T[MAX_RANK];
  auto count = [&](int64_t &P) {
    int nt = 0;
    P = 1;
    for (int k = 0; k < nr; ++k) {
      P *= T[k];
      if (T[k] > 1) ++nt;
    }
    return nt;
  };
  auto take = [&](const int *ord, int64_t target) {
    int64_t run = 1;
    for (int i = 0; i < nr; ++i) {
      const int k = ord[i];
      if (run * d.n[k] <= target) {
        T[k] = d.n[k];
        run *= d.n[k];
        continue;
      }
      T[k] = std::max<int32_t>(T[k], (int32_t)std::min<int64_t>(d.n[k], (target + run - 1) / run));
      return;
    }
  };
  bool found = false;
  int64_t P = 1;
  for (int64_t target = 128; target >= PERM_RUN_MIN && !found; target /= 2) {
    for (int k = 0; k < nr; ++k) T[k] = 1;
    take(in_ord, target);
    take(out_ord, target);
    found = count(P) <= PERM_MAXT && P <= PERM_LDS;
  }
  if (!found) return 0;
  // grow: the first not-whole dimension of each side, alternately
  for (bool grew = true; grew && 2 * P <= PERM_LDS;) {
    grew = false;
    for (const int *ord : {in_ord, out_ord}) {
      for (int i = 0; i < nr; ++i) {
        const int k = ord[i];
        if (T[k] == d.n[k]) continue;
        const int32_t nt_k = std::min<int32_t>(d.n[k], std::max<int32_t>(2, 2 * T[k]));
        int64_t P2 = 0;
        const int32_t old = T[k];
        T[k] = nt_k;
        if (count(P2) <= PERM_MAXT && P2 <= PERM_LDS) {
          P = P2;
          grew = true;
        } else {
          T[k] = old;
        }
        break;
      }
    }
  }
  int64_t tiles = 1;
  int nt = 0;
  for (int k = 0; k < nr; ++k) {
    d.t[k] = T[k];
    d.c[k] = (d.n[k] + T[k] - 1) / T[k];
    tiles *= d.c[k];
  }
  for (int j = 0; j < PERM_MAXT; ++j) d.rd[j] = d.wd[j] = -1, d.lt[j] = 0;
  for (int i = 0, j = 0; i < nr; ++i)
    if (T[in_ord[i]] > 1) d.rd[j++] = in_ord[i];
  int32_t stride = 1;
  for (int i = 0; i < nr; ++i)
    if (T[out_ord[i]] > 1) {
      d.wd[nt] = out_ord[i];
      d.lt[nt] = stride;
      stride *= T[out_ord[i]];
      ++nt;
    }
  d.ptile = (int32_t)P;
  return tiles;
}

static void build_transpose(TransposePlan &P, const tnx_blocks *a, const tnx_leg *legs, const int32_t *perm,
                            hipStream_t s, bool graded = false) {
  TNX_REQUIRE(a && legs && perm, TNX_ERR_ARG, "null argument");
  const int r = a->rank, nb = a->n_blocks;
  TNX_REQUIRE(r >= 1 && r <= MAX_RANK, TNX_ERR_ARG, "rank out of range");
  std::vector<int> seen(r, 0);
  uint64_t space = 1;
  for (int d = 0; d < r; ++d) {
    TNX_REQUIRE(perm[d] >= 0 && perm[d] < r && !seen[perm[d]], TNX_ERR_ARG, "perm is not a permutation");
    seen[perm[d]] = 1;
    space *= (uint64_t)legs[perm[d]].n_charges;
    TNX_REQUIRE(space < (uint64_t(1) << 62), TNX_ERR_CAPACITY, "key space too large");
  }
  P.kind = PlanKind::Transpose;
  P.out_rank = r;
  P.out_nblocks = nb;
  if (nb == 0) return;
  LegTable lt;
  lt.build(legs, r, s);
  DevBuf<int32_t> kA, dperm;
  kA.upload(a->keys, (size_t)nb * r, s);
  std::vector<int32_t> hp(perm, perm + r);
  dperm.upload(hp.data(), r, s);
  DevBuf<uint64_t> code(nb), code_s(nb);
  DevBuf<uint32_t> ids(nb), order(nb);
  DevBuf<char> tmp;
  hipLaunchKernelGGL(transpose_code_kernel, dim3(nblk(nb)), dim3(256), 0, s, nb, r, kA.p, dperm.p, lt.d_legs.p,
                     lt.d_ch.p, code.p, ids.p);
  TNX_CHECK_HIP(hipGetLastError());
  dev_sort_pairs(code.p, code_s.p, ids.p, order.p, nb, bits_for(space), s, tmp);
  std::vector<uint32_t> h_order(nb);
  order.download(h_order.data(), nb, s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  // output tables in sorted order; per-block descriptors
  P.out_keys.resize((size_t)nb * r);
  P.out_shapes.resize((size_t)nb * r);
  P.out_offsets.resize(nb);
  std::vector<PDesc> desc(nb);
  int64_t off = 0, wg = 0;
  for (int k = 0; k < nb; ++k) {
    const int i = (int)h_order[k];
    PDesc d{};
    d.neg = graded ? koszul_negative(a->keys + (int64_t)i * r, perm, r) : 0;
    d.in_off = a->offsets[i];
    d.out_off = off;
    int64_t st[MAX_RANK];
    int64_t acc = 1;
    for (int l = r - 1; l >= 0; --l) {
      st[l] = acc;
      acc *= a->shapes[(int64_t)i * r + l];
    }
    TNX_REQUIRE(acc < (int64_t(1) << 31) - 4096, TNX_ERR_CAPACITY, "a block of 2^31 - 4096 or more elements");
    // output dimensions with their input strides: size-1 ones dropped,
    // neighbours merged where the input holds them adjacent and in order
    int nr = 0;
    for (int dd = 0; dd < r; ++dd) {
      P.out_keys[(size_t)k * r + dd] = a->keys[(int64_t)i * r + perm[dd]];
      const int32_t n = a->shapes[(int64_t)i * r + perm[dd]];
      P.out_shapes[(size_t)k * r + dd] = n;
      if (n == 1) continue;
      const int32_t sd = (int32_t)st[perm[dd]];
      if (nr > 0 && (int64_t)d.s[nr - 1] == (int64_t)sd * n) {
        d.n[nr - 1] *= n;
        d.s[nr - 1] = sd;
      } else {
        d.n[nr] = n;
        d.s[nr] = sd;
        ++nr;
      }
    }
    if (nr == 0) {
      d.n[0] = 1;
      d.s[0] = 1;
      nr = 1;
    }
    d.rank = nr;
    d.o[nr - 1] = 1;
    for (int l = nr - 2; l >= 0; --l) d.o[l] = d.o[l + 1] * d.n[l + 1];
    d.size = (int32_t)acc;
    d.mode = 0;
    int64_t tasks = (acc + 256 * PERM_EPT - 1) / (256 * PERM_EPT);
    if (acc > 0 && nr > 1 && !(d.s[nr - 1] == 1 && d.n[nr - 1] >= PERM_RUN)) {
      const int64_t nt = choose_perm_tile(d);
      if (nt > 0) {
        d.mode = 1;
        tasks = nt;
      }
    }
    d.task0 = wg;
    wg += tasks;
    P.out_offsets[k] = off;
    desc[k] = d;
    off += acc;
  }
  TNX_REQUIRE(wg < (int64_t(1) << 31), TNX_ERR_CAPACITY, "too many permutation workgroups");
  P.out_total = off;
  P.bytes = 16.0 * off;
  P.desc.upload(desc.data(), nb, s);
  P.h_desc = desc;
  P.n_desc = nb;
  P.n_wg = wg;
  P.n_tasks = (int32_t)wg;
  TNX_CHECK_HIP(hipStreamSynchronize(s));
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int tnx_version(void) { return 1000; }

int tnx_last_error(char *buf, size_t n) {
  if (!buf || n == 0) return TNX_ERR_ARG;
  std::snprintf(buf, n, "%s", g_last_error.c_str());
  return TNX_OK;
}

int tnx_device_sync(void) {
  return guarded([&] { TNX_CHECK_HIP(hipDeviceSynchronize()); });
}

int tnx_valid_blocks(const tnx_leg *legs, int32_t rank, int32_t sym_n, int32_t has_target, int32_t target,
                     int32_t capacity, int32_t *keys_out, int32_t *shapes_out, int32_t *n_out, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(legs && n_out, TNX_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    KeySet V;
    enumerate_keys(legs, rank, sym_n, false, has_target != 0, target, s, V);
    TNX_REQUIRE(V.n < (int64_t(1) << 31), TNX_ERR_CAPACITY, "too many sectors");
    *n_out = (int32_t)V.n;
    if (V.n > capacity) throw Error{TNX_ERR_CAPACITY, "valid block count exceeds capacity"};
    if (V.n) {
      TNX_REQUIRE(keys_out && shapes_out, TNX_ERR_ARG, "null output");
      V.keys.download(keys_out, V.n * rank, s);
      V.shapes.download(shapes_out, V.n * rank, s);
      TNX_CHECK_HIP(hipStreamSynchronize(s));
    }
  });
}

int tnx_plan_pairwise(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst, int32_t b_kfirst,
                      int32_t n_contracted, const tnx_leg *out_legs, int32_t sym_n, int32_t mode,
                      int32_t has_target, int32_t target, void *stream, tnx_plan **plan) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan pointer");
    std::unique_ptr<PairwisePlan> P(new PairwisePlan());
    build_pairwise(*P, a, b, a_kfirst, b_kfirst, n_contracted, out_legs, sym_n, mode, has_target != 0, target,
                   (hipStream_t)stream);
    *plan = P.release();
  });
}

int tnx_plan_pairwise_permuted(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst, int32_t b_kfirst,
                               int32_t n_contracted, const tnx_leg *out_legs, int32_t sym_n, int32_t mode,
                               int32_t has_target, int32_t target, const int32_t *out_perm, void *stream,
                               tnx_plan **plan) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan pointer");
    std::unique_ptr<PairwisePlan> P(new PairwisePlan());
    build_pairwise(*P, a, b, a_kfirst, b_kfirst, n_contracted, out_legs, sym_n, mode, has_target != 0, target,
                   (hipStream_t)stream);
    apply_output_perm(*P, a->rank - n_contracted, out_perm, (hipStream_t)stream);
    if (validation_enabled()) validate_pairwise(*P, (hipStream_t)stream);
    *plan = P.release();
  });
}

int tnx_plan_pairwise_graded(const tnx_blocks *a, const tnx_blocks *b, int32_t a_kfirst, int32_t b_kfirst,
                             int32_t n_contracted, const tnx_leg *out_legs, int32_t sym_n, int32_t mode,
                             int32_t has_target, int32_t target, int32_t n_sign_pairs, const int32_t *sign_pairs,
                             void *stream, tnx_plan **plan) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan pointer");
    std::unique_ptr<PairwisePlan> P(new PairwisePlan());
    build_pairwise(*P, a, b, a_kfirst, b_kfirst, n_contracted, out_legs, sym_n, mode, has_target != 0, target,
                   (hipStream_t)stream, n_sign_pairs, sign_pairs);
    *plan = P.release();
  });
}

int tnx_plan_output_info(const tnx_plan *plan, int32_t *n_blocks, int32_t *rank, int64_t *total_elems) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan");
    if (n_blocks) *n_blocks = plan->out_nblocks;
    if (rank) *rank = plan->out_rank;
    if (total_elems) *total_elems = plan->out_total;
  });
}

int tnx_plan_output_copy(const tnx_plan *plan, int32_t *keys, int32_t *shapes, int64_t *offsets) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan");
    const size_t n = (size_t)plan->out_nblocks, r = (size_t)plan->out_rank;
    if (keys && n) std::memcpy(keys, plan->out_keys.data(), n * r * sizeof(int32_t));
    if (shapes && n) std::memcpy(shapes, plan->out_shapes.data(), n * r * sizeof(int32_t));
    if (offsets && n) std::memcpy(offsets, plan->out_offsets.data(), n * sizeof(int64_t));
  });
}

int tnx_plan_stats(const tnx_plan *plan, double *flops, double *bytes, int32_t *n_pairs, int32_t *n_tasks) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan");
    if (flops) *flops = plan->flops;
    if (bytes) *bytes = plan->bytes;
    if (n_pairs) *n_pairs = plan->n_pairs;
    if (n_tasks) *n_tasks = plan->n_tasks;
  });
}

int tnx_plan_free(tnx_plan *plan) {
  return guarded([&] {
    if (plan)
      TNX_REQUIRE(plan->live_graphs.load() == 0, TNX_ERR_ARG,
                  "the plan still has captured graphs (tnx_lanczos_graph_create); free them first");
    delete plan;
  });
}

static void validate_plan(const tnx_plan *plan, hipStream_t s) {
  switch (plan->kind) {
    case PlanKind::Pairwise: validate_pairwise(static_cast<const PairwisePlan &>(*plan), s); break;
    case PlanKind::Matvec:
    case PlanKind::EnvLeft:
    case PlanKind::EnvRight: validate_fused(plan, s); break;
    default: break;  // transpose plans run no GEMM
  }
}

static GemmProgram *stage_program(tnx_plan *plan, int stage) {
  TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan");
  if (plan->kind == PlanKind::Pairwise) {
    TNX_REQUIRE(stage == 0, TNX_ERR_ARG, "a pairwise plan has one GEMM stage (0)");
    return &static_cast<PairwisePlan &>(*plan).prog;
  }
  if (plan->kind == PlanKind::BatchPairwise) {
    TNX_REQUIRE(stage == 0, TNX_ERR_ARG, "a batch of pairwise plans has one GEMM stage (0)");
    return &static_cast<BatchPairwisePlan &>(*plan).prog;
  }
  TNX_REQUIRE(plan->kind == PlanKind::Matvec || plan->kind == PlanKind::EnvLeft, TNX_ERR_ARG,
              "plan has no grouped-GEMM stage");
  return fused_program(plan, stage);
}

int tnx_plan_set_tiling(tnx_plan *plan, int32_t stage, int32_t shape, int32_t kchunk, int32_t xcd, void *stream) {
  return guarded([&] {
    GemmProgram *g = stage_program(plan, stage);
    TileCfg c;
    c.shape = shape;
    c.kchunk = kchunk;
    c.xcd = xcd ? 1 : 0;
    set_tiling(*g, c, (hipStream_t)stream);
    ++plan->generation;
    if (validation_enabled()) validate_plan(plan, (hipStream_t)stream);
  });
}

int tnx_plan_validate(const tnx_plan *plan, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan");
    validate_plan(plan, (hipStream_t)stream);
  });
}

int tnx_plan_get_tiling(const tnx_plan *plan, int32_t stage, int32_t *shape, int32_t *kchunk, int32_t *xcd,
                        int32_t *n_tasks) {
  return guarded([&] {
    const GemmProgram *g = stage_program(const_cast<tnx_plan *>(plan), stage);
    if (shape) *shape = g->tiled ? g->cfg.shape : 0;
    if (kchunk) *kchunk = g->tiled ? g->cfg.kchunk : 0;
    if (xcd) *xcd = g->tiled ? g->cfg.xcd : 0;
    if (n_tasks) *n_tasks = g->tiled ? g->n_vtasks : g->n_tasks;
  });
}

int tnx_contract_execute(const tnx_plan *plan, const double *a, const double *b, double *c, int32_t compute,
                         void *stream) {
  return guarded([&] {
    TNX_REQUIRE(plan && plan->kind == PlanKind::Pairwise, TNX_ERR_ARG, "not a pairwise plan");
    const PairwisePlan &P = static_cast<const PairwisePlan &>(*plan);
    if (P.prog.n_tasks == 0) return;
    TNX_REQUIRE(a && b && c, TNX_ERR_ARG, "null buffer");
    launch_grouped_gemm(P.prog, a, b, c, compute, false, (hipStream_t)stream);
  });
}

int tnx_contract_execute_perblock(const tnx_plan *plan, const double *a, const double *b, double *c,
                                  int32_t compute, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(plan && plan->kind == PlanKind::Pairwise, TNX_ERR_ARG, "not a pairwise plan");
    PairwisePlan &P = const_cast<PairwisePlan &>(static_cast<const PairwisePlan &>(*plan));
    TNX_REQUIRE(!P.prog.permuted, TNX_ERR_ARG, "the per-block path writes the natural output order only");
    hipStream_t s = (hipStream_t)stream;
    ensure_perblock(P, s);
    TNX_REQUIRE(a && b && c || P.prog.n_pairs == 0, TNX_ERR_ARG, "null buffer");
    for (int k = 0; k < P.prog.n_pairs; ++k)
      launch_prepared(P.prog.a_lay, P.prog.b_lay, P.prog.signs, compute, P.pb_task_cnt[k], s, P.pb_tasks.p + P.pb_task_off[k],
                      P.pb_outs.p, P.prog.pairs.p, a, b, c, !P.pb_first[k]);
  });
}

int tnx_contract_vjp(const tnx_plan *plan, const double *g_c, const double *a, const double *b, double *g_a,
                     double *g_b, int32_t compute, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(plan && plan->kind == PlanKind::Pairwise, TNX_ERR_ARG, "not a pairwise plan");
    PairwisePlan &P = const_cast<PairwisePlan &>(static_cast<const PairwisePlan &>(*plan));
    TNX_REQUIRE(compute >= 0 && compute <= 2, TNX_ERR_DTYPE, "unknown compute mode");
    TNX_REQUIRE(!P.prog.permuted, TNX_ERR_ARG, "no adjoint for a plan with a permuted output order");
    hipStream_t s = (hipStream_t)stream;
    ensure_adjoint(P, s);
    if (g_b) {
      TNX_REQUIRE(g_c && a || P.adj_b.n_pairs == 0, TNX_ERR_ARG, "null buffer");
      zero_fill(g_b, P.b_total, s);
      launch_grouped_gemm(P.adj_b, a, g_c, g_b, compute, false, s);
    }
    if (g_a) {
      TNX_REQUIRE(g_c && b || P.adj_a.n_pairs == 0, TNX_ERR_ARG, "null buffer");
      zero_fill(g_a, P.a_total, s);
      P.bt_list.run(b, P.bt.p, s);
      if (P.prog.a_lay == LAY_MK) {
        launch_grouped_gemm(P.adj_a, g_c, P.bt.p, g_a, compute, false, s);
      } else {
        launch_grouped_gemm(P.adj_a, g_c, P.bt.p, P.at.p, compute, false, s);
        P.at_list.run(P.at.p, g_a, s);
      }
    }
    TNX_CHECK_HIP(hipGetLastError());
  });
}

int tnx_plan_transpose(const tnx_blocks *a, const tnx_leg *legs, const int32_t *perm, void *stream,
                       tnx_plan **plan) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan pointer");
    std::unique_ptr<TransposePlan> P(new TransposePlan());
    build_transpose(*P, a, legs, perm, (hipStream_t)stream);
    *plan = P.release();
  });
}

int tnx_plan_transpose_graded(const tnx_blocks *a, const tnx_leg *legs, const int32_t *perm, int32_t graded,
                              void *stream, tnx_plan **plan) {
  return guarded([&] {
    TNX_REQUIRE(plan, TNX_ERR_ARG, "null plan pointer");
    std::unique_ptr<TransposePlan> P(new TransposePlan());
    build_transpose(*P, a, legs, perm, (hipStream_t)stream, graded != 0);
    *plan = P.release();
  });
}

// Lowest address among the instances' buffers of one kind, and each
// buffer's element offset from it (buffers 8-byte aligned).
static uintptr_t lowest(const void *const *ptr, const std::vector<int> &live) {
  uintptr_t lo = UINTPTR_MAX;
  for (int i : live) {
    TNX_REQUIRE(ptr[i], TNX_ERR_ARG, "null buffer in a batch");
    TNX_REQUIRE(((uintptr_t)ptr[i] & 7) == 0, TNX_ERR_ARG, "batch buffers must be 8-byte aligned");
    lo = std::min(lo, (uintptr_t)ptr[i]);
  }
  return lo;
}

// The output ranges [ptr, ptr + 8 * len) of the instances must not overlap.
static void check_disjoint(const void *const *ptr, const std::vector<int64_t> &len, const std::vector<int> &live) {
  std::vector<std::pair<uintptr_t, uintptr_t>> r;
  for (int i : live)
    if (len[i] > 0) r.push_back({(uintptr_t)ptr[i], (uintptr_t)ptr[i] + 8 * (uintptr_t)len[i]});
  std::sort(r.begin(), r.end());
  for (size_t k = 1; k < r.size(); ++k)
    TNX_REQUIRE(r[k].first >= r[k - 1].second, TNX_ERR_ARG, "the outputs of a batch overlap");
}

int tnx_batch_pairwise(const tnx_plan *const *plans, int32_t n, const double *const *a, const double *const *b,
                       double *const *c, void *stream, tnx_plan **batch) {
  return guarded([&] {
    TNX_REQUIRE(batch && plans && a && b && c && n >= 1, TNX_ERR_ARG, "null argument or empty batch");
    const hipStream_t s = (hipStream_t)stream;
    std::vector<int> live;
    std::vector<int64_t> out_len(n, 0);
    int a_lay = -1;
    for (int i = 0; i < n; ++i) {
      TNX_REQUIRE(plans[i] && plans[i]->kind == PlanKind::Pairwise, TNX_ERR_ARG, "batch entry is not a pairwise plan");
      const PairwisePlan &P = static_cast<const PairwisePlan &>(*plans[i]);
      out_len[i] = P.out_total;
      if (P.prog.n_outs == 0 || P.prog.n_tasks == 0) continue;
      TNX_REQUIRE(a_lay < 0 || a_lay == P.prog.a_lay, TNX_ERR_ARG, "the plans of a batch mix A operand layouts");
      a_lay = P.prog.a_lay;
      live.push_back(i);
    }
    auto B = std::make_unique<BatchPairwisePlan>();
    B->kind = PlanKind::BatchPairwise;
    if (live.empty()) {
      *batch = B.release();
      return;
    }
    check_disjoint((const void *const *)c, out_len, live);
    const uintptr_t amin = lowest((const void *const *)a, live), bmin = lowest((const void *const *)b, live),
                    cmin = lowest((const void *const *)c, live);
    GemmProgram &G = B->prog;
    G.a_lay = a_lay;
    G.b_lay = LAY_KN;
    std::vector<GemmPair> pairs;
    std::vector<GemmOut> outs;
    std::vector<GemmStore> stores;
    bool permuted = false;
    for (int i : live) permuted = permuted || static_cast<const PairwisePlan &>(*plans[i]).prog.permuted;
    for (int i : live) {
      const GemmProgram &g = static_cast<const PairwisePlan &>(*plans[i]).prog;
      if (permuted) {
        if (g.permuted)
          stores.insert(stores.end(), g.h_stores.begin(), g.h_stores.end());
        else
          stores.resize(stores.size() + g.h_outs.size(), GemmStore{});  // nr == 0: plain row-major store
      }
      const int64_t da = (int64_t)(((uintptr_t)a[i] - amin) / 8), db = (int64_t)(((uintptr_t)b[i] - bmin) / 8),
                    dc = (int64_t)(((uintptr_t)c[i] - cmin) / 8);
      const int32_t p0 = (int32_t)pairs.size();
      for (GemmPair p : g.h_pairs) {
        p.a_off += da;
        p.b_off += db;
        pairs.push_back(p);
      }
      for (GemmOut o : g.h_outs) {
        o.c_off += dc;
        o.pair_begin += p0;
        o.pair_end += p0;
        outs.push_back(o);
      }
      TNX_REQUIRE(pairs.size() < (size_t(1) << 31) && outs.size() < (size_t(1) << 31), TNX_ERR_CAPACITY,
                  "batch too large");
    }
    G.n_pairs = (int32_t)pairs.size();
    G.n_outs = (int32_t)outs.size();
    G.pairs.upload(pairs.data(), pairs.size(), s);
    G.outs.upload(outs.data(), outs.size(), s);
    if (permuted) {
      TNX_REQUIRE(stores.size() == outs.size(), TNX_ERR_ARG, "store list does not match the outputs");
      G.h_stores = stores;
      G.stores.upload(G.h_stores.data(), G.h_stores.size(), s);
      G.permuted = true;
    }
    finalize_program(G, s);
    B->A = (const double *)amin;
    B->B = (const double *)bmin;
    B->C = (double *)cmin;
    B->flops = G.flops;
    B->n_pairs = G.n_pairs;
    B->n_tasks = G.n_tasks;
    *batch = B.release();
  });
}

int tnx_batch_transpose(const tnx_plan *const *plans, int32_t n, const double *const *in, double *const *out,
                        void *stream, tnx_plan **batch) {
  return guarded([&] {
    TNX_REQUIRE(batch && plans && in && out && n >= 1, TNX_ERR_ARG, "null argument or empty batch");
    std::vector<int> live;
    std::vector<int64_t> out_len(n, 0);
    for (int i = 0; i < n; ++i) {
      TNX_REQUIRE(plans[i] && plans[i]->kind == PlanKind::Transpose, TNX_ERR_ARG,
                  "batch entry is not a transpose plan");
      const TransposePlan &P = static_cast<const TransposePlan &>(*plans[i]);
      out_len[i] = P.out_total;
      if (P.n_wg > 0) live.push_back(i);
    }
    auto B = std::make_unique<BatchTransposePlan>();
    B->kind = PlanKind::BatchTranspose;
    if (live.empty()) {
      *batch = B.release();
      return;
    }
    check_disjoint((const void *const *)out, out_len, live);
    const uintptr_t imin = lowest((const void *const *)in, live), omin = lowest((const void *const *)out, live);
    std::vector<PDesc> desc;
    int64_t wg = 0;
    for (int i : live) {
      const TransposePlan &P = static_cast<const TransposePlan &>(*plans[i]);
      const int64_t di = (int64_t)(((uintptr_t)in[i] - imin) / 8), dout = (int64_t)(((uintptr_t)out[i] - omin) / 8);
      for (PDesc d : P.h_desc) {
        d.in_off += di;
        d.out_off += dout;
        d.task0 += wg;
        desc.push_back(d);
      }
      wg += P.n_wg;
    }
    TNX_REQUIRE(wg < (int64_t(1) << 31) && desc.size() < (size_t(1) << 31), TNX_ERR_CAPACITY, "batch too large");
    B->desc.upload(desc.data(), desc.size(), (hipStream_t)stream);
    B->n_desc = (int32_t)desc.size();
    B->n_wg = wg;
    B->in = (const double *)imin;
    B->out = (double *)omin;
    B->n_tasks = (int32_t)wg;
    TNX_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    *batch = B.release();
  });
}

int tnx_batch_execute(const tnx_plan *batch, int32_t compute, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(batch, TNX_ERR_ARG, "null plan");
    const hipStream_t s = (hipStream_t)stream;
    if (batch->kind == PlanKind::BatchPairwise) {
      const BatchPairwisePlan &B = static_cast<const BatchPairwisePlan &>(*batch);
      if (B.prog.n_tasks == 0) return;
      launch_grouped_gemm(B.prog, B.A, B.B, B.C, compute, false, s);
    } else if (batch->kind == PlanKind::BatchTranspose) {
      const BatchTransposePlan &B = static_cast<const BatchTransposePlan &>(*batch);
      if (B.n_wg == 0) return;
      hipLaunchKernelGGL(permute_kernel, dim3((unsigned)B.n_wg), dim3(256), 0, s, B.desc.p, B.n_desc, B.in, B.out);
      TNX_CHECK_HIP(hipGetLastError());
    } else {
      throw Error{TNX_ERR_ARG, "not a batch plan"};
    }
  });
}

int tnx_copy2d_execute(const tnx_copy2d *desc, int32_t n, int64_t n_wg, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(n >= 0 && n_wg >= 0, TNX_ERR_ARG, "negative count");
    if (n == 0 || n_wg == 0) return;
    TNX_REQUIRE(desc, TNX_ERR_ARG, "null descriptor array");
    TNX_REQUIRE(n_wg < (int64_t(1) << 31), TNX_ERR_CAPACITY, "too many copy workgroups");
    hipLaunchKernelGGL(copy2d_kernel, dim3((unsigned)n_wg), dim3(256), 0, (hipStream_t)stream, desc, n);
    TNX_CHECK_HIP(hipGetLastError());
  });
}

int tnx_transpose_execute(const tnx_plan *plan, const double *a, double *out, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(plan && plan->kind == PlanKind::Transpose, TNX_ERR_ARG, "not a transpose plan");
    const TransposePlan &P = static_cast<const TransposePlan &>(*plan);
    if (P.n_wg == 0) return;
    TNX_REQUIRE(a && out, TNX_ERR_ARG, "null buffer");
    hipLaunchKernelGGL(permute_kernel, dim3((unsigned)P.n_wg), dim3(256), 0, (hipStream_t)stream, P.desc.p, P.n_desc,
                       a, out);
    TNX_CHECK_HIP(hipGetLastError());
  });
}

}  // extern "C"
