// Counter-based initial-state generator: Threefry-2x32 (20 rounds; Salmon,
// Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11)
// with the key derivation and bit layouts of jax.random, which the reference
// uses for its initial tensors (src/tenax/core/tensor.py:655-660,
// src/tenax/algorithms/dmrg.py:1484, ipeps.py:2315-2316):
//   PRNGKey(seed)  = (seed >> 32, seed & 0xffffffff)
//   fold_in(k, d)  = threefry2x32(k, (0, d))
//   64-bit words   layout 1 ("partitionable"): word i = threefry2x32(k, (0, i))
//                  layout 0 ("original"):      word i = (y0_i << 32) | y1_i with
//                                              (y0_i, y1_i) = threefry2x32(k, (i, n + i))
//   uniform double u = bits(52 mantissa) in [0, 1), mapped to [nextafter(-1, 0), 1)
//   normal         = sqrt(2) * erfinv(u)
// erfinv: Giles' single-precision approximation ("Approximating the erfinv
// function", GPU Computing Gems, 2011) refined by Newton steps on erf / erfc
// in double precision.
#include "tnx_internal.h"

#include <algorithm>
#include <cmath>

#include "../../include/tnx_random.h"

namespace tnx {

__host__ __device__ inline uint32_t rotl32(uint32_t v, uint32_t r) { return (v << r) | (v >> (32u - r)); }

__host__ __device__ inline void threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t &o0,
                                             uint32_t &o1) {
  const uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  const uint32_t rot[2][4] = {{13u, 15u, 26u, 6u}, {17u, 29u, 16u, 24u}};
  uint32_t x0 = c0 + ks[0], x1 = c1 + ks[1];
  for (int g = 0; g < 5; ++g) {
    for (int r = 0; r < 4; ++r) {
      x0 += x1;
      x1 = rotl32(x1, rot[g & 1][r]);
      x1 ^= x0;
    }
    x0 += ks[(g + 1) % 3];
    x1 += ks[(g + 2) % 3] + (uint32_t)(g + 1);
  }
  o0 = x0;
  o1 = x1;
}

__device__ inline uint64_t random_word(uint32_t k0, uint32_t k1, int64_t i, int64_t n, int layout) {
  uint32_t a, b;
  if (layout == 1) {
    threefry2x32(k0, k1, (uint32_t)((uint64_t)i >> 32), (uint32_t)i, a, b);
  } else {
    threefry2x32(k0, k1, (uint32_t)i, (uint32_t)(n + i), a, b);
  }
  return ((uint64_t)a << 32) | b;
}

__device__ inline double erfinv_d(double y) {
  double w = -log((1.0 - y) * (1.0 + y));
  double p;
  if (w < 5.0) {
    w -= 2.5;
    p = 2.81022636e-08;
    p = 3.43273939e-07 + p * w;
    p = -3.5233877e-06 + p * w;
    p = -4.39150654e-06 + p * w;
    p = 0.00021858087 + p * w;
    p = -0.00125372503 + p * w;
    p = -0.00417768164 + p * w;
    p = 0.246640727 + p * w;
    p = 1.50140941 + p * w;
  } else {
    w = sqrt(w) - 3.0;
    p = -0.000200214257;
    p = 0.000100950558 + p * w;
    p = 0.00134934322 + p * w;
    p = -0.00367342844 + p * w;
    p = 0.00573950773 + p * w;
    p = -0.0076224613 + p * w;
    p = 0.00943887047 + p * w;
    p = 1.00167406 + p * w;
    p = 2.83297682 + p * w;
  }
  double x = p * y;
  const double ay = fabs(y);
  const double two_over_sqrt_pi = 1.1283791670955125739;
  for (int it = 0; it < 3; ++it) {
    // residual erf(x) - y; in the tails through erfc, where 1 -+ y is exact
    double r;
    if (ay <= 0.5)
      r = erf(x) - y;
    else if (y > 0)
      r = (1.0 - y) - erfc(x);
    else
      r = erfc(-x) - (1.0 + y);
    x -= r / (two_over_sqrt_pi * exp(-x * x));
  }
  return x;
}

__device__ inline double normal_from_word(uint64_t bits) {
  const double f = __longlong_as_double((long long)((bits >> 12) | 0x3FF0000000000000ull)) - 1.0;
  const double lo = -1.0 + 0x1p-53;  // nextafter(-1, 0)
  // f * (max - min) + min, rounded as two operations (no fused multiply-add)
  double u = __dadd_rn(__dmul_rn(f, 1.0 - lo), lo);
  u = u > lo ? u : lo;
  return 1.4142135623730951 * erfinv_d(u);
}

// One key per block: element e of the packed buffer belongs to the block whose
// [off, off + size) contains it; its index inside the block picks the word.
__global__ void random_normal_blocks_kernel(int32_t nb, const uint32_t *__restrict__ keys,
                                            const int64_t *__restrict__ off, const int64_t *__restrict__ size,
                                            int64_t total, int layout, double scale, double *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const int64_t i = e - off[lo];
    if (i >= size[lo]) continue;  // gap between blocks (not produced by packed layouts)
    out[e] = scale * normal_from_word(random_word(keys[2 * lo], keys[2 * lo + 1], i, size[lo], layout));
  }
}

__global__ void random_bits_kernel(uint32_t k0, uint32_t k1, int64_t n, int layout, uint64_t *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = random_word(k0, k1, i, n, layout);
}

}  // namespace tnx

using namespace tnx;

extern "C" {

int tnx_threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t *out2) {
  return guarded([&] {
    TNX_REQUIRE(out2, TNX_ERR_ARG, "null output");
    threefry2x32(k0, k1, c0, c1, out2[0], out2[1]);
  });
}

int tnx_random_normal_blocks(int32_t n_blocks, const uint32_t *keys, const int64_t *offsets, const int64_t *sizes,
                             int32_t layout, double stddev, double *out, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(n_blocks >= 0, TNX_ERR_ARG, "negative block count");
    TNX_REQUIRE(layout == 0 || layout == 1, TNX_ERR_ARG, "layout must be 0 (original) or 1 (partitionable)");
    if (n_blocks == 0) return;
    TNX_REQUIRE(keys && offsets && sizes && out, TNX_ERR_ARG, "null argument");
    int64_t total = 0;
    for (int b = 0; b < n_blocks; ++b) {
      TNX_REQUIRE(sizes[b] >= 0 && offsets[b] >= total, TNX_ERR_ARG, "blocks must be packed in ascending order");
      total = offsets[b] + sizes[b];
    }
    if (total == 0) return;
    hipStream_t s = (hipStream_t)stream;
    DevBuf<uint32_t> dk;
    DevBuf<int64_t> doff, dsz;
    dk.upload(keys, 2 * (size_t)n_blocks, s);
    doff.upload(offsets, n_blocks, s);
    dsz.upload(sizes, n_blocks, s);
    const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(random_normal_blocks_kernel, dim3(grid), dim3(256), 0, s, n_blocks, dk.p, doff.p, dsz.p,
                       total, layout, stddev, out);
    TNX_CHECK_HIP(hipGetLastError());
    TNX_CHECK_HIP(hipStreamSynchronize(s));  // the key/offset tables are freed on return
  });
}

int tnx_random_bits(uint32_t k0, uint32_t k1, int64_t n, int32_t layout, uint64_t *out, void *stream) {
  return guarded([&] {
    TNX_REQUIRE(layout == 0 || layout == 1, TNX_ERR_ARG, "layout must be 0 (original) or 1 (partitionable)");
    if (n <= 0) return;
    TNX_REQUIRE(out, TNX_ERR_ARG, "null output");
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(random_bits_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, k0, k1, n, layout, out);
    TNX_CHECK_HIP(hipGetLastError());
  });
}

}  // extern "C"
