/*
 * tnx_random.h -- initial-state generator of libtnx (C ABI).
 *
 * Threefry-2x32 (20 rounds; Salmon et al., SC'11) with jax.random's key
 * derivation and 64-bit word layouts, the generator the reference draws its
 * initial tensors from (src/tenax/core/tensor.py:655-660
 * jax.random.fold_in/normal, src/tenax/algorithms/dmrg.py:1484 PRNGKey).
 *   PRNGKey(seed) = (seed >> 32, seed & 0xffffffff); fold_in(k, d) = threefry(k, (0, d))
 *   layout 1: word i = threefry(k, (i >> 32, i));  layout 0: (y0 << 32) | y1 with
 *   (y0, y1) = threefry(k, (i, n + i)) for an n-word draw
 *   normal = sqrt(2) * erfinv(u), u = 52 mantissa bits mapped to [nextafter(-1, 0), 1)
 */
#ifndef TNX_RANDOM_H
#define TNX_RANDOM_H

#include "tnx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One Threefry-2x32 block on the host: out2 = threefry((k0, k1), (c0, c1)). */
int tnx_threefry2x32(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t *out2);

/* Packed N(0, stddev) blocks on the device: block b (elements offsets[b] ..
 * offsets[b] + sizes[b] of out, ascending and non-overlapping) is
 * jax.random.normal(key b, (sizes[b],)) * stddev with key b = keys[2b], keys[2b+1]. */
int tnx_random_normal_blocks(int32_t n_blocks, const uint32_t *keys, const int64_t *offsets,
                             const int64_t *sizes, int32_t layout, double stddev, double *out,
                             void *stream);

/* The raw 64-bit words of an n-word draw (device buffer of n uint64). */
int tnx_random_bits(uint32_t k0, uint32_t k1, int64_t n, int32_t layout, uint64_t *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TNX_RANDOM_H */
