// Host-side validation of the device work lists (debug aid and test
// instrument): every address range a grouped-GEMM launch can touch is checked
// against the block layouts of its operands before anything runs.  Enabled
// for every plan by TNX_VALIDATE=1, and on demand through tnx_plan_validate /
// tnx_check_gemm_worklist.  No counterpart in the reference (its per-block
// einsums index Python dicts).
#include "tnx_internal.h"

#include <algorithm>
#include <cstdlib>

namespace tnx {

void Extents::add(int64_t off, int64_t size) {
  TNX_REQUIRE(off >= 0 && size >= 0, TNX_ERR_BAD_STRUCTURE, "negative block offset or size");
  if (size == 0) return;
  lo.push_back(off);
  hi.push_back(off + size);
  total = std::max(total, off + size);
}

void Extents::finish() {
  std::vector<size_t> ord(lo.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return lo[a] < lo[b]; });
  std::vector<int64_t> l(ord.size()), h(ord.size());
  for (size_t i = 0; i < ord.size(); ++i) {
    l[i] = lo[ord[i]];
    h[i] = hi[ord[i]];
    TNX_REQUIRE(i == 0 || l[i] >= h[i - 1], TNX_ERR_BAD_STRUCTURE, "operand blocks overlap");
  }
  lo.swap(l);
  hi.swap(h);
}

bool Extents::covers(int64_t a, int64_t b) const {
  if (a < 0 || b < a) return false;
  if (a == b) return true;
  auto it = std::upper_bound(lo.begin(), lo.end(), a);
  if (it == lo.begin()) return false;
  const size_t i = (size_t)(it - lo.begin()) - 1;
  return b <= hi[i];
}

Extents extents_of(const int32_t *shapes, const int64_t *offsets, int32_t n, int32_t rank) {
  Extents e;
  for (int32_t i = 0; i < n; ++i) {
    int64_t sz = 1;
    for (int32_t l = 0; l < rank; ++l) sz *= shapes[(int64_t)i * rank + l];
    e.add(offsets[i], sz);
  }
  e.finish();
  return e;
}

static std::string at(const char *kind, int64_t i) { return std::string(kind) + " " + std::to_string(i); }

void check_worklist(int a_lay, const GemmPair *pairs, int64_t n_pairs, const GemmOut *outs, int64_t n_outs,
                    const GemmTask *tasks, int64_t n_tasks, int bm, int bn, int64_t ws_tiles, const Extents &A,
                    const Extents &B, const Extents &C, const std::string &what) {
  auto fail = [&](const std::string &m) { throw Error{TNX_ERR_BAD_STRUCTURE, what + ": " + m}; };
  TNX_REQUIRE(bm > 0 && bn > 0, TNX_ERR_ARG, "tile size must be positive");
  std::vector<int64_t> ktiles(n_outs, 0);
  for (int64_t o = 0; o < n_outs; ++o) {
    const GemmOut &g = outs[o];
    if (g.M < 0 || g.N < 0) fail(at("output", o) + " has a negative extent");
    if (g.M == 0 || g.N == 0) continue;
    if (g.ldc < g.N) fail(at("output", o) + " row pitch below N");
    if (!C.covers(g.c_off, g.c_off + (int64_t)(g.M - 1) * g.ldc + g.N))
      fail(at("output", o) + " range is not inside one block of C");
    if (g.pair_begin < 0 || g.pair_end < g.pair_begin || g.pair_end > n_pairs)
      fail(at("output", o) + " pair range outside the pair list");
    for (int k = g.pair_begin; k < g.pair_end; ++k) {
      const GemmPair &p = pairs[k];
      if (p.K < 0) fail(at("pair", k) + " has negative K");
      if (p.K == 0) continue;
      ktiles[o] += (p.K + GEMM_BK - 1) / GEMM_BK;
      int64_t a_end;
      if (a_lay == LAY_MK) {
        if (p.lda < p.K) fail(at("pair", k) + " A row pitch below K");
        a_end = p.a_off + (int64_t)(g.M - 1) * p.lda + p.K;
      } else {
        if (p.lda < g.M) fail(at("pair", k) + " A row pitch below M");
        a_end = p.a_off + (int64_t)(p.K - 1) * p.lda + g.M;
      }
      if (!A.covers(p.a_off, a_end)) fail(at("pair", k) + " A range is not inside one block of A");
      if (p.ldb < g.N) fail(at("pair", k) + " B row pitch below N");
      if (!B.covers(p.b_off, p.b_off + (int64_t)(p.K - 1) * p.ldb + g.N))
        fail(at("pair", k) + " B range is not inside one block of B");
    }
  }
  for (int64_t t = 0; t < n_tasks; ++t) {
    const GemmTask &q = tasks[t];
    if (q.out < 0) continue;  // padding entry of an interleaved list
    if (q.out >= n_outs) fail(at("task", t) + " names a missing output");
    const GemmOut &g = outs[q.out];
    if (q.m0 < 0 || q.n0 < 0 || q.m0 % bm || q.n0 % bn || q.m0 >= std::max(g.M, 1) || q.n0 >= std::max(g.N, 1))
      fail(at("task", t) + " is not a tile of its output");
    if (q.kt0 < 0 || q.kt1 <= q.kt0) fail(at("task", t) + " has an empty K range");
    if (q.kt0 > 0 && q.kt0 >= ktiles[q.out]) fail(at("task", t) + " starts past its output's K tiles");
    if (q.slot >= 0 && q.kt1 > ktiles[q.out]) fail(at("task", t) + " ends past its output's K tiles");
    if (q.slot >= 0 && q.slot >= ws_tiles) fail(at("task", t) + " writes past the split-K workspace");
    if (q.slot < 0 && q.kt0 != 0) fail(at("task", t) + " covers part of K but has no workspace slot");
  }
}

void check_program(const GemmProgram &prog, const Extents &A, const Extents &B, const Extents &C,
                   const std::string &what, hipStream_t s) {
  if (prog.n_outs == 0) return;
  TNX_REQUIRE((int64_t)prog.h_outs.size() == prog.n_outs && (int64_t)prog.h_pairs.size() == prog.n_pairs,
              TNX_ERR_ARG, "program has no host mirror");
  std::vector<GemmTask> t(std::max<int32_t>(prog.n_tasks, 0));
  if (!t.empty()) prog.tasks.download(t.data(), t.size(), s);
  std::vector<GemmTask> vt(prog.tiled ? std::max<int32_t>(prog.n_vtasks, 0) : 0);
  if (!vt.empty()) prog.vtasks.download(vt.data(), vt.size(), s);
  TNX_CHECK_HIP(hipStreamSynchronize(s));
  check_worklist(prog.a_lay, prog.h_pairs.data(), prog.n_pairs, prog.h_outs.data(), prog.n_outs, t.data(),
                 (int64_t)t.size(), GEMM_BM, GEMM_BN, 0, A, B, C, what + " (default tiles)");
  if (prog.tiled) {
    const int bm = (prog.cfg.shape & 1) ? 128 : 64, bn = (prog.cfg.shape & 2) ? 128 : 64;
    const int64_t ws_tiles = (int64_t)(prog.ws.n / ((size_t)bm * bn));
    check_worklist(prog.a_lay, prog.h_pairs.data(), prog.n_pairs, prog.h_outs.data(), prog.n_outs, vt.data(),
                   (int64_t)vt.size(), bm, bn, ws_tiles, A, B, C, what + " (tiling)");
  }
}

bool validation_enabled() {
  const char *v = std::getenv("TNX_VALIDATE");
  return v && *v && std::string(v) != "0";
}

}  // namespace tnx

using namespace tnx;

extern "C" int tnx_check_gemm_worklist(int32_t a_layout, const void *pairs, int32_t n_pairs, const void *outs,
                                       int32_t n_outs, const void *tasks, int32_t n_tasks, int32_t bm, int32_t bn,
                                       int64_t ws_tiles, const int64_t *a_blocks, int32_t n_a,
                                       const int64_t *b_blocks, int32_t n_b, const int64_t *c_blocks,
                                       int32_t n_c) {
  return guarded([&] {
    TNX_REQUIRE(a_layout == LAY_MK || a_layout == LAY_KM, TNX_ERR_ARG, "a_layout must be 0 (MK) or 1 (KM)");
    TNX_REQUIRE(n_pairs >= 0 && n_outs >= 0 && n_tasks >= 0 && n_a >= 0 && n_b >= 0 && n_c >= 0, TNX_ERR_ARG,
                "negative count");
    auto ext = [](const int64_t *blk, int32_t n) {
      Extents e;
      for (int32_t i = 0; i < n; ++i) e.add(blk[2 * i], blk[2 * i + 1]);
      e.finish();
      return e;
    };
    check_worklist(a_layout, (const GemmPair *)pairs, n_pairs, (const GemmOut *)outs, n_outs,
                   (const GemmTask *)tasks, n_tasks, bm, bn, ws_tiles, ext(a_blocks, n_a), ext(b_blocks, n_b),
                   ext(c_blocks, n_c), "work list");
  });
}
