"""Two-site DMRG with the block-sparse device backend.

Mirrors reference src/tenax/algorithms/dmrg.py: ``DMRGConfig`` (44-77),
``DMRGResult`` (80-95), ``SweepOps`` (98-110), the sweep schedule of ``dmrg``
(125-305: L->R with left-env updates, rebuild all right envs, R->L with
right-env updates, energy = last local eigenvalue, early stop on
|dE| < convergence_tol after the first sweep), the environment lists (355-398),
``_svd_and_truncate_site`` (808-882), boundary padding (924-975), the
symmetric SweepOps (890-1365) and the builders (1373-1578).

What runs where:
* theta = contract(A_i, A_{i+1}): device plan + grouped block GEMM;
* every matvec of the Lanczos solve: ONE fused device plan
  (L.theta grouped GEMM -> MPO coefficient mix -> .R grouped GEMM), and the
  whole Lanczos loop (FP64 alpha/beta, vector updates, optional
  re-orthogonalisation, tridiagonal eigensolve, Ritz vector) on the device;
* environment updates: fused device plans of the same three-stage form;
* SVD: per-sector FP64 SVD on the device with the global truncation rule.

Krylov vectors live in the full sector layout of the padded two-site tensor
(absent blocks are zero, which changes no sum); the set of blocks the
reference would carry (union of keys, tensor.py:923-952) is tracked from the
plan's reachability edges and applied to the Lanczos result.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass
from typing import Any, NamedTuple

import numpy as np
import torch

from tenax_amd import _lib
from tenax_amd.algorithms.auto_mpo import build_auto_mpo
from tenax_amd.contraction import plans
from tenax_amd.contraction.contractor import contract, get_compute
from tenax_amd.core.index import FlowDirection, TensorIndex
from tenax_amd.core.symmetry import U1Symmetry
from tenax_amd.core.tensor import DenseTensor, PRNGKey, SymmetricTensor, Tensor, _device, valid_blocks
from tenax_amd.linalg import truncated_svd_symmetric
from tenax_amd.network.network import TensorNetwork


@dataclass
class DMRGConfig:
    """dmrg.py:44-77 (``noise`` is kept for signature parity; the reference never reads it).

    Extra fields: ``lanczos_reorth`` (one classical Gram-Schmidt pass per
    Lanczos step, FP64; off like the reference), ``compute`` ("f64" FP64 MFMA,
    "f32", "split_f32"; None = the global contraction setting) and
    ``per_block`` (run the sweep as the plain per-block GPU path, the
    reference's loop structure on the device: one launch per block product in
    matvecs and environment updates, Lanczos scalars through the host, one SVD
    call per sector; the comparison baseline) and ``lanczos_graph`` (each
    two-site update's whole Lanczos solve, fused matvecs included, replayed
    from a captured hipGraph rebound to the update's problem; None = for
    problems of at least GRAPH_MIN_ELEMS theta elements, where the replay
    plus re-recording measured faster than the eager solve).
    """

    max_bond_dim: int = 100
    num_sweeps: int = 10
    convergence_tol: float = 1e-10
    num_states: int = 1
    two_site: bool = True
    lanczos_max_iter: int = 50
    lanczos_tol: float = 1e-12
    noise: float = 0.0
    svd_trunc_err: float | None = None
    target_charge: int | None = None
    verbose: bool = False
    lanczos_reorth: bool = False
    compute: str | None = None
    per_block: bool = False
    lanczos_graph: bool | None = None


class DMRGResult(NamedTuple):
    energy: float
    energies_per_sweep: list
    mps: TensorNetwork
    truncation_errors: list
    converged: bool


class SweepOps(NamedTuple):
    build_trivial_left_env: Callable[..., Tensor]
    build_trivial_right_env: Callable[..., Tensor]
    update_left_env: Callable[[Tensor, Tensor, Tensor], Tensor]
    update_right_env: Callable[[Tensor, Tensor, Tensor], Tensor]
    two_site_update: Callable[..., tuple]
    one_site_update: Callable[..., tuple]


# ------------------------------------------------------------------ helpers
def _compute(config=None) -> int:
    mode = (config.compute if config is not None and config.compute else None) or get_compute()
    return _lib.COMPUTE[mode]


def _assert_symmetric(*tensors, context: str) -> None:
    """dmrg.py:1154-1162."""
    for i, t in enumerate(tensors):
        if not isinstance(t, SymmetricTensor):
            raise TypeError(f"{context}: expected SymmetricTensor for input {i}, got {type(t).__name__}. "
                            f"The symmetric DMRG path must never fall back to dense.")


def _trivial_env(flows, labels) -> SymmetricTensor:
    sym = U1Symmetry()
    z = np.zeros(1, np.int32)
    idx = tuple(TensorIndex(sym, z, f, l) for f, l in zip(flows, labels))
    data = torch.ones(1, dtype=torch.float64, device=_device())
    return SymmetricTensor._wrap(idx, np.zeros((1, 3), np.int32), np.ones((1, 3), np.int32), np.zeros(1, np.int64),
                                 data)


def _build_trivial_left_env_symmetric(dtype=None) -> SymmetricTensor:
    """dmrg.py:890-905."""
    _lib.load()
    return _trivial_env((FlowDirection.IN, FlowDirection.IN, FlowDirection.OUT),
                        ("env_mps_l", "env_mpo_l", "env_mps_conj_l"))


def _build_trivial_right_env_symmetric(dtype=None) -> SymmetricTensor:
    """dmrg.py:908-921."""
    _lib.load()
    return _trivial_env((FlowDirection.OUT, FlowDirection.OUT, FlowDirection.IN),
                        ("env_mps_r", "env_mpo_r", "env_mps_conj_r"))


def _pad_boundary_symmetric(t: SymmetricTensor, pad_left: bool) -> SymmetricTensor:
    """dmrg.py:924-952: add a trivial charge-0 leg (no data movement)."""
    sym = U1Symmetry()
    z = np.zeros(1, np.int32)
    n = t.n_blocks
    if pad_left:
        idx = (TensorIndex(sym, z, FlowDirection.IN, "_pad_l"),) + t.indices
        keys = np.concatenate([np.zeros((n, 1), np.int32), t.keys], axis=1)
        shapes = np.concatenate([np.ones((n, 1), np.int32), t.shapes], axis=1)
    else:
        idx = t.indices + (TensorIndex(sym, z, FlowDirection.OUT, "_pad_r"),)
        keys = np.concatenate([t.keys, np.zeros((n, 1), np.int32)], axis=1)
        shapes = np.concatenate([t.shapes, np.ones((n, 1), np.int32)], axis=1)
    return SymmetricTensor._wrap(idx, keys, shapes, t.offsets, t.data)


def _unpad_boundary_symmetric(t: SymmetricTensor, pad_left: bool) -> SymmetricTensor:
    """dmrg.py:955-975."""
    if pad_left:
        return SymmetricTensor._wrap(t.indices[1:], t.keys[:, 1:], t.shapes[:, 1:], t.offsets, t.data)
    return SymmetricTensor._wrap(t.indices[:-1], t.keys[:, :-1], t.shapes[:, :-1], t.offsets, t.data)


def _pad_site(A: SymmetricTensor) -> SymmetricTensor:
    if A.ndim == 2:
        lab = A.labels()
        return _pad_boundary_symmetric(A, pad_left=isinstance(lab[0], str) and lab[0].startswith("p"))
    return A


_MPO_MIRROR: dict = {}


def _mirrored_mpo(mpo: SymmetricTensor) -> SymmetricTensor:
    """W[e,p,x,b] -> W'[b,p,x,e] (cached per MPO tensor)."""
    key = id(mpo.data)
    hit = _MPO_MIRROR.get(key)
    if hit is None or hit[0] is not mpo.data:
        hit = (mpo.data, mpo.transpose((3, 1, 2, 0)))
        _MPO_MIRROR[key] = hit
    return hit[1]


def _env_update(side: int, env: SymmetricTensor, site: SymmetricTensor, mpo: SymmetricTensor, config=None):
    A = _pad_site(site)
    if side == 0:
        out_idx = (A.indices[2], mpo.indices[3], A.indices[2])
        W = mpo
    else:
        # R[a,b,c] B[d,p,a] W[e,p,x,b] B[f,x,c] is the left form with
        # B' = B(a,p,d) and W' = W(b,p,x,e): same sum, same output legs.
        out_idx = (A.indices[0], mpo.indices[0], A.indices[0])
        A = A.transpose((2, 1, 0))
        W = _mirrored_mpo(mpo)
    plan, keys, shapes, offs, tot = plans.env_plan(0, env, A, W, out_idx)
    out = torch.empty(max(tot, 0), dtype=torch.float64, device=env.data.device)
    lib = _lib.load()
    fn = lib.tnx_env_execute_perblock if (config is not None and config.per_block) else lib.tnx_env_execute
    _lib.check(fn(plan.ptr, _lib.dptr(env.data), _lib.dptr(A.data), _lib.dptr(W.data),
                                   _lib.dptr(out), _compute(config), _lib.stream_handle()))
    return SymmetricTensor._wrap(out_idx, keys, shapes, offs, out)


def _update_left_env_symmetric(left_env, mps_site, mpo_site) -> SymmetricTensor:
    """new_L[d,e,f] = L[a,b,c] A[a,p,d] W[b,p,x,e] A*[c,x,f]  (dmrg.py:1165-1198)."""
    _assert_symmetric(left_env, mps_site, mpo_site, context="_update_left_env_symmetric")
    with _phase("env"):
        return _env_update(0, left_env, mps_site, mpo_site, _ACTIVE_CONFIG.get("cfg"))


def _update_right_env_symmetric(right_env, mps_site, mpo_site) -> SymmetricTensor:
    """new_R[d,e,f] = R[a,b,c] B[d,p,a] W[e,p,x,b] B*[f,x,c]  (dmrg.py:1201-1232)."""
    _assert_symmetric(right_env, mps_site, mpo_site, context="_update_right_env_symmetric")
    with _phase("env"):
        return _env_update(1, right_env, mps_site, mpo_site, _ACTIVE_CONFIG.get("cfg"))


_ACTIVE_CONFIG: dict = {}

# Optional wall-clock accounting of the sweep phases (device-synchronised at
# phase boundaries while enabled; used by the benchmark).
TIMERS: dict = {"enabled": False}
# Krylov steps taken by the most recent two-site update (either path).
LAST_STEPS: dict = {"n": 0}
# Wall-clock seconds of each sweep of the most recent dmrg() call (device
# synchronised at sweep ends only when the phase timers are enabled).
LAST_RUN: dict = {"sweep_seconds": []}
# Progress lines on stderr every `every` two-site updates (0 = silent).
PROGRESS: dict = {"every": 0}


def _progress(sweep, i, direction, energy):
    n = PROGRESS["every"]
    if n and i % n == 0:
        import sys

        print(f"[dmrg] sweep {sweep + 1} {direction} bond {i} E={energy:.12f}", file=sys.stderr, flush=True)


class _phase:
    def __init__(self, name):
        self.name = name

    def __enter__(self):
        if TIMERS["enabled"]:
            import time

            torch.cuda.synchronize()
            self.t0 = time.perf_counter()
        return self

    def __exit__(self, *exc):
        if TIMERS["enabled"]:
            import time

            torch.cuda.synchronize()
            TIMERS[self.name] = TIMERS.get(self.name, 0.0) + time.perf_counter() - self.t0
        return False


def _uniform_target(t: SymmetricTensor):
    flows = np.array([int(ix.flow) for ix in t.indices], np.int64)
    v = (t.keys.astype(np.int64) * flows).sum(axis=1)
    return int(v[0]) if len(v) and np.all(v == v[0]) else None


def _ranges(offsets, shapes, idx, device) -> torch.Tensor:
    """Device index of the concatenated element ranges of blocks ``idx``."""
    from tenax_amd.core.tensor import _block_ranges_index

    return _block_ranges_index(offsets, shapes, idx, device)


class TwoSiteProblem:
    """Device data of one two-site update: full theta layout, matvec plan, reach edges."""

    def __init__(self, theta4: SymmetricTensor, left_env, mpo_l, mpo_r, right_env):
        tgt = _uniform_target(theta4)
        keys, shapes = valid_blocks(theta4.indices, None if not tgt else tgt)
        offs, tot = SymmetricTensor._layout(shapes)
        self.indices = theta4.indices
        self.keys, self.shapes, self.offsets, self.total = keys, shapes, offs, tot
        kidx = {tuple(int(q) for q in k): i for i, k in enumerate(keys)}
        self.theta_pos = np.array([kidx[tuple(int(q) for q in k)] for k in theta4.keys], np.int64)
        self.plan, self.src, self.dst = plans.matvec_plan(left_env, keys, shapes, offs, mpo_l, mpo_r, right_env)
        self.envs = (left_env, mpo_l, mpo_r, right_env)
        dev = theta4.data.device
        x = torch.zeros(tot, dtype=torch.float64, device=dev)
        if theta4.n_blocks:
            di = _ranges(offs, shapes, self.theta_pos, dev)
            x[di] = theta4.data
        self.x = x
        if not self.plan.tuned:
            self._tune()
            self.plan.tuned = True

    def _tune(self):
        """Pick the tiling of both grouped-GEMM stages on this problem's buffers."""
        from tenax_amd.contraction import autotune

        f = (_lib.C.c_double * 4)()
        lib = _lib.load()
        _lib.check(lib.tnx_fused_stage_flops(self.plan.ptr, f))
        if not autotune.ENABLED or max(f[1], f[3]) < autotune.MIN_FLOPS:
            return
        hits = [autotune.cached(st, f[st], "matvec") for st in (1, 3)]
        if all(h is not None for h in hits):
            # structurally similar plan seen before: apply its tilings, no timing runs
            for st, cfg in zip((1, 3), hits):
                self.plan.set_tiling(st, *cfg)
            autotune.STATS["reused"] += 2
            return
        L, Wl, Wr, R = self.envs
        y = torch.empty_like(self.x)

        def stage(st):
            return lambda: _lib.check(lib.tnx_matvec_execute_stage(
                self.plan.ptr, st, _lib.dptr(L.data), _lib.dptr(Wl.data), _lib.dptr(Wr.data), _lib.dptr(R.data),
                _lib.dptr(self.x), _lib.dptr(y), 0, _lib.stream_handle()))

        for st in range(4):  # intermediates in place for the stage-alone runs
            stage(st)()
        autotune.tune(self.plan, 1, stage(1), f[1], "matvec")
        autotune.tune(self.plan, 3, stage(3), f[3], "matvec")

    def present_after(self, n_steps: int) -> np.ndarray:
        present = np.zeros(len(self.keys), bool)
        present[self.theta_pos] = True
        for _ in range(max(0, n_steps - 1)):
            grow = present.copy()
            grow[self.dst[present[self.src]]] = True
            if np.array_equal(grow, present):
                break
            present = grow
        return present

    def matvec(self, x: torch.Tensor, compute: int, per_block: bool = False) -> torch.Tensor:
        L, Wl, Wr, R = self.envs
        y = torch.empty_like(x)
        lib = _lib.load()
        fn = lib.tnx_matvec_execute_perblock if per_block else lib.tnx_matvec_execute
        _lib.check(fn(self.plan.ptr, _lib.dptr(L.data), _lib.dptr(Wl.data), _lib.dptr(Wr.data), _lib.dptr(R.data),
                      _lib.dptr(x), _lib.dptr(y), compute, _lib.stream_handle()))
        return y

    def lanczos(self, max_iter: int, tol: float, reorth: bool, compute: int):
        L, Wl, Wr, R = self.envs
        e = _lib.C.c_double()
        n = _lib.C.c_int32()
        lib = _lib.load()
        _lib.check(lib.tnx_lanczos(self.plan.ptr, _lib.dptr(L.data), _lib.dptr(Wl.data), _lib.dptr(Wr.data),
                                   _lib.dptr(R.data), _lib.dptr(self.x), int(max_iter), float(tol), int(bool(reorth)),
                                   compute, _lib.C.byref(e), _lib.C.byref(n), _lib.stream_handle()))
        return e.value, n.value

    def lanczos_graph(self, max_iter: int, tol: float, reorth: bool, compute: int) -> "LanczosGraph":
        """The device Lanczos solve of :meth:`lanczos` recorded once as a hipGraph
        (tnx_lanczos_graph_create), for repeated solves on this structure."""
        return LanczosGraph(self, max_iter, tol, reorth, compute)

    def extract(self, present: np.ndarray) -> SymmetricTensor:
        sel = np.nonzero(present)[0]
        keys, shapes = self.keys[sel], self.shapes[sel]
        offs, _ = SymmetricTensor._layout(shapes)
        gi = _ranges(self.offsets, self.shapes, sel, self.x.device)
        return SymmetricTensor._wrap(self.indices, keys, shapes, offs, self.x[gi])


class LanczosGraph:
    """A captured device Lanczos solve on a :class:`TwoSiteProblem` (no reference
    counterpart: dmrg.py:978-1045 is a Python loop).  ``launch()`` replays every
    kernel of the solve with one graph launch on the current stream, with the
    problem's vector ``x`` as start vector and output; ``result()`` waits and
    returns (energy, n_steps).  ``update(x=...)`` re-records it for another start
    vector buffer or after the plan's tiling changed.  Bitwise the same as
    :meth:`TwoSiteProblem.lanczos`."""

    def __init__(self, prob: TwoSiteProblem, max_iter: int, tol: float, reorth: bool, compute: int):
        self.prob = prob  # keeps the plan and the operand buffers alive
        self.x = prob.x
        L, Wl, Wr, R = prob.envs
        lib = _lib.load()
        ptr = _lib.C.c_void_p()
        _lib.check(lib.tnx_lanczos_graph_create(prob.plan.ptr, _lib.dptr(L.data), _lib.dptr(Wl.data),
                                                _lib.dptr(Wr.data), _lib.dptr(R.data), _lib.dptr(self.x),
                                                int(max_iter), float(tol), int(bool(reorth)), int(compute),
                                                _lib.stream_handle(), _lib.C.byref(ptr)))
        self.ptr = ptr

    def update(self, x: torch.Tensor | None = None):
        if x is not None:
            if x.shape != self.prob.x.shape or x.dtype != torch.float64 or not x.is_contiguous():
                raise ValueError("start vector must be a contiguous float64 tensor of the problem's size")
            self.x = x
        L, Wl, Wr, R = self.prob.envs
        _lib.check(_lib.load().tnx_lanczos_graph_update(self.ptr, _lib.dptr(L.data), _lib.dptr(Wl.data),
                                                        _lib.dptr(Wr.data), _lib.dptr(R.data), _lib.dptr(self.x),
                                                        _lib.stream_handle()))

    def rebind(self, prob: TwoSiteProblem):
        """Re-record the solve for another problem (the next two-site update of
        a sweep): tnx_lanczos_graph_rebind updates the executable graph in
        place when the launch topology is unchanged."""
        L, Wl, Wr, R = prob.envs
        _lib.check(_lib.load().tnx_lanczos_graph_rebind(self.ptr, prob.plan.ptr, _lib.dptr(L.data),
                                                        _lib.dptr(Wl.data), _lib.dptr(Wr.data),
                                                        _lib.dptr(R.data), _lib.dptr(prob.x),
                                                        _lib.stream_handle()))
        self.prob = prob
        self.x = prob.x

    def launch(self):
        _lib.check(_lib.load().tnx_graph_launch(self.ptr, _lib.stream_handle()))

    def result(self):
        e = _lib.C.c_double()
        n = _lib.C.c_int32()
        _lib.check(_lib.load().tnx_lanczos_graph_result(self.ptr, _lib.stream_handle(), _lib.C.byref(e),
                                                        _lib.C.byref(n)))
        return e.value, n.value

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib.load(require_gpu=False).tnx_graph_free(self.ptr)
            self.ptr = None

    def __del__(self):
        import sys

        if sys.is_finalizing():  # the HIP runtime may already be gone: leave it to process exit
            return
        try:
            self.close()
        except Exception:
            pass


def _theta_padded(site_l, site_r):
    shared = set(site_l.labels()) & set(site_r.labels())
    theta = contract(site_l, site_r) if shared else site_l
    nd = theta.ndim
    pad_left = False
    if nd == 3:
        lab = [ix.label for ix in theta.indices]
        if not any(isinstance(l, str) and l.startswith("v") for l in lab[:1]):
            pad_left = True
            theta = _pad_boundary_symmetric(theta, pad_left=True)
        else:
            theta = _pad_boundary_symmetric(theta, pad_left=False)
    return theta, nd, pad_left


def _two_site_update_symmetric(site_l, site_r, left_env, mpo_l, mpo_r, right_env, config):
    """dmrg.py:1235-1299 with the device Lanczos."""
    _assert_symmetric(site_l, site_r, left_env, mpo_l, mpo_r, right_env, context="_two_site_update_symmetric")
    with _phase("theta"):
        theta, nd, pad_left = _theta_padded(site_l, site_r)
    with _phase("plan"):
        prob = TwoSiteProblem(theta, left_env, mpo_l, mpo_r, right_env)
    with _phase("lanczos"):
        use_graph = config.lanczos_graph
        if use_graph is None:
            use_graph = prob.total >= GRAPH_MIN_ELEMS
        if use_graph:
            energy, n_steps = _graph_solve(prob, config)
        else:
            energy, n_steps = prob.lanczos(config.lanczos_max_iter, config.lanczos_tol, config.lanczos_reorth,
                                           _compute(config))
    LAST_STEPS["n"] = n_steps
    theta_opt = prob.extract(prob.present_after(n_steps))
    if nd == 3:
        theta_opt = _unpad_boundary_symmetric(theta_opt, pad_left=pad_left)
    return theta_opt, energy


# One captured solve per (max_iter, tol, reorth, compute), re-recorded for each
# two-site update of a sweep (lanczos_graph).
_GRAPHS: dict = {}
# Automatic choice (lanczos_graph=None): the captured solve for theta sizes
# from here up.  Measured per two-site update: at chi=2048 (~3e6 elements)
# replay + re-recording 3 ms faster than the eager solve over a 20-bond
# segment; at chi=64 (~3e3) re-recording (0.6-0.9 ms) costs more than the
# replay saves (0.26 ms).
GRAPH_MIN_ELEMS = 1 << 20


def release_graphs():
    """Free the captured solves kept for lanczos_graph (their Krylov workspaces)."""
    for g in _GRAPHS.values():
        g.close()
    _GRAPHS.clear()


def _graph_solve(prob: TwoSiteProblem, config: DMRGConfig):
    """The two-site update's Lanczos solve as a replay of a captured graph
    (tnx_lanczos_graph_*): the graph recorded for the previous update is
    rebound to this problem (an update of the executable graph when the launch
    topology matches) and launched once.  Results are bitwise those of the
    eager solve."""
    key = (int(config.lanczos_max_iter), float(config.lanczos_tol), bool(config.lanczos_reorth), _compute(config))
    g = _GRAPHS.get(key)
    if g is None:
        g = _GRAPHS[key] = LanczosGraph(prob, *key)
    else:
        g.rebind(prob)
    g.launch()
    return g.result()


def _one_site_update_symmetric(*args, **kwargs):
    raise NotImplementedError("one-site DMRG is not on the accelerated path (two_site=True only)")


def _symmetric_ops() -> SweepOps:
    return SweepOps(
        build_trivial_left_env=_build_trivial_left_env_symmetric,
        build_trivial_right_env=_build_trivial_right_env_symmetric,
        update_left_env=_update_left_env_symmetric,
        update_right_env=_update_right_env_symmetric,
        two_site_update=_two_site_update_symmetric,
        one_site_update=_one_site_update_symmetric,
    )


def _svd_and_truncate_site(theta, site: int, config: DMRGConfig, sweep_right: bool = True):
    """dmrg.py:808-882 (s absorbed into the site moving away from the centre)."""
    labels = theta.labels()
    lv = f"v{site - 1}_{site}" if site > 0 else None
    rv = f"v{site + 1}_{site + 2}"
    left = [l for l in labels if l in (lv, f"p{site}") and l is not None]
    right = [l for l in labels if l in (rv, f"p{site + 1}") and l is not None]
    if not left or not right:
        n = len(labels)
        left, right = list(labels[: n // 2]), list(labels[n // 2:])
    bond = f"v{site}_{site + 1}"
    with _phase("svd"):
        A, s, B, s_full = truncated_svd_symmetric(theta, left, right, bond, config.max_bond_dim,
                                                  config.svd_trunc_err, absorb="right" if sweep_right else "left",
                                                  per_sector=config.per_block)
    n_keep = len(s)
    if len(s_full) > n_keep:
        tot = float(torch.sum(s_full ** 2))
        tr = float(torch.sum(s_full[n_keep:] ** 2))
        err = float(np.sqrt(tr / (tot + 1e-15)))
    else:
        err = 0.0
    return A, s, B, err


def _build_left_environments_list(mps, mpo, L, ops):
    envs = [None] * (L + 1)
    envs[0] = ops.build_trivial_left_env()
    for i in range(L - 1):
        envs[i + 1] = ops.update_left_env(envs[i], mps[i], mpo[i])
    return envs


def _build_right_environments_list(mps, mpo, L, ops):
    envs = [None] * (L + 1)
    envs[L] = ops.build_trivial_right_env()
    for i in range(L - 1, 0, -1):
        envs[i] = ops.update_right_env(envs[i + 1], mps[i], mpo[i])
    return envs


def dmrg(hamiltonian: TensorNetwork, initial_mps: TensorNetwork, config: DMRGConfig) -> DMRGResult:
    """Two-site DMRG ground state (dmrg.py:125-305)."""
    L = hamiltonian.n_nodes()
    mps = [initial_mps.get_tensor(i) for i in range(L)]
    mpo = [hamiltonian.get_tensor(i) for i in range(L)]
    all_sym = all(isinstance(t, SymmetricTensor) for t in mps + mpo)
    all_dense = all(isinstance(t, DenseTensor) for t in mps + mpo)
    if not all_sym:
        if all_dense:
            raise NotImplementedError("the dense DMRG backend is not on the accelerated path; "
                                      "use SymmetricTensor MPS and MPO")
        raise TypeError("dmrg() requires uniform tensor types: all DenseTensor or all "
                        "SymmetricTensor. Got mixed types — convert explicitly.")
    if not config.two_site:
        raise NotImplementedError("one-site DMRG is not on the accelerated path (two_site=True only)")
    ops = _symmetric_ops()
    if config.target_charge is not None:
        validate_mps_sector(mps, config.target_charge)
    _ACTIVE_CONFIG["cfg"] = config
    try:
        left = _build_left_environments_list(mps, mpo, L, ops)
        right = _build_right_environments_list(mps, mpo, L, ops)
        energies, errs = [], []
        energy, converged = 0.0, False
        update = _per_block_two_site if config.per_block else ops.two_site_update
        import time as _time

        LAST_RUN["sweep_seconds"] = []
        for sweep in range(config.num_sweeps):
            t_sweep = _time.perf_counter()
            prev = energy
            if sweep > 0:
                left = _build_left_environments_list(mps, mpo, L, ops)
            for i in range(L - 1):
                r = right[i + 2] if right[i + 2] is not None else ops.build_trivial_right_env()
                theta, e = update(mps[i], mps[i + 1], left[i], mpo[i], mpo[i + 1], r, config)
                energy = float(e)
                A, s, B, err = _svd_and_truncate_site(theta, i, config)
                mps[i], mps[i + 1] = A, B
                errs.append(err)
                left[i + 1] = ops.update_left_env(left[i], A, mpo[i])
                _progress(sweep, i, "L->R", energy)
            right = _build_right_environments_list(mps, mpo, L, ops)
            for i in range(L - 2, -1, -1):
                r2 = right[i + 2] if right[i + 2] is not None else ops.build_trivial_right_env()
                theta, e = update(mps[i], mps[i + 1], left[i], mpo[i], mpo[i + 1], r2, config)
                energy = float(e)
                A, s, B, err = _svd_and_truncate_site(theta, i, config, sweep_right=False)
                mps[i], mps[i + 1] = A, B
                errs.append(err)
                right[i + 1] = ops.update_right_env(r2, B, mpo[i + 1])
                _progress(sweep, i, "R->L", energy)
            energies.append(energy)
            LAST_RUN["sweep_seconds"].append(_time.perf_counter() - t_sweep)
            if config.verbose:
                print(f"Sweep {sweep + 1}/{config.num_sweeps}: E = {energy:.10f}")
            if config.target_charge is not None:
                sec = compute_mps_sector(mps)
                if sec != config.target_charge:
                    raise RuntimeError(f"Sector drift detected after sweep {sweep + 1}: "
                                       f"MPS sector={sec}, expected target_charge={config.target_charge}.")
            if sweep > 0 and abs(energy - prev) < config.convergence_tol:
                converged = True
                if config.verbose:
                    print(f"Converged at sweep {sweep + 1}")
                break
    finally:
        _ACTIVE_CONFIG.pop("cfg", None)
        release_graphs()
    out = TensorNetwork(name="DMRG_MPS")
    for i, t in enumerate(mps):
        out.add_node(i, t)
    for i in range(L - 1):
        for lbl in sorted(set(mps[i].labels()) & set(mps[i + 1].labels()), key=str):
            out.connect(i, lbl, i + 1, lbl)
    return DMRGResult(energy, energies, out, errs, converged)


def _per_block_two_site(site_l, site_r, left_env, mpo_l, mpo_r, right_env, config):
    """Two-site update whose matvecs run as the plain per-block GPU path.

    Same Lanczos recurrence as the device solver, driven from the host with
    one launch per block product (the reference's per-combination structure);
    used only as the comparison baseline.
    """
    theta, nd, pad_left = _theta_padded(site_l, site_r)
    prob = TwoSiteProblem(theta, left_env, mpo_l, mpo_r, right_env)
    comp = _compute(config)
    x = prob.x
    m = config.lanczos_max_iter
    V = torch.empty((m + 1, x.numel()), dtype=torch.float64, device=x.device)
    V[0] = x * (1.0 / (float(torch.linalg.vector_norm(x)) + 1e-15))
    alphas, betas = [], [0.0]
    nv = 1
    for step in range(m):
        v = V[nv - 1]
        w = prob.matvec(v, comp, per_block=True)
        a = float(torch.dot(v, w))  # host round trip per scalar, as dmrg.py:1005-1012
        alphas.append(a)
        w = w - v * a
        if step > 0:
            w = w - V[nv - 2] * betas[-1]
        if config.lanczos_reorth:
            c = V[:nv] @ w
            w = w - V[:nv].T @ c
        b = float(torch.linalg.vector_norm(w))
        betas.append(b)
        if b < config.lanczos_tol:
            break
        V[nv] = w * (1.0 / b)
        nv += 1
    basis = [V[i] for i in range(nv)]
    n = len(alphas)
    LAST_STEPS["n"] = n
    if n == 1:
        e, xv = alphas[0], basis[0]
    else:
        T = np.diag(alphas) + np.diag(betas[1:n], 1) + np.diag(betas[1:n], -1)
        ev, evec = np.linalg.eigh(T)
        i = int(np.argmin(ev))
        e = float(ev[i])
        xv = basis[0] * float(evec[0, i])
        for k in range(1, n):
            xv = xv + basis[k] * float(evec[k, i])
        xv = xv * (1.0 / (float(torch.linalg.vector_norm(xv)) + 1e-15))
    prob.x = xv
    theta_opt = prob.extract(prob.present_after(n))
    if nd == 3:
        theta_opt = _unpad_boundary_symmetric(theta_opt, pad_left=pad_left)
    return theta_opt, e


# ------------------------------------------------------------------ builders
def build_mpo_heisenberg(L: int, Jz: float = 1.0, Jxy: float = 1.0, hz: float = 0.0, dtype: Any = None):
    """dmrg.py:1373-1412: U(1)-symmetric XXZ MPO through AutoMPO."""
    terms = []
    for i in range(L - 1):
        terms.append((Jz, "Sz", i, "Sz", i + 1))
        terms.append((Jxy / 2, "Sp", i, "Sm", i + 1))
        terms.append((Jxy / 2, "Sm", i, "Sp", i + 1))
    if hz != 0.0:
        for i in range(L):
            terms.append((hz, "Sz", i))
    if not terms:
        terms.append((0.0, "Sz", 0))
    return build_auto_mpo(terms, L=L, symmetric=True)


def _virtual_charges(bond_dim: int, target_charge: int) -> np.ndarray:
    """dmrg.py:1459-1479."""
    if target_charge == 0:
        req = [-1, 0, 1]
    else:
        req = list(range(min(-1, target_charge - 1), max(1, target_charge + 1) + 1))
    per = max(1, bond_dim // len(req))
    v = np.concatenate([np.full(per, q, np.int32) for q in req])[:bond_dim]
    if len(v) < bond_dim:
        v = np.concatenate([v, np.full(bond_dim - len(v), req[len(req) // 2], np.int32)])
    return v


def build_random_symmetric_mps(L: int, bond_dim: int = 4, dtype: Any = None, seed: int = 42,
                               target_charge: int = 0) -> TensorNetwork:
    """dmrg.py:1415-1523; site i draws from PRNGKey(seed + i)."""
    if target_charge % 2 != L % 2:
        raise ValueError(f"target_charge={target_charge} has parity {target_charge % 2} but "
                         f"L={L} has parity {L % 2}. Each site contributes ±1, so total "
                         f"charge must have the same parity as L.")
    sym = U1Symmetry()
    phys = np.array([1, -1], np.int32)
    virt = _virtual_charges(bond_dim, target_charge)
    mps = TensorNetwork(name=f"symmetric_MPS_L{L}")
    for i in range(L):
        tgt = target_charge if i == L - 1 else None
        if i == 0:
            idx = (TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"),
                   TensorIndex(sym, virt, FlowDirection.OUT, f"v{i}_{i + 1}"))
        elif i == L - 1:
            idx = (TensorIndex(sym, virt, FlowDirection.IN, f"v{i - 1}_{i}"),
                   TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"))
        else:
            idx = (TensorIndex(sym, virt, FlowDirection.IN, f"v{i - 1}_{i}"),
                   TensorIndex(sym, phys, FlowDirection.IN, f"p{i}"),
                   TensorIndex(sym, virt, FlowDirection.OUT, f"v{i}_{i + 1}"))
        mps.add_node(i, SymmetricTensor.random_normal(idx, PRNGKey(seed + i), target=tgt))
    for i in range(L - 1):
        mps.connect(i, f"v{i}_{i + 1}", i + 1, f"v{i}_{i + 1}")
    return mps


def compute_mps_sector(mps_tensors) -> int | None:
    """dmrg.py:1526-1556."""
    r = mps_tensors[-1]
    if not isinstance(r, SymmetricTensor) or r.n_blocks == 0:
        return None
    return _uniform_target(r)


def validate_mps_sector(mps_tensors, target_charge: int) -> None:
    """dmrg.py:1559-1578."""
    sec = compute_mps_sector(mps_tensors)
    if sec is None:
        raise ValueError(f"Cannot determine MPS sector (mixed or no SymmetricTensor blocks). "
                         f"Expected target_charge={target_charge}.")
    if sec != target_charge:
        raise ValueError(f"MPS sector {sec} does not match target_charge={target_charge}.")
