"""Plan cache: device plans keyed by block structure.

A plan depends only on structure (keys, shapes, layout, output legs), never on
values, so one plan serves every call with the same structure -- e.g. every
Lanczos iteration of a DMRG update.  Mirrors the role of the reference's
shape-keyed expression caches (contractor.py:150-166, 405-457; dmrg.py:1099-1107).
"""

from __future__ import annotations

import threading
from collections import OrderedDict

import numpy as np

from tenax_amd import _lib

_MAX_PLANS = 1024
_MAX_FUSED = 48  # fused DMRG plans own intermediate buffers: keep fewer
_cache: "OrderedDict[tuple, object]" = OrderedDict()
_fused: "OrderedDict[tuple, object]" = OrderedDict()
_lock = threading.Lock()


def _get(key, build):
    store, cap = (_fused, _MAX_FUSED) if key[0] in ("env", "matvec") else (_cache, _MAX_PLANS)
    with _lock:
        hit = store.get(key)
        if hit is not None:
            store.move_to_end(key)
            return hit
    val = build()
    with _lock:
        store[key] = val
        while len(store) > cap:
            store.popitem(last=False)
    return val


def clear_plan_cache():
    with _lock:
        _cache.clear()
        _fused.clear()


def live_pairwise_plans():
    """The cached pairwise contraction plans (most recently used last)."""
    with _lock:
        return [v[0] for k, v in _cache.items() if k[0] == "pair"]


def live_env_plans():
    """The cached environment-update plans (most recently used last)."""
    with _lock:
        return [v[0] for k, v in _fused.items() if k[0] == "env"]


def _struct_sig(t):
    return (t.keys.tobytes(), t.shapes.tobytes(), t.keys.shape)


def _legs_sig(indices):
    return tuple((ix.sectors()[0].tobytes(), ix.sectors()[1].tobytes(), int(ix.flow)) for ix in indices)


def pairwise_plan(A, B, a_kfirst: int, b_kfirst: int, nc: int, out_indices, mode: int, target, sign_pairs=(),
                  out_perm=None):
    """Plan C = A . B (device planner); returns (Plan, keys, shapes, offsets, total).

    ``sign_pairs``: fermionic exchange-sign leg pairs (positions in the
    concatenated [A legs, B legs]), evaluated per block pair on the device.
    ``out_perm``: store output leg j as natural leg out_perm[j] (the natural
    order is [A free legs, B free legs], described by ``out_indices``); the
    GEMM then writes that order directly (tnx_plan_pairwise_permuted; raises
    TnxError when the permutation does not fit the store descriptors)."""
    sym = A.indices[0].symmetry if A.indices else out_indices[0].symmetry
    sign_pairs = tuple((int(x), int(y)) for x, y in sign_pairs)
    out_perm = None if out_perm is None else tuple(int(p) for p in out_perm)
    key = ("pair", _struct_sig(A), _struct_sig(B), A.ndim, B.ndim, a_kfirst, b_kfirst, nc,
           _legs_sig(out_indices), sym.sym_n, mode, target, sign_pairs, out_perm)

    def build():
        lib = _lib.load()
        ta = _lib.BlockTable(A.keys, A.shapes, A.offsets, A.ndim)
        tb = _lib.BlockTable(B.keys, B.shapes, B.offsets, B.ndim)
        legs = _lib.LegArray([(*ix.sectors(), int(ix.flow)) for ix in out_indices])
        ptr = _lib.C.c_void_p()
        has_t = target is not None
        if out_perm is not None:
            if sign_pairs:
                raise ValueError("a permuted output store is not available with fermionic signs")
            pm = np.asarray(out_perm, np.int32)
            rc = lib.tnx_plan_pairwise_permuted(_lib.C.byref(ta.s), _lib.C.byref(tb.s), a_kfirst, b_kfirst, nc,
                                                legs.arr, sym.sym_n, mode, int(has_t), int(target) if has_t else 0,
                                                _lib.i32p(pm), _lib.stream_handle(), _lib.C.byref(ptr))
        else:
            sp = np.asarray(sign_pairs, np.int32).reshape(-1)
            rc = lib.tnx_plan_pairwise_graded(_lib.C.byref(ta.s), _lib.C.byref(tb.s), a_kfirst, b_kfirst, nc,
                                              legs.arr, sym.sym_n, mode, int(has_t), int(target) if has_t else 0,
                                              len(sign_pairs), _lib.i32p(sp) if len(sp) else None,
                                              _lib.stream_handle(), _lib.C.byref(ptr))
        _lib.check(rc)
        plan = _lib.Plan(ptr)
        keys, shapes, offs, tot = plan.output()
        return plan, keys, shapes, offs, tot

    return _get(key, build)


def transpose_plan(T, axes, graded: bool = False):
    """Leg permutation plan; ``graded``: multiply each block by its Koszul sign."""
    key = ("transpose", _struct_sig(T), T.ndim, tuple(axes), _legs_sig(T.indices), bool(graded))

    def build():
        lib = _lib.load()
        ta = _lib.BlockTable(T.keys, T.shapes, T.offsets, T.ndim)
        legs = _lib.LegArray([(*ix.sectors(), int(ix.flow)) for ix in T.indices])
        perm = np.asarray(axes, np.int32)
        ptr = _lib.C.c_void_p()
        _lib.check(lib.tnx_plan_transpose_graded(_lib.C.byref(ta.s), legs.arr, _lib.i32p(perm), int(bool(graded)),
                                                 _lib.stream_handle(), _lib.C.byref(ptr)))
        plan = _lib.Plan(ptr)
        keys, shapes, offs, tot = plan.output()
        return plan, keys, shapes, offs, tot

    return _get(key, build)


def _table(T):
    return _lib.BlockTable(T.keys, T.shapes, T.offsets, T.ndim)


def env_plan(side: int, E, A, W, out_indices):
    """Environment-update plan (left: side 0, right: side 1)."""
    key = ("env", side, _struct_sig(E), _struct_sig(A), _struct_sig(W), _legs_sig(out_indices))

    def build():
        lib = _lib.load()
        te, ta, tw = _table(E), _table(A), _table(W)
        legs = _lib.LegArray([(*ix.sectors(), int(ix.flow)) for ix in out_indices])
        ptr = _lib.C.c_void_p()
        _lib.check(lib.tnx_env_plan(side, _lib.C.byref(te.s), _lib.C.byref(ta.s), _lib.C.byref(tw.s), legs.arr,
                                    _lib.stream_handle(), _lib.C.byref(ptr)))
        plan = _lib.Plan(ptr)
        keys, shapes, offs, tot = plan.output()
        return plan, keys, shapes, offs, tot

    return _get(key, build)


def matvec_plan(L, theta_keys, theta_shapes, theta_offsets, Wl, Wr, R):
    """Fused two-site matvec plan over the full theta sector layout.

    Returns (Plan, reach_src, reach_dst): theta block -> output block edges.
    """
    key = ("matvec", _struct_sig(L), theta_keys.tobytes(), theta_shapes.tobytes(), _struct_sig(Wl),
           _struct_sig(Wr), _struct_sig(R))

    def build():
        lib = _lib.load()
        tl, tt, twl, twr, tr = (_table(L), _lib.BlockTable(theta_keys, theta_shapes, theta_offsets, 4), _table(Wl),
                                _table(Wr), _table(R))
        ptr = _lib.C.c_void_p()
        _lib.check(lib.tnx_matvec_plan(_lib.C.byref(tl.s), _lib.C.byref(tt.s), _lib.C.byref(twl.s),
                                       _lib.C.byref(twr.s), _lib.C.byref(tr.s), _lib.stream_handle(),
                                       _lib.C.byref(ptr)))
        plan = _lib.Plan(ptr)
        ne = _lib.C.c_int32()
        cap = 1
        while True:
            src = np.zeros(cap, np.int32)
            dst = np.zeros(cap, np.int32)
            rc = lib.tnx_matvec_reach(plan.ptr, _lib.C.byref(ne), _lib.i32p(src), _lib.i32p(dst), cap)
            if rc == _lib.TNX_ERR_CAPACITY:
                cap = ne.value
                continue
            _lib.check(rc)
            break
        return plan, src[: ne.value].copy(), dst[: ne.value].copy()

    return _get(key, build)
