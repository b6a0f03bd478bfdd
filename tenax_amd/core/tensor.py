"""Tensor protocol, DenseTensor and the packed block-sparse SymmetricTensor.

Mirrors reference src/tenax/core/tensor.py:194-974.  The storage model is the
MI355X one: a ``SymmetricTensor`` is ONE contiguous float64 device buffer
(torch tensor on the GPU) holding its blocks in lexicographically sorted key
order, each block row-major in its own leg order, plus host tables
(keys int32 [n_blocks, rank], shapes, int64 offsets).  ``.blocks`` is a dict
*view* of device slices, so code written against the reference's
``dict[BlockKey, array]`` keeps working.

Sector enumeration for the factories runs on the device
(``tnx_valid_blocks``, replacing ``_compute_valid_blocks``, tensor.py:53-165).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from tenax_amd import _lib
from tenax_amd.core.index import Label, TensorIndex

BlockKey = tuple[int, ...]


def _device():
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


# ------------------------------------------------------------------ generator
# jax.random's 64-bit word layout: 1 = "partitionable" (jax's default since
# 0.5), 0 = the original layout of earlier jax versions.
RANDOM_LAYOUT = 1


class PRNGKey:
    """Seed words of the initial-state generator.

    ``PRNGKey(seed)`` and ``fold_in(key, i)`` mirror jax.random.PRNGKey/fold_in
    as the reference uses them (tensor.py:655-660, dmrg.py:1484): the words
    (seed, i1, i2, ...) stand for PRNGKey(seed) folded with i1, i2, ...; the
    Threefry-2x32 key is derived by libtnx (``key``), and the normals are
    drawn on the device (``tnx_random_normal_blocks``).
    """

    __slots__ = ("words",)

    def __init__(self, seed):
        self.words = tuple(int(w) for w in (seed if isinstance(seed, (tuple, list)) else (seed,)))

    @property
    def key(self) -> tuple[int, int]:
        lib = _lib.load(require_gpu=False)
        s = self.words[0]
        k = ((s >> 32) & 0xFFFFFFFF, s & 0xFFFFFFFF)
        out = (_lib.C.c_uint32 * 2)()
        for w in self.words[1:]:
            _lib.check(lib.tnx_threefry2x32(k[0], k[1], 0, int(w) & 0xFFFFFFFF, out))
            k = (int(out[0]), int(out[1]))
        return k

    def __repr__(self):
        return f"PRNGKey{self.words}"


def fold_in(key: PRNGKey, i: int) -> PRNGKey:
    return PRNGKey(key.words + (int(i),))


def _draw_blocks(keys, sizes, stddev: float = 1.0) -> torch.Tensor:
    """Packed N(0, stddev) blocks, block b from key ``keys[b]`` (device)."""
    lib = _lib.load()
    sizes = np.asarray(sizes, np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes) else sizes
    out = torch.empty(int(sizes.sum()), dtype=torch.float64, device=_device())
    if len(sizes) and out.numel():
        kk = np.ascontiguousarray(np.asarray([k.key for k in keys], np.uint32).reshape(-1))
        _lib.check(lib.tnx_random_normal_blocks(len(sizes), kk.ctypes.data_as(_lib.C.POINTER(_lib.C.c_uint32)),
                                                _lib.i64p(offs), _lib.i64p(sizes), RANDOM_LAYOUT, float(stddev),
                                                _lib.dptr(out), _lib.stream_handle()))
    return out


def normal(key: PRNGKey, shape) -> np.ndarray:
    """jax.random.normal(key, shape, float64) drawn on the device (host copy)."""
    key = key if isinstance(key, PRNGKey) else PRNGKey(key)
    n = int(np.prod(shape)) if len(shape) else 1
    return _draw_blocks([key], [n]).cpu().numpy().reshape(shape)


# ------------------------------------------------------------------ sectors
def valid_blocks(indices: tuple[TensorIndex, ...], target: int | None = None):
    """Sorted valid keys and block shapes, enumerated on the device.

    Same set as reference _compute_valid_blocks (tensor.py:53-165); returned
    in the sorted order the reference uses for storage (tensor.py:576, 651).
    """
    r = len(indices)
    if r == 0:
        return np.zeros((1, 0), np.int32), np.zeros((1, 0), np.int32)
    lib = _lib.load()
    sym = indices[0].symmetry
    legs = _lib.LegArray([(*ix.sectors(), int(ix.flow)) for ix in indices])
    cap = 1
    for ix in indices:
        cap *= len(ix.sectors()[0])
    cap = max(1, min(cap, 1 << 26))
    keys = np.zeros((cap, r), np.int32)
    shapes = np.zeros((cap, r), np.int32)
    n = _lib.C.c_int32()
    has_t = target is not None
    rc = lib.tnx_valid_blocks(legs.arr, r, sym.sym_n, int(has_t), int(target) if has_t else 0, cap,
                              _lib.i32p(keys), _lib.i32p(shapes), _lib.C.byref(n), _lib.stream_handle())
    _lib.check(rc)
    return keys[: n.value].copy(), shapes[: n.value].copy()


def _conserved_target(indices, key) -> int:
    return sum(int(ix.flow) * int(q) for ix, q in zip(indices, key))


# ------------------------------------------------------------------ protocol
class Tensor:
    """Duck-typed tensor protocol (tensor.py:194-293)."""

    @property
    def indices(self) -> tuple[TensorIndex, ...]:
        raise NotImplementedError

    @property
    def ndim(self) -> int:
        return len(self.indices)

    def labels(self) -> tuple[Label, ...]:
        return tuple(ix.label for ix in self.indices)

    def __rmul__(self, scalar):
        return self.__mul__(scalar)

    def __sub__(self, other):
        return self.__add__(other.__mul__(-1.0))

    def __neg__(self):
        return self.__mul__(-1.0)


def _relabel_indices(indices, mapping):
    return tuple(ix.relabel(mapping[ix.label]) if ix.label in mapping else ix for ix in indices)


class DenseTensor(Tensor):
    """Dense device tensor with leg metadata (tensor.py:325-499)."""

    def __init__(self, data, indices):
        t = data if isinstance(data, torch.Tensor) else torch.as_tensor(np.asarray(data))
        if t.dtype not in (torch.float64, torch.float32, torch.complex128, torch.complex64):
            t = t.to(torch.float64)
        t = t.to(_device())
        if t.ndim != len(indices):
            raise ValueError(f"data has {t.ndim} dims but {len(indices)} indices given")
        for i, (d, ix) in enumerate(zip(t.shape, indices)):
            if d != ix.dim:
                raise ValueError(f"data.shape[{i}]={d} but indices[{i}].dim={ix.dim}")
        self._data = t
        self._indices = tuple(indices)

    @property
    def indices(self):
        return self._indices

    @property
    def dtype(self):
        return self._data.dtype

    def todense(self):
        return self._data

    def conj(self):
        return DenseTensor(self._data.conj(), self._indices)

    def dagger(self):
        return DenseTensor(self._data.conj(), tuple(ix.dual() for ix in self._indices))

    def bar(self):
        return DenseTensor(self._data.conj(), tuple(ix.flip_flow() for ix in self._indices))

    def transpose(self, axes):
        return DenseTensor(self._data.permute(*axes).contiguous(), tuple(self._indices[i] for i in axes))

    def norm(self):
        return torch.linalg.vector_norm(self._data)

    def relabel(self, old, new):
        if old not in self.labels():
            raise KeyError(f"Label {old!r} not found in tensor with labels {self.labels()}")
        return DenseTensor(self._data, _relabel_indices(self._indices, {old: new}))

    def relabels(self, mapping):
        return DenseTensor(self._data, _relabel_indices(self._indices, mapping))

    def __mul__(self, scalar):
        return DenseTensor(self._data * scalar, self._indices)

    def __add__(self, other):
        o = other.todense() if not isinstance(other, DenseTensor) else other._data
        return DenseTensor(self._data + o, self._indices)

    def max_abs(self):
        return self._data.abs().max()

    def __repr__(self):
        return f"DenseTensor(shape={tuple(self._data.shape)}, dtype={self.dtype}, labels={self.labels()})"


class SymmetricTensor(Tensor):
    """Packed block-sparse tensor (tensor.py:505-974).

    Construction from a ``dict[key -> array]`` validates conservation like the
    reference (tensor.py:551-568) and packs the blocks onto the device.
    """

    __slots__ = ("_indices", "_keys", "_shapes", "_offsets", "_data", "_kidx")

    def __init__(self, blocks: dict, indices):
        self._indices = tuple(indices)
        self._validate_keys(blocks.keys())
        keys = sorted(tuple(int(q) for q in k) for k in blocks)
        r = len(self._indices)
        arrays = []
        shapes = np.zeros((len(keys), r), np.int32)
        for i, k in enumerate(keys):
            a = blocks[k] if k in blocks else blocks[_find_key(blocks, k)]
            a = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, dtype=np.float64))
            shapes[i] = a.shape
            arrays.append(a.reshape(-1).to(device=_device(), dtype=torch.float64))
        sizes = shapes.prod(axis=1).astype(np.int64) if len(keys) else np.zeros(0, np.int64)
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(keys) else sizes
        data = torch.cat(arrays) if arrays else torch.zeros(0, dtype=torch.float64, device=_device())
        self._keys = np.asarray(keys, np.int32).reshape(len(keys), r)
        self._shapes = shapes
        self._offsets = offsets
        self._data = data
        self._kidx = None

    def _validate_keys(self, keys):
        if not self._indices:
            return
        sym = self._indices[0].symmetry
        ident = sym.identity()
        n = sym.n_values()
        for k in keys:
            v = _conserved_target(self._indices, k)
            if n is not None:
                v %= n
            if v != ident:
                raise ValueError(f"Block {tuple(k)} violates charge conservation: "
                                 f"fused={v}, expected identity={ident}")

    # ------------------------------------------------ packed constructors
    @classmethod
    def _wrap(cls, indices, keys, shapes, offsets, data) -> "SymmetricTensor":
        obj = object.__new__(cls)
        obj._indices = tuple(indices)
        obj._offsets = np.asarray(offsets, np.int64).reshape(-1)
        n = len(obj._offsets)  # one offset per block (rank 0: the scalar block)
        obj._keys = np.asarray(keys, np.int32).reshape(n, len(obj._indices))
        obj._shapes = np.asarray(shapes, np.int32).reshape(n, len(obj._indices))
        obj._data = data
        obj._kidx = None
        return obj

    @staticmethod
    def _layout(shapes: np.ndarray):
        sizes = shapes.prod(axis=1).astype(np.int64) if len(shapes) else np.zeros(0, np.int64)
        offsets = np.zeros(len(sizes), np.int64)
        if len(sizes):
            offsets[1:] = np.cumsum(sizes)[:-1]
        return offsets, int(sizes.sum())

    def _with_data(self, data, indices=None) -> "SymmetricTensor":
        return SymmetricTensor._wrap(self._indices if indices is None else indices, self._keys, self._shapes,
                                     self._offsets, data)

    # ------------------------------------------------ factories
    @classmethod
    def zeros(cls, indices, dtype: Any = torch.float64, target: int | None = None):
        keys, shapes = valid_blocks(tuple(indices), target)
        offs, tot = cls._layout(shapes)
        return cls._wrap(indices, keys, shapes, offs, torch.zeros(tot, dtype=torch.float64, device=_device()))

    @classmethod
    def random_normal(cls, indices, key, dtype: Any = torch.float64, stddev: float = 1.0,
                      target: int | None = None):
        """Block i of the sorted valid keys draws from fold_in(key, i) (tensor.py:629-668)."""
        key = key if isinstance(key, PRNGKey) else PRNGKey(key)
        keys, shapes = valid_blocks(tuple(indices), target)
        offs, tot = cls._layout(shapes)
        sizes = np.prod(shapes, axis=1).astype(np.int64) if len(keys) else np.zeros(0, np.int64)
        data = _draw_blocks([fold_in(key, i) for i in range(len(keys))], sizes, stddev)
        return cls._wrap(indices, keys, shapes, offs, data)

    @classmethod
    def from_dense(cls, data, indices, tol: float = 1e-12):
        """Gather the valid sectors of a dense array (tensor.py:670-732)."""
        indices = tuple(indices)
        dense = data if isinstance(data, torch.Tensor) else torch.as_tensor(np.asarray(data, dtype=np.float64))
        dense = dense.to(device=_device(), dtype=torch.float64)
        if tuple(dense.shape) != tuple(ix.dim for ix in indices):
            raise ValueError(f"data.shape {tuple(dense.shape)} does not match index dims "
                             f"{tuple(ix.dim for ix in indices)}")
        keys, shapes = valid_blocks(indices)
        offs, tot = cls._layout(shapes)
        gidx = _dense_gather_index(indices, keys, shapes)
        flat = dense.reshape(-1)
        out = flat[gidx] if tot else torch.zeros(0, dtype=torch.float64, device=flat.device)
        if tol != float("inf"):
            mask = torch.ones(flat.numel(), dtype=torch.bool, device=flat.device)
            if tot:
                mask[gidx] = False
            bad = flat[mask].abs() > tol
            nbad = int(bad.sum().item()) if flat.numel() else 0
            if nbad:
                mx = float(flat[mask].abs().max().item())
                raise ValueError(f"data has {nbad} non-zero elements outside symmetry-allowed sectors "
                                 f"(max abs value: {mx:.3e})")
        return cls._wrap(indices, keys, shapes, offs, out.contiguous())

    # ------------------------------------------------ protocol
    @property
    def indices(self):
        return self._indices

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def n_blocks(self) -> int:
        return int(self._keys.shape[0])

    @property
    def data(self) -> torch.Tensor:
        """The packed device buffer (float64, 1-D)."""
        return self._data

    @property
    def keys(self) -> np.ndarray:
        return self._keys

    @property
    def shapes(self) -> np.ndarray:
        return self._shapes

    @property
    def offsets(self) -> np.ndarray:
        return self._offsets

    def key_index(self) -> dict:
        if self._kidx is None:
            self._kidx = {tuple(int(q) for q in k): i for i, k in enumerate(self._keys)}
        return self._kidx

    def block(self, i: int) -> torch.Tensor:
        shp = tuple(int(s) for s in self._shapes[i])
        o = int(self._offsets[i])
        return self._data[o: o + int(np.prod(shp))].view(shp)

    @property
    def blocks(self) -> dict:
        """Dict view key -> device block (tensor.py:750-753)."""
        return {k: self.block(i) for k, i in self.key_index().items()}

    def block_shapes(self) -> dict:
        return {tuple(int(q) for q in k): tuple(int(s) for s in sh) for k, sh in zip(self._keys, self._shapes)}

    def todense(self):
        """Scatter into a dense device tensor (tensor.py:760-781; testing only)."""
        shape = tuple(ix.dim for ix in self._indices)
        out = torch.zeros(int(np.prod(shape)) if shape else 1, dtype=torch.float64, device=self._data.device)
        if self.n_blocks:
            out[_dense_gather_index(self._indices, self._keys, self._shapes)] = self._data
        return out.reshape(shape)

    def conj(self):
        return self._with_data(self._data.conj())

    def bar(self):
        return self._with_data(self._data.conj(), tuple(ix.flip_flow() for ix in self._indices))

    def dagger(self):
        """Dual legs, dual keys, conjugated data (tensor.py:783-816); graded
        groups multiply each block by the product of its charges' twist
        phases, (-1)^(number of odd charges)."""
        sym = self._indices[0].symmetry
        new_keys = sym.dual(self._keys.astype(np.int64)).astype(np.int32)
        t = self
        if sym.is_fermionic and self.n_blocks:
            odd = np.asarray(sym.parity(self._keys.reshape(-1)), np.int64).reshape(self._keys.shape).sum(axis=1) % 2
            if odd.any():
                t = self._with_data(self._data * _block_sign_vector(self, odd))
        return _reorder(t, tuple(ix.dual() for ix in self._indices), new_keys, conj=True)

    def transpose(self, axes):
        """Leg permutation through the device transpose plan (tensor.py:832-861);
        for graded (fermionic) groups each block takes its Koszul sign.

        Differentiable: when the data carries a gradient the kernel runs
        behind an autograd node whose adjoint is the inverse permutation."""
        axes = tuple(int(a) for a in axes)
        if axes == tuple(range(self.ndim)):
            return self
        if torch.is_grad_enabled() and self._data.requires_grad:
            from tenax_amd.contraction.plans import transpose_plan

            _, keys, shapes, offs, _ = transpose_plan(self, axes, _graded(self))
            out = _TransposeFn.apply(self._data, self, axes)
            return SymmetricTensor._wrap(tuple(self._indices[i] for i in axes), keys, shapes, offs, out)
        return _transpose_raw(self, axes)

    def norm(self):
        return torch.linalg.vector_norm(self._data) if self._data.numel() else torch.zeros(
            (), dtype=torch.float64, device=self._data.device)

    def relabel(self, old, new):
        if old not in self.labels():
            raise KeyError(f"Label {old!r} not found in tensor with labels {self.labels()}")
        return self._with_data(self._data, _relabel_indices(self._indices, {old: new}))

    def relabels(self, mapping):
        return self._with_data(self._data, _relabel_indices(self._indices, mapping))

    def __mul__(self, scalar):
        return self._with_data(self._data * scalar)

    def __add__(self, other):
        if not isinstance(other, SymmetricTensor):
            raise TypeError(f"Cannot add SymmetricTensor and {type(other).__name__}; "
                            f"convert to matching type first.")
        if np.array_equal(self._keys, other._keys):
            return self._with_data(self._data + other._data)
        return _union_add(self, other)

    def __sub__(self, other):
        if not isinstance(other, SymmetricTensor):
            raise TypeError(f"Cannot subtract {type(other).__name__} from SymmetricTensor.")
        return self.__add__(other.__mul__(-1.0))

    def max_abs(self):
        if not self._data.numel():
            return torch.zeros((), dtype=torch.float64, device=self._data.device)
        return self._data.abs().max()

    def __repr__(self):
        return (f"SymmetricTensor(ndim={self.ndim}, n_blocks={self.n_blocks}, nnz={self._data.numel()}, "
                f"dtype={self.dtype}, labels={self.labels()})")


def _graded(t: "SymmetricTensor") -> bool:
    return bool(t._indices) and t._indices[0].symmetry.is_fermionic


def _block_sign_vector(t: "SymmetricTensor", negative) -> torch.Tensor:
    """+-1 per element of t's packed data, -1 on the blocks flagged in ``negative``."""
    sizes = np.prod(t._shapes, axis=1).astype(np.int64) if t.n_blocks else np.zeros(0, np.int64)
    sign = np.where(np.asarray(negative, bool), -1.0, 1.0)
    dev = t._data.device
    return torch.repeat_interleave(torch.from_numpy(sign).to(dev), torch.from_numpy(sizes).to(dev),
                                   output_size=int(sizes.sum()))


def _transpose_raw(t: "SymmetricTensor", axes, graded=None) -> "SymmetricTensor":
    """Device leg permutation; ``graded`` (default: the group is fermionic)
    applies the Koszul signs, ``graded=False`` moves data only."""
    from tenax_amd.contraction.plans import transpose_plan

    if graded is None:
        graded = _graded(t)
    plan, keys, shapes, offs, tot = transpose_plan(t, axes, graded)
    out = torch.empty(tot, dtype=torch.float64, device=t._data.device)
    if t.n_blocks:
        lib = _lib.load()
        _lib.check(lib.tnx_transpose_execute(plan.ptr, _lib.dptr(t._data.detach()), _lib.dptr(out),
                                             _lib.stream_handle()))
    return SymmetricTensor._wrap(tuple(t._indices[i] for i in axes), keys, shapes, offs, out)


class _TransposeFn(torch.autograd.Function):
    """Block transpose with its adjoint (the inverse permutation)."""

    @staticmethod
    def forward(ctx, data, t, axes):
        res = _transpose_raw(t, axes)
        ctx.meta = (res, axes)
        return res._data

    @staticmethod
    def backward(ctx, g):
        res, axes = ctx.meta
        inv = tuple(int(i) for i in np.argsort(axes))
        back = _transpose_raw(res._with_data(g.contiguous()), inv)
        return back._data, None, None


def _find_key(blocks, k):
    for kk in blocks:
        if tuple(int(q) for q in kk) == k:
            return kk
    raise KeyError(k)


def _dense_gather_index(indices, keys, shapes) -> torch.Tensor:
    """Flat dense positions of every packed element (block order, row-major)."""
    dims = [ix.dim for ix in indices]
    strides = np.ones(len(dims), np.int64)
    for l in range(len(dims) - 2, -1, -1):
        strides[l] = strides[l + 1] * dims[l + 1]
    parts = []
    for k in keys:
        g = np.zeros((), np.int64)
        for l, ix in enumerate(indices):
            pos = ix.positions(int(k[l])).astype(np.int64) * strides[l]
            g = np.add.outer(g, pos)
        parts.append(g.reshape(-1))
    idx = np.concatenate(parts) if parts else np.zeros(0, np.int64)
    return torch.from_numpy(idx).to(_device())


def _block_ranges_index(offsets, shapes, order, device) -> torch.Tensor:
    """Concatenated element ranges of blocks ``order`` (in that order) as a
    device index, built on the device from per-block starts and sizes (two
    small host arrays cross PCIe, not one index per element)."""
    order = np.asarray(order, np.int64)
    if len(order) == 0:
        return torch.zeros(0, dtype=torch.int64, device=device)
    sz = shapes[order].prod(axis=1).astype(np.int64)
    shift = np.asarray(offsets, np.int64)[order] - (np.cumsum(sz) - sz)
    total = int(sz.sum())
    shift_d = torch.repeat_interleave(torch.from_numpy(shift).to(device), torch.from_numpy(sz).to(device),
                                      output_size=total)
    return torch.arange(total, dtype=torch.int64, device=device) + shift_d


def _reorder(t: SymmetricTensor, indices, new_keys, conj=False) -> SymmetricTensor:
    order = np.lexsort(new_keys.T[::-1]) if len(new_keys) else np.zeros(0, np.int64)
    keys = new_keys[order]
    shapes = t._shapes[order]
    offs, tot = SymmetricTensor._layout(shapes)
    gi = _block_ranges_index(t._offsets, t._shapes, order, t._data.device)
    data = t._data[gi] if tot else t._data[:0]
    if conj:
        data = data.conj()
    return SymmetricTensor._wrap(indices, keys, shapes, offs, data)


def _union_add(a: SymmetricTensor, b: SymmetricTensor) -> SymmetricTensor:
    """Sum over the union of keys (tensor.py:923-952)."""
    ka, kb = a.key_index(), b.key_index()
    allk = sorted(set(ka) | set(kb))
    r = a.ndim
    keys = np.asarray(allk, np.int32).reshape(len(allk), r)
    shapes = np.zeros_like(keys)
    for i, k in enumerate(allk):
        shapes[i] = a._shapes[ka[k]] if k in ka else b._shapes[kb[k]]
    offs, tot = SymmetricTensor._layout(shapes)
    out = torch.zeros(tot, dtype=torch.float64, device=a._data.device)
    for src, kx in ((a, ka), (b, kb)):
        dst_order = [i for i, k in enumerate(allk) if k in kx]
        src_order = [kx[allk[i]] for i in dst_order]
        if not dst_order:
            continue
        di = _block_ranges_index(offs, shapes, dst_order, out.device)
        si = _block_ranges_index(src._offsets, src._shapes, src_order, out.device)
        out.index_add_(0, di, src._data[si])
    return SymmetricTensor._wrap(a._indices, keys, shapes, offs, out)


def inner(a: Tensor, b: Tensor):
    """Sum of conj(a)*b over common keys (tensor.py:296-319)."""
    if isinstance(a, DenseTensor) and isinstance(b, DenseTensor):
        return torch.sum(a._data.conj() * b._data)
    if isinstance(a, SymmetricTensor) and isinstance(b, SymmetricTensor):
        if np.array_equal(a._keys, b._keys):
            return torch.dot(a._data.conj(), b._data) if a._data.numel() else torch.zeros(
                (), dtype=torch.float64, device=a._data.device)
        ka, kb = a.key_index(), b.key_index()
        common = [k for k in ka if k in kb]
        if not common:
            return torch.zeros((), dtype=torch.float64, device=a._data.device)
        ia = _block_ranges_index(a._offsets, a._shapes, [ka[k] for k in common], a._data.device)
        ib = _block_ranges_index(b._offsets, b._shapes, [kb[k] for k in common], a._data.device)
        return torch.dot(a._data[ia].conj(), b._data[ib])
    return torch.sum(a.todense().conj() * b.todense())
