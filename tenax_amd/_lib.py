"""ctypes binding of libtnx.so (the C ABI declared in include/tnx.h and include/tnx_dmrg.h).

The library is built in-tree (``make -C tenax_amd/csrc``) and loaded from this
directory.  There is no fallback: if the library or a GPU is missing, every
entry point raises ``TnxUnavailable``.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TNX_LIB_PATH selects another build of the library (e.g. a variant compiled
# with different kernel parameters); the in-tree libtnx.so otherwise.
LIB_PATH = os.environ.get("TNX_LIB_PATH") or os.path.join(_HERE, "libtnx.so")

TNX_OK = 0
TNX_ERR_BAD_STRUCTURE = -1
TNX_ERR_CONSERVATION = -2
TNX_ERR_LABEL_COUNT = -3
TNX_ERR_DTYPE = -4
TNX_ERR_HIP = -5
TNX_ERR_OOM = -6
TNX_ERR_ARG = -7
TNX_ERR_CAPACITY = -8

COMPUTE = {"f64": 0, "f32": 1, "split_f32": 2}
MODE_PAIRWISE_FILTERED = 0
MODE_MULTI_REACHABLE = 1


class TnxUnavailable(RuntimeError):
    """libtnx.so is not built or no GPU is present; there is no CPU fallback."""


class TnxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libtnx error {code}: {msg}")
        self.code = code


class Leg(C.Structure):
    _fields_ = [
        ("n_charges", C.c_int32),
        ("charges", C.POINTER(C.c_int32)),
        ("mult", C.POINTER(C.c_int32)),
        ("flow", C.c_int32),
    ]


class Blocks(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("n_blocks", C.c_int32),
        ("keys", C.POINTER(C.c_int32)),
        ("shapes", C.POINTER(C.c_int32)),
        ("offsets", C.POINTER(C.c_int64)),
    ]


_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PD = C.POINTER(C.c_double)

# name -> (restype, argtypes); the exported surface of include/*.h
SIGNATURES = {
    "tnx_version": (C.c_int, []),
    "tnx_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "tnx_device_sync": (C.c_int, []),
    "tnx_valid_blocks": (C.c_int, [C.POINTER(Leg), _i32, _i32, _i32, _i32, _i32, _PI32, _PI32, _PI32, _p]),
    "tnx_plan_pairwise": (C.c_int, [C.POINTER(Blocks), C.POINTER(Blocks), _i32, _i32, _i32, C.POINTER(Leg),
                                    _i32, _i32, _i32, _i32, _p, C.POINTER(_p)]),
    "tnx_plan_pairwise_graded": (C.c_int, [C.POINTER(Blocks), C.POINTER(Blocks), _i32, _i32, _i32, C.POINTER(Leg),
                                           _i32, _i32, _i32, _i32, _i32, _PI32, _p, C.POINTER(_p)]),
    "tnx_plan_pairwise_permuted": (C.c_int, [C.POINTER(Blocks), C.POINTER(Blocks), _i32, _i32, _i32, C.POINTER(Leg),
                                             _i32, _i32, _i32, _i32, _PI32, _p, C.POINTER(_p)]),
    "tnx_plan_output_info": (C.c_int, [_p, _PI32, _PI32, _PI64]),
    "tnx_plan_output_copy": (C.c_int, [_p, _PI32, _PI32, _PI64]),
    "tnx_plan_stats": (C.c_int, [_p, _PD, _PD, _PI32, _PI32]),
    "tnx_plan_free": (C.c_int, [_p]),
    "tnx_plan_set_tiling": (C.c_int, [_p, _i32, _i32, _i32, _i32, _p]),
    "tnx_plan_get_tiling": (C.c_int, [_p, _i32, _PI32, _PI32, _PI32, _PI32]),
    "tnx_contract_execute": (C.c_int, [_p, _p, _p, _p, _i32, _p]),
    "tnx_contract_execute_perblock": (C.c_int, [_p, _p, _p, _p, _i32, _p]),
    "tnx_contract_vjp": (C.c_int, [_p, _p, _p, _p, _p, _p, _i32, _p]),
    "tnx_plan_transpose": (C.c_int, [C.POINTER(Blocks), C.POINTER(Leg), _PI32, _p, C.POINTER(_p)]),
    "tnx_transpose_execute": (C.c_int, [_p, _p, _p, _p]),
    "tnx_copy2d_execute": (C.c_int, [_p, _i32, _i64, _p]),
    "tnx_batch_pairwise": (C.c_int, [C.POINTER(_p), _i32, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p), _p,
                                     C.POINTER(_p)]),
    "tnx_batch_transpose": (C.c_int, [C.POINTER(_p), _i32, C.POINTER(_p), C.POINTER(_p), _p, C.POINTER(_p)]),
    "tnx_batch_execute": (C.c_int, [_p, _i32, _p]),
    "tnx_plan_validate": (C.c_int, [_p, _p]),
    "tnx_check_gemm_worklist": (C.c_int, [_i32, _p, _i32, _p, _i32, _p, _i32, _i32, _i32, _i64, _PI64, _i32,
                                          _PI64, _i32, _PI64, _i32]),
    "tnx_plan_transpose_graded": (C.c_int, [C.POINTER(Blocks), C.POINTER(Leg), _PI32, _i32, _p, C.POINTER(_p)]),
    # include/tnx_dmrg.h
    "tnx_matvec_plan": (C.c_int, [C.POINTER(Blocks)] * 5 + [_p, C.POINTER(_p)]),
    "tnx_matvec_reach": (C.c_int, [_p, _PI32, _PI32, _PI32, _i32]),
    "tnx_matvec_execute": (C.c_int, [_p, _p, _p, _p, _p, _p, _p, _i32, _p]),
    "tnx_matvec_execute_perblock": (C.c_int, [_p, _p, _p, _p, _p, _p, _p, _i32, _p]),
    "tnx_matvec_execute_stage": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p, _i32, _p]),
    "tnx_fused_stage_flops": (C.c_int, [_p, _PD]),
    "tnx_env_plan": (C.c_int, [_i32, C.POINTER(Blocks), C.POINTER(Blocks), C.POINTER(Blocks), C.POINTER(Leg), _p,
                               C.POINTER(_p)]),
    "tnx_env_execute": (C.c_int, [_p, _p, _p, _p, _p, _i32, _p]),
    "tnx_env_execute_perblock": (C.c_int, [_p, _p, _p, _p, _p, _i32, _p]),
    "tnx_lanczos": (C.c_int, [_p, _p, _p, _p, _p, _p, _i32, C.c_double, _i32, _i32, _PD, _PI32, _p]),
    "tnx_lanczos_noalloc": (C.c_int, [_p, _p, _p, _p, _p, _p, _i32, C.c_double, _i32, _i32, _PD, _PI32, _p]),
    "tnx_lanczos_graph_create": (C.c_int, [_p, _p, _p, _p, _p, _p, _i32, C.c_double, _i32, _i32, _p,
                                           C.POINTER(_p)]),
    "tnx_lanczos_graph_update": (C.c_int, [_p, _p, _p, _p, _p, _p, _p]),
    "tnx_lanczos_graph_rebind": (C.c_int, [_p, _p, _p, _p, _p, _p, _p, _p]),
    "tnx_graph_launch": (C.c_int, [_p, _p]),
    "tnx_lanczos_graph_result": (C.c_int, [_p, _p, _PD, _PI32]),
    "tnx_graph_free": (C.c_int, [_p]),
    # include/tnx_random.h
    "tnx_threefry2x32": (C.c_int, [C.c_uint32] * 4 + [C.POINTER(C.c_uint32)]),
    "tnx_random_normal_blocks": (C.c_int, [_i32, C.POINTER(C.c_uint32), _PI64, _PI64, _i32, C.c_double, _p, _p]),
    "tnx_random_bits": (C.c_int, [C.c_uint32, C.c_uint32, _i64, _i32, _p, _p]),
}

_lib = None
_lock = threading.Lock()


def load(require_gpu: bool = True):
    """Load libtnx.so (raises TnxUnavailable when it is missing)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise TnxUnavailable(
                    f"{LIB_PATH} is not built; run `make -C tenax_amd/csrc` (no CPU fallback exists)")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    if require_gpu:
        import torch

        if not torch.cuda.is_available():
            raise TnxUnavailable("no GPU visible: the tenax_amd compute path runs only on the GPU")
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    _lib.tnx_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != TNX_OK:
        raise TnxError(rc, last_error())


def stream_handle():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def i32p(a: np.ndarray):
    return a.ctypes.data_as(_PI32)


def i64p(a: np.ndarray):
    return a.ctypes.data_as(_PI64)


def dptr(t):
    """Device pointer of a torch tensor (None for an empty tensor)."""
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else C.c_void_p(0)


class LegArray:
    """Keeps numpy buffers alive while a Leg[] array is in use."""

    def __init__(self, legs):
        # legs: iterable of (charges ndarray int32 ascending, mult ndarray int32, flow int)
        self._keep = []
        legs = list(legs)
        self.arr = (Leg * max(1, len(legs)))()
        for i, (ch, mu, fl) in enumerate(legs):
            ch = np.ascontiguousarray(ch, dtype=np.int32)
            mu = np.ascontiguousarray(mu, dtype=np.int32)
            self._keep += [ch, mu]
            self.arr[i] = Leg(len(ch), i32p(ch), i32p(mu), int(fl))
        self.n = len(legs)


class BlockTable:
    """Keeps numpy buffers alive while a Blocks struct is in use."""

    def __init__(self, keys: np.ndarray, shapes: np.ndarray, offsets: np.ndarray, rank: int):
        self.keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1)
        self.shapes = np.ascontiguousarray(shapes, dtype=np.int32).reshape(-1)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64).reshape(-1)
        nb = len(self.offsets)
        self.s = Blocks(rank, nb, i32p(self.keys), i32p(self.shapes), i64p(self.offsets))


class Plan:
    """Owning wrapper of a tnx_plan*."""

    __slots__ = ("ptr", "_lib", "tuned", "__weakref__")

    def __init__(self, ptr):
        self.ptr = ptr
        self._lib = _lib
        self.tuned = False

    def __del__(self):
        if self.ptr and self._lib is not None:
            try:
                self._lib.tnx_plan_free(self.ptr)
            except Exception:
                pass
            self.ptr = None

    def output(self):
        nb, rk, tot = C.c_int32(), C.c_int32(), C.c_int64()
        check(_lib.tnx_plan_output_info(self.ptr, C.byref(nb), C.byref(rk), C.byref(tot)))
        n, r = nb.value, rk.value
        keys = np.zeros((n, r), np.int32)
        shapes = np.zeros((n, r), np.int32)
        offs = np.zeros(n, np.int64)
        if n:
            check(_lib.tnx_plan_output_copy(self.ptr, i32p(keys), i32p(shapes), i64p(offs)))
        return keys, shapes, offs, tot.value

    def set_tiling(self, stage: int, shape: int, kchunk: int = 0, xcd: int = 0):
        """Tiling of one grouped-GEMM stage (0 pairwise; 1 or 3 fused): shape 0-3 =
        64x64, 128x64, 64x128, 128x128; kchunk K tiles per task (0 whole, <0 auto);
        xcd 1 = XCD-aware task order."""
        check(_lib.tnx_plan_set_tiling(self.ptr, stage, shape, kchunk, xcd, stream_handle()))

    def tiling(self, stage: int):
        sh, kc, xc, nt = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(_lib.tnx_plan_get_tiling(self.ptr, stage, C.byref(sh), C.byref(kc), C.byref(xc), C.byref(nt)))
        return {"shape": sh.value, "kchunk": kc.value, "xcd": xc.value, "n_tasks": nt.value}

    def stats(self):
        f, b = C.c_double(), C.c_double()
        npairs, ntasks = C.c_int32(), C.c_int32()
        check(_lib.tnx_plan_stats(self.ptr, C.byref(f), C.byref(b), C.byref(npairs), C.byref(ntasks)))
        return {"flops": f.value, "bytes": b.value, "n_pairs": npairs.value, "n_tasks": ntasks.value}


COPY2D_CHUNK = 4096  # TNX_COPY2D_CHUNK (include/tnx.h)
_COPY2D = np.dtype([("src", np.uint64), ("dst", np.uint64), ("src_ld", np.int64), ("dst_ld", np.int64),
                    ("task0", np.int64), ("rows", np.int32), ("cols", np.int32), ("scale", np.float64)])


def copy2d(copies, device):
    """Batched 2-D strided copies in one launch (tnx_copy2d_execute).

    copies: iterable of (src, src_off, src_ld, dst, dst_off, dst_ld, rows,
    cols, scale) with src / dst contiguous float64 device tensors and element
    offsets / row pitches; each rectangle is checked to lie inside its tensor.
    """
    import torch

    copies = [c for c in copies if c[6] > 0 and c[7] > 0]
    if not copies:
        return
    d = np.zeros(len(copies), _COPY2D)
    wg = 0
    for i, (src, so, sld, dst, do, dld, rows, cols, scale) in enumerate(copies):
        for t, off, ld in ((src, so, sld), (dst, do, dld)):
            if t.dtype != torch.float64 or not t.is_contiguous():
                raise ValueError("copy2d needs contiguous float64 tensors")
            if off < 0 or ld < cols or off + (rows - 1) * ld + cols > t.numel():
                raise ValueError(f"copy2d rectangle {rows}x{cols} at {off} (pitch {ld}) outside a tensor of "
                                 f"{t.numel()} elements")
        d[i] = (src.data_ptr() + 8 * so, dst.data_ptr() + 8 * do, sld, dld, wg, rows, cols, scale)
        wg += -(-rows * cols // COPY2D_CHUNK)
    dev = torch.from_numpy(d.view(np.uint8)).to(device)
    check(load().tnx_copy2d_execute(dptr(dev), len(copies), wg, stream_handle()))
    return dev  # the caller may keep it alive until the copies ran (stream order makes this optional)
