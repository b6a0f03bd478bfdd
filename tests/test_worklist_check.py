"""CPU tests of the work-list validator (tnx_check_gemm_worklist): a valid
grouped-GEMM work list passes, and each kind of out-of-range entry -- the
mistakes that would make a launch read or write outside an operand block --
is caught with a message naming it.  No device work is involved."""

import ctypes as C

import numpy as np
import pytest

from tenax_amd import _lib

PAIR = np.dtype([("a_off", "<i8"), ("b_off", "<i8"), ("K", "<i4"), ("lda", "<i4"), ("ldb", "<i4"), ("neg", "<i4")])
OUT = np.dtype([("c_off", "<i8"), ("M", "<i4"), ("N", "<i4"), ("ldc", "<i4"), ("pair_begin", "<i4"),
                ("pair_end", "<i4"), ("pad", "<i4")])
TASK = np.dtype([("out", "<i4"), ("m0", "<i4"), ("n0", "<i4"), ("kt0", "<i4"), ("kt1", "<i4"), ("slot", "<i4"),
                 ("pad0", "<i4"), ("pad1", "<i4")])
WHOLE = 0x7FFFFFFF


def _valid():
    """C[o] = sum over pairs A_k B_k for two outputs, A in MK layout.

    A blocks: a0 100x40 at 0, a1 100x24 at 4000, a2 70x30 at 6400.
    B blocks: b0 40x90 at 0, b1 24x90 at 3600, b2 30x50 at 5760.
    C blocks: c0 100x90 at 0, c1 70x50 at 9000."""
    a_blocks = np.array([[0, 4000], [4000, 2400], [6400, 2100]], np.int64)
    b_blocks = np.array([[0, 3600], [3600, 2160], [5760, 1500]], np.int64)
    c_blocks = np.array([[0, 9000], [9000, 3500]], np.int64)
    pairs = np.zeros(3, PAIR)
    pairs[0] = (0, 0, 40, 40, 90, 0)
    pairs[1] = (4000, 3600, 24, 24, 90, 0)
    pairs[2] = (6400, 5760, 30, 30, 50, 0)
    outs = np.zeros(2, OUT)
    outs[0] = (0, 100, 90, 90, 0, 2, 0)
    outs[1] = (9000, 70, 50, 50, 2, 3, 0)
    tasks = []
    for o, (M, N) in enumerate([(100, 90), (70, 50)]):
        for m0 in range(0, M, 64):
            for n0 in range(0, N, 64):
                tasks.append((o, m0, n0, 0, WHOLE, -1, 0, 0))
    # a split tile: output 0, tile (0, 0), its 5 K tiles in two slots
    tasks.append((0, 0, 0, 0, 3, 0, 0, 0))
    tasks.append((0, 0, 0, 3, 5, 1, 0, 0))
    return dict(a_layout=0, pairs=pairs, outs=outs, tasks=np.array(tasks, TASK), bm=64, bn=64, ws=2,
                a=a_blocks, b=b_blocks, c=c_blocks)


def _check(w):
    lib = _lib.load(require_gpu=False)
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    i64 = lambda x: np.ascontiguousarray(x, np.int64).reshape(-1).ctypes.data_as(C.POINTER(C.c_int64))
    a, b, c = (np.ascontiguousarray(w[k], np.int64) for k in "abc")
    rc = lib.tnx_check_gemm_worklist(w["a_layout"], p(w["pairs"]), len(w["pairs"]), p(w["outs"]), len(w["outs"]),
                                     p(w["tasks"]), len(w["tasks"]), w["bm"], w["bn"], w["ws"], i64(a), len(a),
                                     i64(b), len(b), i64(c), len(c))
    return rc, _lib.last_error()


def test_valid_worklist_passes():
    rc, msg = _check(_valid())
    assert rc == 0, msg


def _mutate(field, index, value, table="pairs"):
    w = _valid()
    w[table][index][field] = value
    return w


@pytest.mark.parametrize("case,expect", [
    (("a_off", 1, 4001, "pairs"), "pair 1 A range"),          # A segment shifted across a block end
    (("lda", 0, 39, "pairs"), "pair 0 A row pitch"),          # pitch below K
    (("b_off", 2, 5800, "pairs"), "pair 2 B range"),          # B segment past its block
    (("K", 1, 25, "pairs"), "pair 1 A row pitch below K"),    # one K column too many
    (("ldb", 0, 80, "pairs"), "pair 0 B row pitch"),          # pitch below N
    (("c_off", 1, 8999, "outs"), "output 1 range"),           # output straddles two C blocks
    (("M", 0, 101, "outs"), "output 0 range"),                # one row too many
    (("pair_end", 1, 4, "outs"), "output 1 pair range"),      # pair list overrun
    (("m0", 0, 32, "tasks"), "task 0 is not a tile"),         # misaligned tile
    (("n0", 4, 64, "tasks"), "task 4 is not a tile"),         # output 1 has a single tile column
    (("slot", 6, 2, "tasks"), "task 6 writes past the split-K workspace"),
    (("kt1", 7, 6, "tasks"), "task 7 ends past"),             # K range past the output's 5 tiles
    (("out", 2, 5, "tasks"), "task 2 names a missing output"),
])
def test_each_violation_is_caught(case, expect):
    field, index, value, table = case
    rc, msg = _check(_mutate(field, index, value, table))
    assert rc == _lib.TNX_ERR_BAD_STRUCTURE
    assert expect in msg, msg


def test_row_past_block_end_is_caught():
    """K and pitch grown together: the last row runs past the block's end."""
    w = _valid()
    w["pairs"][1]["K"] = 25
    w["pairs"][1]["lda"] = 25
    rc, msg = _check(w)
    assert rc == _lib.TNX_ERR_BAD_STRUCTURE and "pair 1 A range" in msg, msg


def test_overlapping_operand_blocks_rejected():
    w = _valid()
    w["b"] = np.array([[0, 3600], [3500, 2160], [5760, 1500]], np.int64)
    rc, msg = _check(w)
    assert rc == _lib.TNX_ERR_BAD_STRUCTURE and "overlap" in msg


def test_km_layout_extents():
    """KM layout: an A block is K rows of pitch lda >= M."""
    w = _valid()
    w["a_layout"] = 1
    for k, (K, M) in enumerate([(40, 100), (24, 100), (30, 70)]):
        w["pairs"][k]["lda"] = M
    w["a"] = np.array([[0, 4000], [4000, 2400], [6400, 2100]], np.int64)
    rc, msg = _check(w)
    assert rc == 0, msg
    w["pairs"][2]["lda"] = 69
    rc, msg = _check(w)
    assert rc == _lib.TNX_ERR_BAD_STRUCTURE and "pair 2 A row pitch below M" in msg
