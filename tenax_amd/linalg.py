"""Block-diagonal SVD with global truncation for SymmetricTensor.

Mirrors reference src/tenax/linalg.py:34-333 and 665-825:

* sector q of a block = fused flow*charge of its left legs (linalg.py:34-69);
* each sector matrix is decomposed independently in FP64 on the GPU;
* singular values of all sectors are merged with a STABLE sort by -s
  (sector visit order, then index inside the sector; linalg.py:211), the
  first ``min(n, max_singular_values)`` are kept (at least one), with the
  optional truncation-error rule of linalg.py:220-235;
* the new bond's charges follow the global descending order (linalg.py:242),
  so charge-q positions on the bond are the sector's kept singular vectors in
  order; U blocks are keyed ``lk + (q,)``, Vh blocks ``(q,) + rk``.

Sector visit order here: first appearance in the sorted block order (the
reference visits dict insertion order; only exact ties across sectors can
tell the two apart).  Row/column sub-keys inside a sector are likewise in
sorted order.  Blocks enter the sector matrix reshaped in their stored leg
order, exactly as the reference's linalg.py:181 does (see _assemble).  All
sector matrices of a call live in one device buffer filled by one batched
copy launch (tnx_copy2d_execute), and the kept singular vectors are unpacked
into the U and Vh blocks by one launch each, where the reference updates one
block slice at a time (linalg.py:163-187, 268-309); only the spectrum and
its sort order travel to the host, which fixes the new bond's structure.

Decomposition.  Without ``absorb`` each sector matrix goes through
torch.linalg.svd (rocSOLVER gesvd) in float64, and so does the DMRG split
(``absorb``) while every sector is at most SMALL_SECTOR wide (sectors of one
shape in one batched call).  Above that size the split uses the fact that
only one factor has to be an isometry: the other one carries the singular
values.  The isometric side comes from a Gram eigenproblem,
``M M^T = U diag(s^2) U^T`` (absorb="right") or ``M^T M = V diag(s^2) V^T``
(absorb="left").  (Solving the smaller side's Gram matrix instead and
re-orthonormalising M V was tried: its columns for singular values below
~1e-8 s_max are not determined by the product, and the Householder QR that
then has to complete them costs as much as the larger eigenproblem.)  The carried side
is the exact product ``U^T M`` or ``M V``; s is the row (column) norm of that
product.  The carried factor needs no division by s, and ``U U^T M`` is the
optimal rank-k approximation up to eigenvector error eps*s_max^2/gap, which
only matters for the truncation when two singular values straddling the cut
are degenerate to that level.  The eigenproblems of all sectors are solved
in batched torch.linalg.eigh calls (size groups run concurrently on side
streams): Gram matrices are scaled by their trace and padded to a common size
with -1 on the padded diagonal, so padded eigenpairs sit at -1 and the
sector's own ones in [0, 1] come out exactly decoupled.  ``per_sector=True``
instead runs one rocSOLVER SVD per sector in sector order (the reference's
loop, linalg.py:163-190): the per-block comparison path.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from tenax_amd import _lib
from tenax_amd.core.index import FlowDirection, Label, TensorIndex
from tenax_amd.core.tensor import DenseTensor, SymmetricTensor, Tensor


def _sector_layout(t: SymmetricTensor, lax, rax):
    """Group blocks by the fused flow*charge of the left legs (linalg.py:34-69)."""
    sym = t.indices[0].symmetry
    flows = np.array([int(t.indices[i].flow) for i in lax], np.int64)
    qs = (t.keys[:, lax].astype(np.int64) * flows).sum(axis=1)
    n = sym.n_values()
    if n is not None:
        qs = np.mod(qs, n)
    sectors: dict[int, dict] = {}
    for i in range(t.n_blocks):
        q = int(qs[i])
        sec = sectors.setdefault(q, {"rows": {}, "cols": {}, "blocks": []})
        lk = tuple(int(t.keys[i, a]) for a in lax)
        rk = tuple(int(t.keys[i, a]) for a in rax)
        lshape = tuple(int(t.shapes[i, a]) for a in lax)
        rshape = tuple(int(t.shapes[i, a]) for a in rax)
        if lk not in sec["rows"]:
            sec["rows"][lk] = (int(np.prod(lshape)), lshape)
        if rk not in sec["cols"]:
            sec["cols"][rk] = (int(np.prod(rshape)), rshape)
        sec["blocks"].append((i, lk, rk))
    for sec in sectors.values():
        r0 = 0
        sec["row_off"] = {}
        for lk, (nr, _) in sec["rows"].items():
            sec["row_off"][lk] = r0
            r0 += nr
        c0 = 0
        sec["col_off"] = {}
        for rk, (nc, _) in sec["cols"].items():
            sec["col_off"][rk] = c0
            c0 += nc
        sec["R"], sec["C"] = r0, c0
    return sectors


# rocSOLVER's batched eigh costs about the same per call for a batch of 1 or
# of 12 (its tridiagonal reduction is a sequence of launches, ~20 us per row),
# so matrices of similar size are padded to a common size and batched, in
# chunks of at most this many bytes of padded input.  The call also blocks its
# host thread while the GPU is mostly idle, so batches of different sizes run
# concurrently: each from its own host thread on its own stream (default
# priority).  One 11-sector chi=2048 DMRG split: one padded batch 38 ms, size
# groups in parallel ~27 ms.
_MAX_BATCH_BYTES = 4 << 30
_N_STREAMS = 4
_PARALLEL_MIN = 256  # below this largest size one batch is cheapest
_pool = None
_streams: dict = {}


def _batches(sizes):
    """Chunks of matrix indices sharing one padded eigh call, with the pad size.

    Sizes are grouped (largest first) so that a group's padded size is at
    most 1.2x its smallest member; beyond _N_STREAMS groups the smallest are
    merged.  Chunks respect the byte cap."""
    order = sorted(range(len(sizes)), key=lambda j: -sizes[j])
    groups = []
    for j in order:
        if groups and (len(groups) >= _N_STREAMS or sizes[groups[-1][0]] <= 1.2 * sizes[j]
                       or max(sizes) < _PARALLEL_MIN):
            groups[-1].append(j)
        else:
            groups.append([j])
    out = []
    for g in groups:
        P = -(-sizes[g[0]] // 32) * 32
        per = max(1, _MAX_BATCH_BYTES // (P * P * 8))
        out.extend((P, g[i:i + per]) for i in range(0, len(g), per))
    return out


def _eigh_one(grams, P, js, pad_value, normalise, dev):
    stack = torch.zeros((len(js), P, P), dtype=torch.float64, device=dev)
    scales = []
    for b, j in enumerate(js):
        g = grams[j]
        n = g.shape[0]
        sc = (torch.diagonal(g).sum() if normalise else torch.linalg.matrix_norm(g)).clamp_min(1e-300)
        scales.append(sc)
        stack[b, :n, :n] = g / sc
        if n < P:
            pad = torch.arange(n, P, device=dev)
            stack[b, pad, pad] = pad_value
    w, V = torch.linalg.eigh(stack)
    res = []
    for b, j in enumerate(js):
        n = grams[j].shape[0]
        res.append((j, w[b, P - n:].flip(-1) * scales[b], V[b, :n, P - n:].flip(-1)))
    return res


def _eigvecs_desc(grams, pad_value=-1.0, normalise=True):
    """Eigenpairs of symmetric matrices, columns by descending eigenvalue, from
    as few batched torch.linalg.eigh calls as the size cap allows.  Matrices
    are scaled into [-1, 1] (trace for PSD Gram matrices, Frobenius norm
    otherwise) and padded with ``pad_value`` on the diagonal, below every
    eigenvalue of the scaled matrix, so padded pairs separate exactly.
    Returns [(eigenvalues, eigenvectors)] in the original scale."""
    global _pool
    dev = grams[0].device
    out = [None] * len(grams)
    batches = _batches([g.shape[0] for g in grams])
    if len(batches) == 1:
        for j, w, V in _eigh_one(grams, *batches[0], pad_value, normalise, dev):
            out[j] = (w, V)
        return out
    from concurrent.futures import ThreadPoolExecutor

    if _pool is None:
        _pool = ThreadPoolExecutor(max_workers=_N_STREAMS, thread_name_prefix="tnx-eigh")
    main = torch.cuda.current_stream(dev)
    ss = _streams.setdefault(dev.index, [torch.cuda.Stream(device=dev) for _ in range(_N_STREAMS)])

    def run(k, P, js):
        st = ss[k % _N_STREAMS]
        with torch.cuda.device(dev), torch.cuda.stream(st):
            res = _eigh_one(grams, P, js, pad_value, normalise, dev)
            for _, w, V in res:
                w.record_stream(main)
                V.record_stream(main)
        return res

    for st in ss:
        st.wait_stream(main)
    for g in grams:
        for st in ss:
            g.record_stream(st)
    futs = [_pool.submit(run, k, P, js) for k, (P, js) in enumerate(batches)]
    for f in futs:
        for j, w, V in f.result():
            out[j] = (w, V)
    for st in ss:
        main.wait_stream(st)
    return out


# largest min(rows, cols) of any sector for which the DMRG split uses a
# direct SVD per sector instead of the batched Gram eigenproblems
SMALL_SECTOR = 64


def _factor_per_sector(mats, absorb):
    """One rocSOLVER SVD call per sector matrix, in sector order (the loop of
    reference linalg.py:163-190 on the device; the per-block comparison path),
    with s absorbed by a product when asked."""
    out = []
    for M in mats:
        U, sv, Vh = torch.linalg.svd(M, full_matrices=False)
        if absorb == "right":
            out.append((U, sv, Vh * sv[:, None]))
        elif absorb == "left":
            out.append((U * sv[None, :], sv, Vh))
        else:
            out.append((U, sv, Vh))
    return out


def _factor_all(mats, absorb):
    """Per sector matrix M: (U, s, Vh) with columns/rows by descending s.

    absorb=None: rocSOLVER SVD.
    "right": U from eigh(M M^T) (isometry), Vh = U^T M (= diag(s) Vh).
    "left":  Vh from eigh(M^T M) (isometry), U = M V (= U diag(s)).
      s = norms of the rows of U^T M (columns of M V).
    "isometries": both factors as isometries with the SVD's pairing, from the
      eigenvectors of the Jordan-Wielandt matrix [[0, M], [M^T, 0]], whose
      eigenpairs are (+-s_i, (u_i, +-v_i)/sqrt(2)).  Its eigh resolves s with
      absolute error eps*s_max, like a direct SVD (a Gram matrix squares the
      spread and loses the vectors of s below ~1e-8 s_max).
    min(R, C) values per sector.
    """
    if absorb is None:
        return [torch.linalg.svd(M, full_matrices=False) for M in mats]
    if absorb in ("left", "right") and max(min(M.shape) for M in mats) <= SMALL_SECTOR:
        # small problems: a direct SVD per sector (rocSOLVER gesvd, cheap at
        # this size) rounds singular values like LAPACK, which decides
        # near-ties at the truncation cut the same way as the reference;
        # sectors of one shape (the +-q pairs) share one batched call
        out = [None] * len(mats)
        by_shape: dict = {}
        for j, M in enumerate(mats):
            by_shape.setdefault(tuple(M.shape), []).append(j)
        for js in by_shape.values():
            U, sv, Vh = torch.linalg.svd(torch.stack([mats[j] for j in js]), full_matrices=False)
            for b, j in enumerate(js):
                out[j] = ((U[b], sv[b], Vh[b] * sv[b][:, None]) if absorb == "right"
                          else (U[b] * sv[b][None, :], sv[b], Vh[b]))
        return out
    if absorb == "isometries":
        augs = []
        for M in mats:
            R, C = M.shape
            A = torch.zeros((R + C, R + C), dtype=torch.float64, device=M.device)
            A[:R, R:] = M
            A[R:, :R] = M.T
            augs.append(A)
        out = []
        for M, (w, X) in zip(mats, _eigvecs_desc(augs, pad_value=-2.0, normalise=False)):
            R, C = M.shape
            k = min(R, C)
            U = X[:R, :k]
            V = X[R:, :k]
            U = U / torch.linalg.vector_norm(U, dim=0).clamp_min(1e-300)
            V = V / torch.linalg.vector_norm(V, dim=0).clamp_min(1e-300)
            out.append((U, w[:k].clamp_min(0.0), V.T))
        return out
    grams = [M @ M.T for M in mats] if absorb == "right" else [M.T @ M for M in mats]
    out = []
    for M, (_, V) in zip(mats, _eigvecs_desc(grams)):
        k = min(M.shape)
        V = V[:, :k]
        if absorb == "left":
            B = M @ V
            s, perm = torch.sort(torch.linalg.vector_norm(B, dim=0), descending=True, stable=True)
            out.append((B[:, perm], s, V[:, perm].T))
        else:
            B = V.T @ M
            s, perm = torch.sort(torch.linalg.vector_norm(B, dim=1), descending=True, stable=True)
            out.append((V[:, perm], s, B[perm, :]))
    return out


# Leading singular triplets by block subspace iteration, for a cut that is a
# fixed count k (HOTRG's bond dimension): sector matrices at least
# LEADING_MIN_RATIO * k on their short side iterate a block of p = 2k vectors,
#   U = orth(M V),  V R = orth(M^T U),  U^T M V = R^T = X diag(s) Y^T,
# until the Ritz triplets (U X, s, V Y) of the leading k have residual
# |M V Y - U X diag(s)| <= LEADING_TOL * s_0 per column (the other residual
# M^T U X - V Y diag(s) vanishes by construction).  orth is shifted
# Cholesky QR in three passes (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
# "Shifted Cholesky QR for computing the QR factorization of ill-conditioned
# matrices", SIAM J. Sci. Comput. 42 (2020)): batched Gram products, p x p
# Cholesky factors and triangular inverses, orthonormal to working precision
# for condition numbers up to ~1/eps; rocSOLVER's batched Householder QR of
# the same blocks is 50-100x slower.  The small SVD of R^T comes from the
# eigh of its 2p-wide Jordan-Wielandt matrix, so s has the absolute accuracy
# eps * s_0 of a direct SVD and each (u, v) keeps the SVD's pairing.  The
# leading k converge at rate s_2k / s_k per half step (Z2 Ising environments
# at chi = 48, 1152-wide sectors: 2-11 iterations).  A group whose Cholesky
# factorisation breaks down, or that does not converge within
# LEADING_MAX_ITER, falls back to the full Jordan-Wielandt eigh.
LEADING_TOL = 2e-14
LEADING_MAX_ITER = 40
LEADING_MIN_RATIO = 4
# Matrices of nearly equal shape (e.g. the sectors of one step's instances,
# whose charge splits differ by a few states) iterate together, zero-padded
# to the largest shape of their group, when that costs at most this factor
# in elements: zero rows and columns change neither the singular values nor
# the singular vectors' restriction to the matrix, and one batched iteration
# replaces one per shape.
LEADING_PAD = 1.25
LEADING_STATS: dict = {"calls": 0, "matrices": 0, "iterations": 0, "fallback": 0, "fb_orth": 0, "fb_svd": 0,
                       "fb_conv": 0, "complete_eigh_fail": 0, "complete_chol_fail": 0}
# Ritz vectors of singular values below this fraction of the largest are not
# resolved by any FP64 factorisation (LAPACK's to eps * s_0 / gap); the
# iteration requires convergence of the leading ones above it only.
LEADING_RESOLVED = 1e-11


_ORTH_TOL = 1e-12


def _orth(X):
    """Shifted Cholesky QR3 of a batch of tall blocks: (Q, R, ok), X = Q R."""
    B, m, n = X.shape
    eye = torch.eye(n, dtype=X.dtype, device=X.device).expand(B, n, n)
    shift = 11.0 * (m * n + n * (n + 1)) * _EPS * (X * X).sum(dim=(1, 2))
    ok = torch.ones(B, dtype=torch.bool, device=X.device)
    Q, Rt = X, None
    for k in range(3):
        A = torch.bmm(Q.transpose(1, 2), Q)
        if k == 0:
            A = A + shift[:, None, None] * eye
        L, info = torch.linalg.cholesky_ex(A)
        ok &= info == 0
        R = L.transpose(1, 2)
        Q = torch.bmm(Q, torch.linalg.solve_triangular(R, eye, upper=True))
        Rt = R if Rt is None else torch.bmm(R, Rt)
    # beyond condition ~1/eps (numerically rank-deficient blocks) the passes
    # can succeed without orthonormalising: such members are not ok
    ok &= (torch.bmm(Q.transpose(1, 2), Q) - eye).abs().amax(dim=(1, 2)) <= _ORTH_TOL
    return Q, Rt, ok


# Eigenvalues of the Gram matrix of the shifted first pass (in [0, 1]) are
# resolved to ~eps; directions below this (block singular values under
# ~2e-12 |X|, i.e. sector singular values under the noise floor _NOISE) are
# replaced by random directions in _orth_complete.
_NULL_GRAM = 1e-14


def _orth_complete(X, fill):
    """Orthonormal basis of a batch of tall blocks that may be numerically
    rank-deficient: (Q, ok) with span(Q) containing X's resolved column space
    (columns of singular value above ~eps * |X|) and the rest completed from
    ``fill`` (random blocks of X's shape).  Shifted Cholesky pass, eigh of
    its Gram matrix with the resolved directions first, null directions
    replaced, then two plain Cholesky QR passes, which keep the span of each
    leading set of columns."""
    B, m, n = X.shape
    eye = torch.eye(n, dtype=X.dtype, device=X.device).expand(B, n, n)
    shift = 11.0 * (m * n + n * (n + 1)) * _EPS * (X * X).sum(dim=(1, 2))
    L1, info = torch.linalg.cholesky_ex(torch.bmm(X.transpose(1, 2), X) + shift[:, None, None] * eye)
    ok = info == 0
    Q1 = torch.bmm(X, torch.linalg.solve_triangular(L1.transpose(1, 2), eye, upper=True))
    if not bool(ok.all()):
        return Q1, ok
    try:
        lam, W = torch.linalg.eigh(torch.bmm(Q1.transpose(1, 2), Q1))
    except torch.linalg.LinAlgError:
        LEADING_STATS["complete_eigh_fail"] += 1
        return Q1, torch.zeros_like(ok)
    lam, W = lam.flip(-1), W.flip(-1)
    good = lam > _NULL_GRAM
    Q = torch.bmm(Q1, W) * torch.rsqrt(lam.clamp_min(_NULL_GRAM))[:, None, :]
    Q = torch.where(good[:, None, :], Q, fill)
    for _ in range(2):
        L, info = torch.linalg.cholesky_ex(torch.bmm(Q.transpose(1, 2), Q))
        ok &= info == 0
        Q = torch.bmm(Q, torch.linalg.solve_triangular(L.transpose(1, 2), eye, upper=True))
    if not bool(ok.all()):
        LEADING_STATS["complete_chol_fail"] += 1
    return Q, ok


def _small_svd(T):
    """SVD of a batch of square matrices from their Jordan-Wielandt eigh:
    (X, s, Y) with T = X diag(s) Y^T, s descending."""
    B, n, _ = T.shape
    J = torch.zeros((B, 2 * n, 2 * n), dtype=T.dtype, device=T.device)
    J[:, :n, n:] = T
    J[:, n:, :n] = T.transpose(1, 2)
    w, E = torch.linalg.eigh(J)
    top = E[:, :, n:].flip(-1)
    X = top[:, :n, :]
    Y = top[:, n:, :]
    X = X / torch.linalg.vector_norm(X, dim=1, keepdim=True).clamp_min(1e-300)
    Y = Y / torch.linalg.vector_norm(Y, dim=1, keepdim=True).clamp_min(1e-300)
    return X, w[:, n:].flip(-1).clamp_min(0.0), Y


def _orthonormal_ordered(H, gen):
    """Orthonormalise the columns of a batch of blocks in column order (each
    leading set of columns keeps its span): two Cholesky QR passes, or, where
    the columns are numerically dependent (Ritz vectors of singular values at
    the noise level), Gram-Schmidt in two passes per column with dependent
    columns replaced by random directions."""
    B, n, k = H.shape
    eye = torch.eye(k, dtype=H.dtype, device=H.device).expand(B, k, k)
    Q = H
    ok = True
    for _ in range(2):
        L, info = torch.linalg.cholesky_ex(torch.bmm(Q.transpose(1, 2), Q))
        if bool((info != 0).any()):
            ok = False
            break
        Q = torch.bmm(Q, torch.linalg.solve_triangular(L.transpose(1, 2), eye, upper=True))
    if ok and float((torch.bmm(Q.transpose(1, 2), Q) - eye).abs().max()) <= _ORTH_TOL:
        return Q
    Q = torch.zeros_like(H)
    for j in range(k):
        v = H[:, :, j]
        for attempt in range(2):
            for _ in range(2):
                if j:
                    v = v - torch.bmm(Q[:, :, :j], torch.bmm(Q[:, :, :j].transpose(1, 2), v[:, :, None]))[:, :, 0]
            nrm = torch.linalg.vector_norm(v, dim=1)
            weak = nrm < 1e-6
            if attempt == 1 or not bool(weak.any()):
                break
            v = torch.where(weak[:, None], torch.randn(v.shape, generator=gen, dtype=v.dtype, device=v.device), v)
        Q[:, :, j] = v / nrm.clamp_min(1e-300)[:, None]
    return Q


def _leading_isometries(mats, k):
    """(U, s, Vh) of the leading min(k, R, C) singular triplets of each sector
    matrix, s descending; matrices too small for the iteration go through the
    Jordan-Wielandt factorisation and are cut to the same count."""
    out = [None] * len(mats)
    small, iterate = [], []
    big = [j for j, M in enumerate(mats) if min(M.shape) >= LEADING_MIN_RATIO * k]
    norms = {}
    if big:
        nrm = torch.stack([torch.linalg.matrix_norm(mats[j]) for j in big]).cpu().numpy()
        norms = dict(zip(big, nrm))
    for j, M in enumerate(mats):
        # the iteration works on M / |M|_F; zero or non-finite matrices take
        # the full factorisation
        if j in norms and np.isfinite(norms[j]) and norms[j] > 0:
            iterate.append(j)
        else:
            small.append(j)
    # groups [R, C, smallest member's elements, members]: largest matrices
    # first, each joins the first group whose padded shape (the largest rows
    # and columns of its members) stays within LEADING_PAD times the elements
    # of every member
    groups = []
    for j in sorted(iterate, key=lambda j: (-mats[j].shape[0] * mats[j].shape[1], j)):
        R, C = mats[j].shape
        for g in groups:
            Rg, Cg = max(g[0], R), max(g[1], C)
            if Rg * Cg <= LEADING_PAD * min(g[2], R * C):
                g[0], g[1], g[2] = Rg, Cg, min(g[2], R * C)
                g[3].append(j)
                break
        else:
            groups.append([R, C, R * C, [j]])
    redo = list(small)
    LEADING_STATS["calls"] += 1
    for R, C, _, js in groups:
        dev = mats[js[0]].device
        p = 2 * k
        Mb = torch.zeros((len(js), R, C), dtype=torch.float64, device=dev)
        rmask = torch.zeros((len(js), R, 1), dtype=torch.float64, device=dev)
        cmask = torch.zeros((len(js), C, 1), dtype=torch.float64, device=dev)
        for b, j in enumerate(js):
            r, c = mats[j].shape
            Mb[b, :r, :c] = mats[j] / float(norms[j])
            rmask[b, :r] = 1.0
            cmask[b, :c] = 1.0
        Mt = Mb.transpose(1, 2).contiguous()
        gen = torch.Generator(device=dev)
        gen.manual_seed(0x51DE)
        # random directions (start block, completions) only in the rows and
        # columns a member owns, so its padding stays exactly zero
        V, _, ok = _orth(torch.randn((len(js), C, p), generator=gen, dtype=torch.float64, device=dev) * cmask)
        Y = torch.bmm(Mb, V)
        conv = torch.zeros(len(js), dtype=torch.bool, device=dev)
        it = 0
        complete = False  # rank-deficient members seen: complete null directions
        svd_failed = False
        while it < LEADING_MAX_ITER and js:
            if complete:
                U, ok1 = _orth_complete(Y, torch.randn(Y.shape, generator=gen, dtype=Y.dtype, device=dev) * rmask)
                Z = torch.bmm(Mt, U)
                V, ok2 = _orth_complete(Z, torch.randn(Z.shape, generator=gen, dtype=Z.dtype, device=dev) * cmask)
                Rv = torch.bmm(V.transpose(1, 2), Z)  # = V^T M^T U, so R^T = U^T M V
            else:
                U, _, ok1 = _orth(Y)
                V, Rv, ok2 = _orth(torch.bmm(Mt, U))
            ok &= ok1 & ok2
            ok_host = ok.cpu().numpy()
            if not ok_host.all() and not complete:
                # numerically rank-deficient blocks: redo this iteration with completion
                complete = True
                ok = torch.ones(len(js), dtype=torch.bool, device=dev)
                continue
            if not ok_host.all():
                LEADING_STATS["fb_orth"] += int((~ok_host).sum())
                # rank-deficient blocks leave for the fallback; the others
                # restart this iteration from their current V
                keep = torch.as_tensor(np.flatnonzero(ok_host), device=dev)
                redo.extend(j for j, good in zip(js, ok_host) if not good)
                js = [j for j, good in zip(js, ok_host) if good]
                if not js:
                    break
                Mb, Mt, V, rmask, cmask = Mb[keep], Mt[keep], V[keep], rmask[keep], cmask[keep]
                Y = torch.bmm(Mb, V)
                ok = torch.ones(len(js), dtype=torch.bool, device=dev)
                conv = torch.zeros(len(js), dtype=torch.bool, device=dev)
                continue
            it += 1
            try:
                X, s, Yv = _small_svd(Rv.transpose(1, 2).contiguous())
            except torch.linalg.LinAlgError:
                LEADING_STATS["fb_svd"] += len(js)
                svd_failed = True
                conv = torch.zeros_like(conv)
                break
            Y = torch.bmm(Mb, V)
            Xk = X[:, :, :k]
            Yk = Yv[:, :, :k].contiguous()
            res = torch.bmm(Y, Yk) - torch.bmm(U, Xk) * s[:, None, :k]
            r = torch.linalg.vector_norm(res, dim=1) / s[:, :1].clamp_min(1e-300)
            resolved = s[:, :k] >= LEADING_RESOLVED * s[:, :1]
            conv = ((r <= LEADING_TOL) | ~resolved).all(dim=1)
            LEADING_STATS["iterations"] += 1
            if bool(conv.all()):
                break
        LEADING_STATS["matrices"] += len(js)
        if not js:
            continue
        conv_host = conv.cpu().numpy()
        if not svd_failed:
            LEADING_STATS["fb_conv"] += int((~conv_host).sum())
        if not conv_host.any():
            redo.extend(js)
            continue
        # the (x, y) halves of a Jordan-Wielandt eigenvector for s_i resolve
        # to eps * s_0 / s_i: the kept ones are made orthonormal in s order
        Xk = _orthonormal_ordered(Xk.contiguous(), gen)
        Yk = _orthonormal_ordered(Yk, gen)
        Uk = torch.bmm(U, Xk)
        Vhk = torch.bmm(Yk.transpose(1, 2), V.transpose(1, 2))
        sk = s[:, :k] * torch.tensor([norms[j] for j in js], dtype=torch.float64, device=dev)[:, None]
        for b, j in enumerate(js):
            if conv_host[b]:
                r, c = mats[j].shape
                out[j] = (Uk[b, :r], sk[b], Vhk[b, :, :c])
            else:
                redo.append(j)
    if redo:
        LEADING_STATS["fallback"] += len(redo) - len(small)
        for j, (U, sv, Vh) in zip(redo, _factor_all([mats[i] for i in redo], "isometries")):
            n = min(k, sv.numel())
            out[j] = (U[:, :n], sv[:n], Vh[:n, :])
    return out


# singular values below this fraction of the largest are treated as noise
_NOISE = 1e-12


def _orthonormal_columns(X):
    Q, R = torch.linalg.qr(X)
    d = torch.sign(torch.diagonal(R))
    d = torch.where(d == 0, torch.ones_like(d), d)
    return Q * d[None, :]


def _keep_count(s_host, max_singular_values, max_truncation_err):
    n_total = len(s_host)
    n_keep = n_total
    if max_truncation_err is not None and n_total > 0:
        tot = float(np.sum(s_host ** 2))
        if tot > 0:
            acc = 0.0
            for i in range(n_total - 1, 0, -1):
                acc += float(s_host[i]) ** 2
                if acc / tot > max_truncation_err ** 2:
                    n_keep = i + 1
                    break
    if max_singular_values is not None:
        n_keep = min(n_keep, int(max_singular_values))
    return max(1, min(n_keep, n_total))


def _koszul_negative(parities, perm) -> bool:
    """True when the permutation has an odd number of inversions between legs
    of odd parity (tensor.py:32-50)."""
    n = 0
    for i in range(len(perm)):
        for j in range(i + 1, len(perm)):
            if perm[i] > perm[j] and parities[perm[i]] and parities[perm[j]]:
                n += 1
    return n % 2 == 1


def _assemble(t: SymmetricTensor, lax, rax):
    """Sector matrices of t in sector visit order.

    Each stored block is reshaped to (rows, cols) in its STORED leg order, as
    the reference does (linalg.py:181: ``block.reshape(rows, cols)`` with no
    transpose to left+right order).  When the left legs lead the tensor this
    is the textbook matricisation; otherwise the reference factorises a
    reshuffled sector matrix, and so does this.  U and Vh blocks are then
    shaped by the left / right legs (linalg.py:268-309).
    """
    sectors = _sector_layout(t, lax, rax)
    sym = t.indices[0].symmetry
    perm = list(lax) + list(rax)
    q_list, dims = [], []
    for q, sec in sectors.items():
        if sec["R"] and sec["C"]:
            q_list.append(q)
            dims.append((sec["R"], sec["C"]))
    if not q_list:
        raise ValueError("SVD of a tensor with no blocks")
    # all sector matrices in one zeroed buffer, the blocks copied in by one launch
    starts = np.concatenate([[0], np.cumsum([R * C for R, C in dims])]).astype(np.int64)
    buf = torch.zeros(int(starts[-1]), dtype=torch.float64, device=t.data.device)
    data = t.data.detach()
    copies = []
    for j, q in enumerate(q_list):
        sec = sectors[q]
        C = dims[j][1]
        for i, lk, rk in sec["blocks"]:
            nr = sec["rows"][lk][0]
            nc = sec["cols"][rk][0]
            # Koszul sign of the reordering to (left, right) legs (linalg.py:171-186)
            neg = sym.is_fermionic and _koszul_negative(sym.parity(t.keys[i]), perm)
            copies.append((data, int(t.offsets[i]), nc, buf, int(starts[j]) + sec["row_off"][lk] * C
                           + sec["col_off"][rk], C, nr, nc, -1.0 if neg else 1.0))
    _lib.copy2d(copies, buf.device)
    mats = [buf[starts[j]:starts[j + 1]].view(R, C) for j, (R, C) in enumerate(dims)]
    return sectors, q_list, mats


def truncated_svd_symmetric_many(tensors, left_labels, right_labels, new_bond_label, max_singular_values=None,
                                 max_truncation_err=None, normalize=False, absorb=None, per_sector=False,
                                 leading=False):
    """truncated_svd_symmetric over several tensors with the same labels, whose
    sector eigenproblems share one batched eigh (e.g. many HOTRG instances).
    Each tensor is truncated on its own (global order within the tensor).

    ``leading`` (with absorb="isometries", a max_singular_values cut and no
    truncation-error rule): only the leading max_singular_values triplets of
    each sector are computed (block subspace iteration); the returned full
    spectrum then holds those values only."""
    if torch.is_grad_enabled() and any(t.data.requires_grad for t in tensors):
        raise NotImplementedError("autograd through the block-sparse SVD is not supported (the CTM and DMRG "
                                  "paths factorise constants)")
    prepared = []
    all_mats = []
    for t in tensors:
        labels = t.labels()
        lax = [labels.index(l) for l in left_labels]
        rax = [labels.index(l) for l in right_labels]
        sectors, q_list, mats = _assemble(t, lax, rax)
        prepared.append((t, lax, rax, sectors, q_list, len(all_mats)))
        all_mats.extend(mats)
    if per_sector:
        factors = _factor_per_sector(all_mats, absorb)
    elif leading and absorb == "isometries" and max_singular_values is not None and max_truncation_err is None:
        factors = _leading_isometries(all_mats, max(1, int(max_singular_values)))
    else:
        factors = _factor_all(all_mats, absorb)
    if absorb in ("left", "right") and REFINE > 0 and not per_sector:
        _refine_small(prepared, all_mats, factors, absorb, max_singular_values, max_truncation_err)
    return [_truncate_and_pack(t, lax, rax, sectors, q_list, all_mats[j0:j0 + len(q_list)],
                               factors[j0:j0 + len(q_list)], new_bond_label, max_singular_values,
                               max_truncation_err, normalize, absorb)
            for t, lax, rax, sectors, q_list, j0 in prepared]


# The DMRG state depends on the kept subspace of each sector only (the
# carried factor is the exact product).  From a Gram eigenproblem that
# subspace is accurate to eps * s_0^2 / (s_k^2 - s_{k+1}^2) (s_0 the sector's
# largest, s_k its last kept and s_{k+1} its first dropped singular value);
# sectors where this exceeds REFINE are factorised again through the
# Jordan-Wielandt matrix, accurate to eps * s_0 / (s_k - s_{k+1}) like a
# direct SVD.  Off by default (REFINE = 0): at the sizes that take the Gram
# path the cut's singular-value gaps are resolved, and refinement costs ~4x
# the eigh time of each refined sector.
REFINE = 0.0
_EPS = 2.220446049250313e-16


def _refine_small(prepared, mats, factors, absorb, max_singular_values, max_truncation_err):
    redo = []
    for t, lax, rax, sectors, q_list, j0 in prepared:
        s_list = [factors[j0 + j][1] for j in range(len(q_list))]
        s_all = torch.cat(s_list)
        s_sorted, order = torch.sort(s_all, descending=True, stable=True)
        n_keep = _keep_count(s_sorted.cpu().numpy(), max_singular_values, max_truncation_err)
        sec_of = np.concatenate([np.full(len(v), j, np.int64) for j, v in enumerate(s_list)])
        n_per = np.bincount(sec_of[order[:n_keep].cpu().numpy()], minlength=len(s_list))
        for j, sv in enumerate(s_list):
            nq = int(n_per[j])
            if nq == 0 or nq >= sv.numel():
                continue
            h = sv[[0, nq - 1, nq]].cpu().numpy().astype(np.float64)
            gap = h[1] * h[1] - h[2] * h[2]
            if gap <= 0 or _EPS * h[0] * h[0] / gap > REFINE:
                redo.append(j0 + j)
    if not redo:
        return
    exact = _factor_all([mats[j] for j in redo], "isometries")
    for j, (U, sv, Vh) in zip(redo, exact):
        # orthonormal isometric side (noise-level columns completed), exact
        # product on the carried side
        if absorb == "right":
            U = _orthonormal_columns(U)
            factors[j] = (U, sv, U.T @ mats[j])
        else:
            V = _orthonormal_columns(Vh.T)
            factors[j] = (mats[j] @ V, sv, V.T)


def truncated_svd_symmetric(tensor: SymmetricTensor, left_labels, right_labels, new_bond_label,
                            max_singular_values=None, max_truncation_err=None, normalize=False, absorb=None,
                            per_sector=False):
    """Returns (U, s, Vh, s_full).

    ``absorb``: None (plain SVD), "right"/"left" (s carried by Vh / U, the DMRG
    split of dmrg.py:874-880 done in place), or "isometries" (U and Vh both
    isometries as from the SVD, s not absorbed; the HOTRG use, hotrg.py:137-146).
    """
    return truncated_svd_symmetric_many([tensor], left_labels, right_labels, new_bond_label, max_singular_values,
                                        max_truncation_err, normalize, absorb, per_sector)[0]


def _truncate_and_pack(t, lax, rax, sectors, q_list, mats, factors, new_bond_label, max_singular_values,
                       max_truncation_err, normalize, absorb):
    dev = t.data.device
    s_list = [f[1] for f in factors]
    s_cat = torch.cat(s_list)
    s_sorted, order = torch.sort(s_cat, descending=True, stable=True)
    s_host = s_sorted.cpu().numpy()
    order = order.cpu().numpy()
    n_keep = _keep_count(s_host, max_singular_values, max_truncation_err)
    sec_of = np.concatenate([np.full(len(s), j, np.int64) for j, s in enumerate(s_list)])
    kept = order[:n_keep]
    bond_charges = np.array([q_list[sec_of[g]] for g in kept], np.int32)
    n_per = np.bincount(sec_of[kept], minlength=len(q_list))
    s_final = s_sorted[:n_keep]
    s_sum = float(s_final.sum())
    if normalize and s_sum > 0:
        s_final = s_final / s_sum
    sym = t.indices[0].symmetry
    bond_out = TensorIndex(sym, bond_charges, FlowDirection.OUT, new_bond_label)
    bond_in = TensorIndex(sym, bond_charges, FlowDirection.IN, new_bond_label)
    U_items, V_items = [], []  # (key, block shape, rows, cols, factor, first row, first column, scale)
    for j, q in enumerate(q_list):
        nq = int(n_per[j])
        if nq == 0:
            continue
        U, _, Vh = factors[j]
        Uq = U[:, :nq]
        Vq = Vh[:nq, :]
        if absorb == "isometries" and float(factors[j][1][nq - 1]) <= _NOISE * float(s_host[0]):
            # kept vectors at the noise floor (e.g. zero singular values when
            # nothing is truncated): the (u, v) split of their Jordan-Wielandt
            # eigenvectors is undetermined, so complete both sides to
            # orthonormal sets (QR keeps the well-resolved leading columns).
            # Gram and direct-SVD factors are orthonormal by construction.
            Uq = _orthonormal_columns(Uq)
            Vq = _orthonormal_columns(Vq.T).T
        # with absorb the Gram factors already carry s on the absorbing side
        su = sv = 1.0
        if normalize and absorb in ("left", "right") and s_sum > 0:
            if absorb == "left":
                su = 1.0 / s_sum
            else:
                sv = 1.0 / s_sum
        sec = sectors[q]
        for lk, (nr, shp) in sec["rows"].items():
            U_items.append((lk + (q,), shp + (nq,), nr, nq, Uq, sec["row_off"][lk], 0, su))
        for rk, (nc, shp) in sec["cols"].items():
            V_items.append(((q,) + rk, (nq,) + shp, nq, nc, Vq, 0, sec["col_off"][rk], sv))
    U_idx = tuple(t.indices[a] for a in lax) + (bond_out,)
    V_idx = (bond_in,) + tuple(t.indices[a] for a in rax)
    U_t, V_t = _pack_copies(U_idx, U_items, dev), _pack_copies(V_idx, V_items, dev)
    return U_t, s_final, V_t, s_sorted


def _addressable(X):
    """(1-D view over the whole storage of X, element offset of X[0, 0] in it,
    row pitch) for a 2-D float64 X whose columns are contiguous (a copy of X
    is made when they are not)."""
    if X.shape[1] > 1 and X.stride(1) != 1:
        X = X.contiguous()
    pitch = X.stride(0) if X.shape[0] > 1 else X.shape[1]
    n = X.untyped_storage().nbytes() // X.element_size()
    return X.as_strided((n,), (1,), 0), X.storage_offset(), pitch


def _pack_copies(indices, items, device):
    """Packed SymmetricTensor of blocks given as sub-rectangles of factor
    matrices: items (key, block shape, rows, cols, source matrix, first row,
    first column, scale); the blocks are laid out in sorted key order and
    filled by one batched copy."""
    items = sorted(items, key=lambda it: it[0])
    r = len(indices)
    keys = np.array([it[0] for it in items], np.int32).reshape(len(items), r)
    shapes = np.array([it[1] for it in items], np.int32).reshape(len(items), r)
    offs, tot = SymmetricTensor._layout(shapes)
    data = torch.empty(tot, dtype=torch.float64, device=device)
    copies = []
    for (key, shp, rows, cols, X, r0, c0, scale), off in zip(items, offs):
        base, xoff, pitch = _addressable(X)
        copies.append((base, xoff + r0 * pitch + c0, pitch, data, int(off), cols, rows, cols, scale))
    _lib.copy2d(copies, data.device)
    return SymmetricTensor._wrap(indices, keys, shapes, offs, data)


def svd(tensor: Tensor, left_labels: Sequence[Label], right_labels: Sequence[Label], new_bond_label: Label = "bond",
        max_singular_values: int | None = None, max_truncation_err: float | None = None, normalize: bool = False):
    """(U, s, Vh, s_full) of a SymmetricTensor or DenseTensor (linalg.py:665-825)."""
    labels = tensor.labels()
    for l in list(left_labels) + list(right_labels):
        if l not in labels:
            raise ValueError(f"Label {l!r} not in tensor labels {labels}")
    if set(left_labels) & set(right_labels):
        raise ValueError("left_labels and right_labels overlap")
    if len(left_labels) + len(right_labels) != len(labels):
        raise ValueError("left_labels + right_labels must cover every leg")
    if isinstance(tensor, SymmetricTensor):
        U, s, Vh, s_full = truncated_svd_symmetric(tensor, left_labels, right_labels, new_bond_label,
                                                   max_singular_values, max_truncation_err, normalize)
        return U, s, Vh, s_full
    # dense tensors: one FP64 SVD on the device
    lax = [labels.index(l) for l in left_labels]
    rax = [labels.index(l) for l in right_labels]
    d = tensor.todense().permute(*(lax + rax))
    ls = d.shape[: len(lax)]
    rs = d.shape[len(lax):]
    M = d.reshape(int(np.prod(ls)), int(np.prod(rs)))
    U, s, Vh = torch.linalg.svd(M, full_matrices=False)
    n = len(s)
    k = n if max_singular_values is None else max(1, min(n, int(max_singular_values)))
    if max_truncation_err is not None:
        sq = (s ** 2).cpu().numpy()
        tot = float(sq.sum())
        if tot > 0:
            acc = 0.0
            for i in range(n - 1, 0, -1):
                acc += float(sq[i])
                if acc / tot > max_truncation_err ** 2:
                    k = min(k, i + 1)
                    break
    s_full = s
    s = s[:k]
    if normalize and float(s.sum()) > 0:
        s = s / s.sum()
    sym = tensor.indices[0].symmetry
    zero = np.zeros(k, np.int32)
    U_t = DenseTensor(U[:, :k].reshape(tuple(ls) + (k,)),
                      tuple(tensor.indices[i] for i in lax) + (TensorIndex(sym, zero, FlowDirection.OUT, new_bond_label),))
    V_t = DenseTensor(Vh[:k, :].reshape((k,) + tuple(rs)),
                      (TensorIndex(sym, zero, FlowDirection.IN, new_bond_label),) + tuple(tensor.indices[i] for i in rax))
    return U_t, s, V_t, s_full


truncated_svd = svd
