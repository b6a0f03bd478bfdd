"""Oracle: tensor-protocol CTMRG on dense arrays, energies and the AD forward.

Restates, for a DenseTensor site tensor A with legs (u, d, l, r, phys):

* _ctm_tensor.py:346-365  double layer (ket x bar(bra), legs fused (x, X) -> x2
  with the ket index slow, _tensor_utils.py:141-169);
* _ctm_tensor.py:521-578  identity-like initial corners and edges
  (_split_ctm_tensor.py:64-80 for the dense corner);
* _ctm_tensor.py:586-759  the four moves (grown corners and edge, fused legs
  in the order the reference's fuse_indices produces);
* _ctm_tensor.py:233-338  eigh / QR projectors and their application, with
  one deliberate addition: a canonical eigenvector basis (see canonical_top);
* _ctm_tensor.py:767-810  max-abs renormalisation (EPS = 1e-15) and the sweep;
* _ctm_tensor.py:861-917  ctm_tensor with QR warm-up and the corner-SV criterion;
* ad_utils.py:190-229    the QR gauge fix (R's diagonal made non-negative, see
  qr_pos), ad_utils.py:657-685 the AD forward
  (sweep + gauge fix per iteration, no warm-up);
* ipeps.py:1368-1549     double layer with open physical legs, the 2x1 / 1x2
  reduced density matrices and the energy per site.

Pinned by tests/test_oracle_ctm.py: a D=1 product state (the energy is the
closed form 2 <vv|H|vv> / <v|v>^2 for any chi), C4v-symmetric tensors (the
horizontal and vertical bond energies must agree: a transposed leg in either
RDM breaks it), the RDM's trace and hermiticity, chi-convergence of the energy,
and central finite differences of the energy for the gradient.  The
reference's own JAX gradient cannot be produced here: parity unpinned there.
Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import numpy as np

EPS = 1e-15


# ---------------------------------------------------------------- fuse helper
def fuse(x, a, b):
    """_tensor_utils.py:141-169: bring axes a<b together at a's position
    (a slow), reshape."""
    a, b = sorted((a, b))
    nd = x.ndim
    other = [i for i in range(nd) if i not in (a, b)]
    perm = other[:a] + [a, b] + other[a:]
    y = np.transpose(x, perm)
    s = list(y.shape)
    return y.reshape(s[:a] + [s[a] * s[a + 1]] + s[a + 2:])


def double_layer(A):
    """(D^2, D^2, D^2, D^2): legs (u2, d2, l2, r2), fused (ket, bra)."""
    a8 = np.einsum("udlrs,UDLRs->udlrUDLR", A, np.conj(A))
    x = fuse(a8, 0, 4)   # (u2, d, l, r, D, L, R)
    x = fuse(x, 1, 4)    # (u2, d2, l, r, L, R)
    x = fuse(x, 2, 4)    # (u2, d2, l2, r, R)
    return fuse(x, 3, 4)  # (u2, d2, l2, r2)


def double_layer_open(A):
    """ipeps.py:1371-1381: (D^2, D^2, D^2, D^2, d, d)."""
    D, d = A.shape[0], A.shape[4]
    ao = np.einsum("udlrs,UDLRt->uUdDlLrRst", A, np.conj(A))
    return ao.reshape(D * D, D * D, D * D, D * D, d, d)


def init_env(D2, chi):
    """Corners eye(min(chi, D^2)) padded; edges T[i, :, i] = 1 (i < min(chi, D^2))."""
    m = min(chi, D2)
    C = np.zeros((chi, chi))
    C[:m, :m] = np.eye(m)
    T = np.zeros((chi, D2, chi))
    for i in range(m):
        T[i, :, i] = 1.0
    return [C.copy(), C.copy(), C.copy(), C.copy(), T.copy(), T.copy(), T.copy(), T.copy()]


# ---------------------------------------------------------------- projector
CLUSTER_TOL = 1e-10   # eigenvalues closer than this x the largest are one cluster
PIVOT_TOL = 1e-6      # projections of unit vectors shorter than this are skipped


def canonical_signs(P):
    """Each column's first entry with |entry| >= (1 - 1e-8) max|entry| made
    positive (magnitudes equal up to rounding count as ties)."""
    a = np.abs(P)
    idx = np.argmax(a >= (1.0 - 1e-8) * a.max(axis=0, keepdims=True), axis=0)
    sg = np.sign(P[idx, np.arange(P.shape[1])])
    sg[sg == 0] = 1.0
    return P * sg[None, :]


def canonical_top(w, V, k):
    """The top-k eigenvectors (w ascending as eigh returns it) in a canonical
    basis: inside every cluster of (numerically) equal eigenvalues that
    reaches into the kept set, the basis is Gram-Schmidt of the cluster
    projector's columns taken in index order (the projections of e_1, e_2,
    ...), then every column gets canonical signs.

    The reference leaves the basis to the eigensolver, and its moves are not
    covariant under it (the right move relabels C2's projected leg c2_l and
    its unprojected leg c2_d, _ctm_tensor.py:669, while the top move contracts
    c2_d with T2's projected leg): converged energies then move by ~1e-3 with
    mere eigenvector signs, and the identity-like initial environment has
    exactly degenerate spectra.  A canonical basis makes CTM deterministic and
    the same for LAPACK and rocSOLVER."""
    w = w[::-1]
    V = V[:, ::-1]
    n = len(w)
    out = V[:, :k].copy()
    scale = max(abs(w[0]), 1e-300)
    i = 0
    while i < k:
        j = i + 1
        while j < n and abs(w[j] - w[j - 1]) <= CLUSTER_TOL * scale:
            j += 1
        if j - i > 1:
            Vc = V[:, i:j]
            basis = []
            for r in range(Vc.shape[0]):
                v = Vc @ Vc[r, :]
                for b in basis:
                    v = v - (b @ v) * b
                nv = np.linalg.norm(v)
                if nv > PIVOT_TOL:
                    basis.append(v / nv)
                if len(basis) == j - i:
                    break
            B = np.stack(basis, axis=1)
            m = min(j, k) - i
            out[:, i:i + m] = B[:, :m]
        i = j
    return canonical_signs(out)


# Projector freezing: "record" appends every projector computed, "replay"
# returns the recorded ones in order (i counts).  The reference computes its
# projectors under stop_gradient (_ctm_tensor.py:233-296), so the derivative
# its implicit backward takes is that of the CTM map with the projectors held
# fixed; finite differences of the frozen map check it.
FROZEN = {"mode": None, "list": [], "i": 0}


def projector(C1g, C4g, chi, method):
    """_ctm_tensor.py:233-296 (dense path); P: (fused, k), canonical basis."""
    if FROZEN["mode"] == "replay":
        P = FROZEN["list"][FROZEN["i"]]
        FROZEN["i"] += 1
        return P
    P = _projector(C1g, C4g, chi, method)
    if FROZEN["mode"] == "record":
        FROZEN["list"].append(P)
    return P


def _projector(C1g, C4g, chi, method):
    if method not in ("eigh", "qr"):
        raise ValueError(f"Unknown projector_method={method!r}; expected 'eigh' or 'qr'.")
    if method == "qr":
        Q, R = np.linalg.qr(np.concatenate([C1g, C4g], axis=1))
        rho = R @ R.conj().T
        rho = 0.5 * (rho + rho.conj().T)
        w, V = np.linalg.eigh(rho)
        k = min(chi, len(w))
        return canonical_signs(Q @ canonical_top(w, V, k))
    rho = C1g @ C1g.conj().T + C4g @ C4g.conj().T
    rho = 0.5 * (rho + rho.conj().T)
    w, V = np.linalg.eigh(rho)
    k = min(chi, len(w))
    return canonical_top(w, V, k)


def _apply(P, Ca, Cb, Tg):
    """_ctm_tensor.py:299-338: P^T Ca, P^T Cb, P^T Tg P (real tensors)."""
    return P.T @ Ca, P.T @ Cb, np.einsum("fx,fgr,gy->xry", P, Tg, P, optimize=True)


# ---------------------------------------------------------------- moves
def move_left(env, a, chi, method):
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    C1g = fuse(np.einsum("ab,buc->auc", C1, T1), 0, 1)          # (c1_d u2 | t1_r)
    C4g = fuse(np.einsum("gh,hdi->gdi", C4, T3), 0, 1)          # (c4_r d2 | t3_l)
    Ta = np.einsum("alg,udlr->agudr", T4, a, optimize=True)      # t4_d t4_u u2 d2 r2
    Tg = fuse(fuse(Ta, 0, 2), 1, 2)                              # (t4_d u2)(t4_u d2) r2
    P = projector(C1g, C4g, chi, method)
    C1n, C4n, T4n = _apply(P, C1g, C4g, Tg)
    return [C1n, C2, C3, C4n, T1, T2, T3, T4n]


def move_right(env, a, chi, method):
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    C2g = fuse(np.einsum("ce,buc->ebu", C2, T1), 0, 2)          # (c2_d u2 | t1_l)
    C3g = fuse(np.einsum("im,hdi->mhd", C3, T3), 0, 2)          # (c3_l d2 | t3_r)
    Ta = np.einsum("erm,udlr->emudl", T2, a, optimize=True)      # t2_u t2_d u2 d2 l2
    Tg = fuse(fuse(Ta, 0, 2), 1, 2)
    P = projector(C2g, C3g, chi, method)
    C2n, C3n, T2n = _apply(P, C2g, C3g, Tg)
    return [C1, C2n, C3n, C4, T1, T2n, T3, T4]


def move_top(env, a, chi, method):
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    C1g = fuse(np.einsum("ab,alg->blg", C1, T4), 0, 1)          # (c1_r l2 | t4_u)
    C2g = fuse(np.einsum("ce,erm->crm", C2, T2), 0, 1)          # (c2_l r2 | t2_d)
    Ta = np.einsum("buc,udlr->bcdlr", T1, a, optimize=True)      # t1_l t1_r d2 l2 r2
    Tg = fuse(fuse(Ta, 0, 3), 1, 3)                              # (t1_l l2)(t1_r r2) d2
    P = projector(C1g, C2g, chi, method)
    C1n, C2n, T1n = _apply(P, C1g, C2g, Tg)
    return [C1n, C2n, C3, C4, T1n, T2, T3, T4]


def move_bottom(env, a, chi, method):
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    C4g = fuse(np.einsum("gh,alg->hal", C4, T4), 0, 2)          # (c4_u l2 | t4_d)
    C3g = fuse(np.einsum("im,erm->ier", C3, T2), 0, 2)          # (c3_u r2 | t2_u)
    Ta = np.einsum("hdi,udlr->hiulr", T3, a, optimize=True)      # t3_r t3_l u2 l2 r2
    Tg = fuse(fuse(Ta, 0, 3), 1, 3)                              # (t3_r l2)(t3_l r2) u2
    P = projector(C4g, C3g, chi, method)
    C4n, C3n, T3n = _apply(P, C4g, C3g, Tg)
    return [C1, C2, C3n, C4n, T1, T2, T3n, T4]


def renormalize(env):
    return [x / (np.max(np.abs(x)) + EPS) for x in env]


def sweep(env, a, chi, renorm=True, method="eigh"):
    env = move_left(env, a, chi, method)
    env = move_right(env, a, chi, method)
    env = move_top(env, a, chi, method)
    env = move_bottom(env, a, chi, method)
    return renormalize(env) if renorm else env


def sv_diff(s_new, s_old):
    return float(np.max(np.abs(s_new / (np.sum(s_new) + 1e-15) - s_old / (np.sum(s_old) + 1e-15))))


def ctm(A, chi, max_iter=100, conv_tol=1e-8, renorm=True, method="eigh", qr_warmup_steps=3):
    """_ctm_tensor.py:868-917."""
    a = double_layer(A)
    env = init_env(a.shape[0], chi)
    if method == "qr" and qr_warmup_steps > 0:
        w = min(qr_warmup_steps, max_iter)
        for _ in range(w):
            env = sweep(env, a, chi, renorm, "eigh")
        max_iter -= w
    prev = None
    for _ in range(max_iter):
        env = sweep(env, a, chi, renorm, method)
        s = np.linalg.svd(env[0], compute_uv=False)
        if prev is not None and sv_diff(s, prev) < conv_tol:
            break
        prev = s
    return env


def qr_pos(C):
    """QR with R's diagonal made non-negative (the reference takes the raw
    LAPACK factors, whose signs differ between libraries; with the CTM's
    sign sensitivity a fixed convention is needed for reproducible results)."""
    Q, R = np.linalg.qr(C)
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return Q * d[None, :], R * d[:, None]


def gauge_fix(env):
    """ad_utils.py:190-229 (with qr_pos)."""
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    Q1, R1 = qr_pos(C1)
    T1n = np.einsum("ab,bdc->adc", Q1.conj().T, T1)
    T4n = np.einsum("ab,bdc->adc", Q1.conj().T, T4)
    Q2, R2 = qr_pos(C2)
    T1n = np.einsum("adb,bc->adc", T1n, Q2)
    T2n = np.einsum("ab,bdc->adc", Q2.conj().T, T2)
    Q3, R3 = qr_pos(C3)
    T2n = np.einsum("adb,bc->adc", T2n, Q3)
    T3n = np.einsum("adb,bc->adc", T3, Q3)
    Q4, R4 = qr_pos(C4)
    T3n = np.einsum("ab,bdc->adc", Q4.conj().T, T3n)
    T4n = np.einsum("adb,bc->adc", T4n, Q4)
    return [R1, R2, R3, R4, T1n, T2n, T3n, T4n]


def ctm_fixed_point(A, chi, max_iter=100, conv_tol=1e-8, renorm=True, method="eigh"):
    """ad_utils.py:657-685 (the AD forward: sweep + gauge fix, no warm-up)."""
    a = double_layer(A)
    env = init_env(a.shape[0], chi)
    prev = None
    n = 0
    for n in range(1, max_iter + 1):
        env = gauge_fix(sweep(env, a, chi, renorm, method))
        s = np.linalg.svd(env[0], compute_uv=False)
        if prev is not None and sv_diff(s, prev) < conv_tol:
            break
        prev = s
    return env, n


# ---------------------------------------------------------------- energy
def rdm2x1(A, env, d):
    """ipeps.py:1384-1447."""
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    ao = double_layer_open(A)
    UL = np.einsum("ab,buc->auc", C1, T1)
    UR = np.einsum("cuf,fg->cug", T1, C2)
    LL = np.einsum("gi,idj->gdj", C4, T3)
    LR = np.einsum("jdk,mk->jdm", T3, C3)
    Lenv = np.einsum("auc,axg,gdj->ucxdj", UL, T4, LL, optimize=True)
    Renv = np.einsum("cuf,frm,jdm->curjd", UR, T2, LR, optimize=True)
    L1 = np.einsum("ucxdj,udxrst->crjst", Lenv, ao, optimize=True)
    R2 = np.einsum("curjd,udlrtv->cjltv", Renv, ao, optimize=True)
    rdm = np.einsum("crjst,cjruv->stuv", L1, R2, optimize=True).transpose(0, 2, 1, 3)
    m = rdm.reshape(d * d, d * d)
    m = 0.5 * (m + m.conj().T)
    m = m / (np.trace(m) + 1e-15)
    return m.reshape(d, d, d, d)


def rdm1x2(A, env, d):
    """ipeps.py:1450-1516."""
    C1, C2, C3, C4, T1, T2, T3, T4 = env
    ao = double_layer_open(A)
    top = np.einsum("ab,buc,ce->aue", C1, T1, C2, optimize=True)
    row1 = np.einsum("aue,alf,erg->ulfrg", top, T4, T2, optimize=True)
    site1 = np.einsum("ulfrg,udlrst->dfgst", row1, ao, optimize=True)
    T4ao = np.einsum("fmh,pqmnwx->fhpqnwx", T4, ao, optimize=True)
    s12 = np.einsum("abcst,bhaqnwx->chqnstwx", site1, T4ao, optimize=True)
    s12r = np.einsum("chqnstwx,cni->hqistwx", s12, T2, optimize=True)
    bot = np.einsum("hj,jqk,ik->hqi", C4, T3, C3, optimize=True)
    rdm = np.einsum("hqistwx,hqi->stwx", s12r, bot, optimize=True).transpose(0, 2, 1, 3)
    m = rdm.reshape(d * d, d * d)
    m = 0.5 * (m + m.conj().T)
    m = m / (np.trace(m) + 1e-15)
    return m.reshape(d, d, d, d)


def energy(A, env, H):
    """ipeps.py:1519-1549: E = Tr(rho_h H) + Tr(rho_v H)."""
    d = A.shape[4]
    H = H.reshape(d, d, d, d)
    return float(np.real(np.einsum("ijkl,ijkl->", rdm2x1(A, env, d), H)
                         + np.einsum("ijkl,ijkl->", rdm1x2(A, env, d), H)))


def heisenberg_gate():
    """Sz Sz + (S+ S- + S- S+)/2 on two spin-1/2 sites, (d, d, d, d)."""
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T
    H = np.kron(sz, sz) + 0.5 * (np.kron(sp, sm) + np.kron(sm, sp))
    return H.reshape(2, 2, 2, 2)


def loss(A, chi, H, max_iter=100, conv_tol=1e-10, method="qr"):
    """optimize_gs_ad's loss (ipeps.py:2323-2328): normalise, AD forward, energy."""
    An = A / (np.linalg.norm(A) + 1e-10)
    env, _ = ctm_fixed_point(An, chi, max_iter, conv_tol, True, method)
    return energy(An, env, H)


def frozen_loss(A, env_star, projs, chi, H, n_iter, method="qr"):
    """loss() with the projectors held at `projs` (recorded by one sweep at the
    fixed point env_star of the unperturbed tensor): n_iter gauge-fixed
    sweeps of the frozen map from env_star, then the energy."""
    An = A / (np.linalg.norm(A) + 1e-10)
    a = double_layer(An)
    env = env_star
    try:
        FROZEN.update(mode="replay", list=projs, i=0)
        for _ in range(n_iter):
            FROZEN["i"] = 0
            env = gauge_fix(sweep(env, a, chi, True, method))
    finally:
        FROZEN.update(mode=None, list=[], i=0)
    return energy(An, env, H)


def frozen_fixed_point(A, chi, n_iter, method="qr"):
    """Fixed point of the normalised tensor (n_iter gauge-fixed sweeps) and the
    four projectors of one more sweep at it."""
    An = A / (np.linalg.norm(A) + 1e-10)
    env, _ = ctm_fixed_point(An, chi, n_iter, 0.0, True, method)
    try:
        FROZEN.update(mode="record", list=[], i=0)
        sweep(env, double_layer(An), chi, True, method)
        projs = list(FROZEN["list"])
    finally:
        FROZEN.update(mode=None, list=[], i=0)
    return env, projs


def frozen_gradient_fd(A, chi, H, n_iter, h=1e-5, method="qr", directions=None):
    """Central differences (step h) of frozen_loss around A: the full gradient,
    or the directional derivatives along `directions`."""
    env, projs = frozen_fixed_point(A, chi, n_iter, method)

    def d(u):
        return (frozen_loss(A + h * u, env, projs, chi, H, n_iter, method)
                - frozen_loss(A - h * u, env, projs, chi, H, n_iter, method)) / (2 * h)

    if directions is not None:
        return [d(u) for u in directions]
    g = np.zeros_like(A)
    gf = g.reshape(-1)
    for i in range(gf.size):
        e = np.zeros(A.size)
        e[i] = 1.0
        gf[i] = d(e.reshape(A.shape))
    return g


def loss_gradient_fd(A, chi, H, h=1e-5, max_iter=100, conv_tol=1e-12, method="qr"):
    """Central finite differences of loss() in every entry of A (2 * A.size
    converged forward runs): the gradient the reference takes by implicit
    differentiation of ctm_tensor_converge (ad_utils.py:694-764)."""
    g = np.zeros_like(A)
    flat = A.reshape(-1)
    gf = g.reshape(-1)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp = loss((flat + e).reshape(A.shape), chi, H, max_iter, conv_tol, method)
        lm = loss((flat - e).reshape(A.shape), chi, H, max_iter, conv_tol, method)
        gf[i] = (lp - lm) / (2 * h)
    return g


def optimize_gs(A, chi, H, lr, num_steps, gs_conv_tol=1e-8, max_iter=100, conv_tol=1e-12, method="qr",
                grad=None):
    """ipeps.py:2370-2434 (1-site): Adam with optax's defaults (b1=0.9,
    b2=0.999, eps=1e-8, bias-corrected moments) on A, renormalised after each
    update; the best-energy A is returned with its converged energy.
    ``grad(A)`` supplies the gradient of loss() (central differences by
    default)."""
    grad = grad or (lambda x: loss_gradient_fd(x, chi, H, max_iter=max_iter, conv_tol=conv_tol, method=method))
    A = A / (np.linalg.norm(A) + 1e-10)
    m = np.zeros_like(A)
    v = np.zeros_like(A)
    best_e, best_A, prev = float("inf"), A, float("inf")
    energies = []
    for step in range(num_steps):
        e = loss(A, chi, H, max_iter, conv_tol, method)
        energies.append(e)
        g = grad(A)
        if e < best_e:
            best_e, best_A = e, A
        if abs(e - prev) < gs_conv_tol:
            break
        prev = e
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        mh = m / (1 - 0.9 ** (step + 1))
        vh = v / (1 - 0.999 ** (step + 1))
        A = A - lr * mh / (np.sqrt(vh) + 1e-8)
        A = A / (np.linalg.norm(A) + 1e-10)
    Af = best_A / (np.linalg.norm(best_A) + 1e-10)
    return Af, loss(Af, chi, H, max_iter, conv_tol, method), energies
