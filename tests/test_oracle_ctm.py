"""Pins the CTMRG oracle (oracle/ctm.py) against closed forms and symmetries.

* D=1 product state |v>^N: every chi gives E = 2 <vv|H|vv> / <v|v>^2 exactly.
* C4v-symmetric site tensors (averaged over the 8 symmetries of the square
  acting on (u, d, l, r)): the horizontal and vertical bond energies agree up
  to the finite-chi asymmetry of the sweep order (a transposed index in either
  RDM contraction or a mislabelled grown edge gives O(1e-2) differences).
* Horizontal chains: a site tensor with trivial vertical bonds,
  A[u,d,l,r,s] = delta_u0 delta_d0 M_s[l,r], is a stack of independent iMPS.
  Its horizontal two-site RDM is the iMPS RDM from the transfer matrix's
  dominant eigenvectors, and its vertical RDM is rho_1 x rho_1.  Checked with
  a random, site-asymmetric two-site operator, which catches swapped sites,
  a transposed edge or a mis-oriented move.
* chi convergence: the energy of a D=2 tensor changes by < 1e-6 between
  chi = 8 and 12.
* The gauge fix leaves corner spectra unchanged and produces upper-triangular
  corners.
"""

import numpy as np
import pytest

from oracle import ctm as oc


def _bond_energies(A, chi, method="eigh"):
    env = oc.ctm(A, chi, max_iter=200, conv_tol=1e-12, method=method)
    H = oc.heisenberg_gate()
    d = A.shape[4]
    eh = float(np.einsum("ijkl,ijkl->", oc.rdm2x1(A, env, d), H))
    ev = float(np.einsum("ijkl,ijkl->", oc.rdm1x2(A, env, d), H))
    return eh, ev


@pytest.mark.parametrize("chi", [2, 4])
def test_product_state_closed_form(chi):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(2)
    A = v.reshape(1, 1, 1, 1, 2)
    H = oc.heisenberg_gate()
    vv = np.kron(v, v)
    exact = 2 * float(vv @ H.reshape(4, 4) @ vv) / float(v @ v) ** 2
    env = oc.ctm(A, chi, max_iter=10)
    assert abs(oc.energy(A, env, H) - exact) < 1e-13


def _c4v(A):
    perms = []
    # (u, d, l, r) images under the dihedral group of the square
    for p in [(0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
              (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0)]:
        perms.append(np.transpose(A, p + (4,)))
    return sum(perms) / len(perms)


@pytest.mark.parametrize("method", ["eigh", "qr"])
def test_c4v_symmetric_bonds_agree(method):
    rng = np.random.default_rng(5)
    A = _c4v(rng.standard_normal((2, 2, 2, 2, 2)))
    # the sweep order (left, right, top, bottom) breaks the symmetry at finite
    # chi: |E_h - E_v| is 2e-5 at chi=4, 3e-7 at chi=8, 5e-8 at chi=16
    eh, ev = _bond_energies(A, 16, method)
    assert abs(eh - ev) < 1e-6


def _imps_rdms(M):
    """Exact two-site and one-site RDMs of the iMPS M[s, l, r]."""
    d, D, _ = M.shape
    E = np.einsum("slr,smn->lmrn", M, M.conj()).reshape(D * D, D * D)
    w, V = np.linalg.eig(E)
    i = np.argmax(np.abs(w))
    R, lam = np.real(V[:, i]).reshape(D, D), np.real(w[i])
    w, V = np.linalg.eig(E.T)
    L = np.real(V[:, np.argmax(np.abs(w))]).reshape(D, D)
    n = np.einsum("ab,ab->", L, R)
    rho2 = np.einsum("ab,sac,tbd,ucx,vdy,xy->sutv", L, M, M.conj(), M, M.conj(), R) / (lam ** 2 * n)
    rho1 = np.einsum("ab,sac,tbd,cd->st", L, M, M.conj(), R) / (lam * n)
    return rho2, rho1


def test_horizontal_chains_match_exact_imps():
    """D=3 chains (their two-site RDM is not swap-symmetric: |rho_12 - rho_21|
    = 5e-2).  The trivial vertical bonds make the CTM spectra degenerate, so
    accuracy is modest and not monotone in chi (5e-5 at chi=12); 1e-3 still
    separates a swapped or mis-oriented contraction by a factor 50."""
    rng = np.random.default_rng(11)
    rng.standard_normal((2, 2, 2))
    d, D = 2, 3
    M = rng.standard_normal((d, D, D))
    gate = rng.standard_normal((d, d, d, d))
    gate = gate + gate.transpose(2, 3, 0, 1)
    rho2, rho1 = _imps_rdms(M)
    assert np.abs(rho2 - rho2.transpose(1, 0, 3, 2)).max() > 1e-2
    A = np.zeros((D, D, D, D, d))
    A[0, 0] = np.transpose(M, (1, 2, 0))
    env = oc.ctm(A, 12, max_iter=400, conv_tol=1e-13)
    assert np.abs(oc.rdm2x1(A, env, d) - rho2).max() < 1e-3
    assert np.abs(oc.rdm1x2(A, env, d) - np.einsum("ik,jl->ijkl", rho1, rho1)).max() < 1e-3


def test_chi_convergence():
    rng = np.random.default_rng(3)
    A = _c4v(rng.standard_normal((2, 2, 2, 2, 2)))
    A = A / np.linalg.norm(A)
    H = oc.heisenberg_gate()
    e8 = oc.energy(A, oc.ctm(A, 8, max_iter=200, conv_tol=1e-12), H)
    e12 = oc.energy(A, oc.ctm(A, 12, max_iter=200, conv_tol=1e-12), H)
    assert abs(e8 - e12) < 1e-6


def test_gauge_fix_properties():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((2, 2, 2, 2, 2))
    env = oc.ctm(A, 6, max_iter=5)
    fixed = oc.gauge_fix(env)
    for c, f in zip(env[:4], fixed[:4]):
        assert np.allclose(np.linalg.svd(c, compute_uv=False), np.linalg.svd(f, compute_uv=False), atol=1e-13)
        assert np.abs(np.tril(f, -1)).max() < 1e-14


def test_invalid_projector_method():
    with pytest.raises(ValueError, match="Unknown projector_method"):
        oc.projector(np.eye(2), np.eye(2), 2, "svd")
