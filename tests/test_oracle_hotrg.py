"""Pins the HOTRG oracle (oracle/hotrg.py) against what the mathematics fixes.

* Untruncated paired HOTRG steps (U on both pairs) followed by the periodic
  trace give ln Z of a finite torus exactly: checked against brute-force
  enumeration of 2x1, 2x2, 4x2 and 4x4 Ising tori, for the Z2 and the dense
  tensor and both contraction orders (a transposed leg, a swapped copy or a
  wrong normalisation weight fails).  The reference form (Vh^dagger on the
  down pair) is exact only while conj(U) Vh = I, e.g. for a symmetric PSD
  traced environment: true for the first two Z2 steps, checked the same way.
* Onsager's quadrature (trg.py:245-291) against the closed form at T_c,
  ln Z/N = 2G/pi + ln(2)/2 (G = Catalan), and the high-temperature series
  ln 2 + 2 ln cosh K + t^4 + 2 t^6 + 9/2 t^8 (t = tanh K).
* The Z2 site tensor has the 8 blocks the reference's test asserts
  (test_hotrg.py:292-293) and is the Hadamard rotation of the dense one.
* The reference's accuracy tests (test_hotrg.py:122-317): chi=16, 20 steps
  within 0.5% / 1% / 2% of Onsager.
"""

import numpy as np
import pytest

from oracle import hotrg as oh

CATALAN = 0.915965594177219015


def _lnz_untruncated(T, steps, order, paired):
    total, Tf, _ = oh.hotrg(T, 4096, steps, order=order, return_tensor=True, paired=paired)
    return (2 ** steps) * total + np.log(oh.trace_final(Tf))


TORI = [(1, 2, 1), (2, 2, 2), (3, 4, 2), (4, 4, 4)]


@pytest.mark.parametrize("steps,nx,ny", TORI)
@pytest.mark.parametrize("kind", ["z2", "dense"])
@pytest.mark.parametrize("order", ["reference", "chi5"])
def test_untruncated_paired_hotrg_equals_bruteforce_torus(steps, nx, ny, kind, order):
    beta = 0.37
    T = oh.ising_tensor_z2(beta) if kind == "z2" else oh.ising_tensor_dense_symt(beta)
    got = _lnz_untruncated(T, steps, order, paired=True)
    ref = oh.torus_lnz_bruteforce(nx, ny, beta)
    assert abs(got - ref) < 1e-11 * max(1.0, abs(ref))


@pytest.mark.parametrize("steps,nx,ny", TORI[:2])
@pytest.mark.parametrize("order", ["reference", "chi5"])
def test_untruncated_reference_form_exact_while_psd(steps, nx, ny, order):
    beta = 0.37
    got = _lnz_untruncated(oh.ising_tensor_z2(beta), steps, order, paired=False)
    assert abs(got - oh.torus_lnz_bruteforce(nx, ny, beta)) < 1e-11 * abs(got)


def test_bruteforce_checks_itself():
    # beta = 0: Z = 2^N; one spin row with self-links: E = 2 per spin
    assert abs(oh.torus_lnz_bruteforce(2, 2, 0.0) - 4 * np.log(2)) < 1e-14
    assert abs(oh.torus_lnz_bruteforce(1, 1, 0.5) - np.log(2 * np.exp(1.0))) < 1e-14


def test_onsager_at_critical_point():
    """The integrand is pinned on a fine grid; the reference's 10^4-point
    half-offset grid itself is ~1e-5 off at T_c (log singularity at t = 0)."""
    bc = np.log(1 + np.sqrt(2)) / 2
    exact = 2 * CATALAN / np.pi + np.log(2) / 2
    assert abs(-bc * oh.onsager_free_energy(bc, n_pts=10 ** 6) - exact) < 1e-6
    assert abs(-bc * oh.onsager_free_energy(bc) - exact) < 2e-5


@pytest.mark.parametrize("beta", [0.02, 0.05])
def test_onsager_high_temperature_series(beta):
    """Fine grid vs the series; the default grid's end-interval error is
    O(beta^2 / n_pts) and stays below 1e-6 here."""
    t = np.tanh(beta)
    series = np.log(2) + 2 * np.log(np.cosh(beta)) + t ** 4 + 2 * t ** 6 + 4.5 * t ** 8
    assert abs(-beta * oh.onsager_free_energy(beta, n_pts=10 ** 6) - series) < 1e-8
    assert abs(-beta * oh.onsager_free_energy(beta) - series) < 1e-6


def test_z2_tensor_structure():
    T = oh.ising_tensor_z2(0.3)
    assert len(T.blocks) == 8
    for k in T.blocks:
        assert sum(k) % 2 == 0
    H = np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)
    dense = np.einsum("ua,vb,wc,xd,uvwx->abcd", H, H, H, H, oh.ising_tensor_dense(0.3))
    assert np.abs(T.todense() - dense).max() < 1e-14


def test_orders_agree_truncated():
    T = oh.ising_tensor_z2(0.42)
    a = oh.hotrg(T, 6, 6, order="reference")
    b = oh.hotrg(T, 6, 6, order="chi5")
    assert abs(a - b) < 1e-12


@pytest.mark.parametrize("beta,tol", [(0.2, 0.005), (0.3, 0.01), (0.44, 0.02), (0.6, 0.005)])
def test_reference_accuracy_chi16(beta, tol):
    lnz = oh.hotrg(oh.ising_tensor_z2(beta), 16, 20, order="chi5")
    f = -lnz / beta
    ex = oh.onsager_free_energy(beta)
    assert abs(f - ex) / abs(ex) < tol


@pytest.mark.parametrize("steps,nx,ny", TORI)
def test_dense_reference_form_untruncated_equals_bruteforce(steps, nx, ny):
    """The DenseTensor path (legs transposed before the SVD) is exact on tori
    without truncation: there the traced environment, matricised as
    (u,U)|(d,D), is symmetric PSD for the Ising tensor, so conj(U) Vh = I."""
    beta = 0.37
    T = oh.ising_tensor_dense(beta)
    tot = 0.0
    for k in range(steps):
        T, ln = oh.step_dense(T, 4096, k % 2 == 0)
        tot += ln / 2 ** (k + 1)
    got = 2 ** steps * tot + np.log(np.einsum("uull->", T))
    assert abs(got - oh.torus_lnz_bruteforce(nx, ny, beta)) < 1e-11 * abs(got)


def test_dense_and_z2_free_energies_agree_to_truncation_level():
    """Same physics in two bases: chi=16, 20 steps at beta=0.3."""
    a = oh.hotrg_dense(oh.ising_tensor_dense(0.3), 16, 20)
    b = oh.hotrg(oh.ising_tensor_z2(0.3), 16, 20, order="chi5")
    ex = -0.3 * oh.onsager_free_energy(0.3)
    assert abs(a - ex) / ex < 0.01 and abs(b - ex) / ex < 0.01
