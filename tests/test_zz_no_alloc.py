"""GPU test of the allocation-free Lanczos entry point (tnx_lanczos_noalloc).

It is called through the C ABI and must reproduce tnx_lanczos bit for bit on
the same plan and start vector: two solves in a row on one device (the second
on a smaller problem, so the shared scratch holds the first call's larger
contents), with and without re-orthogonalisation.  A solve larger than the
scratch is refused on the host before any device work.
"""

import ctypes as C
import importlib

import numpy as np
import pytest
import torch

from oracle import dmrg as od
from tests.helpers import from_oracle

pytestmark = pytest.mark.gpu


def _problems():
    """Two two-site problems: the centre bond of a chi=64 L=20 chain with the
    converged-state charge profile, then the first bond of an L=8 chain."""
    pd = importlib.import_module("tenax_amd.algorithms.dmrg")
    from tenax_amd import build_mpo_heisenberg
    from tenax_amd.workloads import synthetic_dmrg_mps

    out = []
    L = 20
    mpo_tn, mps_tn = build_mpo_heisenberg(L), synthetic_dmrg_mps(L, 64, seed=2)
    mpo = [mpo_tn.get_tensor(i) for i in range(L)]
    mps = [mps_tn.get_tensor(i) for i in range(L)]
    c = L // 2 - 1
    left = pd._build_trivial_left_env_symmetric()
    for i in range(c):
        left = pd._update_left_env_symmetric(left, mps[i], mpo[i])
        left = left * (1.0 / float(left.max_abs()))
    right = pd._build_trivial_right_env_symmetric()
    for i in range(L - 1, c + 1, -1):
        right = pd._update_right_env_symmetric(right, mps[i], mpo[i])
        right = right * (1.0 / float(right.max_abs()))
    theta, _, _ = pd._theta_padded(mps[c], mps[c + 1])
    out.append(pd.TwoSiteProblem(theta * (1.0 / float(theta.norm())), left, mpo[c], mpo[c + 1], right))

    mpo = od.build_mpo_heisenberg(8)
    mps = od.build_random_symmetric_mps(8, bond_dim=4, seed=3)
    _, _, mps, _, _ = od.dmrg(mpo, mps, od.DMRGConfig(max_bond_dim=12, num_sweeps=1, lanczos_max_iter=8))
    oR = od.trivial_right_env()
    for k in range(7, 1, -1):
        oR = od.update_right_env(oR, mps[k], mpo[k])
    theta, _, _ = pd._theta_padded(from_oracle(mps[0]), from_oracle(mps[1]))
    out.append(pd.TwoSiteProblem(theta, from_oracle(od.trivial_left_env()), from_oracle(mpo[0]),
                                 from_oracle(mpo[1]), from_oracle(oR)))
    assert out[1].total < out[0].total
    return out


def _solve(fn, prob, x, max_iter, reorth):
    from tenax_amd import _lib

    L, Wl, Wr, R = prob.envs
    e = C.c_double()
    n = C.c_int32()
    rc = fn(prob.plan.ptr, _lib.dptr(L.data), _lib.dptr(Wl.data), _lib.dptr(Wr.data), _lib.dptr(R.data),
            _lib.dptr(x), int(max_iter), 1e-12, int(reorth), 0, C.byref(e), C.byref(n), _lib.stream_handle())
    return rc, e.value, n.value


@pytest.mark.parametrize("reorth", [False, True])
def test_noalloc_lanczos_matches_allocating_entry_point(gpu, reorth):
    from tenax_amd import _lib

    lib = _lib.load()
    for prob in _problems():  # larger problem first, then the smaller one
        print(f"noalloc lanczos: n={prob.total} reorth={reorth}", flush=True)
        x_ref = prob.x.clone()
        x_new = prob.x.clone()
        rc, e_ref, n_ref = _solve(lib.tnx_lanczos, prob, x_ref, 30, reorth)
        _lib.check(rc)
        rc, e_new, n_new = _solve(lib.tnx_lanczos_noalloc, prob, x_new, 30, reorth)
        _lib.check(rc)
        torch.cuda.synchronize()
        assert n_new == n_ref
        assert e_new == e_ref
        assert torch.equal(x_new, x_ref)
        assert np.isfinite(e_new) and float(torch.linalg.vector_norm(x_new)) == pytest.approx(1.0, abs=1e-12)


def test_noalloc_lanczos_refuses_oversized_solve(gpu):
    from tenax_amd import _lib

    lib = _lib.load()
    prob = _problems()[0]
    x = prob.x.clone()
    # (max_iter + 1) * n doubles of Krylov basis beyond the 64 MiB scratch
    max_iter = (64 << 20) // (8 * prob.total) + 1
    rc, _, _ = _solve(lib.tnx_lanczos_noalloc, prob, x, max_iter, False)
    assert rc == _lib.TNX_ERR_CAPACITY
    assert "use tnx_lanczos" in _lib.last_error()
    torch.cuda.synchronize()
    assert torch.equal(x, prob.x)
