"""Work-list validation of real plans (tnx_plan_validate): every grouped-GEMM
work list (default and every tiling) and every mix list the device planners
build, over random contraction structures, the environment and matvec plans
of the L=8 chi=12 state whose first right-environment update once faulted the
GPU, and the chi=2048 centre-bond matvec."""

import importlib
import os

import numpy as np
import pytest

from oracle import core as oc
from oracle import dmrg as od
from tests.helpers import from_oracle

pytestmark = pytest.mark.gpu

TILINGS = [(sh, kc, xc) for sh in (0, 1, 2, 3) for kc in (0, 1, -1, -2) for xc in (0, 1)]


@pytest.fixture(autouse=True)
def validate_every_plan():
    """TNX_VALIDATE=1: the library checks each plan when it is built and when its tiling changes."""
    old = os.environ.get("TNX_VALIDATE")
    os.environ["TNX_VALIDATE"] = "1"
    yield
    if old is None:
        os.environ.pop("TNX_VALIDATE", None)
    else:
        os.environ["TNX_VALIDATE"] = old


def _validate(plan):
    from tenax_amd import _lib

    _lib.check(_lib.load().tnx_plan_validate(plan.ptr, _lib.stream_handle()))


@pytest.mark.parametrize("seed", range(6))
def test_random_pairwise_plans_validate(gpu, seed):
    from tenax_amd.contraction import plans
    from tenax_amd.contraction.contractor import contract

    rng = np.random.default_rng(seed)
    sym = oc.U1() if seed % 2 == 0 else oc.Zn(3)

    def leg(lbl, flow):
        ch = rng.integers(-2, 3, size=int(rng.integers(3, 9))) if sym.n is None else rng.integers(0, 3, size=7)
        return oc.Leg(sym, ch, flow, lbl)

    shared = [leg("k0", oc.OUT), leg("k1", oc.IN)]
    A = (leg("a", oc.IN),) + tuple(shared) + (leg("b", oc.OUT),)
    B = tuple(l.flip_flow() for l in shared) + (leg("c", oc.IN), leg("d", oc.OUT))
    oA = oc.SymT.random_normal(A, (seed, 1))
    oB = oc.SymT.random_normal(B, (seed, 2))
    # the list keeps the earlier plans alive, so that a plan the cache evicts
    # meanwhile cannot hand its id to a new one
    kept = plans.live_pairwise_plans()
    before = {id(p) for p in kept}
    C = contract(from_oracle(oA), from_oracle(oB))  # built with validation on
    assert C.n_blocks > 0
    n = 0
    for plan in plans.live_pairwise_plans():
        if id(plan) in before:
            continue
        _validate(plan)
        for t in TILINGS:
            plan.set_tiling(0, *t)  # validated again by the library
        plan.set_tiling(0, 0, 0, 0)
        n += 1
    assert n > 0


def test_environment_and_matvec_plans_of_the_faulted_state_validate(gpu):
    """The L=8, chi=12, seed-3 state after one oracle sweep (the structure of
    the first right-environment update that faulted): every left and right
    environment plan and every site's matvec plan under every tiling."""
    pd = importlib.import_module("tenax_amd.algorithms.dmrg")
    from tenax_amd.contraction import plans

    mpo = od.build_mpo_heisenberg(8)
    mps = od.build_random_symmetric_mps(8, bond_dim=4, seed=3)
    _, _, mps, _, _ = od.dmrg(mpo, mps, od.DMRGConfig(max_bond_dim=12, num_sweeps=1, lanczos_max_iter=8))
    P = [from_oracle(t) for t in mps]
    W = [from_oracle(t) for t in mpo]
    ops = pd._symmetric_ops()
    kept = plans.live_env_plans()  # alive: their ids cannot be reused
    before = {id(p) for p in kept}
    left = pd._build_left_environments_list(P, W, 8, ops)
    right = pd._build_right_environments_list(P, W, 8, ops)
    new = [p for p in plans.live_env_plans() if id(p) not in before]
    for plan in new:
        _validate(plan)
    for i in range(7):
        r = right[i + 2] if right[i + 2] is not None else ops.build_trivial_right_env()
        theta, _, _ = pd._theta_padded(P[i], P[i + 1])
        prob = pd.TwoSiteProblem(theta, left[i], W[i], W[i + 1], r)
        _validate(prob.plan)
        for st in (1, 3):
            for t in TILINGS:
                prob.plan.set_tiling(st, *t)
            prob.plan.set_tiling(st, 0, 0, 0)


def test_chi2048_centre_matvec_plan_validates(gpu):
    from tenax_amd.workloads import centre_bond_problem

    prob, _, _, _ = centre_bond_problem(100, 2048, 2)
    _validate(prob.plan)
    for st in (1, 3):
        for t in [(0, 0, 1), (0, -1, 0), (0, -1, 1), (1, 0, 1), (2, -1, 1), (3, 0, 1)]:
            prob.plan.set_tiling(st, *t)
        prob.plan.set_tiling(st, 0, 0, 0)
