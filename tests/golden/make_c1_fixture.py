"""Writes tests/golden/c1_oracle.json: the FP64 oracle's DMRG trajectory for the
L=20 Heisenberg chain (Jz=Jxy=1, U(1) MPO), initial MPS build_random_symmetric_mps(
20, bond_dim=4, seed=0), max_bond_dim=64, 10 sweeps, lanczos_max_iter=50.

Runs the oracle (oracle/dmrg.py, dense evaluation mode) with and without the
FP64 Lanczos re-orthogonalisation.  Usage: python tests/golden/make_c1_fixture.py
"""

import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import dmrg as od  # noqa: E402


def run(reorth):
    od.MODE["eval"] = "dense"
    t = time.time()
    cfg = od.DMRGConfig(max_bond_dim=64, num_sweeps=10, lanczos_max_iter=50, lanczos_reorth=reorth)
    e, es, mps, errs, conv = od.dmrg(od.build_mpo_heisenberg(20), od.build_random_symmetric_mps(20, 4, seed=0), cfg)
    bonds = []
    for i in range(19):
        leg = mps[i].legs[-1]
        q, m = __import__("numpy").unique(leg.charges, return_counts=True)
        bonds.append({str(int(a)): int(b) for a, b in zip(q, m)})
    return {"energy": e, "energies_per_sweep": es, "converged": conv, "truncation_errors_last_sweep": errs[-38:],
            "bond_multiplicities": bonds, "seconds": time.time() - t}


if __name__ == "__main__":
    out = {"config": {"L": 20, "Jz": 1.0, "Jxy": 1.0, "bond_dim0": 4, "seed": 0, "max_bond_dim": 64,
                      "num_sweeps": 10, "lanczos_max_iter": 50, "lanczos_tol": 1e-12, "convergence_tol": 1e-10},
           "ed_anchor_2sz0": -8.6824733343990,
           "reorth": run(True), "no_reorth": run(False)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: (v["energies_per_sweep"] if isinstance(v, dict) and "energies_per_sweep" in v else v)
                      for k, v in out.items()}, indent=1))
