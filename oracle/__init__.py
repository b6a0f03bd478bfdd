"""CPU FP64 oracle for the block-sparse contraction path.

TEST INFRASTRUCTURE ONLY. Nothing in ``tenax_amd`` imports, links or calls
this package; only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg use it, and only as the checker.

It is a NumPy restatement of the reference algorithms (tenax, pure Python on
JAX), one function per reference function, each citing the reference
file:line it follows.  Values are float64 NumPy arrays; block structure is
plain Python dicts, exactly as the reference keeps it.

Pinning (what fixes each function independently of its own formula):

* ``core.compute_valid_blocks`` / ``block_slices``: brute-force enumeration of
  every charge tuple on tiny legs, and the nonzero pattern of ``todense``.
* ``contract.contract_symmetric`` / ``blockwise_contract``: dense ``np.einsum``
  of the densified operands (the reference's own cross-check,
  tests/test_contraction.py:329-350, tests/test_dmrg.py:574-618).
* ``linalg.truncated_svd_symmetric``: reconstruction U·diag(s)·Vh against the
  dense SVD spectrum of the densified tensor.
* ``dmrg.dmrg``: exact diagonalisation energies (L=4, L=6, L=4 with 2Sz=2)
  from an independently built Kronecker-product Hamiltonian.
* ``mpo.build_mpo_heisenberg``: known MPO bond charges [0,0,2,-2,0]
  (tests/test_auto_mpo.py:628-656) and dense MPO contraction == ED matrix.
* ``hotrg``: Onsager's closed form for the 2D Ising free energy.

Parity unpinned (an import of JAX would be needed, and JAX is absent):
the exact jax.random initial tensors, LAPACK gauge choices, and the
reference's own per-sweep energy trajectory.  The oracle and the GPU path
share one documented generator (``core.normal``) instead.
"""
