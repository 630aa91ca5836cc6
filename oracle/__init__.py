"""CPU oracle for mpsim's MPS gate-application path.

TEST INFRASTRUCTURE ONLY. This package is a plain-numpy restatement of the
reference algorithm (mpsim/core.py, mpsim/gates.py and the TensorNetwork 0.2.1
SVD split it calls). It exists so that tests, the smoke check and the CPU
baseline leg of bench.py have something to compare the HIP path against. The
product package (mpsim_amd) never imports it; the product path fails loudly
when its HIP library is missing instead of falling back to this code.

Modules:
  mps_oracle   restated MPS class, gates, Haar recipe, TN SVD split semantics
  statevector  FP64 dense state-vector simulator (the independent check)

Pinning (what checks each function against something other than itself):
  - MPS init / max bond dims: reference test expectations restated in
    tests/test_oracle.py (core_test.py:86-163).
  - one/two-site updates, wavefunction, SWAP network: FP64 statevector on
    random circuits plus closed forms (CNOT/SWAP truth tables, Bell, GHZ, QFT),
    core_test.py:409-1258 restated.
  - truncation (TN split): np.linalg.svd spectra, Eckart-Young optimality,
    Bell maxsvals=1 / maxsvals=0 / fraction known answers (core_test.py:905-1174,
    README.md:45-57), bond-dimension growth pins (core_test.py:932-973).
  - inner_product / norm / renormalize: dense <psi|phi> from wavefunctions.
  - haar_random_unitary: unitarity, and the Mezzadri phase fix checked by
    re-factoring the Gaussian draw (Q^H G must be upper-triangular with a
    positive real diagonal).
No oracle function is left "parity unpinned".
"""
