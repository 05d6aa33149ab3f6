"""CPU oracle for the block-sparse symmetric tensor path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain Python + numpy/scipy, the algorithms of the reference's tensor
core (cpp977/Xped: src/Core/Qbasis.cpp, FusionTree.cpp, treepair.cpp, Tensor.cpp,
TensorBase.cpp, include/Xped/Symmetry/*.hpp, include/Xped/Util/Permutations.hpp,
Constfct.hpp).  Every function cites the reference file:line it follows.

It is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path (``xped_amd``) never calls
into this package and fails loudly when its HIP library is missing.

How each part is pinned (see the tests under tests/):
  * Wigner 3j/6j/9j (exact rational Racah sums): sympy-generated fixtures
    (tests/golden/wigner_fixtures.json) plus orthogonality / special-value identities.
  * U(1) and fermionic coefficients: closed forms, exhaustive small-q tables.
  * Bases, fusion trees, offsets: the reference's own invariants (tests/basis.cpp:44-142,
    tree_tests.hpp:1-167) and an independent brute-force enumerator on tiny bases.
  * Abelian permute: dense-tensor equivalence (tensor_tests.hpp:1-48 pattern) against
    numpy.transpose of the dense expansion.
  * SU(2) permute: round trips p then p^-1 and unitarity of the recoupling maps.
  * GEMM / SVD / QR / eigh: textbook per-block numpy/LAPACK calls plus invariants
    (reconstruction, orthogonality, ordering) and closed-form spectra.
  * mt19937 / uniform_real_distribution / normal_distribution: the C++ standard's known
    answer (10000th output of the default-seeded engine is 4123659995).
"""
from . import rng, symmetry, structure, wigner  # noqa: F401
