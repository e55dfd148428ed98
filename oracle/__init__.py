"""CPU oracle for the exact-transform engine -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the reference algorithms of the one hot
path the engine accelerates (the separable per-axis DP transforms over a
mixed-radix product space, the support-set walks and the dense-field
queries).  It is the checker, never the thing measured or shipped:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
  of ``bench.py`` may import it;
* the product path (``coldcore_amd``) never imports it and fails loudly
  when its HIP library is missing.

Modules
-------
``space``  mixed-radix indexing (reference src/coldcore/space.py:10-45)
``ref``    O(|X|^2) brute force from the four coverage predicates
           (reference src/coldcore/reference/base.py:22-97, covering.py:16-17,
           lee.py:22-25, torus.py:22-25, grid.py:23-25)
``dp``     numpy axis-DP restatement of the reference ``cpu`` backend
           (reference src/coldcore/dp.py:66-367, protocol.py:143-219)
``walsh``  an independent integer Walsh-Hadamard oracle for binary Hamming
           cells (pure mathematics, no reference counterpart)

How each part is pinned (a plausible mistake must fail one of these checks):

* ``ref`` is the textbook definition of the four metrics; it is checked
  against closed-form ball sizes (reference patterns.py:146-162, restated
  independently in tests), hand-worked tiny cells and the reference tests'
  identities (test_reference_backend.py:23-81).
* ``dp`` is checked field-for-field against ``ref`` on the reference's own
  16-cell parity matrix (test_dp_backend.py:21-38), mu in {1, 2}.
* ``walsh`` is checked against ``ref`` on small binary cells and against
  ``dp`` at full size.
"""
