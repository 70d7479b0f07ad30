"""CPU oracle for the Tic-tac AGS/WPCD solve path -- TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the reference's CPU algorithm
(seanbsm/Tic-tac, C++/OpenMP/MKL/GSL).  It exists to check the HIP product,
never to be shipped or measured as the product:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only as the checker;
* the product (``tic-tac_amd/``) must never import, link or execute anything
  from here.

Each function cites the reference file:line it restates.  Modules:

  states     PW/FWP state spaces, kinematics, Gauss-Legendre, unique 2N index
  potential  Malfliet-Tjon and the seeded synthetic NN potential
  swp        V in the WP basis, SWP diagonalisation, on-shell bins
  resolvent  diagonal SWP-basis channel resolvent G
  angular    Wigner 3j/6j/9j (exact rationals), SPHARM Plm, A-tilde
  p123       overlap mask, G-tilde, P123 WP-bin quadrature, COO -> CSC
  solve      C^T / VC block tables, A = C^T (2 P123) V C, Pade and dense solves

How each part is pinned (closed forms, library routines, invariants) is stated
in the module headers; a part with no such check says "parity unpinned".
"""
