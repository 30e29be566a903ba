"""CPU oracle for the CCSD residual / energy path (TEST INFRASTRUCTURE ONLY).

This package restates, in FP64 numpy, eT's closed-shell CCSD ground-state
algorithm in the T1-transformed Cholesky basis (reference
src/wavefunctions/ccsd/omega_ccsd.F90 and the routines it calls).  It exists
only to check the HIP implementation: it may be imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
never by the product package ``et_amd``.

Pinning (how each function is checked independently of retyping its formula):

* ``omega`` / ``energy`` / ``t1_cholesky`` / ``fock_t1``: against a brute-force
  determinant-space evaluation of <mu| exp(-T) H exp(T) |HF> with explicit
  T1 and T2 (tests/brute_force_cc.py) on tiny random systems, and against the
  closed forms Omega(t=0) = (F_ai, g_aibj with 1/2 on the diagonal) and
  E(MP2 guess) = E_MP2.
* ``omega_a2_plus_minus``: against the naive ladder sum_cd g_acbd t_cidj.
* ``omega_a2_batched`` (eT's batched ladder with the C gathers and scatter of
  ``ladder_pm.c``): equal to ``omega_a2_plus_minus`` for every batch length.
* ``t1_cholesky``: against the matrix identity L_t1 = X L Y^T.
* ``diis_extrapolate``: against the closed-form two-vector DIIS solution.
"""
