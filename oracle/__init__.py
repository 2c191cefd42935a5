"""CPU oracle for the reconstruction hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy (FP64, with an FP32 mirror of the reference's
event filters), the algorithms of lpz7777777/JSCC_Recon_SPECT_Cylinder that
the GPU engine replaces.  Each function cites the reference file:line it
follows.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import it, and only as the checker or the timed CPU baseline; the product
package (jscc_recon_spect_cylinder_amd) never imports it.

Pinning (how each function is checked against something other than itself):
  layout.py   polar grid / RotMat / RotMatInv vs checksums of the reference's
              Factors/*.csv files (tests/golden/factors_*.npz)
  cone.py     theta/KN closed forms; sigma_pos vs FP64 finite differences of
              beta; detector variances vs Detector.csv layers; rotation
              equivariance of t_c under RotMat; the full strict chain vs golden
              vectors written by the reference's own function
              (tests/golden/strict_cone_small.npz)
  osem.py     brute-force loops of the GEMV weights; the rotation gather and
              scatter against explicitly column-permuted matrices (random
              non-involutive permutations, so a RotMat/RotMatInv swap fails);
              alpha limits of the joint mode (alpha = 2 -> bin mode, 0 -> list
              mode) and its count identity; EM fixed points (with rotations)
              and count preservation; safe/apply update special cases
  sparse.py   hand-computable reduce/expand operators (fine 4, stride 2),
              adjoint identity, reference-written golden projector digests
              (tests/golden/sparse_projector_golden.npz); materialize against a
              per-element loop with NaN/inf/zero/negative rows, non-finite A
              entries, clamped and half-rounded crystal numbers
  scd.py      thinning against the reference's literal expand-and-compare
              algorithm; down-sampling against outputs written by the
              reference's own function (tests/golden/scd_downsample_golden.npz)
              and its count rules; the list-mode sensitivity against its
              scale invariance, total and proportional-case closed forms
All pins live in tests/test_oracle.py, tests/test_oracle_pins.py and
tests/test_sparse_projector.py (CPU only).
"""
