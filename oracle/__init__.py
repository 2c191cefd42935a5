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
  osem.py     textbook GEMV (adjoint identity), EM fixed point and count
              preservation, safe/apply update special cases
  sparse.py   hand-computable reduce/expand operators (fine 4, stride 2),
              adjoint identity, reference-written golden projector
              (tests/golden/sparse_projector_small.npz)
"""
