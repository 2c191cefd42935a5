"""MI355X (gfx950) engine for the reconstruction hot path of the cylindrical
JSCC SPECT code (lpz7777777/JSCC_Recon_SPECT_Cylinder).

Modules named after the reference keep its call signatures:
  process_list_plane_strict     get_compton_backproj_list_single
  process_list_plane_sparse     get_compton_backproj_list_single_sparse
  compton_sparse_ops            ComptonSparseProjector and the coarse/fine operators
  recon_osem_plane              get_weight_single / get_weight_compton / osem_*_mode / run_recon_osem
  recon_osem_plane_sparse       sparse list and joint modes, safe_em_update
  recon_osem_local_cntstat      local SC OSEM with random bin subsets
  recon_osem_local_sparse_jsccsd_only   local sparse JSCCSD OSEM
Engine-facing pieces: engine.ListModeStore (T resident in HBM), ops (tensor
wrappers of the C ABI in include/jscc_hip.h), synth (seeded workloads).
All numerics run in libjscc_hip.so; there is no CPU fallback.
"""

__version__ = "0.1.0"
