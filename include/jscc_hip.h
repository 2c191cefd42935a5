/*
 * libjscc_hip — C ABI of the MI355X (gfx950) reconstruction hot path of the
 * cylindrical-detector JSCC SPECT code: the Compton cone back-projection that
 * builds list-mode rows T from two-hit events, and the OSEM projections
 * A.x, A^T(y/Ax) and T^T(1/Tx).
 *
 * Conventions
 *   - every pointer is device memory unless stated; "stream" is a hipStream_t;
 *   - every entry point returns 0 on success and a negative code on failure,
 *     with a message available from jscc_last_error() (thread-local);
 *   - the library never allocates device memory: scratch space is passed in
 *     as (workspace, workspace_bytes), sized by the *_workspace_bytes queries
 *     (jscc_proj_em_noalloc alone uses a static buffer of the library);
 *   - index tables that mirror the reference's 1-based tensors (RotMat,
 *     RotMatInv) are int64 with an element stride, exactly as torch stores a
 *     column of an [P, R] int64 tensor;
 *   - "err" words are device int32 flags OR-ed by kernels when an index is out
 *     of range (1: crystal number, 2: image gather, 4: weight gather,
 *     8: sensitivity gather, 16: thinning position, 32: grid barrier time-out
 *     of the persistent OSEM kernel).  The host raises IndexError when set.
 *   - storage dtypes: 0 = FP32, 1 = BF16, 2 = FP8 e4m3fn (per-row scale),
 *     3 = FP16 (IEEE half, per-row scale).
 *
 * Each entry point names the reference interface (file:line in
 * lpz7777777/JSCC_Recon_SPECT_Cylinder) whose work it performs.
 */
#ifndef JSCC_HIP_H
#define JSCC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */
int jscc_abi_version(void);
const char* jscc_last_error(void);
int jscc_device_cu_count(int* out);

/* ---- list events -------------------------------------------------------- */

/* Energy smearing, energy window, kinematic-domain and same-layer filters,
 * crystal position variances and the per-event cone parameters (64-byte
 * records, see jscc_events.h), one record per input event.
 * Replaces process_list_plane_strict.py:152-199 together with :24-36 and
 * :55-96 (and the identical steps of process_list_plane_sparse.py:40-72).
 * list: [n, list_stride] FP32 rows (cpnum1, e1, cpnum2, e2, ...), 1-based
 * crystal numbers; noise1/noise2: the N(0,1) draws of the reference's
 * torch.randn calls (both NULL: no smearing); list_writeback: where the
 * smeared e1/e2 are written back (the reference mutates its argument in
 * place), may be NULL; detector: [n_det, det_stride] FP32 (x, y, z, ...);
 * n_sysmat_rows: rows of the system matrix the first hit indexes (0: no
 * matrix), checked with torch's index rules like sysmat[cpnum1 - 1].
 * det_layers4 / rear_abs_y describe the detector's |y| layers (exactly four
 * distinct |y| values, the largest being the rear layer).  The *_f32
 * arguments are the FP32 roundings the reference's comparisons use. */
int jscc_events_prepare(const float* list, int64_t n, int64_t list_stride, const float* noise1,
                        const float* noise2, float* list_writeback, const float* detector, int64_t n_det,
                        int64_t det_stride, int64_t n_sysmat_rows, int det_layers4, float rear_abs_y,
                        float delta_r1, float delta_r2,
                        float e0_f32, float ene_res_f32, float thr_max_f32, float thr_min_f32, float thr_sum_f32,
                        double e0, double ene_res, void* events_out, int32_t* flags_out, int32_t* err_out,
                        void* stream);

/* Order-preserving stream compaction: out_idx[k] = k-th i with flags[i] != 0;
 * *out_count = number of set flags.  Replaces the boolean-mask indexing
 * t[flag] of process_list_plane_strict.py:172-198, 244-248. */
int64_t jscc_compact_workspace_bytes(int64_t n);
int jscc_compact_flags(const int32_t* flags, int64_t n, int32_t* out_idx, int32_t* out_count, void* workspace,
                       int64_t workspace_bytes, void* stream);

/* ---- cone rows ---------------------------------------------------------- */

/* One row of T per event ev_idx[r] (events from jscc_events_prepare):
 * t_c[r, p] = exp(-(beta - theta)^2 / (2 sigma^2)) * (KN - sin^2 beta), times
 * the first-hit system-matrix factor, for the n_vox voxels (SoA coordinates,
 * 16-byte aligned, padded to a multiple of 4).  sigma_model 0 = the strict
 * model (energy + propagated position sigma, process_list_plane_strict.py:
 * 39-52, 99-134, 201-228), 1 = a fixed angular sigma (rad).
 * System-matrix factor: sysmat NULL -> 1; sysmat with tap tables NULL ->
 * sysmat[arow, p] (process_list_plane_strict.py:236-240); with tap tables ->
 * w0[p]*sysmat[arow, i0[p]] + w1[p]*sysmat[arow, i1[p]] (the coarse-grid row
 * of process_list_plane_sparse.py:105).
 * Outputs (each optional): out_t [n_rows, ld_out] unnormalised t in FP32 or
 * BF16; out_tc [n_rows, ld_tc] FP32 t_c; stats [n_rows, 4] FP64 =
 * (sum t, sum t^2, sum t_c, max t) -- the inputs of the NaN / zero-sum /
 * effective-support filters of process_list_plane_strict.py:241-248,137-148. */
int jscc_cone_rows(const void* events, const int32_t* ev_idx, int64_t n_rows, const float* coor_x,
                   const float* coor_y, const float* coor_z, int64_t n_vox, const float* sysmat, int64_t ld_sysmat,
                   const int32_t* tap_i0, const int32_t* tap_i1, const float* tap_w0, const float* tap_w1,
                   int sigma_model, double fixed_sigma, void* out_t, int out_dtype, int64_t ld_out, float* out_tc,
                   int64_t ld_tc, double* stats, void* stream);

/* jscc_cone_rows with the row count read on the device: n_rows_dev (int32)
 * is the count of the event compaction that produced ev_idx, so no host read
 * of it precedes the launch.  The grid covers n_rows_max rows (ev_idx, out_t,
 * stats hold that many); rows [count, n_rows_max) get NaN statistics, which
 * jscc_rows_keep_flags drops, and no row values.  Same computation as
 * jscc_cone_rows for the rows below the count. */
int jscc_cone_rows_counted(const void* events, const int32_t* ev_idx, const int32_t* n_rows_dev, int64_t n_rows_max,
                           const float* coor_x, const float* coor_y, const float* coor_z, int64_t n_vox,
                           const float* sysmat, int64_t ld_sysmat, const int32_t* tap_i0, const int32_t* tap_i1,
                           const float* tap_w0, const float* tap_w1, int sigma_model, double fixed_sigma, void* out_t,
                           int out_dtype, int64_t ld_out, float* out_tc, int64_t ld_tc, double* stats, void* stream);

/* jscc_cone_rows_counted (n_rows_dev may be NULL: all n_rows_max rows) that
 * also returns the list-mode sensitivity of the rows it writes:
 *   sens[p] = sum over kept rows i of T_i[p] / S_i
 * with keep_i and S_i as jscc_rows_keep_flags decides them at min_support --
 * torch.sum(t, dim=0) of the normalised, kept T that main_plane.py:348-355
 * scales into Sensi_d (and recon_osem_plane.py:253-256 divides by) -- taken
 * from the rows as stored.  Each row is read back by the thread that wrote it,
 * so the sums cost no second pass over T in HBM.  out_dtype FP32 or BF16: rows
 * as jscc_cone_rows stores them, read back during the cone arithmetic of the
 * work-group's next rows.  out_dtype FP16: the FP32 values of a group wait in
 * a slot of the workspace and, right after the group's statistics, are read
 * back as the codes of jscc_quantize_rows (half(t / q), q = max / 2^14) with
 * row_scale[i] = q / S_i, row_scale * code ~ t / S; a dropped row's codes are
 * zeros; sens is then the column sums of row_scale * code, the rows the OSEM
 * projects.  Persistent grid of jscc_cone_sens_grid(n_vox) work-groups taking
 * groups of rows (8; 4 for FP16) from a ticket counter: the grouping of the
 * FP32 sums, and so the last bits of sens, can differ between runs (rows,
 * codes and statistics do not).  First-hit factor
 * sysmat[arow] or none (no tap tables), no t_c output, n_vox <= 32768.
 * workspace: jscc_cone_sens_workspace_bytes(n_vox, out_dtype) bytes, 16-byte
 * aligned.  Statistics equal jscc_cone_rows's up to the order of the per-row
 * sums (another work-group shape); the row values are the same bits. */
int jscc_cone_sens_grid(int64_t n_vox);
int64_t jscc_cone_sens_workspace_bytes(int64_t n_vox, int out_dtype);
int jscc_cone_rows_sens(const void* events, const int32_t* ev_idx, const int32_t* n_rows_dev, int64_t n_rows_max,
                        const float* coor_x, const float* coor_y, const float* coor_z, int64_t n_vox,
                        const float* sysmat, int64_t ld_sysmat, int sigma_model, double fixed_sigma, void* out_t,
                        int out_dtype, int64_t ld_out, float* row_scale, double* stats, double min_support,
                        float* sens, void* workspace, int64_t workspace_bytes, void* stream);

/* keep[i] = S finite and nonzero and S^2/Q >= min_support
 * (process_list_plane_strict.py:241-246, 145-146); inv_sum (optional) = 1/S
 * for kept rows (the row normalisation of :138, applied at use). */
int jscc_rows_keep_flags(const double* stats, int64_t n, double min_support, int32_t* flags, float* inv_sum,
                         void* stream);

/* dst[i, col_off + p] = src[idx[i], p] / stats[idx[i], stat_field] (FP64 divide),
 * optional dst[i, 0] = lead[idx[i]] (packed sparse rows).  Produces the
 * row-normalised, compacted T / t_compton of process_list_plane_strict.py:
 * 137-148, 244-253 and the packed rows of compton_sparse_ops.py:359-368. */
int jscc_rows_gather_normalize(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, const int32_t* idx,
                               int64_t n_out, const double* stats, int stat_field, const float* lead, float* dst,
                               int64_t ld_dst, int64_t col_off, void* stream);

/* t_single rows: dst[i, :] = sysmat[arow[i], :] / sum (process_list_plane_strict.py:237, 140). */
int jscc_arows_normalize(const float* sysmat, int64_t lda, int64_t n_cols, const int32_t* arow, int64_t n_out,
                         float* dst, int64_t ld_dst, void* stream);

/* FP8 e4m3fn storage of rows with per-row scale (row_scale * code ~ t / S,
 * S and max from stats); stats NULL: rows of a system matrix, scaled by
 * their own largest |value| and not normalised (row_scale * code ~ a).
 * No counterpart in the reference (precision mode of the MI355X engine). */
int jscc_quantize_fp8_rows(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, int64_t n_rows,
                           const double* stats, uint8_t* dst, int64_t ld_dst, float* row_scale, void* stream);

/* The same for dst_dtype 2 (FP8 e4m3fn, row max coded as 448) or 3 (IEEE half,
 * row max coded as 2^14, so values down to 2^-28 of the row max stay normal).
 * No counterpart in the reference (precision modes of the MI355X engine). */
int jscc_quantize_rows(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, int64_t n_rows,
                       const double* stats, void* dst, int dst_dtype, int64_t ld_dst, float* row_scale,
                       void* stream);

/* ---- coarse/fine polar-grid operators ----------------------------------- */

/* out[r, j] = sum_k val[k] * in[r, col[k]], k in [ptr[j], ptr[j+1]): the
 * CSR form of compton_sparse_ops.py:293-356 (reduce_fine_rows_to_coarse,
 * upsample_coarse_rows_to_fine, project_fine_rows_to_coarse_adjoint,
 * reduce_fine_image_to_coarse, upsample_coarse_weight_to_fine). */
int jscc_rows_csr_map(const float* in, int64_t ld_in, int64_t n_rows, const int32_t* ptr, const int32_t* col,
                      const float* val, int64_t n_out, float* out, int64_t ld_out, void* stream);

/* materialize_sparse_event_rows_to_fine (compton_sparse_ops.py:380-400):
 * out[r, :] = nan_to_num(expand(t_c) * sysmat[clamp(round(c)-1)]), divided by
 * its row sum when normalize != 0; flags[r] = row sum finite and > 0.
 * normalize = 0 gives the unnormalised rows of the grouped
 * get_weight_compton_sparse (recon_osem_plane_sparse.py:22-57). */
int jscc_materialize_rows(const float* packed, int64_t ld_packed, int64_t n_rows, int64_t n_coarse,
                          const int32_t* ptr, const int32_t* col, const float* val, int64_t n_fine,
                          const float* sysmat, int64_t lda, int64_t M, float* out, int64_t ld_out, int normalize,
                          int32_t* flags, void* stream);

/* Fused list-mode weight of packed coarse rows [cpnum1, t_c(n_zc*n_lc)], the
 * sparse production mode, with no fine row written:
 *   t = expand(t_c) * sysmat[clamp(round(cpnum1)-1, 0, n_arows-1)], expand =
 *   z_expand (x) layer_expand given as two taps per fine index: z_idx/z_w
 *   [n_zf][2] (coarse z index, weight), l_idx/l_w [n_lf][2] (coarse layer index,
 *   weight; l_idx = NULL: identity layer map);
 *   normalize != 0: compton_sparse_ops.py:380-400 then
 *     recon_osem_local_sparse_jsccsd_only.py:13-19 (element-wise nan_to_num, rows
 *     with a non-finite or non-positive sum dropped, t/sum, clamped ratio);
 *   normalize == 0: get_weight_compton_sparse of recon_osem_plane_sparse.py:22-57
 *     (unnormalised rows, clamped ratio).
 * w[j] (+)= out_scale * nan_to_num(sum_i t_i[src(j)] / max(t_i . x_rot, eps)),
 * x_rot / src as in jscc_proj.  rows (optional): int32 order in which the
 * n_rows packed rows are taken (row i of the call is packed row rows[i]).
 * za_idx/za_w [n_zc][n_adj] (optional): for each coarse z slice the fine z
 * slices with a tap on it and their weights (the transpose of z_expand,
 * padded with weight 0).  With an identity layer map, n_adj = 4 and at most
 * half as many coarse as fine pixels (up to 25600), the rows of one first-hit
 * crystal are weighted in grouped form (the adjoints E^T(x*A[c]) and E^T A[c]
 * once per run of equal crystal numbers in the row order, as the reference's
 * grouped variant does); rows whose products could be non-finite are
 * weighted product by product.  Fine pixels n_zf*n_lf <= 32768 (a multiple
 * of 4), coarse pixels <= 16384.  Error word bits: 2 x_perm, 4 inv_perm,
 * 64 tap table index out of range. */
int64_t jscc_coarse_weight_workspace_bytes(int64_t n_rows, int64_t n_fine);
int jscc_coarse_weight(const float* packed, int64_t ld_packed, const int32_t* rows, int64_t n_rows,
                       const int32_t* z_idx, const float* z_w, int n_zf, int n_zc, const int32_t* za_idx,
                       const float* za_w, int n_adj, const int32_t* l_idx, const float* l_w, int n_lf, int n_lc,
                       const float* sysmat, int64_t lda, int64_t n_arows, const float* x, const int64_t* x_perm,
                       int64_t x_perm_stride, int64_t x_len, float eps, int use_eps, int normalize,
                       const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate, float* w,
                       int64_t w_len, void* workspace, int64_t workspace_bytes, int32_t* err, void* stream);

/* ---- OSEM projections --------------------------------------------------- */

/* Fused one-pass projection of a block of rows (A or T), replacing
 * get_weight_single (recon_osem_plane.py:28-30; clamped twin
 * recon_osem_local_cntstat.py:6-8) and get_weight_compton
 * (recon_osem_plane.py:33-42), together with the index_select rotation
 * gather/scatter and alpha weighting of osem_*_mode (recon_osem_plane.py:45-125).
 *   mode 0: w (+)= out_scale * G^T [ s * (y / max(s * (M x_rot), eps)) ] rows
 *   mode 1: fwd_out[i] = s_i * (M x_rot)_i
 *   mode 2: w (+)= out_scale * G^T [ s * coef ]
 *   mode 3: w (+)= out_scale * G^T [ s ]  (column sums)
 * where row i of the call is stored row (rows ? rows[row_base+i] : row_base+i),
 * s = row_scale[stored row] (or 1), x_rot[p] = x[x_perm[p]-1] (or x), and the
 * gather G picks partial[inv_perm[j]-1] for output j (or j).  y may be NULL
 * (numerator 1).  use_eps = 0 reproduces the unclamped reference division.
 * flags bit 0: y is indexed by the stored row id (random bin subsets given as
 * a row list, recon_osem_local_cntstat.py:11-25, without copying A);
 * bit 1: the block weight is passed through nan_to_num(0,0,0) before out_scale
 * (recon_osem_local_sparse_jsccsd_only.py:13-19).
 * Any n_cols up to jscc_proj_max_cols: rows are read in whole vectors of
 * jscc_proj_vec_width(dtype) elements, so ld must be a multiple of it that
 * covers n_cols rounded up to it, and every row must extend that far in
 * memory; what the row holds past n_cols is never used (the ragged last
 * vector is cleared past n_cols before the dot products). */
int jscc_proj_vec_width(int dtype);
int jscc_proj_max_cols(int dtype);
/* The largest number of work-groups (and partial rows in the workspace) a
 * projection launch over n_rows rows uses: one per CU, at least 4 rows each
 * (FP8; other dtypes take at least 8). */
int jscc_proj_grid(int64_t n_rows);
int64_t jscc_proj_workspace_bytes(int64_t n_rows, int64_t n_cols);
int jscc_proj(int mode, const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows, int64_t n_rows,
              int64_t row_base, const float* row_scale, const float* x, const int64_t* x_perm, int64_t x_perm_stride,
              int64_t x_len, const float* y, int64_t y_stride, const float* coef, float eps, int use_eps,
              float* fwd_out, const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate,
              int flags, float* w, int64_t w_len, void* workspace, int64_t workspace_bytes, int32_t* err,
              void* stream);

/* One OSEM subset step in one call: the mode-0 back-projection of jscc_proj
 * into w, then, in the same reduction pass, em_x[j] = update(em_x[j],
 * em_c1*w[j] (+ em_c2*em_w2[j]), em_s[j]) with the modes of jscc_em_update,
 * and w (and em_w2) cleared afterwards when zero_w != 0.  Replaces, for one
 * subset, get_weight_compton / get_weight_single + "img = img * weight /
 * s_map" of osem_bin_mode / osem_list_mode (recon_osem_plane.py:45-80) and,
 * with em_w2 = the single-photon weight, the "(2 - alpha) * weight_compton +
 * alpha * weight_single" update of osem_joint_mode (:117-118).  em_x may be
 * the projected image x; em_w2 may be NULL. */
int jscc_proj_em(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows, int64_t n_rows,
                 int64_t row_base, const float* row_scale, const float* x, const int64_t* x_perm,
                 int64_t x_perm_stride, int64_t x_len, const float* y, int64_t y_stride, float eps, int use_eps,
                 const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate, int flags,
                 float* w, int64_t w_len, float* em_x, const float* em_s, int em_mode, float em_eps, float em_c1,
                 float* em_w2, float em_c2, int zero_w, void* workspace, int64_t workspace_bytes, int32_t* err,
                 void* stream);

/* jscc_proj (modes 0 and 3) and jscc_proj_em over subset sub_i of sub_k of
 * the first n rows past row_base, where n = *count (device int32) clamped to
 * [0, count_cap]: call rows [row_base + n*sub_i/sub_k, row_base +
 * n*(sub_i+1)/sub_k), the split_tensor_even bounds of recon_osem_plane.py:8-18
 * taken on the device.  The launch is sized by count_cap (rows, and a
 * workspace of jscc_proj_workspace_bytes(ceil(count_cap / sub_k), n_cols)), so
 * a chain of subset steps over a row list whose length stays on the device
 * (the kept-row compaction) needs no host synchronisation.  Numerators y, if
 * given, must be indexed by the stored row (flags bit 0).  An empty subset
 * back-projects zero (jscc_proj_em then applies its update with w = 0, as
 * osem_list_mode does for an empty split). */
int jscc_proj_counted(int mode, const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                      const int32_t* count, int64_t count_cap, int sub_i, int sub_k, int64_t row_base,
                      const float* row_scale, const float* x, const int64_t* x_perm, int64_t x_perm_stride,
                      int64_t x_len, const float* y, int64_t y_stride, float eps, int use_eps, const int64_t* inv_perm,
                      int64_t inv_perm_stride, float out_scale, int accumulate, int flags, float* w, int64_t w_len,
                      void* workspace, int64_t workspace_bytes, int32_t* err, void* stream);
int jscc_proj_em_counted(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                         const int32_t* count, int64_t count_cap, int sub_i, int sub_k, int64_t row_base,
                         const float* row_scale, const float* x, const int64_t* x_perm, int64_t x_perm_stride,
                         int64_t x_len, const float* y, int64_t y_stride, float eps, int use_eps,
                         const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate, int flags,
                         float* w, int64_t w_len, float* em_x, const float* em_s, int em_mode, float em_eps,
                         float em_c1, float* em_w2, float em_c2, int zero_w, void* workspace,
                         int64_t workspace_bytes, int32_t* err, void* stream);

/* jscc_proj_em without the workspace arguments: the partial back-projections
 * go to a static device buffer of JSCC_NOALLOC_SCRATCH_BYTES that the library
 * holds on each GPU it runs on, so the call allocates nothing (no hipMalloc,
 * no caller workspace).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Subsets whose
 * jscc_proj_workspace_bytes(n_rows, n_cols) exceed the buffer are rejected
 * (24 MiB: on a 256-CU GPU, whose full BF16 / FP16 / FP32 projection grid is
 * 192 work-groups plus 12 second-stage rows, widths up to 30840 columns fit at
 * any row count -- FP8 rows, whose grid is 256 + 16, up to 23130 -- and wider
 * rows in subsets small enough for a smaller grid).  All calls on one
 * GPU share its buffer: calls that overlap on one GPU (other streams, other
 * threads) race.  The stream must belong to the current device. */
#define JSCC_NOALLOC_SCRATCH_BYTES (24LL << 20)
int jscc_proj_em_noalloc(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                         int64_t n_rows, int64_t row_base, const float* row_scale, const float* x,
                         const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, const float* y,
                         int64_t y_stride, float eps, int use_eps, const int64_t* inv_perm, int64_t inv_perm_stride,
                         float out_scale, int accumulate, int flags, float* w, int64_t w_len, float* em_x,
                         const float* em_s, int em_mode, float em_eps, float em_c1, float* em_w2, float em_c2,
                         int zero_w, int32_t* err, void* stream);

/* One JSCC-SD subset step of osem_joint_mode (recon_osem_plane.py:100-125) in
 * ONE projection launch: the single-photon rows (mat, rows [row_base,
 * row_base + n_rows) of the stored-row list or of the matrix when rows is
 * NULL, numerators y, indexed by stored row when y_by_row) and the Compton rows
 * (mat2, same dtype and n_cols, numerators y2, ones when NULL) are streamed by
 * the same work-groups into one partial sum, then the reduction applies
 *   w (+)= c1 * A^T(y / A x) + c2 * T^T(y2 / T x),
 *   em_x[j] = update(em_x[j], w[j], em_s[j])      (modes of jscc_em_update)
 * and clears w when zero_w != 0.  Replaces the pair get_weight_single +
 * get_weight_compton and "img * ((2 - alpha) * weight_compton + alpha *
 * weight_single) / s_map" (:117-118) with c1 = alpha, c2 = 2 - alpha.  Either
 * segment may be empty, not both; the workspace is
 * jscc_proj_workspace_bytes(n_rows + n_rows2, n_cols). */
int jscc_proj_joint_em(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows, int64_t n_rows,
                       int64_t row_base, const float* row_scale, const float* y, int64_t y_stride, int y_by_row,
                       float c1, const void* mat2, int64_t ld2, const int32_t* rows2, int64_t n_rows2,
                       int64_t row_base2, const float* row_scale2, const float* y2, int64_t y2_stride, int y2_by_row,
                       float c2, const float* x, float eps, int use_eps, int accumulate, float* w, int64_t w_len,
                       float* em_x, const float* em_s, int em_mode, float em_eps, int zero_w, void* workspace,
                       int64_t workspace_bytes, int32_t* err, void* stream);

/* Every subset of n_iter list-mode OSEM iterations in ONE cooperative launch
 * (one work-group per CU): per subset the rows rows[sub_off[k] ..
 * sub_off[k+1]) are streamed once (mode 0 of jscc_proj with y = 1, no
 * rotation), the work-groups meet at a grid barrier, each sums the partial
 * back-projections of its own pixels and applies the EM update (modes of
 * jscc_em_update, c1 = 1) to x, and after a second barrier all reload x.  The
 * first rows of the next subset are requested before the barriers.
 * Equivalent to n_iter x K calls of jscc_proj_em (osem_list_mode,
 * recon_osem_plane.py:63-80, one rotation and one energy).  sub_off: device
 * int64 [n_sub + 1].  A grid barrier that does not complete within its
 * polling bound sets err bit 32 and ends the launch (results invalid). */
int64_t jscc_osem_persistent_workspace_bytes(int64_t n_cols);
int jscc_osem_persistent(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                         const float* row_scale, const int64_t* sub_off, int n_sub, int n_iter, float* x,
                         const float* sens, float eps, int use_eps, int em_mode, float em_eps, void* workspace,
                         int64_t workspace_bytes, int32_t* err, void* stream);

/* EM update with w = c1*w1 (+ c2*w2):
 *   mode 0: x = x * w / s                        (recon_osem_plane.py:54, 73, 117-118)
 *   mode 1: safe_em_update                        (recon_osem_plane_sparse.py:60-66)
 *   mode 2: apply_em_update                       (distributed/python/recon_osem_dist.py:46-51)
 * zero_after (optional) is cleared for the next subset. */
int jscc_em_update(float* x, const float* w1, float c1, const float* w2, float c2, const float* s, int64_t n, int mode,
                   float eps, float* zero_after, void* stream);

/* Sensitivity s[j] = (1/R) sum_r sum_i A[i, inv[j, r]-1] in FP64
 * (main_plane.py:320-323; main_local_cntstat.py:140-146). inv may be NULL. */
int64_t jscc_sensitivity_workspace_bytes(int64_t M, int64_t P);
int jscc_sensitivity(const float* A, int64_t M, int64_t lda, int64_t P, const int64_t* inv, int64_t inv_row_stride,
                     int64_t inv_col_stride, int R, float* out, void* workspace, int64_t workspace_bytes,
                     int32_t* err, void* stream);

/* No-store cone projection: the rows of T for events ev_idx[0..n_rows) are
 * evaluated in registers (same model as jscc_cone_rows, first-hit factor
 * sysmat[arow, p] or 1) and consumed at once, so T is never written.  Rows
 * are divided by their sums when normalize != 0 (the reference's T).
 *   mode 0: w (+)= out_scale * sum_i T_i / max(T_i x, eps)   (T^T(1/Tx) of
 *           recon_osem_plane.py:33-42 for a T that is not stored)
 *   mode 1: fwd_out[i] = T_i x
 *   mode 2: w (+)= out_scale * sum_i coef_i T_i
 * dsig_out (optional, fixed-sigma model): d(T_i x)/d sigma.  Used for lists
 * whose T exceeds HBM, the lazy per-iteration recompute of
 * distributed/python/recon_osem_dist_sparse_jsccsd_only_lazy.py:46-125, and
 * the adjoint of the differentiable cone operator (no reference counterpart).
 * The kernels run on whole voxel quads: P4 = n_vox rounded up to 4.  The
 * coordinates, x and each sysmat row hold P4 values (16-byte aligned; x past
 * n_vox finite); a count that is not a multiple of 4 needs a sysmat whose
 * rows are zero past n_vox (ld_sysmat 0: one row for every event, e.g. ones
 * then zeros), so the padded voxels carry zero weight; w gets n_vox values. */
int jscc_cone_proj_grid(int64_t n_rows);
int64_t jscc_cone_proj_workspace_bytes(int64_t n_rows, int64_t n_vox);
int jscc_cone_proj(int mode, const void* events, const int32_t* ev_idx, int64_t n_rows, const float* coor_x,
                   const float* coor_y, const float* coor_z, int64_t n_vox, const float* sysmat, int64_t ld_sysmat,
                   int sigma_model, double fixed_sigma, const float* x, const float* coef, float eps, int use_eps,
                   int normalize, float* fwd_out, float* dsig_out, float out_scale, int accumulate, float* w,
                   int64_t w_len, void* workspace, int64_t workspace_bytes, int32_t* err, void* stream);

/* Batched no-store cone projection: B reconstructions (each with its own
 * image and events) in one launch.  ev_idx lists, reconstruction by
 * reconstruction, the event records of the current subset's rows:
 * reconstruction b owns ev_idx[row_off[b] .. row_off[b+1]) (row_off: device
 * int64 [B+1], non-decreasing, row_off[0] = 0, row_off[B] = n_rows).  The
 * work-groups split the n_rows rows evenly whatever the list lengths.
 *   mode 0 (OSEM step): x[b, :] <- update(x[b, :], T_b^T (1 / T_b x_b),
 *           sens[b, :]) with the modes of jscc_em_update, for every b at once
 *           (osem_list_mode's subset update, recon_osem_plane.py:63-80, for B
 *           independent reconstructions); w (optional) receives the weights;
 *   mode 1 (column sums): w[b, :] = sum of the normalised rows of b (the
 *           list-mode sensitivity of main_plane.py:348-355).
 * x, sens, w: FP32 [B, ld*] with 16-byte aligned rows; x rows hold n_vox
 * rounded up to 4 finite values, and a count that is not a multiple of 4
 * needs a sysmat zero past n_vox, as for jscc_cone_proj (updates and sums
 * cover the first n_vox columns).  The launch shape is
 * what an autotuner varies: groups_per_cu, work-groups per CU (0: 4), and
 * tile, one of jscc_cone_batch_tiles(n_vox) tile shapes (threads per work-
 * group, voxel quads per thread, rows per step; jscc_cone_batch_tile_shape
 * describes shape `tile`; 0 is the default).  No counterpart in the reference
 * (it reconstructs one image per process). */
int jscc_cone_batch_tiles(int64_t n_vox);
int jscc_cone_batch_tile_shape(int64_t n_vox, int tile, int* threads, int* quads, int* rows_per_step);
int jscc_cone_batch_groups(int64_t n_rows, int groups_per_cu);
int64_t jscc_cone_batch_workspace_bytes(int64_t n_rows, int n_recon, int64_t n_vox, int groups_per_cu);
int jscc_cone_proj_batched(int mode, const void* events, const int32_t* ev_idx, int64_t n_rows, const int64_t* row_off,
                           int n_recon, const float* coor_x, const float* coor_y, const float* coor_z, int64_t n_vox,
                           const float* sysmat, int64_t ld_sysmat, int sigma_model, double fixed_sigma, float* x,
                           int64_t ldx, float eps, int use_eps, int normalize, const float* sens, int64_t ld_sens,
                           int em_mode, float em_eps, float* w, int64_t ldw, int groups_per_cu, int tile,
                           void* workspace, int64_t workspace_bytes, int32_t* err, void* stream);

/* Count thinning: out[j * out_stride] = number of positions[0..n_pos) that fall
 * in bin j, where bin j owns the copy positions [cum[j-1], cum[j]) (cum: the
 * inclusive prefix sums of round(count)).  With positions = the first k
 * entries of a random permutation of all copies this is the equal-count
 * projection of main_plane.py:300-310 and the count downsampling of
 * sparse_main_utils.py:130-150.  workspace: int32 [n_bins]; err bit 16: a
 * position outside [0, cum[n_bins-1]). */
int jscc_thin_counts(const int64_t* cum_counts, int64_t n_bins, const int64_t* positions, int64_t n_pos, float* out,
                     int64_t out_stride, int32_t* workspace, int32_t* err, void* stream);

/* FP64 sum of n strided FP32 values into *out (device), fixed order: the
 * per-iteration count / likelihood reductions (main_plane.py:312-313). */
int jscc_sum_f64(const float* v, int64_t n, int64_t stride, double* out, void* stream);

/* Self-test of the hardware FP8 decoder: out512[0..255] hardware decode of
 * every e4m3fn code, out512[256..511] the library's software decode. */
int jscc_fp8_selftest(float* out512, void* stream);

/* ---- K frames sharing one BF16 system matrix (MFMA) ---------------------- */

/* No counterpart in the reference, which reconstructs one data set per call:
 * K single-photon data sets y_k with one system matrix (the time frames of a
 * dynamic acquisition, bootstrap replicates) each run the per-subset weight
 * get_weight_single (recon_osem_plane.py:28-30) and the EM update of
 * osem_bin_mode (:45-54).  Batched, the subset products are GEMMs on the
 * matrix cores (v_mfma_f32_16x16x32_bf16):
 *   forward  F[i][f] = sum_j R[i][j] X[j][f]   a = R, row-major [rows x lda]
 *   back     W[j][f] = sum_i R[i][j] C[i][f]   a = R^T, [P x lda] (the matrix
 *                                              kept a second time, transposed)
 * The frame operand is FP32, passed as two BF16 terms hi = bf16(v),
 * lo = bf16(v - hi), frame-major [k_frames x ldb] with the reduced index
 * contiguous; frames-minor FP32 arrays ([n x ld], frame f of item i at
 * i * ld + f) hold images X [P x K], data Y [M x K] and forward projections.
 * k_frames is 16, 32 or 64 (pad with frames whose data are zero). */

/* dst[c][r] = src[r][c] for a BF16 matrix [rows x cols] (row stride ld_src);
 * dst rows are ld_dst >= rows long and their columns [rows, ld_dst) zeroed. */
int jscc_frames_transpose_bf16(const uint16_t* src, int64_t rows, int64_t cols, int64_t ld_src, uint16_t* dst,
                               int64_t ld_dst, void* stream);

/* hi/lo[f][i] = terms of x[i * ldx + f] for i < n, zero for n <= i < n_pad. */
int jscc_frames_split(const float* x, int64_t n, int64_t n_pad, int k_frames, int64_t ldx, uint16_t* hi,
                      uint16_t* lo, int64_t ldt, void* stream);

/* Chunks of the reduced index for a GEMM of m output rows on `cus` compute
 * units (about one 64-row work-group per unit), and the partial-sum bytes of
 * a GEMM. */
int jscc_frames_split_count(int64_t m, int64_t kdim, int cus);
int64_t jscc_frames_partial_bytes(int64_t m, int k_frames, int n_split);

/* partial[sp][i][f] = sum over chunk sp of the reduced index k < kdim of
 * a[i * lda + k] * (b_hi[f * ldb + k] + b_lo[f * ldb + k]), i < m; kdim a
 * multiple of 128 (zero-padded operands), 16-byte aligned operands with
 * strides a multiple of 8.  n_split must be the value jscc_frames_split_count
 * returns for (m, kdim). */
int jscc_frames_gemm(const uint16_t* a, int64_t lda, int64_t m, int64_t kdim, const uint16_t* b_hi,
                     const uint16_t* b_lo, int64_t ldb, int k_frames, int n_split, float* partial,
                     int64_t partial_bytes, void* stream);

/* The ratio step of a subset of n rows: F[i][f] = sum_sp partial[sp][i][f],
 * c = y[i * ldy + f] / F (/ max(F, eps) when use_eps), written as the hi/lo
 * terms of C^T at columns c_off + i of [k_frames x ldc]; the other columns of
 * [0, n_pad) are zeroed.  f_out [n x k_frames] (F) may be NULL. */
int jscc_frames_ratio(const float* partial, int n_split, int64_t n, int64_t n_pad, int64_t c_off, int k_frames,
                      const float* y, int64_t ldy, float eps, int use_eps, uint16_t* c_hi, uint16_t* c_lo,
                      int64_t ldc, float* f_out, void* stream);

/* The EM update of every frame: W[j][f] = sum_sp partial[sp][j][f] (zero when
 * partial is NULL, an empty subset), x[j * ldx + f] = update(x, W, s[j]) with
 * the modes of jscc_em_update, and the hi/lo terms of the updated image at
 * t[f][j] when t_hi is given. */
int jscc_frames_update(const float* partial, int n_split, int64_t n_pix, int k_frames, const float* s, float* x,
                       int64_t ldx, int em_mode, float em_eps, uint16_t* t_hi, uint16_t* t_lo, int64_t ldt,
                       void* stream);

/* Built-in kernel timer (no counterpart in the reference).  While enabled
 * (per calling thread), the main kernel of every launch is bracketed by a HIP
 * event pair on its own stream, by tag: 0 cone rows (jscc_cone_rows),
 * 1 projection modes 0-2 (jscc_proj / jscc_proj_em), 2 projection mode 3
 * (column sums), 3 no-store cone projection (jscc_cone_proj), 4 fused coarse
 * row weight (jscc_coarse_weight), 5 multi-frame GEMM (jscc_frames_gemm).
 * jscc_prof_read synchronises the pairs of one tag and returns their count and
 * summed milliseconds. */
int jscc_prof_enable(int on);
int jscc_prof_read(int tag, int* count, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* JSCC_HIP_H */
