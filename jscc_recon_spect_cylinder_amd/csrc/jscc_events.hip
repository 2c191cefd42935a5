// Event preparation (energy smearing, thresholds, kinematic and same-layer
// filters, per-event cone parameters) and order-preserving stream compaction.
//
// Follows process_list_plane_strict.py:24-52, 55-96, 151-199 (and the same
// steps in process_list_plane_sparse.py:40-72).  The filter predicates are
// evaluated with the reference's FP32 operation order and explicit
// round-to-nearest intrinsics (no contraction), so the surviving event set is
// bit-identical to the reference's for the same list and noise.  The
// continuous per-event quantities (theta, KN, the energy sigmas) are then
// computed in FP64 from the smeared FP32 energies.
#include "jscc_common.h"
#include "jscc_events.h"

#include <math.h>

namespace {

constexpr double kElectronRestMeV = 0.511;  // process_list_plane_strict.py:7

// reference clamp of cos(theta) / cos(beta): torch clamps an FP32 tensor to
// (-1 + 1e-7, 1 - 1e-7), i.e. to the FP32 roundings of those bounds.
__device__ __forceinline__ double cos_clamp_f32bound() { return (double)(float)(1.0 - 1e-7); }

// Correctly rounded FP32 square root: the FP64 root rounded once to FP32 (53 >=
// 2*24 + 2 bits, so the double rounding cannot differ from a direct one).  The
// device's FP32 sqrt (sqrtf / __fsqrt_rn) is not correctly rounded on gfx950.
__device__ __forceinline__ float sqrt_rn_f32(float x) { return (float)__dsqrt_rn((double)x); }

__device__ __forceinline__ double theta_from_e1(double e1, double e0) {
  // process_list_plane_strict.py:24-27
  double c = 1.0 - (kElectronRestMeV * e1) / ((e0 - e1) * e0);
  const double cm = cos_clamp_f32bound();
  c = c < -cm ? -cm : (c > cm ? cm : c);
  return acos(c);
}

// process_list_plane_strict.py:55-96 ; one crystal's position variance.
__device__ __forceinline__ void crystal_sigma_sq(float abs_y, int layers4, float rear_abs_y,
                                                 float extra, float* sx, float* sy, float* sz) {
  if (!layers4) {
    float v = extra > 0.0f ? extra : 0.0f;
    v = v * v;
    *sx = v; *sy = v; *sz = v;
    return;
  }
  if (abs_y == rear_abs_y) {  // rear layer 2 x 6 x 2 mm
    *sx = (float)(4.0 / 12.0); *sy = (float)(36.0 / 12.0); *sz = (float)(4.0 / 12.0);
  } else {                    // front layers 3 x 3 x 3 mm
    *sx = (float)(9.0 / 12.0); *sy = (float)(9.0 / 12.0); *sz = (float)(9.0 / 12.0);
  }
  if (extra > 0.0f) {
    float e2 = (float)((double)extra * (double)extra);
    *sx = __fadd_rn(*sx, e2); *sy = __fadd_rn(*sy, e2); *sz = __fadd_rn(*sz, e2);
  }
}

__global__ void __launch_bounds__(256) k_events_prepare(
    const float* __restrict__ list, int64_t n, int64_t ls,
    const float* __restrict__ noise1, const float* __restrict__ noise2,
    float* __restrict__ writeback,
    const float* __restrict__ det, int64_t ndet, int64_t ds, int64_t n_arows,
    int layers4, float rear_abs_y, float dr1, float dr2,
    float e0f, float resf, float thr_max, float thr_min, float thr_sum,
    double e0d, double resd,
    JsccEvent* __restrict__ out, int32_t* __restrict__ flags, int32_t* __restrict__ err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* row = list + i * ls;
  const float c1f = row[0];
  float e1 = row[1];
  const float c2f = row[2];
  float e2 = row[3];

  // process_list_plane_strict.py:157-161 : sigma_k = e_k*R/2.355*(E0/e_k)**0.5 ;
  // torch evaluates E0/e_k as reciprocal(e_k)*E0 and **0.5 as sqrt.
  if (noise1 != nullptr) {
    float s1 = __fmul_rn(__fdiv_rn(__fmul_rn(e1, resf), 2.355f),
                         sqrt_rn_f32(__fmul_rn(__fdiv_rn(1.0f, e1), e0f)));
    float s2 = __fmul_rn(__fdiv_rn(__fmul_rn(e2, resf), 2.355f),
                         sqrt_rn_f32(__fmul_rn(__fdiv_rn(1.0f, e2), e0f)));
    e1 = __fadd_rn(e1, __fmul_rn(s1, noise1[i]));
    e2 = __fadd_rn(e2, __fmul_rn(s2, noise2[i]));
    if (writeback != nullptr) {  // the reference mutates list columns 1 and 3 in place
      writeback[i * ls + 1] = e1;
      writeback[i * ls + 3] = e2;
    }
  }

  // :163-171 energy window
  bool keep = (e1 < thr_max) && (e1 > thr_min) && (e2 > thr_min) && (__fadd_rn(e1, e2) > thr_sum);
  // :30-36 kinematic domain
  if (keep) {
    const float ee = 0.511f;
    float cr = __fsub_rn(1.0f, __fdiv_rn(__fmul_rn(ee, e1), __fmul_rn(__fsub_rn(e0f, e1), e0f)));
    keep = (cr > (float)(-1.0 + 1e-6)) && (cr < (float)(1.0 - 1e-6));
  }
  JsccEvent ev;
  ev.src = (int32_t)i;
  if (keep) {
    // :185-189 gather crystal positions; torch index semantics (negative wraps)
    bool bad = !(c1f == c1f) || !(c2f == c2f) || fabsf(c1f) > 2.0e9f || fabsf(c2f) > 2.0e9f;
    int64_t i1 = bad ? 0 : (int64_t)(int32_t)c1f - 1;
    int64_t i2 = bad ? 0 : (int64_t)(int32_t)c2f - 1;
    if (bad || i1 < -ndet || i1 >= ndet || i2 < -ndet || i2 >= ndet) {
      atomicOr(err, 1);
      keep = false;
    } else {
      if (i1 < 0) i1 += ndet;
      if (i2 < 0) i2 += ndet;
      const float* q1 = det + i1 * ds;
      const float* q2 = det + i2 * ds;
      const float y1 = q1[1], y2 = q2[1];
      keep = fabsf(__fsub_rn(y1, y2)) > 0.1f;  // :189 same-layer filter
      if (keep) {
        ev.p1x = q1[0]; ev.p1y = y1; ev.p1z = q1[2];
        ev.ax = __fsub_rn(q2[0], q1[0]);
        ev.ay = __fsub_rn(y2, y1);
        ev.az = __fsub_rn(q2[2], q1[2]);
        double dsq = (double)ev.ax * ev.ax + (double)ev.ay * ev.ay + (double)ev.az * ev.az;
        ev.d12sq = (float)dsq;
        // :209-210 theta and the Klein-Nishina factor, FP64
        const double e1d = (double)e1, e0 = e0d;
        const double th = theta_from_e1(e1d, e0);
        ev.theta = (float)th;
        ev.kn = (float)(e0 / (e0 - e1d) + (e0 - e1d) / e0);
        // :39-49 asymmetric energy sigma
        const double se = e1d * resd / 2.355 * sqrt(e0 / e1d);
        const double lo_b = 1e-7, hi_b = e0 - 1e-7;
        double el = e1d - se, eh = e1d + se;
        el = el < lo_b ? lo_b : (el > hi_b ? hi_b : el);
        eh = eh < lo_b ? lo_b : (eh > hi_b ? hi_b : eh);
        const double sm = th - theta_from_e1(el, e0);
        const double sp = theta_from_e1(eh, e0) - th;
        ev.sig_minus = (float)(sm > 1e-7 ? sm : 1e-7);
        ev.sig_plus = (float)(sp > 1e-7 ? sp : 1e-7);
        // :59-96 position variances of both hits
        float ax1, ay1, az1, ax2, ay2, az2;
        crystal_sigma_sq(fabsf(y1), layers4, rear_abs_y, dr1, &ax1, &ay1, &az1);
        crystal_sigma_sq(fabsf(y2), layers4, rear_abs_y, dr2, &ax2, &ay2, &az2);
        ev.wx = __fadd_rn(ax1, ax2);
        ev.wy = __fadd_rn(ay1, ay2);
        ev.wz = __fadd_rn(az1, az2);
        // :237-240 sysmat[cpnum1 - 1] (torch wrap relative to the matrix rows)
        int64_t ia = (int64_t)(int32_t)c1f - 1;
        if (n_arows > 0) {
          if (ia < -n_arows || ia >= n_arows) {
            atomicOr(err, 1);
            keep = false;
          } else if (ia < 0) {
            ia += n_arows;
          }
        } else {
          ia = i1;
        }
        ev.arow = (int32_t)ia;
        // a non-finite hit position makes every weight of the row NaN in the reference
        // (:201-228), which then drops the row (:241-246): drop the event here
        const float fin = ev.p1x * 0.0f + ev.p1y * 0.0f + ev.p1z * 0.0f + ev.ax * 0.0f + ev.ay * 0.0f +
                          ev.az * 0.0f + ev.d12sq * 0.0f;
        if (!(fin == 0.0f)) keep = false;
      }
    }
  }
  if (keep) out[i] = ev;
  flags[i] = keep ? 1 : 0;
}

// ------------------------------------------------------------------ compaction
constexpr int kCompactBlock = 1024;

__global__ void __launch_bounds__(kCompactBlock) k_flag_counts(const int32_t* __restrict__ flags, int64_t n,
                                                               int32_t* __restrict__ counts) {
  __shared__ int32_t wsum[kCompactBlock / JSCC_WAVE];
  int64_t i = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
  int f = (i < n) ? (flags[i] != 0) : 0;
  unsigned long long m = __ballot(f);
  if (jscc_lane() == 0) wsum[jscc_wave_id()] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kCompactBlock / JSCC_WAVE; ++w) s += wsum[w];
    counts[blockIdx.x] = s;
  }
}

// single work-group exclusive scan over the per-block counts; total -> *total
__global__ void __launch_bounds__(1024) k_scan_counts(int32_t* __restrict__ counts, int nb,
                                                      int32_t* __restrict__ total) {
  __shared__ int32_t carry;
  __shared__ int32_t wsum[16];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int idx = base + threadIdx.x;
    int v = idx < nb ? counts[idx] : 0;
    // inclusive wave scan
    int x = v;
#pragma unroll
    for (int off = 1; off < JSCC_WAVE; off <<= 1) {
      int y = __shfl_up(x, off, JSCC_WAVE);
      if (jscc_lane() >= off) x += y;
    }
    if (jscc_lane() == JSCC_WAVE - 1) wsum[jscc_wave_id()] = x;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < jscc_wave_id(); ++w) woff += wsum[w];
    int excl = carry + woff + x - v;
    __syncthreads();
    if (idx < nb) counts[idx] = excl;
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kCompactBlock) k_flag_scatter(const int32_t* __restrict__ flags, int64_t n,
                                                                const int32_t* __restrict__ offsets,
                                                                int32_t* __restrict__ out_idx) {
  __shared__ int32_t wsum[kCompactBlock / JSCC_WAVE];
  int64_t i = (int64_t)blockIdx.x * kCompactBlock + threadIdx.x;
  int f = (i < n) ? (flags[i] != 0) : 0;
  unsigned long long m = __ballot(f);
  int lane = jscc_lane();
  int before = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[jscc_wave_id()] = __popcll(m);
  __syncthreads();
  int woff = 0;
  for (int w = 0; w < jscc_wave_id(); ++w) woff += wsum[w];
  if (f) out_idx[offsets[blockIdx.x] + woff + before] = (int32_t)i;
}

// stats [n,4] = (S, Q, Sc, max) -> keep flag (process_list_plane_strict.py:241-248, 137-148)
__global__ void __launch_bounds__(256) k_rows_keep(const double* __restrict__ stats, int64_t n,
                                                   double min_support, int32_t* __restrict__ flags,
                                                   float* __restrict__ inv_sum) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double S = stats[i * 4 + 0];
  const double Q = stats[i * 4 + 1];
  const double M = stats[i * 4 + 3];
  const bool keep = jscc_row_keep(S, Q, M, min_support);
  flags[i] = keep ? 1 : 0;
  if (inv_sum) inv_sum[i] = keep ? (float)(1.0 / S) : 0.0f;
}

}  // namespace

extern "C" int jscc_events_prepare(const float* list, int64_t n, int64_t list_stride,
                                   const float* noise1, const float* noise2, float* list_writeback,
                                   const float* detector, int64_t n_det, int64_t det_stride, int64_t n_sysmat_rows,
                                   int det_layers4, float rear_abs_y, float delta_r1, float delta_r2,
                                   float e0_f32, float ene_res_f32, float thr_max_f32, float thr_min_f32,
                                   float thr_sum_f32, double e0, double ene_res,
                                   void* events_out, int32_t* flags_out, int32_t* err_out, void* stream) {
  JSCC_REQUIRE(n >= 0, "jscc_events_prepare: negative event count");
  if (n == 0) return 0;
  JSCC_REQUIRE(list && detector && events_out && flags_out && err_out, "jscc_events_prepare: null pointer");
  JSCC_REQUIRE(list_stride >= 4 && det_stride >= 3 && n_det > 0 && n_sysmat_rows >= 0,
               "jscc_events_prepare: bad strides");
  JSCC_REQUIRE((noise1 == nullptr) == (noise2 == nullptr), "jscc_events_prepare: noise1/noise2 must both be set or unset");
  hipStream_t s = (hipStream_t)stream;
  int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_events_prepare, dim3((unsigned)nb), dim3(256), 0, s, list, n, list_stride, noise1, noise2,
                     list_writeback, detector, n_det, det_stride, n_sysmat_rows, det_layers4, rear_abs_y, delta_r1,
                     delta_r2,
                     e0_f32, ene_res_f32, thr_max_f32, thr_min_f32, thr_sum_f32, e0, ene_res,
                     (JsccEvent*)events_out, flags_out, err_out);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int64_t jscc_compact_workspace_bytes(int64_t n) {
  int64_t nb = (n + kCompactBlock - 1) / kCompactBlock;
  return (nb + 1) * (int64_t)sizeof(int32_t);
}

extern "C" int jscc_compact_flags(const int32_t* flags, int64_t n, int32_t* out_idx, int32_t* out_count,
                                  void* workspace, int64_t workspace_bytes, void* stream) {
  JSCC_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "jscc_compact_flags: n out of range");
  JSCC_REQUIRE(out_count != nullptr, "jscc_compact_flags: null count");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    JSCC_CHECK_HIP(hipMemsetAsync(out_count, 0, sizeof(int32_t), s));
    return 0;
  }
  JSCC_REQUIRE(flags && out_idx && workspace, "jscc_compact_flags: null pointer");
  JSCC_REQUIRE(workspace_bytes >= jscc_compact_workspace_bytes(n), "jscc_compact_flags: workspace too small");
  int64_t nb = (n + kCompactBlock - 1) / kCompactBlock;
  int32_t* counts = (int32_t*)workspace;
  hipLaunchKernelGGL(k_flag_counts, dim3((unsigned)nb), dim3(kCompactBlock), 0, s, flags, n, counts);
  JSCC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, (int)nb, out_count);
  JSCC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)nb), dim3(kCompactBlock), 0, s, flags, n, counts, out_idx);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_rows_keep_flags(const double* stats, int64_t n, double min_support, int32_t* flags,
                                    float* inv_sum, void* stream) {
  JSCC_REQUIRE(n >= 0, "jscc_rows_keep_flags: negative n");
  if (n == 0) return 0;
  JSCC_REQUIRE(stats && flags, "jscc_rows_keep_flags: null pointer");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rows_keep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, stats, n, min_support, flags,
                     inv_sum);
  JSCC_CHECK_LAUNCH();
  return 0;
}
