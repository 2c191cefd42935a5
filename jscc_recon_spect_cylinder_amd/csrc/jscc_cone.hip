// Compton cone back-projection: one row of the list-mode matrix T per event.
//
// Reference: process_list_plane_strict.py:99-134 (sigma_pos), :39-52
// (sigma_ene), :201-248 (cone weight, first-hit weighting, row statistics);
// process_list_plane_sparse.py:78-114 (same on coarse voxels with the
// 2-tap-reduced system-matrix row).
//
// Layout / mapping (gfx950):
//   * a work-group owns E consecutive events (rows) and sweeps all P voxels in
//     tiles of 4*256; every thread evaluates 4 consecutive voxels for each of
//     the E events.  Event parameters are wave-uniform, so they live in scalar
//     registers (s_load), not LDS.
//   * voxel coordinates are SoA float arrays padded to a multiple of 4, read as
//     float4 (one 16-byte load per component per thread, contiguous across the
//     wave); the first-hit system-matrix row is read the same way.
//   * beta from |v01 x v12| and v01 . v12 (the half-angle form of atan2, a
//     hand-fitted polynomial): well conditioned at the cone axis, unlike acos
//     of an FP32 cosine; |v01 x v12|^2 by Lagrange's identity away from the
//     axis, by the cross product near it.
//   * rows are written unnormalised (FP32 or BF16) and the row statistics
//     S = sum t, Q = sum t^2, Sc = sum t_c and max t are reduced in FP64; the
//     normalisation 1/S is applied by whoever consumes the row.
#include <type_traits>

#include "jscc_common.h"
#include "jscc_events.h"

namespace {

enum { SIG_STRICT = 0, SIG_FIXED = 1 };
enum { AF_NONE = 0, AF_ROW = 1, AF_TAPS = 2 };

constexpr int kConeThreads = 256;
constexpr int kConeVec = 4;

// 2 atan(r) on [0, 1] as r*P(r^2); coefficients fitted to minimise the
// absolute error (2.3e-7 rad in FP32 Horner form, the FP32 rounding floor of
// values up to pi/2).
__device__ __forceinline__ float two_atan_01(float r) {
  const float s = r * r;
  float p = -8.105187127e-03f;
  p = fmaf(p, s, 4.371275015e-02f);
  p = fmaf(p, s, -1.118076552e-01f);
  p = fmaf(p, s, 1.928333648e-01f);
  p = fmaf(p, s, -2.781694243e-01f);
  p = fmaf(p, s, 3.989309530e-01f);
  p = fmaf(p, s, -6.665972215e-01f);
  p = fmaf(p, s, 1.999998673e+00f);
  return r * p;
}

struct ConeConsts {
  float beta_min;            // acos of the reference's FP32 clamp bound (beta_max = pi - beta_min)
  float s2_min;              // 1 - cos_max^2 : sin^2 floor implied by the clamp
  float fixed_hl2;           // log2(e) / (2 sigma^2) for the fixed-sigma model
  float inv_sig3;            // 1 / sigma^3 (d log t / d sigma = delta^2 / sigma^3)
};

ConeConsts make_consts(int sigma_model, double fixed_sigma) {
  ConeConsts cc;
  const double cmax = (double)(float)(1.0 - 1e-7);  // the reference's FP32 clamp of cos(beta)
  cc.beta_min = (float)acos(cmax);
  cc.s2_min = (float)(1.0 - cmax * cmax);
  const bool fixed = sigma_model == SIG_FIXED && fixed_sigma > 0.0;
  cc.fixed_hl2 = fixed ? (float)(0.5 * 1.4426950408889634 / (fixed_sigma * fixed_sigma)) : 0.0f;
  cc.inv_sig3 = fixed ? (float)(1.0 / (fixed_sigma * fixed_sigma * fixed_sigma)) : 0.0f;
  return cc;
}

// wave-uniform parameters of one event (scalar registers)
struct ConeEv {
  float p1x, p1y, p1z, ax, ay, az, d12sq, th, pi_m_th, kn, sp, sm, wx, wy, wz;
  int64_t arow_off;
};

__device__ __forceinline__ ConeEv load_event(const JsccEvent& q, int64_t lda) {
  ConeEv e;
  e.p1x = q.p1x; e.p1y = q.p1y; e.p1z = q.p1z;
  e.ax = q.ax; e.ay = q.ay; e.az = q.az;
  e.d12sq = q.d12sq;
  e.th = q.theta; e.kn = q.kn;
  e.pi_m_th = 3.14159265358979324f - q.theta;
  e.sp = q.sig_plus; e.sm = q.sig_minus;
  e.wx = q.wx; e.wy = q.wy; e.wz = q.wz;
  e.arow_off = (int64_t)q.arow * lda;
  return e;
}

// sin^2(beta) below which |v01 x v12|^2 is taken from the cross product: above
// it, d01^2 d12^2 - (v01 . v12)^2 loses at most ~3 eps cot(beta) rad of beta
// (7e-7 rad at the bound), below it the difference cancels toward the cone axis.
constexpr float kLagrangeMinS2 = 0.0625f;

// t_c of one event at 4 voxels.  For the fixed-sigma model and DSIG,
// dlog[k] = d log(t_c) / d sigma = delta^2 / sigma^3.
//
// |v01 x v12|^2 = d01^2 d12^2 - (v01 . v12)^2 (Lagrange's identity: 2 VALU
// instead of 9 for the cross product and its norm).  A lane whose 4 voxels
// include one with sin^2(beta) < kLagrangeMinS2 (or a negative difference)
// recomputes all 4 from the cross product; a wave skips that branch unless one
// of its lanes takes it (about 1 wave in 7 on a 64 x 256 polar grid).  The
// choice depends only on the lane's own 4 voxels, so every kernel that tiles
// voxels in aligned quads gets the same bits for the same pair.
template <int SIG, bool DSIG>
__device__ __forceinline__ void cone_tc4(const ConeEv& e, const float* cxv, const float* cyv, const float* czv,
                                         const ConeConsts& cc, float* tc, float* dlog) {
  float vx[4], vy[4], vz[4], d01sq[4], dot[4], msq[4], rp[4], s2r[4];
  bool near_axis = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    vx[k] = e.p1x - cxv[k];
    vy[k] = e.p1y - cyv[k];
    vz[k] = e.p1z - czv[k];
    d01sq[k] = fmaf(vx[k], vx[k], fmaf(vy[k], vy[k], vz[k] * vz[k]));
    dot[k] = fmaf(vx[k], e.ax, fmaf(vy[k], e.ay, vz[k] * e.az));
    const float pr = d01sq[k] * e.d12sq;
    msq[k] = fmaf(-dot[k], dot[k], pr);
    rp[k] = __builtin_amdgcn_rcpf(pr);
    s2r[k] = msq[k] * rp[k];
    near_axis |= s2r[k] < kLagrangeMinS2;
  }
  if (near_axis) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float mx = fmaf(vy[k], e.az, -vz[k] * e.ay);
      const float my = fmaf(vz[k], e.ax, -vx[k] * e.az);
      const float mz = fmaf(vx[k], e.ay, -vy[k] * e.ax);
      msq[k] = fmaf(mx, mx, fmaf(my, my, mz * mz));
      s2r[k] = msq[k] * rp[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // Half-angle form of beta = atan2(|v01 x v12|, v01 . v12) with h = d01 d12:
    // b = 2 atan(|v01 x v12| / (h + |v01 . v12|)) is beta for a non-negative dot
    // product and pi - beta otherwise, and the ratio lies in [0, 1] without a
    // min/max swap or quadrant fix-up.  Hardware sqrt / rsq / rcp / exp2 (about
    // 1 ulp): no range-reduction branches in the loop.
    const float y = __builtin_amdgcn_sqrtf(msq[k]);
    const float h = __builtin_amdgcn_rsqf(rp[k]);
    float b = two_atan_01(y * __builtin_amdgcn_rcpf(h + fabsf(dot[k])));
    // the clamp of beta to [beta_min, pi - beta_min] is the clamp of b to
    // [beta_min, pi/2] (b <= pi/2 already); one median-of-three (b is finite:
    // events with non-finite hits are dropped before this kernel and voxel
    // coordinates are checked finite)
    b = __builtin_amdgcn_fmed3f(b, cc.beta_min, 2.0f);
    const bool fwd = dot[k] >= 0.0f;
    // sin^2(beta) of the clamped beta: never below the clamp's floor
    const float s2 = __builtin_amdgcn_fmed3f(s2r[k], cc.s2_min, 2.0f);
    // delta = beta - theta; the fixed-sigma model needs only delta^2 = (b - theta')^2
    // with theta' = theta or pi - theta
    const float delta = SIG == SIG_STRICT ? (fwd ? b - e.th : e.pi_m_th - b) : b - (fwd ? e.th : e.pi_m_th);
    float hl2;  // log2(e) / (2 sigma^2)
    if (SIG == SIG_STRICT) {
      const float sige = delta >= 0.0f ? e.sp : e.sm;
      const float qx = fmaf(e.d12sq, vx[k], -dot[k] * e.ax);
      const float qy = fmaf(e.d12sq, vy[k], -dot[k] * e.ay);
      const float qz = fmaf(e.d12sq, vz[k], -dot[k] * e.az);
      const float num = fmaf(qx * qx, e.wx, fmaf(qy * qy, e.wy, qz * qz * e.wz));
      const float den = e.d12sq * e.d12sq * fmaxf(msq[k], 1e-12f * d01sq[k] * e.d12sq);
      const float spos2 = fmaxf(num * __builtin_amdgcn_rcpf(den), 1e-12f);
      const float sig2 = fmaxf(fmaf(sige, sige, spos2), 1e-12f);
      hl2 = 0.72134752044448170368f * __builtin_amdgcn_rcpf(sig2);
    } else {
      hl2 = cc.fixed_hl2;
    }
    const float d2 = delta * delta;
    if (DSIG) dlog[k] = d2 * cc.inv_sig3;
    tc[k] = __builtin_amdgcn_exp2f(-d2 * hl2) * (e.kn - s2);
  }
}

template <int AF>
__device__ __forceinline__ void afactor4(const ConeEv& e, const float* __restrict__ A, int base, const int* ti0,
                                         const int* ti1, const float* tw0, const float* tw1, float* av) {
  if (AF == AF_ROW) {
    const float4 a4 = *reinterpret_cast<const float4*>(A + e.arow_off + base);
    av[0] = a4.x; av[1] = a4.y; av[2] = a4.z; av[3] = a4.w;
  } else if (AF == AF_TAPS) {
    const float* Ar = A + e.arow_off;
#pragma unroll
    for (int k = 0; k < 4; ++k) av[k] = fmaf(tw1[k], Ar[ti1[k]], tw0[k] * Ar[ti0[k]]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) av[k] = 1.0f;
  }
}

template <int AF>
__device__ __forceinline__ void load_taps(int base, int P, const int32_t* __restrict__ tap_i0,
                                          const int32_t* __restrict__ tap_i1, const float* __restrict__ tap_w0,
                                          const float* __restrict__ tap_w1, int* ti0, int* ti1, float* tw0,
                                          float* tw1) {
  if (AF == AF_TAPS) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int p = base + k < P ? base + k : P - 1;
      ti0[k] = tap_i0[p]; ti1[k] = tap_i1[p]; tw0[k] = tap_w0[p]; tw1[k] = tap_w1[p];
    }
  }
}

// One tile of 4 voxels for the E events of a work-group: the row values, stored in rows rrow[e], and the row
// statistics (S, Q, Sc, max) accumulated per thread in FP32.  A slot past the row count carries the event of
// the last row and stores into that row (rrow = the last row): the same values to the same addresses by the
// same thread, so the stores need no condition.
template <int SIG, int AF, int OUT_DT, bool WANT_TC, int E>
__device__ __forceinline__ void cone_rows_tile_at(const ConeEv* evs, const int* rrow, int base, int P,
                                                  float4 vx4, float4 vy4, float4 vz4, const float* __restrict__ A,
                                                  const int32_t* __restrict__ tap_i0,
                                                  const int32_t* __restrict__ tap_i1,
                                                  const float* __restrict__ tap_w0, const float* __restrict__ tap_w1,
                                                  const ConeConsts& cc, void* __restrict__ out_t, int64_t ldt,
                                                  float* __restrict__ out_tc, int64_t ldtc, float* accS, float* accQ,
                                                  float* accC, float* accM) {
  const float vcx[4] = {vx4.x, vx4.y, vx4.z, vx4.w};
  const float vcy[4] = {vy4.x, vy4.y, vy4.z, vy4.w};
  const float vcz[4] = {vz4.x, vz4.y, vz4.z, vz4.w};
  float tw0[4], tw1[4];
  int ti0[4], ti1[4];
  load_taps<AF>(base, P, tap_i0, tap_i1, tap_w0, tap_w1, ti0, ti1, tw0, tw1);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    float av[4];
    afactor4<AF>(evs[e], A, base, ti0, ti1, tw0, tw1, av);
    float tv[4], tcv[4];
    cone_tc4<SIG, false>(evs[e], vcx, vcy, vcz, cc, tcv, nullptr);
#pragma unroll
    for (int k = 0; k < 4; ++k) tv[k] = tcv[k] * av[k];
    if (base + 4 > P) {  // the padded tail of the last tile (P not a multiple of 4)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (base + k >= P) { tcv[k] = 0.0f; tv[k] = 0.0f; }
    }
    {
      const int64_t r = (int64_t)rrow[e];
      // row start (uniform) plus a 32-bit byte offset: scalar base + vector offset addressing
      if (OUT_DT == JSCC_DT_F32) {
        char* rowp = reinterpret_cast<char*>(out_t) + r * ldt * (int64_t)sizeof(float);
        *reinterpret_cast<float4*>(rowp + (uint32_t)base * 4u) = make_float4(tv[0], tv[1], tv[2], tv[3]);
      } else if (OUT_DT == JSCC_DT_BF16) {
        uint2 pk;
        pk.x = jscc_pack_bf16x2(tv[0], tv[1]);
        pk.y = jscc_pack_bf16x2(tv[2], tv[3]);
        char* rowp = reinterpret_cast<char*>(out_t) + r * ldt * (int64_t)sizeof(uint16_t);
        *reinterpret_cast<uint2*>(rowp + (uint32_t)base * 2u) = pk;
      }
      if (WANT_TC) {
        *reinterpret_cast<float4*>(out_tc + r * ldtc + base) = make_float4(tcv[0], tcv[1], tcv[2], tcv[3]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      accS[e] += tv[k];
      accQ[e] = fmaf(tv[k], tv[k], accQ[e]);
      if (AF != AF_NONE) accC[e] += tcv[k];  // without an A factor t = t_c: Sc = S (set below)
      accM[e] = fmaxf(accM[e], tv[k]);  // fmaxf drops NaN; NaN rows are caught by S
    }
  }
}

template <int SIG, int AF, int OUT_DT, bool WANT_TC, int E>
__device__ __forceinline__ void cone_rows_tile(const ConeEv* evs, const int* rrow, int base, int P,
                                               const float* __restrict__ cx, const float* __restrict__ cy,
                                               const float* __restrict__ cz, const float* __restrict__ A,
                                               const int32_t* __restrict__ tap_i0, const int32_t* __restrict__ tap_i1,
                                               const float* __restrict__ tap_w0, const float* __restrict__ tap_w1,
                                               const ConeConsts& cc, void* __restrict__ out_t, int64_t ldt,
                                               float* __restrict__ out_tc, int64_t ldtc, float* accS, float* accQ,
                                               float* accC, float* accM) {
  cone_rows_tile_at<SIG, AF, OUT_DT, WANT_TC, E>(evs, rrow, base, P, *reinterpret_cast<const float4*>(cx + base),
                                                 *reinterpret_cast<const float4*>(cy + base),
                                                 *reinterpret_cast<const float4*>(cz + base), A, tap_i0, tap_i1,
                                                 tap_w0, tap_w1, cc, out_t, ldt, out_tc, ldtc, accS, accQ, accC, accM);
}

// Rows whose largest value is below kQ64Below: FP32 squares of their tail would fall under the normal range,
// so their sum of squares is recomputed in FP64 (per thread over its tiles, DPP within the wave, waves in
// order) from the same row values.  Every thread of the work-group calls it (it holds two barriers); wq is
// LDS for one value per wave.
constexpr double kQ64Below = 0x1p-40;

template <int SIG, int AF, int THREADS>
__device__ double cone_row_q64(const ConeEv& ev, int P, const float* __restrict__ cx, const float* __restrict__ cy,
                               const float* __restrict__ cz, const float* __restrict__ A,
                               const int32_t* __restrict__ tap_i0, const int32_t* __restrict__ tap_i1,
                               const float* __restrict__ tap_w0, const float* __restrict__ tap_w1,
                               const ConeConsts& cc, double* wq) {
  double q = 0.0;
  for (int base = threadIdx.x * kConeVec; base < P; base += THREADS * kConeVec) {
    const float4 vx4 = *reinterpret_cast<const float4*>(cx + base);
    const float4 vy4 = *reinterpret_cast<const float4*>(cy + base);
    const float4 vz4 = *reinterpret_cast<const float4*>(cz + base);
    const float vcx[4] = {vx4.x, vx4.y, vx4.z, vx4.w};
    const float vcy[4] = {vy4.x, vy4.y, vy4.z, vy4.w};
    const float vcz[4] = {vz4.x, vz4.y, vz4.z, vz4.w};
    float tw0[4], tw1[4];
    int ti0[4], ti1[4];
    load_taps<AF>(base, P, tap_i0, tap_i1, tap_w0, tap_w1, ti0, ti1, tw0, tw1);
    float av[4], tcv[4];
    afactor4<AF>(ev, A, base, ti0, ti1, tw0, tw1, av);
    cone_tc4<SIG, false>(ev, vcx, vcy, vcz, cc, tcv, nullptr);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float t = base + k < P ? tcv[k] * av[k] : 0.0f;
      q = fma((double)t, (double)t, q);
    }
  }
  q = jscc_wave_sum(q);
  if (jscc_lane() == 0) wq[jscc_wave_id()] = q;
  __syncthreads();
  double t = wq[0];
  for (int w = 1; w < THREADS / JSCC_WAVE; ++w) t += wq[w];
  __syncthreads();
  return t;
}

template <int SIG, int AF, int OUT_DT, bool WANT_TC, int E>
__global__ void __launch_bounds__(kConeThreads) k_cone_rows(
    const JsccEvent* __restrict__ ev, const int32_t* __restrict__ ev_idx, int n_rows,
    const float* __restrict__ cx, const float* __restrict__ cy, const float* __restrict__ cz, int P,
    const float* __restrict__ A, int64_t lda, const int32_t* __restrict__ tap_i0,
    const int32_t* __restrict__ tap_i1, const float* __restrict__ tap_w0, const float* __restrict__ tap_w1,
    ConeConsts cc, void* __restrict__ out_t, int64_t ldt, float* __restrict__ out_tc, int64_t ldtc,
    double* __restrict__ stats, const int32_t* __restrict__ n_rows_dev) {
  const int row0 = blockIdx.x * E;
  const int n_grid = n_rows;  // rows the grid covers
  if (n_rows_dev) {
    // the row count read on the device (at most the grid's rows): the grid's rows past it get NaN
    // statistics, which the keep decision drops, and no row values
    const int nd = *n_rows_dev;
    n_rows = nd < 0 ? 0 : (nd < n_rows ? nd : n_rows);
    if (row0 >= n_rows) {
      if (threadIdx.x < E * 4 && row0 + (int)threadIdx.x / 4 < n_grid)
        stats[(int64_t)row0 * 4 + threadIdx.x] = __builtin_nan("");
      return;
    }
  }
  ConeEv evs[E];
  bool valid[E];
  int rrow[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    valid[e] = (row0 + e) < n_rows;
    rrow[e] = valid[e] ? (row0 + e) : (n_rows - 1);
    evs[e] = load_event(ev[ev_idx[rrow[e]]], lda);
  }
  // per-thread partial sums in FP32 (at most P/1024 tiles x 4 values each),
  // combined across threads in FP64.  A row whose largest value is below 2^-40
  // gets its sum of squares again in FP64 (cone_row_q64): its effective support
  // S^2/Q must match the reference's scale-free 1/sum((t/S)^2) even where the
  // squares of its tail fall under the FP32 normal range.
  float accS[E], accQ[E], accC[E], accM[E];
#pragma unroll
  for (int e = 0; e < E; ++e) { accS[e] = 0.0f; accQ[e] = 0.0f; accC[e] = 0.0f; accM[e] = 0.0f; }

  for (int base = threadIdx.x * kConeVec; base < P; base += kConeThreads * kConeVec)
    cone_rows_tile<SIG, AF, OUT_DT, WANT_TC, E>(evs, rrow, base, P, cx, cy, cz, A, tap_i0, tap_i1, tap_w0,
                                                tap_w1, cc, out_t, ldt, out_tc, ldtc, accS, accQ, accC, accM);

  if (AF == AF_NONE) {
#pragma unroll
    for (int e = 0; e < E; ++e) accC[e] = accS[e];
  }
  // block reduction of the E x 4 statistics
  __shared__ double red[4][kConeThreads / JSCC_WAVE][E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    double s = jscc_wave_sum((double)accS[e]);
    double q = jscc_wave_sum((double)accQ[e]);
    double c = jscc_wave_sum((double)accC[e]);
    float m = jscc_wave_max(accM[e]);
    if (jscc_lane() == 0) {
      red[0][jscc_wave_id()][e] = s;
      red[1][jscc_wave_id()][e] = q;
      red[2][jscc_wave_id()][e] = c;
      red[3][jscc_wave_id()][e] = (double)m;
    }
  }
  __syncthreads();
  __shared__ double q64[E], wq[kConeThreads / JSCC_WAVE];
  bool fix[E];
  bool any_fix = false;
#pragma unroll
  for (int e = 0; e < E; ++e) {  // the row maxima, read by every thread: a uniform decision
    double m = red[3][0][e];
    for (int w = 1; w < kConeThreads / JSCC_WAVE; ++w) m = fmax(m, red[3][w][e]);
    fix[e] = valid[e] && !(m >= kQ64Below);
    any_fix |= fix[e];
  }
  if (any_fix) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (fix[e]) {
        const double q = cone_row_q64<SIG, AF, kConeThreads>(evs[e], P, cx, cy, cz, A, tap_i0, tap_i1, tap_w0, tap_w1,
                                                             cc, wq);
        if (threadIdx.x == 0) q64[e] = q;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < E * 4) {
    const int e = threadIdx.x / 4, k = threadIdx.x % 4;
    if (row0 + e < n_rows) {
      double v = red[k][0][e];
      for (int w = 1; w < kConeThreads / JSCC_WAVE; ++w) v = (k == 3) ? fmax(v, red[k][w][e]) : v + red[k][w][e];
      if (k == 1 && fix[e]) v = q64[e];
      stats[(int64_t)(row0 + e) * 4 + k] = v;
    } else if (row0 + e < n_grid) {
      stats[(int64_t)(row0 + e) * 4 + k] = __builtin_nan("");
    }
  }
}

template <int SIG, int AF, int OUT_DT, bool WANT_TC>
void launch_cone(int n_rows, hipStream_t s, const JsccEvent* ev, const int32_t* idx, const float* cx,
                 const float* cy, const float* cz, int P, const float* A, int64_t lda, const int32_t* ti0,
                 const int32_t* ti1, const float* tw0, const float* tw1, ConeConsts cc, void* out_t,
                 int64_t ldt, float* out_tc, int64_t ldtc, double* stats, const int32_t* n_dev) {
  constexpr int E = 4;
  dim3 grid((unsigned)((n_rows + E - 1) / E));
  hipLaunchKernelGGL((k_cone_rows<SIG, AF, OUT_DT, WANT_TC, E>), grid, dim3(kConeThreads), 0, s, ev, idx, n_rows,
                     cx, cy, cz, P, A, lda, ti0, ti1, tw0, tw1, cc, out_t, ldt, out_tc, ldtc, stats, n_dev);
}

template <int SIG, int AF>
int dispatch_out(int out_dtype, bool want_tc, int n_rows, hipStream_t s, const JsccEvent* ev, const int32_t* idx,
                 const float* cx, const float* cy, const float* cz, int P, const float* A, int64_t lda,
                 const int32_t* ti0, const int32_t* ti1, const float* tw0, const float* tw1, ConeConsts cc,
                 void* out_t, int64_t ldt, float* out_tc, int64_t ldtc, double* stats, const int32_t* n_dev) {
#define JSCC_CONE_CASE(DT, TC)                                                                              \
  launch_cone<SIG, AF, DT, TC>(n_rows, s, ev, idx, cx, cy, cz, P, A, lda, ti0, ti1, tw0, tw1, cc, out_t, ldt, \
                               out_tc, ldtc, stats, n_dev)
  if (out_dtype == JSCC_DT_F32) {
    if (want_tc) JSCC_CONE_CASE(JSCC_DT_F32, true); else JSCC_CONE_CASE(JSCC_DT_F32, false);
  } else if (out_dtype == JSCC_DT_BF16) {
    if (want_tc) JSCC_CONE_CASE(JSCC_DT_BF16, true); else JSCC_CONE_CASE(JSCC_DT_BF16, false);
  } else if (out_dtype == -1) {  // statistics (+ optional t_c) only
    if (want_tc) JSCC_CONE_CASE(-1, true); else JSCC_CONE_CASE(-1, false);
  } else {
    jscc_set_error("jscc_cone_rows: unsupported output dtype %d", out_dtype);
    return -2;
  }
#undef JSCC_CONE_CASE
  return 0;
}

// ----------------------------------------------------------------------------
// Cone rows with the list-mode sensitivity in the same pass.  A persistent
// grid (kSensBlocksPerCu work-groups of 512 threads per CU; 1 of 1024 above
// 16384 voxels) takes groups of 8 rows (two halves of 4 events) from a ticket
// counter, so work-groups on slower CUs simply take fewer groups (a static
// split measured 25 % between the fastest and slowest work-group).  Every row
// is evaluated and stored as in k_cone_rows, its statistics reduced (one pair
// of barriers per group), the keep decision of jscc_rows_keep_flags taken
// (f = 1/S for a kept row, 0 otherwise).  The rows of group g are read back
// (each thread its own stored values, the same bits the column-sum pass would
// read) during the tile loop of the work-group's next group, one voxel quad
// per tile, so the loads have a whole tile of cone arithmetic to arrive;
// f x value is added into the work-group's LDS column sums (each thread owns
// its quads: no barrier).  At the end every work-group writes its column sums
// and k_sens_reduce adds them in work-group order.  The split of groups
// between work-groups follows the ticket order, so the FP32 grouping of the
// sums (the last bits of sens) can differ between runs.
// ----------------------------------------------------------------------------
constexpr int kSensBlocksPerCu = 2;  // 512-thread work-groups per CU (64 KB of column sums each)
constexpr int kSensHalves = 2;       // groups of 2 x 4 rows: one pair of barriers per 8 rows

template <int THREADS>
constexpr int sens_blocks_per_cu() { return THREADS == 512 ? kSensBlocksPerCu : 1; }

// the ticket counter to 0 and the zero row (read back in place of dropped rows) to zeros
__global__ void k_sens_prep(int* ticket, float* zrow, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *ticket = 0;
  for (int j = i; j < n; j += gridDim.x * blockDim.x) zrow[j] = 0.0f;
}

// FP16 rows (OUT_DT = JSCC_DT_FP16; groups of 4 rows): the cone values of a
// group go as FP32 into a one-group slot per work-group, and right after the
// group's statistics (while the slot is still in the L2 / Infinity Cache) they
// are quantised like jscc_quantize_rows (code = half(t / q), q = max / 2^14,
// row_scale = q / S), the codes stored and row_scale x code added -- the column
// sums of the rows the OSEM will project.  A dropped row is read from the zero
// row (its codes are stored as zeros; the OSEM never reads it).  Reading back
// one group later, as for BF16 rows, kept 2 x 8 rows per work-group live (512 MB
// at 16384 voxels, beyond the Infinity Cache): 5.23 against 4.32 ms per pass.
template <int SIG, int AF, int OUT_DT, int THREADS>
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_cone_rows_sens(
    const JsccEvent* __restrict__ ev, const int32_t* __restrict__ ev_idx, int n_grid,
    const int32_t* __restrict__ n_rows_dev, const float* __restrict__ cx, const float* __restrict__ cy,
    const float* __restrict__ cz, int P, const float* __restrict__ A, int64_t lda, ConeConsts cc,
    void* __restrict__ out_t, int64_t ldt, double* __restrict__ stats, double min_support,
    float* __restrict__ partial, int64_t ldp, int* __restrict__ ticket, const float* __restrict__ zrow,
    float* __restrict__ ring, float* __restrict__ row_scale) {  // ring: the FP16 rows' group slots
  constexpr int E = 4;              // events evaluated together (scalar registers)
  constexpr bool Q16 = OUT_DT == JSCC_DT_FP16;
  constexpr int H = Q16 ? 1 : kSensHalves;  // E-event halves per group: one barrier round per H * E rows
  constexpr int R = H * E;          // rows per group
  constexpr int NW = THREADS / JSCC_WAVE;
  constexpr int TILE_DT = Q16 ? JSCC_DT_F32 : OUT_DT;  // what the tile loop stores
  using Raw = typename std::conditional<TILE_DT == JSCC_DT_F32, float4, uint2>::type;
  extern __shared__ float4 sacc[];  // [ldp / 4] column sums; quad q is owned by thread q % THREADS
  __shared__ double red[4][NW][R];
  __shared__ double q64[R], wq[NW];
  __shared__ float4 fsh4[H];  // the factors f of the rows of the work-group's last finished group
  __shared__ float4 qsh4[H];  // FP16 rows: 1 / q of those rows
  float* fsh = reinterpret_cast<float*>(fsh4);
  float* qsh = reinterpret_cast<float*>(qsh4);
  __shared__ int next_g;
  const int tid = threadIdx.x;
  int n_rows = n_grid;
  if (n_rows_dev) {
    const int nd = *n_rows_dev;
    n_rows = nd < 0 ? 0 : (nd < n_grid ? nd : n_grid);
  }
  // rows of the grid past the device count: NaN statistics (dropped by the keep decision)
  for (int64_t r = (int64_t)n_rows + (int64_t)blockIdx.x * THREADS + tid; r < n_grid;
       r += (int64_t)gridDim.x * THREADS) {
#pragma unroll
    for (int k = 0; k < 4; ++k) stats[r * 4 + k] = __builtin_nan("");
  }
  const int n_quads = (int)(ldp / 4);
  for (int q = tid; q < n_quads; q += THREADS) sacc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int n_groups = (n_rows + R - 1) / R;
  if (tid < H) {  // no previous group: nothing is added
    fsh4[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
    qsh4[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (tid == 0) next_g = atomicAdd(ticket, 1);
  __syncthreads();
  int g = next_g;
  __syncthreads();
  // the previous group of this work-group: its first row and the factors f of its rows (uniform)
  int prev_row0 = -1;
  float pf[R];
#pragma unroll
  for (int e = 0; e < R; ++e) pf[e] = 0.0f;
  float pf_last = 0.0f;  // f of the previous group's last row below the count (FP16 rows: of this group)

  while (g < n_groups) {
    const int row0 = g * R;
    bool valid[R];
    float accS[R], accQ[R], accC[R], accM[R];
#pragma unroll
    for (int e = 0; e < R; ++e) {
      valid[e] = (row0 + e) < n_rows;
      accS[e] = 0.0f; accQ[e] = 0.0f; accC[e] = 0.0f; accM[e] = 0.0f;
    }
    float* ring_cur = Q16 ? ring + (int64_t)blockIdx.x * R * ldp : nullptr;  // one slot: read back in the group
#pragma unroll
    for (int h = 0; h < H; ++h) {
      ConeEv evs[E];
      int rrow[E];  // rows the tile loop stores to (rows of the group's slot for FP16 rows)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = valid[h * E + e] ? (row0 + h * E + e) : (n_rows - 1);
        evs[e] = load_event(ev[ev_idx[r]], lda);
        rrow[e] = Q16 ? r - row0 : r;
      }
      // rows of the previous group's half h, as stored (not for FP16 rows, read back within their group); a
      // dropped row (f = 0, possibly holding NaN) and every row before the first group read the zero row
      // instead, so the additions below need no branch
      const Raw* prow[E];
#pragma unroll
      for (int e = 0; e < E; ++e)
        prow[e] = !Q16 && pf[h * E + e] != 0.0f
                      ? reinterpret_cast<const Raw*>(reinterpret_cast<const char*>(out_t) +
                                                     (int64_t)(prev_row0 + h * E + e) * ldt * (int64_t)(sizeof(Raw) / 4))
                      : reinterpret_cast<const Raw*>(zrow);
      for (int q = tid; q * kConeVec < P; q += THREADS) {
        const int base = q * kConeVec;
        const uint32_t coff = (uint32_t)base * 4u;
        const float4 vx4 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(cx) + coff);
        const float4 vy4 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(cy) + coff);
        const float4 vz4 = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(cz) + coff);
        // the read-back loads are issued after the coordinate loads (memory clobber: no reordering), so
        // waiting for the coordinates does not wait for them (loads complete in order); unconditional, since
        // a branch here would make that wait drain them too
        asm volatile("" ::: "memory");
        Raw rb[E];
        const uint32_t qoff = (uint32_t)q * (uint32_t)sizeof(Raw);  // one 32-bit offset from scalar row bases
        if constexpr (!Q16) {
#pragma unroll
          for (int e = 0; e < E; ++e)
            rb[e] = *reinterpret_cast<const Raw*>(reinterpret_cast<const char*>(prow[e]) + qoff);
        }
        cone_rows_tile_at<SIG, AF, TILE_DT, false, E>(evs, rrow, base, P, vx4, vy4, vz4, A, nullptr, nullptr,
                                                      nullptr, nullptr, cc, Q16 ? (void*)ring_cur : out_t,
                                                      Q16 ? ldp : ldt, nullptr, 0, accS + h * E, accQ + h * E,
                                                      accC + h * E, accM + h * E);
        if constexpr (Q16) continue;  // FP16 rows are read back after the group's statistics (below)
        float4 a = sacc[q];
        float w[4] = {a.x, a.y, a.z, a.w};
        // f of the previous group's rows read from LDS (a vector register operand; the scalar registers are
        // taken by the event parameters and row bases).  fsh holds them until the next group's finalisation.
        const float4 f4 = fsh4[h];
        const float fv[E] = {f4.x, f4.y, f4.z, f4.w};
#pragma unroll
        for (int e = 0; e < E; ++e) {
          float v[4];
          if constexpr (TILE_DT == JSCC_DT_F32) {
            v[0] = rb[e].x; v[1] = rb[e].y; v[2] = rb[e].z; v[3] = rb[e].w;
          } else {
            v[0] = __uint_as_float(rb[e].x << 16); v[1] = __uint_as_float(rb[e].x & 0xffff0000u);
            v[2] = __uint_as_float(rb[e].y << 16); v[3] = __uint_as_float(rb[e].y & 0xffff0000u);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) w[k] = fmaf(fv[e], v[k], w[k]);
        }
        sacc[q] = make_float4(w[0], w[1], w[2], w[3]);
      }
    }
    if (AF == AF_NONE) {
#pragma unroll
      for (int e = 0; e < R; ++e) accC[e] = accS[e];
    }
#pragma unroll
    for (int e = 0; e < R; ++e) {
      const double sS = jscc_wave_sum((double)accS[e]);
      const double sQ = jscc_wave_sum((double)accQ[e]);
      const double sC = jscc_wave_sum((double)accC[e]);
      const float sM = jscc_wave_max(accM[e]);
      if (jscc_lane() == 0) {
        red[0][jscc_wave_id()][e] = sS;
        red[1][jscc_wave_id()][e] = sQ;
        red[2][jscc_wave_id()][e] = sC;
        red[3][jscc_wave_id()][e] = (double)sM;
      }
    }
    if (tid == 0) next_g = atomicAdd(ticket, 1);
    __syncthreads();
    const int g_next = next_g;
    bool fix[R];
    bool any_fix = false;
#pragma unroll
    for (int e = 0; e < R; ++e) {  // rows below 2^-40: the sum of squares again in FP64 (uniform decision)
      double m = red[3][0][e];
      for (int w = 1; w < NW; ++w) m = fmax(m, red[3][w][e]);
      fix[e] = valid[e] && !(m >= kQ64Below);
      any_fix |= fix[e];
    }
    if (any_fix) {
#pragma unroll
      for (int e = 0; e < R; ++e) {
        if (fix[e]) {
          const ConeEv evf = load_event(ev[ev_idx[row0 + e]], lda);
          const double q = cone_row_q64<SIG, AF, THREADS>(evf, P, cx, cy, cz, A, nullptr, nullptr, nullptr, nullptr,
                                                          cc, wq);
          if (tid == 0) q64[e] = q;
        }
      }
      __syncthreads();
    }
    if (tid < R) {
      const int e = tid;
      float f = 0.0f, inv_q = 0.0f;
      if (row0 + e < n_rows) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          double t = red[k][0][e];
          for (int w = 1; w < NW; ++w) t = (k == 3) ? fmax(t, red[k][w][e]) : t + red[k][w][e];
          if (k == 1 && fix[e]) t = q64[e];
          v[k] = t;
          stats[(int64_t)(row0 + e) * 4 + k] = t;
        }
        const bool keep = jscc_row_keep(v[0], v[1], v[3], min_support);
        if constexpr (Q16) {  // the per-row scale of jscc_quantize_rows (row max at 2^14)
          const double q = v[3] > 0.0 ? v[3] / 16384.0 : 1.0;
          row_scale[row0 + e] = (float)(q / v[0]);
          if (keep) {  // a dropped row reads the zero row: 1 / q stays 0 (it may not be finite for such a row)
            f = (float)(q / v[0]);
            inv_q = (float)(1.0 / q);
          }
        } else {
          if (keep) f = (float)(1.0 / v[0]);
        }
      }
      fsh[e] = f;
      qsh[e] = inv_q;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e) pf[e] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(fsh[e])));
    {
      const int last = n_rows - 1 - row0 < R - 1 ? n_rows - 1 - row0 : R - 1;
      pf_last = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(fsh[last])));
    }
    if constexpr (Q16) {
      // FP16 rows: this group's FP32 values, still in the L2 / Infinity Cache, quantised, stored and summed
      // now; a row past the count repeats the stores of the last row (the same codes) and adds nothing
      const int last = n_rows - 1 - row0;
      float iq[R], fr[R];
      const float* src[R];
      uint16_t* dst[R];
#pragma unroll
      for (int e = 0; e < R; ++e) {
        const int le = e < last ? e : last;
        fr[e] = e < last ? pf[e] : pf_last;
        iq[e] = qsh[le];
        src[e] = fr[e] != 0.0f ? ring_cur + (int64_t)le * ldp : zrow;
        dst[e] = reinterpret_cast<uint16_t*>(out_t) + (int64_t)(row0 + le) * ldt;
        fr[e] = e <= last ? fr[e] : 0.0f;  // rows past the count add nothing
      }
      for (int q = tid; q * kConeVec < P; q += THREADS) {
        float4 rv[R];
#pragma unroll
        for (int e = 0; e < R; ++e) rv[e] = reinterpret_cast<const float4*>(src[e])[q];
        float4 a = sacc[q];
#pragma unroll
        for (int e = 0; e < R; ++e) {
          _Float16 c[4];
          c[0] = (_Float16)(rv[e].x * iq[e]); c[1] = (_Float16)(rv[e].y * iq[e]);
          c[2] = (_Float16)(rv[e].z * iq[e]); c[3] = (_Float16)(rv[e].w * iq[e]);
          uint2 pk;
          pk.x = (uint32_t)__builtin_bit_cast(uint16_t, c[0]) | ((uint32_t)__builtin_bit_cast(uint16_t, c[1]) << 16);
          pk.y = (uint32_t)__builtin_bit_cast(uint16_t, c[2]) | ((uint32_t)__builtin_bit_cast(uint16_t, c[3]) << 16);
          *reinterpret_cast<uint2*>(reinterpret_cast<char*>(dst[e]) + (uint32_t)q * 8u) = pk;
          a.x = fmaf(fr[e], (float)c[0], a.x); a.y = fmaf(fr[e], (float)c[1], a.y);
          a.z = fmaf(fr[e], (float)c[2], a.z); a.w = fmaf(fr[e], (float)c[3], a.w);
        }
        sacc[q] = a;
      }
    }
    prev_row0 = row0;
    g = g_next;
  }
  // the rows of the work-group's last group (FP16 rows: already read back within it)
  if (!Q16 && prev_row0 >= 0) {
#pragma unroll
    for (int e = 0; e < R; ++e) {
      if (prev_row0 + e < n_rows && pf[e] != 0.0f) {
        const Raw* pr = reinterpret_cast<const Raw*>(reinterpret_cast<const char*>(out_t) +
                                                     (int64_t)(prev_row0 + e) * ldt * (int64_t)(sizeof(Raw) / 4));
        for (int q = tid; q * kConeVec < P; q += THREADS) {
          const Raw rw = pr[q];
          float v[4];
          if constexpr (TILE_DT == JSCC_DT_F32) {
            v[0] = rw.x; v[1] = rw.y; v[2] = rw.z; v[3] = rw.w;
          } else {
            v[0] = __uint_as_float(rw.x << 16); v[1] = __uint_as_float(rw.x & 0xffff0000u);
            v[2] = __uint_as_float(rw.y << 16); v[3] = __uint_as_float(rw.y & 0xffff0000u);
          }
          float4 a = sacc[q];
          a.x = fmaf(pf[e], v[0], a.x); a.y = fmaf(pf[e], v[1], a.y);
          a.z = fmaf(pf[e], v[2], a.z); a.w = fmaf(pf[e], v[3], a.w);
          sacc[q] = a;
        }
      }
    }
  }
  __syncthreads();
  float4* dst = reinterpret_cast<float4*>(partial + (int64_t)blockIdx.x * ldp);
  for (int q = tid; q < n_quads; q += THREADS) dst[q] = sacc[q];
}

// sens[p] = sum over work-groups b (in order) of partial[b, p]: 64 columns per block, 16 groups of 16 threads
// sum every 16th partial row, then one thread per column quad adds the 16 group sums in order.
__global__ void __launch_bounds__(256) k_sens_reduce(const float* __restrict__ partial, int G, int64_t ldp, int P,
                                                     float* __restrict__ sens) {
  __shared__ float4 acc[16][16];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int col = blockIdx.x * 64 + tx * 4;
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  if (col < P) {
    for (int b = ty; b < G; b += 16) {
      const float4 v = *reinterpret_cast<const float4*>(partial + (int64_t)b * ldp + col);
      sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
    }
  }
  acc[ty][tx] = sum;
  __syncthreads();
  if (ty == 0 && col < P) {
    float4 t = acc[0][tx];
#pragma unroll
    for (int k = 1; k < 16; ++k) {
      t.x += acc[k][tx].x; t.y += acc[k][tx].y; t.z += acc[k][tx].z; t.w += acc[k][tx].w;
    }
    const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (col + k < P) sens[col + k] = tv[k];
  }
}

// threads per work-group of the fused pass at P voxels (0: too many voxels)
int sens_threads(int64_t P) { return P <= 16384 ? 512 : (P <= 32768 ? 1024 : 0); }
int sens_grid(int64_t P) {
  return (sens_threads(P) == 512 ? kSensBlocksPerCu : 1) * jscc_num_cus();
}
size_t sens_lds(int64_t P) { return (size_t)((P + 3) / 4) * sizeof(float4); }

// workspace: the work-groups' column sums, the ticket counter (256 bytes), the zero row, and for FP16 rows a
// slot of 4 FP32 rows per work-group
int64_t sens_workspace(int64_t P, int out_dtype) {
  const int64_t ldp = (P + 3) / 4 * 4, G = sens_grid(P);
  return (G + 1) * ldp * (int64_t)sizeof(float) + 256 +
         (out_dtype == JSCC_DT_FP16 ? G * 4 * ldp * (int64_t)sizeof(float) : 0);
}

template <int SIG, int AF, int OUT_DT, int THREADS>
int launch_sens_t(int G, hipStream_t s, const JsccEvent* ev, const int32_t* idx, int n_grid, const int32_t* n_dev,
                  const float* cx, const float* cy, const float* cz, int P, const float* A, int64_t lda,
                  ConeConsts cc, void* out_t, int64_t ldt, double* stats, double min_support, float* partial,
                  int64_t ldp, int* ticket, float* row_scale) {
  static bool attr = false;  // more than 64 KB of dynamic LDS needs the attribute (set once)
  if (sens_lds(P) > 65536 && !attr) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cone_rows_sens<SIG, AF, OUT_DT, THREADS>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sens_lds(32768)));
    attr = true;
  }
  float* zrow = reinterpret_cast<float*>(reinterpret_cast<char*>(ticket) + 256);
  float* ring = zrow + ldp;
  hipLaunchKernelGGL(k_sens_prep, dim3(16), dim3(256), 0, s, ticket, zrow, (int)ldp);
  hipLaunchKernelGGL((k_cone_rows_sens<SIG, AF, OUT_DT, THREADS>), dim3(G), dim3(THREADS), sens_lds(P), s, ev, idx,
                     n_grid, n_dev, cx, cy, cz, P, A, lda, cc, out_t, ldt, stats, min_support, partial, ldp, ticket,
                     zrow, ring, row_scale);
  return 0;
}

template <int SIG, int AF, int OUT_DT>
int launch_sens(int G, hipStream_t s, const JsccEvent* ev, const int32_t* idx, int n_grid, const int32_t* n_dev,
                const float* cx, const float* cy, const float* cz, int P, const float* A, int64_t lda, ConeConsts cc,
                void* out_t, int64_t ldt, double* stats, double min_support, float* partial, int64_t ldp,
                int* ticket, float* row_scale) {
  if (sens_threads(P) == 512)
    return launch_sens_t<SIG, AF, OUT_DT, 512>(G, s, ev, idx, n_grid, n_dev, cx, cy, cz, P, A, lda, cc, out_t, ldt,
                                               stats, min_support, partial, ldp, ticket, row_scale);
  return launch_sens_t<SIG, AF, OUT_DT, 1024>(G, s, ev, idx, n_grid, n_dev, cx, cy, cz, P, A, lda, cc, out_t, ldt,
                                              stats, min_support, partial, ldp, ticket, row_scale);
}

}  // namespace

static int cone_rows_impl(const void* events, const int32_t* ev_idx, int64_t n_rows, const int32_t* n_rows_dev,
                          const float* coor_x, const float* coor_y, const float* coor_z, int64_t n_vox,
                          const float* sysmat, int64_t ld_sysmat, const int32_t* tap_i0, const int32_t* tap_i1,
                          const float* tap_w0, const float* tap_w1, int sigma_model, double fixed_sigma,
                          void* out_t, int out_dtype, int64_t ld_out, float* out_tc, int64_t ld_tc, double* stats,
                          void* stream) {
  JSCC_REQUIRE(n_rows >= 0 && n_rows < ((int64_t)1 << 31), "jscc_cone_rows: n_rows out of range");
  if (n_rows == 0) return 0;
  JSCC_REQUIRE(events && ev_idx && coor_x && coor_y && coor_z && stats, "jscc_cone_rows: null pointer");
  JSCC_REQUIRE(n_vox > 0 && n_vox < ((int64_t)1 << 30), "jscc_cone_rows: bad voxel count");
  JSCC_REQUIRE(((uintptr_t)coor_x | (uintptr_t)coor_y | (uintptr_t)coor_z) % 16 == 0,
               "jscc_cone_rows: coordinate arrays must be 16-byte aligned and padded to a multiple of 4");
  JSCC_REQUIRE(sigma_model == SIG_STRICT || sigma_model == SIG_FIXED, "jscc_cone_rows: bad sigma model");
  if (out_t) {
    JSCC_REQUIRE(ld_out >= n_vox && ld_out % 4 == 0, "jscc_cone_rows: ld_out must be >= n_vox and a multiple of 4");
    JSCC_REQUIRE((uintptr_t)out_t % 16 == 0, "jscc_cone_rows: out_t must be 16-byte aligned");
  } else {
    out_dtype = -1;
  }
  if (out_tc) {
    JSCC_REQUIRE(ld_tc >= n_vox && ld_tc % 4 == 0 && (uintptr_t)out_tc % 16 == 0,
                 "jscc_cone_rows: out_tc must be 16-byte aligned with ld a multiple of 4");
  }
  const bool taps = tap_i0 != nullptr;
  if (taps) JSCC_REQUIRE(tap_i1 && tap_w0 && tap_w1 && sysmat, "jscc_cone_rows: incomplete tap tables");
  if (sysmat && !taps) {
    JSCC_REQUIRE(ld_sysmat >= n_vox && ld_sysmat % 4 == 0 && (uintptr_t)sysmat % 16 == 0,
                 "jscc_cone_rows: sysmat rows must be 16-byte aligned with ld a multiple of 4");
  }
  if (sigma_model == SIG_FIXED) JSCC_REQUIRE(fixed_sigma > 0.0, "jscc_cone_rows: fixed sigma must be positive");

  const ConeConsts cc = make_consts(sigma_model, fixed_sigma);

  hipStream_t s = (hipStream_t)stream;
  const JsccEvent* ev = (const JsccEvent*)events;
  const bool want_tc = out_tc != nullptr;
  const int P = (int)n_vox;
  const int n = (int)n_rows;
  int rc;
  jscc_prof_begin(JSCC_TAG_CONE, s);
#define JSCC_CONE_AF(SIG)                                                                                   \
  if (taps)                                                                                                   \
    rc = dispatch_out<SIG, AF_TAPS>(out_dtype, want_tc, n, s, ev, ev_idx, coor_x, coor_y, coor_z, P, sysmat,  \
                                    ld_sysmat, tap_i0, tap_i1, tap_w0, tap_w1, cc, out_t, ld_out, out_tc,      \
                                    ld_tc, stats, n_rows_dev);                                                \
  else if (sysmat)                                                                                            \
    rc = dispatch_out<SIG, AF_ROW>(out_dtype, want_tc, n, s, ev, ev_idx, coor_x, coor_y, coor_z, P, sysmat,   \
                                   ld_sysmat, tap_i0, tap_i1, tap_w0, tap_w1, cc, out_t, ld_out, out_tc,       \
                                   ld_tc, stats, n_rows_dev);                                                 \
  else                                                                                                        \
    rc = dispatch_out<SIG, AF_NONE>(out_dtype, want_tc, n, s, ev, ev_idx, coor_x, coor_y, coor_z, P, sysmat,  \
                                    ld_sysmat, tap_i0, tap_i1, tap_w0, tap_w1, cc, out_t, ld_out, out_tc,      \
                                    ld_tc, stats, n_rows_dev);
  if (sigma_model == SIG_STRICT) {
    JSCC_CONE_AF(SIG_STRICT)
  } else {
    JSCC_CONE_AF(SIG_FIXED)
  }
#undef JSCC_CONE_AF
  jscc_prof_end(JSCC_TAG_CONE, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_cone_rows(const void* events, const int32_t* ev_idx, int64_t n_rows, const float* coor_x,
                              const float* coor_y, const float* coor_z, int64_t n_vox, const float* sysmat,
                              int64_t ld_sysmat, const int32_t* tap_i0, const int32_t* tap_i1, const float* tap_w0,
                              const float* tap_w1, int sigma_model, double fixed_sigma, void* out_t,
                              int out_dtype, int64_t ld_out, float* out_tc, int64_t ld_tc, double* stats,
                              void* stream) {
  return cone_rows_impl(events, ev_idx, n_rows, nullptr, coor_x, coor_y, coor_z, n_vox, sysmat, ld_sysmat, tap_i0,
                        tap_i1, tap_w0, tap_w1, sigma_model, fixed_sigma, out_t, out_dtype, ld_out, out_tc, ld_tc,
                        stats, stream);
}

extern "C" int jscc_cone_rows_counted(const void* events, const int32_t* ev_idx, const int32_t* n_rows_dev,
                                      int64_t n_rows_max, const float* coor_x, const float* coor_y,
                                      const float* coor_z, int64_t n_vox, const float* sysmat, int64_t ld_sysmat,
                                      const int32_t* tap_i0, const int32_t* tap_i1, const float* tap_w0,
                                      const float* tap_w1, int sigma_model, double fixed_sigma, void* out_t,
                                      int out_dtype, int64_t ld_out, float* out_tc, int64_t ld_tc, double* stats,
                                      void* stream) {
  JSCC_REQUIRE(n_rows_dev != nullptr, "jscc_cone_rows_counted: null row count");
  return cone_rows_impl(events, ev_idx, n_rows_max, n_rows_dev, coor_x, coor_y, coor_z, n_vox, sysmat, ld_sysmat,
                        tap_i0, tap_i1, tap_w0, tap_w1, sigma_model, fixed_sigma, out_t, out_dtype, ld_out, out_tc,
                        ld_tc, stats, stream);
}

extern "C" int jscc_cone_sens_grid(int64_t n_vox) { return sens_threads(n_vox) ? sens_grid(n_vox) : 0; }

extern "C" int64_t jscc_cone_sens_workspace_bytes(int64_t n_vox, int out_dtype) {
  return sens_threads(n_vox) ? sens_workspace(n_vox, out_dtype) : 0;
}

extern "C" int jscc_cone_rows_sens(const void* events, const int32_t* ev_idx, const int32_t* n_rows_dev,
                                   int64_t n_rows_max, const float* coor_x, const float* coor_y, const float* coor_z,
                                   int64_t n_vox, const float* sysmat, int64_t ld_sysmat, int sigma_model,
                                   double fixed_sigma, void* out_t, int out_dtype, int64_t ld_out, float* row_scale,
                                   double* stats, double min_support, float* sens, void* workspace,
                                   int64_t workspace_bytes, void* stream) {
  JSCC_REQUIRE(n_rows_max >= 0 && n_rows_max < ((int64_t)1 << 31), "jscc_cone_rows_sens: n_rows out of range");
  JSCC_REQUIRE(events && ev_idx && coor_x && coor_y && coor_z && stats && out_t && sens,
               "jscc_cone_rows_sens: null pointer");
  JSCC_REQUIRE(n_vox > 0 && sens_threads(n_vox) > 0, "jscc_cone_rows_sens: %lld voxels (1 to 32768)",
               (long long)n_vox);
  JSCC_REQUIRE(((uintptr_t)coor_x | (uintptr_t)coor_y | (uintptr_t)coor_z) % 16 == 0,
               "jscc_cone_rows_sens: coordinate arrays must be 16-byte aligned and padded to a multiple of 4");
  JSCC_REQUIRE(sigma_model == SIG_STRICT || sigma_model == SIG_FIXED, "jscc_cone_rows_sens: bad sigma model");
  JSCC_REQUIRE(out_dtype == JSCC_DT_F32 || out_dtype == JSCC_DT_BF16 || out_dtype == JSCC_DT_FP16,
               "jscc_cone_rows_sens: rows are stored as FP32, BF16 or FP16 (per-row scale)");
  JSCC_REQUIRE(out_dtype != JSCC_DT_FP16 || row_scale != nullptr, "jscc_cone_rows_sens: FP16 rows need row_scale");
  JSCC_REQUIRE(ld_out >= n_vox && ld_out % 4 == 0 && (uintptr_t)out_t % 16 == 0,
               "jscc_cone_rows_sens: out_t must be 16-byte aligned with ld >= n_vox, a multiple of 4");
  if (sysmat) {
    JSCC_REQUIRE(ld_sysmat >= n_vox && ld_sysmat % 4 == 0 && (uintptr_t)sysmat % 16 == 0,
                 "jscc_cone_rows_sens: sysmat rows must be 16-byte aligned with ld a multiple of 4");
  }
  if (sigma_model == SIG_FIXED) JSCC_REQUIRE(fixed_sigma > 0.0, "jscc_cone_rows_sens: fixed sigma must be positive");
  JSCC_REQUIRE(workspace && workspace_bytes >= sens_workspace(n_vox, out_dtype) && (uintptr_t)workspace % 16 == 0,
               "jscc_cone_rows_sens: workspace too small or misaligned");
  hipStream_t s = (hipStream_t)stream;
  const int P = (int)n_vox;
  const int64_t ldp = (n_vox + 3) / 4 * 4;
  if (n_rows_max == 0) {  // no rows: zero sums (the reduction over no partial rows)
    hipLaunchKernelGGL(k_sens_reduce, dim3((unsigned)((P + 63) / 64)), dim3(256), 0, s, (const float*)workspace, 0,
                       ldp, P, sens);
    JSCC_CHECK_LAUNCH();
    return 0;
  }
  const ConeConsts cc = make_consts(sigma_model, fixed_sigma);
  const JsccEvent* ev = (const JsccEvent*)events;
  const int G = sens_grid(n_vox);
  const int n = (int)n_rows_max;
  float* partial = (float*)workspace;
  int* ticket = (int*)((char*)workspace + (int64_t)G * ldp * (int64_t)sizeof(float));
  int rc = 0;
  jscc_prof_begin(JSCC_TAG_CONE, s);
#define JSCC_SENS(SIG, AF, DT)                                                                                        \
  rc = launch_sens<SIG, AF, DT>(G, s, ev, ev_idx, n, n_rows_dev, coor_x, coor_y, coor_z, P, sysmat, ld_sysmat, cc, \
                                out_t, ld_out, stats, min_support, partial, ldp, ticket, row_scale)
#define JSCC_SENS_DT(SIG, AF)                                   \
  if (out_dtype == JSCC_DT_F32) JSCC_SENS(SIG, AF, JSCC_DT_F32); \
  else if (out_dtype == JSCC_DT_BF16) JSCC_SENS(SIG, AF, JSCC_DT_BF16); \
  else JSCC_SENS(SIG, AF, JSCC_DT_FP16);
  if (sigma_model == SIG_STRICT) {
    if (sysmat) { JSCC_SENS_DT(SIG_STRICT, AF_ROW) } else { JSCC_SENS_DT(SIG_STRICT, AF_NONE) }
  } else {
    if (sysmat) { JSCC_SENS_DT(SIG_FIXED, AF_ROW) } else { JSCC_SENS_DT(SIG_FIXED, AF_NONE) }
  }
#undef JSCC_SENS_DT
#undef JSCC_SENS
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sens_reduce, dim3((unsigned)((P + 63) / 64)), dim3(256), 0, s, (const float*)partial, G, ldp, P,
                     sens);
  jscc_prof_end(JSCC_TAG_CONE, s);
  JSCC_CHECK_LAUNCH();
  return 0;
}

// ============================================================================
// No-store cone projection: rows are evaluated in registers and consumed at
// once, so T never touches HBM.  Used when T does not fit in HBM (lists of
// many millions of events), by the differentiable operator's backward pass
// ("T recomputed in the backward") and for lazy event lists.
//   CP_RATIO: w += T_i^T / max(T_i x, eps)         (list-mode OSEM weight)
//   CP_FWD:   fwd[i] = T_i x                       (+ d(T_i x)/d sigma)
//   CP_COEF:  w += coef_i T_i^T                    (+ d(T_i x)/d sigma)
// T_i is the row divided by its sum when `normalize` (the reference's rows).
// ============================================================================
namespace {

enum { CP_RATIO = 0, CP_FWD = 1, CP_COEF = 2 };

struct ConeProjArgs {
  const JsccEvent* ev;
  const int32_t* ev_idx;
  int64_t n_rows;
  const float* cx;
  const float* cy;
  const float* cz;
  int P;
  const float* A;
  int64_t lda;
  ConeConsts cc;
  const float* x;
  const float* coef;
  float eps;
  int use_eps;
  int normalize;
  float* fwd_out;
  float* dsig_out;
  float* partial;
};

// Rows of more than 16384 voxels (8 quads per thread) keep their values in the
// thread's own LDS slots (NCH x BLOCK float4, dynamic) instead of registers:
// held next to the accumulators they would spill.
template <int BLOCK, int NCH>
constexpr bool cp_row_in_lds() { return NCH >= 8; }

template <int SIG, int AF, int BLOCK, int NCH, int MODE, bool DSIG>
__global__ void __launch_bounds__(BLOCK) k_cone_proj(ConeProjArgs a) {
  constexpr int NW = BLOCK / JSCC_WAVE;
  constexpr bool TRL = cp_row_in_lds<BLOCK, NCH>();
  const int tid = threadIdx.x;
  extern __shared__ float4 cp_rows[];
  float tr[TRL ? 1 : NCH][4];
  float wr[NCH][4];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int k = 0; k < 4; ++k) wr[c][k] = 0.0f;
  const int64_t per = a.n_rows / gridDim.x, rem = a.n_rows % gridDim.x;
  const int64_t r_begin = blockIdx.x * per + (blockIdx.x < rem ? blockIdx.x : rem);
  const int64_t r_end = r_begin + per + (blockIdx.x < rem ? 1 : 0);
  __shared__ double red[2][4][NW];
  int buf = 0;
  for (int64_t i = r_begin; i < r_end; ++i) {
    const ConeEv e = load_event(a.ev[a.ev_idx[i]], a.lda);
    float S = 0.0f, D = 0.0f, dS = 0.0f, dD = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      int base = (c * BLOCK + tid) * 4;
      // formed at use: the chunks' loads are not hoisted ahead of their arithmetic (which would hold
      // every chunk's coordinates in registers at once)
      asm volatile("" : "+v"(base));
      if (base < a.P) {
        const float4 vx4 = *reinterpret_cast<const float4*>(a.cx + base);
        const float4 vy4 = *reinterpret_cast<const float4*>(a.cy + base);
        const float4 vz4 = *reinterpret_cast<const float4*>(a.cz + base);
        const float4 x4 = *reinterpret_cast<const float4*>(a.x + base);
        const float vcx[4] = {vx4.x, vx4.y, vx4.z, vx4.w};
        const float vcy[4] = {vy4.x, vy4.y, vy4.z, vy4.w};
        const float vcz[4] = {vz4.x, vz4.y, vz4.z, vz4.w};
        const float xv[4] = {x4.x, x4.y, x4.z, x4.w};
        float av[4];
        afactor4<AF>(e, a.A, base, nullptr, nullptr, nullptr, nullptr, av);
        float tcv[4], dl[4];
        cone_tc4<SIG, DSIG>(e, vcx, vcy, vcz, a.cc, tcv, dl);
        float tq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float t = tcv[k] * av[k];
          tq[k] = t;
          S += t;
          D = fmaf(t, xv[k], D);
          if (DSIG) {
            const float dt = t * dl[k];
            dS += dt;
            dD = fmaf(dt, xv[k], dD);
          }
        }
        if constexpr (TRL) {
          cp_rows[c * BLOCK + tid] = make_float4(tq[0], tq[1], tq[2], tq[3]);
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) tr[TRL ? 0 : c][k] = tq[k];
        }
      } else {
        if constexpr (TRL) {
          cp_rows[c * BLOCK + tid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) tr[TRL ? 0 : c][k] = 0.0f;
        }
      }
    }
    // wave sums in FP32 by DPP (64 partials of <= 32 values each), across waves in FP64
    double vS = (double)jscc_wave_sum_dpp(S);
    double vD = (double)jscc_wave_sum_dpp(D);
    double vdS = DSIG ? (double)jscc_wave_sum_dpp(dS) : 0.0;
    double vdD = DSIG ? (double)jscc_wave_sum_dpp(dD) : 0.0;
    if (jscc_lane() == 0) {
      red[buf][0][jscc_wave_id()] = vS;
      red[buf][1][jscc_wave_id()] = vD;
      red[buf][2][jscc_wave_id()] = vdS;
      red[buf][3][jscc_wave_id()] = vdD;
    }
    __syncthreads();
    double tS = 0.0, tD = 0.0, tdS = 0.0, tdD = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      tS += red[buf][0][w];
      tD += red[buf][1][w];
      if (DSIG) {
        tdS += red[buf][2][w];
        tdD += red[buf][3][w];
      }
    }
    buf ^= 1;
    const double inv = a.normalize ? 1.0 / tS : 1.0;
    const double f = tD * inv;
    if (tid == 0) {
      if (MODE == CP_FWD) a.fwd_out[i] = (float)f;
      if (DSIG && a.dsig_out)
        a.dsig_out[i] = (float)(a.normalize ? (tdD * tS - tD * tdS) * inv * inv : tdD);
    }
    if (MODE != CP_FWD) {
      float coef;
      if (MODE == CP_RATIO) {
        const double den = a.use_eps ? (f > (double)a.eps ? f : (double)a.eps) : f;
        coef = (float)(inv / den);
      } else {
        coef = (float)((double)a.coef[i] * inv);
      }
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        float tq[4];
        if constexpr (TRL) {
          const float4 v = cp_rows[c * BLOCK + tid];
          tq[0] = v.x; tq[1] = v.y; tq[2] = v.z; tq[3] = v.w;
        } else {
#pragma unroll
          for (int k = 0; k < 4; ++k) tq[k] = tr[TRL ? 0 : c][k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) wr[c][k] = fmaf(coef, tq[k], wr[c][k]);
      }
    }
  }
  if (MODE != CP_FWD) {
    float* dst = a.partial + (int64_t)blockIdx.x * a.P;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int base = (c * BLOCK + tid) * 4;
      if (base < a.P)
        *reinterpret_cast<float4*>(dst + base) = make_float4(wr[c][0], wr[c][1], wr[c][2], wr[c][3]);
    }
  }
}

// dynamic LDS of a launch (the row values of cp_row_in_lds shapes), with the
// attribute that allows more than 64 KB set once per kernel
template <typename K>
int lds_launch_attr(K* fn, size_t lds, bool* done) {
  if (lds > 65536 && !*done) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds));
    *done = true;
  }
  return 0;
}

template <int SIG, int AF, int BLOCK, int NCH, int MODE, bool DSIG>
int launch_cp_one(const ConeProjArgs& a, int G, hipStream_t s) {
  constexpr size_t lds = cp_row_in_lds<BLOCK, NCH>() ? (size_t)NCH * BLOCK * sizeof(float4) : 0;
  static bool attr = false;
  const int rc = lds_launch_attr(&k_cone_proj<SIG, AF, BLOCK, NCH, MODE, DSIG>, lds, &attr);
  if (rc) return rc;
  hipLaunchKernelGGL((k_cone_proj<SIG, AF, BLOCK, NCH, MODE, DSIG>), dim3(G), dim3(BLOCK), lds, s, a);
  return 0;
}

template <int SIG, int AF, int BLOCK, int NCH, bool DSIG>
int launch_cp_mode(int mode, const ConeProjArgs& a, int G, hipStream_t s) {
  if (mode == CP_RATIO) return launch_cp_one<SIG, AF, BLOCK, NCH, CP_RATIO, DSIG>(a, G, s);
  if (mode == CP_FWD) return launch_cp_one<SIG, AF, BLOCK, NCH, CP_FWD, DSIG>(a, G, s);
  return launch_cp_one<SIG, AF, BLOCK, NCH, CP_COEF, DSIG>(a, G, s);
}

template <int SIG, int AF, bool DSIG>
int launch_cp_size(int mode, const ConeProjArgs& a, int G, hipStream_t s) {
  // the row lives in registers (4 * NCH voxels per thread), or in LDS above 16384 voxels
  if (a.P <= 4 * 256 * 4) return launch_cp_mode<SIG, AF, 256, 4, DSIG>(mode, a, G, s);
  if (a.P <= 4 * 512 * 4) return launch_cp_mode<SIG, AF, 512, 4, DSIG>(mode, a, G, s);
  if (a.P <= 4 * 1024 * 4) return launch_cp_mode<SIG, AF, 1024, 4, DSIG>(mode, a, G, s);
  if (a.P <= 4 * 1024 * 8) return launch_cp_mode<SIG, AF, 1024, 8, DSIG>(mode, a, G, s);
  jscc_set_error("jscc_cone_proj: %d voxels exceed the row limit (32768)", a.P);
  return -3;
}

}  // namespace

extern "C" int jscc_cone_proj_grid(int64_t n_rows) {
  int64_t g = 4 * (int64_t)jscc_num_cus();
  if (g > n_rows) g = n_rows;
  return (int)(g < 1 ? 1 : g);
}

extern "C" int64_t jscc_cone_proj_workspace_bytes(int64_t n_rows, int64_t n_vox) {
  return jscc_partials_bytes(jscc_cone_proj_grid(n_rows), (n_vox + 3) / 4 * 4);
}

extern "C" int jscc_cone_proj(int mode, const void* events, const int32_t* ev_idx, int64_t n_rows,
                              const float* coor_x, const float* coor_y, const float* coor_z, int64_t n_vox,
                              const float* sysmat, int64_t ld_sysmat, int sigma_model, double fixed_sigma,
                              const float* x, const float* coef, float eps, int use_eps, int normalize,
                              float* fwd_out, float* dsig_out, float out_scale, int accumulate, float* w,
                              int64_t w_len, void* workspace, int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(mode >= 0 && mode <= 2, "jscc_cone_proj: bad mode");
  // the kernels run on whole voxel quads: P4 = n_vox rounded up to 4 (image, factor rows and partial sums
  // extend that far); a ragged count needs a factor row that is zero past n_vox, so the padded voxels of the
  // last quad carry zero weight
  JSCC_REQUIRE(n_rows >= 0 && n_rows < ((int64_t)1 << 31) && n_vox > 0, "jscc_cone_proj: bad shape");
  JSCC_REQUIRE(n_vox % 4 == 0 || sysmat != nullptr,
               "jscc_cone_proj: %lld voxels (not a multiple of 4) need a factor row that is zero past them",
               (long long)n_vox);
  const int64_t P4 = (n_vox + 3) / 4 * 4;
  JSCC_REQUIRE(sigma_model == SIG_STRICT || sigma_model == SIG_FIXED, "jscc_cone_proj: bad sigma model");
  JSCC_REQUIRE(dsig_out == nullptr || sigma_model == SIG_FIXED,
               "jscc_cone_proj: d/d sigma exists for the fixed-sigma model only");
  JSCC_REQUIRE(events && ev_idx && coor_x && coor_y && coor_z && x && err, "jscc_cone_proj: null pointer");
  JSCC_REQUIRE(((uintptr_t)coor_x | (uintptr_t)coor_y | (uintptr_t)coor_z | (uintptr_t)x) % 16 == 0,
               "jscc_cone_proj: coordinates and image must be 16-byte aligned");
  if (sysmat)
    JSCC_REQUIRE(ld_sysmat % 4 == 0 && (ld_sysmat == 0 || ld_sysmat >= P4) && (uintptr_t)sysmat % 16 == 0,
                 "jscc_cone_proj: sysmat alignment");
  if (mode == CP_FWD) JSCC_REQUIRE(fwd_out != nullptr, "jscc_cone_proj: null forward output");
  if (mode == CP_COEF) JSCC_REQUIRE(coef != nullptr, "jscc_cone_proj: null coefficients");
  if (mode != CP_FWD) JSCC_REQUIRE(w != nullptr && w_len == n_vox, "jscc_cone_proj: weight must have n_vox entries");
  if (sigma_model == SIG_FIXED) JSCC_REQUIRE(fixed_sigma > 0.0, "jscc_cone_proj: fixed sigma must be positive");
  hipStream_t s = (hipStream_t)stream;
  if (n_rows == 0) {
    if (mode != CP_FWD && !accumulate) JSCC_CHECK_HIP(hipMemsetAsync(w, 0, w_len * sizeof(float), s));
    return 0;
  }
  const int G = jscc_cone_proj_grid(n_rows);
  if (mode != CP_FWD)
    JSCC_REQUIRE(workspace && workspace_bytes >= jscc_partials_bytes(G, P4), "jscc_cone_proj: workspace too small");
  ConeProjArgs a;
  a.ev = (const JsccEvent*)events; a.ev_idx = ev_idx; a.n_rows = n_rows;
  a.cx = coor_x; a.cy = coor_y; a.cz = coor_z; a.P = (int)P4; a.A = sysmat; a.lda = ld_sysmat;
  a.cc = make_consts(sigma_model, fixed_sigma);
  a.x = x; a.coef = coef; a.eps = eps; a.use_eps = use_eps; a.normalize = normalize;
  a.fwd_out = fwd_out; a.dsig_out = dsig_out; a.partial = (float*)workspace;
  const bool dsig = dsig_out != nullptr;
  int rc;
  jscc_prof_begin(JSCC_TAG_CONE_PROJ, s);
  if (sigma_model == SIG_STRICT) {
    rc = sysmat ? launch_cp_size<SIG_STRICT, AF_ROW, false>(mode, a, G, s)
                : launch_cp_size<SIG_STRICT, AF_NONE, false>(mode, a, G, s);
  } else if (dsig) {
    rc = sysmat ? launch_cp_size<SIG_FIXED, AF_ROW, true>(mode, a, G, s)
                : launch_cp_size<SIG_FIXED, AF_NONE, true>(mode, a, G, s);
  } else {
    rc = sysmat ? launch_cp_size<SIG_FIXED, AF_ROW, false>(mode, a, G, s)
                : launch_cp_size<SIG_FIXED, AF_NONE, false>(mode, a, G, s);
  }
  jscc_prof_end(JSCC_TAG_CONE_PROJ, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  if (mode == CP_FWD) return 0;
  const JsccEm none = {nullptr, nullptr, 0, 0.0f, 0, 1.0f, nullptr, 0.0f};
  return jscc_reduce_partials((float*)workspace, G, (int)n_vox, P4, nullptr, 0, out_scale, accumulate, 0, w, w_len,
                              err, s,
                              none);
}

// ============================================================================
// Batched no-store projection: many reconstructions (each its own image and
// event list) in one launch.  The rows of all reconstructions for the current
// subset are concatenated reconstruction by reconstruction (row_off[b] ..
// row_off[b+1]); the work-groups split the concatenation evenly, so lists of
// very different lengths share the chip without idle work-groups.  A
// work-group flushes its accumulators whenever it crosses into the next
// reconstruction, into partial slot g + b (unique, since both the work-group
// ranges and the reconstructions are contiguous and ordered).  The column
// reduction then sums, for reconstruction b, the slots of the work-groups
// whose range meets [row_off[b], row_off[b+1]) and applies either the EM
// update of image b (OSEM step) or writes the weight (column sums).
// ============================================================================
namespace {

enum { BP_OSEM = 0, BP_SUM = 1 };

struct ConeBatchArgs {
  const JsccEvent* ev;
  const int32_t* ev_idx;
  int64_t n_rows;
  const int64_t* row_off;  // [n_recon + 1]
  int n_recon;
  const float* cx;
  const float* cy;
  const float* cz;
  int P;
  const float* A;
  int64_t lda;
  ConeConsts cc;
  const float* x;  // [n_recon, ldx]
  int64_t ldx;
  float eps;
  int use_eps;
  int normalize;
  float* partial;  // [G + n_recon - 1, P]
};

// rows of work-group g under the even split of n rows over G groups
__device__ __forceinline__ void wg_rows(int64_t n, int G, int g, int64_t* b, int64_t* e) {
  const int64_t per = n / G, rem = n % G;
  *b = g * per + (g < rem ? g : rem);
  *e = *b + per + (g < rem ? 1 : 0);
}

__device__ __forceinline__ int wg_of_row(int64_t n, int G, int64_t r) {
  const int64_t per = n / G, rem = n % G;
  const int64_t big = rem * (per + 1);
  if (r < big) return (int)(r / (per + 1));
  return (int)(rem + (r - big) / per);
}

// Steps of 8 or more quads per thread (R rows x NCH quads at 1024 threads)
// keep the row values in the thread's own LDS slots: in registers next to the
// accumulators they would spill.
template <int BLOCK, int NCH, int R>
constexpr bool bp_rows_in_lds() { return BLOCK >= 1024 && NCH * R >= 8; }

// R rows per work-group step: the rows are evaluated into registers, their
// sums reduced by DPP within each wave, and one barrier serves all R rows;
// lane q of every wave then finishes row q (cross-wave FP64 sum and
// coefficient), which is read back as a scalar.
// (the pointers are __restrict__ parameters so that the wave-uniform event
// records and offsets are read with scalar loads into scalar registers)
template <int SIG, int AF, int BLOCK, int NCH, int R, int MODE>
__global__ void __launch_bounds__(BLOCK)
    k_cone_proj_batched(const JsccEvent* __restrict__ ev, const int32_t* __restrict__ ev_idx,
                        const int64_t* __restrict__ row_off, const float* __restrict__ cx,
                        const float* __restrict__ cy, const float* __restrict__ cz, const float* __restrict__ A,
                        const float* __restrict__ x, float* __restrict__ partial, ConeBatchArgs a) {
  a.ev = ev; a.ev_idx = ev_idx; a.row_off = row_off; a.cx = cx; a.cy = cy; a.cz = cz; a.A = A; a.x = x;
  a.partial = partial;
  constexpr int NW = BLOCK / JSCC_WAVE;
  constexpr bool TRL = bp_rows_in_lds<BLOCK, NCH, R>();
  extern __shared__ float4 bp_rows[];  // TRL: slot (q * NCH + c) * BLOCK + tid
  const int tid = threadIdx.x;
  const int lane = jscc_lane();
  const int g = blockIdx.x;
  int64_t r_begin, r_end;
  wg_rows(a.n_rows, gridDim.x, g, &r_begin, &r_end);
  if (r_begin >= r_end) return;
  // reconstruction of the first row: last b with row_off[b] <= r_begin
  int lo = 0, hi = a.n_recon;  // invariant row_off[lo] <= r_begin < row_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.row_off[mid] <= r_begin) lo = mid; else hi = mid;
  }
  int b = lo;
  __shared__ float red[2][2][R][NW];
  int buf = 0;
  float wr[NCH][4];
  int64_t i = r_begin;
  while (i < r_end) {
    while (a.row_off[b + 1] <= i) ++b;  // skip reconstructions without rows here
    const int64_t seg_end = a.row_off[b + 1] < r_end ? a.row_off[b + 1] : r_end;
    const float* xb = MODE == BP_OSEM ? a.x + (int64_t)b * a.ldx : nullptr;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) wr[c][k] = 0.0f;
    for (; i < seg_end; i += R) {
      float tr[TRL ? 1 : R][TRL ? 1 : NCH][4];
      float S[R], D[R];
      ConeEv e[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int64_t iq = i + q < seg_end ? i + q : i;  // a slot past the segment repeats row i (weight 0)
        e[q] = load_event(a.ev[a.ev_idx[iq]], a.lda);
        S[q] = 0.0f;
        D[q] = 0.0f;
      }
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        int base = (c * BLOCK + tid) * 4;
        // reloaded for every step (L2 hits) instead of hoisted out of the row loop, which
        // would pin 12 registers per chunk for the whole kernel
        asm volatile("" : "+v"(base));
        if (base < a.P) {
          const float4 vx4 = *reinterpret_cast<const float4*>(a.cx + base);
          const float4 vy4 = *reinterpret_cast<const float4*>(a.cy + base);
          const float4 vz4 = *reinterpret_cast<const float4*>(a.cz + base);
          const float vcx[4] = {vx4.x, vx4.y, vx4.z, vx4.w};
          const float vcy[4] = {vy4.x, vy4.y, vy4.z, vy4.w};
          const float vcz[4] = {vz4.x, vz4.y, vz4.z, vz4.w};
          float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
          if (MODE == BP_OSEM) {
            const float4 x4 = *reinterpret_cast<const float4*>(xb + base);
            xv[0] = x4.x; xv[1] = x4.y; xv[2] = x4.z; xv[3] = x4.w;
          }
#pragma unroll
          for (int q = 0; q < R; ++q) {
            float av[4];
            afactor4<AF>(e[q], a.A, base, nullptr, nullptr, nullptr, nullptr, av);
            float tcv[4];
            cone_tc4<SIG, false>(e[q], vcx, vcy, vcz, a.cc, tcv, nullptr);
            float tq[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const float t = tcv[k] * av[k];
              tq[k] = t;
              S[q] += t;
              if (MODE == BP_OSEM) D[q] = fmaf(t, xv[k], D[q]);
            }
            if constexpr (TRL) {
              bp_rows[(q * NCH + c) * BLOCK + tid] = make_float4(tq[0], tq[1], tq[2], tq[3]);
            } else {
#pragma unroll
              for (int k = 0; k < 4; ++k) tr[q][c][k] = tq[k];
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < R; ++q) {
            if constexpr (TRL) {
              bp_rows[(q * NCH + c) * BLOCK + tid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            } else {
#pragma unroll
              for (int k = 0; k < 4; ++k) tr[q][c][k] = 0.0f;
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < R; ++q) {
        S[q] = jscc_wave_sum_dpp(S[q]);
        if (MODE == BP_OSEM) D[q] = jscc_wave_sum_dpp(D[q]);
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          red[buf][0][q][jscc_wave_id()] = S[q];
          red[buf][1][q][jscc_wave_id()] = D[q];
        }
      }
      __syncthreads();
      float cl = 0.0f;
      if (lane < R) {
        double tS = 0.0, tD = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          tS += (double)red[buf][0][lane][w];
          if (MODE == BP_OSEM) tD += (double)red[buf][1][lane][w];
        }
        const double inv = a.normalize ? 1.0 / tS : 1.0;
        if (MODE == BP_OSEM) {
          const double f = tD * inv;
          const double den = a.use_eps ? (f > (double)a.eps ? f : (double)a.eps) : f;
          cl = (float)(inv / den);
        } else {
          cl = (float)inv;
        }
        if (i + lane >= seg_end) cl = 0.0f;
      }
      buf ^= 1;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const float coef = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cl), q));
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          float tq[4];
          if constexpr (TRL) {
            const float4 v = bp_rows[(q * NCH + c) * BLOCK + tid];
            tq[0] = v.x; tq[1] = v.y; tq[2] = v.z; tq[3] = v.w;
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) tq[k] = tr[q][c][k];
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) wr[c][k] = fmaf(coef, tq[k], wr[c][k]);
        }
      }
    }
    float* dst = a.partial + (int64_t)(g + b) * a.P;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int base = (c * BLOCK + tid) * 4;
      if (base < a.P) *reinterpret_cast<float4*>(dst + base) = make_float4(wr[c][0], wr[c][1], wr[c][2], wr[c][3]);
    }
    i = seg_end;
  }
}

// Column reduction per reconstruction (blockIdx.y = b), 64 columns per block:
// 16 groups of 16 threads sum every 16th contributing slot, then combine in
// LDS in a fixed order.  OSEM: x_b <- em(x_b, w, s_b); SUM: w_out_b <- w.
// P: the partial row stride (whole quads); only the first n_vox columns are written.
__global__ void __launch_bounds__(256) k_reduce_batched(const float* __restrict__ part, int G, int64_t n_rows,
                                                        const int64_t* __restrict__ row_off, int P, int n_vox,
                                                        int mode,
                                                        float* __restrict__ x, int64_t ldx,
                                                        const float* __restrict__ sens, int64_t lds, int em_mode,
                                                        float em_eps, float* __restrict__ w, int64_t ldw) {
  __shared__ float4 acc[16][16];
  const int b = blockIdx.y;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int col = blockIdx.x * 64 + tx * 4;
  const int64_t lo = row_off[b], hi = row_off[b + 1];
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  if (hi > lo && col < P) {
    const int g0 = wg_of_row(n_rows, G, lo), g1 = wg_of_row(n_rows, G, hi - 1);
    for (int gg = g0 + ty; gg <= g1; gg += 16) {
      const float4 v = *reinterpret_cast<const float4*>(part + (int64_t)(gg + b) * P + col);
      sum.x += v.x; sum.y += v.y; sum.z += v.z; sum.w += v.w;
    }
  }
  acc[ty][tx] = sum;
  __syncthreads();
  if (ty == 0 && col < P) {
    float4 t = acc[0][tx];
#pragma unroll
    for (int k = 1; k < 16; ++k) {
      t.x += acc[k][tx].x; t.y += acc[k][tx].y; t.z += acc[k][tx].z; t.w += acc[k][tx].w;
    }
    const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t j = col + k;
      if (j >= n_vox) break;
      if (w) w[(int64_t)b * ldw + j] = tv[k];
      if (mode == BP_OSEM) {
        float* xp = x + (int64_t)b * ldx + j;
        *xp = jscc_em_value(*xp, tv[k], sens[(int64_t)b * lds + j], em_mode, em_eps);
      }
    }
  }
}

template <int SIG, int AF, int BLOCK, int NCH, int R, int MODE>
int launch_bp_one(const ConeBatchArgs& a, int G, hipStream_t s) {
  constexpr size_t lds = bp_rows_in_lds<BLOCK, NCH, R>() ? (size_t)R * NCH * BLOCK * sizeof(float4) : 0;
  static bool attr = false;
  const int rc = lds_launch_attr(&k_cone_proj_batched<SIG, AF, BLOCK, NCH, R, MODE>, lds, &attr);
  if (rc) return rc;
  hipLaunchKernelGGL((k_cone_proj_batched<SIG, AF, BLOCK, NCH, R, MODE>), dim3(G), dim3(BLOCK), lds, s, a.ev,
                     a.ev_idx, a.row_off, a.cx, a.cy, a.cz, a.A, a.x, a.partial, a);
  return 0;
}

template <int SIG, int AF, int BLOCK, int NCH, int R>
int launch_bp_shape(int mode, const ConeBatchArgs& a, int G, hipStream_t s) {
  if (mode == BP_OSEM) return launch_bp_one<SIG, AF, BLOCK, NCH, R, BP_OSEM>(a, G, s);
  return launch_bp_one<SIG, AF, BLOCK, NCH, R, BP_SUM>(a, G, s);
}

// Tile shapes (threads per work-group, voxel quads per thread, rows per step) by voxel count; shape 0 is
// the default, the others are the autotuner's alternatives (every shape covers the row: BLOCK x NCH x 4 >= P).
constexpr int kBpShapes[4][3][3] = {
    {{256, 4, 2}, {256, 4, 1}, {512, 2, 2}},     // P <= 4096
    {{512, 4, 2}, {512, 4, 1}, {1024, 2, 2}},    // P <= 8192
    {{1024, 4, 2}, {1024, 4, 1}, {512, 8, 1}},   // P <= 16384
    {{1024, 8, 1}, {0, 0, 0}, {0, 0, 0}}};       // P <= 32768

int bp_class(int64_t P) { return P <= 4096 ? 0 : (P <= 8192 ? 1 : (P <= 16384 ? 2 : (P <= 32768 ? 3 : -1))); }
int bp_shape_count(int64_t P) { const int c = bp_class(P); return c < 0 ? 0 : (c == 3 ? 1 : 3); }

template <int SIG, int AF>
int launch_bp_size(int mode, const ConeBatchArgs& a, int G, hipStream_t s, int tile) {
  switch (bp_class(a.P) * 3 + tile) {
    case 0: return launch_bp_shape<SIG, AF, 256, 4, 2>(mode, a, G, s);
    case 1: return launch_bp_shape<SIG, AF, 256, 4, 1>(mode, a, G, s);
    case 2: return launch_bp_shape<SIG, AF, 512, 2, 2>(mode, a, G, s);
    case 3: return launch_bp_shape<SIG, AF, 512, 4, 2>(mode, a, G, s);
    case 4: return launch_bp_shape<SIG, AF, 512, 4, 1>(mode, a, G, s);
    case 5: return launch_bp_shape<SIG, AF, 1024, 2, 2>(mode, a, G, s);
    case 6: return launch_bp_shape<SIG, AF, 1024, 4, 2>(mode, a, G, s);
    case 7: return launch_bp_shape<SIG, AF, 1024, 4, 1>(mode, a, G, s);
    case 8: return launch_bp_shape<SIG, AF, 512, 8, 1>(mode, a, G, s);
    case 9: return launch_bp_shape<SIG, AF, 1024, 8, 1>(mode, a, G, s);
    default: break;
  }
  jscc_set_error("jscc_cone_proj_batched: %d voxels exceed the row limit (32768)", a.P);
  return -3;
}

}  // namespace

extern "C" int jscc_cone_batch_tiles(int64_t n_vox) { return bp_shape_count(n_vox); }

extern "C" int jscc_cone_batch_tile_shape(int64_t n_vox, int tile, int* threads, int* quads, int* rows_per_step) {
  JSCC_REQUIRE(threads && quads && rows_per_step, "jscc_cone_batch_tile_shape: null pointer");
  JSCC_REQUIRE(tile >= 0 && tile < bp_shape_count(n_vox), "jscc_cone_batch_tile_shape: tile %d of %d for %lld voxels",
               tile, bp_shape_count(n_vox), (long long)n_vox);
  const int* sh = kBpShapes[bp_class(n_vox)][tile];
  *threads = sh[0];
  *quads = sh[1];
  *rows_per_step = sh[2];
  return 0;
}

extern "C" int jscc_cone_batch_groups(int64_t n_rows, int groups_per_cu) {
  int64_t g = (int64_t)(groups_per_cu > 0 ? groups_per_cu : 4) * jscc_num_cus();
  if (g > n_rows) g = n_rows;
  return (int)(g < 1 ? 1 : g);
}

extern "C" int64_t jscc_cone_batch_workspace_bytes(int64_t n_rows, int n_recon, int64_t n_vox, int groups_per_cu) {
  const int64_t G = jscc_cone_batch_groups(n_rows, groups_per_cu);
  return (G + (n_recon > 0 ? n_recon : 1) - 1) * ((n_vox + 3) / 4 * 4) * (int64_t)sizeof(float);
}

extern "C" int jscc_cone_proj_batched(int mode, const void* events, const int32_t* ev_idx, int64_t n_rows,
                                      const int64_t* row_off, int n_recon, const float* coor_x, const float* coor_y,
                                      const float* coor_z, int64_t n_vox, const float* sysmat, int64_t ld_sysmat,
                                      int sigma_model, double fixed_sigma, float* x, int64_t ldx, float eps,
                                      int use_eps, int normalize, const float* sens, int64_t ld_sens, int em_mode,
                                      float em_eps, float* w, int64_t ldw, int groups_per_cu, int tile,
                                      void* workspace, int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(mode == BP_OSEM || mode == BP_SUM, "jscc_cone_proj_batched: bad mode");
  JSCC_REQUIRE(n_recon >= 1 && n_recon <= 65535, "jscc_cone_proj_batched: bad reconstruction count");
  // whole voxel quads, as jscc_cone_proj: a ragged count needs a factor row that is zero past it
  JSCC_REQUIRE(n_rows >= 0 && n_rows < ((int64_t)1 << 31) && n_vox > 0, "jscc_cone_proj_batched: bad shape");
  JSCC_REQUIRE(n_vox % 4 == 0 || sysmat != nullptr,
               "jscc_cone_proj_batched: %lld voxels (not a multiple of 4) need a factor row that is zero past them",
               (long long)n_vox);
  const int64_t P4 = (n_vox + 3) / 4 * 4;
  JSCC_REQUIRE(sigma_model == SIG_STRICT || sigma_model == SIG_FIXED, "jscc_cone_proj_batched: bad sigma model");
  if (sigma_model == SIG_FIXED) JSCC_REQUIRE(fixed_sigma > 0.0, "jscc_cone_proj_batched: fixed sigma must be positive");
  JSCC_REQUIRE(events && ev_idx && row_off && coor_x && coor_y && coor_z && err, "jscc_cone_proj_batched: null pointer");
  JSCC_REQUIRE(((uintptr_t)coor_x | (uintptr_t)coor_y | (uintptr_t)coor_z) % 16 == 0,
               "jscc_cone_proj_batched: coordinates must be 16-byte aligned");
  if (sysmat)
    JSCC_REQUIRE(ld_sysmat % 4 == 0 && (ld_sysmat == 0 || ld_sysmat >= P4) && (uintptr_t)sysmat % 16 == 0,
                 "jscc_cone_proj_batched: sysmat alignment");
  if (mode == BP_OSEM) {
    JSCC_REQUIRE(x && sens && ldx >= P4 && ld_sens >= n_vox && ldx % 4 == 0 && (uintptr_t)x % 16 == 0,
                 "jscc_cone_proj_batched: images [n_recon, ldx] (16-byte aligned rows) and sensitivities needed");
    JSCC_REQUIRE(em_mode >= 0 && em_mode <= 2, "jscc_cone_proj_batched: bad EM mode");
  } else {
    JSCC_REQUIRE(w != nullptr && ldw >= n_vox, "jscc_cone_proj_batched: null weight output");
  }
  JSCC_REQUIRE(tile >= 0 && tile < (bp_shape_count(n_vox) > 0 ? bp_shape_count(n_vox) : 1),
               "jscc_cone_proj_batched: tile %d of %d for %lld voxels", tile, bp_shape_count(n_vox), (long long)n_vox);
  if (n_rows == 0) return 0;
  const int G = jscc_cone_batch_groups(n_rows, groups_per_cu);
  JSCC_REQUIRE(workspace && workspace_bytes >= jscc_cone_batch_workspace_bytes(n_rows, n_recon, n_vox, groups_per_cu),
               "jscc_cone_proj_batched: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  ConeBatchArgs a;
  a.ev = (const JsccEvent*)events; a.ev_idx = ev_idx; a.n_rows = n_rows; a.row_off = row_off; a.n_recon = n_recon;
  a.cx = coor_x; a.cy = coor_y; a.cz = coor_z; a.P = (int)P4; a.A = sysmat; a.lda = ld_sysmat;
  a.cc = make_consts(sigma_model, fixed_sigma);
  a.x = x; a.ldx = ldx; a.eps = eps; a.use_eps = use_eps; a.normalize = normalize; a.partial = (float*)workspace;
  int rc;
  jscc_prof_begin(JSCC_TAG_CONE_PROJ, s);
  if (sigma_model == SIG_STRICT) {
    rc = sysmat ? launch_bp_size<SIG_STRICT, AF_ROW>(mode, a, G, s, tile)
                : launch_bp_size<SIG_STRICT, AF_NONE>(mode, a, G, s, tile);
  } else {
    rc = sysmat ? launch_bp_size<SIG_FIXED, AF_ROW>(mode, a, G, s, tile)
                : launch_bp_size<SIG_FIXED, AF_NONE>(mode, a, G, s, tile);
  }
  jscc_prof_end(JSCC_TAG_CONE_PROJ, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_reduce_batched, dim3((unsigned)((P4 + 63) / 64), (unsigned)n_recon), dim3(256), 0, s,
                     (const float*)workspace, G, n_rows, row_off, (int)P4, (int)n_vox, mode, x, ldx, sens, ld_sens,
                     em_mode, em_eps, w, ldw);
  JSCC_CHECK_LAUNCH();
  return 0;
}
