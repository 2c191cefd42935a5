// Row utilities around the cone builder and the sparse (coarse-grid) T format.
//
// Reference: process_list_plane_strict.py:137-148, 236-253 (row normalisation,
// returned t / t_compton / t_single); compton_sparse_ops.py:293-356 (row and
// image transforms between the fine and coarse polar grids), :359-400
// (packed rows [cpnum1, t_c...] and their materialisation on the fine grid).
//
// The reference applies its coarse<->fine operators as dense (block-diagonal)
// matmuls; here every operator is a small CSR table (at most 4 taps per output
// pixel for the expand, 2 for the reduce) applied by a gather kernel.
#include "jscc_common.h"

namespace {

constexpr int kRowThreads = 256;

template <typename T>
__device__ __forceinline__ float load_as_f32(const T* p, int64_t i);
template <>
__device__ __forceinline__ float load_as_f32<float>(const float* p, int64_t i) { return p[i]; }
template <>
__device__ __forceinline__ float load_as_f32<uint16_t>(const uint16_t* p, int64_t i) {
  return jscc_bf16_to_f32(p[i]);
}

// dst[i, off + p] = src[idx[i], p] / denom(idx[i]) ; optional dst[i, 0] = lead[idx[i]]
template <typename T>
__global__ void __launch_bounds__(kRowThreads) k_gather_normalize(const T* __restrict__ src, int64_t lds, int64_t P,
                                                                  const int32_t* __restrict__ idx, int64_t n_out,
                                                                  const double* __restrict__ stats, int field,
                                                                  const float* __restrict__ lead,
                                                                  float* __restrict__ dst, int64_t ldd,
                                                                  int64_t col_off) {
  const int64_t i = blockIdx.x;
  if (i >= n_out) return;
  const int64_t r = idx ? idx[i] : i;
  const double den = stats ? stats[r * 4 + field] : 1.0;
  const T* s = src + r * lds;
  float* d = dst + i * ldd;
  if (lead && threadIdx.x == 0) d[0] = lead[r];
  for (int64_t p = threadIdx.x; p < P; p += kRowThreads) {
    d[col_off + p] = (float)((double)load_as_f32<T>(s, p) / den);
  }
}

// t_single: dst[i, :] = A[arow[src_ev[idx[i]]], :] / sum(...)   (process_list_plane_strict.py:237,140)
__global__ void __launch_bounds__(kRowThreads) k_arows_normalize(const float* __restrict__ A, int64_t lda, int64_t P,
                                                                 const int32_t* __restrict__ arow,
                                                                 int64_t n_out, float* __restrict__ dst,
                                                                 int64_t ldd) {
  __shared__ double red[kRowThreads / JSCC_WAVE];
  const int64_t i = blockIdx.x;
  if (i >= n_out) return;
  const float* a = A + (int64_t)arow[i] * lda;
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < P; p += kRowThreads) s += (double)a[p];
  s = jscc_wave_sum(s);
  if (jscc_lane() == 0) red[jscc_wave_id()] = s;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < kRowThreads / JSCC_WAVE; ++w) tot += red[w];
  float* d = dst + i * ldd;
  for (int64_t p = threadIdx.x; p < P; p += kRowThreads) d[p] = (float)((double)a[p] / tot);
}

// Low-precision rows with a per-row scale: stored = code(t / q) with
// q = max / TOP (TOP = 448 for e4m3fn, 2^14 for IEEE half), row_scale = q / S
// so that row_scale * stored ~ t / S.  Without statistics (a system-matrix
// row): max = the row's largest |value|, S = 1.
template <typename T, int OUT>
__global__ void __launch_bounds__(kRowThreads) k_quantize(const T* __restrict__ src, int64_t lds, int64_t P,
                                                          int64_t n, const double* __restrict__ stats,
                                                          void* __restrict__ dst, int64_t ldd,
                                                          float* __restrict__ row_scale) {
  __shared__ float red[kRowThreads / JSCC_WAVE];
  const int64_t r = blockIdx.x;
  if (r >= n) return;
  const T* s = src + r * lds;
  double S = 1.0, mx;
  if (stats) {
    S = stats[r * 4 + 0];
    mx = stats[r * 4 + 3];
  } else {
    float m = 0.0f;
    for (int64_t p = threadIdx.x; p < P; p += kRowThreads) m = fmaxf(m, fabsf(load_as_f32<T>(s, p)));
    m = jscc_wave_max(m);
    if (jscc_lane() == 0) red[jscc_wave_id()] = m;
    __syncthreads();
    m = red[0];
    for (int w = 1; w < kRowThreads / JSCC_WAVE; ++w) m = fmaxf(m, red[w]);
    mx = m;
  }
  constexpr double kTop = OUT == JSCC_DT_FP8 ? 448.0 : 16384.0;
  const double q = mx > 0.0 ? mx / kTop : 1.0;
  const float inv_q = (float)(1.0 / q);
  for (int64_t p = threadIdx.x; p < ldd; p += kRowThreads) {
    const float v = p < P ? load_as_f32<T>(s, p) * inv_q : 0.0f;
    if (OUT == JSCC_DT_FP8) {
      reinterpret_cast<uint8_t*>(dst)[r * ldd + p] = (uint8_t)jscc_f32_to_fp8(v);
    } else {
      reinterpret_cast<_Float16*>(dst)[r * ldd + p] = (_Float16)v;  // round to nearest even
    }
  }
  if (threadIdx.x == 0) row_scale[r] = (float)(q / S);
}

// 8 consecutive source values (16-byte aligned) as FP32
template <typename T>
__device__ __forceinline__ void load8_as_f32(const T* p, float* v);
template <>
__device__ __forceinline__ void load8_as_f32<float>(const float* p, float* v) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
template <>
__device__ __forceinline__ void load8_as_f32<uint16_t>(const uint16_t* p, float* v) {
  const uint4 a = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(w[k] << 16);
    v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// k_quantize with 8 values per thread and step: 16- or 32-byte loads, 8- (FP8)
// or 16-byte (FP16) stores.  Needs 16-byte aligned rows on both sides
// (checked by the launcher); values past P are written as zeros, as there.
template <typename T, int OUT>
__global__ void __launch_bounds__(kRowThreads) k_quantize_v8(const T* __restrict__ src, int64_t lds, int64_t P,
                                                             int64_t n, const double* __restrict__ stats,
                                                             void* __restrict__ dst, int64_t ldd,
                                                             float* __restrict__ row_scale) {
  __shared__ float red[kRowThreads / JSCC_WAVE];
  const int64_t r = blockIdx.x;
  if (r >= n) return;
  const T* s = src + r * lds;
  double S = 1.0, mx;
  if (stats) {
    S = stats[r * 4 + 0];
    mx = stats[r * 4 + 3];
  } else {
    float m = 0.0f;
    for (int64_t p = threadIdx.x * 8; p < P; p += kRowThreads * 8) {
      float v[8];
      load8_as_f32<T>(s + p, v);
#pragma unroll
      for (int k = 0; k < 8; ++k) m = (p + k < P) ? fmaxf(m, fabsf(v[k])) : m;
    }
    m = jscc_wave_max(m);
    if (jscc_lane() == 0) red[jscc_wave_id()] = m;
    __syncthreads();
    m = red[0];
    for (int w = 1; w < kRowThreads / JSCC_WAVE; ++w) m = fmaxf(m, red[w]);
    mx = m;
  }
  constexpr double kTop = OUT == JSCC_DT_FP8 ? 448.0 : 16384.0;
  const double q = mx > 0.0 ? mx / kTop : 1.0;
  const float inv_q = (float)(1.0 / q);
#pragma unroll 2
  for (int64_t p = threadIdx.x * 8; p < ldd; p += kRowThreads * 8) {
    // values past P are zeroed before the scaling, so that the product feeds
    // the conversion directly: for FP16 both round in one step
    // (v_fma_mix), exactly as in k_quantize
    float x[8], v[8];
    if (p < P) {
      load8_as_f32<T>(s + p, x);
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = p + k < P ? x[k] : 0.0f;
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = x[k] * inv_q;
    if (OUT == JSCC_DT_FP8) {
      uint2 o;
      o.x = jscc_f32_to_fp8(v[0]) | (jscc_f32_to_fp8(v[1]) << 8) | (jscc_f32_to_fp8(v[2]) << 16) |
            (jscc_f32_to_fp8(v[3]) << 24);
      o.y = jscc_f32_to_fp8(v[4]) | (jscc_f32_to_fp8(v[5]) << 8) | (jscc_f32_to_fp8(v[6]) << 16) |
            (jscc_f32_to_fp8(v[7]) << 24);
      *reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(dst) + r * ldd + p) = o;
    } else {
      uint4 o;
      uint32_t* ow = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
      for (int k = 0; k < 4; ++k)  // round to nearest even
        ow[k] = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v[2 * k]) |
                ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v[2 * k + 1]) << 16);
      *reinterpret_cast<uint4*>(reinterpret_cast<_Float16*>(dst) + r * ldd + p) = o;
    }
  }
  if (threadIdx.x == 0) row_scale[r] = (float)(q / S);
}

template <int OUT>
int launch_quantize(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, int64_t n_rows,
                    const double* stats, void* dst, int64_t ld_dst, float* row_scale, hipStream_t s) {
  // 8 values per thread and step when both row layouts keep 16-byte alignment
  const bool v8 = ld_src % 8 == 0 && ld_dst % 8 == 0 && ((uintptr_t)src % 16) == 0 && ((uintptr_t)dst % 16) == 0;
  if (v8 && src_dtype == JSCC_DT_F32) {
    hipLaunchKernelGGL((k_quantize_v8<float, OUT>), dim3((unsigned)n_rows), dim3(kRowThreads), 0, s,
                       (const float*)src, ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale);
  } else if (v8 && src_dtype == JSCC_DT_BF16) {
    hipLaunchKernelGGL((k_quantize_v8<uint16_t, OUT>), dim3((unsigned)n_rows), dim3(kRowThreads), 0, s,
                       (const uint16_t*)src, ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale);
  } else if (src_dtype == JSCC_DT_F32) {
    hipLaunchKernelGGL((k_quantize<float, OUT>), dim3((unsigned)n_rows), dim3(kRowThreads), 0, s, (const float*)src,
                       ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale);
  } else if (src_dtype == JSCC_DT_BF16) {
    hipLaunchKernelGGL((k_quantize<uint16_t, OUT>), dim3((unsigned)n_rows), dim3(kRowThreads), 0, s,
                       (const uint16_t*)src, ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale);
  } else {
    jscc_set_error("jscc_quantize_rows: unsupported source dtype %d", src_dtype);
    return -2;
  }
  JSCC_CHECK_LAUNCH();
  return 0;
}

// out[r, j] = sum_k val[k] * in[r, col[k]], k in [ptr[j], ptr[j+1])   (row-wise sparse linear map)
__global__ void __launch_bounds__(kRowThreads) k_rows_csr_map(const float* __restrict__ in, int64_t ldi,
                                                              int64_t n_rows, const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ col,
                                                              const float* __restrict__ val, int64_t n_out,
                                                              float* __restrict__ out, int64_t ldo) {
  const int64_t r = blockIdx.y;
  const int64_t j = (int64_t)blockIdx.x * kRowThreads + threadIdx.x;
  if (r >= n_rows || j >= n_out) return;
  const float* a = in + r * ldi;
  float s = 0.0f;
  for (int k = ptr[j]; k < ptr[j + 1]; ++k) s = fmaf(val[k], a[col[k]], s);
  out[r * ldo + j] = s;
}

__device__ __forceinline__ float zero_nonfinite(float v) {
  return (v == v && fabsf(v) != __builtin_inff()) ? v : 0.0f;
}

// compton_sparse_ops.py:380-400 : t_fine = nan_to_num(expand(t_c) * A[cp]) ; row-normalise ; valid
// Rows are written normalised (in place, by the same threads) together with a flag.
// A non-finite coarse value makes the whole fine row non-finite in the
// reference's dense expansion (0 * inf = NaN): the row is written as zeros and
// flagged invalid (normalize = 1) or written as NaN (normalize = 0, the
// unnormalised rows of the grouped weight, whose products are not filtered).
__global__ void __launch_bounds__(kRowThreads) k_materialize(const float* __restrict__ packed, int64_t ldp,
                                                             int64_t n, int64_t Pc, const int32_t* __restrict__ ptr,
                                                             const int32_t* __restrict__ col,
                                                             const float* __restrict__ val, int64_t Pf,
                                                             const float* __restrict__ A, int64_t lda, int64_t M,
                                                             float* __restrict__ out, int64_t ldo,
                                                             int normalize, int32_t* __restrict__ flags) {
  __shared__ double red[kRowThreads / JSCC_WAVE];
  const int64_t r = blockIdx.x;
  if (r >= n) return;
  const float* row = packed + r * ldp;
  int64_t cp = (int64_t)rintf(row[0]) - 1;
  cp = cp < 0 ? 0 : (cp > M - 1 ? M - 1 : cp);
  const float* tc = row + 1;
  const float* a = A + cp * lda;
  float* o = out + r * ldo;
  int bad = 0;
  for (int64_t k = threadIdx.x; k < Pc; k += kRowThreads) {
    const float v = tc[k];
    bad |= (v == v && fabsf(v) != __builtin_inff()) ? 0 : 1;
  }
  bad = __syncthreads_or(bad);
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < Pf; p += kRowThreads) {
    float u = 0.0f;
    for (int k = ptr[p]; k < ptr[p + 1]; ++k) u = fmaf(val[k], tc[col[k]], u);
    float t = u * a[p];
    if (bad) t = normalize ? 0.0f : __builtin_nanf("");
    else if (normalize) t = zero_nonfinite(t);
    o[p] = t;
    s += (double)t;
  }
  s = jscc_wave_sum(s);
  if (jscc_lane() == 0) red[jscc_wave_id()] = s;
  __syncthreads();
  double tot = 0.0;
  for (int w = 0; w < kRowThreads / JSCC_WAVE; ++w) tot += red[w];
  const bool valid = !bad && (tot == tot) && fabs(tot) != __builtin_inf() && tot > 0.0;
  if (valid && normalize) {
    for (int64_t p = threadIdx.x; p < Pf; p += kRowThreads) o[p] = (float)((double)o[p] / tot);
  }
  if (threadIdx.x == 0) flags[r] = valid ? 1 : 0;
}

}  // namespace

extern "C" int jscc_rows_gather_normalize(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols,
                                          const int32_t* idx, int64_t n_out, const double* stats, int stat_field,
                                          const float* lead, float* dst, int64_t ld_dst, int64_t col_off,
                                          void* stream) {
  JSCC_REQUIRE(n_out >= 0 && n_cols >= 0, "jscc_rows_gather_normalize: bad shape");
  if (n_out == 0 || n_cols == 0) return 0;
  JSCC_REQUIRE(src && dst, "jscc_rows_gather_normalize: null pointer");
  JSCC_REQUIRE(stat_field >= 0 && stat_field < 4, "jscc_rows_gather_normalize: bad stat field");
  hipStream_t s = (hipStream_t)stream;
  if (src_dtype == JSCC_DT_F32) {
    hipLaunchKernelGGL(k_gather_normalize<float>, dim3((unsigned)n_out), dim3(kRowThreads), 0, s, (const float*)src,
                       ld_src, n_cols, idx, n_out, stats, stat_field, lead, dst, ld_dst, col_off);
  } else if (src_dtype == JSCC_DT_BF16) {
    hipLaunchKernelGGL(k_gather_normalize<uint16_t>, dim3((unsigned)n_out), dim3(kRowThreads), 0, s,
                       (const uint16_t*)src, ld_src, n_cols, idx, n_out, stats, stat_field, lead, dst, ld_dst,
                       col_off);
  } else {
    jscc_set_error("jscc_rows_gather_normalize: unsupported dtype %d", src_dtype);
    return -2;
  }
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_arows_normalize(const float* A, int64_t lda, int64_t n_cols, const int32_t* arow, int64_t n_out,
                                    float* dst, int64_t ld_dst, void* stream) {
  JSCC_REQUIRE(n_out >= 0, "jscc_arows_normalize: bad shape");
  if (n_out == 0) return 0;
  JSCC_REQUIRE(A && arow && dst, "jscc_arows_normalize: null pointer");
  hipLaunchKernelGGL(k_arows_normalize, dim3((unsigned)n_out), dim3(kRowThreads), 0, (hipStream_t)stream, A, lda,
                     n_cols, arow, n_out, dst, ld_dst);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_quantize_rows(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, int64_t n_rows,
                                  const double* stats, void* dst, int dst_dtype, int64_t ld_dst, float* row_scale,
                                  void* stream) {
  JSCC_REQUIRE(n_rows >= 0 && ld_dst >= n_cols && ld_src >= n_cols, "jscc_quantize_rows: bad shape");
  if (n_rows == 0) return 0;
  JSCC_REQUIRE(src && dst && row_scale, "jscc_quantize_rows: null pointer");
  JSCC_REQUIRE(n_rows <= 0x7fffffffLL, "jscc_quantize_rows: too many rows");
  hipStream_t s = (hipStream_t)stream;
  if (dst_dtype == JSCC_DT_FP8)
    return launch_quantize<JSCC_DT_FP8>(src, src_dtype, ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale, s);
  if (dst_dtype == JSCC_DT_FP16)
    return launch_quantize<JSCC_DT_FP16>(src, src_dtype, ld_src, n_cols, n_rows, stats, dst, ld_dst, row_scale, s);
  jscc_set_error("jscc_quantize_rows: unsupported storage dtype %d", dst_dtype);
  return -2;
}

extern "C" int jscc_quantize_fp8_rows(const void* src, int src_dtype, int64_t ld_src, int64_t n_cols, int64_t n_rows,
                                      const double* stats, uint8_t* dst, int64_t ld_dst, float* row_scale,
                                      void* stream) {
  return jscc_quantize_rows(src, src_dtype, ld_src, n_cols, n_rows, stats, dst, JSCC_DT_FP8, ld_dst, row_scale,
                            stream);
}

extern "C" int jscc_rows_csr_map(const float* in, int64_t ld_in, int64_t n_rows, const int32_t* ptr,
                                 const int32_t* col, const float* val, int64_t n_out, float* out, int64_t ld_out,
                                 void* stream) {
  JSCC_REQUIRE(n_rows >= 0 && n_out >= 0 && n_rows < 65536 * 32768LL, "jscc_rows_csr_map: bad shape");
  if (n_rows == 0 || n_out == 0) return 0;
  JSCC_REQUIRE(in && ptr && col && val && out, "jscc_rows_csr_map: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const int64_t chunk = 65535;
  for (int64_t r0 = 0; r0 < n_rows; r0 += chunk) {
    const int64_t nr = (n_rows - r0) < chunk ? (n_rows - r0) : chunk;
    hipLaunchKernelGGL(k_rows_csr_map, dim3((unsigned)((n_out + kRowThreads - 1) / kRowThreads), (unsigned)nr),
                       dim3(kRowThreads), 0, s, in + r0 * ld_in, ld_in, nr, ptr, col, val, n_out, out + r0 * ld_out,
                       ld_out);
    JSCC_CHECK_LAUNCH();
  }
  return 0;
}

extern "C" int jscc_materialize_rows(const float* packed, int64_t ld_packed, int64_t n_rows, int64_t n_coarse,
                                     const int32_t* ptr, const int32_t* col, const float* val, int64_t n_fine,
                                     const float* A, int64_t lda, int64_t M, float* out, int64_t ld_out,
                                     int normalize, int32_t* flags, void* stream) {
  JSCC_REQUIRE(n_rows >= 0 && ld_packed >= n_coarse + 1 && M > 0, "jscc_materialize_rows: bad shape");
  if (n_rows == 0) return 0;
  JSCC_REQUIRE(packed && ptr && col && val && A && out && flags, "jscc_materialize_rows: null pointer");
  hipLaunchKernelGGL(k_materialize, dim3((unsigned)n_rows), dim3(kRowThreads), 0, (hipStream_t)stream, packed,
                     ld_packed, n_rows, n_coarse, ptr, col, val, n_fine, A, lda, M, out, ld_out, normalize, flags);
  JSCC_CHECK_LAUNCH();
  return 0;
}

// Decodes all 256 e4m3fn codes with the hardware converter and with the
// software decoder; out[0..255] = hardware, out[256..511] = software.
__global__ void k_fp8_selftest(float* out) {
  const int c = threadIdx.x;
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v hw = __builtin_amdgcn_cvt_pk_f32_fp8(c, false);
  out[c] = hw.x;
  out[256 + c] = jscc_fp8_to_f32((uint32_t)c);
}

extern "C" int jscc_fp8_selftest(float* out512, void* stream) {
  JSCC_REQUIRE(out512 != nullptr, "jscc_fp8_selftest: null pointer");
  hipLaunchKernelGGL(k_fp8_selftest, dim3(1), dim3(256), 0, (hipStream_t)stream, out512);
  JSCC_CHECK_LAUNCH();
  return 0;
}

// ---------------------------------------------------------------------------
// Count thinning (main_plane.py:300-310; sparse_main_utils.py:130-150): the
// reference expands bin j into round(count_j) copies of j, draws a random
// permutation of the copies and keeps the first k.  Here the expansion is
// implicit: copy position p belongs to the bin whose cumulative count range
// contains p (upper bound in the inclusive prefix sums), and the kept copies
// are histogrammed with integer atomics (exact, order independent).
// ---------------------------------------------------------------------------
namespace {

__global__ void __launch_bounds__(256) k_fill_i32(int32_t* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = 0;
}

__global__ void __launch_bounds__(256) k_thin_hist(const int64_t* __restrict__ cum, int64_t n_bins,
                                                   const int64_t* __restrict__ pos, int64_t n_pos,
                                                   int32_t* __restrict__ hist, int32_t* __restrict__ err) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pos) return;
  const int64_t p = pos[i];
  if (p < 0 || n_bins == 0 || p >= cum[n_bins - 1]) {
    atomicOr(err, 16);
    return;
  }
  int64_t lo = 0, hi = n_bins - 1;  // first j with cum[j] > p
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (cum[mid] > p) hi = mid; else lo = mid + 1;
  }
  atomicAdd(hist + lo, 1);
}

__global__ void __launch_bounds__(256) k_i32_to_f32(const int32_t* __restrict__ h, int64_t n, float* __restrict__ out,
                                                    int64_t stride) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i * stride] = (float)h[i];
}

}  // namespace

extern "C" int jscc_thin_counts(const int64_t* cum_counts, int64_t n_bins, const int64_t* positions, int64_t n_pos,
                                float* out, int64_t out_stride, int32_t* workspace, int32_t* err, void* stream) {
  JSCC_REQUIRE(n_bins >= 0 && n_pos >= 0, "jscc_thin_counts: bad sizes");
  if (n_bins == 0) return 0;
  JSCC_REQUIRE(cum_counts && out && workspace && err && (n_pos == 0 || positions), "jscc_thin_counts: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const unsigned gb = (unsigned)((n_bins + 255) / 256);
  hipLaunchKernelGGL(k_fill_i32, dim3(gb), dim3(256), 0, s, workspace, n_bins);
  JSCC_CHECK_LAUNCH();
  if (n_pos > 0) {
    hipLaunchKernelGGL(k_thin_hist, dim3((unsigned)((n_pos + 255) / 256)), dim3(256), 0, s, cum_counts, n_bins,
                       positions, n_pos, workspace, err);
    JSCC_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_i32_to_f32, dim3(gb), dim3(256), 0, s, workspace, n_bins, out, out_stride > 0 ? out_stride : 1);
  JSCC_CHECK_LAUNCH();
  return 0;
}
