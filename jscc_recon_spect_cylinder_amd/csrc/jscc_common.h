// Shared device helpers for the gfx950 kernels of libjscc_hip.
//
// Everything here is CDNA-specific: 64-lane wavefronts, DPP/bpermute wave
// reductions, BF16 / OCP-FP8 (e4m3fn) storage formats decoded to FP32 in
// registers.  No CUDA, no dual paths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define JSCC_WAVE 64

// ---------------------------------------------------------------------------
// error reporting (host side; one thread-local message per calling thread)
// ---------------------------------------------------------------------------
void jscc_set_error(const char* fmt, ...);

#define JSCC_CHECK_HIP(expr)                                                     \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      jscc_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),      \
                     __FILE__, __LINE__);                                        \
      return -1;                                                                 \
    }                                                                            \
  } while (0)

#define JSCC_REQUIRE(cond, ...)                                                  \
  do {                                                                           \
    if (!(cond)) {                                                               \
      jscc_set_error(__VA_ARGS__);                                               \
      return -2;                                                                 \
    }                                                                            \
  } while (0)

// after a launch: catch launch-configuration errors without synchronising
#define JSCC_CHECK_LAUNCH()                                                      \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess) {                                                      \
      jscc_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),  \
                     __FILE__, __LINE__);                                        \
      return -1;                                                                 \
    }                                                                            \
  } while (0)

// storage dtype codes shared with the C ABI
enum {
  JSCC_DT_F32 = 0,
  JSCC_DT_BF16 = 1,
  JSCC_DT_FP8 = 2,  // OCP e4m3fn, per-row scale carried separately
  JSCC_DT_FP16 = 3, // IEEE half, per-row scale carried separately (row max at 2^14)
};

// ---------------------------------------------------------------------------
// storage formats
// ---------------------------------------------------------------------------
__device__ __forceinline__ float jscc_bf16_to_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// two FP32 -> packed BF16x2 with the gfx950 converter (round-to-nearest-even)
typedef float jscc_f2_t __attribute__((ext_vector_type(2)));
typedef __bf16 jscc_bf2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t jscc_pack_bf16x2(float lo, float hi) {
  jscc_f2_t v;
  v.x = lo;
  v.y = hi;
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, jscc_bf2_t));
}

// e4m3fn -> FP32.  The 7 magnitude bits are placed at the top of the FP32
// exponent/mantissa field and rescaled by 2^120 (= 2^(127-7)), which handles
// normals and subnormals with one multiply.  0x7f / 0xff are NaN.
__device__ __forceinline__ float jscc_fp8_to_f32(uint32_t b) {
  uint32_t mag = b & 0x7fu;
  float v = __uint_as_float(mag << 20) * 1.329227995784916e36f;  // 2^120
  if (mag == 0x7fu) v = __uint_as_float(0x7fc00000u);
  return (b & 0x80u) ? -v : v;
}

// FP32 -> e4m3fn, round-to-nearest-even, saturating to +-448.
__device__ __forceinline__ uint32_t jscc_f32_to_fp8(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t sign = (u >> 24) & 0x80u;
  float a = fabsf(f);
  if (!(a == a)) return 0x7fu | sign;  // NaN
  if (a >= 448.0f) return 0x7eu | sign;  // saturate (e4m3fn max = 448)
  if (a < 0.0009765625f) {  // below half of the smallest subnormal 2^-9 -> maybe 0
    // subnormal grid is k * 2^-9, k in 0..7 ; handled below as well
  }
  // exponent of a (unbiased); a>0 here unless zero
  if (a == 0.0f) return sign;
  int e = (int)((__float_as_uint(a) >> 23) & 0xff) - 127;
  if (e < -6) {
    // subnormal in e4m3: value = k * 2^-9
    float k = rintf(a * 512.0f);  // RNE on the 2^-9 grid
    uint32_t kk = (uint32_t)k;    // 0..8 ; 8 becomes the smallest normal
    return sign | kk;             // kk==8 -> 0b0001000 = exp 1, man 0 (2^-6): correct
  }
  // normal: keep 3 mantissa bits with RNE
  uint32_t m = __float_as_uint(a) & 0x7fffffu;
  uint32_t keep = m >> 20;
  uint32_t rem = m & 0xfffffu;
  uint32_t half = 0x80000u;
  if (rem > half || (rem == half && (keep & 1u))) keep += 1;
  uint32_t ee = (uint32_t)(e + 7);
  if (keep == 8) { keep = 0; ee += 1; }
  uint32_t code = (ee << 3) | keep;
  if (code > 0x7eu) code = 0x7eu;
  return sign | code;
}

// ---------------------------------------------------------------------------
// wave / block reductions (64-wide wavefronts)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T jscc_wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, JSCC_WAVE);
  return v;
}

template <typename T>
__device__ __forceinline__ T jscc_wave_max(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    T o = __shfl_xor(v, off, JSCC_WAVE);
    v = (o > v) ? o : v;
  }
  return v;
}

// Wave sum with DPP lane moves folded into the adds (quad swaps, half-row and
// row mirrors, then the gfx9 row broadcasts); the total is read from lane 63
// and is uniform across the wave.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float jscc_dpp_add(float v) {
  const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false);
  return v + __int_as_float(moved);
}

__device__ __forceinline__ float jscc_wave_sum_dpp(float v) {
  v = jscc_dpp_add<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = jscc_dpp_add<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v = jscc_dpp_add<0x141, 0xf>(v);  // row_half_mirror: sums of 8
  v = jscc_dpp_add<0x140, 0xf>(v);  // row_mirror: sums of 16
  v = jscc_dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = jscc_dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ int jscc_lane() { return threadIdx.x & (JSCC_WAVE - 1); }
__device__ __forceinline__ int jscc_wave_id() { return threadIdx.x >> 6; }

// number of compute units of the current device (cached per process)
int jscc_num_cus();
// compute units a launch on `stream` can use: the popcount of its CU mask
// (hipExtStreamCreateWithCUMask), else the device's count; persistent grids
// (one work-group per CU, all resident) are sized by it
int jscc_stream_cus(hipStream_t stream);

// EM updates: 0 plain x*w/s (recon_osem_plane.py:54); 1 safe_em_update
// (recon_osem_plane_sparse.py:60-66); 2 apply_em_update
// (distributed/python/recon_osem_dist.py:46-51).
__device__ __forceinline__ float jscc_nan_to_num(float v, float nan_v, float pinf, float ninf) {
  if (v != v) return nan_v;
  if (v == __builtin_inff()) return pinf;
  if (v == -__builtin_inff()) return ninf;
  return v;
}

__device__ __forceinline__ float jscc_em_value(float xv, float w, float sv, int mode, float eps) {
  if (mode == 0) return (xv * w) / sv;
  if (mode == 1) {
    const float ws = jscc_nan_to_num(w, 0.0f, 0.0f, 0.0f);
    const float ss = jscc_nan_to_num(sv, 0.0f, 0.0f, 0.0f);
    return ss > eps ? (xv * fmaxf(ws, 0.0f)) / ss : 0.0f;
  }
  const float ss = (sv != sv) ? sv : fmaxf(sv, eps);
  float u = w / ss;
  u = jscc_nan_to_num(u, 1.0f, 1e6f, 0.0f);
  return jscc_nan_to_num(xv * u, 0.0f, 1e12f, 0.0f);
}

// Optional EM update applied where a subset weight is finished:
// x[j] <- em(x[j], c1*w[j] (+ c2*w2[j]), s[j]) (modes of jscc_em_update); w[j]
// (and w2[j]) are zeroed after when zero_w is set.  x == nullptr: no update.
struct JsccEm {
  float* x;
  const float* s;
  int mode;
  float eps;
  int zero_w;
  float c1;
  float* w2;
  float c2;
};

// Deterministic reduction of per-work-group partial back-projections [G, P]
// (row stride ldp: a multiple of 4 covering P) into w (inverse-rotation
// gather, scale, accumulate, optional nan_to_num, optional EM update); the
// buffer must hold jscc_partials_bytes(G, ldp).
int jscc_reduce_partials(float* part, int G, int P, int64_t ldp, const int64_t* inv_perm, int64_t ips, float scale,
                         int accumulate, int sanitize, float* w, int64_t w_len, int32_t* err, hipStream_t s,
                         const JsccEm& em);
int64_t jscc_partials_bytes(int G, int64_t ldp);

// Built-in kernel timer: when enabled, hot kernels are bracketed by HIP event
// pairs recorded on their launch stream (tag = kernel family).
enum {
  JSCC_TAG_CONE = 0,
  JSCC_TAG_PROJ = 1,
  JSCC_TAG_PROJ_SUM = 2,
  JSCC_TAG_CONE_PROJ = 3,
  JSCC_TAG_COARSE = 4,
  JSCC_TAG_FRAMES = 5,
  JSCC_NUM_TAGS = 6
};
bool jscc_prof_on();
void jscc_prof_begin(int tag, hipStream_t s);
void jscc_prof_end(int tag, hipStream_t s);
