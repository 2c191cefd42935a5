// Projections of K frames (images) that share one BF16 system matrix, on the
// gfx950 matrix cores.  K single-photon reconstructions with the same matrix
// (the time frames of a dynamic acquisition, bootstrap replicates of one data
// set) each run get_weight_single (recon_osem_plane.py:28-30) and the EM
// update of osem_bin_mode (:45-54); batched, the per-subset products are GEMMs
//   forward  F[i][f] = sum_j R[i][j] X[j][f]        (rows i, pixels j, frames f)
//   back     W[j][f] = sum_i R[i][j] C[i][f],       C = Y / max(F, eps)
// each reducing over an index that is contiguous in its first operand when the
// matrix is kept twice: R row-major [rows x P] for the forward product and its
// transpose Rt [P x rows] for the back-projection (BF16 both).
//
// The frame operand (images X, coefficients C) is FP32.  It is carried as two
// BF16 terms hi = bf16(v), lo = bf16(v - hi), stored frame-major [K x n] with
// the reduced index contiguous, and every product is two MFMAs (R hi + R lo):
// R is exact in BF16 (it is stored so), v is kept to 16 significant bits, the
// accumulation is FP32 (v_mfma_f32_16x16x32_bf16).
//
// GEMM layout: a wave owns 32 output rows (two 16-row tiles) and all K frames
// (NT tiles of 16) over one of n_split contiguous chunks of the reduced index;
// its partial sums go to [n_split][rows][K] and the element-wise kernel that
// follows adds them in a fixed order.  Operands are 16-byte buffer loads whose
// range ends at the operand's last row (rows past it read zero).
#include "jscc_common.h"

namespace {

typedef __bf16 jscc_bf8 __attribute__((ext_vector_type(8)));
typedef float jscc_mf4 __attribute__((ext_vector_type(4)));

constexpr int kFrWaves = 4;        // independent waves per work-group
constexpr int kFrRows = 32;        // output rows per wave (two 16-row MFMA tiles)
constexpr int kFrK = 32;           // reduced indices per MFMA step
constexpr int kFrUnroll = 4;       // MFMA steps whose loads are issued together
constexpr int kBufWord3 = 0x00020000;

__device__ __forceinline__ uint16_t bf16_rne(float v) {
  // round to nearest even (the operands are finite or NaN; NaN stays NaN)
  const uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ jscc_bf8 load_frag(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(jscc_bf8, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// partial[sp][i][f] = sum over the reduced indices of chunk sp of a[i][k] * (bh[f][k] + bl[f][k])
template <int NT>
__global__ void __launch_bounds__(kFrWaves * JSCC_WAVE)
    k_frames_gemm(const uint16_t* __restrict__ a, int64_t lda, int64_t m, int64_t kdim,
                  const uint16_t* __restrict__ bh, const uint16_t* __restrict__ bl, int64_t ldb, int64_t kchunk,
                  int n_split, float* __restrict__ partial) {
  const int lane = jscc_lane();
  const int64_t wave = (int64_t)blockIdx.x * kFrWaves + jscc_wave_id();
  const int64_t m_tiles = (m + kFrRows - 1) / kFrRows;
  const int64_t mt = wave % m_tiles, sp = wave / m_tiles;
  if (sp >= n_split) return;
  const int64_t m0 = mt * kFrRows;
  const int64_t k0 = sp * kchunk;
  const int64_t k1 = k0 + kchunk < kdim ? k0 + kchunk : kdim;
  const int64_t rows = m - m0 < kFrRows ? m - m0 : kFrRows;
  constexpr int KF = 16 * NT;

  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(a + m0 * lda), (short)0,
                                                                      (int)(rows * lda * 2), kBufWord3);
  const __amdgpu_buffer_rsrc_t rh =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(bh), (short)0, (int)(KF * ldb * 2), kBufWord3);
  const __amdgpu_buffer_rsrc_t rl =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(bl), (short)0, (int)(KF * ldb * 2), kBufWord3);
  // lane l supplies row (l & 15) of a 16-row tile and the 8 reduced indices 8 (l >> 4) .. +7 of a step
  const int r16 = lane & 15, g8 = 8 * (lane >> 4);
  const int va0 = (int)((r16 * lda + g8) * 2), va1 = (int)(((16 + r16) * lda + g8) * 2);
  int vb[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) vb[u] = (int)(((16 * u + r16) * ldb + g8) * 2);

  jscc_mf4 acc[2][NT];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[t][u] = jscc_mf4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int64_t k = k0; k < k1; k += kFrK * kFrUnroll) {
    jscc_bf8 fa[kFrUnroll][2], fh[kFrUnroll][NT], fl[kFrUnroll][NT];
#pragma unroll
    for (int s = 0; s < kFrUnroll; ++s) {
      // kchunk and kdim are multiples of kFrK * kFrUnroll: every step of the group lies in the chunk
      const int so = (int)((k + s * kFrK) * 2);
      fa[s][0] = load_frag(ra, va0, so);
      fa[s][1] = load_frag(ra, va1, so);
#pragma unroll
      for (int u = 0; u < NT; ++u) {
        fh[s][u] = load_frag(rh, vb[u], so);
        fl[s][u] = load_frag(rl, vb[u], so);
      }
    }
#pragma unroll
    for (int s = 0; s < kFrUnroll; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[s][t], fh[s][u], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[s][t], fl[s][u], acc[t][u], 0, 0, 0);
        }
  }
  // lane l holds rows 4 (l >> 4) + j (j < 4) of column l & 15 of each 16 x 16 tile
  float* out = partial + (sp * m + m0) * KF;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * t + 4 * (lane >> 4) + j;
      if (r < rows) {
#pragma unroll
        for (int u = 0; u < NT; ++u) out[(int64_t)r * KF + 16 * u + r16] = acc[t][u][j];
      }
    }
}

// hi/lo BF16 terms of v (frame-major, reduced index contiguous)
__device__ __forceinline__ void split_store(float v, uint16_t* hi, uint16_t* lo, int64_t at) {
  const uint16_t h = bf16_rne(v);
  hi[at] = h;
  lo[at] = bf16_rne(v - bf16_val(h));
}

// t_hi/t_lo[f][off + i] = split(x[i][f]) for i < n, zero for i in [n, n_pad) (the reduced-index padding)
__global__ void k_frames_split(const float* __restrict__ x, int64_t n, int64_t n_pad, int KF, int64_t ldx,
                               uint16_t* __restrict__ hi, uint16_t* __restrict__ lo, int64_t ldt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pad * KF) return;
  const int64_t i = e / KF;
  const int f = (int)(e - i * KF);
  split_store(i < n ? x[i * ldx + f] : 0.0f, hi, lo, (int64_t)f * ldt + i);
}

// c = y / max(F, eps) for the rows [0, n) of a subset (F = the sum of the partials), written as the hi/lo
// terms of C^T at columns c_off + i; the columns [0, n_pad) outside [c_off, c_off + n) are zeroed
__global__ void k_frames_ratio(const float* __restrict__ partial, int n_split, int64_t n, int64_t n_pad,
                               int64_t c_off, int KF, const float* __restrict__ y, int64_t ldy, float eps,
                               int use_eps, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo, int64_t ldc,
                               float* __restrict__ f_out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_pad * KF) return;
  const int64_t ip = e / KF;
  const int f = (int)(e - ip * KF);
  const int64_t i = ip - c_off;
  float c = 0.0f;
  if (i >= 0 && i < n) {
    float F = 0.0f;
    for (int s = 0; s < n_split; ++s) F += partial[((int64_t)s * n + i) * KF + f];
    if (f_out) f_out[i * KF + f] = F;
    float den = F;
    if (use_eps && den < eps) den = eps;
    c = y[i * ldy + f] / den;
  }
  split_store(c, hi, lo, (int64_t)f * ldc + ip);
}

// x[j][f] <- update(x[j][f], W[j][f], s[j]) (W = the sum of the partials; modes of jscc_em_update) and the
// hi/lo terms of the updated image at t[f][j]
__global__ void k_frames_update(const float* __restrict__ partial, int n_split, int64_t P, int KF,
                                const float* __restrict__ s, float* __restrict__ x, int64_t ldx, int em_mode,
                                float em_eps, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo, int64_t ldt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * KF) return;
  const int64_t j = e / KF;
  const int f = (int)(e - j * KF);
  float W = 0.0f;
  if (partial)
    for (int sp = 0; sp < n_split; ++sp) W += partial[((int64_t)sp * P + j) * KF + f];
  const float v = jscc_em_value(x[j * ldx + f], W, s[j], em_mode, em_eps);
  x[j * ldx + f] = v;
  if (hi) split_store(v, hi, lo, (int64_t)f * ldt + j);
}

// dst[c][r] = src[r][c] (BF16), 64 x 64 tiles through LDS; dst rows are ld_dst long and the columns
// [rows, ld_dst) are zeroed
__global__ void __launch_bounds__(256) k_frames_transpose(const uint16_t* __restrict__ src, int64_t rows,
                                                          int64_t cols, int64_t ld_src, uint16_t* __restrict__ dst,
                                                          int64_t ld_dst) {
  __shared__ uint16_t tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int rr = ty; rr < 64; rr += 4) {
    const int64_t r = r0 + rr, c = c0 + tx;
    tile[rr][tx] = (r < rows && c < cols) ? src[r * ld_src + c] : (uint16_t)0;
  }
  __syncthreads();
  for (int cc = ty; cc < 64; cc += 4) {
    const int64_t c = c0 + cc, r = r0 + tx;
    if (c < cols && r < ld_dst) dst[c * ld_dst + r] = tile[tx][cc];
  }
}

int frames_grid(int64_t waves) { return (int)((waves + kFrWaves - 1) / kFrWaves); }

}  // namespace

extern "C" int jscc_frames_transpose_bf16(const uint16_t* src, int64_t rows, int64_t cols, int64_t ld_src,
                                          uint16_t* dst, int64_t ld_dst, void* stream) {
  JSCC_REQUIRE(src && dst && rows > 0 && cols > 0 && ld_src >= cols && ld_dst >= rows,
               "jscc_frames_transpose_bf16: bad shape");
  const dim3 grid((unsigned)((cols + 63) / 64), (unsigned)((ld_dst + 63) / 64));
  hipLaunchKernelGGL(k_frames_transpose, grid, dim3(256), 0, (hipStream_t)stream, src, rows, cols, ld_src, dst,
                     ld_dst);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_frames_split(const float* x, int64_t n, int64_t n_pad, int k_frames, int64_t ldx, uint16_t* hi,
                                 uint16_t* lo, int64_t ldt, void* stream) {
  JSCC_REQUIRE(x && hi && lo && n >= 0 && n_pad >= n && k_frames > 0 && ldx >= k_frames && ldt >= n_pad,
               "jscc_frames_split: bad shape");
  if (n_pad == 0) return 0;
  const int64_t tot = n_pad * k_frames;
  hipLaunchKernelGGL(k_frames_split, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, n,
                     n_pad, k_frames, ldx, hi, lo, ldt);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_frames_split_count(int64_t m, int64_t kdim, int cus) {
  // about 16 waves per compute unit, each over at least 8 MFMA steps of the reduced index
  const int64_t m_tiles = (m + kFrRows - 1) / kFrRows;
  const int64_t want = (16LL * (cus > 0 ? cus : 256) + m_tiles - 1) / m_tiles;
  const int64_t most = kdim / (kFrK * kFrUnroll * 2) > 0 ? kdim / (kFrK * kFrUnroll * 2) : 1;
  int64_t ns = want < most ? want : most;
  return (int)(ns < 1 ? 1 : ns);
}

extern "C" int64_t jscc_frames_partial_bytes(int64_t m, int k_frames, int n_split) {
  return m * (int64_t)k_frames * n_split * 4;
}

extern "C" int jscc_frames_gemm(const uint16_t* a, int64_t lda, int64_t m, int64_t kdim, const uint16_t* b_hi,
                                const uint16_t* b_lo, int64_t ldb, int k_frames, int n_split, float* partial,
                                int64_t partial_bytes, void* stream) {
  JSCC_REQUIRE(a && b_hi && b_lo && partial, "jscc_frames_gemm: null pointer");
  JSCC_REQUIRE(k_frames == 16 || k_frames == 32 || k_frames == 64, "jscc_frames_gemm: %d frames (16, 32 or 64)",
               k_frames);
  JSCC_REQUIRE(m > 0 && kdim > 0 && n_split >= 1, "jscc_frames_gemm: bad shape");
  JSCC_REQUIRE(kdim % (kFrK * kFrUnroll) == 0 && lda >= kdim && ldb >= kdim,
               "jscc_frames_gemm: the reduced length %lld must be a multiple of %d within both strides",
               (long long)kdim, kFrK * kFrUnroll);
  JSCC_REQUIRE((uintptr_t)a % 16 == 0 && (uintptr_t)b_hi % 16 == 0 && (uintptr_t)b_lo % 16 == 0 && lda % 8 == 0 &&
                   ldb % 8 == 0,
               "jscc_frames_gemm: operands must be 16-byte aligned with strides a multiple of 8");
  JSCC_REQUIRE((double)kFrRows * lda * 2 < 2147483647.0 && (double)k_frames * ldb * 2 < 2147483647.0,
               "jscc_frames_gemm: operand tile beyond the 32-bit buffer range");
  JSCC_REQUIRE(partial_bytes >= jscc_frames_partial_bytes(m, k_frames, n_split), "jscc_frames_gemm: partial too small");
  int64_t kchunk = (kdim + n_split - 1) / n_split;
  kchunk = (kchunk + kFrK * kFrUnroll - 1) / (kFrK * kFrUnroll) * (kFrK * kFrUnroll);
  const int ns = (int)((kdim + kchunk - 1) / kchunk);
  JSCC_REQUIRE(ns == n_split, "jscc_frames_gemm: %d chunks of the reduced index do not split %lld evenly into "
               "multiples of %d (use jscc_frames_split_count)", n_split, (long long)kdim, kFrK * kFrUnroll);
  const int64_t waves = (m + kFrRows - 1) / kFrRows * n_split;
  hipStream_t s = (hipStream_t)stream;
  jscc_prof_begin(JSCC_TAG_FRAMES, s);
#define JSCC_FR(NTV)                                                                                         \
  hipLaunchKernelGGL(k_frames_gemm<NTV>, dim3(frames_grid(waves)), dim3(kFrWaves * JSCC_WAVE), 0, s, a, lda, m, \
                     kdim, b_hi, b_lo, ldb, kchunk, n_split, partial)
  if (k_frames == 16) JSCC_FR(1);
  else if (k_frames == 32) JSCC_FR(2);
  else JSCC_FR(4);
#undef JSCC_FR
  jscc_prof_end(JSCC_TAG_FRAMES, s);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_frames_ratio(const float* partial, int n_split, int64_t n, int64_t n_pad, int64_t c_off,
                                 int k_frames, const float* y, int64_t ldy, float eps, int use_eps, uint16_t* c_hi,
                                 uint16_t* c_lo, int64_t ldc, float* f_out, void* stream) {
  JSCC_REQUIRE(partial && y && c_hi && c_lo && n >= 1 && c_off >= 0 && c_off + n <= n_pad && ldc >= n_pad &&
                   ldy >= k_frames && k_frames > 0 && n_split >= 1,
               "jscc_frames_ratio: bad shape");
  const int64_t tot = n_pad * k_frames;
  hipLaunchKernelGGL(k_frames_ratio, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, partial,
                     n_split, n, n_pad, c_off, k_frames, y, ldy, eps, use_eps, c_hi, c_lo, ldc, f_out);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_frames_update(const float* partial, int n_split, int64_t n_pix, int k_frames, const float* s,
                                  float* x, int64_t ldx, int em_mode, float em_eps, uint16_t* t_hi, uint16_t* t_lo,
                                  int64_t ldt, void* stream) {
  JSCC_REQUIRE(s && x && n_pix >= 1 && k_frames > 0 && ldx >= k_frames && em_mode >= 0 && em_mode <= 2 &&
                   (partial == nullptr || n_split >= 1) && ((t_hi == nullptr) == (t_lo == nullptr)) &&
                   (t_hi == nullptr || ldt >= n_pix),
               "jscc_frames_update: bad arguments");
  const int64_t tot = n_pix * k_frames;
  hipLaunchKernelGGL(k_frames_update, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, partial,
                     n_split, n_pix, k_frames, s, x, ldx, em_mode, em_eps, t_hi, t_lo, ldt);
  JSCC_CHECK_LAUNCH();
  return 0;
}
