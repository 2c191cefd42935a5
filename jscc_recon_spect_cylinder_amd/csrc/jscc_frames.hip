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
// GEMM layout: a work-group owns 64 output rows and all K frames over one of
// n_split contiguous chunks of the reduced index; its partial sums go to
// [n_split][rows][K] and the element-wise kernel that follows adds them in a
// fixed order.  Operands reach LDS by 16-byte memory -> LDS copies of
// contiguous row segments whose range ends at the operand's last row (rows
// past it read zero), three or four stages in flight.
#include "jscc_common.h"

namespace {

typedef __bf16 jscc_bf8 __attribute__((ext_vector_type(8)));
typedef float jscc_mf4 __attribute__((ext_vector_type(4)));

constexpr int kFrWaves = 4;        // waves per work-group
constexpr int kFrRowsWg = 64;      // output rows per work-group (a 16-row MFMA tile per wave)
constexpr int kFrK = 32;           // reduced indices per MFMA step
constexpr int kFrAlign = 256;      // the reduced length is a multiple of every stage length
constexpr int kBufWord3 = 0x00020000;
typedef __attribute__((address_space(3))) void* lds_vp;
typedef int jscc_i4v __attribute__((ext_vector_type(4)));

// s_waitcnt immediate of gfx9 waiting for vmcnt <= n only (expcnt and lgkmcnt at their maxima)
constexpr int jscc_vmcnt_enc(int n) { return (n & 0xf) | ((n >> 4) << 14) | (0x7 << 4) | (0xf << 8); }

__device__ __forceinline__ uint16_t bf16_rne(float v) {
  // round to nearest even (the operands are finite or NaN; NaN stays NaN)
  const uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }


// Stage geometry of the GEMM: KS reduced indices per stage (512 B of a BF16 row for 16 frames, 256 B
// otherwise), NSTAGE stages in flight, LDS = NSTAGE * (64 rows + 2 * 16 NT frame rows) * KS * 2 bytes.
template <int NT> struct FrGeom {
  static constexpr int KS = NT == 1 ? 256 : 128;
  static constexpr int NSTAGE = NT == 2 ? 4 : 3;
  static constexpr int RB = KS * 2;                     // bytes of one row in a stage
  static constexpr int CH = RB / 16;                    // 16-byte chunks of a row in a stage
  static constexpr int A_BYTES = kFrRowsWg * RB;        // the matrix tile of a stage
  static constexpr int B_BYTES = 16 * NT * RB;          // one frame term of a stage
  static constexpr int STAGE = A_BYTES + 2 * B_BYTES;
  static constexpr int LDS = NSTAGE * STAGE;
  static constexpr int A_COPIES = A_BYTES / 1024 / kFrWaves;  // 1 KB copies per wave and stage
  static constexpr int B_COPIES = B_BYTES / 1024 / kFrWaves;  // per frame term
};

// 16-byte memory -> LDS copy of one lane (buffer range checked); inline asm, so that the compiler does not
// wait for it before every later LDS read (the stage barriers wait for it instead); no register outputs
__device__ __forceinline__ void dma16(const jscc_i4v& rsrc, int voff, int soff, uint32_t lds_byte) {
  asm volatile("buffer_load_dwordx4 %0, %1, %2 offen lds" ::"v"(voff), "s"(rsrc), "s"(soff), "{m0}"(lds_byte)
               : "memory");
}

__device__ __forceinline__ jscc_i4v rsrc_of(const void* base, int64_t bytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(base);
  jscc_i4v r;
  r.x = (int)(uint32_t)a;
  r.y = (int)((uint32_t)(a >> 32) & 0xffffu);
  r.z = (int)bytes;
  r.w = kBufWord3;
  return r;
}

// partial[sp][i][f] = sum over the reduced indices of chunk sp of a[i][k] * (bh[f][k] + bl[f][k]).
// A work-group of 4 waves owns 64 rows (wave w the 16-row MFMA tile w) and all 16 NT frames.  Each stage
// brings the rows' next KS reduced indices and both frame terms into LDS by 16-byte copies of contiguous
// row segments (a row of a stage is RB contiguous bytes in memory); chunk c of row r sits at chunk slot
// c ^ (r & 15), so that the 16 rows of an MFMA fragment read 16 different bank groups.
template <int NT>
__global__ void __launch_bounds__(kFrWaves * JSCC_WAVE)
    k_frames_gemm(const uint16_t* __restrict__ a, int64_t lda, int64_t m, int64_t kdim,
                  const uint16_t* __restrict__ bh, const uint16_t* __restrict__ bl, int64_t ldb, int64_t kchunk,
                  int n_split, float* __restrict__ partial) {
  using Gm = FrGeom<NT>;
  extern __shared__ __attribute__((aligned(16))) unsigned char frlds[];
  const int lane = jscc_lane(), wv = jscc_wave_id();
  const int64_t m_tiles = (m + kFrRowsWg - 1) / kFrRowsWg;
  const int64_t mt = blockIdx.x % m_tiles, sp = blockIdx.x / m_tiles;
  const int64_t m0 = mt * kFrRowsWg;
  const int64_t k0 = sp * kchunk;
  const int64_t k1 = k0 + kchunk < kdim ? k0 + kchunk : kdim;
  const int n_st = (int)((k1 - k0) / Gm::KS);
  const int64_t rows = m - m0 < kFrRowsWg ? m - m0 : kFrRowsWg;
  constexpr int KF = 16 * NT;

  const jscc_i4v ra = rsrc_of(a + m0 * lda, rows * lda * 2);
  const jscc_i4v rh = rsrc_of(bh, (int64_t)KF * ldb * 2), rl = rsrc_of(bl, (int64_t)KF * ldb * 2);
  const uint32_t lds0 = (uint32_t)(uintptr_t)(lds_vp)frlds;
  // copy q of a wave fills LDS bytes [1 KB * (q * waves + w), +1 KB) of a tile: lane j -> row
  // (that KB * 1024 + 16 j) / RB, slot (16 j / 16) % CH, which holds the row's chunk slot ^ (row & 15)
  auto row_of = [&](int kb) { return (kb * 1024 + 16 * lane) / Gm::RB; };
  auto chunk_of = [&](int kb) { return ((lane % Gm::CH) ^ (row_of(kb) & 15)); };
  auto issue = [&](int st) {
    const int buf = st % Gm::NSTAGE;
    const uint32_t sb = lds0 + buf * Gm::STAGE;
    const int so = (int)((k0 + (int64_t)st * Gm::KS) * 2);
#pragma unroll
    for (int q = 0; q < Gm::A_COPIES; ++q) {
      const int kb = q * kFrWaves + wv;
      dma16(ra, (int)(row_of(kb) * lda * 2) + chunk_of(kb) * 16, so, sb + kb * 1024);
    }
#pragma unroll
    for (int q = 0; q < Gm::B_COPIES; ++q) {
      const int kb = q * kFrWaves + wv;
      const int vo = (int)(row_of(kb) * ldb * 2) + chunk_of(kb) * 16;
      dma16(rh, vo, so, sb + Gm::A_BYTES + kb * 1024);
      dma16(rl, vo, so, sb + Gm::A_BYTES + Gm::B_BYTES + kb * 1024);
    }
  };
  constexpr int PER_STAGE = Gm::A_COPIES + 2 * Gm::B_COPIES;  // copies per wave and stage

  jscc_mf4 acc[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) acc[u] = jscc_mf4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int st = 0; st < Gm::NSTAGE - 1 && st < n_st; ++st) issue(st);

  const int r16 = lane & 15, g = lane >> 4;
  for (int st = 0; st < n_st; ++st) {
    // stage st has landed once at most the copies of the later stages already issued are outstanding
    const int later = n_st - 1 - st < Gm::NSTAGE - 2 ? n_st - 1 - st : Gm::NSTAGE - 2;
    if (later >= 2) __builtin_amdgcn_s_waitcnt(jscc_vmcnt_enc(2 * PER_STAGE));
    else if (later == 1) __builtin_amdgcn_s_waitcnt(jscc_vmcnt_enc(PER_STAGE));
    else __builtin_amdgcn_s_waitcnt(jscc_vmcnt_enc(0));
    __syncthreads();  // every wave's copies of stage st; every wave is past stage st - 1 (its buffer is free)
    if (st + Gm::NSTAGE - 1 < n_st) issue(st + Gm::NSTAGE - 1);
    const unsigned char* sa = frlds + (st % Gm::NSTAGE) * Gm::STAGE;
    const unsigned char* sbh = sa + Gm::A_BYTES;
    const unsigned char* sbl = sbh + Gm::B_BYTES;
#pragma unroll
    for (int s = 0; s < Gm::KS / kFrK; ++s) {
      const int slot = (4 * s + g) ^ r16;
      const jscc_bf8 fa = *reinterpret_cast<const jscc_bf8*>(sa + (16 * wv + r16) * Gm::RB + slot * 16);
#pragma unroll
      for (int u = 0; u < NT; ++u) {
        const int off = (16 * u + r16) * Gm::RB + slot * 16;
        const jscc_bf8 fh = *reinterpret_cast<const jscc_bf8*>(sbh + off);
        const jscc_bf8 fl = *reinterpret_cast<const jscc_bf8*>(sbl + off);
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fh, acc[u], 0, 0, 0);
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fl, acc[u], 0, 0, 0);
      }
    }
  }
  // lane l holds rows 4 (l >> 4) + j (j < 4) of column l & 15 of each 16 x 16 tile
  float* out = partial + (sp * m + m0) * KF;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = 16 * wv + 4 * g + j;
    if (r < rows) {
#pragma unroll
      for (int u = 0; u < NT; ++u) out[(int64_t)r * KF + 16 * u + r16] = acc[u][j];
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

template <int NT>
int launch_frames_gemm(hipStream_t s, int64_t G, const uint16_t* a, int64_t lda, int64_t m, int64_t kdim,
                       const uint16_t* bh, const uint16_t* bl, int64_t ldb, int64_t kchunk, int n_split,
                       float* partial) {
  static bool attr_set = false;
  if (!attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_frames_gemm<NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, FrGeom<NT>::LDS));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_frames_gemm<NT>, dim3((unsigned)G), dim3(kFrWaves * JSCC_WAVE), FrGeom<NT>::LDS, s, a, lda,
                     m, kdim, bh, bl, ldb, kchunk, n_split, partial);
  return 0;
}

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
  // about one work-group per compute unit (one is resident: 96-144 KB of LDS), each over at least four
  // stages of the reduced index (measured on MI355X at 8192 x 16384 and 16384 x 8192, 16 frames: one
  // work-group per unit 54.9 / 52.0 us, two 60.5 / 54.2 us, four 61.0-64.4 / 60.8 us)
  const int64_t m_tiles = (m + kFrRowsWg - 1) / kFrRowsWg;
  const int64_t want = ((cus > 0 ? cus : 256) + m_tiles - 1) / m_tiles;
  const int64_t most = kdim / (4 * kFrAlign) > 0 ? kdim / (4 * kFrAlign) : 1;
  int64_t ns = want < most ? want : most;
  ns = ns < 1 ? 1 : ns;
  // the count the GEMM accepts: chunks of a multiple of kFrAlign that split kdim into exactly ns
  int64_t kchunk = (kdim + ns - 1) / ns;
  kchunk = (kchunk + kFrAlign - 1) / kFrAlign * kFrAlign;
  return (int)((kdim + kchunk - 1) / kchunk);
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
  JSCC_REQUIRE(kdim % kFrAlign == 0 && lda >= kdim && ldb >= kdim,
               "jscc_frames_gemm: the reduced length %lld must be a multiple of %d within both strides",
               (long long)kdim, kFrAlign);
  JSCC_REQUIRE((uintptr_t)a % 16 == 0 && (uintptr_t)b_hi % 16 == 0 && (uintptr_t)b_lo % 16 == 0 && lda % 8 == 0 &&
                   ldb % 8 == 0,
               "jscc_frames_gemm: operands must be 16-byte aligned with strides a multiple of 8");
  JSCC_REQUIRE((double)kFrRowsWg * lda * 2 < 2147483647.0 && (double)k_frames * ldb * 2 < 2147483647.0,
               "jscc_frames_gemm: operand tile beyond the 32-bit buffer range");
  JSCC_REQUIRE(partial_bytes >= jscc_frames_partial_bytes(m, k_frames, n_split), "jscc_frames_gemm: partial too small");
  int64_t kchunk = (kdim + n_split - 1) / n_split;
  kchunk = (kchunk + kFrAlign - 1) / kFrAlign * kFrAlign;
  const int ns = (int)((kdim + kchunk - 1) / kchunk);
  JSCC_REQUIRE(ns == n_split, "jscc_frames_gemm: %d chunks of the reduced index do not split %lld evenly into "
               "multiples of %d (use jscc_frames_split_count)", n_split, (long long)kdim, kFrAlign);
  const int64_t G = (m + kFrRowsWg - 1) / kFrRowsWg * n_split;
  hipStream_t s = (hipStream_t)stream;
  int rc;
  jscc_prof_begin(JSCC_TAG_FRAMES, s);
  if (k_frames == 16) rc = launch_frames_gemm<1>(s, G, a, lda, m, kdim, b_hi, b_lo, ldb, kchunk, n_split, partial);
  else if (k_frames == 32) rc = launch_frames_gemm<2>(s, G, a, lda, m, kdim, b_hi, b_lo, ldb, kchunk, n_split, partial);
  else rc = launch_frames_gemm<4>(s, G, a, lda, m, kdim, b_hi, b_lo, ldb, kchunk, n_split, partial);
  jscc_prof_end(JSCC_TAG_FRAMES, s);
  if (rc != 0) return rc;
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
