// OSEM projection kernels.
//
// Reference: recon_osem_plane.py:28-42 (get_weight_single / get_weight_compton),
// :45-125 (osem_*_mode: rotation gathers/scatters, alpha weighting, update),
// recon_osem_local_cntstat.py:6-8 (clamped twin), recon_osem_plane_sparse.py:60-66
// (safe_em_update), distributed/python/recon_osem_dist.py:46-51 (apply_em_update),
// main_plane.py:320-323 (sensitivity).
//
// One pass over the matrix computes both projections: a work-group keeps the
// (permuted) image and its partial back-projection in registers, streams its
// share of rows once from HBM, reduces each row's dot product across the
// work-group, forms the ratio y/max(Ax,eps) and accumulates ratio*row into the
// registers.  Per-work-group partial back-projections are summed by a
// two-level deterministic column reduction that also applies the inverse
// rotation gather and the alpha weight.  Rows are FP32, BF16 or FP8-e4m3fn
// (with a per-row scale), always accumulated in FP32.
#include "jscc_common.h"
#include "jscc_hip.h"

#include <mutex>

namespace {

enum { MODE_RATIO = 0, MODE_FWD = 1, MODE_COEF = 2, MODE_ONES = 3 };

constexpr int kProjThreads = 1024;  // one work-group per CU, 4 waves per SIMD
constexpr int kProjWaves = kProjThreads / JSCC_WAVE;

typedef float jscc_f4v __attribute__((ext_vector_type(4)));
typedef unsigned int jscc_u4v __attribute__((ext_vector_type(4)));
typedef float jscc_f2v __attribute__((ext_vector_type(2)));

template <int DT> struct Vec;
template <> struct Vec<JSCC_DT_F32> { static constexpr int N = 4; using T = jscc_f4v; };
template <> struct Vec<JSCC_DT_BF16> { static constexpr int N = 8; using T = jscc_u4v; };
template <> struct Vec<JSCC_DT_FP8> { static constexpr int N = 16; using T = jscc_u4v; };
template <> struct Vec<JSCC_DT_FP16> { static constexpr int N = 8; using T = jscc_u4v; };

template <int DT>
__device__ __forceinline__ void decode(const typename Vec<DT>::T& raw, float* v);

template <>
__device__ __forceinline__ void decode<JSCC_DT_F32>(const jscc_f4v& r, float* v) {
  v[0] = r.x; v[1] = r.y; v[2] = r.z; v[3] = r.w;
}
template <>
__device__ __forceinline__ void decode<JSCC_DT_BF16>(const jscc_u4v& r, float* v) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(r[k] << 16);
    v[2 * k + 1] = __uint_as_float(r[k] & 0xffff0000u);
  }
}
// IEEE half: the conversion folds into the consuming FMA (v_fma_mix_f32 selects
// either half of the register), so an FP16 element costs one instruction
template <>
__device__ __forceinline__ void decode<JSCC_DT_FP16>(const jscc_u4v& r, float* v) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = (float)__builtin_bit_cast(_Float16, (uint16_t)(r[k] & 0xffffu));
    v[2 * k + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(r[k] >> 16));
  }
}
// gfx950 converts OCP e4m3fn in hardware, two bytes per instruction
template <>
__device__ __forceinline__ void decode<JSCC_DT_FP8>(const jscc_u4v& r, float* v) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const jscc_f2v lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)r[k], false);
    const jscc_f2v hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)r[k], true);
    v[4 * k] = lo.x; v[4 * k + 1] = lo.y; v[4 * k + 2] = hi.x; v[4 * k + 3] = hi.y;
  }
}

// The second row segment of a joint launch: a matrix of the same storage type
// and width whose rows start `off` elements from the first matrix, its own
// stored-row list, scales and numerator, and a factor on its numerators.
struct ProjSeg {
  int64_t off;
  int64_t ld;
  const int32_t* rows;
  int64_t row_base;
  const float* row_scale;
  const float* y;
  int64_t y_stride;
  int y_by_row;
  float num_scale;
};

struct ProjArgs {
  const void* M;             // matrix rows
  int64_t ld;                // row stride (elements)
  int P;                     // columns (any count; rows are read in whole vectors past it)
  int ldp;                   // row stride of the partial sums (P rounded up to 16)
  const int32_t* rows;       // optional stored-row index list
  int64_t n_rows;            // rows in this call
  int64_t row_base;          // offset into rows / stored row ids
  const float* row_scale;    // optional per stored row scale
  const float* x;            // image [P_img]
  const int64_t* x_perm;     // optional 1-based gather x_rot[p] = x[perm[p]-1]
  int64_t x_perm_stride;
  int64_t x_len;             // length of x (bounds for the gather)
  const float* y;            // optional numerator per call row (ones if null)
  int64_t y_stride;
  const float* coef;         // MODE_COEF: per call row coefficient
  float eps;                 // clamp of A x (when use_eps)
  int use_eps;
  float* fwd_out;            // MODE_FWD: per call row A x
  float* partial;            // [gridDim.x, P]
  int32_t* err;              // index error flag
  int y_by_row;              // numerator indexed by stored row id instead of call row
  // Joint launches: call rows [0, n_rows1) come from the operands above with
  // their numerators scaled by num_scale1, rows [n_rows1, n_rows) from seg2
  // (n_rows1 == n_rows and num_scale1 == 1 otherwise).
  int64_t n_rows1;
  float num_scale1;
  ProjSeg seg2;
  // Counted launches: when cnt is set, the call rows are [lo, hi) past
  // row_base with lo = n * sub_i / sub_k, hi = n * (sub_i + 1) / sub_k and
  // n = *cnt clamped to [0, cnt_cap] -- the split_tensor_even bounds of a row
  // count that stays on the device (n_rows then only sizes the grid).
  const int32_t* cnt;
  int64_t cnt_cap;
  int sub_i, sub_k;
};

// Row metadata staged in LDS for a chunk of a work-group's rows: stored row
// id, row scale and numerator (y, or the MODE_COEF coefficient).  Reading them
// from LDS inside the row loop never waits on the row loads in flight.
constexpr int kStage = 128;

// gfx9 buffer resource word 3 (raw, untyped) and the nontemporal cache policy
constexpr int kBufWord3 = 0x00020000;
constexpr int kBufNT = 2;

struct RowMeta {
  int64_t off[kStage];  // element offset of the row from the matrix start
  float sc[kStage];
  float num[kStage];
};

// R rows held in their storage format (decoded at use)
template <int DT, int NCH, int R>
struct RowBuf {
  typename Vec<DT>::T raw[R][NCH];
};

// A column count that is not a multiple of the vector width leaves a ragged last
// vector per row: the thread that holds it keeps only the first nb bytes of it
// (nb = 16 everywhere else), so the row's padding -- memory a store need not
// have written -- cannot reach the sums as 0 * inf or NaN.
template <typename VT>
__device__ __forceinline__ void clear_tail(VT& v, int nb) {
  jscc_u4v u = __builtin_bit_cast(jscc_u4v, v);
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const int k = nb - 4 * d;
    u[d] &= k >= 4 ? 0xffffffffu : (k <= 0 ? 0u : (1u << (8 * k)) - 1u);
  }
  v = __builtin_bit_cast(VT, u);
}

// Issue the loads of rows [i0, i0 + R) of the staged chunk starting at c0.
// Buffer loads: the resource's byte range is the row (or empty for slots at or
// past c1), so out-of-range lanes and dead slots read zero without a branch or
// a memory access, and the loads stay unconditional for the wait counters.
template <int DT, int NCH, int R>
__device__ __forceinline__ void fetch_rows(const ProjArgs& a, const RowMeta& meta, int64_t i0, int64_t c0,
                                           int64_t c1, int nvec, int tid, RowBuf<DT, NCH, R>& b) {
  using VT = typename Vec<DT>::T;
  constexpr int EB = DT == JSCC_DT_F32 ? 4 : (DT == JSCC_DT_FP8 ? 1 : 2);
  // The load offsets pass through an empty volatile asm, which stays ordered
  // after the one closing the previous block's accumulation: the next
  // buffer's loads cannot be hoisted above the current rows' arithmetic (which
  // would keep three buffers live).
  int voff[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    voff[c] = (c * kProjThreads + tid) * 16;
    asm volatile("" : "+v"(voff[c]));
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int64_t i = i0 + q;
    const bool live = i < c1;
    const int64_t off = meta.off[live ? (int)(i - c0) : 0];
    const char* rowp = reinterpret_cast<const char*>(a.M) + off * EB;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(rowp), (short)0, live ? nvec * 16 : 0, kBufWord3);
#pragma unroll
    for (int c = 0; c < NCH; ++c)
      b.raw[q][c] = __builtin_bit_cast(VT, __builtin_amdgcn_raw_buffer_load_b128(rs, voff[c], 0, kBufNT));
  }
}

// One block of R rows: dot products with the LDS image, work-group reduction,
// ratio / coefficient, and accumulation of coef * row into the registers.
// TAIL (a column count that is not a multiple of the vector width, modes with a dot product): tail_c is the chunk
// holding the rows' ragged last vector (uniform), tail_nb the bytes of it this thread keeps.  A separate
// instantiation, so the loop of every other launch carries none of it.
template <int DT, int NCH, int R, int MODE, bool TAIL>
__device__ __forceinline__ void process_rows(const ProjArgs& a, const RowMeta& meta, int64_t i0, int64_t c0,
                                             int64_t c1, int tid, RowBuf<DT, NCH, R>& b, const jscc_f4v* xs4,
                                             float (&red)[2][R][kProjWaves], int& buf,
                                             float (&wr)[NCH][Vec<DT>::N], int tail_c, int tail_nb) {
  constexpr int VN = Vec<DT>::N;
  float coefq[R];
  if constexpr (TAIL) {
    // every chunk is masked (all ones outside the tail chunk): a test of c against tail_c would let the
    // compiler index the row buffer dynamically, i.e. move it to scratch memory.  The masks are formed here
    // for each buffer (from an opaque copy of tail_nb), not kept live across the loop.
    int nb = tail_nb;
    asm volatile("" : "+v"(nb));
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int nbc = NCH == 1 || c == tail_c ? nb : 16;  // one chunk: it is the tail's
#pragma unroll
      for (int q = 0; q < R; ++q) clear_tail(b.raw[q][c], nbc);
    }
  }
  if (MODE == MODE_RATIO || MODE == MODE_FWD) {
    float d[R];
#pragma unroll
    for (int q = 0; q < R; ++q) d[q] = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
      for (int kk = 0; kk < VN / 4; ++kk) {
        const jscc_f4v t = xs4[(c * (VN / 4) + kk) * kProjThreads + tid];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          float v[VN];
          decode<DT>(b.raw[q][c], v);
          d[q] = fmaf(v[4 * kk], t.x, d[q]);
          d[q] = fmaf(v[4 * kk + 1], t.y, d[q]);
          d[q] = fmaf(v[4 * kk + 2], t.z, d[q]);
          d[q] = fmaf(v[4 * kk + 3], t.w, d[q]);
        }
      }
    }
    // wave sums (DPP, result uniform) -> LDS -> one lane per row finishes the
    // work-group sum and its coefficient, which is then read back as a scalar
#pragma unroll
    for (int q = 0; q < R; ++q) d[q] = jscc_wave_sum_dpp(d[q]);
    if (jscc_lane() == 0) {
#pragma unroll
      for (int q = 0; q < R; ++q) red[buf][q][jscc_wave_id()] = d[q];
    }
    __syncthreads();
  }
  const int lane = jscc_lane();
  float cl = 0.0f;
  if (lane < R) {
    const int64_t i = i0 + lane;
    const int m = i < c1 ? (int)(i - c0) : 0;
    const float sc = meta.sc[m];
    if (MODE == MODE_RATIO || MODE == MODE_FWD) {
      const jscc_f4v* rp = reinterpret_cast<const jscc_f4v*>(&red[buf][lane][0]);
      jscc_f4v acc = rp[0];
#pragma unroll
      for (int w4 = 1; w4 < kProjWaves / 4; ++w4) acc += rp[w4];
      const float f = ((acc.x + acc.y) + (acc.z + acc.w)) * sc;
      if (MODE == MODE_FWD) {
        if (jscc_wave_id() == 0 && i < c1) a.fwd_out[i] = f;
      } else {
        const float den = a.use_eps ? fmaxf(f, a.eps) : f;
        cl = (i < c1) ? sc * (meta.num[m] / den) : 0.0f;
      }
    } else {  // MODE_COEF, MODE_ONES: num holds the coefficient (1 for ONES)
      cl = (i < c1) ? sc * meta.num[m] : 0.0f;
    }
  }
  if (MODE == MODE_RATIO || MODE == MODE_FWD) buf ^= 1;
#pragma unroll
  for (int q = 0; q < R; ++q) coefq[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cl), q));
  if (MODE != MODE_FWD) {
    if (DT != JSCC_DT_F32) {
      // Packed rows are decoded a second time after the reduction instead of
      // keeping the decoded floats live across it (halves the row registers).
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < NCH; ++c) asm volatile("" : "+v"(b.raw[q][c]));
    }
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        float v[VN];
        decode<DT>(b.raw[q][c], v);
#pragma unroll
        for (int k = 0; k < VN; ++k) wr[c][k] = fmaf(coefq[q], v[k], wr[c][k]);
      }
    // the accumulation completes here (see fetch_rows)
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int k = 0; k < VN; ++k) asm volatile("" : "+v"(wr[c][k]));
  }
}

// The pointers are separate __restrict__ parameters (not struct members) so
// the compiler knows the inputs are not written by this kernel.
template <int DT, int NCH, int R, int MODE, bool TAIL>
__global__ void __launch_bounds__(kProjThreads)
    k_proj_fused(const void* __restrict__ M, const int32_t* __restrict__ rows, const float* __restrict__ row_scale,
                 const float* __restrict__ x, const int64_t* __restrict__ x_perm, const float* __restrict__ y,
                 const float* __restrict__ coef, float* __restrict__ fwd_out, float* __restrict__ partial,
                 int32_t* __restrict__ err, ProjArgs a) {
  a.M = M; a.rows = rows; a.row_scale = row_scale; a.x = x; a.x_perm = x_perm; a.y = y; a.coef = coef;
  a.fwd_out = fwd_out; a.partial = partial; a.err = err;
  if (a.cnt) {
    int64_t n = *a.cnt;
    n = n < 0 ? 0 : (n > a.cnt_cap ? a.cnt_cap : n);
    const int64_t lo = n * a.sub_i / a.sub_k, hi = n * (a.sub_i + 1) / a.sub_k;
    a.row_base += lo;
    a.n_rows = hi - lo;
    a.n_rows1 = a.n_rows;
  }
  constexpr int VN = Vec<DT>::N;
  constexpr int EB = DT == JSCC_DT_F32 ? 4 : (DT == JSCC_DT_FP8 ? 1 : 2);
  const int tid = threadIdx.x;
  const int nvec = (a.P + VN - 1) / VN;
  const int last_cols = a.P - (nvec - 1) * VN;  // columns of the last vector
  const int tail_c = TAIL ? (nvec - 1) / kProjThreads : -1;
  const int tail_nb = TAIL && tid == (nvec - 1) % kProjThreads ? last_cols * EB : 16;

  // this work-group's contiguous share of the rows
  const int64_t per = a.n_rows / gridDim.x, rem = a.n_rows % gridDim.x;
  const int64_t r_begin = blockIdx.x * per + (blockIdx.x < rem ? blockIdx.x : rem);
  const int64_t r_end = r_begin + per + (blockIdx.x < rem ? 1 : 0);

  __shared__ RowMeta meta;
  __shared__ __align__(16) float red[2][R][kProjWaves];
  int buf = 0;
  // Row metadata of a chunk into LDS (ids, scales, numerators).
  auto stage = [&](int64_t c0, int64_t c1) {
    __syncthreads();  // the previous chunk's metadata is no longer read
    if (tid < c1 - c0) {
      const int64_t i = c0 + tid;
      float num = 1.0f;
      if (i < a.n_rows1) {
        const int64_t sid = rows ? (int64_t)rows[a.row_base + i] : a.row_base + i;
        meta.off[tid] = sid * a.ld;
        meta.sc[tid] = row_scale ? row_scale[sid] : 1.0f;
        if (MODE == MODE_COEF) {
          num = coef[i];
        } else if (MODE == MODE_RATIO && y) {
          num = y[(a.y_by_row ? sid : i) * a.y_stride];
        }
        num *= a.num_scale1;
      } else {  // second segment of a joint launch (ratio mode)
        const ProjSeg& g = a.seg2;
        const int64_t j = i - a.n_rows1;
        const int64_t sid = g.rows ? (int64_t)g.rows[g.row_base + j] : g.row_base + j;
        meta.off[tid] = g.off + sid * g.ld;
        meta.sc[tid] = g.row_scale ? g.row_scale[sid] : 1.0f;
        if (g.y) num = g.y[(g.y_by_row ? sid : j) * g.y_stride];
        num *= g.num_scale;
      }
      meta.num[tid] = num;
    }
    __syncthreads();
  };

  // Two register buffers in ping-pong: the next R rows are in flight while
  // the current R rows are reduced and accumulated.  Every fetch is issued
  // (slots past the chunk read nothing), so the loop body is straight-line.
  // The first rows are requested before the image is loaded into LDS, so that
  // their HBM latency overlaps the image load.
  RowBuf<DT, NCH, R> ba, bb;
  if (r_begin < r_end) {
    stage(r_begin, r_begin + kStage < r_end ? r_begin + kStage : r_end);
    fetch_rows<DT, NCH, R>(a, meta, r_begin, r_begin, r_begin + kStage < r_end ? r_begin + kStage : r_end, nvec,
                           tid, ba);
  }
  // The image (gathered through the rotation) is kept in LDS as per-thread
  // storage: each thread only reads back the positions it wrote, in the same
  // chunk layout as the rows, so no barrier is needed and the registers go to
  // rows in flight instead.
  extern __shared__ jscc_f4v xs4[];
  float wr[NCH][VN];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int k = 0; k < VN; ++k) wr[c][k] = 0.0f;
  if (MODE == MODE_RATIO || MODE == MODE_FWD) {
    // All of this thread's image loads are issued before any is used (a
    // rotation gather first loads its indices, then the pixels).
    if (x_perm) {
      int32_t src[NCH][VN];
      int bad = 0;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int vi = c * kProjThreads + tid;
#pragma unroll
        for (int k = 0; k < VN; ++k) {
          if constexpr (TAIL) {  // per column: the last vector is cut by P (src -1: a zero pixel)
            const int64_t p = (int64_t)vi * VN + k;
            const bool in = p < a.P;
            const int64_t v = x_perm[(in ? p : 0) * a.x_perm_stride];
            const bool ok = v >= 1 && v <= a.x_len;
            bad |= (ok || !in) ? 0 : 1;
            src[c][k] = ok && in ? (int32_t)(v - 1) : -1;
          } else {  // per vector: P is a multiple of VN
            const int64_t p = (int64_t)(vi < nvec ? vi : 0) * VN + k;
            const int64_t v = x_perm[p * a.x_perm_stride];
            const bool ok = v >= 1 && v <= a.x_len;
            bad |= (ok || vi >= nvec) ? 0 : 1;
            src[c][k] = ok ? (int32_t)(v - 1) : 0;
          }
        }
      }
      if (bad) atomicOr(err, 2);
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int vi = c * kProjThreads + tid;
        float xv[VN];
#pragma unroll
        for (int k = 0; k < VN; ++k) xv[k] = (TAIL ? src[c][k] >= 0 : vi < nvec) ? x[src[c][k]] : 0.0f;
#pragma unroll
        for (int kk = 0; kk < VN / 4; ++kk) {
          jscc_f4v t;
          t.x = xv[4 * kk]; t.y = xv[4 * kk + 1]; t.z = xv[4 * kk + 2]; t.w = xv[4 * kk + 3];
          xs4[(c * (VN / 4) + kk) * kProjThreads + tid] = t;
        }
      }
    } else {
      // x is P contiguous floats, 16-byte aligned: whole quads below P, zeros past it
      const jscc_f4v* x4 = reinterpret_cast<const jscc_f4v*>(x);
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int vi = c * kProjThreads + tid;
#pragma unroll
        for (int kk = 0; kk < VN / 4; ++kk) {
          const int q = vi * (VN / 4) + kk;
          xs4[(c * (VN / 4) + kk) * kProjThreads + tid] =
              (TAIL ? (q + 1) * 4 <= a.P : vi < nvec) ? x4[q] : jscc_f4v(0.0f);
        }
      }
      if (TAIL && (a.P & 3)) {  // the quad that P cuts, read by element in the thread whose slot holds it
        const int q = a.P / 4, vi = q / (VN / 4);
        if (tid == vi % kProjThreads) {
          jscc_f4v t = jscc_f4v(0.0f);
          t.x = x[4 * q];
          if (4 * q + 1 < a.P) t.y = x[4 * q + 1];
          if (4 * q + 2 < a.P) t.z = x[4 * q + 2];
          xs4[((vi / kProjThreads) * (VN / 4) + q % (VN / 4)) * kProjThreads + tid] = t;
        }
      }
    }
  }


  for (int64_t c0 = r_begin; c0 < r_end; c0 += kStage) {
    const int64_t c1 = c0 + kStage < r_end ? c0 + kStage : r_end;
    if (c0 != r_begin) {
      stage(c0, c1);
      fetch_rows<DT, NCH, R>(a, meta, c0, c0, c1, nvec, tid, ba);
    }
    const int64_t n_steps = (c1 - c0 + R - 1) / R;
    int64_t i0 = c0;
    for (int64_t st = 0; st + 2 <= n_steps; st += 2) {
      fetch_rows<DT, NCH, R>(a, meta, i0 + R, c0, c1, nvec, tid, bb);
      process_rows<DT, NCH, R, MODE, TAIL>(a, meta, i0, c0, c1, tid, ba, xs4, red, buf, wr, tail_c, tail_nb);
      fetch_rows<DT, NCH, R>(a, meta, i0 + 2 * R, c0, c1, nvec, tid, ba);
      process_rows<DT, NCH, R, MODE, TAIL>(a, meta, i0 + R, c0, c1, tid, bb, xs4, red, buf, wr, tail_c, tail_nb);
      i0 += 2 * R;
    }
    if (n_steps & 1) process_rows<DT, NCH, R, MODE, TAIL>(a, meta, i0, c0, c1, tid, ba, xs4, red, buf, wr, tail_c, tail_nb);
  }

  if (MODE != MODE_FWD) {
    float* dst = partial + (int64_t)blockIdx.x * a.ldp;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int vi = c * kProjThreads + tid;
      if (vi < nvec) {
#pragma unroll
        for (int k = 0; k < VN; k += 4)
          *reinterpret_cast<float4*>(dst + (int64_t)vi * VN + k) =
              make_float4(wr[c][k], wr[c][k + 1], wr[c][k + 2], wr[c][k + 3]);
      }
    }
  }
}

// ============================================================================
// Persistent list-mode OSEM: every subset of every iteration in ONE
// cooperative launch (one work-group per CU, all resident).  Per subset:
//   A. the work-groups stream their share of the subset's rows exactly as
//      k_proj_fused does (ratio mode) and write their partial back-projection;
//      each then requests the first rows of the NEXT subset (rows do not depend
//      on the image), so their HBM latency hides behind the two barriers;
//   B. grid barrier; work-group g sums the partials of its own columns in a
//      fixed order and applies the EM update to those pixels of x;
//   C. grid barrier; every work-group reloads the updated image into LDS.
// The launch-to-launch gaps, per-launch ramp and the separate reduction
// launch of the one-launch-per-subset path disappear.  The grid barrier is a
// two-level counter tree plus a generation word with agent-scope
// release/acquire (the fences write back / invalidate the per-XCD L2 caches),
// and a bounded wait: a barrier that does not complete within spin_limit polls
// sets error bit 32 and every work-group leaves the kernel.
// ============================================================================
// Barrier words (unsigned, each on its own 128-byte line): [kBarGroups]
// group counters at i * kBarStride, the group-leader counter and the
// generation after them.  Work-group g arrives on counter g % kBarGroups; the
// last arrival of a group arrives on the leader counter, and the last leader
// advances the generation: at most G / kBarGroups + kBarGroups serialised
// atomics on any one word instead of G.
constexpr int kBarGroups = 8;
constexpr int kBarStride = 32;
constexpr int kBarLead = kBarGroups * kBarStride;
constexpr int kBarGen = kBarLead + kBarStride;
constexpr int kBarWords = kBarGen + kBarStride;

__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned n_groups, int64_t spin_limit, int32_t* err) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();  // release this group's writes device-wide
    const unsigned gen = __hip_atomic_load(bar + kBarGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned grp = blockIdx.x % kBarGroups;
    const unsigned in_grp = (n_groups - grp + kBarGroups - 1) / kBarGroups;  // groups g with g % kBarGroups == grp
    const unsigned leaders = n_groups < (unsigned)kBarGroups ? n_groups : (unsigned)kBarGroups;
    unsigned* cg = bar + grp * kBarStride;
    if (__hip_atomic_fetch_add(cg, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == in_grp) {
      __hip_atomic_store(cg, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__hip_atomic_fetch_add(bar + kBarLead, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u == leaders) {
        __hip_atomic_store(bar + kBarLead, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(bar + kBarGen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    int64_t spins = 0;
    // relaxed polling: an acquire load per poll would invalidate this XCD's L2 every time;
    // the fence below is the one acquire
    while (__hip_atomic_load(bar + kBarGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > spin_limit) {
        atomicOr(err, 32);
        ok = 0;
        break;
      }
    }
    __threadfence();  // acquire: what the other groups wrote is visible from here on
  }
  __syncthreads();
  return ok != 0;
}

struct OsemArgs {
  int64_t ld;
  int P;
  int n_sub;
  int n_iter;
  float eps;
  int use_eps;
  int em_mode;
  float em_eps;
  int64_t spin_limit;
};

// rows [lo, hi) of the kept list split evenly over the work-groups
__device__ __forceinline__ void share_of(int64_t lo, int64_t hi, int g, int G, int64_t* rb, int64_t* re) {
  const int64_t n = hi - lo, per = n / G, rem = n % G;
  *rb = lo + g * per + (g < rem ? g : rem);
  *re = *rb + per + (g < rem ? 1 : 0);
}

template <int DT, int NCH, int R>
__global__ void __launch_bounds__(kProjThreads)
    k_osem_persistent(const void* __restrict__ M, const int32_t* __restrict__ rows,
                      const float* __restrict__ row_scale, float* __restrict__ x, const float* __restrict__ sens,
                      float* __restrict__ partial, const int64_t* __restrict__ sub_off, unsigned* __restrict__ bar,
                      int32_t* __restrict__ err, OsemArgs o) {
  constexpr int VN = Vec<DT>::N;
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int G = gridDim.x;
  const int nvec = o.P / VN;
  ProjArgs a;
  a.M = M; a.ld = o.ld; a.P = o.P; a.ldp = o.P; a.rows = rows; a.n_rows = 0; a.row_base = 0; a.row_scale = row_scale;
  a.x = x; a.x_perm = nullptr; a.x_perm_stride = 0; a.x_len = o.P; a.y = nullptr; a.y_stride = 0; a.coef = nullptr;
  a.eps = o.eps; a.use_eps = o.use_eps; a.fwd_out = nullptr; a.partial = partial; a.err = err; a.y_by_row = 0;
  a.n_rows1 = 0; a.num_scale1 = 1.0f; a.seg2 = ProjSeg{};
  a.cnt = nullptr; a.cnt_cap = 0; a.sub_i = 0; a.sub_k = 1;

  extern __shared__ jscc_f4v xs4[];
  __shared__ RowMeta meta;
  __shared__ __align__(16) float red[2][R][kProjWaves];
  int buf = 0;
  auto stage = [&](int64_t c0, int64_t c1) {
    __syncthreads();
    if (tid < c1 - c0) {
      const int64_t sid = (int64_t)rows[c0 + tid];
      meta.off[tid] = sid * o.ld;
      meta.sc[tid] = row_scale ? row_scale[sid] : 1.0f;
      meta.num[tid] = 1.0f;
    }
    __syncthreads();
  };
  auto load_image = [&]() {
    const jscc_f4v* x4 = reinterpret_cast<const jscc_f4v*>(x);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const int vi = c * kProjThreads + tid;
#pragma unroll
      for (int kk = 0; kk < VN / 4; ++kk)
        xs4[(c * (VN / 4) + kk) * kProjThreads + tid] = vi < nvec ? x4[vi * (VN / 4) + kk] : jscc_f4v(0.0f);
    }
  };

  RowBuf<DT, NCH, R> ba, bb;
  float wr[NCH][VN];
  const int total = o.n_sub * o.n_iter;
  int64_t rb, re;
  share_of(sub_off[0], sub_off[1], g, G, &rb, &re);
  if (rb < re) {
    stage(rb, rb + kStage < re ? rb + kStage : re);
    fetch_rows<DT, NCH, R>(a, meta, rb, rb, rb + kStage < re ? rb + kStage : re, nvec, tid, ba);
  }
  load_image();
  for (int step = 0; step < total; ++step) {
    const int k = step % o.n_sub;
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int j = 0; j < VN; ++j) wr[c][j] = 0.0f;
    // ---- A: this group's rows of subset k (the first R are already in flight)
    for (int64_t c0 = rb; c0 < re; c0 += kStage) {
      const int64_t c1 = c0 + kStage < re ? c0 + kStage : re;
      if (c0 != rb) {
        stage(c0, c1);
        fetch_rows<DT, NCH, R>(a, meta, c0, c0, c1, nvec, tid, ba);
      }
      const int64_t n_steps = (c1 - c0 + R - 1) / R;
      int64_t i0 = c0;
      for (int64_t st = 0; st + 2 <= n_steps; st += 2) {
        fetch_rows<DT, NCH, R>(a, meta, i0 + R, c0, c1, nvec, tid, bb);
        process_rows<DT, NCH, R, MODE_RATIO, false>(a, meta, i0, c0, c1, tid, ba, xs4, red, buf, wr, -1, 16);
        fetch_rows<DT, NCH, R>(a, meta, i0 + 2 * R, c0, c1, nvec, tid, ba);
        process_rows<DT, NCH, R, MODE_RATIO, false>(a, meta, i0 + R, c0, c1, tid, bb, xs4, red, buf, wr, -1, 16);
        i0 += 2 * R;
      }
      if (n_steps & 1) process_rows<DT, NCH, R, MODE_RATIO, false>(a, meta, i0, c0, c1, tid, ba, xs4, red, buf, wr, -1, 16);
    }
    {
      float* dst = partial + (int64_t)g * o.P;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        const int vi = c * kProjThreads + tid;
        if (vi < nvec) {
#pragma unroll
          for (int j = 0; j < VN; j += 4)
            *reinterpret_cast<float4*>(dst + (int64_t)vi * VN + j) =
                make_float4(wr[c][j], wr[c][j + 1], wr[c][j + 2], wr[c][j + 3]);
        }
      }
    }
    // the next subset's first rows go out now
    if (step + 1 < total) {
      const int kn = (step + 1) % o.n_sub;
      share_of(sub_off[kn], sub_off[kn + 1], g, G, &rb, &re);
      if (rb < re) {
        stage(rb, rb + kStage < re ? rb + kStage : re);
        fetch_rows<DT, NCH, R>(a, meta, rb, rb, rb + kStage < re ? rb + kStage : re, nvec, tid, ba);
      }
    }
    if (!grid_sync(bar, (unsigned)G, o.spin_limit, err)) return;
    // ---- B: column sums of this group's pixels over all partials, then the EM update.
    // 1024 threads = 64 groups x 16 float4 columns; the image LDS is the scratch (reloaded after).
    {
      const int tx = tid & 15, ty = tid >> 4;
      const int cpw = ((o.P + G - 1) / G + 3) / 4 * 4;  // columns per work-group, multiple of 4
      const int col_begin = g * cpw;
      jscc_f4v* acc = xs4;  // [64][16]
      for (int cb = col_begin; cb < col_begin + cpw && cb < o.P; cb += 64) {
        const int col = cb + tx * 4;
        jscc_f4v sum = jscc_f4v(0.0f);
        if (col < o.P && col < col_begin + cpw) {
          for (int gg = ty; gg < G; gg += 64)
            sum += *reinterpret_cast<const jscc_f4v*>(partial + (int64_t)gg * o.P + col);
        }
        __syncthreads();
        acc[ty * 16 + tx] = sum;
        __syncthreads();
        if (ty == 0 && col < o.P && col < col_begin + cpw) {
          jscc_f4v t = acc[tx];
          for (int q = 1; q < 64; ++q) t += acc[q * 16 + tx];
          const float tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
          for (int j = 0; j < 4; ++j) x[col + j] = jscc_em_value(x[col + j], tv[j], sens[col + j], o.em_mode, o.em_eps);
        }
      }
    }
    if (!grid_sync(bar, (unsigned)G, o.spin_limit, err)) return;
    load_image();
  }
}

// partial [G, P] (row stride ldp) -> partial2 [ceil(G/16), P] (same stride): 16
// partial rows per output row, so that enough work-groups stream the partials at
// full rate
constexpr int kRedRows = 16;
__global__ void __launch_bounds__(256) k_reduce_stage1(const float* __restrict__ part, int G, int P, int64_t ldp,
                                                       float* __restrict__ out) {
  const int col = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (col >= P) return;
  const int g0 = blockIdx.y * kRedRows;
  const int g1 = min(G, g0 + kRedRows);
  float4 v[kRedRows];
#pragma unroll
  for (int k = 0; k < kRedRows; ++k)
    v[k] = g0 + k < g1 ? *reinterpret_cast<const float4*>(part + (int64_t)(g0 + k) * ldp + col)
                       : make_float4(0.f, 0.f, 0.f, 0.f);
  float4 s = v[0];
#pragma unroll
  for (int k = 1; k < kRedRows; ++k) {
    s.x += v[k].x; s.y += v[k].y; s.z += v[k].z; s.w += v[k].w;
  }
  *reinterpret_cast<float4*>(out + (int64_t)blockIdx.y * ldp + col) = s;
}

typedef JsccEm EmEpilogue;

// The weight combination and accumulation are rounded operation by operation, as the reference's separate
// tensor operations are ((2 - alpha) * weight_compton + alpha * weight_single, weight += ...): no contraction
// into fused multiply-adds, which the compiler could otherwise place differently in each kernel.
__device__ __forceinline__ void finish_weight(float sum, int64_t j, float scale, int accumulate, int sanitize,
                                              float* __restrict__ w, const EmEpilogue& em) {
#pragma clang fp contract(off)
  if (sanitize && !(sum == sum && fabsf(sum) != __builtin_inff())) sum = 0.0f;  // nan_to_num(0, 0, 0)
  const float add = scale == 1.0f ? sum : scale * sum;
  const float wj = accumulate ? w[j] + add : add;
  if (em.x) {
    // the weight combination of k_em_update: c1*w (+ c2*w2), same order
    float wc = wj;
    if (em.w2) {
      wc = em.c1 * wc;
      wc = wc + em.c2 * em.w2[j];
    } else if (em.c1 != 1.0f) {
      wc = em.c1 * wc;
    }
    em.x[j] = jscc_em_value(em.x[j], wc, em.s[j], em.mode, em.eps);
    w[j] = em.zero_w ? 0.0f : wj;
    if (em.w2 && em.zero_w) em.w2[j] = 0.0f;
  } else {
    w[j] = wj;
  }
}

// finish_weight with the operands (w, x, s, w2 of column j) already loaded (the k_reduce_direct form: they
// are loaded before the partial sums, so that their round trip overlaps the reduction's)
__device__ __forceinline__ void finish_weight_pre(float sum, int64_t j, float scale, int accumulate, int sanitize,
                                                  float* __restrict__ w, const EmEpilogue& em, float wv, float xv,
                                                  float sv, float w2v) {
#pragma clang fp contract(off)
  if (sanitize && !(sum == sum && fabsf(sum) != __builtin_inff())) sum = 0.0f;  // nan_to_num(0, 0, 0)
  const float add = scale == 1.0f ? sum : scale * sum;
  const float wj = accumulate ? wv + add : add;
  if (em.x) {
    float wc = wj;
    if (em.w2) {
      wc = em.c1 * wc;
      wc = wc + em.c2 * w2v;
    } else if (em.c1 != 1.0f) {
      wc = em.c1 * wc;
    }
    em.x[j] = jscc_em_value(xv, wc, sv, em.mode, em.eps);
    w[j] = em.zero_w ? 0.0f : wj;
    if (em.w2 && em.zero_w) em.w2[j] = 0.0f;
  } else {
    w[j] = wj;
  }
}

// w[j] = (accumulate ? w[j] : 0) + scale * sum_g part[g][src(j)],
// src(j) = inv_perm[j]-1 (the reference's index_select with RotMatInv)
__global__ void __launch_bounds__(256) k_reduce_stage2(const float* __restrict__ part, int G, int P, int64_t ldp,
                                                       const int64_t* __restrict__ inv_perm, int64_t ips,
                                                       float scale, int accumulate, int sanitize,
                                                       float* __restrict__ w, int64_t w_len,
                                                       int32_t* __restrict__ err, EmEpilogue em) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= w_len) return;
  int64_t src = j;
  if (inv_perm) {
    src = inv_perm[j * ips] - 1;
    if (src < 0 || src >= P) { atomicOr(err, 4); src = 0; }
  } else if (j >= P) {
    return;
  }
  float sum = 0.0f;
#pragma unroll 8
  for (int g = 0; g < G; ++g) sum += part[(int64_t)g * ldp + src];
  finish_weight(sum, j, scale, accumulate, sanitize, w, em);
}

// Without a rotation gather the columns are reduced in one pass: a block owns
// 64 columns; 16 row groups of 16 threads (one float4 column each) sum every
// 16th partial row, then the groups are combined in LDS in a fixed order.  The
// partial rows (stride ldp, a multiple of 4) extend past P to whole quads.
constexpr int kDirCols = 64;
__global__ void __launch_bounds__(256) k_reduce_direct(const float* __restrict__ part, int G, int P, int64_t ldp,
                                                       float scale, int accumulate, int sanitize,
                                                       float* __restrict__ w, EmEpilogue em) {
  constexpr int TX = kDirCols / 4, TY = 256 / TX;
  __shared__ jscc_f4v acc[TY][TX];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int col = blockIdx.x * kDirCols + tx * 4;
  // the operands of the update, loaded before the partial sums so that their round trip overlaps them
  float pw[4] = {0.f, 0.f, 0.f, 0.f}, px[4] = {0.f, 0.f, 0.f, 0.f}, ps[4] = {1.f, 1.f, 1.f, 1.f};
  float pw2[4] = {0.f, 0.f, 0.f, 0.f};
  if (ty == 0 && col < P) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (col + k < P) {
        if (accumulate) pw[k] = w[col + k];
        if (em.x) {
          px[k] = em.x[col + k];
          ps[k] = em.s[col + k];
          if (em.w2) pw2[k] = em.w2[col + k];
        }
      }
    }
  }
  jscc_f4v sum = jscc_f4v(0.0f);
  if (col < P) {
#pragma unroll 8
    for (int g = ty; g < G; g += TY) sum += *reinterpret_cast<const jscc_f4v*>(part + (int64_t)g * ldp + col);
  }
  acc[ty][tx] = sum;
  __syncthreads();
  if (ty == 0 && col < P) {
    jscc_f4v t = acc[0][tx];
#pragma unroll
    for (int k = 1; k < TY; ++k) t += acc[k][tx];
    finish_weight_pre(t.x, col, scale, accumulate, sanitize, w, em, pw[0], px[0], ps[0], pw2[0]);
    if (col + 1 < P) finish_weight_pre(t.y, col + 1, scale, accumulate, sanitize, w, em, pw[1], px[1], ps[1], pw2[1]);
    if (col + 2 < P) finish_weight_pre(t.z, col + 2, scale, accumulate, sanitize, w, em, pw[2], px[2], ps[2], pw2[2]);
    if (col + 3 < P) finish_weight_pre(t.w, col + 3, scale, accumulate, sanitize, w, em, pw[3], px[3], ps[3], pw2[3]);
  }
}

__global__ void __launch_bounds__(256) k_em_update(float* __restrict__ x, const float* __restrict__ w1, float c1,
                                                   const float* __restrict__ w2, float c2,
                                                   const float* __restrict__ s, int64_t n, int mode, float eps,
                                                   float* __restrict__ zero_after) {
#pragma clang fp contract(off)
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  // w = c1*w1 (+ c2*w2 when w2 is given), evaluated like the reference's
  // (2-alpha)*weight_compton + alpha*weight_single.
  float w = w1[j];
  if (w2) {
    w = c1 * w;
    w = w + c2 * w2[j];
  } else if (c1 != 1.0f) {
    w = c1 * w;
  }
  x[j] = jscc_em_value(x[j], w, s[j], mode, eps);
  if (zero_after) zero_after[j] = 0.0f;
}

// FP64 column sums of a row-major FP32 matrix: partial [gridDim.y, P]
__global__ void __launch_bounds__(256) k_colsum_f64(const float* __restrict__ A, int64_t M, int64_t lda, int64_t P,
                                                    int64_t rows_per, double* __restrict__ part) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= P) return;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < M ? r0 + rows_per : M;
  double s = 0.0;
  for (int64_t r = r0; r < r1; ++r) s += (double)A[r * lda + col];
  part[(int64_t)blockIdx.y * P + col] = s;
}

// sens[j] = (1/R) * sum_r colsum[inv[j, r]-1]   (main_plane.py:320-323)
__global__ void __launch_bounds__(256) k_sens_gather(const double* __restrict__ part, int G, int64_t P,
                                                     const int64_t* __restrict__ inv, int64_t inv_rs,
                                                     int64_t inv_cs, int R, float* __restrict__ out,
                                                     int32_t* __restrict__ err) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= P) return;
  double acc = 0.0;
  for (int r = 0; r < R; ++r) {
    int64_t src = inv ? inv[j * inv_rs + r * inv_cs] - 1 : j;
    if (src < 0 || src >= P) { atomicOr(err, 8); src = 0; }
    double cs = 0.0;
    for (int g = 0; g < G; ++g) cs += part[(int64_t)g * P + src];
    acc += cs;
  }
  out[j] = (float)(acc / (double)R);
}

template <int DT, int NCH, int R, int MODE, bool TAIL = false>
int launch_one(const ProjArgs& a, int G, hipStream_t s) {
  // image slots in LDS: NCH chunks x 512 threads x VN floats (only for the modes that read x)
  const size_t lds = (MODE == MODE_RATIO || MODE == MODE_FWD) ? (size_t)NCH * kProjThreads * Vec<DT>::N * 4 : 0;
  static bool attr_set = false;
  if (lds > 65536 && !attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_proj_fused<DT, NCH, R, MODE, TAIL>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_proj_fused<DT, NCH, R, MODE, TAIL>), dim3(G), dim3(kProjThreads), lds, s, a.M, a.rows, a.row_scale,
                     a.x, a.x_perm, a.y, a.coef, a.fwd_out, a.partial, a.err, a);
  return 0;
}

// Modes with a dot product over a ragged last vector take the TAIL instantiation; the column-sum and
// coefficient modes only carry the padding into partial columns past P, which the reduction never reads out.
// (The TAIL instantiations hold at most 4 rows per buffer: the masks of 8 spill registers.)
template <int DT, int NCH, int R>
int launch_fused(int mode, const ProjArgs& a, int G, hipStream_t s) {
  const bool tail = a.P % Vec<DT>::N != 0;
  constexpr int RT = R > 4 ? 4 : R;
  switch (mode) {
    case MODE_RATIO:
      return tail ? launch_one<DT, NCH, RT, MODE_RATIO, true>(a, G, s) : launch_one<DT, NCH, R, MODE_RATIO>(a, G, s);
    case MODE_FWD:
      return tail ? launch_one<DT, NCH, RT, MODE_FWD, true>(a, G, s) : launch_one<DT, NCH, R, MODE_FWD>(a, G, s);
    case MODE_COEF: return launch_one<DT, NCH, R, MODE_COEF>(a, G, s);
    case MODE_ONES: return launch_one<DT, NCH, R, MODE_ONES>(a, G, s);
    default: jscc_set_error("jscc_proj: bad mode %d", mode); return -2;
  }
}

// Rows per register buffer: two buffers of R rows x NCH 16-byte vectors take
// at most 64 VGPRs next to the NCH x VN accumulators (128 per lane in all);
// FP8 decodes 16 values per vector and gets half as many rows (one from two
// vectors per thread on: with two its decoded rows would spill).  From two
// vectors per thread on, a buffer holds 4 / NCH rows (64 KB in flight per
// 1024-thread work-group): with more the memory system is oversubscribed --
// 10 x 8 OSEM subsets over 164758 x 16384 took 8.76 ms with 2 BF16 rows per
// buffer against 9.00 ms with 4, and 16.40 against 16.72 ms with 1 FP32 row
// against 2 (192 work-groups); a streaming read peaks at 32-64 KB in flight
// per CU.
constexpr int rows_per_buf(int vn, int nch) {
  return nch * vn >= 32 ? 1 : (nch == 1 ? (vn == 16 ? 4 : 8) : (4 / nch > 1 ? 4 / nch : 1));
}

template <int DT>
int dispatch_nch(int mode, const ProjArgs& a, int G, hipStream_t s) {
  constexpr int VN = Vec<DT>::N;
  // chunk counts up to the widest row jscc_proj_max_cols admits (32768 columns: 8 FP32, 4 BF16 / FP16 and
  // 2 FP8 vectors per thread); no wider shape is instantiated
  constexpr int kMaxNch = 32768 / VN / kProjThreads;
  const int nch = ((a.P + VN - 1) / VN + kProjThreads - 1) / kProjThreads;
  if (nch <= 1) return launch_fused<DT, 1, rows_per_buf(VN, 1)>(mode, a, G, s);
  if (nch <= 2) return launch_fused<DT, 2, rows_per_buf(VN, 2)>(mode, a, G, s);
  if constexpr (kMaxNch >= 4) {
    if (nch <= 4) return launch_fused<DT, 4, rows_per_buf(VN, 4)>(mode, a, G, s);
  }
  if constexpr (kMaxNch >= 8) {
    if (nch <= 8) return launch_fused<DT, 8, 1>(mode, a, G, s);
  }
  jscc_set_error("jscc_proj: %d columns exceed the register-resident limit for this dtype", a.P);
  return -3;
}

__global__ void __launch_bounds__(256) k_zero_u32(unsigned* __restrict__ v, int n) {
  for (int i = threadIdx.x; i < n; i += 256) v[i] = 0u;
}

template <int DT, int NCH, int R>
int launch_persistent(const void* M, const int32_t* rows, const float* row_scale, float* x, const float* sens,
                      float* partial, const int64_t* sub_off, unsigned* bar, int32_t* err, OsemArgs o, int G,
                      hipStream_t s) {
  const size_t img = (size_t)NCH * kProjThreads * Vec<DT>::N * 4;
  const size_t lds = img > 16384 ? img : 16384;  // the column-sum scratch of phase B needs 16 KB
  const void* fn = reinterpret_cast<const void*>(&k_osem_persistent<DT, NCH, R>);
  if (lds > 65536) JSCC_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  JSCC_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kProjThreads, lds));
  JSCC_REQUIRE(per_cu >= 1, "jscc_osem_persistent: the kernel does not fit on a compute unit");
  void* args[] = {(void*)&M, (void*)&rows, (void*)&row_scale, (void*)&x, (void*)&sens, (void*)&partial,
                  (void*)&sub_off, (void*)&bar, (void*)&err, (void*)&o};
  JSCC_CHECK_HIP(hipLaunchCooperativeKernel(fn, dim3(G), dim3(kProjThreads), args, (unsigned)lds, s));
  return 0;
}

template <int DT>
int dispatch_persistent(const void* M, const int32_t* rows, const float* row_scale, float* x, const float* sens,
                        float* partial, const int64_t* sub_off, unsigned* bar, int32_t* err, OsemArgs o, int G,
                        hipStream_t s) {
  constexpr int VN = Vec<DT>::N;
  const int nch = (o.P / VN + kProjThreads - 1) / kProjThreads;
  // half the rows per buffer of k_proj_fused: the prefetched buffer stays live across the barriers
#define JSCC_PERSIST(N)                                                                                      \
  return launch_persistent<DT, N, (rows_per_buf(VN, N) > 1 ? rows_per_buf(VN, N) / 2 : 1)>(                  \
      M, rows, row_scale, x, sens, partial, sub_off, bar, err, o, G, s)
  if (nch <= 1) JSCC_PERSIST(1);
  if (nch <= 2) JSCC_PERSIST(2);
  if constexpr (32768 / VN / kProjThreads >= 4) {
    if (nch <= 4) JSCC_PERSIST(4);
  }
#undef JSCC_PERSIST
  jscc_set_error("jscc_osem_persistent: %d columns exceed the persistent kernel's limit for this dtype", o.P);
  return -3;
}

}  // namespace

extern "C" int jscc_proj_vec_width(int dtype) {
  return dtype == JSCC_DT_F32 ? 4 : ((dtype == JSCC_DT_BF16 || dtype == JSCC_DT_FP16) ? 8 : (dtype == JSCC_DT_FP8 ? 16 : -1));
}

extern "C" int jscc_proj_max_cols(int dtype) {
  const int vw = jscc_proj_vec_width(dtype);
  if (vw < 0) return -1;
  // the widest instantiated chunk count: register-resident accumulators of up
  // to 8 (FP8: 4) vectors per thread, and an image of the ratio / forward
  // modes in LDS (NCH x threads x vector x 4 bytes) within the 160 KB of a
  // work-group -- 32768 columns for every storage type
  int nch = dtype == JSCC_DT_FP8 ? 4 : 8;
  while (nch > 1 && (size_t)nch * kProjThreads * vw * 4 > 160 * 1024) nch /= 2;
  return nch * kProjThreads * vw;
}

// Work-groups of one projection launch: at most one per CU, each with at least
// rows_min rows.  Below a full grid (small subsets, rows resident in the
// Infinity Cache) a launch is latency-bound and fewer, longer work-groups
// write fewer partials: 8 rows per group measured 10.4 vs 10.9 us per
// 1024 x 4096 FP32 subset (FP8 rows, a quarter of the threads busy per row at
// 4096 columns, prefer 4).
static int proj_grid_rows(int64_t n_rows, int rows_min, bool all_cus, hipStream_t s) {
  const int all = jscc_stream_cus(s);
  const int q = all_cus ? all : (3 * all) / 4;
  const int cus = q > 0 ? q : 1;
  int64_t g = (n_rows + rows_min - 1) / rows_min;
  if (g > cus) g = cus;
  if (g < 1) g = 1;
  return (int)g;
}

// The grid of a launch over rows of `dtype`.  BF16 / FP16 / FP32 rows take at
// most 3/4 of the CUs: fewer concurrent row streams keep HBM more efficient --
// 10 x 8 OSEM subsets over 164758 BF16 rows x 16384 took 9.00 ms at 192
// work-groups against 9.47 ms at 256 (FP32: 16.8 vs 18.0 ms), the column-sum
// pass 0.82 vs 0.86 ms.  FP8 rows (half the bytes in flight per row) keep every
// CU: the JSCCSD FP8 iteration took 3.07 ms at 192 against 2.52 ms at 256.
static int proj_grid_for(int dtype, int64_t n_rows, hipStream_t s) {
  const bool fp8 = dtype == JSCC_DT_FP8;
  return proj_grid_rows(n_rows, fp8 ? 4 : 8, fp8, s);
}

// row stride of a projection launch's partial sums: P rounded up to the widest vector (16 columns), so that a
// ragged last vector of any storage type writes inside its row
static int64_t proj_ldp(int64_t P) { return (P + 15) / 16 * 16; }

// workspace bytes a launch of `dtype` over n_rows rows on stream s needs
static int64_t proj_need(int dtype, int64_t n_rows, int64_t P, hipStream_t s) {
  return jscc_partials_bytes(proj_grid_for(dtype, n_rows, s), proj_ldp(P));
}

extern "C" int jscc_proj_grid(int64_t n_rows) {
  return proj_grid_rows(n_rows, 4, true, nullptr);  // the most work-groups any dtype's launch uses
}

extern "C" int64_t jscc_proj_workspace_bytes(int64_t n_rows, int64_t P) {
  return jscc_partials_bytes(jscc_proj_grid(n_rows), proj_ldp(P));
}

// Fused projection of one block of rows.
//   mode 0: partial <- sum_i scale_i * row_i * y_i / max(scale_i * row_i . x_rot, eps)
//   mode 1: fwd_out[i] <- scale_i * row_i . x_rot
//   mode 2: partial <- sum_i scale_i * coef_i * row_i
//   mode 3: partial <- sum_i scale_i * row_i
// then (modes 0, 2, 3) w[j] (+)= out_scale * partial[inv_perm[j]-1].
// flags bit 0: the numerator y is indexed by the stored row id; bit 1: the
// block's weight goes through nan_to_num(0, 0, 0) before scaling.
static int proj_impl(int mode, const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                     int64_t n_rows, int64_t row_base, const float* row_scale, const float* x,
                     const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, const float* y,
                     int64_t y_stride, const float* coef, float eps, int use_eps, float* fwd_out,
                     const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate,
                     int flags, float* w, int64_t w_len, void* workspace, int64_t workspace_bytes,
                     int32_t* err, void* stream, const JsccEm& em, int64_t n_rows2 = 0,
                     const ProjSeg* seg2 = nullptr, float num_scale1 = 1.0f, const int32_t* cnt = nullptr,
                     int64_t cnt_cap = 0, int sub_i = 0, int sub_k = 1) {
  JSCC_REQUIRE(mode >= 0 && mode <= 3, "jscc_proj: bad mode");
  JSCC_REQUIRE(n_rows >= 0 && n_cols > 0, "jscc_proj: bad shape");
  JSCC_REQUIRE(err != nullptr, "jscc_proj: null error word");
  const int vw = jscc_proj_vec_width(dtype);
  JSCC_REQUIRE(vw > 0, "jscc_proj: bad dtype %d", dtype);
  JSCC_REQUIRE(ld % vw == 0 && ld >= (n_cols + vw - 1) / vw * vw,
               "jscc_proj: the row stride (%lld) must be a multiple of %d covering the %lld columns", (long long)ld,
               vw, (long long)n_cols);
  JSCC_REQUIRE(n_cols <= jscc_proj_max_cols(dtype), "jscc_proj: too many columns (%lld)", (long long)n_cols);
  JSCC_REQUIRE(dtype != JSCC_DT_FP8 || row_scale != nullptr, "jscc_proj: FP8 rows need a per-row scale");
  hipStream_t s = (hipStream_t)stream;
  if (mode != 1) {
    JSCC_REQUIRE(w != nullptr && w_len > 0, "jscc_proj: null weight output");
    JSCC_REQUIRE(inv_perm != nullptr || w_len == n_cols, "jscc_proj: weight length mismatch");
  }
  if (mode == 0 || mode == 1) JSCC_REQUIRE(x != nullptr && x_len > 0, "jscc_proj: null image");
  if (mode == 0 || mode == 1) JSCC_REQUIRE(x_perm != nullptr || x_len == n_cols, "jscc_proj: image length mismatch");
  if (mode == 1) JSCC_REQUIRE(fwd_out != nullptr, "jscc_proj: null forward output");
  if (mode == 2) JSCC_REQUIRE(coef != nullptr, "jscc_proj: null coefficients");

  const int64_t n_total = n_rows + n_rows2;  // rows of both segments of a joint launch
  if (n_total == 0) {
    if (mode != 1 && !accumulate) JSCC_CHECK_HIP(hipMemsetAsync(w, 0, w_len * sizeof(float), s));
    return 0;
  }
  JSCC_REQUIRE(mat != nullptr && ((uintptr_t)mat % 16) == 0, "jscc_proj: matrix must be 16-byte aligned");
  const int G = proj_grid_for(dtype, n_total, s);
  const int G2 = (G + kRedRows - 1) / kRedRows;
  if (mode != 1) {
    JSCC_REQUIRE(workspace != nullptr && workspace_bytes >= proj_need(dtype, n_total, n_cols, s),
                 "jscc_proj: workspace too small");
  }
  ProjArgs a;
  a.M = mat; a.ld = ld; a.P = (int)n_cols; a.ldp = (int)proj_ldp(n_cols); a.rows = rows; a.n_rows = n_total; a.row_base = row_base;
  a.n_rows1 = n_rows; a.num_scale1 = num_scale1; a.seg2 = seg2 ? *seg2 : ProjSeg{};
  a.row_scale = row_scale; a.x = x; a.x_perm = x_perm; a.x_perm_stride = x_perm_stride; a.x_len = x_len;
  a.y = y; a.y_stride = y_stride; a.coef = coef; a.eps = eps; a.use_eps = use_eps; a.fwd_out = fwd_out;
  a.partial = (float*)workspace; a.err = err; a.y_by_row = (flags & 1) ? 1 : 0;
  a.cnt = cnt; a.cnt_cap = cnt_cap; a.sub_i = sub_i; a.sub_k = sub_k;
  int rc;
  const int tag = mode == 3 ? JSCC_TAG_PROJ_SUM : JSCC_TAG_PROJ;
  jscc_prof_begin(tag, s);
  if (dtype == JSCC_DT_F32) rc = dispatch_nch<JSCC_DT_F32>(mode, a, G, s);
  else if (dtype == JSCC_DT_BF16) rc = dispatch_nch<JSCC_DT_BF16>(mode, a, G, s);
  else if (dtype == JSCC_DT_FP16) rc = dispatch_nch<JSCC_DT_FP16>(mode, a, G, s);
  else rc = dispatch_nch<JSCC_DT_FP8>(mode, a, G, s);
  jscc_prof_end(tag, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  if (mode == 1) return 0;

  return jscc_reduce_partials((float*)workspace, G, (int)n_cols, proj_ldp(n_cols), inv_perm, inv_perm_stride,
                              out_scale, accumulate, (flags & 2) ? 1 : 0, w, w_len, err, s, em);
}

extern "C" int jscc_proj(int mode, const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                         int64_t n_rows, int64_t row_base, const float* row_scale, const float* x,
                         const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, const float* y,
                         int64_t y_stride, const float* coef, float eps, int use_eps, float* fwd_out,
                         const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale, int accumulate,
                         int flags, float* w, int64_t w_len, void* workspace, int64_t workspace_bytes,
                         int32_t* err, void* stream) {
  const JsccEm none = {nullptr, nullptr, 0, 0.0f, 0, 1.0f, nullptr, 0.0f};
  return proj_impl(mode, mat, dtype, ld, n_cols, rows, n_rows, row_base, row_scale, x, x_perm, x_perm_stride, x_len,
                   y, y_stride, coef, eps, use_eps, fwd_out, inv_perm, inv_perm_stride, out_scale, accumulate, flags,
                   w, w_len, workspace, workspace_bytes, err, stream, none);
}

// One OSEM subset step: jscc_proj's ratio back-projection into w followed, in
// the same reduction pass, by the EM update of jscc_em_update on em_x with the
// weight em_c1*w (+ em_c2*em_w2) and sensitivity em_s; w (and em_w2) are
// zeroed afterwards when zero_w is set.  em_x may be the projected image x.
extern "C" int jscc_proj_em(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                            int64_t n_rows, int64_t row_base, const float* row_scale, const float* x,
                            const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, const float* y,
                            int64_t y_stride, float eps, int use_eps, const int64_t* inv_perm,
                            int64_t inv_perm_stride, float out_scale, int accumulate, int flags, float* w,
                            int64_t w_len, float* em_x, const float* em_s, int em_mode, float em_eps, float em_c1,
                            float* em_w2, float em_c2, int zero_w, void* workspace, int64_t workspace_bytes,
                            int32_t* err, void* stream) {
  JSCC_REQUIRE(em_x != nullptr && em_s != nullptr, "jscc_proj_em: null image or sensitivity");
  JSCC_REQUIRE(em_mode >= 0 && em_mode <= 2, "jscc_proj_em: bad EM mode %d", em_mode);
  JSCC_REQUIRE(n_rows > 0, "jscc_proj_em: empty subset (use jscc_em_update)");
  const JsccEm em = {em_x, em_s, em_mode, em_eps, zero_w, em_c1, em_w2, em_c2};
  return proj_impl(0, mat, dtype, ld, n_cols, rows, n_rows, row_base, row_scale, x, x_perm, x_perm_stride, x_len, y,
                   y_stride, nullptr, eps, use_eps, nullptr, inv_perm, inv_perm_stride, out_scale, accumulate, flags,
                   w, w_len, workspace, workspace_bytes, err, stream, em);
}

// ------------------------------------------------------------------ counted launches
// Subset sub_i of sub_k of the first n rows past row_base, n = *count (device
// int32, clamped to [0, count_cap]): the launch is sized by count_cap and the
// bounds are taken on the device, so it needs no host read of the count.
static int counted_check(const int32_t* count, int64_t count_cap, int sub_i, int sub_k, const float* y, int flags,
                         const char* what) {
  JSCC_REQUIRE(count != nullptr, "%s: null row count", what);
  JSCC_REQUIRE(count_cap > 0 && count_cap < ((int64_t)1 << 31), "%s: count_cap out of range", what);
  JSCC_REQUIRE(sub_k >= 1 && sub_i >= 0 && sub_i < sub_k, "%s: subset %d of %d", what, sub_i, sub_k);
  JSCC_REQUIRE(y == nullptr || (flags & 1), "%s: numerators must be indexed by the stored row", what);
  return 0;
}

extern "C" int jscc_proj_counted(int mode, const void* mat, int dtype, int64_t ld, int64_t n_cols,
                                 const int32_t* rows, const int32_t* count, int64_t count_cap, int sub_i, int sub_k,
                                 int64_t row_base, const float* row_scale, const float* x, const int64_t* x_perm,
                                 int64_t x_perm_stride, int64_t x_len, const float* y, int64_t y_stride, float eps,
                                 int use_eps, const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale,
                                 int accumulate, int flags, float* w, int64_t w_len, void* workspace,
                                 int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(mode == 0 || mode == 3, "jscc_proj_counted: ratio (0) or column-sum (3) mode");
  const int rc0 = counted_check(count, count_cap, sub_i, sub_k, y, flags, "jscc_proj_counted");
  if (rc0) return rc0;
  const JsccEm none = {nullptr, nullptr, 0, 0.0f, 0, 1.0f, nullptr, 0.0f};
  const int64_t n_max = (count_cap + sub_k - 1) / sub_k;
  return proj_impl(mode, mat, dtype, ld, n_cols, rows, n_max, row_base, row_scale, x, x_perm, x_perm_stride, x_len,
                   y, y_stride, nullptr, eps, use_eps, nullptr, inv_perm, inv_perm_stride, out_scale, accumulate,
                   flags, w, w_len, workspace, workspace_bytes, err, stream, none, 0, nullptr, 1.0f, count,
                   count_cap, sub_i, sub_k);
}

extern "C" int jscc_proj_em_counted(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                                    const int32_t* count, int64_t count_cap, int sub_i, int sub_k, int64_t row_base,
                                    const float* row_scale, const float* x, const int64_t* x_perm,
                                    int64_t x_perm_stride, int64_t x_len, const float* y, int64_t y_stride, float eps,
                                    int use_eps, const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale,
                                    int accumulate, int flags, float* w, int64_t w_len, float* em_x,
                                    const float* em_s, int em_mode, float em_eps, float em_c1, float* em_w2,
                                    float em_c2, int zero_w, void* workspace, int64_t workspace_bytes, int32_t* err,
                                    void* stream) {
  JSCC_REQUIRE(em_x != nullptr && em_s != nullptr, "jscc_proj_em_counted: null image or sensitivity");
  JSCC_REQUIRE(em_mode >= 0 && em_mode <= 2, "jscc_proj_em_counted: bad EM mode %d", em_mode);
  const int rc0 = counted_check(count, count_cap, sub_i, sub_k, y, flags, "jscc_proj_em_counted");
  if (rc0) return rc0;
  const JsccEm em = {em_x, em_s, em_mode, em_eps, zero_w, em_c1, em_w2, em_c2};
  const int64_t n_max = (count_cap + sub_k - 1) / sub_k;
  return proj_impl(0, mat, dtype, ld, n_cols, rows, n_max, row_base, row_scale, x, x_perm, x_perm_stride, x_len, y,
                   y_stride, nullptr, eps, use_eps, nullptr, inv_perm, inv_perm_stride, out_scale, accumulate, flags,
                   w, w_len, workspace, workspace_bytes, err, stream, em, 0, nullptr, 1.0f, count, count_cap, sub_i,
                   sub_k);
}

// ------------------------------------------------------------------ allocation-free subset step
// The partial back-projections of jscc_proj_em_noalloc live in a file-scope
// device array (each GPU holds its own copy, at its own address, for the life
// of the process).  In host code `scratch` is only a stand-in for the symbol:
// its device address is looked up once per device with hipGetSymbolAddress
// and only that address is handed to kernels.  Every work-group writes its
// whole partial row and the reduction writes every row it reads, so nothing
// in the buffer needs clearing between calls.  Calls that overlap on one GPU
// (on different streams, or from different threads) share the buffer and race.
static __device__ __align__(256) unsigned char scratch[JSCC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// The scratch address on the device of `stream` (which must be the current device).
int noalloc_scratch(hipStream_t stream, void** out) {
  int dev = 0;
  JSCC_CHECK_HIP(hipGetDevice(&dev));
  JSCC_REQUIRE(dev >= 0 && dev < kMaxDevices, "jscc_proj_em_noalloc: device %d beyond the table of %d", dev,
               kMaxDevices);
  if (stream != nullptr) {
    int sdev = -1;
    JSCC_CHECK_HIP(hipStreamGetDevice(stream, &sdev));
    JSCC_REQUIRE(sdev == dev, "jscc_proj_em_noalloc: the stream belongs to device %d, the current device is %d",
                 sdev, dev);
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    JSCC_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    JSCC_REQUIRE(p != nullptr, "jscc_proj_em_noalloc: no device address for the scratch buffer");
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return 0;
}
}  // namespace

extern "C" int jscc_proj_em_noalloc(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                                    int64_t n_rows, int64_t row_base, const float* row_scale, const float* x,
                                    const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, const float* y,
                                    int64_t y_stride, float eps, int use_eps, const int64_t* inv_perm,
                                    int64_t inv_perm_stride, float out_scale, int accumulate, int flags, float* w,
                                    int64_t w_len, float* em_x, const float* em_s, int em_mode, float em_eps,
                                    float em_c1, float* em_w2, float em_c2, int zero_w, int32_t* err,
                                    void* stream) {
  JSCC_REQUIRE(n_rows > 0 && n_cols > 0, "jscc_proj_em_noalloc: empty subset or no columns");
  // the bytes this launch uses (jscc_proj_workspace_bytes bounds every dtype and stream from above)
  const int vw = jscc_proj_vec_width(dtype);
  JSCC_REQUIRE(vw > 0, "jscc_proj_em_noalloc: bad dtype %d", dtype);
  const int64_t need = proj_need(dtype, n_rows, n_cols, (hipStream_t)stream);
  JSCC_REQUIRE(need <= (int64_t)JSCC_NOALLOC_SCRATCH_BYTES,
               "jscc_proj_em_noalloc: %lld rows x %lld columns need %lld bytes of scratch, above the %lld bytes "
               "of the library's static buffer (use jscc_proj_em with a workspace)",
               (long long)n_rows, (long long)n_cols, (long long)need, (long long)JSCC_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  const int rc = noalloc_scratch((hipStream_t)stream, &ws);
  if (rc != 0) return rc;
  return jscc_proj_em(mat, dtype, ld, n_cols, rows, n_rows, row_base, row_scale, x, x_perm, x_perm_stride, x_len, y,
                      y_stride, eps, use_eps, inv_perm, inv_perm_stride, out_scale, accumulate, flags, w, w_len,
                      em_x, em_s, em_mode, em_eps, em_c1, em_w2, em_c2, zero_w, ws,
                      (int64_t)JSCC_NOALLOC_SCRATCH_BYTES, err, stream);
}

// One JSCC-SD subset step in one projection launch: the rows of the first
// matrix (numerators c1 * y1) and then those of the second (numerators
// c2 * y2, ones when y2 is null) are back-projected with their ratios into the
// same partial sums, so w (+)= c1 A1^T(y1 / A1 x) + c2 A2^T(y2 / A2 x), and
// the reduction applies the EM update of jscc_em_update on em_x with that
// weight and sensitivity em_s (w zeroed afterwards when zero_w is set).
extern "C" int jscc_proj_joint_em(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                                  int64_t n_rows, int64_t row_base, const float* row_scale, const float* y,
                                  int64_t y_stride, int y_by_row, float c1, const void* mat2, int64_t ld2,
                                  const int32_t* rows2, int64_t n_rows2, int64_t row_base2, const float* row_scale2,
                                  const float* y2, int64_t y2_stride, int y2_by_row, float c2, const float* x,
                                  float eps, int use_eps, int accumulate, float* w, int64_t w_len, float* em_x,
                                  const float* em_s, int em_mode, float em_eps, int zero_w, void* workspace,
                                  int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(em_x != nullptr && em_s != nullptr, "jscc_proj_joint_em: null image or sensitivity");
  JSCC_REQUIRE(em_mode >= 0 && em_mode <= 2, "jscc_proj_joint_em: bad EM mode %d", em_mode);
  JSCC_REQUIRE(n_rows >= 0 && n_rows2 >= 0 && n_rows + n_rows2 > 0,
               "jscc_proj_joint_em: empty subset (use jscc_em_update)");
  const int vw = jscc_proj_vec_width(dtype);
  JSCC_REQUIRE(vw > 0, "jscc_proj_joint_em: bad dtype %d", dtype);
  JSCC_REQUIRE(mat != nullptr && ((uintptr_t)mat % 16) == 0, "jscc_proj_joint_em: matrix must be 16-byte aligned");
  JSCC_REQUIRE(n_rows2 == 0 || (mat2 != nullptr && ((uintptr_t)mat2 % 16) == 0),
               "jscc_proj_joint_em: second matrix must be 16-byte aligned");
  JSCC_REQUIRE(n_rows2 == 0 || (ld2 % vw == 0 && ld2 >= (n_cols + vw - 1) / vw * vw),
               "jscc_proj_joint_em: second row stride (%lld) must be a multiple of %d covering the columns",
               (long long)ld2, vw);
  JSCC_REQUIRE(n_rows2 == 0 || dtype != JSCC_DT_FP8 || row_scale2 != nullptr,
               "jscc_proj_joint_em: FP8 rows need a per-row scale");
  // element offset of the second matrix from the first (both 16-byte aligned,
  // so a whole number of elements of any storage type)
  const int eb = dtype == JSCC_DT_F32 ? 4 : (dtype == JSCC_DT_FP8 ? 1 : 2);
  ProjSeg g{};
  if (n_rows2 > 0) {
    g.off = (int64_t)(reinterpret_cast<intptr_t>(mat2) - reinterpret_cast<intptr_t>(mat)) / eb;
    g.ld = ld2; g.rows = rows2; g.row_base = row_base2; g.row_scale = row_scale2; g.y = y2; g.y_stride = y2_stride;
    g.y_by_row = y2_by_row ? 1 : 0; g.num_scale = c2;
  }
  const JsccEm em = {em_x, em_s, em_mode, em_eps, zero_w, 1.0f, nullptr, 0.0f};
  if (n_rows == 0) {  // the second segment alone
    return proj_impl(0, mat2, dtype, ld2, n_cols, rows2, n_rows2, row_base2, row_scale2, x, nullptr, 0, n_cols, y2,
                     y2_stride, nullptr, eps, use_eps, nullptr, nullptr, 0, 1.0f, accumulate, y2_by_row ? 1 : 0, w,
                     w_len, workspace, workspace_bytes, err, stream, em, 0, nullptr, c2);
  }
  return proj_impl(0, mat, dtype, ld, n_cols, rows, n_rows, row_base, row_scale, x, nullptr, 0, n_cols, y, y_stride,
                   nullptr, eps, use_eps, nullptr, nullptr, 0, 1.0f, accumulate, y_by_row ? 1 : 0, w, w_len,
                   workspace, workspace_bytes, err, stream, em, n_rows2, n_rows2 > 0 ? &g : nullptr, c1);
}

// partial [G, P] with row stride ldp (followed by room for ceil(G/16) more rows) -> w, deterministic order
int jscc_reduce_partials(float* part, int G, int P, int64_t ldp, const int64_t* inv_perm, int64_t ips, float scale,
                         int accumulate, int sanitize, float* w, int64_t w_len, int32_t* err, hipStream_t s,
                         const JsccEm& em) {
  JSCC_REQUIRE(ldp % 4 == 0 && ldp >= (P + 3) / 4 * 4, "jscc_reduce_partials: partial row stride");
  if (!inv_perm) {
    JSCC_REQUIRE(w_len == P, "jscc_reduce_partials: weight length mismatch");
    hipLaunchKernelGGL(k_reduce_direct, dim3((unsigned)((P + kDirCols - 1) / kDirCols)), dim3(256), 0, s, part, G, P,
                       ldp, scale, accumulate, sanitize, w, em);
    JSCC_CHECK_LAUNCH();
    return 0;
  }
  int Gfin = G;
  if (G > kRedRows) {
    const int G2 = (G + kRedRows - 1) / kRedRows;
    float* part2 = part + (int64_t)G * ldp;
    dim3 grid1((unsigned)(((P + 3) / 4 + 255) / 256), (unsigned)G2);
    hipLaunchKernelGGL(k_reduce_stage1, grid1, dim3(256), 0, s, part, G, P, ldp, part2);
    JSCC_CHECK_LAUNCH();
    part = part2;
    Gfin = G2;
  }
  hipLaunchKernelGGL(k_reduce_stage2, dim3((unsigned)((w_len + 255) / 256)), dim3(256), 0, s, part, Gfin, P, ldp,
                     inv_perm, ips, scale, accumulate, sanitize, w, w_len, err, em);
  JSCC_CHECK_LAUNCH();
  return 0;
}

int64_t jscc_partials_bytes(int G, int64_t ldp) {
  const int64_t G2 = (G + kRedRows - 1) / kRedRows;
  return ((int64_t)G + G2) * ldp * (int64_t)sizeof(float);
}

// ------------------------------------------------------------------ FP64 sums
__global__ void __launch_bounds__(1024) k_sum_f64(const float* __restrict__ v, int64_t n, int64_t stride,
                                                  double* __restrict__ out) {
  __shared__ double red[1024 / JSCC_WAVE];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += (double)v[i * stride];
  s = jscc_wave_sum(s);
  if (jscc_lane() == 0) red[jscc_wave_id()] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 1024 / JSCC_WAVE; ++k) t += red[k];
    *out = t;
  }
}

extern "C" int jscc_sum_f64(const float* v, int64_t n, int64_t stride, double* out, void* stream) {
  JSCC_REQUIRE(n >= 0 && out != nullptr && (n == 0 || v != nullptr), "jscc_sum_f64: bad arguments");
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, (hipStream_t)stream, v, n, stride > 0 ? stride : 1, out);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int jscc_em_update(float* x, const float* w1, float c1, const float* w2, float c2, const float* s,
                              int64_t n, int mode, float eps, float* zero_after, void* stream) {
  JSCC_REQUIRE(n >= 0 && mode >= 0 && mode <= 2, "jscc_em_update: bad arguments");
  if (n == 0) return 0;
  JSCC_REQUIRE(x && w1 && s, "jscc_em_update: null pointer");
  hipLaunchKernelGGL(k_em_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, w1, c1, w2,
                     c2, s, n, mode, eps, zero_after);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int64_t jscc_sensitivity_workspace_bytes(int64_t M, int64_t P) {
  int64_t G = (M + 255) / 256;
  if (G > 64) G = 64;
  if (G < 1) G = 1;
  return G * P * (int64_t)sizeof(double);
}

// sens[j] = (1/R) * sum_r sum_i A[i, inv[j, r]-1]  (inv may be null: plain column sums, R = 1)
extern "C" int jscc_sensitivity(const float* A, int64_t M, int64_t lda, int64_t P, const int64_t* inv,
                                int64_t inv_row_stride, int64_t inv_col_stride, int R, float* out,
                                void* workspace, int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(M >= 0 && P > 0 && R >= 1 && lda >= P, "jscc_sensitivity: bad shape");
  JSCC_REQUIRE(A && out && err && workspace, "jscc_sensitivity: null pointer");
  JSCC_REQUIRE(workspace_bytes >= jscc_sensitivity_workspace_bytes(M, P), "jscc_sensitivity: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  int64_t G = (M + 255) / 256;
  if (G > 64) G = 64;
  if (G < 1) G = 1;
  const int64_t rows_per = (M + G - 1) / G;
  hipLaunchKernelGGL(k_colsum_f64, dim3((unsigned)((P + 255) / 256), (unsigned)G), dim3(256), 0, s, A, M, lda, P,
                     rows_per > 0 ? rows_per : 1, (double*)workspace);
  JSCC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sens_gather, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, (const double*)workspace,
                     (int)G, P, inv, inv_row_stride, inv_col_stride, R, out, err);
  JSCC_CHECK_LAUNCH();
  return 0;
}

extern "C" int64_t jscc_osem_persistent_workspace_bytes(int64_t n_cols) {
  return (int64_t)kBarWords * 4 + (int64_t)jscc_num_cus() * n_cols * (int64_t)sizeof(float);
}

extern "C" int jscc_osem_persistent(const void* mat, int dtype, int64_t ld, int64_t n_cols, const int32_t* rows,
                                    const float* row_scale, const int64_t* sub_off, int n_sub, int n_iter, float* x,
                                    const float* sens, float eps, int use_eps, int em_mode, float em_eps,
                                    void* workspace, int64_t workspace_bytes, int32_t* err, void* stream) {
  JSCC_REQUIRE(n_sub >= 1 && n_iter >= 0, "jscc_osem_persistent: bad subset / iteration count");
  JSCC_REQUIRE(mat && rows && sub_off && x && sens && err && workspace, "jscc_osem_persistent: null pointer");
  const int vw = jscc_proj_vec_width(dtype);
  JSCC_REQUIRE(vw > 0, "jscc_osem_persistent: bad dtype %d", dtype);
  JSCC_REQUIRE(n_cols > 0 && n_cols % vw == 0 && ld % vw == 0 && ld >= n_cols,
               "jscc_osem_persistent: columns and row stride must be multiples of %d", vw);
  JSCC_REQUIRE(dtype != JSCC_DT_FP8 || row_scale != nullptr, "jscc_osem_persistent: FP8 rows need a per-row scale");
  JSCC_REQUIRE(((uintptr_t)mat | (uintptr_t)x) % 16 == 0, "jscc_osem_persistent: matrix and image must be 16-byte aligned");
  JSCC_REQUIRE(em_mode >= 0 && em_mode <= 2, "jscc_osem_persistent: bad EM mode");
  JSCC_REQUIRE(workspace_bytes >= jscc_osem_persistent_workspace_bytes(n_cols), "jscc_osem_persistent: workspace too small");
  if (n_iter == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  unsigned* bar = (unsigned*)workspace;
  float* partial = (float*)((char*)workspace + kBarWords * 4);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, bar, kBarWords);
  JSCC_CHECK_LAUNCH();
  OsemArgs o;
  o.ld = ld; o.P = (int)n_cols; o.n_sub = n_sub; o.n_iter = n_iter; o.eps = eps; o.use_eps = use_eps;
  o.em_mode = em_mode; o.em_eps = em_eps; o.spin_limit = 4000000;  // ~0.1-1 s of polling: far above any phase
  const int G = jscc_stream_cus(s);  // every work-group resident on the CUs the stream may use
  int rc;
  jscc_prof_begin(JSCC_TAG_PROJ, s);
  if (dtype == JSCC_DT_F32) rc = dispatch_persistent<JSCC_DT_F32>(mat, rows, row_scale, x, sens, partial, sub_off, bar, err, o, G, s);
  else if (dtype == JSCC_DT_BF16) rc = dispatch_persistent<JSCC_DT_BF16>(mat, rows, row_scale, x, sens, partial, sub_off, bar, err, o, G, s);
  else if (dtype == JSCC_DT_FP16) rc = dispatch_persistent<JSCC_DT_FP16>(mat, rows, row_scale, x, sens, partial, sub_off, bar, err, o, G, s);
  else rc = dispatch_persistent<JSCC_DT_FP8>(mat, rows, row_scale, x, sens, partial, sub_off, bar, err, o, G, s);
  jscc_prof_end(JSCC_TAG_PROJ, s);
  return rc;
}
