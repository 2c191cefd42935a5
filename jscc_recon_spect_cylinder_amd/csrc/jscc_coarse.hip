// Fused list-mode weight of packed coarse event rows (the sparse production
// mode): each row [cpnum1, t_c(P_c)] is expanded to the fine grid with index
// arithmetic, multiplied by its first-hit system-matrix row and consumed by
// T^T(1/max(T x, eps)) in the same pass; no fine row is ever written.
//
//   expand: fine(zf, lf) = sum_a wz_a(zf) sum_b wl_b(lf) t_c[zc_a(zf), lc_b(lf)]
//           (z_expand (x) layer_expand, at most two taps per axis; the layer
//           sum first, then the z sum, as rows . layer_expand^T . z_expand^T in
//           compton_sparse_ops.py:331-342)
//   normalize = 1 (materialise variant, compton_sparse_ops.py:380-400 followed by
//     recon_osem_local_sparse_jsccsd_only.py:13-19):
//       t = nan_to_num(expand(t_c) * A[cp]) element-wise; rows whose sum S is not
//       finite and positive are dropped; t_norm = t / S;
//       w += t_norm / max(t_norm . x, eps)
//   normalize = 0 (grouped variant, recon_osem_plane_sparse.py:22-57, which equals
//     T^T(1/max(T x, eps)) for the unnormalised rows t = expand(t_c) * A[cp]).
//   A non-finite coarse value spreads over the whole fine row in the reference's
//   dense expansion (0 * inf = NaN): such a row is dropped (normalize = 1) or
//   turns the block's weight into NaN (normalize = 0), which the final
//   nan_to_num of the block weight then zeroes, as in the reference.  The
//   unnormalised products are not filtered element-wise (the grouped variant
//   has no per-element nan_to_num).
//
// Layout: one work-group of 1024 threads per compute unit streams a contiguous
// share of rows.  Thread t owns fine pixels j = t + k*1024 (k < K): their image
// values and weight accumulators live in registers.  The coarse row is staged
// in LDS (double-buffered: the next row's loads are in flight while the
// current row is consumed), the axis tap tables sit in LDS for the whole
// kernel.  The per-work-group weight goes to a partial row and is reduced by
// the shared partial reduction (with the rotation gather and nan_to_num).
#include "jscc_common.h"

namespace {

constexpr int kCoarseThreads = 1024;
constexpr int kCoarseWaves = kCoarseThreads / JSCC_WAVE;
constexpr int kRowSlots = 16;  // coarse values per thread and row: n_coarse <= 16384
constexpr int kMaxCoarse = kRowSlots * kCoarseThreads;

// gfx9 raw buffer resource word 3
constexpr int kBufWord3 = 0x00020000;
typedef __attribute__((address_space(3))) void* lds_vp;
typedef float jscc_f4v __attribute__((ext_vector_type(4)));
typedef int jscc_i4v __attribute__((ext_vector_type(4)));

struct AxisTap {
  int i0, i1;    // coarse indices (z: already multiplied by the coarse layer count)
  float w0, w1;  // interpolation weights (w1 = 0, i1 = i0 for a single tap)
};

__device__ __forceinline__ bool finite_f(float v) { return v == v && fabsf(v) != __builtin_inff(); }

template <int K, bool LID, bool ZU, bool NORM>
__global__ void __launch_bounds__(kCoarseThreads)
    k_coarse_weight(const float* __restrict__ packed, int64_t ldp, int64_t n_rows, int Pc,
                    const int32_t* __restrict__ z_idx, const float* __restrict__ z_w, int Zf, int Zc,
                    const int32_t* __restrict__ l_idx, const float* __restrict__ l_w, int Lf, int Lc,
                    const float* __restrict__ A, int64_t lda, int64_t M, const float* __restrict__ x,
                    const int64_t* __restrict__ x_perm, int64_t xps, int64_t x_len, float eps, int use_eps,
                    float* __restrict__ partial, int32_t* __restrict__ err, const int32_t* __restrict__ row_list,
                    const int32_t* __restrict__ row_count, int add_partial) {
  extern __shared__ float lds[];
  const int tid = threadIdx.x;
  const int P = Zf * Lf;
  const int pc_pad = (Pc + 3) & ~3;
  float* rowbuf = lds;                                                    // [2][pc_pad]
  AxisTap* zt = reinterpret_cast<AxisTap*>(lds + 2 * pc_pad);             // [Zf]
  AxisTap* lt = zt + Zf;                                                  // [Lf] (general layer map)
  float* red = reinterpret_cast<float*>(LID ? lt : lt + Lf);              // [3][3][kCoarseWaves], 16-byte aligned
  (void)kMaxCoarse;

  // ---- tap tables (indices checked once; a bad table sets bit 64 and reads index 0)
  int bad_tab = 0;
  for (int f = tid; f < Zf; f += kCoarseThreads) {
    int a0 = z_idx[2 * f], a1 = z_idx[2 * f + 1];
    if (a0 < 0 || a0 >= Zc || a1 < 0 || a1 >= Zc) { bad_tab = 1; a0 = a0 < 0 || a0 >= Zc ? 0 : a0; a1 = a1 < 0 || a1 >= Zc ? 0 : a1; }
    zt[f] = AxisTap{a0 * Lc, a1 * Lc, z_w[2 * f], z_w[2 * f + 1]};
  }
  if (!LID) {
    for (int f = tid; f < Lf; f += kCoarseThreads) {
      int a0 = l_idx[2 * f], a1 = l_idx[2 * f + 1];
      if (a0 < 0 || a0 >= Lc || a1 < 0 || a1 >= Lc) { bad_tab = 1; a0 = a0 < 0 || a0 >= Lc ? 0 : a0; a1 = a1 < 0 || a1 >= Lc ? 0 : a1; }
      lt[f] = AxisTap{a0, a1, l_w[2 * f], l_w[2 * f + 1]};
    }
  }
  if (bad_tab) atomicOr(err, 64);

  // ---- this work-group's rows (of the row list and its device count, when given)
  if (row_count) n_rows = *row_count;
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t per = n_rows / G, rem = n_rows % G;
  const int64_t rb = g * per + (g < rem ? g : rem);
  const int64_t re = rb + per + (g < rem ? 1 : 0);

  // ---- image (rotation gather) and accumulators
  float xr[K], wr[K], t[K];
  {
    int bad = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = tid + k * kCoarseThreads;
      float v = 0.0f;
      if (j < P) {
        int64_t src = j;
        if (x_perm) {
          const int64_t pv = x_perm[(int64_t)j * xps];
          if (pv < 1 || pv > x_len) { bad = 1; src = 0; } else { src = pv - 1; }
        }
        v = x[src];
      }
      xr[k] = v;
      wr[k] = 0.0f;
    }
    if (bad) atomicOr(err, 2);
  }
  const int zf0 = tid / Lf, lf0 = tid - (tid / Lf) * Lf;
  const int zstep = kCoarseThreads / Lf, lstep = kCoarseThreads - (kCoarseThreads / Lf) * Lf;
  // Wave-uniform z taps (64 | Lf): lane q < K of every wave holds those of the
  // wave's slot q: the coarse word offsets of both z taps at the slot's first
  // layer pixel (identity layer map) or of both z slices (general map, with the
  // slot's first layer pixel apart), 16 bits each (n_coarse <= 16384), and the weights.
  int tz_off = 0, tz_lf = 0;
  float tz_w0 = 0.0f, tz_w1 = 0.0f;
  if (ZU && jscc_lane() < K) {
    const int j0 = (tid & ~(JSCC_WAVE - 1)) + jscc_lane() * kCoarseThreads;
    int zf = j0 / Lf;
    const int lfw = j0 - zf * Lf;
    if (zf >= Zf) zf = Zf - 1;
    int a0 = z_idx[2 * zf], a1 = z_idx[2 * zf + 1];
    a0 = a0 < 0 ? 0 : (a0 >= Zc ? Zc - 1 : a0);
    a1 = a1 < 0 ? 0 : (a1 >= Zc ? Zc - 1 : a1);
    tz_w0 = z_w[2 * zf];
    tz_w1 = z_w[2 * zf + 1];
    const int base = LID ? lfw : 0;
    tz_off = (a0 * Lc + base) | ((a1 * Lc + base) << 16);
    tz_lf = lfw;
  }

  // ---- coarse rows go straight from memory into LDS (no registers): wave w
  // copies values [w*64 + m*1024, +64) of the row; lanes past the row end are
  // masked off
  const int lane = jscc_lane(), wv = jscc_wave_id();
  // (runtime loops: unrolled, their per-slot offsets would be hoisted into
  // registers held across the row loop)
  const int slots = (Pc + kCoarseThreads - 1) / kCoarseThreads;
  auto row_id = [&](int64_t r) -> int64_t { return row_list ? (int64_t)row_list[r] : r; };
  auto fetch = [&](int64_t r, int b) {
    const float* src = packed + row_id(r) * ldp + 1;
    float* dst = rowbuf + b * pc_pad;
#pragma unroll 1
    for (int m = 0; m < slots; ++m) {
      const int cb = wv * JSCC_WAVE + m * kCoarseThreads;
      if (cb + lane < Pc) __builtin_amdgcn_global_load_lds(src + cb + lane, (lds_vp)(dst + cb), 4, 0, 0);
    }
  };
  if (rb < re) fetch(rb, 0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  for (int64_t r = rb; r < re; ++r) {
    const int buf = (int)((r - rb) & 1);
    const float* c = rowbuf + buf * pc_pad;
    float* rd = red + buf * 3 * kCoarseWaves;  // wave sums, double-buffered: a fast wave may write the
                                               // next row's while a slow one still reads this row's
    // a non-finite coarse value anywhere in the row
    int bad = 0;
#pragma unroll 1
    for (int q = tid; q * 4 < Pc; q += kCoarseThreads) {
      const jscc_f4v v = reinterpret_cast<const jscc_f4v*>(c)[q];  // the padding past Pc is not checked
      bad |= finite_f(v.x) ? 0 : 1;
      bad |= (q * 4 + 1 >= Pc || finite_f(v.y)) ? 0 : 1;
      bad |= (q * 4 + 2 >= Pc || finite_f(v.z)) ? 0 : 1;
      bad |= (q * 4 + 3 >= Pc || finite_f(v.w)) ? 0 : 1;
    }
    if (r + 1 < re) fetch(r + 1, buf ^ 1);

    // first-hit row of A: clamp(round(cpnum1) - 1, 0, M - 1) (compton_sparse_ops.py:390-391);
    // buffer loads over the row's P values (offsets past P read zero).  The
    // offset base passes through an empty asm each row so that the K offsets
    // are formed at use instead of being held in registers across rows.
    const float cpf = rintf(packed[row_id(r) * ldp]) - 1.0f;
    int64_t cp = 0;
    if (cpf >= (float)(M - 1)) cp = M - 1;
    else if (cpf > 0.0f) cp = (int64_t)cpf;
    const __amdgpu_buffer_rsrc_t ars =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A + cp * lda), (short)0, P * 4, kBufWord3);
    int vb = tid * 4;
    asm volatile("" : "+v"(vb));
#pragma unroll
    for (int k = 0; k < K; ++k)
      t[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ars, vb + k * kCoarseThreads * 4, 0, 0));

    // phase 1: expand, weight by A, row sum and forward projection.  Branch
    // free: pixels past P read a clamped tap and have t = 0 (buffer range) and x = 0.
    // The normalised variant's element-wise nan_to_num is deferred: a
    // non-finite product makes the row sum non-finite, and only then (rare)
    // are the products filtered and summed again.
    float S = 0.0f, D = 0.0f;
    auto consume = [&](int k, float u) {
#pragma clang fp contract(off)  // S sums the rounded products, as the filtered pass does
      const float v = u * t[k];
      t[k] = v;
      S += v;
      D = fmaf(v, xr[k], D);
    };
    if (ZU) {
      // 64 | Lf: the 64 pixels of a wave's slot share one fine z slice; the
      // slot's z taps are wave-uniform and come from lane k of the wave's table
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int off = __builtin_amdgcn_readlane(tz_off, k);
        const float w0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tz_w0), k));
        const float w1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tz_w1), k));
        const int o0 = off & 0xffff, o1 = (int)((unsigned)off >> 16);
        float u;
        if (LID) {
          u = w0 * c[o0 + lane];
          u = fmaf(w1, c[o1 + lane], u);
        } else {
          const AxisTap ll = lt[__builtin_amdgcn_readlane(tz_lf, k) + lane];
          float g0 = ll.w0 * c[o0 + ll.i0];
          g0 = fmaf(ll.w1, c[o0 + ll.i1], g0);
          float g1 = ll.w0 * c[o1 + ll.i0];
          g1 = fmaf(ll.w1, c[o1 + ll.i1], g1);
          u = w0 * g0;
          u = fmaf(w1, g1, u);
        }
        consume(k, u);
      }
    } else {
      int zf = zf0, lf = lf0;
      asm volatile("" : "+v"(zf), "+v"(lf));  // pixel coordinates are stepped per row, not held per pixel
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const AxisTap zz = zt[zf < Zf ? zf : Zf - 1];
        float u;
        if (LID) {
          u = zz.w0 * c[zz.i0 + lf];
          u = fmaf(zz.w1, c[zz.i1 + lf], u);
        } else {
          const AxisTap ll = lt[lf];
          float g0 = ll.w0 * c[zz.i0 + ll.i0];
          g0 = fmaf(ll.w1, c[zz.i0 + ll.i1], g0);
          float g1 = ll.w0 * c[zz.i1 + ll.i0];
          g1 = fmaf(ll.w1, c[zz.i1 + ll.i1], g1);
          u = zz.w0 * g0;
          u = fmaf(zz.w1, g1, u);
        }
        consume(k, u);
        zf += zstep;
        lf += lstep;
        if (lf >= Lf) { lf -= Lf; zf += 1; }
      }
    }

    // work-group sums (fixed order) and the row's bad flag
    S = jscc_wave_sum_dpp(S);
    D = jscc_wave_sum_dpp(D);
    const int wbad = __builtin_amdgcn_ballot_w64(bad != 0) != 0;
    if (jscc_lane() == 0) {
      rd[jscc_wave_id()] = S;
      rd[kCoarseWaves + jscc_wave_id()] = D;
      rd[2 * kCoarseWaves + jscc_wave_id()] = wbad ? 1.0f : 0.0f;
    }
    // one barrier per row: after it the wave sums are complete, the next row
    // has landed in LDS (every wave waited for its own copies) and no wave
    // reads this row's buffer any more (the copy into it is issued next row)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    float St = 0.0f, Dt = 0.0f, Bt = 0.0f;
    auto block_sums = [&](const float* rs) {
      St = 0.0f; Dt = 0.0f; Bt = 0.0f;
      const jscc_f4v* red4 = reinterpret_cast<const jscc_f4v*>(rs);  // 16-byte reads, waves in order
#pragma unroll
      for (int q = 0; q < kCoarseWaves / 4; ++q) {
        const jscc_f4v a = red4[q], b = red4[kCoarseWaves / 4 + q], f = red4[kCoarseWaves / 2 + q];
        St += a.x; St += a.y; St += a.z; St += a.w;
        Dt += b.x; Dt += b.y; Dt += b.z; Dt += b.w;
        Bt += f.x + f.y + f.z + f.w;
      }
    };
    block_sums(rd);
    if (NORM && Bt == 0.0f && !finite_f(St)) {
      // some product is not finite (or the sum overflowed): nan_to_num the
      // products and sum again (the same decision in every thread: St is read
      // in the same order everywhere).  The third sums buffer is not the
      // next row's, which a fast wave may already be writing.
      float* rf = red + 2 * 3 * kCoarseWaves;
      S = 0.0f;
      D = 0.0f;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float v = finite_f(t[k]) ? t[k] : 0.0f;
        t[k] = v;
        S += v;
        D = fmaf(v, xr[k], D);
      }
      S = jscc_wave_sum_dpp(S);
      D = jscc_wave_sum_dpp(D);
      if (jscc_lane() == 0) {
        rf[jscc_wave_id()] = S;
        rf[kCoarseWaves + jscc_wave_id()] = D;
        rf[2 * kCoarseWaves + jscc_wave_id()] = 0.0f;
      }
      __syncthreads();
      block_sums(rf);
    }

    float cf;
    bool use = true;
    if (NORM) {
      if (Bt != 0.0f || !finite_f(St) || !(St > 0.0f)) {
        use = false;
        cf = 0.0f;
      } else {
        float den = Dt / St;
        if (use_eps && den < eps) den = eps;
        cf = (1.0f / den) / St;
      }
    } else if (Bt != 0.0f) {
      cf = __builtin_nanf("");
    } else {
      float den = Dt;
      if (use_eps && den < eps) den = eps;
      cf = 1.0f / den;
    }
    // phase 2: w += t * cf (an invalid normalised row contributes nothing)
    if (use) {
#pragma unroll
      for (int k = 0; k < K; ++k) wr[k] = fmaf(t[k], cf, wr[k]);
    }
  }

  float* out = partial + (int64_t)g * P;
  // the output offsets are formed here: left to the compiler, they are shared
  // with the prologue's and spilled across the row loop
  int jb = tid;
  asm volatile("" : "+v"(jb));
  if (add_partial) {
    // rows handed over by the grouped kernel: added to its partial of the same
    // work-group index (a group without rows leaves that partial as it is)
    if (rb < re) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int j = jb + k * kCoarseThreads;
        if (j < P) out[j] += wr[k];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = jb + k * kCoarseThreads;
      if (j < P) out[j] = wr[k];
    }
  }
}

// ============================================================================
// Prefetching per-row form (identity layer map, 256 | Lf, P <= 2048 * K4): the
// same weight as k_coarse_weight, arranged so that no row waits for memory it
// issues itself and every memory access moves 16 bytes per lane.  While row i
// is consumed, the copy of row i+1 into LDS and -- when row i+1 has another
// first-hit crystal -- the A row of that crystal into registers are in
// flight; both are waited for at the row's barrier.  They are issued a pixel
// step at a time inside the first pass over row i, so no wave stalls on a
// full queue of loads before it computes.  Rows of one crystal in a row share
// its A values (no reload); the crystal numbers of the work-group's rows sit
// in an LDS table (read without a memory wait).
//
// Thread t owns the pixel quads q = t + 512 k (pixels 4q .. 4q+3): image
// values, weight accumulators, the row's products and the A values of the
// next row in registers (16 K4 of its 256).  A wave's 256 pixels of one step
// share a fine z slice (256 | Lf), so their two z taps are wave-uniform: one
// broadcast LDS read per step from a per-wave tap table; the coarse values of
// a quad are two 16-byte LDS reads.  Memory -> LDS copies of 16 bytes per lane read the packed rows at
// any 4-byte alignment (the buffer range zeroes the words past the row), A
// rows are 16-byte buffer loads whose range is the row.
// ============================================================================
constexpr int kPfThreads = 512;
constexpr int kPfWaves = kPfThreads / JSCC_WAVE;
constexpr int kPfChunk = 4 * JSCC_WAVE;  // coarse words of one wave's 16-byte copy
constexpr int kPfMaxK4 = 13;             // pixel quads per thread: P <= 26624

// LDS words of the prefetching form: two row buffers of pc_pad words (a multiple of kPfChunk), the wave
// sums [3 buffers][3 sums][waves], the tap table [waves][K4] of int4, then the crystal table of the
// work-group's rows
__host__ __device__ constexpr int pf_pc_pad(int Pc) { return (Pc + kPfChunk - 1) / kPfChunk * kPfChunk; }
__host__ __device__ constexpr int pf_red_words() { return 3 * 3 * kPfWaves; }
// 16-byte copies each wave issues per row: the coarse row holds at most pf_copies(K4) * 2048 values
__host__ __device__ constexpr int pf_copies(int K4) { return (K4 + 1) / 2; }

template <int K4, bool NORM>
__global__ void __launch_bounds__(kPfThreads)
    k_coarse_pf(const float* __restrict__ packed, int64_t ldp, int64_t n_rows, int Pc,
                const int32_t* __restrict__ z_idx, const float* __restrict__ z_w, int Zf, int Zc, int Lf,
                const float* __restrict__ A, int64_t lda, int64_t M, const float* __restrict__ x,
                const int64_t* __restrict__ x_perm, int64_t xps, int64_t x_len, float eps, int use_eps,
                float* __restrict__ partial, int32_t* __restrict__ err, const int32_t* __restrict__ row_list) {
  extern __shared__ float lds[];
  const int tid = threadIdx.x, lane = jscc_lane(), wv = jscc_wave_id();
  const int P = Zf * Lf;
  const int pc_pad = pf_pc_pad(Pc);
  float* rowbuf = lds;                                                     // [2][pc_pad]
  float* red = lds + 2 * pc_pad;                                           // [3][3][kPfWaves]
  int4* taps = reinterpret_cast<int4*>(red + pf_red_words());             // [kPfWaves][K4]
  int32_t* cps = reinterpret_cast<int32_t*>(taps + kPfWaves * K4);       // [rows of this group]

  // ---- this work-group's rows and their first-hit crystals clamp(round(cpnum1) - 1, 0, M - 1)
  // (compton_sparse_ops.py:390-391)
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t per = n_rows / G, rem = n_rows % G;
  const int64_t rb = g * per + (g < rem ? g : rem);
  const int nr = (int)(per + (g < rem ? 1 : 0));
  auto row_id = [&](int i) -> int64_t { return row_list ? (int64_t)row_list[rb + i] : rb + i; };
  for (int i = tid; i < nr; i += kPfThreads) {
    const float cpf = rintf(packed[row_id(i) * ldp]) - 1.0f;
    int cp = 0;
    if (cpf >= (float)(M - 1)) cp = (int)(M - 1);
    else if (cpf > 0.0f) cp = (int)cpf;
    cps[i] = cp;
  }

  // ---- z tap table (indices checked once; a bad table sets bit 64 and reads a clamped index)
  {
    int bad_tab = 0;
    for (int f = tid; f < Zf; f += kPfThreads) {
      const int a0 = z_idx[2 * f], a1 = z_idx[2 * f + 1];
      if (a0 < 0 || a0 >= Zc || a1 < 0 || a1 >= Zc) bad_tab = 1;
    }
    if (bad_tab) atomicOr(err, 64);
  }
  // the taps of pixel step k of wave w, taps[w * K4 + k] = (coarse word offsets of both taps, their weights)
  for (int e = tid; e < kPfWaves * K4; e += kPfThreads) {
    const int w = e / K4, k = e - w * K4;
    const int j0 = 4 * (w * JSCC_WAVE + k * kPfThreads);
    int zf = j0 / Lf;
    const int lfw = j0 - zf * Lf;  // before the clamp: a step past P still reads inside the row
    if (zf >= Zf) zf = Zf - 1;
    int a0 = z_idx[2 * zf], a1 = z_idx[2 * zf + 1];
    a0 = a0 < 0 ? 0 : (a0 >= Zc ? Zc - 1 : a0);
    a1 = a1 < 0 ? 0 : (a1 >= Zc ? Zc - 1 : a1);
    taps[e] = make_int4((a0 * Lf + lfw) >> 2, (a1 * Lf + lfw) >> 2, __float_as_int(z_w[2 * zf]),
                        __float_as_int(z_w[2 * zf + 1]));
  }
  const int wvu = __builtin_amdgcn_readfirstlane(wv);

  // ---- image (rotation gather) and accumulators
  float xr[4 * K4], wr[4 * K4], t[4 * K4], an[4 * K4];
  {
    int bad = 0;
#pragma unroll
    for (int k = 0; k < 4 * K4; ++k) {
      const int j = 4 * (tid + (k >> 2) * kPfThreads) + (k & 3);
      float v = 0.0f;
      if (j < P) {
        int64_t src = j;
        if (x_perm) {
          const int64_t pv = x_perm[(int64_t)j * xps];
          if (pv < 1 || pv > x_len) { bad = 1; src = 0; } else { src = pv - 1; }
        }
        v = x[src];
      }
      xr[k] = v;
      wr[k] = 0.0f;
    }
    if (bad) atomicOr(err, 2);
  }

  // A values of crystal cp for pixel step k into an: one 16-byte buffer load whose range is the row's P
  // values (quads past P read zero)
  auto a_rsrc = [&](int cp) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A + (int64_t)cp * lda), (short)0, P * 4, kBufWord3);
  };
  // (the step's offset is the scalar offset, so all steps share one lane-offset register)
  const int lane16 = tid * 16;
  auto load_a = [&](__amdgpu_buffer_rsrc_t ars, int k) {
    const jscc_f4v v =
        __builtin_bit_cast(jscc_f4v, __builtin_amdgcn_raw_buffer_load_b128(ars, lane16, k * kPfThreads * 16, 0));
    an[4 * k] = v.x;
    an[4 * k + 1] = v.y;
    an[4 * k + 2] = v.z;
    an[4 * k + 3] = v.w;
  };
  // chunk ch (256 words) of coarse row i into LDS buffer b: a 16-byte buffer-to-LDS copy whose range is the
  // row's Pc values (the padding up to pc_pad reads zero).  Issued by inline asm: the compiler would wait
  // for every such copy before the next LDS read (it cannot tell the buffers apart); the row's barrier
  // waits for them instead.  The asm has no register outputs.
  const int n_chunks = pc_pad / kPfChunk;
  auto r_rsrc = [&](int i) {
    const uint64_t a = reinterpret_cast<uint64_t>(packed + row_id(i) * ldp + 1);
    jscc_i4v r;
    r.x = (int)(uint32_t)a;
    r.y = (int)((uint32_t)(a >> 32) & 0xffffu);  // stride 0
    r.z = Pc * 4;                                // range in bytes
    r.w = kBufWord3;
    return r;
  };
  auto copy_chunk = [&](const jscc_i4v& rrs, int b, int ch) {
    const uint32_t la = (uint32_t)(uintptr_t)(lds_vp)(rowbuf + b * pc_pad + ch * kPfChunk);
    asm volatile("buffer_load_dwordx4 %0, %1, %2 offen lds" ::"v"(lane * 16), "s"(rrs), "s"(ch * kPfChunk * 4),
                 "{m0}"(la)
                 : "memory");
  };

  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();  // the crystal table
  if (nr > 0) {
    const __amdgpu_buffer_rsrc_t ars = a_rsrc(cps[0]);
    const jscc_i4v rrs = r_rsrc(0);
#pragma unroll
    for (int k = 0; k < K4; ++k) load_a(ars, k);
    for (int ch = wvu; ch < n_chunks; ch += kPfWaves) copy_chunk(rrs, 0, ch);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();

  for (int i = 0; i < nr; ++i) {
    const int b = i & 1;
    // this row's A values become its products in place; the next row's A values go into an (a repeated
    // crystal keeps them there) and its coarse values into the other LDS buffer, a chunk per step of pass 1,
    // so that no wave stalls on a queue of copies before it computes
    const bool next = i + 1 < nr;  // the last row copies itself into the other buffer (no branch per step)
    const int cp = __builtin_amdgcn_readfirstlane(cps[i]);
    const int cpn = next ? __builtin_amdgcn_readfirstlane(cps[i + 1]) : cp;
    const bool need_a = cpn != cp;
    const jscc_i4v rrs = r_rsrc(next ? i + 1 : i);
#pragma unroll
    for (int k = 0; k < 4 * K4; ++k) t[k] = an[k];
    if (need_a) {
      const __amdgpu_buffer_rsrc_t ars = a_rsrc(cpn);
#pragma unroll
      for (int k = 0; k < K4; ++k) load_a(ars, k);
    }
    const float* c = rowbuf + b * pc_pad;
    float* rd = red + b * 3 * kPfWaves;  // wave sums, double-buffered

    // a non-finite coarse value anywhere in the row (0 * v is NaN for v = +-inf or NaN; the padding is zero)
    float zsum = 0.0f;
#pragma unroll 1
    for (int q = tid; q < pc_pad / 4; q += kPfThreads) {
      const jscc_f4v v = reinterpret_cast<const jscc_f4v*>(c)[q];
      zsum = fmaf(v.x, 0.0f, zsum);
      zsum = fmaf(v.y, 0.0f, zsum);
      zsum = fmaf(v.z, 0.0f, zsum);
      zsum = fmaf(v.w, 0.0f, zsum);
    }
    const int bad = zsum != 0.0f;

    // pass 1: products (expanded coarse values times the A row, rounded as stored), their row sum and the
    // forward projection
    float S = 0.0f, D = 0.0f;
    const jscc_f4v* c4 = reinterpret_cast<const jscc_f4v*>(c) + lane;  // this lane's quads (offsets are multiples of 4)
#pragma unroll
      for (int k = 0; k < K4; ++k) {
#pragma clang fp contract(off)
        const int4 tp = taps[wvu * K4 + k];  // one broadcast read
        const jscc_f4v c0 = c4[tp.x];
        const jscc_f4v c1 = c4[tp.y];
        const float w0 = __int_as_float(tp.z), w1 = __int_as_float(tp.w);
        const float cv0[4] = {c0.x, c0.y, c0.z, c0.w}, cv1[4] = {c1.x, c1.y, c1.z, c1.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float u = w0 * cv0[e];
          u = fmaf(w1, cv1[e], u);
          const float v = u * t[4 * k + e];
          t[4 * k + e] = v;
          S += v;
          D = fmaf(v, xr[4 * k + e], D);
        }
        // one copy per step of the first pf_copies(K4) steps, without a branch: the chunk index is clamped to
        // the row's last chunk (a repeated copy writes the same values again)
        if (k < pf_copies(K4)) copy_chunk(rrs, b ^ 1, min(wvu + k * kPfWaves, n_chunks - 1));
      }
    S = jscc_wave_sum_dpp(S);
    D = jscc_wave_sum_dpp(D);
    const int wbad = __builtin_amdgcn_ballot_w64(bad != 0) != 0;
    if (lane == 0) {
      rd[wv] = S;
      rd[kPfWaves + wv] = D;
      rd[2 * kPfWaves + wv] = wbad ? 1.0f : 0.0f;
    }
    // one barrier per row: after it the wave sums are complete, the next row's copy and A values have landed
    // (every wave waited for its own loads) and no wave reads this row's buffer any more
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    float St = 0.0f, Dt = 0.0f, Bt = 0.0f;
    auto block_sums = [&](const float* rs) {
      St = 0.0f; Dt = 0.0f; Bt = 0.0f;
#pragma unroll
      for (int q = 0; q < kPfWaves; ++q) {
        St += rs[q];
        Dt += rs[kPfWaves + q];
        Bt += rs[2 * kPfWaves + q];
      }
    };
    block_sums(rd);
    if (NORM && Bt == 0.0f && !finite_f(St)) {
      // some product is not finite (or the sum overflowed): nan_to_num the products and sum again (the
      // same decision in every thread); the third sums buffer is not the next row's
      float* rf = red + 2 * 3 * kPfWaves;
      S = 0.0f;
      D = 0.0f;
#pragma unroll
      for (int k = 0; k < 4 * K4; ++k) {
        const float v = finite_f(t[k]) ? t[k] : 0.0f;
        t[k] = v;
        S += v;
        D = fmaf(v, xr[k], D);
      }
      S = jscc_wave_sum_dpp(S);
      D = jscc_wave_sum_dpp(D);
      if (lane == 0) {
        rf[wv] = S;
        rf[kPfWaves + wv] = D;
        rf[2 * kPfWaves + wv] = 0.0f;
      }
      __syncthreads();
      block_sums(rf);
    }

    float cf;
    bool use = true;
    if (NORM) {
      if (Bt != 0.0f || !finite_f(St) || !(St > 0.0f)) {
        use = false;
        cf = 0.0f;
      } else {
        float den = Dt / St;
        if (use_eps && den < eps) den = eps;
        cf = (1.0f / den) / St;
      }
    } else if (Bt != 0.0f) {
      cf = __builtin_nanf("");
    } else {
      float den = Dt;
      if (use_eps && den < eps) den = eps;
      cf = 1.0f / den;
    }
    // pass 2: w += t * cf (an invalid normalised row contributes nothing)
    if (use) {
#pragma unroll
      for (int k = 0; k < 4 * K4; ++k) wr[k] = fmaf(t[k], cf, wr[k]);
    }
  }

  float* out = partial + (int64_t)g * P;
  int jb = 4 * tid;
  asm volatile("" : "+v"(jb));
#pragma unroll
  for (int k = 0; k < K4; ++k) {
    const int j = jb + k * 4 * kPfThreads;
    if (j < P) {
      jscc_f4v v;
      v.x = wr[4 * k];
      v.y = wr[4 * k + 1];
      v.z = wr[4 * k + 2];
      v.w = wr[4 * k + 3];
      *reinterpret_cast<jscc_f4v*>(out + j) = v;
    }
  }
}

// ============================================================================
// Grouped form (identity layer map).  The rows of one first-hit crystal c
// share A[c]; with the expand E = z_expand (x) I,
//   t_i . x = t_c,i . u_c,   u_c = E^T (A[c] * x)      (coarse, once per group)
//   S_i     = t_c,i . v_c,   v_c = E^T A[c]
//   w      += A[c] * E (sum over the group of coef_i t_c,i)
// -- the algebra of get_weight_compton_sparse (recon_osem_plane_sparse.py:
// 39-51) for both variants.  A[c] is read once per group instead of once per
// row, and the row loop streams coarse rows with two dot products and one
// update per value.  Runs of equal crystal numbers in the launch's row order
// form the groups (rows ordered by crystal give the longest runs).
//
// The materialised variant's element-wise nan_to_num of t = E(t_c) * A[c]
// matters only where a product is not finite: a row is taken here when its
// coarse values are finite (a non-finite value poisons the reference's whole
// fine row, which is then dropped), A[c] is finite and no product or partial
// sum can overflow (max|t_c| max|A[c]| max(1, max|x|) 4 P < FLT_MAX); every
// other row is appended to a device list and weighted exactly, one product at
// a time, by k_coarse_weight, which adds into the same partial sums.
//
// Layout: one 1024-thread work-group per CU over a contiguous share of rows;
// thread t owns fine pixels f = t + 1024 k (k < KF) and coarse pixels
// q = t + 1024 k (k < KC): coarse row values, u, v and the group's coarse
// weight in registers, its fine weight in registers.  One LDS buffer of P
// floats carries A*x and then A for the two adjoints of a group, and the
// group's coarse weight for its expansion.
// ============================================================================
constexpr int kGrpThreads = 1024;
constexpr int kGrpWaves = kGrpThreads / JSCC_WAVE;
constexpr float kFltMax = 3.402823466e38f;

struct GrpZTap {  // fine z slice: word offsets of its two coarse taps from the fine index, and weights
  int d0, d1;
  float w0, w1;
};
struct GrpZAdj {  // coarse z slice: word offset of one contributing fine slice from the coarse index, weight
  int d;
  float w;
};

// a work-group-uniform value in a scalar register
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float abs_or_inf(float v) {
  const float a = fabsf(v);
  return a <= kFltMax ? a : __builtin_inff();  // NaN -> inf
}

// wave maximum of non-negative values, uniform across the wave
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_max_nonneg(float v) {
  const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false);
  return fmaxf(v, __int_as_float(moved));
}
__device__ __forceinline__ float wave_max_nonneg(float v) {
  v = dpp_max_nonneg<0xB1, 0xf>(v);
  v = dpp_max_nonneg<0x4E, 0xf>(v);
  v = dpp_max_nonneg<0x141, 0xf>(v);
  v = dpp_max_nonneg<0x140, 0xf>(v);
  v = dpp_max_nonneg<0x142, 0xa>(v);
  v = dpp_max_nonneg<0x143, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__global__ void __launch_bounds__(256) k_grp_prepare(const float* __restrict__ x, const int64_t* __restrict__ x_perm,
                                                     int64_t xps, int64_t x_len, int P, float* __restrict__ x_rot,
                                                     int32_t* __restrict__ count, int32_t* __restrict__ err) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j == 0) *count = 0;
  if (x_perm == nullptr || j >= P) return;
  const int64_t pv = x_perm[(int64_t)j * xps];
  int64_t src = pv - 1;
  if (pv < 1 || pv > x_len) {
    atomicOr(err, 2);
    src = 0;
  }
  x_rot[j] = x[src];
}

// word offset of the reduction scratch behind the buffer and the tap tables (a multiple of 4 words)
__host__ __device__ __forceinline__ int grouped_red_offset(int P, int Zf, int Zc, int na) {
  const int tab_words = ((P + 3) & ~3) + Zf * (int)(sizeof(GrpZTap) / 4) + Zc * na * (int)(sizeof(GrpZAdj) / 4);
  return (tab_words + 3) & ~3;
}

// kGrpAdj contributions per coarse slice: linear z expansion at z stride 2 (the reference's default) gives
// at most 4 fine slices per coarse slice; other strides take the per-product kernel
constexpr int kGrpAdj = 4;

template <int KF, bool NORM>
__global__ void __launch_bounds__(kGrpThreads)
    k_coarse_grouped(const float* __restrict__ packed, int64_t ldp, const int32_t* __restrict__ rows,
                     int64_t n_rows, int L, int Zf, int Zc, const int32_t* __restrict__ z_idx,
                     const float* __restrict__ z_w, const int32_t* __restrict__ za_idx,
                     const float* __restrict__ za_w, const float* __restrict__ A, int64_t lda, int64_t M,
                     const float* __restrict__ x, float eps, int use_eps, float* __restrict__ partial,
                     int32_t* __restrict__ cx_list, int32_t* __restrict__ cx_count, int32_t* __restrict__ err) {
  constexpr int KC = (KF + 1) / 2;
  constexpr int na = kGrpAdj;
  extern __shared__ float lds[];
  const int tid = threadIdx.x, lane = jscc_lane(), wv = jscc_wave_id();
  const int P = Zf * L, Pc = Zc * L;
  float* F = lds;                                                       // [P]
  GrpZTap* zt = reinterpret_cast<GrpZTap*>(lds + ((P + 3) & ~3));       // [Zf]
  GrpZAdj* za = reinterpret_cast<GrpZAdj*>(zt + Zf);                    // [Zc][na]
  float* red = lds + grouped_red_offset(P, Zf, Zc, na);                 // [2][3][kGrpWaves], 16-byte aligned

  // ---- tap tables, as word offsets from the fine (or coarse) index: the fine
  // pixel (zf, l) reads coarse (zc_a, l) = f + (zc_a - zf) L, the coarse pixel
  // (zc, l) reads fine (zf_j, l) = q + (zf_j - zc) L
  int bad_tab = 0;
  for (int f = tid; f < Zf; f += kGrpThreads) {
    int a0 = z_idx[2 * f], a1 = z_idx[2 * f + 1];
    if (a0 < 0 || a0 >= Zc || a1 < 0 || a1 >= Zc) { bad_tab = 1; a0 = a1 = f < Zc ? f : 0; }
    zt[f] = GrpZTap{(a0 - f) * L, (a1 - f) * L, z_w[2 * f], z_w[2 * f + 1]};
  }
  for (int e = tid; e < Zc * na; e += kGrpThreads) {
    const int zc = e / na;
    int fj = za_idx[e];
    float wj = za_w[e];
    if (fj < 0 || fj >= Zf) { bad_tab = 1; fj = zc; wj = 0.0f; }
    za[e] = GrpZAdj{(fj - zc) * L, wj};
  }
  if (bad_tab) atomicOr(err, 64);

  // ---- rows of this work-group
  const int G = gridDim.x, g = blockIdx.x;
  const int64_t per = n_rows / G, rem = n_rows % G;
  const int64_t rb = g * per + (g < rem ? g : rem);
  const int64_t re = rb + per + (g < rem ? 1 : 0);

  // pixel slots: slot k of this thread is (z_k, l_k) with z_0 = tid / L, stepped by 1024 = zstep L + lstep
  const int z0 = tid / L, l0 = tid - z0 * L;
  const int zstep = kGrpThreads / L, lstep = kGrpThreads - zstep * L;

  // Buffer loads over one row each (the range check covers the vector offset, which holds the whole
  // offset: slots past the row's end read 0); the slot offsets and coordinates are formed at each use from
  // a base passed through an empty asm: held as loop invariants they would take 3 KF registers across
  // the row loop.
  const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(x), (short)0, P * 4,
                                                                       kBufWord3);
  auto ld = [&](__amdgpu_buffer_rsrc_t rs, int vbase, int k) -> float {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, vbase + k * kGrpThreads * 4, 0, 0));
  };
  auto pinned_base = [&]() -> int {
    int v = tid * 4;
    asm volatile("" : "+v"(v));
    return v;
  };

  // ---- the guard on products: max|x| over the image (every pixel is some thread's slot)
  float xm = 0.0f;
  {
    const int vb = pinned_base();
#pragma unroll
    for (int k = 0; k < KF; ++k) xm = fmaxf(xm, abs_or_inf(ld(xrs, vb, k)));
  }
  xm = wave_max_nonneg(xm);
  if (lane == 0) red[wv] = xm;
  __syncthreads();  // also publishes the tap tables
  float xmax = 0.0f;
#pragma unroll
  for (int q = 0; q < kGrpWaves; ++q) xmax = fmaxf(xmax, red[q]);
  // rows pass the guard when max|t_c| * max|A[c]| < gthr
  const float gthr = uniform_f(kFltMax / (4.0f * (float)P * fmaxf(1.0f, uniform_f(xmax))));
  int rbuf = 1;  // reduction scratch in use (the prologue used the first half of buffer 0)

  float w[KF], wc[KC], u[KC], v[KC];
#pragma unroll
  for (int k = 0; k < KF; ++k) w[k] = 0.0f;
#pragma unroll
  for (int k = 0; k < KC; ++k) { wc[k] = 0.0f; u[k] = 0.0f; v[k] = 0.0f; }

  // coarse row values by buffer loads over the row's Pc values (slots past the end read 0)
  auto row_of = [&](int64_t r) -> int64_t { return rows ? (int64_t)rows[r] : r; };
  auto load_row = [&](int64_t r, float (&tc)[KC]) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(packed + row_of(r) * ldp + 1), (short)0, Pc * 4, kBufWord3);
    const int vb = pinned_base();
#pragma unroll
    for (int k = 0; k < KC; ++k) tc[k] = ld(rs, vb, k);
  };
  auto crystal_of = [&](int64_t r) -> int64_t {
    const float cpf = rintf(packed[row_of(r) * ldp]) - 1.0f;  // clamp(round(cpnum1) - 1, 0, M - 1)
    if (cpf >= (float)(M - 1)) return M - 1;
    if (cpf > 0.0f) return (int64_t)cpf;
    return 0;
  };
  // out[k] = sum_j za[z_k][j].w * F[q_k + za[z_k][j].d] over the coarse slots
  auto adjoint = [&](float (&out)[KC]) {
    int z = z0, l = l0, qb = tid;
    asm volatile("" : "+v"(z), "+v"(l), "+v"(qb));
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int q = qb + k * kGrpThreads;
      float acc = 0.0f;
      if (q < Pc) {
        const GrpZAdj* t = za + z * na;
#pragma unroll
        for (int j = 0; j < na; ++j) acc = fmaf(t[j].w, F[q + t[j].d], acc);
      }
      out[k] = acc;
      z += zstep;
      l += lstep;
      if (l >= L) { l -= L; z += 1; }
    }
  };

  int64_t cur = -1;   // crystal of the open group
  float amax = 0.0f;  // max|A[cur]|, inf when A[cur] holds a non-finite value

  // u, v of crystal c through the LDS buffer (A * x, then A; A is read twice, the second time from L2)
  auto form = [&](int64_t c) {
    const __amdgpu_buffer_rsrc_t ars =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A + c * lda), (short)0, P * 4, kBufWord3);
    __syncthreads();  // no thread reads the buffer any more (the previous group's expansion)
    float am = 0.0f;
    int fb = tid;
    asm volatile("" : "+v"(fb));
    int vb = pinned_base();
#pragma unroll
    for (int k = 0; k < KF; ++k) {
      const float a = ld(ars, vb, k);
      am = fmaxf(am, abs_or_inf(a));
      const int f = fb + k * kGrpThreads;
      if (f < P) F[f] = a * ld(xrs, vb, k);
    }
    am = wave_max_nonneg(am);
    float* rd = red + rbuf * 3 * kGrpWaves;
    if (lane == 0) rd[wv] = am;
    __syncthreads();
    adjoint(u);
    amax = 0.0f;
#pragma unroll
    for (int q = 0; q < kGrpWaves; ++q) amax = fmaxf(amax, rd[q]);
    amax = uniform_f(amax);
    rbuf ^= 1;
    if (NORM) {  // the row sums (the grouped variant has no normalisation)
      __syncthreads();
      fb = tid;
      asm volatile("" : "+v"(fb));
      vb = pinned_base();
#pragma unroll
      for (int k = 0; k < KF; ++k) {
        const float a = ld(ars, vb, k);
        const int f = fb + k * kGrpThreads;
        if (f < P) F[f] = a;
      }
      __syncthreads();
      adjoint(v);
    }
  };

  // w += A[c] * E(wc) and wc = 0.  A group whose A row is not finite handed all its rows over (its wc is
  // zero) and adds nothing (A * 0 would be NaN where A is infinite).
  auto flush = [&](int64_t c) {
    if (NORM && amax == __builtin_inff()) return;
    // (every thread finished the group's reads of the buffer before the barrier of its first row)
    int qb = tid;
    asm volatile("" : "+v"(qb));
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int q = qb + k * kGrpThreads;
      if (q < Pc) F[q] = wc[k];
      wc[k] = 0.0f;
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t ars =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(A + c * lda), (short)0, P * 4, kBufWord3);
    int z = z0, l = l0, fb = tid;
    asm volatile("" : "+v"(z), "+v"(l), "+v"(fb));
    const int vb = pinned_base();
#pragma unroll
    for (int k = 0; k < KF; ++k) {
      const int f = fb + k * kGrpThreads;
      const float a = ld(ars, vb, k);
      if (f < P) {
        const GrpZTap t = zt[z];
        float e = t.w0 * F[f + t.d0];
        e = fmaf(t.w1, F[f + t.d1], e);
        w[k] = fmaf(a, e, w[k]);
      }
      z += zstep;
      l += lstep;
      if (l >= L) { l -= L; z += 1; }
    }
  };

  // one row: dot products, work-group sums (fixed order), coefficient, coarse weight update
  auto consume = [&](int64_t r, const float (&tc)[KC]) {
    const int64_t c = crystal_of(r);
    if (c != cur) {  // block-uniform: every thread reads the same row id
      if (cur >= 0) flush(cur);
      form(c);
      cur = c;
    }
    float D = 0.0f, S = 0.0f, m = 0.0f;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      D = fmaf(tc[k], u[k], D);
      if (NORM) S = fmaf(tc[k], v[k], S);
      m = fmaxf(m, abs_or_inf(tc[k]));
    }
    D = jscc_wave_sum_dpp(D);
    if (NORM) S = jscc_wave_sum_dpp(S);
    m = wave_max_nonneg(m);
    float* rd = red + rbuf * 3 * kGrpWaves;  // double-buffered: the next row's sums may land before slow
    if (lane == 0) {                           // waves have read these
      rd[wv] = D;
      rd[kGrpWaves + wv] = S;
      rd[2 * kGrpWaves + wv] = m;
    }
    __syncthreads();
    // the wave sums in a fixed order, one 16-byte group of each at a time (read all at once, they
    // would add 48 registers on top of the row buffers)
    float Dt = 0.0f, St = 0.0f, mt = 0.0f;
    const jscc_f4v* rd4 = reinterpret_cast<const jscc_f4v*>(rd);
#pragma unroll
    for (int q = 0; q < kGrpWaves / 4; ++q) {
      const jscc_f4v a = rd4[q], b = rd4[kGrpWaves / 4 + q], e = rd4[kGrpWaves / 2 + q];
      Dt += a.x; Dt += a.y; Dt += a.z; Dt += a.w;
      St += b.x; St += b.y; St += b.z; St += b.w;
      mt = fmaxf(fmaxf(mt, fmaxf(e.x, e.y)), fmaxf(e.z, e.w));
      asm volatile("" : "+v"(Dt), "+v"(St), "+v"(mt));
    }
    Dt = uniform_f(Dt);
    St = uniform_f(St);
    mt = uniform_f(mt);
    rbuf ^= 1;
    float coef = 0.0f;
    bool apply = true;  // (a dropped or handed-over row adds nothing: 0 * inf would be NaN)
    if (NORM) {
      if (mt == __builtin_inff()) {
        apply = false;  // a non-finite coarse value: the reference's row is all NaN, then zero, then dropped
      } else if (!(mt * amax < gthr)) {
        apply = false;  // weighted exactly by the per-product kernel
        if (tid == 0) cx_list[atomicAdd(cx_count, 1)] = (int32_t)row_of(r);
      } else if (!(St > 0.0f)) {
        apply = false;  // zero-sum row: dropped
      } else {
        float den = Dt / St;
        if (use_eps && den < eps) den = eps;
        coef = (1.0f / den) / St;
      }
    } else if (mt == __builtin_inff()) {
      coef = __builtin_nanf("");  // a non-finite coarse value turns the block weight into NaN
    } else {
      float den = Dt;
      if (use_eps && den < eps) den = eps;
      coef = 1.0f / den;
    }
    if (apply) {
#pragma unroll
      for (int k = 0; k < KC; ++k) wc[k] = fmaf(coef, tc[k], wc[k]);
    }
  };

  // ---- the next row is in flight while one is consumed; one call site of the row step (two would
  // duplicate the group code), the prefetched row moved into place by register copies after the step
  float ta[KC], tb[KC];
  if (rb < re) load_row(rb, ta);
#pragma unroll 1
  for (int64_t r = rb; r < re; ++r) {
    if (r + 1 < re) load_row(r + 1, tb);
    consume(r, ta);
#pragma unroll
    for (int k = 0; k < KC; ++k) ta[k] = tb[k];
  }
  if (cur >= 0) flush(cur);

  float* out = partial + (int64_t)g * P;
  int jb = tid;
  asm volatile("" : "+v"(jb));
#pragma unroll
  for (int k = 0; k < KF; ++k) {
    const int f = jb + k * kGrpThreads;
    if (f < P) out[f] = w[k];
  }
}

size_t grouped_lds(int P, int Zf, int Zc, int na) {
  return (size_t)grouped_red_offset(P, Zf, Zc, na) * 4 + 2 * 3 * kGrpWaves * sizeof(float);
}

template <int KF, bool NORM>
int launch_grouped(size_t lds, int G, hipStream_t s, const float* packed, int64_t ldp, const int32_t* rows,
                   int64_t n_rows, int L, int Zf, int Zc, const int32_t* z_idx, const float* z_w,
                   const int32_t* za_idx, const float* za_w, const float* A, int64_t lda, int64_t M, const float* x,
                   float eps, int use_eps, float* partial, int32_t* cx_list, int32_t* cx_count, int32_t* err) {
  static bool attr_set = false;
  if (!attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_grouped<KF, NORM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_coarse_grouped<KF, NORM>), dim3(G), dim3(kGrpThreads), lds, s, packed, ldp, rows, n_rows, L,
                     Zf, Zc, z_idx, z_w, za_idx, za_w, A, lda, M, x, eps, use_eps, partial, cx_list, cx_count, err);
  return 0;
}

// LDS bytes of the prefetching form with a crystal table of `rows` rows
size_t pf_lds(int Pc, int64_t rows) {
  return ((size_t)2 * pf_pc_pad(Pc) + pf_red_words() + (size_t)4 * kPfWaves * kPfMaxK4 + (size_t)rows) * 4;
}

template <int K4, bool NORM>
int launch_pf(size_t lds, int G, hipStream_t s, const float* packed, int64_t ldp, int64_t n_rows, int Pc,
              const int32_t* z_idx, const float* z_w, int Zf, int Zc, int Lf, const float* A, int64_t lda, int64_t M,
              const float* x, const int64_t* x_perm, int64_t xps, int64_t x_len, float eps, int use_eps,
              float* partial, int32_t* err, const int32_t* row_list) {
  static bool attr_set = false;
  if (!attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_pf<K4, NORM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_coarse_pf<K4, NORM>), dim3(G), dim3(kPfThreads), lds, s, packed, ldp, n_rows, Pc, z_idx, z_w,
                     Zf, Zc, Lf, A, lda, M, x, x_perm, xps, x_len, eps, use_eps, partial, err, row_list);
  return 0;
}

int coarse_grid(int64_t n_rows) {
  const int cus = jscc_num_cus();
  int64_t g = n_rows < cus ? n_rows : cus;
  return g < 1 ? 1 : (int)g;
}

size_t coarse_lds(int Pc, int Zf, int Lf, bool lid) {
  const size_t pc_pad = (size_t)((Pc + 3) & ~3);
  return 2 * pc_pad * 4 + (size_t)Zf * sizeof(AxisTap) + (lid ? 0 : (size_t)Lf * sizeof(AxisTap)) +
         3 * 3 * kCoarseWaves * sizeof(float);
}

template <int K, bool LID, bool ZU, bool NORM>
int launch_coarse(size_t lds, int G, hipStream_t s, const float* packed, int64_t ldp, int64_t n_rows, int Pc,
                  const int32_t* z_idx, const float* z_w, int Zf, int Zc, const int32_t* l_idx, const float* l_w,
                  int Lf, int Lc, const float* A, int64_t lda, int64_t M, const float* x, const int64_t* x_perm,
                  int64_t xps, int64_t x_len, float eps, int use_eps, float* partial, int32_t* err,
                  const int32_t* row_list, const int32_t* row_count, int add_partial) {
  static bool attr_set = false;
  if (!attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_weight<K, LID, ZU, NORM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_coarse_weight<K, LID, ZU, NORM>), dim3(G), dim3(kCoarseThreads), lds, s, packed, ldp, n_rows,
                     Pc, z_idx, z_w, Zf, Zc, l_idx, l_w, Lf, Lc, A, lda, M, x, x_perm, xps, x_len, eps, use_eps,
                     partial, err, row_list, row_count, add_partial);
  return 0;
}

// the template flags of one call: identity layer map, wave-uniform z slices (64 | Lf), normalised rows
template <int K>
int dispatch_coarse(bool lid, bool zu, bool norm, size_t lds, int G, hipStream_t s, const float* packed, int64_t ldp,
                    int64_t n_rows, int Pc, const int32_t* z_idx, const float* z_w, int Zf, int Zc,
                    const int32_t* l_idx, const float* l_w, int Lf, int Lc, const float* A, int64_t lda, int64_t M,
                    const float* x, const int64_t* x_perm, int64_t xps, int64_t x_len, float eps, int use_eps,
                    float* partial, int32_t* err, const int32_t* row_list, const int32_t* row_count,
                    int add_partial) {
#define JSCC_COARSE_ARGS                                                                                          \
  lds, G, s, packed, ldp, n_rows, Pc, z_idx, z_w, Zf, Zc, l_idx, l_w, Lf, Lc, A, lda, M, x, x_perm, xps, x_len, eps, \
      use_eps, partial, err, row_list, row_count, add_partial
  const int code = (lid ? 4 : 0) | (zu ? 2 : 0) | (norm ? 1 : 0);
  switch (code) {
    case 0: return launch_coarse<K, false, false, false>(JSCC_COARSE_ARGS);
    case 1: return launch_coarse<K, false, false, true>(JSCC_COARSE_ARGS);
    case 2: return launch_coarse<K, false, true, false>(JSCC_COARSE_ARGS);
    case 3: return launch_coarse<K, false, true, true>(JSCC_COARSE_ARGS);
    case 4: return launch_coarse<K, true, false, false>(JSCC_COARSE_ARGS);
    case 5: return launch_coarse<K, true, false, true>(JSCC_COARSE_ARGS);
    case 6: return launch_coarse<K, true, true, false>(JSCC_COARSE_ARGS);
    default: return launch_coarse<K, true, true, true>(JSCC_COARSE_ARGS);
  }
#undef JSCC_COARSE_ARGS
}

}  // namespace

// workspace: partial rows | rotated image [P] | count of handed-over rows | their row ids [n_rows]
static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

extern "C" int64_t jscc_coarse_weight_workspace_bytes(int64_t n_rows, int64_t n_fine) {
  return align256(jscc_partials_bytes(coarse_grid(n_rows), n_fine)) + align256(n_fine * 4) + 256 +
         align256((n_rows > 0 ? n_rows : 1) * 4);
}

extern "C" int jscc_coarse_weight(const float* packed, int64_t ld_packed, const int32_t* rows, int64_t n_rows,
                                  const int32_t* z_idx, const float* z_w, int n_zf, int n_zc, const int32_t* za_idx,
                                  const float* za_w, int n_adj, const int32_t* l_idx, const float* l_w, int n_lf,
                                  int n_lc, const float* A, int64_t lda, int64_t n_arows, const float* x,
                                  const int64_t* x_perm, int64_t x_perm_stride, int64_t x_len, float eps, int use_eps,
                                  int normalize, const int64_t* inv_perm, int64_t inv_perm_stride, float out_scale,
                                  int accumulate, float* w, int64_t w_len, void* workspace, int64_t workspace_bytes,
                                  int32_t* err, void* stream) {
  JSCC_REQUIRE(n_rows >= 0 && n_zf > 0 && n_zc > 0 && n_lf > 0 && n_lc > 0, "jscc_coarse_weight: bad shape");
  const int64_t P = (int64_t)n_zf * n_lf, Pc = (int64_t)n_zc * n_lc;
  JSCC_REQUIRE(Pc <= kMaxCoarse, "jscc_coarse_weight: %lld coarse pixels exceed %d", (long long)Pc, kMaxCoarse);
  JSCC_REQUIRE(P <= 32 * kCoarseThreads, "jscc_coarse_weight: %lld fine pixels exceed %d", (long long)P,
               32 * kCoarseThreads);
  JSCC_REQUIRE(P % 4 == 0, "jscc_coarse_weight: the fine pixel count must be a multiple of 4");
  JSCC_REQUIRE(ld_packed >= Pc + 1 && lda >= P && n_arows > 0, "jscc_coarse_weight: bad strides");
  JSCC_REQUIRE(l_idx != nullptr || n_lf == n_lc, "jscc_coarse_weight: an identity layer map needs n_lf == n_lc");
  JSCC_REQUIRE(z_idx && z_w && (l_idx == nullptr || l_w != nullptr) && A && x && w && err,
               "jscc_coarse_weight: null pointer");
  JSCC_REQUIRE((za_idx == nullptr) == (za_w == nullptr) && (za_idx == nullptr || n_adj >= 1),
               "jscc_coarse_weight: incomplete z adjoint table");
  JSCC_REQUIRE(x_perm != nullptr || x_len == P, "jscc_coarse_weight: image length mismatch");
  JSCC_REQUIRE(inv_perm != nullptr || w_len == P, "jscc_coarse_weight: weight length mismatch");
  hipStream_t s = (hipStream_t)stream;
  if (n_rows == 0) {
    if (!accumulate) JSCC_CHECK_HIP(hipMemsetAsync(w, 0, w_len * sizeof(float), s));
    return 0;
  }
  JSCC_REQUIRE(packed != nullptr, "jscc_coarse_weight: null rows");
  const int G = coarse_grid(n_rows);
  JSCC_REQUIRE(workspace != nullptr && workspace_bytes >= jscc_coarse_weight_workspace_bytes(n_rows, P),
               "jscc_coarse_weight: workspace too small");
  float* partial = (float*)workspace;
  char* tail = (char*)workspace + align256(jscc_partials_bytes(G, P));
  float* x_rot = (float*)tail;
  int32_t* cx_count = (int32_t*)(tail + align256(P * 4));
  int32_t* cx_list = (int32_t*)(tail + align256(P * 4) + 256);
  const bool lid = l_idx == nullptr;
  const size_t lds = coarse_lds((int)Pc, n_zf, n_lf, lid);
  JSCC_REQUIRE(lds <= 160 * 1024, "jscc_coarse_weight: the coarse row buffers and tap tables need %zu bytes of LDS",
               lds);
  const int kk = (int)((P + kCoarseThreads - 1) / kCoarseThreads);
  const bool zu = n_lf % JSCC_WAVE == 0;
  // the grouped form: identity layer map, an adjoint table, at most half as many coarse as fine pixels
  // (the adjoint table has kGrpAdj entries per coarse slice, padded with zero weights; up to 25 fine slots)
  const size_t glds = za_idx ? grouped_lds((int)P, n_zf, n_zc, kGrpAdj) : 0;
  const bool grouped = lid && za_idx != nullptr && n_adj == kGrpAdj && 2 * Pc <= P && kk <= 25 &&
                       glds <= 160 * 1024;
  // the prefetching per-row form: identity layer map with 256 | Lf, up to 26624 fine pixels, both row
  // buffers and the crystal table of a work-group's rows within the LDS
  const int64_t group_rows = (n_rows + G - 1) / G;
  const int k4 = (int)((P + 4 * kPfThreads - 1) / (4 * kPfThreads));
  const int kv = k4 <= 2 ? 2 : (k4 <= 4 ? 4 : (k4 <= 8 ? 8 : kPfMaxK4));  // the instantiated quad count
  const size_t plds = pf_lds((int)Pc, group_rows);
  const bool pf = lid && n_lf % 256 == 0 && k4 <= kPfMaxK4 && Pc <= pf_copies(kv) * kPfWaves * kPfChunk &&
                  plds <= 160 * 1024 && n_arows <= INT32_MAX;
  int rc;
  jscc_prof_begin(JSCC_TAG_COARSE, s);
#define JSCC_COARSE(KV, XP, LIST, CNT, ADD)                                                                         \
  rc = dispatch_coarse<KV>(lid, zu, normalize != 0, lds, G, s, packed, ld_packed, n_rows, (int)Pc, z_idx, z_w, n_zf, \
                           n_zc, l_idx, l_w, n_lf, n_lc, A, lda, n_arows, XP ? x : x_rot, XP ? x_perm : nullptr,     \
                           x_perm_stride, XP ? x_len : P, eps, use_eps, partial, err, LIST, CNT, ADD)
#define JSCC_COARSE_K(XP, LIST, CNT, ADD)        \
  if (kk <= 4) JSCC_COARSE(4, XP, LIST, CNT, ADD);       \
  else if (kk <= 8) JSCC_COARSE(8, XP, LIST, CNT, ADD);  \
  else if (kk <= 16) JSCC_COARSE(16, XP, LIST, CNT, ADD); \
  else if (kk <= 25) JSCC_COARSE(25, XP, LIST, CNT, ADD); \
  else JSCC_COARSE(32, XP, LIST, CNT, ADD);
  if (!grouped && pf) {
#define JSCC_PF(KV, NV)                                                                                           \
  rc = launch_pf<KV, NV>(plds, G, s, packed, ld_packed, n_rows, (int)Pc, z_idx, z_w, n_zf, n_zc, n_lf, A, lda,      \
                         n_arows, x, x_perm, x_perm_stride, x_len, eps, use_eps, partial, err, rows)
#define JSCC_PF_N(KV) \
  if (normalize) JSCC_PF(KV, true); else JSCC_PF(KV, false);
    if (kv == 2) { JSCC_PF_N(2) }
    else if (kv == 4) { JSCC_PF_N(4) }
    else if (kv == 8) { JSCC_PF_N(8) }
    else { JSCC_PF_N(kPfMaxK4) }
#undef JSCC_PF_N
#undef JSCC_PF
  } else if (!grouped) {
    JSCC_COARSE_K(true, rows, nullptr, 0)
  } else {
    // rotated image (when x_perm is given) and a cleared count of handed-over rows
    hipLaunchKernelGGL(k_grp_prepare, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, x, x_perm, x_perm_stride,
                       x_len, (int)P, x_rot, cx_count, err);
    JSCC_CHECK_LAUNCH();
    const float* xg = x_perm ? x_rot : x;
#define JSCC_GRP(KV, NV)                                                                                           \
  rc = launch_grouped<KV, NV>(glds, G, s, packed, ld_packed, rows, n_rows, n_lf, n_zf, n_zc, z_idx, z_w,             \
                              za_idx, za_w, A, lda, n_arows, xg, eps, use_eps, partial, cx_list, cx_count, err)
#define JSCC_GRP_N(KV) \
  if (normalize) JSCC_GRP(KV, true); else JSCC_GRP(KV, false);
    if (kk <= 4) { JSCC_GRP_N(4) }
    else if (kk <= 8) { JSCC_GRP_N(8) }
    else if (kk <= 16) { JSCC_GRP_N(16) }
    else { JSCC_GRP_N(25) }
#undef JSCC_GRP_N
#undef JSCC_GRP
    if (rc == 0 && normalize) {
      JSCC_CHECK_LAUNCH();
      // the handed-over rows, one product at a time, added into the same partials
      if (x_perm) {
        JSCC_COARSE_K(false, cx_list, cx_count, 1)
      } else {
        JSCC_COARSE_K(true, cx_list, cx_count, 1)
      }
    }
  }
#undef JSCC_COARSE_K
#undef JSCC_COARSE
  jscc_prof_end(JSCC_TAG_COARSE, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  const JsccEm none = {nullptr, nullptr, 0, 0.0f, 0, 1.0f, nullptr, 0.0f};
  return jscc_reduce_partials(partial, G, (int)P, P, inv_perm, inv_perm_stride, out_scale, accumulate, 1, w, w_len,
                              err, s, none);
}
