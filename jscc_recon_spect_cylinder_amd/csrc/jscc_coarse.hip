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
                    float* __restrict__ partial, int32_t* __restrict__ err) {
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

  // ---- this work-group's rows
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
  auto fetch = [&](int64_t r, int b) {
    const float* src = packed + r * ldp + 1;
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
    const float cpf = rintf(packed[r * ldp]) - 1.0f;
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
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = jb + k * kCoarseThreads;
    if (j < P) out[j] = wr[k];
  }
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
                  int64_t xps, int64_t x_len, float eps, int use_eps, float* partial, int32_t* err) {
  static bool attr_set = false;
  if (!attr_set) {
    JSCC_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_weight<K, LID, ZU, NORM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_coarse_weight<K, LID, ZU, NORM>), dim3(G), dim3(kCoarseThreads), lds, s, packed, ldp, n_rows,
                     Pc, z_idx, z_w, Zf, Zc, l_idx, l_w, Lf, Lc, A, lda, M, x, x_perm, xps, x_len, eps, use_eps,
                     partial, err);
  return 0;
}

// the template flags of one call: identity layer map, wave-uniform z slices (64 | Lf), normalised rows
template <int K>
int dispatch_coarse(bool lid, bool zu, bool norm, size_t lds, int G, hipStream_t s, const float* packed, int64_t ldp,
                    int64_t n_rows, int Pc, const int32_t* z_idx, const float* z_w, int Zf, int Zc,
                    const int32_t* l_idx, const float* l_w, int Lf, int Lc, const float* A, int64_t lda, int64_t M,
                    const float* x, const int64_t* x_perm, int64_t xps, int64_t x_len, float eps, int use_eps,
                    float* partial, int32_t* err) {
#define JSCC_COARSE_ARGS                                                                                          \
  lds, G, s, packed, ldp, n_rows, Pc, z_idx, z_w, Zf, Zc, l_idx, l_w, Lf, Lc, A, lda, M, x, x_perm, xps, x_len, eps, \
      use_eps, partial, err
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

extern "C" int64_t jscc_coarse_weight_workspace_bytes(int64_t n_rows, int64_t n_fine) {
  return jscc_partials_bytes(coarse_grid(n_rows), n_fine);
}

extern "C" int jscc_coarse_weight(const float* packed, int64_t ld_packed, int64_t n_rows, const int32_t* z_idx,
                                  const float* z_w, int n_zf, int n_zc, const int32_t* l_idx, const float* l_w,
                                  int n_lf, int n_lc, const float* A, int64_t lda, int64_t n_arows, const float* x,
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
  JSCC_REQUIRE(x_perm != nullptr || x_len == P, "jscc_coarse_weight: image length mismatch");
  JSCC_REQUIRE(inv_perm != nullptr || w_len == P, "jscc_coarse_weight: weight length mismatch");
  hipStream_t s = (hipStream_t)stream;
  if (n_rows == 0) {
    if (!accumulate) JSCC_CHECK_HIP(hipMemsetAsync(w, 0, w_len * sizeof(float), s));
    return 0;
  }
  JSCC_REQUIRE(packed != nullptr, "jscc_coarse_weight: null rows");
  const int G = coarse_grid(n_rows);
  JSCC_REQUIRE(workspace != nullptr && workspace_bytes >= jscc_partials_bytes(G, P),
               "jscc_coarse_weight: workspace too small");
  const bool lid = l_idx == nullptr;
  const size_t lds = coarse_lds((int)Pc, n_zf, n_lf, lid);
  JSCC_REQUIRE(lds <= 160 * 1024, "jscc_coarse_weight: the coarse row buffers and tap tables need %zu bytes of LDS",
               lds);
  const int kk = (int)((P + kCoarseThreads - 1) / kCoarseThreads);
  int rc;
  jscc_prof_begin(JSCC_TAG_COARSE, s);
  const bool zu = n_lf % JSCC_WAVE == 0;
#define JSCC_COARSE(KV)                                                                                          \
  rc = dispatch_coarse<KV>(lid, zu, normalize != 0, lds, G, s, packed, ld_packed, n_rows, (int)Pc, z_idx, z_w, n_zf, \
                           n_zc, l_idx, l_w, n_lf, n_lc, A, lda, n_arows, x, x_perm, x_perm_stride, x_len, eps,      \
                           use_eps, (float*)workspace, err)
  if (kk <= 4) JSCC_COARSE(4);
  else if (kk <= 8) JSCC_COARSE(8);
  else if (kk <= 16) JSCC_COARSE(16);
  else if (kk <= 25) JSCC_COARSE(25);
  else JSCC_COARSE(32);
#undef JSCC_COARSE
  jscc_prof_end(JSCC_TAG_COARSE, s);
  if (rc != 0) return rc;
  JSCC_CHECK_LAUNCH();
  const JsccEm none = {nullptr, nullptr, 0, 0.0f, 0, 1.0f, nullptr, 0.0f};
  return jscc_reduce_partials((float*)workspace, G, (int)P, inv_perm, inv_perm_stride, out_scale, accumulate, 1, w,
                              w_len, err, s, none);
}
