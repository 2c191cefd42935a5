// Per-event record shared by the event-preparation and cone kernels.
#pragma once
#include <stdint.h>

// 64 bytes, one per list event (indexed by the event's input position).
// Geometry in mm, angles in rad.  Written by jscc_events_prepare, read with
// wave-uniform (scalar) loads by the cone kernels.
struct __attribute__((aligned(16))) JsccEvent {
  float p1x, p1y, p1z;   // first-hit crystal centre
  float ax, ay, az;      // v12 = second hit - first hit
  float d12sq;           // |v12|^2
  float theta;           // Compton angle from the smeared E1 (computed in FP64)
  float kn;              // E0/(E0-E1) + (E0-E1)/E0
  float sig_plus;        // angular sigma used where beta >= theta
  float sig_minus;       // angular sigma used where beta <  theta
  float wx, wy, wz;      // sigma_pos1^2 + sigma_pos2^2 per axis (mm^2)
  int32_t arow;          // 0-based system-matrix row of the first hit (cpnum1-1, wrapped)
  int32_t src;           // input position of the event
};
static_assert(sizeof(JsccEvent) == 64, "JsccEvent must be 64 bytes");

// The NaN / zero-sum / effective-support decision of process_list_plane_strict.py:241-246, 145-146 on the
// FP64 row statistics (S = sum t, Q = sum t^2, M = max t).  A row lying entirely below the smallest normal
// FP32 value is a zero row: its subnormal tail depends on the exponential's subnormal handling.  Shared by
// every kernel that takes the decision, so they all keep the same rows.
__device__ __forceinline__ bool jscc_row_keep(double S, double Q, double M, double min_support) {
#pragma clang fp contract(off)
  bool keep = (S == S) && (S != 0.0) && (M >= 1.17549435082228750797e-38);
  if (keep) {
    const double support = (S * S) / Q;  // 1 / sum((t/S)^2)
    keep = support >= min_support;       // NaN compares false
  }
  return keep;
}
