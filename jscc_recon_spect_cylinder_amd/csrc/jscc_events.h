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
