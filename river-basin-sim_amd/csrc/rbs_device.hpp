// rbs_device.hpp -- device helpers shared by the librbs kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include "rbs_internal.hpp"

namespace rbs {

// Warp's scalar max/min (a > b ? a : b); kept explicit so signed zeros and
// comparisons follow the reference's semantics, not fmaxf's.
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fclamp(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// friction.py:104-117 manning_denominator
__device__ __forceinline__ float manning_den(float q_prev, float h_flow, float n, float g, float dt) {
  return 1.0f + g * dt * n * n * fabsf(q_prev) / powf(h_flow, kSevenThirds);
}

// sources.py:274-290 kahan_add
__device__ __forceinline__ void kahan_add(float &total, float &comp, float inc) {
  float y = inc - comp;
  float t = total + y;
  comp = (t - total) - y;
  total = t;
}

// massbalance.py:39-50 _Kahan.add in float64
__device__ __forceinline__ void kahan_add64(double &total, double &comp, double x) {
  double y = x - comp;
  double t = total + y;
  comp = (t - total) - y;
  total = t;
}

// local_inertial.py:72-99 (update_qx) / :102-129 (update_qy): one interior face.
// a = the left/top cell, b = the right/bottom cell.
__device__ __forceinline__ float li_face(float q, float eta_a, float eta_b, float z_a, float z_b,
                                         float n_a, float n_b, float dx, float dt) {
  const float g = kG;
  float h_flow = fmx(eta_a, eta_b) - fmx(z_a, z_b);  // face_h_flow :50-58
  if (h_flow < kHDry) return 0.0f;
  float n_face = 0.5f * (n_a + n_b);
  float slope = (eta_b - eta_a) / dx;
  float num = q - g * h_flow * dt * slope;
  float den = manning_den(q, h_flow, n_face, g, dt);
  return num / den;
}

// boundaries.py:94-106 _drain
__device__ __forceinline__ float drain(float h_edge, float q_out, float dxdt) {
  float removed = dxdt * q_out;
  if (removed > h_edge) removed = h_edge;
  if (removed < 0.0f) removed = 0.0f;
  return h_edge - removed;
}

// sediment.py:161-207: kinematic shear, Shields number, MPM capacity, signed face flux.
__device__ __forceinline__ float face_capacity(float q, float h, float radius, float n, float d50) {
  const float sg = (float)1.65, th_c = (float)0.047;
  if (h < kHDry || d50 <= 0.0f || q == 0.0f) return 0.0f;
  float shear = kG * n * n * q * q / (h * h * powf(radius, (float)(1.0 / 3.0)));
  float theta = shear / (sg * kG * d50);
  if (theta <= th_c) return 0.0f;
  float excess = theta - th_c;
  float qs = 8.0f * powf(excess, 1.5f) * sqrtf(sg * kG * d50 * d50 * d50);
  return q < 0.0f ? -qs : qs;
}

// Map a launch-order block id to a tile id so that the tiles one XCD executes are
// contiguous: blocks go round-robin over the 8 XCDs, and neighbouring tiles share
// halo rows that are then served from the same XCD's L2.  Bijective for any n.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  constexpr int X = 8;
  const int q = n / X, r = n % X;
  const int x = b % X, k = b / X;
  return x * q + (x < r ? x : r) + k;
}

// ---- wave64 / block reductions (fixed order => deterministic) ---------------------
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmx(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reduction of (max f32, 3 x sum f64) over kThreads threads; result valid on
// thread 0.  `scratch` must hold kThreads/64 * 4 doubles.
__device__ __forceinline__ void block_reduce4(float &m, double &a, double &b, double &c,
                                              double *scratch) {
  m = wave_max(m);
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int W = kThreads / 64;
  __syncthreads();
  if (lane == 0) {
    scratch[wid] = (double)m;
    scratch[W + wid] = a;
    scratch[2 * W + wid] = b;
    scratch[3 * W + wid] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float mm = (float)scratch[0];
    double aa = scratch[W], bb = scratch[2 * W], cc = scratch[3 * W];
    for (int w = 1; w < W; ++w) {
      mm = fmx(mm, (float)scratch[w]);
      aa += scratch[W + w];
      bb += scratch[2 * W + w];
      cc += scratch[3 * W + w];
    }
    m = mm;
    a = aa;
    b = bb;
    c = cc;
  }
}

// scheduler.py:48-72 fill_span, float64
__device__ __forceinline__ double fill_span_dev(double span, double dt_raw) {
  if (!isfinite(span)) return dt_raw;
  if (span <= dt_raw) return span;
  double n = ceil(span / dt_raw);
  double dt = span / n;
  if (dt > dt_raw) dt = span / (n + 1.0);
  return dt;
}

}  // namespace rbs
