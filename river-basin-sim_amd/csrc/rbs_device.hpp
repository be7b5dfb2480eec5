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

// Pointers in the global address space.  A kernel argument's pointers are known to be
// global; pointers read from memory (a batch's per-member argument blocks) are
// generic, and their accesses compile to flat instructions on 64-bit per-lane
// addresses.  Cast once, every access is a global load or store with a scalar base.
template <typename T>
using gptr = T __attribute__((address_space(1))) *;
template <typename T>
__device__ __forceinline__ gptr<T> glob(T *p) {
  return (gptr<T>)p;
}

// Powers of a positive float32 (wp.pow in friction.py:31, :57, sediment.py:168, :192).
// The reference passes the exponents as Python floats into a float32 pow, so they are
// float32(7/3) = 7/3 - 7.9e-8, float32(4/3) = 4/3 + 4.0e-8 and float32(1/3) = 1/3 +
// 9.9e-9, and 1.5 exactly.  A libm powf is good to about an ulp and its last bit
// differs between math libraries; at wet/dry switches that bit grows.  So the solver
// evaluates x^e, e = float32(p/3) = p/3 - d, as the correctly rounded float32 power,
// by one fixed sequence of IEEE basic operations (the build sets -ffp-contract=off, so
// the only fused multiply-adds are the fmaf/fma written here, correctly rounded
// everywhere), bit-identical on every IEEE machine:
//  1. y ~ x^(-1/3) from a bit-pattern seed and one float32 fourth-order step (|1 - x
//     y^3| below 8.6e-5 for every normal x); then r = 1 - x y^3 exactly in float64 (x
//     y and y^2 are exact);
//  2. x^(p/3 - d) = x^q y^2 (1 - r)^(-2/3) x^(-d), since x^(-2/3) = y^2 (1 - r)^(-2/3),
//     with q = 3, 2, 1 for p = 7, 4, 1: the binomial series of (1 - r)^(-2/3) to r^3 (the next term
//     is below 2.5e-17) and x^(-d) = exp(-u) = 1 - u + u^2/2, u = d ln x (|u| < 1e-5 over
//     the float32 range), merged into one factor 1 + g, with ln x to 3e-8 absolute
//     (below), so u is good to 2.4e-15 relative;
//  3. x^q y^2 (1 + g) in float64, rounded once to float32.
// Float64 here buys the last float32 bit of a power; every field stays float32.
// The seed y0 (|1 - x y0^3| < 0.13) times the series of (1 - r0)^(-1/3) = 1 + r0/3 +
// 2 r0^2/9 + 14 r0^3/81, r0 = 1 - x y0^3: seven float32 operations (two Newton steps
// take ten and leave |r| below 5.6e-5, which moves no power in [1e-3, 1e4)).
__device__ __forceinline__ float cbrt_seed(float x) {  // x^(-1/3) to 2.9e-5
  const uint32_t j = 0x54a2fa8cu - __float_as_uint(x) / 3u;
  const float y = __uint_as_float(j);
  const float r = fmaf(-(x * y), y * y, 1.0f);
  const float t = fmaf(fmaf(r, (float)(14.0 / 81.0), (float)(2.0 / 9.0)), r, (float)(1.0 / 3.0));
  return fmaf(y * r, t, y);
}
// ln x for a positive normal float32 x = 2^k m, m in [sqrt(1/2), sqrt(2)) (k shifted by
// kadj), to 3e-8 absolute: the bit pattern offset by that of sqrt(1/2) gives k and m;
// ln(1 + t) = t - t^2/2 + t^3 R(t), t = m - 1 exact, with R a degree-6 polynomial fitted
// to the remainder on that range (error 7e-9), in float32; k ln 2 added in float64.
__device__ __forceinline__ double ln_coarse(float x, int kadj) {
  const int32_t ix = (int32_t)(__float_as_uint(x) - 0x3f3504f3u);
  const int k = (ix >> 23) + kadj;
  const float t = __uint_as_float(((uint32_t)ix & 0x007fffffu) + 0x3f3504f3u) - 1.0f;
  float r = 0.08672088310927706f;
  r = fmaf(r, t, -0.14379001534630725f);
  r = fmaf(r, t, 0.1497802208269855f);
  r = fmaf(r, t, -0.16564426496452503f);
  r = fmaf(r, t, 0.19954873155337902f);
  r = fmaf(r, t, -0.2500169873952298f);
  r = fmaf(r, t, 0.3333416597960291f);
  const float s = fmaf(r, t, -0.5f);
  const float l = fmaf(t * t, s, t);
  return fma((double)k, 0.6931471805599453, (double)l);
}

// x^e for e = float32(P/3), P in {1, 4, 7}, x >= 0.  SUB admits subnormal x (scaled
// by 2^48, a power of 2 divisible by 3, first); without it x must be 0 or normal.
template <int P, bool SUB>
__device__ __forceinline__ float pow_third(float x) {
  // d = P/3 - float32(P/3); P - 3 float32(P/3) is exact in float64
  constexpr double d = ((double)P - 3.0 * (double)(float)((double)P / 3.0)) / 3.0;
  float xs = x;
  int kadj = 0;
  bool tiny = false;
  if (SUB) {
    tiny = x < 0x1p-126f;
    if (tiny) {
      xs = x * 0x1p48f;
      kadj = -48;
    }
  }
  const double a = (double)xs, b = (double)cbrt_seed(xs);
  const double b2 = b * b;
  const double r = fma(-(a * b), b2, 1.0);
  // a product and a sum (not an fma) for the r^2 coefficient: one constant per
  // instruction, none kept in vector registers
  const double s2 = r * fma(r * (40.0 / 81.0) + 5.0 / 9.0, r, 2.0 / 3.0);
  const double u = d * ln_coarse(xs, kadj);
  const double g = fma(-u, fma(-0.5, u, s2) + 1.0, s2);
  double A;
  if (P == 7) A = ((a * a) * a) * b2;
  else if (P == 4) A = (a * a) * b2;
  else A = a * b2;
  double p = fma(A, g, A);
  if (SUB && tiny) p *= (P == 7 ? 0x1p-112 : P == 4 ? 0x1p-64 : 0x1p-16);
  return (float)p;
}
// wp.pow(h, 7.0/3.0) for a depth h >= H_DRY (friction.py:31)
__device__ __forceinline__ float pow_e73(float x) { return pow_third<7, false>(x); }
// wp.pow(R, 4.0/3.0) and wp.pow(R, 1.0/3.0) of a hydraulic radius (friction.py:57,
// sediment.py:168), which a narrow channel can make arbitrarily small
__device__ __forceinline__ float pow_e43(float x) { return pow_third<4, true>(x); }
__device__ __forceinline__ float pow_e13(float x) { return pow_third<1, true>(x); }

// A wave-uniform vote: one ballot of the lanes for which `p` fails.
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }

__device__ __forceinline__ float pow15(float x) {  // x^(3/2) = x * sqrt(x)
  const double a = (double)x;
  return (float)(a * sqrt(a));
}

// friction.py:18-31 manning_denominator, 1 + g dt n^2 |q| / h^e, e = float32(7/3), in
// the reference's float32 order: the product left to right, one division, one sum.
__device__ __forceinline__ float manning_den(float q_prev, float h_flow, float n, float g, float dt) {
  return 1.0f + g * dt * n * n * fabsf(q_prev) / pow_e73(h_flow);
}

// sources.py:66-82 kahan_add
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

// a / b for a wave-uniform divisor b, given rb = RN(1/b) (an IEEE float division on
// the host or device).  Markstein's correction q1 = q0 + (a - b q0) rb, with the
// remainder exact in an FMA, is the correctly rounded quotient whenever a/b stays
// in the normal range (checked against IEEE division on 8.7e8 operands, every
// mantissa at 80 exponents included); a wave with a zero, tiny or huge dividend in
// any lane takes the plain division, which also keeps the sign of zero.
__device__ __forceinline__ float div_uniform(float a, float b, float rb) {
  const float q0 = a * rb;
  const float r = __builtin_fmaf(-b, q0, a);
  const float q1 = __builtin_fmaf(r, rb, q0);
  const float m = fabsf(a);
  const bool normal = m >= 0x1p-100f && m <= 0x1p+100f;
  if (wave_all(normal)) return q1;
  return normal ? q1 : a / b;
}

// c / p of an LI face's friction term (friction.py:31: g dt n^2 |q| / h^e; the plain face
// and a channel face's floodplain component), for c >= 0 and p = h^e with h >= kHDry.  The compiler's IEEE float division is this sequence
// with v_div_scale around it, and v_div_scale changes nothing unless c < 2^-103, c/p
// is subnormal or above 2^96, or p > 2^125.  The first two give c/p below 2^-80
// (p >= 1e-7), where the division's last bits vanish in 1 + c/p = 1; the others need
// |q| beyond 1e19 m^2/s or a flow depth beyond 1e16 m.  v_div_fixup keeps the special
// cases (0, inf, NaN) of the IEEE quotient.  Two instructions fewer than the division.
__device__ __forceinline__ float friction_ratio(float c, float p) {
  float y = __builtin_amdgcn_rcpf(p);
  const float e = __builtin_fmaf(-p, y, 1.0f);
  y = __builtin_fmaf(e, y, y);
  float q = c * y;
  float r = __builtin_fmaf(-p, q, c);
  q = __builtin_fmaf(r, y, q);
  r = __builtin_fmaf(-p, q, c);
  q = __builtin_fmaf(r, y, q);
  return __builtin_amdgcn_div_fixupf(q, p, c);
}

// local_inertial.py:72-99 (update_qx) / :102-129 (update_qy): one interior face.
// a = the left/top cell, b = the right/bottom cell; rdx = RN(1/dx).
// The friction term is manning_den's, left to right, with friction_ratio's quotient.
__device__ __forceinline__ float li_face(float q, float eta_a, float eta_b, float z_a, float z_b,
                                         float n_a, float n_b, float dx, float rdx, float dt) {
  const float g = kG;
  float h_flow = fmx(eta_a, eta_b) - fmx(z_a, z_b);  // face_h_flow :50-58
  if (h_flow < kHDry) return 0.0f;
  float n_face = 0.5f * (n_a + n_b);
  float slope = div_uniform(eta_b - eta_a, dx, rdx);
  float num = q - g * h_flow * dt * slope;
  float den = 1.0f + friction_ratio(g * dt * n_face * n_face * fabsf(q), pow_e73(h_flow));
  return num / den;
}

// ---- sub-grid channels (channels.py:59-138, local_inertial.py:266-500) -------------
// The storage curve h -> eta (channels.py:59-72), in float32 IEEE operations; the
// device forms divide by the uniform dx with div_uniform (the same correctly rounded
// quotient), so host and device agree bit for bit.
__host__ __device__ inline float eta_subgrid(float h, float z, float w, float d, float dx) {
  float eta = h + z;
  if (w > 0.0f && d > 0.0f) {
    const float h_bf = w * d / dx;
    if (h <= h_bf) eta = z - d + h * dx / w;
    else eta = z + (h - h_bf);
  }
  return eta;
}
__device__ __forceinline__ float eta_subgrid_u(float h, float z, float w, float d, float dx,
                                               float rdx) {
  float eta = h + z;
  if (w > 0.0f && d > 0.0f) {
    const float h_bf = div_uniform(w * d, dx, rdx);
    if (h <= h_bf) eta = z - d + h * dx / w;
    else eta = z + (h - h_bf);
  }
  return eta;
}
// channels.py:75-85 column_depth: eta - z_ch, the depth the CFL bound sees.
__device__ __forceinline__ float column_depth_u(float h, float w, float d, float dx, float rdx) {
  float col = h;
  if (w > 0.0f && d > 0.0f) {
    const float h_bf = div_uniform(w * d, dx, rdx);
    if (h <= h_bf) col = h * dx / w;
    else col = d + (h - h_bf);
  }
  return col;
}

// One interior face of a channelled state (update_qx_channels :334-381 /
// update_qy_channels :384-429): a = left/top cell, b = right/bottom cell; the
// floodplain component (_floodplain_flux :315-331) and the channel component
// (_channel_flux :286-312, friction.py:35-57) from their previous values, both
// written back; returns the total per unit cell width.
struct ChanCell {
  float eta, z, n, w, d, nc;
};
__device__ __forceinline__ float chan_face(const ChanCell &a, const ChanCell &b, float &q_ch,
                                           float &q_fp, float dx, float rdx, float dt) {
  const float g = kG;
  const float eta_max = fmx(a.eta, b.eta);
  const float slope = div_uniform(b.eta - a.eta, dx, rdx);
  const float z_bank = fmx(a.z, b.z);
  float fp = 0.0f;
  const float h_fp = eta_max - z_bank;
  if (!(h_fp < kHDry)) {
    const float num = q_fp - g * h_fp * dt * slope;
    const float n_fp = 0.5f * (a.n + b.n);  // manning_den's expression, friction_ratio's quotient
    fp = num / (1.0f + friction_ratio(g * dt * n_fp * n_fp * fabsf(q_fp), pow_e73(h_fp)));
  }
  const float w_face = fmn(a.w, b.w);
  float ch = 0.0f;
  if (!(w_face <= 0.0f)) {
    const float z_ch = fmx(a.z - a.d, b.z - b.d);
    const float h_ch = fmn(eta_max, z_bank) - z_ch;
    if (!(h_ch < kHDry)) {
      const float radius = w_face * h_ch / (w_face + 2.0f * h_ch);
      const float nc = 0.5f * (a.nc + b.nc);
      const float num = q_ch - g * h_ch * dt * slope;
      const float den = 1.0f + g * dt * nc * nc * fabsf(q_ch) / (h_ch * pow_e43(radius));
      ch = num / den;
    }
  }
  q_fp = fp;
  q_ch = ch;
  const float frac = div_uniform(w_face, dx, rdx);
  return fp * (1.0f - frac) + ch * frac;
}

// limit_qx_channels / limit_qy_channels (:432-484) for one face: a non-zero total
// takes its donor's factor (lo: the left/top cell's, for positive flow; hi: the
// right/bottom cell's, for negative), and so do both components.  `q` is the
// unlimited total.  A zero total's components are multiplied by 1 (the same bits),
// so the face needs selects and no branch.
__device__ __forceinline__ void limit_components(float q, float &q_ch, float &q_fp, float b_lo,
                                                 float b_hi) {
  const float b = q == 0.0f ? 1.0f : (q < 0.0f ? b_hi : b_lo);
  q_ch = q_ch * b;
  q_fp = q_fp * b;
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
  float shear = kG * n * n * q * q / (h * h * pow_e13(radius));
  float theta = shear / (sg * kG * d50);
  if (theta <= th_c) return 0.0f;
  float excess = theta - th_c;
  float qs = 8.0f * pow15(excess) * sqrtf(sg * kG * d50 * d50 * d50);
  return q < 0.0f ? -qs : qs;
}

// sediment.py:276-302 _combined_capacity: the floodplain and channel components of a
// face's bedload flux, each from its own depth and hydraulic radius, mixed by the
// face's width fraction (per unit cell width).
__device__ __forceinline__ float combined_capacity(float q_ch, float q_fp, const ChanCell &a,
                                                   const ChanCell &b, float d50, float dx,
                                                   float rdx) {
  const float eta_max = fmx(a.eta, b.eta), z_bank = fmx(a.z, b.z);
  const float h_fp = eta_max - z_bank;
  const float qs_fp = face_capacity(q_fp, h_fp, h_fp, 0.5f * (a.n + b.n), d50);
  const float w_face = fmn(a.w, b.w);
  float qs_ch = 0.0f;
  if (w_face > 0.0f) {
    const float h_ch = fmn(eta_max, z_bank) - fmx(a.z - a.d, b.z - b.d);
    if (h_ch >= kHDry)
      qs_ch = face_capacity(q_ch, h_ch, w_face * h_ch / (w_face + 2.0f * h_ch),
                            0.5f * (a.nc + b.nc), d50);
  }
  const float frac = div_uniform(w_face, dx, rdx);
  return qs_fp * (1.0f - frac) + qs_ch * frac;
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

// Block reduction of (max f32, 3 x sum f64) over NT threads; result valid on
// thread 0.  `scratch` must hold NT/64 * 4 doubles.
template <int NT = kThreads>
__device__ __forceinline__ void block_reduce4(float &m, double &a, double &b, double &c,
                                              double *scratch) {
  m = wave_max(m);
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int W = NT / 64;
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
