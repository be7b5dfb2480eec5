// rbs_check.hip -- the device's own proof that the short arithmetic routes of the
// kernels give the bits of the exact ones (rbs_check_arithmetic).
//
// The kernels evaluate the reference's float32 powers and divisions correctly rounded
// (rbs_device.hpp).  Two shorter routes stand in for the long sequences where they are
// exact: pow_third_hw (hardware log2/exp2 seeds and a float64 correction, with a
// midpoint guard that hands doubtful inputs to pow_third) and div_rn (reciprocal, a
// Newton step and two Markstein corrections, where the operands keep every step
// normal).  Their exactness rests on the hardware's v_log_f32, v_exp_f32 and v_rcp_f32,
// so this check runs on the GPU that will run the solver:
//   - every float32 x in [2^-11, 2^21], for each exponent float32(7/3), (4/3), (1/3):
//     where the short route claims exactness it must equal pow_third;
//   - v_rcp_f32(m 2^k) = v_rcp_f32(m) 2^-k for every significand m, k in [-100, 100]
//     (a quotient's significand pair then fixes its bits at any exponents that keep the
//     steps normal);
//   - div_rn against the IEEE division on every significand of b with `a_count`
//     significands of a (all 2^23 of them prove the identity for the whole range);
//   - div_rn on 2^31 random pairs of the ranges its callers admit, and the callers'
//     helpers (one_plus_ratio, div_by_den) against the plain IEEE expressions on random
//     operands over the whole float32 range (zeros, subnormals, infinities, NaN).
#include "rbs_device.hpp"

namespace rbs {
namespace {

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ float ieee_div(float a, float b) { return a / b; }

// counts[0] mismatches, counts[1] inputs left to pow_third
template <int P>
__global__ void check_pow(uint32_t lo_bits, uint32_t n, unsigned long long *counts) {
  unsigned long long bad = 0, slow = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = __uint_as_float(lo_bits + i);
    bool exact;
    const float f = pow_third_hw<P>(x, exact);
    const float e = (P == 7) ? pow_third<7, false>(x) : pow_third<P, true>(x);
    if (!exact) ++slow;
    else if (__float_as_uint(f) != __float_as_uint(e)) ++bad;
  }
  if (bad) atomicAdd(&counts[0], bad);
  if (slow) atomicAdd(&counts[1], slow);
}

__global__ void check_rcp_scale(unsigned long long *count) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (1u << 23)) return;
  const float B = __uint_as_float(0x3f800000u | m);
  const float rB = __builtin_amdgcn_rcpf(B);
  unsigned long long bad = 0;
  for (int k = -100; k <= 100; ++k)
    if (__float_as_uint(__builtin_amdgcn_rcpf(ldexpf(B, k))) != __float_as_uint(ldexpf(rB, -k)))
      ++bad;
  if (bad) atomicAdd(count, bad);
}

// b = 1.m (one per thread, all 2^23), a = 1.(j * stride) for j in [j0, j0 + nj)
__global__ void check_div_pairs(uint32_t j0, uint32_t nj, uint32_t stride,
                                unsigned long long *count) {
  const uint32_t mb = blockIdx.x * blockDim.x + threadIdx.x;
  if (mb >= (1u << 23)) return;
  const float b = __uint_as_float(0x3f800000u | mb);
  unsigned long long bad = 0;
  for (uint32_t j = j0; j < j0 + nj; ++j) {
    const float a = __uint_as_float(0x3f800000u | (j * stride));
    if (__float_as_uint(div_rn(a, b)) != __float_as_uint(ieee_div(a, b))) ++bad;
  }
  if (bad) atomicAdd(count, bad);
}

__device__ __forceinline__ float rand_float(uint32_t h, int e0, int e1, bool sign) {
  const int e = e0 + (int)((h >> 8) % (uint32_t)(e1 - e0 + 1));
  const uint32_t m = mix(h) & 0x7fffffu;
  return __uint_as_float(((uint32_t)(e + 127) << 23) | m | (sign ? (h & 0x80000000u) : 0u));
}

// counts[0]: div_rn on the callers' ranges; counts[1]: one_plus_ratio and div_by_den
// against the IEEE expressions on any operands
__global__ void check_div_random(uint32_t seed, int iters, unsigned long long *counts) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long bad = 0, bad_helpers = 0;
  for (int it = 0; it < iters; ++it) {
    const uint32_t h1 = mix(seed ^ mix(t * 2654435761u + (uint32_t)it));
    const uint32_t h2 = mix(h1 + 0x9e3779b9u), h3 = mix(h2 + 0x85ebca6bu);
    // div_by_den's range, then one_plus_ratio's
    float a = rand_float(h1, -96, 96, true), b = rand_float(h2, 0, 24, false);
    if (__float_as_uint(div_rn(a, b)) != __float_as_uint(ieee_div(a, b))) ++bad;
    a = rand_float(h2, -50, 45, false);
    b = rand_float(h3, -24, 47, false);
    if (__float_as_uint(div_rn(a, b)) != __float_as_uint(ieee_div(a, b))) ++bad;
    // any bit patterns (every exponent, zeros, subnormals, infinities, NaN): the helpers
    // must give the IEEE expressions' bits (NaN results compare as NaN)
    const uint32_t k = h3 & 7u;
    float c = __uint_as_float(mix(h3) & 0x7fffffffu), p = __uint_as_float(mix(h1 ^ h3) & 0x7fffffffu);
    if (k == 0) c = 0.0f;
    if (k == 1) c = rand_float(h2, -60, -40, false);
    if (k == 2) p = rand_float(h1, -30, 50, false);
    if (k == 3) c = rand_float(h1, 40, 50, false);
    const float r1 = one_plus_ratio(c, p), e1 = 1.0f + ieee_div(c, p);
    if (__float_as_uint(r1) != __float_as_uint(e1) && !(r1 != r1 && e1 != e1)) ++bad_helpers;
    // the Manning form: a depth in [H_DRY, 2^21] and its friction power
    const float hd = k < 4 ? rand_float(h2, -10, 20, false) : fmx(kHDry, rand_float(h1, -10, -9, false));
    const float ph = pow_e73(hd);
    const float r3 = one_plus_ratio_h(c, ph, hd), e3 = 1.0f + ieee_div(c, ph);
    if (__float_as_uint(r3) != __float_as_uint(e3) && !(r3 != r3 && e3 != e3)) ++bad_helpers;
    float num = __uint_as_float(mix(h2 ^ 0x51ed27u));
    if (k == 4) num = -0.0f;
    if (k == 5) num = rand_float(h3, -100, -90, true);
    const float den = k == 6 ? 1.0f + rand_float(h1, -40, 30, false) : r1;
    const float r2 = div_by_den(num, den), e2 = ieee_div(num, den);
    if (__float_as_uint(r2) != __float_as_uint(e2) && !(r2 != r2 && e2 != e2)) ++bad_helpers;
  }
  if (bad) atomicAdd(&counts[0], bad);
  if (bad_helpers) atomicAdd(&counts[1], bad_helpers);
}

}  // namespace
}  // namespace rbs

using namespace rbs;

extern "C" int rbs_check_arithmetic(int64_t a_count, uint64_t *counts) {
  if (!counts || a_count < 1 || a_count > (1 << 23) || ((1 << 23) % a_count) != 0) {
    set_error("rbs_check_arithmetic: a_count must divide 2^23 and counts must be given");
    return RBS_E_ARG;
  }
  RBS_API_SHARED;
  unsigned long long *d = nullptr;
  RBS_HIP(hipMalloc(&d, 12 * sizeof(unsigned long long)));
  hipError_t err = hipMemset(d, 0, 12 * sizeof(unsigned long long));
  const uint32_t lo = 0x3a000000u, hi = 0x4a000000u;  // [2^-11, 2^21)
  if (err == hipSuccess) {
    check_pow<7><<<8192, 256>>>(lo, hi - lo, d + 0);
    check_pow<4><<<8192, 256>>>(lo, hi - lo, d + 2);
    check_pow<1><<<8192, 256>>>(lo, hi - lo, d + 4);
    check_rcp_scale<<<(1 << 23) / 256, 256>>>(d + 6);
    const uint32_t stride = (uint32_t)((1 << 23) / a_count);
    for (uint32_t j0 = 0; j0 < (uint32_t)a_count; j0 += 4096) {
      const uint32_t nj = (uint32_t)a_count - j0 < 4096u ? (uint32_t)a_count - j0 : 4096u;
      check_div_pairs<<<(1 << 23) / 256, 256>>>(j0, nj, stride, d + 7);
    }
    check_div_random<<<16384, 256>>>(20261204u, 256, d + 8);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  unsigned long long h[12] = {};
  if (err == hipSuccess) err = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(d);
  if (err != hipSuccess) return hip_fail(err, "rbs_check_arithmetic");
  // counts: pow mismatches and exact-path inputs for 7/3, 4/3, 1/3; rcp scale
  // mismatches; division significand-pair mismatches; random-pair mismatches; helper
  // mismatches; then the numbers checked
  counts[0] = h[0];
  counts[1] = h[2];
  counts[2] = h[4];
  counts[3] = h[1];
  counts[4] = h[3];
  counts[5] = h[5];
  counts[6] = h[6];
  counts[7] = h[7];
  counts[8] = h[8];
  counts[9] = h[9];
  counts[10] = hi - lo;
  counts[11] = (uint64_t)a_count << 23;
  return RBS_OK;
}
