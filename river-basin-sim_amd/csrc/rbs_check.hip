// rbs_check.hip -- the device's own proof that the local-inertial face's short
// arithmetic routes give the bits of the long ones (rbs_check_arithmetic).
//
// The LI face evaluates the reference's friction power h^float32(7/3) correctly rounded
// and its two divisions as IEEE quotients (rbs_device.hpp: manning_update).  Shorter
// routes stand in for the long sequences where they are exact: pow_e73_hw (hardware
// log2/exp2 seeds and a float64 correction, with a midpoint guard that hands doubtful
// inputs to pow_third) and div_rn (reciprocal, a Newton step and two Markstein
// corrections, where the operands keep every step normal).  Their exactness rests on
// the hardware's v_log_f32, v_exp_f32 and v_rcp_f32, so this check runs on the GPU that
// will run the solver:
//   - every float32 h in [2^-11, 2^21]: where the short power claims exactness it must
//     equal pow_third;
//   - v_rcp_f32(m 2^k) = v_rcp_f32(m) 2^-k for every significand m, k in [-100, 100]
//     (a quotient's significand pair then fixes its bits at any exponents that keep the
//     steps normal);
//   - div_rn against the IEEE division on every significand of b with `a_count`
//     significands of a (all 2^23 of them prove the identity for the whole range);
//   - div_rn on 2^31 random pairs of the face's ranges, and manning_update against the
//     plain IEEE expression num / (1 + c / h^float32(7/3)) on random operands over the
//     whole float32 range (zeros, subnormals, infinities, NaN).
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
__global__ void check_pow(uint32_t lo_bits, uint32_t n, unsigned long long *counts) {
  unsigned long long bad = 0, slow = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float x = __uint_as_float(lo_bits + i);
    bool exact;
    const float f = pow_e73_hw(x, exact);
    if (!exact) ++slow;
    else if (__float_as_uint(f) != __float_as_uint(pow_third<7, false>(x))) ++bad;
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

// counts[0]: div_rn on the face's ranges; counts[1]: manning_update against the IEEE
// expression
__global__ void check_div_random(uint32_t seed, int iters, unsigned long long *counts) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long bad = 0, bad_face = 0;
  for (int it = 0; it < iters; ++it) {
    const uint32_t h1 = mix(seed ^ mix(t * 2654435761u + (uint32_t)it));
    const uint32_t h2 = mix(h1 + 0x9e3779b9u), h3 = mix(h2 + 0x85ebca6bu);
    // num / den, then c / p
    float a = rand_float(h1, -96, 96, true), b = rand_float(h2, 0, 24, false);
    if (__float_as_uint(div_rn(a, b)) != __float_as_uint(ieee_div(a, b))) ++bad;
    a = rand_float(h2, -50, 45, false);
    b = rand_float(h3, -24, 47, false);
    if (__float_as_uint(div_rn(a, b)) != __float_as_uint(ieee_div(a, b))) ++bad;
    // the face: a depth of at least H_DRY (now and then far beyond 2^20), c >= +0 and num
    // of any bits, with zeros, tiny and huge values mixed in
    const uint32_t k = h3 & 15u;
    const float h = k == 0 ? rand_float(h1, 19, 40, false)
                           : fmx(kHDry, rand_float(h1 ^ h3, -10, 20, false));
    float c = __uint_as_float(mix(h3) & 0x7fffffffu);
    if (k == 1) c = 0.0f;
    if (k == 2) c = rand_float(h2, -60, -40, false);
    if (k == 3) c = rand_float(h2, 40, 50, false);
    if (k >= 8) c = rand_float(h2, -30, 10, false);
    float num = __uint_as_float(mix(h2 ^ 0x51ed27u));
    if (k == 4) num = -0.0f;
    if (k == 5) num = rand_float(h3, -100, -90, true);
    if (k >= 8) num = rand_float(h3, -20, 10, true);
    const float q = manning_update(num, c, h);
    const float e = num / (1.0f + ieee_div(c, pow_third<7, false>(h)));
    if (__float_as_uint(q) != __float_as_uint(e) && !(q != q && e != e)) ++bad_face;
  }
  if (bad) atomicAdd(&counts[0], bad);
  if (bad_face) atomicAdd(&counts[1], bad_face);
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
  RBS_HIP(hipMalloc(&d, 8 * sizeof(unsigned long long)));
  hipError_t err = hipMemset(d, 0, 8 * sizeof(unsigned long long));
  const uint32_t lo = 0x3a000000u, hi = 0x4a000000u;  // [2^-11, 2^21)
  if (err == hipSuccess) {
    check_pow<<<8192, 256>>>(lo, hi - lo, d + 0);
    check_rcp_scale<<<(1 << 23) / 256, 256>>>(d + 2);
    const uint32_t stride = (uint32_t)((1 << 23) / a_count);
    for (uint32_t j0 = 0; j0 < (uint32_t)a_count; j0 += 4096) {
      const uint32_t nj = (uint32_t)a_count - j0 < 4096u ? (uint32_t)a_count - j0 : 4096u;
      check_div_pairs<<<(1 << 23) / 256, 256>>>(j0, nj, stride, d + 3);
    }
    check_div_random<<<16384, 256>>>(20261204u, 256, d + 4);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipDeviceSynchronize();
  unsigned long long h[8] = {};
  if (err == hipSuccess) err = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(d);
  if (err != hipSuccess) return hip_fail(err, "rbs_check_arithmetic");
  for (int i = 0; i < 6; ++i) counts[i] = h[i];
  counts[6] = hi - lo;
  counts[7] = (uint64_t)a_count << 23;
  return RBS_OK;
}
