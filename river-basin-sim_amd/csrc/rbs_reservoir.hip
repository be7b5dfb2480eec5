// rbs_reservoir.hip -- a reservoir's release on the slow clock, in one launch.
//
// ReservoirOperator.advance (processes/reservoir.py:219-270) reads the pool stage at
// the gauge cell, asks the release rule for a discharge (io/config.py:239-255), and
// moves V = min(Q dt_slow, pool volume) from the pool to the outlet:
//   _withdraw_pool (:83-105): each pool cell loses scale*h, scale = f32(V / stored),
//     and the exact float32 change of every cell is summed in float64;
//   _deliver_release (:108-130): each outlet cell gains f32(removed / A / n), and the
//     float64 difference between that share and what the float32 addition landed is
//     banked as a sink.
// The reference does this in three host round trips (the gauge read, the pool sum,
// the sum of what left).  Here one block of 1024 threads does it all on the device,
// with fixed-order float64 block sums; only the 32-byte record comes back.
#include <cmath>

#include "rbs_device.hpp"

namespace rbs {

constexpr int kResvThreads = 1024;

struct ResvArgs {
  float *h;
  const float *z;
  DevCtl *ctl;
  rbs_reservoir r;
  double t, dt_slow, area;
  int64_t pitch;
  rbs_release *rec;
};

// Sum of one double per thread, the same order every launch; the result is returned
// to every thread.
__device__ double block_sum(double v, double *s_part) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();  // s_part may still be read by the previous sum
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double total = 0.0;
  for (int w = 0; w < kResvThreads / 64; ++w) total += s_part[w];
  return total;
}

// Structure.discharge_at (config.py:239-255); stage NaN = a dry pool.
__device__ double release_rule(const rbs_reservoir &r, double stage) {
  if (isnan(stage)) return 0.0;
  if (r.rule == RBS_RULE_FIXED) return r.release_m3_s;
  if (r.rule != RBS_RULE_TARGET_STAGE) return 0.0;
  const double span = r.crest_m - r.target_stage_m;
  const double frac = (stage - r.target_stage_m) / span;
  const double lo = 0.0 > frac ? 0.0 : frac;  // Python's max(frac, 0.0)
  return r.release_max_m3_s * (1.0 < lo ? 1.0 : lo);  // and min(lo, 1.0)
}

__global__ __launch_bounds__(kResvThreads) void reservoir_kernel(ResvArgs a) {
  __shared__ double s_part[kResvThreads / 64];
  const rbs_reservoir &r = a.r;
  const int64_t P = a.pitch;
  const int tid = threadIdx.x;
  const int pr = r.pool[2] - r.pool[0] + 1, pc = r.pool[3] - r.pool[1] + 1;
  const int orr = r.outlet[2] - r.outlet[0] + 1, oc = r.outlet[3] - r.outlet[1] + 1;
  const int64_t n_pool = (int64_t)pr * pc, n_out = (int64_t)orr * oc;
  auto pool_at = [&](int64_t k) {
    const int64_t i = k / pc, j = k - i * pc;
    return (r.pool[0] + i) * P + (r.pool[1] + j);
  };
  auto out_at = [&](int64_t k) {
    const int64_t i = k / oc, j = k - i * oc;
    return (r.outlet[0] + i) * P + (r.outlet[1] + j);
  };

  // pool_stage (reservoir.py:198-217): z + h at the gauge, None (NaN) when h <= H_DRY
  const int64_t og = (int64_t)r.gauge_row * P + r.gauge_col;
  const double hg = (double)a.h[og];
  const double stage = hg <= kHDry64 ? NAN : (double)a.z[og] + hg;
  const double q = release_rule(r, stage);
  double volume = 0.0;
  if (q > 0.0) {  // block-uniform
    const double requested = q * a.dt_slow;
    double s = 0.0;
    for (int64_t k = tid; k < n_pool; k += kResvThreads) s += (double)a.h[pool_at(k)];
    const double stored = block_sum(s, s_part) * a.area;
    if (stored > 0.0) {
      const double x = requested / stored;
      const float scale = (float)(x < 1.0 ? x : 1.0);  // min(1.0, x), then float32
      double taken = 0.0;
      for (int64_t k = tid; k < n_pool; k += kResvThreads) {
        const int64_t o = pool_at(k);
        const float avail = a.h[o];
        const float h_new = avail - avail * scale;
        a.h[o] = h_new;
        taken += (double)avail - (double)h_new;
      }
      const double removed = block_sum(taken, s_part) * a.area;
      if (removed > 0.0) {
        const double add = removed / a.area / (double)n_out;
        const float add32 = (float)add;
        double shortfall = 0.0;
        for (int64_t k = tid; k < n_out; k += kResvThreads) {
          const int64_t o = out_at(k);
          const float avail = a.h[o];
          const float h_new = avail + add32;
          a.h[o] = h_new;
          shortfall += add - ((double)h_new - (double)avail);
        }
        const double banked = block_sum(shortfall, s_part);
        if (tid == 0) kahan_add64(a.ctl->release_out, a.ctl->release_out_c, banked * a.area);
        volume = removed;
      }
    }
  }
  if (tid == 0) {
    a.rec->time = a.t;
    a.rec->stage = stage;
    a.rec->discharge_m3_s = q;
    a.rec->volume_m3 = volume;
  }
}

}  // namespace rbs

using namespace rbs;

extern "C" int rbs_reservoir_release(rbs_state *s, const rbs_reservoir *r, double t,
                                     double dt_slow, rbs_release *out) {
  RBS_API_SHARED;
  if (!s || !r || !out) return set_error("null argument"), RBS_E_ARG;
  auto box_ok = [&](const int32_t *b) {
    return 0 <= b[0] && b[0] <= b[2] && b[2] < s->ny && 0 <= b[1] && b[1] <= b[3] &&
           b[3] < s->nx;
  };
  if (!box_ok(r->pool)) return set_error("reservoir pool box is outside the grid"), RBS_E_ARG;
  if (!box_ok(r->outlet)) return set_error("reservoir outlet box is outside the grid"), RBS_E_ARG;
  if (!(r->outlet[2] < r->pool[0] || r->outlet[0] > r->pool[2] || r->outlet[3] < r->pool[1] ||
        r->outlet[1] > r->pool[3]))
    return set_error("reservoir outlet overlaps its pool"), RBS_E_ARG;
  if (r->gauge_row < r->pool[0] || r->gauge_row > r->pool[2] || r->gauge_col < r->pool[1] ||
      r->gauge_col > r->pool[3])
    return set_error("reservoir gauge cell is outside its pool"), RBS_E_ARG;
  if (r->rule != RBS_RULE_FIXED && r->rule != RBS_RULE_TARGET_STAGE)
    return set_error("unknown release rule %d", r->rule), RBS_E_ARG;
  if (!(dt_slow >= 0.0) || !std::isfinite(dt_slow)) return set_error("bad dt_slow %g", dt_slow), RBS_E_ARG;
  if (!s->resv_rec) RBS_HIP(hipMalloc(&s->resv_rec, sizeof(rbs_release)));
  rbs_release *d_rec = s->resv_rec;
  ResvArgs a{};
  a.h = s->h[s->cur];
  a.z = s->z;
  a.ctl = s->ctl;
  a.r = *r;
  a.t = t;
  a.dt_slow = dt_slow;
  a.area = s->dx * s->dx;
  a.pitch = s->pitch;
  a.rec = d_rec;
  hipLaunchKernelGGL(reservoir_kernel, dim3(1), dim3(kResvThreads), 0, s->stream, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_rec, sizeof(rbs_release), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(e, "reservoir release");
  s->speed_scheme = -1;  // depths moved: the cached CFL reduction and volume are stale
  return RBS_OK;
}
