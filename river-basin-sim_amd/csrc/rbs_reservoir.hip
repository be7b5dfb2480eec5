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
// the sum of what left).  Here four launches on the state's stream do it on the
// device: the pool sum and the withdrawal over many blocks (the pool's rows dealt
// round-robin, per-block partials), the scale and the delivery in one block each, all
// sums float64 in a fixed order; only the 32-byte record comes back.
#include <cmath>
#include <mutex>

#include "rbs_device.hpp"

namespace rbs {

constexpr int kResvThreads = 256;   // the pool passes: many blocks of one row stride
constexpr int kResvBlocks = 2048;   // at most; the pool's rows are dealt round-robin
constexpr int kDeliverThreads = 1024;

struct ResvArgs {
  float *h;
  const float *z;
  DevCtl *ctl;
  rbs_reservoir r;
  double t, dt_slow, area;
  int64_t pitch;
};

// What the launches of one release hand on to each other.
struct ResvScratch {
  double stage, q;
  float scale;
  int32_t go;  // 1: the pool holds water and the rule asks for some
};

// Fixed-order sum of one double per thread over a block of NT; every thread gets it.
template <int NT>
__device__ double block_sum(double v, double *s_part) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();  // s_part may still be read by a previous sum
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double total = 0.0;
  for (int w = 0; w < NT / 64; ++w) total += s_part[w];
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

// pool_stage (reservoir.py:198-217): z + h at the gauge, None (NaN) when h <= H_DRY.
__device__ double gauge_stage(const ResvArgs &a) {
  const int64_t og = (int64_t)a.r.gauge_row * a.pitch + a.r.gauge_col;
  const double hg = (double)a.h[og];
  return hg <= kHDry64 ? NAN : (double)a.z[og] + hg;
}

// 1. Partial sums of f64(h) over the pool, the pool's rows dealt round-robin to blocks.
__global__ __launch_bounds__(kResvThreads) void resv_sum_kernel(ResvArgs a, double *part) {
  __shared__ double s_part[kResvThreads / 64];
  const rbs_reservoir &r = a.r;
  double s = 0.0;
  if (release_rule(r, gauge_stage(a)) > 0.0) {  // block-uniform
    for (int i = r.pool[0] + blockIdx.x; i <= r.pool[2]; i += gridDim.x) {
      const float *row = a.h + (int64_t)i * a.pitch;
      for (int j = r.pool[1] + threadIdx.x; j <= r.pool[3]; j += kResvThreads) s += (double)row[j];
    }
  }
  s = block_sum<kResvThreads>(s, s_part);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// 2. One block: the stage, the rule, the pool volume and the withdrawal factor.
__global__ __launch_bounds__(kResvThreads) void resv_scale_kernel(ResvArgs a, const double *part,
                                                                  int nb, ResvScratch *sc) {
  __shared__ double s_part[kResvThreads / 64];
  const double stage = gauge_stage(a);
  const double q = release_rule(a.r, stage);
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += kResvThreads) s += part[k];
  const double stored = block_sum<kResvThreads>(s, s_part) * a.area;
  if (threadIdx.x == 0) {
    sc->stage = stage;
    sc->q = q;
    sc->go = 0;
    sc->scale = 0.0f;
    if (q > 0.0 && stored > 0.0) {
      const double x = (q * a.dt_slow) / stored;
      sc->scale = (float)(x < 1.0 ? x : 1.0);  // min(1.0, x), then float32
      sc->go = 1;
    }
  }
}

// 3. _withdraw_pool (:83-105): h -= h * scale per pool cell, partial sums of what left.
__global__ __launch_bounds__(kResvThreads) void resv_withdraw_kernel(ResvArgs a,
                                                                     const ResvScratch *sc,
                                                                     double *part) {
  __shared__ double s_part[kResvThreads / 64];
  const rbs_reservoir &r = a.r;
  double taken = 0.0;
  if (sc->go) {
    const float scale = sc->scale;
    for (int i = r.pool[0] + blockIdx.x; i <= r.pool[2]; i += gridDim.x) {
      float *row = a.h + (int64_t)i * a.pitch;
      for (int j = r.pool[1] + threadIdx.x; j <= r.pool[3]; j += kResvThreads) {
        const float avail = row[j];
        const float h_new = avail - avail * scale;
        row[j] = h_new;
        taken += (double)avail - (double)h_new;
      }
    }
  }
  taken = block_sum<kResvThreads>(taken, s_part);
  if (threadIdx.x == 0) part[blockIdx.x] = taken;
}

// 4. One block: what left the pool, _deliver_release (:108-130) over the outlet with
// its rounding banked, and the record.
__global__ __launch_bounds__(kDeliverThreads) void resv_deliver_kernel(ResvArgs a,
                                                                       const double *part, int nb,
                                                                       const ResvScratch *sc,
                                                                       rbs_release *rec) {
  __shared__ double s_part[kDeliverThreads / 64];
  const rbs_reservoir &r = a.r;
  double volume = 0.0;
  if (sc->go) {  // block-uniform
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += kDeliverThreads) s += part[k];
    const double removed = block_sum<kDeliverThreads>(s, s_part) * a.area;
    if (removed > 0.0) {
      const int oc = r.outlet[3] - r.outlet[1] + 1;
      const int64_t n_out = (int64_t)(r.outlet[2] - r.outlet[0] + 1) * oc;
      const double add = removed / a.area / (double)n_out;
      const float add32 = (float)add;
      double shortfall = 0.0;
      for (int i = r.outlet[0]; i <= r.outlet[2]; ++i) {
        float *row = a.h + (int64_t)i * a.pitch;
        for (int j = r.outlet[1] + threadIdx.x; j <= r.outlet[3]; j += kDeliverThreads) {
          const float avail = row[j];
          const float h_new = avail + add32;
          row[j] = h_new;
          shortfall += add - ((double)h_new - (double)avail);
        }
      }
      const double banked = block_sum<kDeliverThreads>(shortfall, s_part);
      if (threadIdx.x == 0) kahan_add64(a.ctl->release_out, a.ctl->release_out_c, banked * a.area);
      volume = removed;
    }
  }
  if (threadIdx.x == 0) {
    rec->time = a.t;
    rec->stage = sc->stage;
    rec->discharge_m3_s = sc->q;
    rec->volume_m3 = volume;
  }
}

// The no-allocation variant's scratch: the record, the hand-off between launches and
// the block partials of one release, in a file-scope device array (one per GPU, held
// until the process ends).  Host code uses only the per-device address that
// hipGetSymbolAddress returns for it, never the array's name.
constexpr size_t kResvScratchMax = 32768;
static __device__ __align__(256) unsigned char scratch[kResvScratchMax];

// Bytes of scratch one release uses: the record, the hand-off, one partial per block.
static size_t release_scratch_bytes(int nb) {
  return sizeof(rbs_release) + sizeof(ResvScratch) + sizeof(double) * (size_t)nb;
}

}  // namespace rbs

using namespace rbs;

namespace {

int check_release(const rbs_state *s, const rbs_reservoir *r, double dt_slow,
                  const rbs_release *out) {
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
  return RBS_OK;
}

// Blocks of the pool passes: the pool's rows dealt round-robin, at most kResvBlocks.
int release_blocks(const rbs_reservoir *r) {
  const int pool_rows = r->pool[2] - r->pool[0] + 1;
  return pool_rows < kResvBlocks ? pool_rows : kResvBlocks;
}

// The four launches of one release on the state's stream, with `base` the device
// scratch of release_scratch_bytes(nb) bytes (record, hand-off, partials), then the
// record back to the host.
int launch_release(rbs_state *s, const rbs_reservoir *r, double t, double dt_slow,
                   rbs_release *out, void *base) {
  rbs_release *d_rec = static_cast<rbs_release *>(base);
  ResvScratch *sc = reinterpret_cast<ResvScratch *>(d_rec + 1);
  double *part = reinterpret_cast<double *>(sc + 1);
  ResvArgs a{};
  a.h = s->h[s->cur];
  a.z = s->z;
  a.ctl = s->ctl;
  a.r = *r;
  a.t = t;
  a.dt_slow = dt_slow;
  a.area = s->dx * s->dx;
  a.pitch = s->pitch;
  const int nb = release_blocks(r);
  hipLaunchKernelGGL(resv_sum_kernel, dim3(nb), dim3(kResvThreads), 0, s->stream, a, part);
  hipLaunchKernelGGL(resv_scale_kernel, dim3(1), dim3(kResvThreads), 0, s->stream, a, part, nb, sc);
  hipLaunchKernelGGL(resv_withdraw_kernel, dim3(nb), dim3(kResvThreads), 0, s->stream, a, sc, part);
  hipLaunchKernelGGL(resv_deliver_kernel, dim3(1), dim3(kDeliverThreads), 0, s->stream, a, part,
                     nb, sc, d_rec);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_rec, sizeof(rbs_release), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return hip_fail(e, "reservoir release");
  s->speed_scheme = -1;  // depths moved: the cached CFL reduction and volume are stale
  return RBS_OK;
}

// The address of `scratch` on each device, looked up once with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_at[kMaxDevices] = {};

int scratch_on(int device, void **out) {
  if (device < 0 || device >= kMaxDevices)
    return set_error("device %d is beyond the %d the library tracks", device, kMaxDevices),
           RBS_E_RANGE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_at[device]) {
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != device) e = hipSetDevice(device);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != device) {
      const hipError_t e2 = hipSetDevice(prev);
      if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error("no device address for the release scratch"), RBS_E_STATE;
    g_scratch_at[device] = p;
  }
  *out = g_scratch_at[device];
  return RBS_OK;
}

}  // namespace

extern "C" int rbs_reservoir_release(rbs_state *s, const rbs_reservoir *r, double t,
                                     double dt_slow, rbs_release *out) {
  RBS_API_SHARED;
  if (int rc = check_release(s, r, dt_slow, out)) return rc;
  if (!s->resv_rec) {
    // the record, the scratch and the block partials, kept for the state's life
    RBS_HIP(hipMalloc(&s->resv_rec, release_scratch_bytes(kResvBlocks)));
  }
  return launch_release(s, r, t, dt_slow, out, s->resv_rec);
}

extern "C" int rbs_reservoir_release_noalloc(rbs_state *s, const rbs_reservoir *r, double t,
                                             double dt_slow, rbs_release *out) {
  RBS_API_SHARED;
  if (int rc = check_release(s, r, dt_slow, out)) return rc;
  const size_t need = release_scratch_bytes(release_blocks(r));
  if (need > kResvScratchMax)
    return set_error("a release needs %zu bytes of scratch, more than the %zu the no-allocation "
                     "entry point holds",
                     need, kResvScratchMax),
           RBS_E_RANGE;
  void *base = nullptr;
  if (int rc = scratch_on(s->device, &base)) return rc;
  return launch_release(s, r, t, dt_slow, out, base);
}
