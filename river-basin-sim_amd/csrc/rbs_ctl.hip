// rbs_ctl.hip -- device clock, CFL reduction and float64 ledger.
//
// Replaces the per-step host round trip of the reference (compute_dt's
// atomic_max + h_max.numpy(), local_inertial.py:515-544 / hllc.py:972-990, and
// the scheduler's fill_span, scheduler.py:48-72) with device work: the step
// kernels leave per-block partials, and finalize_kernel reduces them in a fixed
// order, updates the Kahan-compensated float64 ledger (massbalance.py:39-158),
// advances the float64 clock and -- in a device-driven run -- chooses the next dt.
#include <type_traits>

#include "rbs_device.hpp"

namespace rbs {

// Maximum-only reduction of the CFL speed of the current state, into partials.
// LI: h (local_inertial.py:503-512), or the water-column depth when the state has
// sub-grid channels (cw != null, :487-500); HLLC: max(|u|,|v|) + sqrt(g h)
// (hllc.py:948-966).  Also sums f64(h) so a prime leaves the ledger volume current.
__global__ __launch_bounds__(kThreads) void reduce_speed_kernel(const float *h, const float *hu,
                                                                const float *hv, const float *cw,
                                                                const float *cd, double dx64,
                                                                int64_t pitch, int ny, int nx,
                                                                int scheme, Partials part,
                                                                int rows_per_blk) {
  const float dx = (float)dx64;
  const float rdx = 1.0f / dx;
  __shared__ double s_red[4 * (kThreads / 64)];
  const int r0 = blockIdx.x * rows_per_blk;
  float m = 0.0f;
  double vol = 0.0, z0 = 0.0, z1 = 0.0;
  for (int r = r0; r < r0 + rows_per_blk && r < ny; ++r) {
    for (int c = threadIdx.x; c < nx; c += kThreads) {
      const int64_t o = (int64_t)r * pitch + c;
      const float hh = h[o];
      vol += (double)hh;
      if (scheme == RBS_SCHEME_LI) {
        m = fmx(m, cw ? column_depth_u(hh, cw[o], cd[o], dx, rdx) : hh);
      } else {
        const float cc = sqrtf(kG * fmx(hh, 0.0f));
        float sp = cc;
        if (hh > kHDry) {
          const float au = fabsf(hu[o]) / hh;
          const float av = fabsf(hv[o]) / hh;
          sp = fmx(au, av) + cc;
        }
        m = fmx(m, sp);
      }
    }
  }
  block_reduce4(m, vol, z0, z1, s_red);
  if (threadIdx.x == 0) {
    part.speed_max[blockIdx.x] = m;
    part.vol[blockIdx.x] = vol;
    part.infil[blockIdx.x] = 0.0;
    part.edge[blockIdx.x] = 0.0;
  }
}

// Host-driven step: write the scalars the step kernels read.
__global__ void set_scalars_kernel(DevCtl *ctl, double t, double dt, float rain, float scale,
                                   float s0, float s1, float s2, float s3) {
  ctl->t = t;
  ctl->dt = dt;
  ctl->rain = rain;
  ctl->rain_scale = scale;
  ctl->stage[0] = s0;
  ctl->stage[1] = s1;
  ctl->stage[2] = s2;
  ctl->stage[3] = s3;
  ctl->active = 1;
}

// boundaries.py:49-70 stage_at: piecewise linear in float64, held outside its range.
__device__ double stage_at_dev(const double *t, const double *lv, int n, double x) {
  if (x <= t[0]) return lv[0];
  if (x >= t[n - 1]) return lv[n - 1];
  for (int k = 0; k + 1 < n; ++k) {
    const double t0 = t[k], t1 = t[k + 1];
    if (t0 <= x && x <= t1) {
      if (t1 == t0) return lv[k + 1];
      return lv[k] + (lv[k + 1] - lv[k]) * (x - t0) / (t1 - t0);
    }
  }
  return lv[n - 1];
}

// Next step of a device-driven run (scheduler.py:171-229 + run.py:378-388).
__device__ void plan_next(DevCtl *ctl, const double *sync, double dx) {
  const double t = ctl->t;
  if (!(t < ctl->t_stop - kEpsT) || ctl->steps >= ctl->step_limit) {
    ctl->active = 0;
    return;
  }
  double dt_raw;
  const double sm = (double)ctl->speed_max;
  if (ctl->run_scheme == RBS_SCHEME_LI) {
    // local_inertial.py:540-544
    if (sm <= kHDry64) {
      dt_raw = ctl->dt_max;
    } else {
      dt_raw = ctl->alpha * dx / sqrt(kGravity * sm);
      dt_raw = dt_raw < ctl->dt_max ? dt_raw : ctl->dt_max;
    }
  } else {
    // hllc.py:987-990
    if (sm <= 0.0) {
      dt_raw = ctl->dt_max;
    } else {
      dt_raw = ctl->alpha * dx / sm;
      dt_raw = dt_raw < ctl->dt_max ? dt_raw : ctl->dt_max;
    }
  }
  int k = ctl->sync_idx;
  while (k < ctl->n_sync && !(sync[k] > t + kEpsT)) ++k;  // next_sync (scheduler.py:165-168)
  ctl->sync_idx = k;
  const double ns = k < ctl->n_sync ? sync[k] : INFINITY;
  const double dt = fill_span_dev(ns - t, dt_raw);
  if (!(dt_raw > 0.0) || !(dt > 0.0)) {  // scheduler.py:221-227: fail loudly
    ctl->active = -1;
    return;
  }
  ctl->dt = dt;
  const bool raining = t < ctl->rain_duration - kEpsT;  // run.py:378
  ctl->rain = ctl->run_field ? 0.0f : (raining ? (float)ctl->rain_m_s : 0.0f);
  ctl->rain_scale = raining ? 1.0f : 0.0f;
  // hllc.py:1308-1312: stages at the step midpoint, narrowed to float32
  for (int e = 0; e < 4; ++e)
    if (ctl->stage_n[e] > 0)
      ctl->stage[e] = (float)stage_at_dev(ctl->stage_t[e], ctl->stage_lv[e], ctl->stage_n[e],
                                          t + 0.5 * dt);
  ctl->active = 1;
}

// Reduce the step's (or prime's) partials in fixed order and update clock + ledger.
// mode: FIN_PRIME (no step was taken), FIN_EXPLICIT (host-driven step),
// FIN_DRIVEN (device-driven step; plans the next one).
// Thread t folds partials t, t + NT, t + 2 NT, ... in that order (sixteen loads per
// array in flight per round: an 8192^2 HLLC state has 72 704 blocks; then four, two
// and one, so that a thread's last few partials do not wait for one another), then the
// block reduces the NT results in a fixed tree.
template <int NT>
__device__ void finalize_body(DevCtl *ctl, Partials part, int nblk, double area, double n_cells,
                              const double *sync, double dx, int mode, int flip) {
  __shared__ double s_red[4 * (NT / 64)];
  // the partials are read before the run's state is known (reads only, in bounds), so
  // the two loads' latencies overlap
  const int active = ctl->active;
  float m = 0.0f;
  double vol = 0.0, inf = 0.0, edge = 0.0;
  int b = threadIdx.x;
  // U of the thread's partials loaded together, then folded in index order (the same
  // order as one at a time): sixteen, four, two, one
  auto fold = [&](auto u_tag) {
    constexpr int U = decltype(u_tag)::value;
    for (; b + (U - 1) * NT < nblk; b += U * NT) {
      float mk[U];
      double vk[U], ik[U], ek[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        mk[j] = part.speed_max[b + j * NT];
        vk[j] = part.vol[b + j * NT];
        ik[j] = part.infil[b + j * NT];
        ek[j] = part.edge[b + j * NT];
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        m = fmx(m, mk[j]);
        vol += vk[j];
        inf += ik[j];
        edge += ek[j];
      }
    }
  };
  fold(std::integral_constant<int, 16>{});
  fold(std::integral_constant<int, 4>{});
  fold(std::integral_constant<int, 2>{});
  fold(std::integral_constant<int, 1>{});
  if (mode == FIN_DRIVEN && active != 1) return;  // block-uniform
  block_reduce4<NT>(m, vol, inf, edge, s_red);
  if (threadIdx.x != 0) return;
  ctl->speed_max = m;
  ctl->stale = 0;
  ctl->volume = vol * area;
  if (mode == FIN_PRIME) return;
  if (mode == FIN_PRIME_DRIVEN) {
    plan_next(ctl, sync, dx);
    return;
  }
  kahan_add64(ctl->infil_out, ctl->infil_out_c, inf * area);
  kahan_add64(ctl->edge_out, ctl->edge_out_c, edge * area);
  const double dt = ctl->dt;
  if (mode == FIN_DRIVEN) {
    // run.py:380-388: the ledger banks the analytic rain volume of the step.
    if (ctl->run_field) {
      if (ctl->rain_scale != 0.0f) kahan_add64(ctl->rain_in, ctl->rain_in_c, ctl->rain_field_sum * dt * area);
    } else {
      const double r = ctl->rain_scale != 0.0f ? ctl->rain_m_s : 0.0;
      if (r > 0.0) kahan_add64(ctl->rain_in, ctl->rain_in_c, r * dt * area * n_cells);
    }
  }
  ctl->t = ctl->t + dt;
  ctl->steps += 1;
  if (flip) ctl->cur ^= 1;  // the LI step wrote the other buffer
  if (mode == FIN_DRIVEN) plan_next(ctl, sync, dx);
  else ctl->active = 0;
}

// four waves: at 4096^2 (756 partials) 5.5 us per finalize against 6.5 us for sixteen
constexpr int kFinThreads = 256;

// The first level of the fold for states with many blocks: one block reads at about
// one CU's bandwidth, 44 us for the 72 704 blocks of an 8192^2 HLLC state.  Block g
// folds partials [g c, g c + c) (c = ceil(nblk / kFoldBlocks)) in the same fixed order
// as finalize_body, into entry g of `out`; the finalize then folds the kFoldBlocks
// entries.  The order is fixed for a given block count, so the result is deterministic.
__global__ __launch_bounds__(kThreads) void fold_partials_kernel(const DevCtl *ctl, Partials part,
                                                                 int nblk, int mode,
                                                                 Partials out) {
  __shared__ double s_red[4 * (kThreads / 64)];
  if (mode == FIN_DRIVEN && ctl->active != 1) return;
  const int chunk = (nblk + kFoldBlocks - 1) / kFoldBlocks;
  const int b0 = blockIdx.x * chunk;
  const int b1 = b0 + chunk < nblk ? b0 + chunk : nblk;
  float m = 0.0f;
  double vol = 0.0, inf = 0.0, edge = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += kThreads) {
    m = fmx(m, part.speed_max[b]);
    vol += part.vol[b];
    inf += part.infil[b];
    edge += part.edge[b];
  }
  block_reduce4<kThreads>(m, vol, inf, edge, s_red);
  if (threadIdx.x == 0) {
    out.speed_max[blockIdx.x] = m;
    out.vol[blockIdx.x] = vol;
    out.infil[blockIdx.x] = inf;
    out.edge[blockIdx.x] = edge;
  }
}


__global__ __launch_bounds__(kFinThreads) void finalize_kernel(DevCtl *ctl, Partials part, int nblk,
                                                               double area, double n_cells,
                                                               const double *sync, double dx,
                                                               int mode, int flip) {
  finalize_body<kFinThreads>(ctl, part, nblk, area, n_cells, sync, dx, mode, flip);
}

// The finalize of `nblk` block partials: through the two-level fold above
// kFoldMin blocks, directly below.
static int launch_finalize(rbs_state *s, int nblk, int mode, int flip) {
  Partials part = s->part;
  if (nblk > kFoldMin) {
    hipLaunchKernelGGL(fold_partials_kernel, dim3(kFoldBlocks), dim3(kThreads), 0, s->stream,
                       s->ctl, s->part, nblk, mode, s->fold);
    RBS_HIP(hipGetLastError());
    part = s->fold;
    nblk = kFoldBlocks;
  }
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kFinThreads), 0, s->stream, s->ctl, part,
                     nblk, s->dx * s->dx, (double)s->ny * (double)s->nx, s->d_sync, s->dx, mode,
                     flip);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

// One block per batch member: that member's device-driven finalize, then its
// status word (bit 0: still active, bit 1: current LI buffer) for the host.  The grid
// covers the batch's capacity; blocks past the live member count return at once.
__global__ __launch_bounds__(kThreads) void batch_finalize_kernel(const FinArgs *fin,
                                                                  const BatchDims *dims,
                                                                  int32_t *status) {
  if ((int)blockIdx.x >= dims->members) return;
  const FinArgs f = fin[blockIdx.x];
  finalize_body<kThreads>(f.ctl, f.part, f.nblk, f.area, f.n_cells, f.sync, f.dx, FIN_DRIVEN, 1);
  __syncthreads();
  if (threadIdx.x == 0) status[blockIdx.x] = (f.ctl->active == 1 ? 1 : 0) | (f.ctl->cur << 1) |
                                             (f.ctl->active == -1 ? 4 : 0);
}

int batch_finalize_launch(const FinArgs *fin, const BatchDims *dims, int grid, int32_t *status,
                          hipStream_t stream) {
  hipLaunchKernelGGL(batch_finalize_kernel, dim3(grid), dim3(kThreads), 0, stream, fin, dims,
                     status);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

int ensure_partials(rbs_state *s, int nblk) {
  if (!s->fold.speed_max) {
    RBS_HIP(hipMalloc(&s->fold.speed_max, sizeof(float) * kFoldBlocks));
    RBS_HIP(hipMalloc(&s->fold.vol, sizeof(double) * kFoldBlocks));
    RBS_HIP(hipMalloc(&s->fold.infil, sizeof(double) * kFoldBlocks));
    RBS_HIP(hipMalloc(&s->fold.edge, sizeof(double) * kFoldBlocks));
  }
  if (nblk <= s->part_cap) return RBS_OK;
  if (s->part.speed_max) {
    RBS_HIP(hipFree(s->part.speed_max));
    RBS_HIP(hipFree(s->part.vol));
    RBS_HIP(hipFree(s->part.infil));
    RBS_HIP(hipFree(s->part.edge));
    RBS_HIP(hipFree(s->quiet));
    RBS_HIP(hipFree(s->hquiet));
  }
  RBS_HIP(hipMalloc(&s->part.speed_max, sizeof(float) * nblk));
  RBS_HIP(hipMalloc(&s->part.vol, sizeof(double) * nblk));
  RBS_HIP(hipMalloc(&s->part.infil, sizeof(double) * nblk));
  RBS_HIP(hipMalloc(&s->part.edge, sizeof(double) * nblk));
  RBS_HIP(hipMalloc(&s->quiet, sizeof(uint32_t) * nblk));
  RBS_HIP(hipMalloc(&s->hquiet, sizeof(uint32_t) * nblk));
  s->part_cap = nblk;
  return quiet_reset(s, s->stream);
}

int quiet_reset(rbs_state *s, hipStream_t stream) {
  if (s->quiet && s->part_cap > 0)
    RBS_HIP(hipMemsetAsync(s->quiet, 0, sizeof(uint32_t) * s->part_cap, stream));
  if (s->hquiet && s->part_cap > 0)
    RBS_HIP(hipMemsetAsync(s->hquiet, 0, sizeof(uint32_t) * s->part_cap, stream));
  return RBS_OK;
}

static int reduce_blocks(const rbs_state *s, int *rows_per_blk) {
  // about 4 rows of 4096 per block; at least one row
  int rpb = (16384 + s->nx - 1) / s->nx;
  if (rpb < 1) rpb = 1;
  *rows_per_blk = rpb;
  return (s->ny + rpb - 1) / rpb;
}

int reduce_speed_launch(rbs_state *s, int scheme, int mode) {
  int rpb = 1;
  const int nblk = reduce_blocks(s, &rpb);
  if (nblk > s->part_cap) {
    set_error("partial buffer too small");
    return RBS_E_STATE;
  }
  const float *h = s->h[s->cur];
  const bool column = scheme == RBS_SCHEME_LI && (s->armed & RBS_ARM_CHAN);
  hipLaunchKernelGGL(reduce_speed_kernel, dim3(nblk), dim3(kThreads), 0, s->stream, h, s->hu,
                     s->hv, column ? s->ch_w : nullptr, column ? s->ch_d : nullptr, s->dx,
                     s->pitch, s->ny, s->nx, scheme, s->part, rpb);
  RBS_HIP(hipGetLastError());
  return launch_finalize(s, nblk, mode, 0);
}

int finalize_launch(rbs_state *s, int mode, int scheme) {
  const int nblk = scheme == RBS_SCHEME_HLLC ? hllc_num_blocks(s) : li_num_blocks(s);
  return launch_finalize(s, nblk, mode, scheme == RBS_SCHEME_LI ? 1 : 0);
}

int set_scalars_launch(rbs_state *s, double t, double dt, float rain, float rain_scale,
                       const float stage[4]) {
  const float z4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const float *st = stage ? stage : z4;
  hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(1), 0, s->stream, s->ctl, t, dt, rain,
                     rain_scale, st[0], st[1], st[2], st[3]);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

int reduce_partials_needed(const rbs_state *s) {
  int rpb = 1;
  return reduce_blocks(s, &rpb);
}

}  // namespace rbs
