// rbs_sources.hip -- point sources and the Exner bed update.
//
// Point sources (processes/inflow.py:91-200): each source cell receives the
// float32 depth f32(Q(t + dt/2) * dt / A), Kahan-compensated per source entry; the
// ledger banks the float32 request.  In a device-driven run the hydrographs are
// evaluated on the device from the clock, so no host copy happens per step.
//
// Morphology activation (processes/morphology.py:223-307): exner_update
// (sediment.py:383-426) in float64, rebuild z = f32(f64(z0) + dz_cum)
// (sediment.py:429-441), clear the transport integrals but keep their Kahan debt
// (sediment.py:513-524).  eta is not stored on the device (every step forms it
// from h + z), so the reference's refresh of eta has nothing to do here.
#include <vector>

#include "rbs_device.hpp"

namespace rbs {

// config.py:111-121 Inflow.discharge_at: piecewise linear, zero outside the curve.
__device__ double discharge_at(const double *kt, const double *kq, int n, double t) {
  if (n < 1 || t < kt[0] || t > kt[n - 1]) return 0.0;
  for (int k = 0; k + 1 < n; ++k) {
    const double t0 = kt[k], t1 = kt[k + 1];
    if (t0 <= t && t <= t1) {
      if (t1 == t0) return kq[k + 1];
      return kq[k] + (kq[k + 1] - kq[k]) * (t - t0) / (t1 - t0);
    }
  }
  return kq[n - 1];
}

// One member's point sources for one step (the body of both launches below).
__device__ void inflow_body(const InflowArgs &a, int driven, double t_host, double dt_host,
                            double *out) {
  const DevCtl *ctl = a.ctl;
  if (driven && ctl->active != 1) return;
  // the buffer holding the current depth: LI ping-pongs (parity on the device);
  // HLLC states pass the same buffer twice
  float *h = ctl->cur ? a.h1 : a.h0;
  const double t = driven ? ctl->t : t_host;
  const double dt = driven ? ctl->dt : dt_host;
  const double t_mid = t + 0.5 * dt;  // inflow.py:180
  for (int k = threadIdx.x; k < a.count; k += blockDim.x) {
    const int n = a.off[k + 1] - a.off[k];
    const double v = discharge_at(a.kt + a.off[k], a.kq + a.off[k], n, t_mid) * dt;
    const float add = (float)(v / a.area);  // inflow.py:182
    const int64_t o = (int64_t)a.rows[k] * a.pitch + a.cols[k];
    float hh = h[o], cp = a.comp[k];
    kahan_add(hh, cp, add);  // _inject_point_sources_c (:91-113)
    h[o] = hh;
    a.comp[k] = cp;
    a.vol[k] = (double)add;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < a.count; ++k) s += a.vol[k];
    const double requested = s * a.area;  // inflow.py:200
    if (driven)
      kahan_add64(a.ctl->point_in, a.ctl->point_in_c, requested);
    else
      *out = requested;
  }
}

__global__ __launch_bounds__(kThreads) void inflow_kernel(InflowArgs a, int driven, double t_host,
                                                          double dt_host, double *out) {
  inflow_body(a, driven, t_host, dt_host, out);
}

// A batch's point sources: one block per member that has any (device-driven); the
// grid covers the batch's capacity and blocks past the live count return at once.
__global__ __launch_bounds__(kThreads) void inflow_batch_kernel(const InflowArgs *list,
                                                                const BatchDims *dims) {
  if ((int)blockIdx.x >= dims->inflow_members) return;
  inflow_body(list[blockIdx.x], 1, 0.0, 0.0, nullptr);
}

void inflow_fill_args(const rbs_state *s, InflowArgs *out) {
  const Inflows &in = s->inflow;
  InflowArgs a{};
  // an HLLC state keeps its depth in h[cur] without ping-pong
  a.h0 = (s->armed & RBS_ARM_HLLC) ? s->h[s->cur] : s->h[0];
  a.h1 = (s->armed & RBS_ARM_HLLC) ? s->h[s->cur] : s->h[1];
  a.pitch = s->pitch;
  a.count = in.count;
  a.rows = in.d_rows;
  a.cols = in.d_cols;
  a.off = in.d_off;
  a.kt = in.d_kt;
  a.kq = in.d_kq;
  a.comp = in.d_comp;
  a.vol = in.d_vol;
  a.ctl = s->ctl;
  a.area = s->dx * s->dx;
  *out = a;
}

int inflow_batch_launch(const InflowArgs *list, const BatchDims *dims, int grid,
                        hipStream_t stream) {
  if (grid < 1) return RBS_OK;
  hipLaunchKernelGGL(inflow_batch_kernel, dim3(grid), dim3(kThreads), 0, stream, list, dims);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

int inflow_launch(rbs_state *s, bool driven, double t, double dt) {
  Inflows &in = s->inflow;
  if (in.count < 1) return RBS_OK;
  InflowArgs a;
  inflow_fill_args(s, &a);
  hipLaunchKernelGGL(inflow_kernel, dim3(1), dim3(kThreads), 0, s->stream, a, driven ? 1 : 0, t,
                     dt, in.d_vol + in.count);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

// sediment.py:383-426 exner_update + :429-441 rebuild_bed, one cell per thread.
__global__ void exner_kernel(const float *qx, const float *qy, const float *lo, const float *hi,
                             const float *z0, double *dz_cum, double *dz_un, float *z,
                             int64_t pitch, int ny, int nx, double dx64, double iomp64) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= nx || r >= ny) return;
  const int64_t o = (int64_t)r * pitch + c;
  const double div_x = (double)qx[o + 1] - (double)qx[o];
  const double div_y = (double)qy[o + pitch] - (double)qy[o];
  const double want = dz_cum[o] - iomp64 * (div_x + div_y) / dx64;
  double nw = want;
  const double l = (double)lo[o], u = (double)hi[o];
  if (nw < l) nw = l;
  if (nw > u) nw = u;
  if (nw != want) dz_un[o] = dz_un[o] + (want - nw);
  dz_cum[o] = nw;
  z[o] = (float)((double)z0[o] + nw);
}

__global__ void zero_kernel(float *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

int exner_launch(rbs_state *s, double porosity) {
  // exner_update takes dx and 1/(1-p) as float32 kernel parameters (sediment.py:389-390)
  const double dx64 = (double)(float)s->dx;
  const double iomp64 = (double)(float)(1.0 / (1.0 - porosity));
  dim3 grid((s->nx + 255) / 256, s->ny);
  hipLaunchKernelGGL(exner_kernel, grid, dim3(256), 0, s->stream, s->qs_int_x, s->qs_int_y,
                     s->dz_lo, s->dz_hi, s->z0, s->dz_cum, s->dz_unapplied, s->z, s->pitch,
                     s->ny, s->nx, dx64, iomp64);
  RBS_HIP(hipGetLastError());
  // clear_transport_integral (sediment.py:513-524): integrals only, debt kept
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, s->stream, s->qs_int_x,
                     (size_t)s->ny * s->pitch);
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, s->stream, s->qs_int_y,
                     (size_t)(s->ny + 1) * s->pitch);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

// Bed-change totals {sum dz_cum, sum |dz_cum|, min, max, sum dz_unapplied}.  Block b
// folds rows [b*rpb, (b+1)*rpb) thread-strided, the block folds its threads in a
// fixed tree, and one block folds the blocks in order: the same bits every call.
constexpr int kSedThreads = 256;
constexpr int kSedQ = 5;

struct SedAcc {
  double v[kSedQ];
  __device__ SedAcc() : v{0.0, 0.0, INFINITY, -INFINITY, 0.0} {}
  __device__ void fold(const double *o) {
    v[0] += o[0];
    v[1] += o[1];
    v[2] = fmin(v[2], o[2]);
    v[3] = fmax(v[3], o[3]);
    v[4] += o[4];
  }
};

__device__ __forceinline__ void sed_block_fold(SedAcc &acc, double *sh) {
  const int t = threadIdx.x;
  for (int q = 0; q < kSedQ; ++q) sh[q * kSedThreads + t] = acc.v[q];
  __syncthreads();
  for (int w = kSedThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      double o[kSedQ], m[kSedQ];
      for (int q = 0; q < kSedQ; ++q) {
        m[q] = sh[q * kSedThreads + t];
        o[q] = sh[q * kSedThreads + t + w];
      }
      SedAcc x;
      for (int q = 0; q < kSedQ; ++q) x.v[q] = m[q];
      x.fold(o);
      for (int q = 0; q < kSedQ; ++q) sh[q * kSedThreads + t] = x.v[q];
    }
    __syncthreads();
  }
  for (int q = 0; q < kSedQ; ++q) acc.v[q] = sh[q * kSedThreads];
}

__global__ __launch_bounds__(kSedThreads) void sed_totals_kernel(const double *dz, const double *un,
                                                                 int64_t pitch, int ny, int nx,
                                                                 int rpb, double *part) {
  __shared__ double sh[kSedQ * kSedThreads];
  SedAcc acc;
  const int r0 = blockIdx.x * rpb;
  for (int r = r0; r < r0 + rpb && r < ny; ++r)
    for (int c = threadIdx.x; c < nx; c += kSedThreads) {
      const int64_t o = (int64_t)r * pitch + c;
      const double d = dz[o];
      const double x[kSedQ] = {d, fabs(d), d, d, un[o]};
      acc.fold(x);
    }
  sed_block_fold(acc, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < kSedQ; ++q) part[kSedQ * blockIdx.x + q] = acc.v[q];
}

__global__ __launch_bounds__(kSedThreads) void sed_totals_final(const double *part, int nblk,
                                                                double *out) {
  __shared__ double sh[kSedQ * kSedThreads];
  SedAcc acc;
  for (int b = threadIdx.x; b < nblk; b += kSedThreads) acc.fold(part + kSedQ * b);
  sed_block_fold(acc, sh);
  if (threadIdx.x == 0)
    for (int q = 0; q < kSedQ; ++q) out[q] = acc.v[q];
}

}  // namespace rbs

using namespace rbs;

extern "C" {

int rbs_set_inflows(rbs_state *s, int32_t count, const int32_t *rows, const int32_t *cols,
                    const int32_t *offsets, const double *knot_t, const double *knot_q) {
  RBS_API_SHARED;
  if (!s || count < 0) return set_error("bad arguments"), RBS_E_ARG;
  if (count > 0 && (!rows || !cols || !offsets || !knot_t || !knot_q))
    return set_error("null inflow table"), RBS_E_ARG;
  std::vector<int64_t> seen;
  for (int k = 0; k < count; ++k) {
    if (rows[k] < 0 || rows[k] >= s->ny || cols[k] < 0 || cols[k] >= s->nx)
      return set_error("inflow cell (%d, %d) is outside the %dx%d grid", rows[k], cols[k], s->ny,
                       s->nx), RBS_E_ARG;
    const int64_t key = (int64_t)rows[k] * s->nx + cols[k];
    for (int64_t q : seen)
      if (q == key)  // inflow.py:127-139: concurrent same-cell sources would race
        return set_error("duplicate inflow cell (%d, %d); merge their hydrographs into one "
                         "[[inflow]] entry (concurrent same-cell sources would race)",
                         rows[k], cols[k]), RBS_E_ARG;
    seen.push_back(key);
    if (offsets[k + 1] < offsets[k]) return set_error("bad hydrograph offsets"), RBS_E_ARG;
    for (int j = offsets[k] + 1; j < offsets[k + 1]; ++j)
      if (knot_t[j] < knot_t[j - 1])
        return set_error("hydrograph times must be non-decreasing"), RBS_E_ARG;
  }
  Inflows &in = s->inflow;
  void *old[] = {in.d_rows, in.d_cols, in.d_off, in.d_kt, in.d_kq, in.d_comp, in.d_vol};
  RBS_HIP(hipStreamSynchronize(s->stream));
  for (void *p : old)
    if (p) (void)hipFree(p);
  in = Inflows{};
  if (count > 0) {
    const int nk = offsets[count];
    RBS_HIP(hipMalloc(&in.d_rows, sizeof(int) * count));
    RBS_HIP(hipMalloc(&in.d_cols, sizeof(int) * count));
    RBS_HIP(hipMalloc(&in.d_off, sizeof(int) * (count + 1)));
    RBS_HIP(hipMalloc(&in.d_kt, sizeof(double) * (nk > 0 ? nk : 1)));
    RBS_HIP(hipMalloc(&in.d_kq, sizeof(double) * (nk > 0 ? nk : 1)));
    RBS_HIP(hipMalloc(&in.d_comp, sizeof(float) * count));
    RBS_HIP(hipMalloc(&in.d_vol, sizeof(double) * (count + 1)));
    RBS_HIP(hipMemcpy(in.d_rows, rows, sizeof(int) * count, hipMemcpyHostToDevice));
    RBS_HIP(hipMemcpy(in.d_cols, cols, sizeof(int) * count, hipMemcpyHostToDevice));
    RBS_HIP(hipMemcpy(in.d_off, offsets, sizeof(int) * (count + 1), hipMemcpyHostToDevice));
    if (nk > 0) {
      RBS_HIP(hipMemcpy(in.d_kt, knot_t, sizeof(double) * nk, hipMemcpyHostToDevice));
      RBS_HIP(hipMemcpy(in.d_kq, knot_q, sizeof(double) * nk, hipMemcpyHostToDevice));
    }
    RBS_HIP(hipMemset(in.d_comp, 0, sizeof(float) * count));
    in.count = count;
    in.n_knots = nk;
  }
  for (int k = 0; k < 2; ++k)
    if (s->graph[k]) {
      (void)hipGraphExecDestroy(s->graph[k]);
      s->graph[k] = nullptr;
    }
  s->graph_key = -1;
  return RBS_OK;
}

int rbs_inflow_apply(rbs_state *s, double t, double dt, double *volume_out) {
  RBS_API_SHARED;
  if (!s || !volume_out) return set_error("null argument"), RBS_E_ARG;
  *volume_out = 0.0;
  if (s->inflow.count < 1) return RBS_OK;
  int rc = inflow_launch(s, false, t, dt);
  if (rc) return rc;
  RBS_HIP(hipMemcpyAsync(volume_out, s->inflow.d_vol + s->inflow.count, sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  s->speed_scheme = -1;  // h changed: the next compute_dt must reduce again
  return RBS_OK;
}

int rbs_sed_totals(rbs_state *s, double *totals) {
  RBS_API_SHARED;
  if (!s || !totals) return set_error("null argument"), RBS_E_ARG;
  if (!(s->armed & RBS_ARM_SED))
    return set_error("bed-change totals need a state armed with RBS_ARM_SED"), RBS_E_STATE;
  const int rpb = s->nx >= 4096 ? 1 : (4096 + s->nx - 1) / s->nx;
  const int nblk = (s->ny + rpb - 1) / rpb;
  if (!s->sed_part) {
    const size_t bytes = sizeof(double) * ((size_t)kSedQ * nblk + kSedQ);
    RBS_HIP(hipMalloc(&s->sed_part, bytes));
    s->bytes += bytes;
  }
  double *out = s->sed_part + (size_t)kSedQ * nblk;
  hipLaunchKernelGGL(sed_totals_kernel, dim3(nblk), dim3(kSedThreads), 0, s->stream, s->dz_cum,
                     s->dz_unapplied, s->pitch, s->ny, s->nx, rpb, s->sed_part);
  hipLaunchKernelGGL(sed_totals_final, dim3(1), dim3(kSedThreads), 0, s->stream, s->sed_part,
                     nblk, out);
  RBS_HIP(hipGetLastError());
  RBS_HIP(hipMemcpyAsync(totals, out, sizeof(double) * kSedQ, hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  return RBS_OK;
}

int rbs_exner_activate(rbs_state *s, double porosity) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (!(s->armed & RBS_ARM_SED))
    return set_error("morphology needs a state armed with RBS_ARM_SED"), RBS_E_STATE;
  if (!(porosity > 0.0 && porosity < 1.0))
    return set_error("porosity must be in (0, 1), got %g", porosity), RBS_E_ARG;
  int rc = exner_launch(s, porosity);
  if (rc) return rc;
  RBS_HIP(hipStreamSynchronize(s->stream));
  return RBS_OK;
}

}  // extern "C"
