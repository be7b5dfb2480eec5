// rbs_state.hip -- the librbs C-ABI: state lifecycle, field transfer, the scheme
// seam (compute_dt / step) and device-driven runs.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>

#include "rbs_device.hpp"

namespace rbs {

static thread_local char g_err[512] = "";

std::shared_mutex &api_mutex() {
  static std::shared_mutex m;
  return m;
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char *what) {
  set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return RBS_E_HIP;
}

// kernels of this unit ------------------------------------------------------------
__global__ void clock_reset_kernel(DevCtl *c, double t) {
  c->t = t;
  c->dt = 0.0;
  c->t_stop = 0.0;
  c->active = 0;
  c->sync_idx = 0;
  c->steps = 0;
  c->rain = 0.0f;
  c->rain_scale = 0.0f;
  c->rain_in = c->rain_in_c = 0.0;
  c->point_in = c->point_in_c = 0.0;
  c->infil_out = c->infil_out_c = 0.0;
  c->edge_out = c->edge_out_c = 0.0;
  c->release_out = c->release_out_c = 0.0;
}

__global__ void stage_config_kernel(DevCtl *c, int n0, int n1, int n2, int n3, const double *t0,
                                    const double *t1, const double *t2, const double *t3,
                                    const double *l0, const double *l1, const double *l2,
                                    const double *l3) {
  c->stage_n[0] = n0;
  c->stage_n[1] = n1;
  c->stage_n[2] = n2;
  c->stage_n[3] = n3;
  c->stage_t[0] = t0;
  c->stage_t[1] = t1;
  c->stage_t[2] = t2;
  c->stage_t[3] = t3;
  c->stage_lv[0] = l0;
  c->stage_lv[1] = l1;
  c->stage_lv[2] = l2;
  c->stage_lv[3] = l3;
}

__global__ void run_config_kernel(DevCtl *c, int scheme, int limit, int field, int n_sync,
                                  double alpha, double dt_max, double rain_m_s,
                                  double rain_duration, double field_sum, double t_stop,
                                  int64_t step_limit) {
  c->run_scheme = scheme;
  c->run_limit = limit;
  c->run_field = field;
  c->n_sync = n_sync;
  c->alpha = alpha;
  c->dt_max = dt_max;
  c->rain_m_s = rain_m_s;
  c->rain_duration = rain_duration;
  c->rain_field_sum = field_sum;
  c->t_stop = t_stop;
  c->sync_idx = 0;
  c->step_limit = step_limit;
  c->active = 0;
}

// field table ---------------------------------------------------------------------
struct FieldView {
  void *dev;
  int rows, cols, elem;
};

static int field_view(rbs_state *s, int f, FieldView *v) {
  const int ny = s->ny, nx = s->nx;
  const int cur = s->cur;
  auto cell = [&](void *p, int elem = 4) { *v = {p, ny, nx, elem}; };
  switch (f) {
    case RBS_F_H: cell(s->h[cur]); break;
    case RBS_F_Z: cell(s->z); break;
    case RBS_F_N: cell(s->n); break;
    case RBS_F_QX: *v = {s->qx[cur], ny, nx + 1, 4}; break;
    case RBS_F_QY: *v = {s->qy[cur], ny + 1, nx, 4}; break;
    case RBS_F_HU: cell(s->hu); break;
    case RBS_F_HV: cell(s->hv); break;
    case RBS_F_RAIN: cell(s->rain); break;
    case RBS_F_INFIL: cell(s->infil); break;
    case RBS_F_H_COMP: cell(s->h_comp); break;
    case RBS_F_D50: cell(s->d50); break;
    case RBS_F_QS_INT_X: *v = {s->qs_int_x, ny, nx + 1, 4}; break;
    case RBS_F_QS_COMP_X: *v = {s->qs_comp_x, ny, nx + 1, 4}; break;
    case RBS_F_QS_INT_Y: *v = {s->qs_int_y, ny + 1, nx, 4}; break;
    case RBS_F_QS_COMP_Y: *v = {s->qs_comp_y, ny + 1, nx, 4}; break;
    case RBS_F_Z0: cell(s->z0); break;
    case RBS_F_DZ_CUM: cell(s->dz_cum, 8); break;
    case RBS_F_DZ_UNAPPLIED: cell(s->dz_unapplied, 8); break;
    case RBS_F_DZ_LO: cell(s->dz_lo); break;
    case RBS_F_DZ_HI: cell(s->dz_hi); break;
    case RBS_F_CH_W: cell(s->ch_w); break;
    case RBS_F_CH_D: cell(s->ch_d); break;
    case RBS_F_CH_N: cell(s->ch_n); break;
    case RBS_F_QX_CH: *v = {s->qx_ch[cur], ny, nx + 1, 4}; break;
    case RBS_F_QX_FP: *v = {s->qx_fp[cur], ny, nx + 1, 4}; break;
    case RBS_F_QY_CH: *v = {s->qy_ch[cur], ny + 1, nx, 4}; break;
    case RBS_F_QY_FP: *v = {s->qy_fp[cur], ny + 1, nx, 4}; break;
    default:
      set_error("unknown field id %d", f);
      return RBS_E_ARG;
  }
  if (v->dev == nullptr) {
    set_error("field %d is not armed on this state", f);
    return RBS_E_STATE;
  }
  return RBS_OK;
}

// Pitched allocation of `rows` rows, zero-filled.  f64 fields use a pitch of the
// same element count (so byte pitch doubles).
static int alloc_field(rbs_state *s, void **p, int rows, int elem) {
  const size_t bytes = (size_t)rows * (size_t)s->pitch * (size_t)elem;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return RBS_E_NOMEM;
  }
  RBS_HIP(hipMemsetAsync(*p, 0, bytes, s->stream));
  s->bytes += bytes;
  return RBS_OK;
}

#define ALLOC(ptr, rows, elem) \
  do { int rc_ = alloc_field(s, (void **)&(ptr), (rows), (elem)); if (rc_) return rc_; } while (0)

// Bytes of the pitched fields a state of ny rows armed with `arm` holds, row counts
// and element sizes as rbs_state_create and arm_fields allocate them.
static uint64_t field_bytes(int ny, int64_t pitch, uint32_t arm) {
  struct Spec {
    uint32_t flag;  // 0: every state
    int count, extra_rows, elem;
  };
  static const Spec specs[] = {
      {0, 4, 0, 4},              // h, qx (ping-pong)
      {0, 2, 1, 4},              // qy (ping-pong)
      {0, 2, 0, 4},              // z, n
      {RBS_ARM_COMP, 1, 0, 4},   // h_comp
      {RBS_ARM_RAIN, 1, 0, 4},   // rain
      {RBS_ARM_INFIL, 1, 0, 4},  // infil
      {RBS_ARM_HLLC, 5, 0, 4},   // hu, hv, h1, hu1, hv1
      {RBS_ARM_SED, 6, 0, 4},    // d50, qs_int_x, qs_comp_x, z0, dz_lo, dz_hi
      {RBS_ARM_SED, 2, 1, 4},    // qs_int_y, qs_comp_y
      {RBS_ARM_SED, 2, 0, 8},    // dz_cum, dz_unapplied
      {RBS_ARM_CHAN, 7, 0, 4},   // ch_w, ch_d, ch_n, qx_ch and qx_fp (ping-pong)
      {RBS_ARM_CHAN, 4, 1, 4},   // qy_ch, qy_fp (ping-pong)
  };
  uint64_t total = 0;
  for (const Spec &sp : specs)
    if (sp.flag == 0 || (arm & sp.flag))
      total += (uint64_t)sp.count * (uint64_t)(ny + sp.extra_rows) * (uint64_t)pitch * sp.elem;
  return total;
}

// Block partials (and dry-tile words) for the largest tiling any variant uses.
static int max_blocks(rbs_state *s) {
  const int li0 = s->li_variant, h0 = s->hllc_variant, c0 = s->chan_variant;
  int nblk = reduce_partials_needed(s);
  for (int v = 0; v < li_variant_count(); ++v) {
    s->li_variant = v;
    nblk = std::max(nblk, li_num_blocks(s));
  }
  for (int v = 0; v < hllc_variant_count(); ++v) {
    s->hllc_variant = v;
    nblk = std::max(nblk, hllc_num_blocks(s));
  }
  for (int v = 0; v < chan_variant_count(); ++v) {
    s->chan_variant = v;
    nblk = std::max(nblk, chan_num_blocks(s));
  }
  s->li_variant = li0;
  s->hllc_variant = h0;
  s->chan_variant = c0;
  return nblk;
}

static void invalidate_graph(rbs_state *s) {
  for (int k = 0; k < 2; ++k)
    if (s->graph[k]) {
      (void)hipGraphExecDestroy(s->graph[k]);
      s->graph[k] = nullptr;
    }
  s->graph_key = -1;
}

static int fill_f32(rbs_state *s, float *p, int rows, float v);

static int arm_fields(rbs_state *s, uint32_t arm) {
  const int ny = s->ny;
  if ((arm & RBS_ARM_SED) && (s->armed & RBS_ARM_SED)) {
    set_error("state is already armed for morphology; re-arming would recapture z0 from a "
              "moved bed and discard the cumulative change that moved it");
    return RBS_E_STATE;
  }
  if ((arm & RBS_ARM_COMP) && !s->h_comp) ALLOC(s->h_comp, ny, 4);
  if ((arm & RBS_ARM_RAIN) && !s->rain) ALLOC(s->rain, ny, 4);
  if ((arm & RBS_ARM_INFIL) && !s->infil) ALLOC(s->infil, ny, 4);
  if ((arm & RBS_ARM_HLLC) && !s->hu) {
    ALLOC(s->hu, ny, 4);
    ALLOC(s->hv, ny, 4);
    ALLOC(s->h1, ny, 4);
    ALLOC(s->hu1, ny, 4);
    ALLOC(s->hv1, ny, 4);
  }
  if (arm & RBS_ARM_SED) {
    ALLOC(s->d50, ny, 4);
    ALLOC(s->qs_int_x, ny, 4);
    ALLOC(s->qs_comp_x, ny, 4);
    ALLOC(s->qs_int_y, ny + 1, 4);
    ALLOC(s->qs_comp_y, ny + 1, 4);
    ALLOC(s->z0, ny, 4);
    ALLOC(s->dz_lo, ny, 4);
    ALLOC(s->dz_hi, ny, 4);
    ALLOC(s->dz_cum, ny, 8);
    ALLOC(s->dz_unapplied, ny, 8);
    // z0 = the bed as it is now (sediment.py:697); bounds default to +-inf
    // (morphology.py:95-97).
    RBS_HIP(hipMemcpyAsync(s->z0, s->z, (size_t)ny * s->pitch * 4, hipMemcpyDeviceToDevice,
                           s->stream));
    int rc = fill_f32(s, s->dz_lo, ny, -INFINITY);
    if (rc) return rc;
    rc = fill_f32(s, s->dz_hi, ny, INFINITY);
    if (rc) return rc;
  }
  if ((arm & RBS_ARM_CHAN) && !s->ch_w) {
    ALLOC(s->ch_w, ny, 4);
    ALLOC(s->ch_d, ny, 4);
    ALLOC(s->ch_n, ny, 4);
    for (int k = 0; k < 2; ++k) {
      ALLOC(s->qx_ch[k], ny, 4);
      ALLOC(s->qx_fp[k], ny, 4);
      ALLOC(s->qy_ch[k], ny + 1, 4);
      ALLOC(s->qy_fp[k], ny + 1, 4);
    }
    // the channel roughness defaults to the floodplain's (channels.py:300-307)
    RBS_HIP(hipMemcpyAsync(s->ch_n, s->n, (size_t)ny * s->pitch * 4, hipMemcpyDeviceToDevice,
                           s->stream));
    s->speed_scheme = -1;  // the CFL reduction now runs over the water column
  }
  s->armed |= arm;
  invalidate_graph(s);
  RBS_HIP(hipStreamSynchronize(s->stream));
  return RBS_OK;
}

__global__ void fill_kernel(float *p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static int fill_f32(rbs_state *s, float *p, int rows, float v) {
  const size_t n = (size_t)rows * s->pitch;
  hipLaunchKernelGGL(fill_kernel, dim3(1024), dim3(256), 0, s->stream, p, n, v);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

static int host_speed_needed(rbs_state *s, int scheme) {
  return s->speed_scheme != scheme;
}

}  // namespace rbs

using namespace rbs;

extern "C" {

const char *rbs_last_error(void) { return g_err; }

const char *rbs_version(void) { return "librbs 0.1 (gfx950)"; }

int rbs_device_count(int32_t *count) {
  if (!count) return set_error("null count"), RBS_E_ARG;
  int n = 0;
  RBS_HIP(hipGetDeviceCount(&n));
  *count = n;
  return RBS_OK;
}

static int state_destroy(rbs_state *s);

int rbs_state_create(int32_t ny, int32_t nx, double dx, uint32_t arm, rbs_state **out) {
  RBS_API_SHARED;
  if (!out) return set_error("null out"), RBS_E_ARG;
  *out = nullptr;
  if (ny < 1 || nx < 1 || !(dx > 0.0) || !std::isfinite(dx)) {
    set_error("bad grid ny=%d nx=%d dx=%g", ny, nx, dx);
    return RBS_E_ARG;
  }
  if ((int64_t)ny * (int64_t)nx > (int64_t)1 << 34) {
    set_error("grid %dx%d too large", ny, nx);
    return RBS_E_RANGE;
  }
  rbs_state *s = new rbs_state();
  s->ny = ny;
  s->nx = nx;
  s->dx = dx;
  s->pitch = ((int64_t)nx + 1 + 63) / 64 * 64;
  int rc = RBS_OK;
  auto fail = [&](int code) {
    state_destroy(s);
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return fail(hip_fail(hipErrorNoDevice, "hipGetDevice"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipStreamCreate"));
#define TRY(x) do { rc = (x); if (rc) return fail(rc); } while (0)
  for (int k = 0; k < 2; ++k) {
    TRY(alloc_field(s, (void **)&s->h[k], ny, 4));
    TRY(alloc_field(s, (void **)&s->qx[k], ny, 4));
    TRY(alloc_field(s, (void **)&s->qy[k], ny + 1, 4));
  }
  TRY(alloc_field(s, (void **)&s->z, ny, 4));
  TRY(alloc_field(s, (void **)&s->n, ny, 4));
  if (hipMalloc(&s->ctl, sizeof(DevCtl)) != hipSuccess) return fail(hip_fail(hipErrorOutOfMemory, "hipMalloc(ctl)"));
  if (hipHostMalloc(&s->ctl_host, sizeof(DevCtl)) != hipSuccess)
    return fail(hip_fail(hipErrorOutOfMemory, "hipHostMalloc(ctl)"));
  if (hipMemsetAsync(s->ctl, 0, sizeof(DevCtl), s->stream) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipMemset(ctl)"));
  s->li_variant = li_initial_variant(s->nx);
  s->hllc_variant = hllc_default_variant();
  s->chan_variant = chan_default_variant();
  TRY(ensure_partials(s, max_blocks(s)));
  if (arm) TRY(arm_fields(s, arm));
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(hip_fail(hipErrorUnknown, "sync"));
#undef TRY
  *out = s;
  return RBS_OK;
}

int rbs_state_destroy(rbs_state *s) {
  RBS_API_SHARED;
  return state_destroy(s);
}

static int state_destroy(rbs_state *s) {
  if (!s) return RBS_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  invalidate_graph(s);
  void *ptrs[] = {s->h[0], s->h[1], s->qx[0], s->qx[1], s->qy[0], s->qy[1], s->z, s->n,
                  s->h_comp, s->rain, s->infil, s->hu, s->hv, s->h1, s->hu1, s->hv1,
                  s->d50, s->qs_int_x, s->qs_comp_x, s->qs_int_y, s->qs_comp_y, s->z0,
                  s->dz_lo, s->dz_hi, s->dz_cum, s->dz_unapplied, s->sed_part, s->part.speed_max,
                  s->part.vol, s->part.infil, s->part.edge, s->quiet, s->hquiet, s->fold.speed_max, s->fold.vol,
                  s->fold.infil, s->fold.edge, s->ctl, s->d_sync,
                  s->inflow.d_rows, s->inflow.d_cols, s->inflow.d_off, s->inflow.d_kt,
                  s->inflow.d_kq, s->inflow.d_comp, s->inflow.d_vol, s->ch_w, s->ch_d,
                  s->ch_n, s->qx_ch[0], s->qx_ch[1], s->qx_fp[0], s->qx_fp[1], s->qy_ch[0],
                  s->qy_ch[1], s->qy_fp[0], s->qy_fp[1], s->xfer, s->resv_rec};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &c : s->stage) {
    delete[] c.t;
    delete[] c.level;
    if (c.d_t) (void)hipFree(c.d_t);
    if (c.d_level) (void)hipFree(c.d_level);
  }
  if (s->ctl_host) (void)hipHostFree(s->ctl_host);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return RBS_OK;
}

int rbs_arm(rbs_state *s, uint32_t arm) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (arm & ~0x3Fu) return set_error("unknown arm flags 0x%x", arm), RBS_E_ARG;
  return arm_fields(s, arm);
}

int rbs_armed(const rbs_state *s, uint32_t *arm) {
  if (!s || !arm) return set_error("null argument"), RBS_E_ARG;
  *arm = s->armed;
  return RBS_OK;
}

int rbs_state_footprint(int32_t ny, int32_t nx, uint32_t arm, uint64_t *fields,
                        uint64_t *total) {
  if (!fields || !total) return set_error("null argument"), RBS_E_ARG;
  if (ny < 1 || nx < 1) return set_error("bad grid ny=%d nx=%d", ny, nx), RBS_E_ARG;
  rbs_state probe;  // geometry only: nothing is allocated
  probe.ny = ny;
  probe.nx = nx;
  probe.pitch = ((int64_t)nx + 1 + 63) / 64 * 64;
  probe.li_variant = li_initial_variant(nx);
  probe.hllc_variant = hllc_default_variant();
  probe.chan_variant = chan_default_variant();
  const uint64_t nblk = (uint64_t)max_blocks(&probe);
  *fields = field_bytes(ny, probe.pitch, arm);
  // partials: a float, three doubles and two dry-block words per block; then the
  // control block, point-source tables, reduction scratch and staging the state adds
  *total = *fields + nblk * (sizeof(float) + 3 * sizeof(double) + 2 * sizeof(uint32_t)) +
           (uint64_t(1) << 20);
  return RBS_OK;
}

int rbs_state_bytes(const rbs_state *s, uint64_t *bytes) {
  if (!s || !bytes) return set_error("null argument"), RBS_E_ARG;
  *bytes = s->bytes;
  return RBS_OK;
}

// Host<->device field copies move in blocks of rows of at most kCopyChunk bytes,
// each under the API lock on its own, so a graph capture on another thread never
// waits for more than one block.  A pitched 2-D copy moves row by row (one DMA per
// row); instead each block is packed to (or unpacked from) a dense device staging
// block by a kernel and crosses the bus as one 1-D copy.
constexpr size_t kCopyChunk = size_t(4) << 20;

// rows x row_w 4-byte words between a pitched field and a dense block.
__global__ void pack_rows_kernel(const uint32_t *__restrict__ src, int64_t pitch_w,
                                 uint32_t *__restrict__ dst, int64_t row_w, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / row_w, k = i - r * row_w;
    dst[i] = src[r * pitch_w + k];
  }
}
__global__ void unpack_rows_kernel(const uint32_t *__restrict__ src, int64_t pitch_w,
                                   uint32_t *__restrict__ dst, int64_t row_w, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / row_w, k = i - r * row_w;
    dst[r * pitch_w + k] = src[i];
  }
}

static int copy_rows(rbs_state *s, const FieldView &v, void *host, bool to_device) {
  const size_t row_bytes = (size_t)v.cols * v.elem;
  const size_t dpitch = (size_t)s->pitch * v.elem;
  const int rows_per = (int)std::max<size_t>(1, kCopyChunk / row_bytes);
  const size_t need = std::max(kCopyChunk, row_bytes);
  for (int r0 = 0; r0 < v.rows; r0 += rows_per) {
    const int nr = std::min(rows_per, v.rows - r0);
    char *h = (char *)host + (size_t)r0 * row_bytes;
    char *d = (char *)v.dev + (size_t)r0 * dpitch;
    RBS_API_SHARED;
    if (s->xfer_cap < need) {
      if (s->xfer) RBS_HIP(hipFree(s->xfer));
      s->xfer = nullptr;
      s->xfer_cap = 0;
      RBS_HIP(hipMalloc(&s->xfer, need));
      s->xfer_cap = need;
    }
    const int64_t row_w = (int64_t)(row_bytes / 4), pitch_w = (int64_t)(dpitch / 4);
    const int64_t n = row_w * nr;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (to_device) {
      RBS_HIP(hipMemcpyAsync(s->xfer, h, (size_t)nr * row_bytes, hipMemcpyHostToDevice,
                             s->stream));
      hipLaunchKernelGGL(unpack_rows_kernel, dim3(grid), dim3(256), 0, s->stream,
                         (const uint32_t *)s->xfer, pitch_w, (uint32_t *)d, row_w, n);
      RBS_HIP(hipGetLastError());
    } else {
      hipLaunchKernelGGL(pack_rows_kernel, dim3(grid), dim3(256), 0, s->stream,
                         (const uint32_t *)d, pitch_w, (uint32_t *)s->xfer, row_w, n);
      RBS_HIP(hipGetLastError());
      RBS_HIP(hipMemcpyAsync(h, s->xfer, (size_t)nr * row_bytes, hipMemcpyDeviceToHost,
                             s->stream));
    }
    RBS_HIP(hipStreamSynchronize(s->stream));
  }
  return RBS_OK;
}

int rbs_upload(rbs_state *s, int32_t field, const void *host, size_t bytes) {
  if (!s || !host) return set_error("null argument"), RBS_E_ARG;
  FieldView v;
  int rc = field_view(s, field, &v);
  if (rc) return rc;
  const size_t want = (size_t)v.rows * v.cols * v.elem;
  if (bytes != want) {
    set_error("field %d expects %zu bytes, got %zu", field, want, bytes);
    return RBS_E_ARG;
  }
  rc = copy_rows(s, v, const_cast<void *>(host), true);
  if (rc) return rc;
  if (field == RBS_F_H || field == RBS_F_QX || field == RBS_F_QY || field == RBS_F_H_COMP ||
      field == RBS_F_QS_COMP_X || field == RBS_F_QS_COMP_Y || field == RBS_F_HU ||
      field == RBS_F_HV) {
    rc = quiet_reset(s, s->stream);
    if (rc) return rc;
  }
  if (field == RBS_F_H || field == RBS_F_HU || field == RBS_F_HV || field == RBS_F_CH_W ||
      field == RBS_F_CH_D)
    s->speed_scheme = -1;
  if (field == RBS_F_RAIN) {
    // run.py:276: the ledger's field-rain inflow uses sum f64(field)
    const float *f = (const float *)host;
    double acc = 0.0;
    for (size_t k = 0; k < (size_t)v.rows * v.cols; ++k) acc += (double)f[k];
    s->rain_field_sum = acc;
  }
  return RBS_OK;
}

static int download_field(rbs_state *s, int32_t field, void *host, size_t bytes);

int rbs_download(rbs_state *s, int32_t field, void *host, size_t bytes) {
  if (!s || !host) return set_error("null argument"), RBS_E_ARG;
  if (field == RBS_F_ETA) {
    // eta = h + z (local_inertial.py:61-69), or the storage curve of a channelled
    // state (channels.py:59-72), formed on the host in float32 from the fields.
    const size_t want = (size_t)s->ny * s->nx * 4;
    if (bytes != want) return set_error("eta expects %zu bytes", want), RBS_E_ARG;
    std::vector<float> z((size_t)s->ny * s->nx);
    int rc = download_field(s, RBS_F_H, host, bytes);
    if (rc) return rc;
    rc = download_field(s, RBS_F_Z, z.data(), bytes);
    if (rc) return rc;
    float *e = (float *)host;
    if (!s->ch_w) {
      for (size_t k = 0; k < z.size(); ++k) e[k] = e[k] + z[k];
      return RBS_OK;
    }
    std::vector<float> w(z.size()), d(z.size());
    rc = download_field(s, RBS_F_CH_W, w.data(), bytes);
    if (rc) return rc;
    rc = download_field(s, RBS_F_CH_D, d.data(), bytes);
    if (rc) return rc;
    const float dx = (float)s->dx;
    for (size_t k = 0; k < z.size(); ++k) e[k] = eta_subgrid(e[k], z[k], w[k], d[k], dx);
    return RBS_OK;
  }
  return download_field(s, field, host, bytes);
}

static int download_field(rbs_state *s, int32_t field, void *host, size_t bytes) {
  FieldView v;
  int rc = field_view(s, field, &v);
  if (rc) return rc;
  const size_t want = (size_t)v.rows * v.cols * v.elem;
  if (bytes != want) {
    set_error("field %d expects %zu bytes, got %zu", field, want, bytes);
    return RBS_E_ARG;
  }
  return copy_rows(s, v, host, false);
}

int rbs_set_boundaries(rbs_state *s, const int32_t edge_types[4]) {
  RBS_API_SHARED;
  if (!s || !edge_types) return set_error("null argument"), RBS_E_ARG;
  for (int e = 0; e < 4; ++e)
    if (edge_types[e] < RBS_BC_CLOSED || edge_types[e] > RBS_BC_FIXED_STAGE)
      return set_error("bad edge type %d", edge_types[e]), RBS_E_ARG;
  for (int e = 0; e < 4; ++e) s->bc[e] = edge_types[e];
  invalidate_graph(s);
  return RBS_OK;
}

int rbs_set_stage_curve(rbs_state *s, int32_t edge, int32_t n, const double *t,
                        const double *level) {
  RBS_API_SHARED;
  if (!s || edge < 0 || edge > 3 || n < 1 || !t || !level)
    return set_error("bad stage curve arguments"), RBS_E_ARG;
  for (int k = 1; k < n; ++k)
    if (t[k] < t[k - 1]) return set_error("stage curve times must be non-decreasing"), RBS_E_ARG;
  StageCurve &c = s->stage[edge];
  delete[] c.t;
  delete[] c.level;
  if (c.d_t) (void)hipFree(c.d_t);
  if (c.d_level) (void)hipFree(c.d_level);
  c.n = n;
  c.t = new double[n];
  c.level = new double[n];
  std::memcpy(c.t, t, sizeof(double) * n);
  std::memcpy(c.level, level, sizeof(double) * n);
  RBS_HIP(hipMalloc(&c.d_t, sizeof(double) * n));
  RBS_HIP(hipMalloc(&c.d_level, sizeof(double) * n));
  RBS_HIP(hipMemcpy(c.d_t, t, sizeof(double) * n, hipMemcpyHostToDevice));
  RBS_HIP(hipMemcpy(c.d_level, level, sizeof(double) * n, hipMemcpyHostToDevice));
  invalidate_graph(s);
  return RBS_OK;
}

int rbs_compute_dt(rbs_state *s, int32_t scheme, double alpha, double dt_max, double *dt_out) {
  RBS_API_SHARED;
  if (!s || !dt_out) return set_error("null argument"), RBS_E_ARG;
  if (scheme != RBS_SCHEME_LI && scheme != RBS_SCHEME_HLLC)
    return set_error("unknown scheme %d", scheme), RBS_E_ARG;
  if (scheme == RBS_SCHEME_HLLC && !(s->armed & RBS_ARM_HLLC)) {
    int rc = arm_fields(s, RBS_ARM_HLLC);  // hllc.compute_dt arms (hllc.py:978)
    if (rc) return rc;
  }
  if (host_speed_needed(s, scheme)) {
    int rc = reduce_speed_launch(s, scheme, FIN_PRIME);
    if (rc) return rc;
    s->speed_scheme = scheme;
  }
  float sm = 0.0f;
  RBS_HIP(hipMemcpyAsync(&sm, (char *)s->ctl + offsetof(DevCtl, speed_max), sizeof(float),
                         hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  const double m = (double)sm;
  if (scheme == RBS_SCHEME_LI) {
    // local_inertial.py:540-544
    if (m <= kHDry64) {
      *dt_out = dt_max;
      return RBS_OK;
    }
    const double dt = alpha * s->dx / std::sqrt(kGravity * m);
    *dt_out = dt < dt_max ? dt : dt_max;
  } else {
    // hllc.py:987-990
    if (m <= 0.0) {
      *dt_out = dt_max;
      return RBS_OK;
    }
    const double dt = alpha * s->dx / m;
    *dt_out = dt < dt_max ? dt : dt_max;
  }
  return RBS_OK;
}

int rbs_li_step(rbs_state *s, double dt, double rain, int32_t limit, double rain_scale,
                double t) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (!(dt > 0.0) || !std::isfinite(dt)) return set_error("non-positive timestep %g", dt), RBS_E_RANGE;
  int rc = set_scalars_launch(s, t, dt, (float)rain, (float)rain_scale);
  if (rc) return rc;
  rc = li_launch_step(s, false, dt, rain, limit, rain_scale);
  if (rc) return rc;
  rc = finalize_launch(s, FIN_EXPLICIT, RBS_SCHEME_LI);
  if (rc) return rc;
  s->speed_scheme = RBS_SCHEME_LI;
  return RBS_OK;
}

// boundaries.py:49-70 stage_at (host, float64)
static double stage_at_host(const StageCurve &c, double x) {
  const int n = c.n;
  if (x <= c.t[0]) return c.level[0];
  if (x >= c.t[n - 1]) return c.level[n - 1];
  for (int k = 0; k + 1 < n; ++k) {
    const double t0 = c.t[k], t1 = c.t[k + 1];
    if (t0 <= x && x <= t1) {
      if (t1 == t0) return c.level[k + 1];
      return c.level[k] + (c.level[k + 1] - c.level[k]) * (x - t0) / (t1 - t0);
    }
  }
  return c.level[n - 1];
}

int rbs_hllc_step(rbs_state *s, double dt, double rain, double rain_scale, double t) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (!(dt > 0.0) || !std::isfinite(dt)) return set_error("non-positive timestep %g", dt), RBS_E_RANGE;
  if (!(s->armed & RBS_ARM_HLLC)) {
    int rc = arm_fields(s, RBS_ARM_HLLC);  // hllc.step arms (hllc.py:1304)
    if (rc) return rc;
  }
  float stage[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int e = 0; e < 4; ++e)
    if (s->bc[e] == RBS_BC_FIXED_STAGE) {
      if (s->stage[e].n < 1) return set_error("fixed_stage edge %d without a stage curve", e), RBS_E_STATE;
      stage[e] = (float)stage_at_host(s->stage[e], t + 0.5 * dt);  // hllc.py:1308-1312
    }
  int rc = set_scalars_launch(s, t, dt, (float)rain, (float)rain_scale, stage);
  if (rc) return rc;
  rc = hllc_launch_step(s, false, dt, rain, rain_scale, t);
  if (rc) return rc;
  rc = finalize_launch(s, FIN_EXPLICIT, RBS_SCHEME_HLLC);
  if (rc) return rc;
  s->speed_scheme = RBS_SCHEME_HLLC;
  return RBS_OK;
}

int rbs_ledger_read(rbs_state *s, rbs_ledger *out) {
  RBS_API_SHARED;
  if (!s || !out) return set_error("null argument"), RBS_E_ARG;
  if (s->speed_scheme == -1) {  // volume of an uploaded state not yet reduced
    int rc = reduce_speed_launch(s, (s->armed & RBS_ARM_HLLC) ? RBS_SCHEME_HLLC : RBS_SCHEME_LI, FIN_PRIME);
    if (rc) return rc;
    s->speed_scheme = (s->armed & RBS_ARM_HLLC) ? RBS_SCHEME_HLLC : RBS_SCHEME_LI;
  }
  RBS_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  const DevCtl &c = *s->ctl_host;
  out->volume = c.volume;
  out->rain_in = c.rain_in;
  out->rain_in_c = c.rain_in_c;
  out->point_in = c.point_in;
  out->point_in_c = c.point_in_c;
  out->infil_out = c.infil_out;
  out->infil_out_c = c.infil_out_c;
  out->edge_out = c.edge_out;
  out->edge_out_c = c.edge_out_c;
  out->release_out = c.release_out;
  out->release_out_c = c.release_out_c;
  out->t = c.t;
  out->dt = c.dt;
  out->steps = c.steps;
  out->speed_max = c.speed_max;
  out->active = c.active;
  return RBS_OK;
}

int rbs_clock_reset(rbs_state *s, double t) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  hipLaunchKernelGGL(clock_reset_kernel, dim3(1), dim3(1), 0, s->stream, s->ctl, t);
  RBS_HIP(hipGetLastError());
  RBS_HIP(hipStreamSynchronize(s->stream));
  return RBS_OK;
}

int rbs_sync(rbs_state *s) {
  if (!s) return set_error("null state"), RBS_E_ARG;
  RBS_HIP(hipStreamSynchronize(s->stream));
  return RBS_OK;
}

int rbs_get_li_variant(const rbs_state *s, int32_t *variant) {
  if (!s || !variant) return set_error("null argument"), RBS_E_ARG;
  *variant = s->li_variant;
  return RBS_OK;
}

int rbs_set_li_variant(rbs_state *s, int32_t variant) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (variant < 0 || variant >= li_variant_count())
    return set_error("unknown LI variant %d", variant), RBS_E_ARG;
  if (variant != s->li_variant) {
    const int rc = quiet_reset(s, s->stream);  // another tiling
    if (rc) return rc;
  }
  s->li_variant = variant;
  invalidate_graph(s);
  return RBS_OK;
}

int rbs_time_li_kernel(rbs_state *s, int32_t reps, double dt, double rain, double rain_scale,
                       double *avg_ms) {
  RBS_API_SHARED;
  if (!s || !avg_ms || reps < 1) return set_error("bad arguments"), RBS_E_ARG;
  if (!(dt > 0.0)) return set_error("non-positive timestep %g", dt), RBS_E_RANGE;
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  int rc = RBS_OK;
  for (auto &e : ev) RBS_HIP(hipEventCreate(&e));
  for (int k = 0; k < reps && rc == RBS_OK; ++k) {
    rc = set_scalars_launch(s, 0.0, dt, (float)rain, (float)rain_scale);
    if (rc) break;
    // the events are stamped by the kernel's dispatch itself (its start and end)
    rc = li_launch_step(s, false, dt, rain, 1, rain_scale, ev[2 * k], ev[2 * k + 1]);
    if (rc) break;
    rc = finalize_launch(s, FIN_EXPLICIT, RBS_SCHEME_LI);
  }
  double total = 0.0;
  if (rc == RBS_OK && hipStreamSynchronize(s->stream) == hipSuccess) {
    for (int k = 0; k < reps; ++k) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) { rc = hip_fail(hipErrorUnknown, "hipEventElapsedTime"); break; }
      total += ms;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->speed_scheme = RBS_SCHEME_LI;
  if (rc == RBS_OK) *avg_ms = total / reps;
  return rc;
}

int rbs_time_hllc_step(rbs_state *s, int32_t reps, double dt, double *avg_ms) {
  RBS_API_SHARED;
  if (!s || !avg_ms || reps < 1) return set_error("bad arguments"), RBS_E_ARG;
  if (!(dt > 0.0)) return set_error("non-positive timestep %g", dt), RBS_E_RANGE;
  if (!(s->armed & RBS_ARM_HLLC)) return set_error("state is not armed for HLLC"), RBS_E_STATE;
  // stage levels as they stand (the ghost values only; timing does not need the curve)
  float stage[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int e = 0; e < 4; ++e)
    if (s->bc[e] == RBS_BC_FIXED_STAGE && s->stage[e].n > 0)
      stage[e] = (float)stage_at_host(s->stage[e], 0.0);
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  int rc = RBS_OK;
  for (auto &e : ev) RBS_HIP(hipEventCreate(&e));
  for (int k = 0; k < reps && rc == RBS_OK; ++k) {
    rc = set_scalars_launch(s, 0.0, dt, 0.0f, 0.0f, stage);
    if (rc) break;
    if (hipEventRecord(ev[2 * k], s->stream) != hipSuccess) { rc = hip_fail(hipErrorUnknown, "hipEventRecord"); break; }
    rc = hllc_launch_step(s, false, dt, 0.0, 0.0, 0.0);  // both SSP-RK2 stage kernels
    if (rc) break;
    if (hipEventRecord(ev[2 * k + 1], s->stream) != hipSuccess) { rc = hip_fail(hipErrorUnknown, "hipEventRecord"); break; }
    rc = finalize_launch(s, FIN_EXPLICIT, RBS_SCHEME_HLLC);
  }
  double total = 0.0;
  if (rc == RBS_OK && hipStreamSynchronize(s->stream) == hipSuccess) {
    for (int k = 0; k < reps; ++k) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) { rc = hip_fail(hipErrorUnknown, "hipEventElapsedTime"); break; }
      total += ms;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  s->speed_scheme = RBS_SCHEME_HLLC;
  if (rc == RBS_OK) *avg_ms = total / reps;
  return rc;
}

int rbs_li_variant_count(int32_t *count) {
  if (!count) return set_error("null count"), RBS_E_ARG;
  *count = li_variant_count();
  return RBS_OK;
}

int rbs_set_hllc_variant(rbs_state *s, int32_t variant) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (variant < 0 || variant >= hllc_variant_count())
    return set_error("HLLC variant %d is not in [0, %d)", variant, hllc_variant_count()), RBS_E_ARG;
  if (variant != s->hllc_variant) {
    const int rc = quiet_reset(s, s->stream);  // another tiling
    if (rc) return rc;
  }
  s->hllc_variant = variant;
  invalidate_graph(s);  // a captured device-driven run launches the old tiles
  return RBS_OK;
}

int rbs_hllc_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *is_default) {
  if (!bx || !by || !is_default) return set_error("null argument"), RBS_E_ARG;
  if (variant < 0 || variant >= hllc_variant_count())
    return set_error("HLLC variant %d is not in [0, %d)", variant, hllc_variant_count()), RBS_E_ARG;
  int x, y;
  hllc_variant_shape(variant, &x, &y);
  *bx = x;
  *by = y;
  *is_default = variant == hllc_default_variant() ? 1 : 0;
  return RBS_OK;
}

int rbs_hllc_variant_count(int32_t *count) {
  if (!count) return set_error("null count"), RBS_E_ARG;
  *count = hllc_variant_count();
  return RBS_OK;
}

int rbs_set_chan_variant(rbs_state *s, int32_t variant) {
  RBS_API_SHARED;
  if (!s) return set_error("null state"), RBS_E_ARG;
  if (variant < 0 || variant >= chan_variant_count())
    return set_error("channel variant %d is not in [0, %d)", variant, chan_variant_count()), RBS_E_ARG;
  s->chan_variant = variant;
  invalidate_graph(s);
  return RBS_OK;
}

int rbs_chan_variant_count(int32_t *count) {
  if (!count) return set_error("null count"), RBS_E_ARG;
  *count = chan_variant_count();
  return RBS_OK;
}

int rbs_chan_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *is_default) {
  if (!bx || !by || !is_default) return set_error("null argument"), RBS_E_ARG;
  if (variant < 0 || variant >= chan_variant_count())
    return set_error("channel variant %d is not in [0, %d)", variant, chan_variant_count()), RBS_E_ARG;
  int x, y;
  chan_variant_shape(variant, &x, &y);
  *bx = x;
  *by = y;
  *is_default = variant == chan_default_variant() ? 1 : 0;
  return RBS_OK;
}

int rbs_li_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *family,
                        int32_t *is_default) {
  if (!bx || !by || !family || !is_default) return set_error("null argument"), RBS_E_ARG;
  if (variant < 0 || variant >= li_variant_count())
    return set_error("LI variant %d is not in [0, %d)", variant, li_variant_count()), RBS_E_ARG;
  int x, y, f;
  li_variant_shape(variant, &x, &y, &f);
  *bx = x;
  *by = y;
  *family = f;
  *is_default = variant == li_default_variant() ? 1 : 0;
  return RBS_OK;
}

}  // extern "C"

// ---- device-driven runs ----------------------------------------------------------
namespace rbs {

// One device-driven step, as enqueued on the stream (and captured in the graph).
static int enqueue_driven_step(rbs_state *s, int scheme, int limit) {
  int rc;
  if (s->inflow.count > 0) {
    rc = inflow_launch(s, true, 0.0, 0.0);
    if (rc) return rc;
  }
  if (scheme == RBS_SCHEME_LI)
    rc = li_launch_step(s, true, 0.0, 0.0, limit, 0.0);
  else
    rc = hllc_launch_step(s, true, 0.0, 0.0, 0.0, 0.0);
  if (rc) return rc;
  return finalize_launch(s, FIN_DRIVEN, scheme);
}

static int build_graph(rbs_state *s, int scheme, int limit, int nsteps) {
  const int saved = s->cur;
  RBS_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = RBS_OK;
  for (int k = 0; k < nsteps && rc == RBS_OK; ++k) rc = enqueue_driven_step(s, scheme, limit);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->cur = saved;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&s->graph[0], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  return RBS_OK;
}

}  // namespace rbs

namespace rbs {
// Everything a device-driven run needs before its first step: validation, the sync
// table, run parameters, stage curves, and the priming reduction + first plan.
int driven_setup(rbs_state *s, const rbs_run_params *p, double t_stop, int64_t max_steps,
                 int64_t *steps0_out) {
  if (!s || !p) return set_error("null argument"), RBS_E_ARG;
  if (p->scheme != RBS_SCHEME_LI && p->scheme != RBS_SCHEME_HLLC)
    return set_error("unknown scheme %d", p->scheme), RBS_E_ARG;
  if (p->n_sync < 1 || !p->sync_times) return set_error("a run needs sync times"), RBS_E_ARG;
  for (int k = 1; k < p->n_sync; ++k)
    if (p->sync_times[k] < p->sync_times[k - 1])
      return set_error("sync times must be sorted ascending"), RBS_E_ARG;
  if (!(p->alpha > 0.0) || !(p->dt_max > 0.0)) return set_error("alpha and dt_max must be > 0"), RBS_E_ARG;
  if (p->scheme == RBS_SCHEME_HLLC && !(s->armed & RBS_ARM_HLLC)) {
    int rc = arm_fields(s, RBS_ARM_HLLC);
    if (rc) return rc;
  }
  if (p->rain_is_field && !(s->armed & RBS_ARM_RAIN))
    return set_error("rain_is_field needs an armed rain field"), RBS_E_STATE;
  if (max_steps <= 0) max_steps = INT64_MAX / 4;
  // sync table
  if (p->n_sync > s->sync_cap) {
    if (s->d_sync) RBS_HIP(hipFree(s->d_sync));
    s->d_sync = nullptr;
    RBS_HIP(hipMalloc(&s->d_sync, sizeof(double) * p->n_sync));
    s->sync_cap = p->n_sync;
    invalidate_graph(s);  // captured finalize kernels hold the table pointer
  }
  RBS_HIP(hipMemcpyAsync(s->d_sync, p->sync_times, sizeof(double) * p->n_sync,
                         hipMemcpyHostToDevice, s->stream));
  // steps so far, for the step limit
  RBS_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  const int64_t steps0 = s->ctl_host->steps;
  hipLaunchKernelGGL(run_config_kernel, dim3(1), dim3(1), 0, s->stream, s->ctl, p->scheme,
                     p->limit, p->rain_is_field, p->n_sync, p->alpha, p->dt_max, p->rain_m_s,
                     p->rain_duration, p->rain_field_sum != 0.0 ? p->rain_field_sum : s->rain_field_sum,
                     t_stop, steps0 + max_steps);
  RBS_HIP(hipGetLastError());
  {
    int ns[4];
    for (int e = 0; e < 4; ++e) {
      ns[e] = (s->bc[e] == RBS_BC_FIXED_STAGE) ? s->stage[e].n : 0;
      if (s->bc[e] == RBS_BC_FIXED_STAGE && ns[e] < 1)
        return set_error("fixed_stage edge %d without a stage curve", e), RBS_E_STATE;
    }
    hipLaunchKernelGGL(stage_config_kernel, dim3(1), dim3(1), 0, s->stream, s->ctl, ns[0], ns[1],
                       ns[2], ns[3], s->stage[0].d_t, s->stage[1].d_t, s->stage[2].d_t,
                       s->stage[3].d_t, s->stage[0].d_level, s->stage[1].d_level,
                       s->stage[2].d_level, s->stage[3].d_level);
    RBS_HIP(hipGetLastError());
  }
  // prime: CFL reduction of the current state + the first plan
  int rc = reduce_speed_launch(s, p->scheme, FIN_PRIME_DRIVEN);
  if (rc) return rc;
  *steps0_out = steps0;
  return RBS_OK;
}
}  // namespace rbs

extern "C" int rbs_run_until(rbs_state *s, const rbs_run_params *p, double t_stop,
                             int64_t max_steps, int64_t *steps_taken) {
  std::shared_lock<std::shared_mutex> api(api_mutex());
  int64_t steps0 = 0;
  if (max_steps <= 0) max_steps = INT64_MAX / 4;
  int rc = driven_setup(s, p, t_stop, max_steps, &steps0);
  if (rc) return rc;
  const int key = p->scheme * 2 + (p->limit ? 1 : 0);
  if (s->graph_key != key) {
    invalidate_graph(s);
    s->graph_key = key;
  }
  // The step kernels read the ping-pong parity from the control block, so one
  // captured graph serves every starting parity.
  const int chunk = s->graph_chunk;
  int64_t taken = 0;
  for (;;) {
    RBS_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s->stream));
    RBS_HIP(hipStreamSynchronize(s->stream));
    const DevCtl &c = *s->ctl_host;
    taken = c.steps - steps0;
    s->cur = c.cur;
    if (c.active == -1) {
      set_error("non-positive timestep at t=%.9g", c.t);
      rc = RBS_E_RANGE;
      break;
    }
    if (c.active != 1) break;
    // launch about the remaining number of steps before looking again
    const double est = c.dt > 0.0 ? (t_stop - c.t) / c.dt : 1.0;
    int64_t want = (int64_t)std::ceil(std::min(est, 1e9)) + 1;
    want = std::min<int64_t>(want, steps0 + max_steps - c.steps);
    // whole chunks from the graph, and a remainder shorter than a chunk as direct
    // launches, so that a step limit is met without replaying steps that find the
    // limit reached and return at once
    int64_t launches = std::min<int64_t>(want / chunk, 64);
    const int64_t rest = launches == 64 ? 0 : want % chunk;
    if (launches > 0 && !s->graph[0]) {
      api.unlock();
      {
        RBS_API_CAPTURE;  // no other thread's runtime calls during the capture
        rc = build_graph(s, p->scheme, p->limit, chunk);
      }
      api.lock();
      if (rc) return rc;
    }
    for (int64_t L = 0; L < launches; ++L) RBS_HIP(hipGraphLaunch(s->graph[0], s->stream));
    for (int64_t k = 0; k < rest; ++k) {
      rc = enqueue_driven_step(s, p->scheme, p->limit);
      if (rc) return rc;
    }
    if (launches == 0 && rest == 0) {  // nothing estimated: one step to move the clock
      rc = enqueue_driven_step(s, p->scheme, p->limit);
      if (rc) return rc;
    }
  }
  s->speed_scheme = p->scheme;
  if (steps_taken) *steps_taken = taken;
  return rc;
}
