// rbs_frames.hip -- output frames at the output cadence without stalling the steps.
//
// The reference's output step (run.py:389-400) reads the state back to the host
// synchronously: h.numpy(), State.velocities_numpy() (state.py:232-256, cell-centred
// u, v from the face discharges or the HLLC momentum, 0 where h < H_DRY) and, with
// morphology, the float64 bed change narrowed to float32 (zarr_writer.py:143-147),
// then appends them to the result store.
//
// Here a frame is one kernel on the state's own stream, enqueued after the steps that
// lead to it: it forms depth, u, v and bed change into a device slot, and records the
// clock.  A second stream copies the slot to pinned host memory while the state's
// stream goes on stepping; the host waits for a frame only when it writes it out.
// The frame arithmetic is the reference's float32 numpy arithmetic, so frames equal
// what velocities_numpy computes from the same fields, bit for bit.
#include <mutex>
#include <vector>

#include "rbs_device.hpp"

namespace rbs {

struct FrameArgs {
  const float *h[2];   // the current depth is h[ctl->cur]
  const float *qx[2];  // LI faces (ping-pong), or null for an HLLC state
  const float *qy[2];
  const float *hu, *hv;  // HLLC momentum, or null
  const double *dz_cum;  // bed change, or null when not requested
  const DevCtl *ctl;
  float *depth, *u, *v, *bed;  // device slot, dense (ny, nx) planes; null = not requested
  double *t;                   // the frame's clock
  int64_t pitch;
  int32_t ny, nx;
};

// One row of cells per block-row: coalesced along x.
__global__ __launch_bounds__(kThreads) void frame_kernel(FrameArgs a) {
  const int cur = a.ctl->cur;
  const float *__restrict__ h = a.h[cur];
  const int64_t P = a.pitch;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *a.t = a.ctl->t;
  for (int r = blockIdx.y; r < a.ny; r += gridDim.y) {
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < a.nx; c += gridDim.x * kThreads) {
      const int64_t o = (int64_t)r * P + c;
      const int64_t od = (int64_t)r * a.nx + c;  // dense: one plane is one copy
      const float hh = h[o];
      if (a.depth) a.depth[od] = hh;
      if (a.u) {
        const bool wet = hh >= kHDry;  // state.py:244
        float mx, my;
        if (a.hu) {  // HLLC: cell-centred momentum (:248-250)
          mx = a.hu[o];
          my = a.hv[o];
        } else {  // LI: the two bounding face discharges averaged (:252-255)
          const float *__restrict__ qx = a.qx[cur];
          const float *__restrict__ qy = a.qy[cur];
          mx = 0.5f * (qx[o] + qx[o + 1]);
          my = 0.5f * (qy[o] + qy[o + P]);
        }
        a.u[od] = wet ? mx / hh : 0.0f;
        a.v[od] = wet ? my / hh : 0.0f;
      }
      if (a.bed) a.bed[od] = (float)a.dz_cum[o];  // zarr_writer.py:147
    }
  }
}

}  // namespace rbs

struct rbs_frames {
  rbs_state *st = nullptr;
  uint32_t fields = 0;
  int slots = 0;
  int nfield = 0;
  hipStream_t copy = nullptr;
  // per slot: device staging and pinned host buffers (nfield dense planes + the
  // clock each), events, state
  std::vector<float *> dev;
  std::vector<double *> dev_t;
  std::vector<float *> host;
  std::vector<double *> host_t;
  std::vector<hipEvent_t> formed, copied;
  std::vector<int> busy;  // 1: captured and not yet released (guarded by `m`)
  std::mutex m;           // capture may run on one host thread, wait/release on another
  uint64_t bytes = 0;
};

namespace rbs {

static int frames_free(rbs_frames *q) {
  if (!q) return RBS_OK;
  if (q->copy) (void)hipStreamSynchronize(q->copy);
  for (auto p : q->dev)
    if (p) (void)hipFree(p);
  for (auto p : q->dev_t)
    if (p) (void)hipFree(p);
  for (auto p : q->host)
    if (p) (void)hipHostFree(p);
  for (auto p : q->host_t)
    if (p) (void)hipHostFree(p);
  for (auto e : q->formed)
    if (e) (void)hipEventDestroy(e);
  for (auto e : q->copied)
    if (e) (void)hipEventDestroy(e);
  if (q->copy) (void)hipStreamDestroy(q->copy);
  delete q;
  return RBS_OK;
}

// The plane index of each requested field inside a slot, in a fixed order.
static int plane_of(uint32_t fields, uint32_t which) {
  int k = 0;
  for (uint32_t f : {RBS_FRAME_DEPTH, RBS_FRAME_U, RBS_FRAME_V, RBS_FRAME_BED_CHANGE}) {
    if (f == which) return (fields & f) ? k : -1;
    if (fields & f) ++k;
  }
  return -1;
}

static int capture_into(rbs_frames *q, int k);

}  // namespace rbs

using namespace rbs;

extern "C" {

int rbs_frames_create(rbs_state *s, uint32_t fields, int32_t slots, rbs_frames **out) {
  RBS_API_SHARED;
  if (!s || !out) return set_error("null argument"), RBS_E_ARG;
  *out = nullptr;
  const uint32_t all = RBS_FRAME_DEPTH | RBS_FRAME_U | RBS_FRAME_V | RBS_FRAME_BED_CHANGE;
  if (fields == 0 || (fields & ~all)) return set_error("bad frame field mask 0x%x", fields), RBS_E_ARG;
  if (((fields & RBS_FRAME_U) != 0) != ((fields & RBS_FRAME_V) != 0))
    return set_error("u and v are written together (state.py:232-256)"), RBS_E_ARG;
  if (slots < 1 || slots > 64) return set_error("slots must be in [1, 64], got %d", slots), RBS_E_ARG;
  if ((fields & RBS_FRAME_BED_CHANGE) && !(s->armed & RBS_ARM_SED))
    return set_error("bed-change frames need a state armed with RBS_ARM_SED"), RBS_E_STATE;
  rbs_frames *q = new rbs_frames();
  q->st = s;
  q->fields = fields;
  q->slots = slots;
  for (uint32_t f = fields; f; f &= f - 1) ++q->nfield;
  auto fail = [&](int code) {
    frames_free(q);
    return code;
  };
  if (hipStreamCreateWithFlags(&q->copy, hipStreamNonBlocking) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipStreamCreate"));
  const size_t plane_dev = (size_t)s->ny * s->nx * sizeof(float);
  const size_t plane_host = plane_dev;
  q->dev.assign(slots, nullptr);
  q->dev_t.assign(slots, nullptr);
  q->host.assign(slots, nullptr);
  q->host_t.assign(slots, nullptr);
  q->formed.assign(slots, nullptr);
  q->copied.assign(slots, nullptr);
  q->busy.assign(slots, 0);
  for (int k = 0; k < slots; ++k) {
    if (hipMalloc(&q->dev[k], plane_dev * q->nfield) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for frame slot %d failed", plane_dev * q->nfield, k);
      return fail(RBS_E_NOMEM);
    }
    if (hipMalloc(&q->dev_t[k], sizeof(double)) != hipSuccess)
      return fail(hip_fail(hipErrorOutOfMemory, "hipMalloc(frame clock)"));
    if (hipHostMalloc(&q->host[k], plane_host * q->nfield) != hipSuccess) {
      set_error("hipHostMalloc of %zu pinned bytes for frame slot %d failed",
                plane_host * q->nfield, k);
      return fail(RBS_E_NOMEM);
    }
    if (hipHostMalloc(&q->host_t[k], sizeof(double)) != hipSuccess)
      return fail(hip_fail(hipErrorOutOfMemory, "hipHostMalloc(frame clock)"));
    if (hipEventCreateWithFlags(&q->formed[k], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&q->copied[k], hipEventDisableTiming) != hipSuccess)
      return fail(hip_fail(hipErrorUnknown, "hipEventCreate"));
    q->bytes += plane_dev * q->nfield + sizeof(double);
  }
  *out = q;
  return RBS_OK;
}

int rbs_frames_destroy(rbs_frames *q) {
  RBS_API_SHARED;
  return frames_free(q);
}

int rbs_frames_capture(rbs_frames *q, int32_t *slot_out) {
  RBS_API_SHARED;
  if (!q || !slot_out) return set_error("null argument"), RBS_E_ARG;
  int k = 0;
  {
    std::lock_guard<std::mutex> g(q->m);
    while (k < q->slots && q->busy[k]) ++k;
    if (k == q->slots)
      return set_error("all %d frame slots are held: release one first", q->slots), RBS_E_STATE;
    q->busy[k] = 1;  // held from here; a failed enqueue below hands it back
  }
  int rc = capture_into(q, k);
  if (rc) {
    std::lock_guard<std::mutex> g(q->m);
    q->busy[k] = 0;
    return rc;
  }
  *slot_out = k;
  return RBS_OK;
}

}  // extern "C"

namespace rbs {
// The frame kernel and the copies of slot k (held by the caller).
static int capture_into(rbs_frames *q, int k) {
  rbs_state *s = q->st;
  FrameArgs a{};
  const bool hllc = (s->armed & RBS_ARM_HLLC) != 0;
  for (int p = 0; p < 2; ++p) {
    // an HLLC state keeps its depth in h[cur] without ping-pong
    a.h[p] = hllc ? s->h[s->cur] : s->h[p];
    a.qx[p] = s->qx[p];
    a.qy[p] = s->qy[p];
  }
  if (hllc) {
    a.hu = s->hu;
    a.hv = s->hv;
  }
  a.dz_cum = (q->fields & RBS_FRAME_BED_CHANGE) ? s->dz_cum : nullptr;
  a.ctl = s->ctl;
  const size_t plane = (size_t)s->ny * s->nx;
  auto at_plane = [&](uint32_t f) -> float * {
    const int p = plane_of(q->fields, f);
    return p < 0 ? nullptr : q->dev[k] + (size_t)p * plane;
  };
  a.depth = at_plane(RBS_FRAME_DEPTH);
  a.u = at_plane(RBS_FRAME_U);
  a.v = at_plane(RBS_FRAME_V);
  a.bed = at_plane(RBS_FRAME_BED_CHANGE);
  a.t = q->dev_t[k];
  a.pitch = s->pitch;
  a.ny = s->ny;
  a.nx = s->nx;
  const int gx = (s->nx + kThreads - 1) / kThreads;
  const int gy = s->ny < 4096 ? s->ny : 4096;
  hipLaunchKernelGGL(frame_kernel, dim3(gx, gy), dim3(kThreads), 0, s->stream, a);
  RBS_HIP(hipGetLastError());
  RBS_HIP(hipEventRecord(q->formed[k], s->stream));
  // the copy stream waits for the frame, not for the steps enqueued after it
  RBS_HIP(hipStreamWaitEvent(q->copy, q->formed[k], 0));
  // one 1-D copy of all planes (a 2-D copy would move row by row)
  RBS_HIP(hipMemcpyAsync(q->host[k], q->dev[k], plane * sizeof(float) * q->nfield,
                         hipMemcpyDeviceToHost, q->copy));
  RBS_HIP(hipMemcpyAsync(q->host_t[k], q->dev_t[k], sizeof(double), hipMemcpyDeviceToHost,
                         q->copy));
  RBS_HIP(hipEventRecord(q->copied[k], q->copy));
  return RBS_OK;
}
}  // namespace rbs

extern "C" {

int rbs_frames_ready(rbs_frames *q, int32_t slot, int32_t *ready) {
  RBS_API_SHARED;
  if (!q || !ready || slot < 0 || slot >= q->slots) return set_error("bad arguments"), RBS_E_ARG;
  {
    std::lock_guard<std::mutex> g(q->m);
    if (!q->busy[slot]) return set_error("frame slot %d holds no frame", slot), RBS_E_STATE;
  }
  const hipError_t e = hipEventQuery(q->copied[slot]);
  if (e == hipSuccess) *ready = 1;
  else if (e == hipErrorNotReady) *ready = 0;
  else return hip_fail(e, "hipEventQuery");
  return RBS_OK;
}

int rbs_frames_wait(rbs_frames *q, int32_t slot, double *t, float **depth, float **u, float **v,
                    float **bed_change) {
  RBS_API_SHARED;
  if (!q || slot < 0 || slot >= q->slots) return set_error("bad arguments"), RBS_E_ARG;
  {
    std::lock_guard<std::mutex> g(q->m);
    if (!q->busy[slot]) return set_error("frame slot %d holds no frame", slot), RBS_E_STATE;
  }
  RBS_HIP(hipEventSynchronize(q->copied[slot]));
  const rbs_state *s = q->st;
  const size_t plane = (size_t)s->ny * s->nx;
  auto host_plane = [&](uint32_t f) -> float * {
    const int p = plane_of(q->fields, f);
    return p < 0 ? nullptr : q->host[slot] + (size_t)p * plane;
  };
  if (t) *t = *q->host_t[slot];
  if (depth) *depth = host_plane(RBS_FRAME_DEPTH);
  if (u) *u = host_plane(RBS_FRAME_U);
  if (v) *v = host_plane(RBS_FRAME_V);
  if (bed_change) *bed_change = host_plane(RBS_FRAME_BED_CHANGE);
  return RBS_OK;
}

int rbs_frames_release(rbs_frames *q, int32_t slot) {
  RBS_API_SHARED;
  if (!q || slot < 0 || slot >= q->slots) return set_error("bad arguments"), RBS_E_ARG;
  {
    std::lock_guard<std::mutex> g(q->m);
    if (!q->busy[slot]) return set_error("frame slot %d holds no frame", slot), RBS_E_STATE;
  }
  RBS_HIP(hipEventSynchronize(q->copied[slot]));
  std::lock_guard<std::mutex> g(q->m);
  q->busy[slot] = 0;
  return RBS_OK;
}

int rbs_frames_bytes(const rbs_frames *q, uint64_t *bytes) {
  if (!q || !bytes) return set_error("null argument"), RBS_E_ARG;
  *bytes = q->bytes;
  return RBS_OK;
}

}  // extern "C"
