// rbs_batch.hip -- many local-inertial scenarios advanced by one launch per step.
//
// The reference runs one scenario per process (run.py:140).  A batch holds up to
// `capacity` states of any grid sizes; each step is one batched fused-LI launch
// over the flat tile space of all members (a tile table maps every block back to
// its member and tile) and one batched finalize with a block per member, so every
// member keeps its own clock, CFL dt, sync table and float64 ledger.  Members that
// reached their stop time idle: their tiles return at once.  Steps are replayed
// from a captured HIP graph; adding or removing a member rebuilds the tables and
// the graph between replays, which is what continuous batching needs.
#include <algorithm>
#include <chrono>
#include <vector>

#include "rbs_device.hpp"

using namespace rbs;

struct rbs_batch {
  int capacity = 0;
  std::vector<rbs_state *> slot;  // slot -> member (nullptr = free)
  hipStream_t stream = nullptr;
  // the bedload and wide segments' launches run on side streams, forked from and
  // joined to `stream` inside the captured graph, beside the plain segment's
  hipStream_t side = nullptr, side_wide = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, join_wide_ev = nullptr;
  int variant = li_default_batch_variant();    // members without the bedload integral
  int variant_sed = li_default_sed_variant();  // members with it
  int chunk = 16;
  // wide members without the bedload integral (at least li_wide_batch_min_nx()
  // columns) run the row-block kernel in their own segment, as they would alone
  // device tables (rebuilt when membership changes)
  void *d_args = nullptr;      // LiArgs[capacity]
  FinArgs *d_fin = nullptr;    // [capacity]
  InflowArgs *d_inflow = nullptr;  // [capacity]: members with point sources
  int n_inflow = 0;
  int32_t *d_status = nullptr; // [capacity]
  int32_t *h_status = nullptr; // pinned mirror
  int *d_tile_member = nullptr, *d_tile_first = nullptr;
  int tile_cap = 0;
  int total_tiles = 0;
  BatchDims *d_dims = nullptr;  // the live counts the captured kernels read
  char *h_stage = nullptr;      // pinned staging of the tables (one rebuild's worth)
  size_t stage_cap = 0;
  bool any_sed = false, any_wide = false, any_inflow = false;
  // what the current graph was captured for: launch sizes and optional nodes
  int graph_tiles = 0;
  bool graph_sed = false, graph_wide = false, graph_inflow = false;
  std::vector<int> order;      // compact index -> slot
  std::vector<int> active;     // per slot: 1 running, 0 done
  hipGraphExec_t graph = nullptr;
  bool dirty = true;
  // host-side cost of membership changes
  int64_t rebuilds = 0, graph_launches = 0;
  double rebuild_ms = 0.0;
  // an advance in flight (launched, not yet collected); the event pair times it
  hipEvent_t start_ev = nullptr, done_ev = nullptr;
  double gpu_ms = 0.0;
  bool in_flight = false;
  int in_flight_n = 0;
};

static void drop_graph(rbs_batch *b) {
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
}

// Rewrite the member tables for the current membership.  The graph is kept: its
// kernels read the live counts from d_dims.  It is dropped (and captured again at the
// next launch) only when the membership outgrows what it was captured for.
static int rebuild(rbs_batch *b) {
  b->order.clear();
  for (int k = 0; k < b->capacity; ++k)
    if (b->slot[k] && b->active[k]) b->order.push_back(k);
  const int n = (int)b->order.size();
  b->total_tiles = 0;
  b->n_inflow = 0;
  b->dirty = false;
  const size_t asz = li_args_size();
  std::vector<char> args(asz * (n > 0 ? n : 1));
  std::vector<FinArgs> fin(n > 0 ? n : 1);
  // three segments of the flat tile space, each launched with its own kernel: plain
  // members on the batch's variant, wide plain members on the row blocks, then the
  // members with the bedload integral; first[i] is member i's first tile within its
  // segment
  std::vector<int> first(n > 0 ? n : 1), seg_member[kBatchSegments];
  std::vector<InflowArgs> inflow;
  for (int i = 0; i < n; ++i) {
    rbs_state *s = b->slot[b->order[i]];
    const bool sed = (s->armed & RBS_ARM_SED) != 0;
    const bool wide = !sed && s->nx >= li_wide_batch_min_nx();
    const int sg = sed ? 1 : (wide ? 2 : 0);
    const int v = sg == 1 ? b->variant_sed : (sg == 2 ? li_wide_batch_variant() : b->variant);
    if (v != s->li_variant) {
      // another tiling: clear the dry-tile words in the batch's stream order
      const int rc = quiet_reset(s, b->stream);
      if (rc) return rc;
    }
    s->li_variant = v;
    li_fill_args(s, 1, reinterpret_cast<LiArgs *>(args.data() + asz * i));
    const int nt = li_num_blocks(s);
    std::vector<int> &seg = seg_member[sg];
    first[i] = (int)seg.size();
    for (int t = 0; t < nt; ++t) seg.push_back(i);
    if (sed) b->any_sed = true;
    if (wide) b->any_wide = true;
    if (s->inflow.count > 0) {
      inflow.emplace_back();
      inflow_fill_args(s, &inflow.back());
      b->any_inflow = true;
    }
    fin[i].ctl = s->ctl;
    fin[i].part = s->part;
    fin[i].sync = s->d_sync;
    fin[i].area = s->dx * s->dx;
    fin[i].n_cells = (double)s->ny * (double)s->nx;
    fin[i].dx = s->dx;
    fin[i].nblk = nt;
  }
  b->n_inflow = (int)inflow.size();
  // flat order: plain, wide, bedload
  BatchDims dims{};
  dims.seg_first[0] = 0;
  dims.seg_tiles[0] = (int)seg_member[0].size();
  dims.seg_first[2] = dims.seg_tiles[0];
  dims.seg_tiles[2] = (int)seg_member[2].size();
  dims.seg_first[1] = dims.seg_first[2] + dims.seg_tiles[2];
  dims.seg_tiles[1] = (int)seg_member[1].size();
  dims.members = n;
  dims.inflow_members = b->n_inflow;
  std::vector<int> member = seg_member[0];
  member.insert(member.end(), seg_member[2].begin(), seg_member[2].end());
  member.insert(member.end(), seg_member[1].begin(), seg_member[1].end());
  b->total_tiles = (int)member.size();
  if (b->total_tiles > b->tile_cap) {  // grow by half again: few reallocations
    RBS_HIP(hipStreamSynchronize(b->stream));
    if (b->d_tile_member) RBS_HIP(hipFree(b->d_tile_member));
    b->d_tile_member = nullptr;
    const int cap = std::max(b->total_tiles, b->tile_cap + b->tile_cap / 2);
    RBS_HIP(hipMalloc(&b->d_tile_member, sizeof(int) * cap));
    b->tile_cap = cap;
    drop_graph(b);  // it holds the old table
  }
  if (b->graph && (b->total_tiles > b->graph_tiles || (b->any_sed && !b->graph_sed) ||
                   (b->any_wide && !b->graph_wide) || (b->any_inflow && !b->graph_inflow)))
    drop_graph(b);
  // stage every table in pinned memory, then copy them with DMA in stream order
  struct Piece {
    void *dev;
    const void *src;
    size_t bytes;
  };
  const Piece pieces[] = {
      {b->d_args, args.data(), n > 0 ? asz * n : 0},
      {b->d_fin, fin.data(), sizeof(FinArgs) * n},
      {b->d_tile_first, first.data(), sizeof(int) * n},
      {b->d_tile_member, member.data(), sizeof(int) * member.size()},
      {b->d_inflow, inflow.data(), sizeof(InflowArgs) * inflow.size()},
      {b->d_dims, &dims, sizeof(dims)},
  };
  size_t need = 0;
  for (const Piece &q : pieces) need += (q.bytes + 255) & ~size_t(255);
  if (need > b->stage_cap) {
    if (b->h_stage) RBS_HIP(hipHostFree(b->h_stage));
    b->h_stage = nullptr;
    const size_t cap = std::max(need, b->stage_cap + b->stage_cap / 2);
    RBS_HIP(hipHostMalloc((void **)&b->h_stage, cap));
    b->stage_cap = cap;
  }
  size_t off = 0;
  for (const Piece &q : pieces) {
    if (q.bytes == 0) continue;
    std::memcpy(b->h_stage + off, q.src, q.bytes);
    RBS_HIP(hipMemcpyAsync(q.dev, b->h_stage + off, q.bytes, hipMemcpyHostToDevice, b->stream));
    off += (q.bytes + 255) & ~size_t(255);
  }
  RBS_HIP(hipStreamSynchronize(b->stream));  // the staging is reused by the next rebuild
  return RBS_OK;
}

// Capture `chunk` batched steps with launch sizes covering the tile capacity and the
// member capacity; the kernels read the live counts from d_dims.  The segments touch
// disjoint members, so the bedload and wide segments' launches fork off to side
// streams and join before the finalize: in the graph the three run side by side, and
// a step of a mixed batch takes the longest of them rather than their sum.
static int capture(rbs_batch *b) {
  const int grid_tiles = b->tile_cap;
  const bool sed = b->any_sed, wide = b->any_wide, inflow = b->any_inflow;
  RBS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = RBS_OK;
  auto ok = [&](hipError_t e, const char *what) {
    if (rc == RBS_OK && e != hipSuccess) rc = hip_fail(e, what);
  };
  for (int k = 0; k < b->chunk && rc == RBS_OK; ++k) {
    // point sources of the members that have them (inflow.py:167-200), one launch
    if (inflow) rc = inflow_batch_launch(b->d_inflow, b->d_dims, b->capacity, b->stream);
    if (rc == RBS_OK && (sed || wide))
      ok(hipEventRecord(b->fork_ev, b->stream), "hipEventRecord(fork)");
    if (rc == RBS_OK && sed) {
      ok(hipStreamWaitEvent(b->side, b->fork_ev, 0), "hipStreamWaitEvent(fork)");
      if (rc == RBS_OK)
        rc = li_batch_launch(b->variant_sed, 1, (const LiArgs *)b->d_args, b->d_tile_member,
                             b->d_tile_first, b->d_dims, grid_tiles, b->side);
      ok(hipEventRecord(b->join_ev, b->side), "hipEventRecord(join)");
    }
    if (rc == RBS_OK && wide) {
      ok(hipStreamWaitEvent(b->side_wide, b->fork_ev, 0), "hipStreamWaitEvent(fork)");
      if (rc == RBS_OK)
        rc = li_batch_launch(li_wide_batch_variant(), 2, (const LiArgs *)b->d_args,
                             b->d_tile_member, b->d_tile_first, b->d_dims, grid_tiles,
                             b->side_wide);
      ok(hipEventRecord(b->join_wide_ev, b->side_wide), "hipEventRecord(join)");
    }
    if (rc == RBS_OK)
      rc = li_batch_launch(b->variant, 0, (const LiArgs *)b->d_args, b->d_tile_member,
                           b->d_tile_first, b->d_dims, grid_tiles, b->stream);
    if (sed) ok(hipStreamWaitEvent(b->stream, b->join_ev, 0), "hipStreamWaitEvent(join)");
    if (wide) ok(hipStreamWaitEvent(b->stream, b->join_wide_ev, 0), "hipStreamWaitEvent(join)");
    if (rc == RBS_OK)
      rc = batch_finalize_launch(b->d_fin, b->d_dims, b->capacity, b->d_status, b->stream);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  b->graph_tiles = grid_tiles;
  b->graph_sed = sed;
  b->graph_wide = wide;
  b->graph_inflow = inflow;
  return RBS_OK;
}

extern "C" {

static int batch_destroy(rbs_batch *b);

int rbs_batch_create(int32_t capacity, rbs_batch **out) {
  RBS_API_SHARED;
  if (!out || capacity < 1 || capacity > 65536) return set_error("bad batch capacity"), RBS_E_ARG;
  rbs_batch *b = new rbs_batch();
  b->capacity = capacity;
  b->slot.assign(capacity, nullptr);
  b->active.assign(capacity, 0);
  auto fail = [&](int rc) {
    batch_destroy(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side_wide, hipStreamNonBlocking) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipStreamCreate"));
  if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join_wide_ev, hipEventDisableTiming) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipEventCreate(fork/join)"));
  if (hipEventCreate(&b->start_ev) != hipSuccess || hipEventCreate(&b->done_ev) != hipSuccess)
    return fail(hip_fail(hipErrorUnknown, "hipEventCreate"));
  if (hipMalloc(&b->d_args, li_args_size() * capacity) != hipSuccess ||
      hipMalloc(&b->d_fin, sizeof(FinArgs) * capacity) != hipSuccess ||
      hipMalloc(&b->d_inflow, sizeof(InflowArgs) * capacity) != hipSuccess ||
      hipMalloc(&b->d_dims, sizeof(BatchDims)) != hipSuccess ||
      hipMalloc(&b->d_status, sizeof(int32_t) * capacity) != hipSuccess ||
      hipMalloc(&b->d_tile_first, sizeof(int) * capacity) != hipSuccess)
    return fail(hip_fail(hipErrorOutOfMemory, "hipMalloc(batch tables)"));
  if (hipHostMalloc(&b->h_status, sizeof(int32_t) * capacity) != hipSuccess)
    return fail(hip_fail(hipErrorOutOfMemory, "hipHostMalloc(batch status)"));
  *out = b;
  return RBS_OK;
}

int rbs_batch_destroy(rbs_batch *b) {
  RBS_API_SHARED;
  return batch_destroy(b);
}

static int batch_destroy(rbs_batch *b) {
  if (!b) return RBS_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  void *p[] = {b->d_args, b->d_fin, b->d_inflow, b->d_dims, b->d_status, b->d_tile_member,
               b->d_tile_first};
  for (void *q : p)
    if (q) (void)hipFree(q);
  if (b->h_status) (void)hipHostFree(b->h_status);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  if (b->done_ev) (void)hipEventDestroy(b->done_ev);
  if (b->start_ev) (void)hipEventDestroy(b->start_ev);
  if (b->join_wide_ev) (void)hipEventDestroy(b->join_wide_ev);
  if (b->join_ev) (void)hipEventDestroy(b->join_ev);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->side_wide) (void)hipStreamDestroy(b->side_wide);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RBS_OK;
}

int rbs_batch_add(rbs_batch *b, rbs_state *st, const rbs_run_params *p, double t_stop,
                  int32_t *slot) {
  RBS_API_SHARED;
  if (!b || !st || !p || !slot) return set_error("null argument"), RBS_E_ARG;
  if (b->in_flight) return set_error("an advance is in flight: collect it first"), RBS_E_STATE;
  if (p->scheme != RBS_SCHEME_LI) return set_error("a batch advances local-inertial scenarios"), RBS_E_ARG;
  if (!p->limit) return set_error("batched steps always run the donor limiter"), RBS_E_ARG;
  if (st->armed & RBS_ARM_CHAN)
    return set_error("states with sub-grid channels run on their own (rbs_run_until), not in a "
                     "batch"), RBS_E_STATE;
  for (int k = 0; k < b->capacity; ++k)
    if (b->slot[k] == st) return set_error("state is already in the batch"), RBS_E_STATE;
  int k = 0;
  while (k < b->capacity && b->slot[k]) ++k;
  if (k == b->capacity) return set_error("batch is full (%d members)", b->capacity), RBS_E_NOMEM;
  RBS_HIP(hipStreamSynchronize(b->stream));
  int64_t steps0 = 0;
  int rc = driven_setup(st, p, t_stop, 0, &steps0);
  if (rc) return rc;
  RBS_HIP(hipStreamSynchronize(st->stream));
  b->slot[k] = st;
  b->active[k] = 1;
  b->dirty = true;
  *slot = k;
  return RBS_OK;
}

int rbs_batch_remove(rbs_batch *b, int32_t slot) {
  RBS_API_SHARED;
  if (!b || slot < 0 || slot >= b->capacity || !b->slot[slot])
    return set_error("no member in slot %d", slot), RBS_E_ARG;
  if (b->in_flight) return set_error("an advance is in flight: collect it first"), RBS_E_STATE;
  RBS_HIP(hipStreamSynchronize(b->stream));
  rbs_state *s = b->slot[slot];
  RBS_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s->stream));
  RBS_HIP(hipStreamSynchronize(s->stream));
  s->cur = s->ctl_host->cur;
  s->speed_scheme = RBS_SCHEME_LI;
  b->slot[slot] = nullptr;
  b->active[slot] = 0;
  b->dirty = true;
  return RBS_OK;
}

int rbs_batch_launch(rbs_batch *b, int32_t steps) {
  if (!b || steps < 1) return set_error("bad arguments"), RBS_E_ARG;
  if (b->in_flight) return set_error("an advance is in flight: collect it first"), RBS_E_STATE;
  const auto t0 = std::chrono::steady_clock::now();
  bool rebuilt = false;
  // Every runtime call below runs under the library lock, shared, as in every other
  // entry point, so that no capture on another thread overlaps it; the capture takes
  // the lock exclusively.
  std::shared_lock<std::shared_mutex> api(api_mutex());
  if (b->dirty) {
    int rc = rebuild(b);
    if (rc) return rc;
    rebuilt = true;
  }
  if (!b->order.empty() && !b->graph) {
    api.unlock();
    int rc;
    {
      RBS_API_CAPTURE;  // no other thread's runtime calls during the capture
      rc = capture(b);
    }
    api.lock();
    if (rc) return rc;
    rebuilt = true;
  }
  const int n = (int)b->order.size();
  if (n == 0) return RBS_OK;
  if (rebuilt) {
    b->rebuilds += 1;
    b->rebuild_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  const int launches = (steps + b->chunk - 1) / b->chunk;
  RBS_HIP(hipEventRecord(b->start_ev, b->stream));
  for (int L = 0; L < launches; ++L) RBS_HIP(hipGraphLaunch(b->graph, b->stream));
  b->graph_launches += launches;
  RBS_HIP(hipMemcpyAsync(b->h_status, b->d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                         b->stream));
  RBS_HIP(hipEventRecord(b->done_ev, b->stream));
  b->in_flight = true;
  b->in_flight_n = n;
  return RBS_OK;
}

int rbs_batch_poll(rbs_batch *b, int32_t *done) {
  RBS_API_SHARED;
  if (!b || !done) return set_error("null argument"), RBS_E_ARG;
  if (!b->in_flight) {
    *done = 1;
    return RBS_OK;
  }
  const hipError_t e = hipEventQuery(b->done_ev);
  if (e == hipErrorNotReady) {
    *done = 0;
    return RBS_OK;
  }
  RBS_HIP(e);
  *done = 1;
  return RBS_OK;
}

int rbs_batch_collect(rbs_batch *b, int32_t *n_active) {
  if (!b) return set_error("null batch"), RBS_E_ARG;
  if (b->in_flight) {
    RBS_API_SHARED;  // as rbs_run_until waits on its stream
    RBS_HIP(hipEventSynchronize(b->done_ev));
    b->in_flight = false;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->start_ev, b->done_ev) == hipSuccess) b->gpu_ms += ms;
    bool changed = false;
    for (int i = 0; i < b->in_flight_n; ++i) {
      const int k = b->order[i];
      const int32_t w = b->h_status[i];
      b->slot[k]->cur = (w >> 1) & 1;
      if (w & 4) return set_error("member %d: non-positive timestep", k), RBS_E_RANGE;
      if (!(w & 1)) {
        b->active[k] = 0;
        changed = true;
      }
    }
    if (changed) b->dirty = true;  // finished members drop out of the tile space
  }
  int c = 0;
  for (int k = 0; k < b->capacity; ++k) c += (b->slot[k] && b->active[k]) ? 1 : 0;
  if (n_active) *n_active = c;
  return RBS_OK;
}

int rbs_batch_advance(rbs_batch *b, int32_t steps, int32_t *n_active) {
  int rc = rbs_batch_launch(b, steps);
  if (rc) return rc;
  return rbs_batch_collect(b, n_active);
}

int rbs_batch_member_active(rbs_batch *b, int32_t slot, int32_t *active) {
  if (!b || !active || slot < 0 || slot >= b->capacity || !b->slot[slot])
    return set_error("no member in slot %d", slot), RBS_E_ARG;
  *active = b->active[slot];
  return RBS_OK;
}

int rbs_batch_stats(rbs_batch *b, int64_t *rebuilds, double *rebuild_ms,
                    int64_t *graph_launches, double *gpu_ms) {
  if (!b || !rebuilds || !rebuild_ms || !graph_launches || !gpu_ms)
    return set_error("null argument"), RBS_E_ARG;
  *rebuilds = b->rebuilds;
  *rebuild_ms = b->rebuild_ms;
  *graph_launches = b->graph_launches;
  *gpu_ms = b->gpu_ms;
  return RBS_OK;
}

int rbs_batch_get_variant(rbs_batch *b, int32_t *variant, int32_t *variant_sed) {
  if (!b || !variant || !variant_sed) return set_error("null argument"), RBS_E_ARG;
  *variant = b->variant;
  *variant_sed = b->variant_sed;
  return RBS_OK;
}

int rbs_batch_set_variant(rbs_batch *b, int32_t variant, int32_t variant_sed) {
  if (!b) return set_error("null batch"), RBS_E_ARG;
  if (b->in_flight) return set_error("an advance is in flight: collect it first"), RBS_E_STATE;
  for (int v : {variant, variant_sed})
    if (v < 0 || v >= li_variant_count())
      return set_error("LI variant %d is not in [0, %d)", v, li_variant_count()), RBS_E_ARG;
  if (variant != b->variant || variant_sed != b->variant_sed) {
    b->variant = variant;
    b->variant_sed = variant_sed;
    b->dirty = true;  // tables follow the new tile shapes ...
    drop_graph(b);    // ... and the graph launches the new kernels
  }
  return RBS_OK;
}

}  // extern "C"
