// rbs_internal.hpp -- shared definitions of librbs (gfx950).
//
// Device layout: every 2-D field is row-major with a padded pitch (a multiple of
// 64 elements, so each row starts on a 256-byte boundary).  Cells use columns
// [0, nx), x-faces [0, nx], rows [0, ny) for cells and x-faces, [0, ny] for
// y-faces.  All fields of one state share the pitch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <string>

#include "../../include/rbs.h"

namespace rbs {

// solver/core/grid.py:44,49 as the float32 kernels see them.
constexpr float kG = 9.81f;
constexpr float kHDry = 1.0e-3f;
constexpr double kGravity = 9.81;
constexpr double kHDry64 = 1.0e-3;
// solver/scheduler.py:45
constexpr double kEpsT = 1e-6;

constexpr int kThreads = 256;  // block size of every tiled kernel
// Block partials beyond kFoldMin are folded in two levels: kFoldBlocks blocks first,
// then the one-block finalize over their kFoldBlocks results.
constexpr int kFoldBlocks = 128;
constexpr int kFoldMin = 4096;

// ---- threads ----------------------------------------------------------------
// Distinct handles may be driven from distinct host threads.  A stream capture
// (device-driven runs, batches) must not overlap another thread's allocation, copy
// or launch through the runtime, so every entry point that issues HIP work holds this
// lock shared and the capture sections hold it exclusively (for a few milliseconds).
std::shared_mutex &api_mutex();
#define RBS_API_SHARED std::shared_lock<std::shared_mutex> rbs_api_lock_(::rbs::api_mutex())
#define RBS_API_CAPTURE std::unique_lock<std::shared_mutex> rbs_api_lock_(::rbs::api_mutex())

// ---- error plumbing ---------------------------------------------------------
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define RBS_HIP(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return ::rbs::hip_fail(e_, #call); \
  } while (0)

// ---- device control block ---------------------------------------------------
// One per state, in device memory.  The step kernels read their scalars from it,
// so one captured graph serves every step of a device-driven run.
struct DevCtl {
  // clock (scheduler.py:171-246), float64
  double t;        // simulated time at the start of the next step
  double dt;       // dt of the next step
  double t_stop;   // device-driven run target
  int32_t active;  // 1 while the next step is due
  int32_t sync_idx;
  int32_t cur;     // LI ping-pong buffer that holds the current state
  int32_t pad0;
  int64_t steps;
  // per-step forcing (run.py:378-388)
  float rain;        // uniform rate handed to the step (float32)
  float rain_scale;  // field multiplier (1 raining, 0 not)
  float speed_max;   // LI: max h;  HLLC: max wave speed (of the current state)
  int32_t stale;     // speed_max needs a fresh reduction
  // ledger, volumes in m^3
  double volume;
  double rain_in, rain_in_c;
  double point_in, point_in_c;
  double infil_out, infil_out_c;
  double edge_out, edge_out_c;
  double release_out, release_out_c;  // reservoir delivery rounding, banked
  // device-driven run parameters (copied from rbs_run_params)
  int32_t run_scheme, run_limit, run_field, n_sync;
  double alpha, dt_max, rain_m_s, rain_duration, rain_field_sum;
  int64_t step_limit;  // device-driven run stops after this many steps in total
  // fixed_stage edges (hllc.py:1308-1312): the prescribed surface of the next step,
  // float32 as the kernels see it, and the curves a device-driven run evaluates.
  float stage[4];
  int32_t stage_n[4];
  const double *stage_t[4];
  const double *stage_lv[4];
};

// Per-block partial results of one step, reduced in fixed order by the finalize
// kernel (deterministic; no float atomics).
struct Partials {
  float *speed_max;   // [nblk]
  double *vol;        // [nblk] sum f64(h_new)
  double *infil;      // [nblk] sum infiltration depth removed
  double *edge;       // [nblk] sum edge drain / banked depth (signed)
};

// What the batched finalize needs of one member.
// The size of a batch's current membership, read by the batched kernels from device
// memory: a captured graph launches grids sized for the batch's capacity, the blocks
// past these counts return at once, and a membership change only rewrites tables.
// The flat tile space of a batch has three segments, each launched with its own kernel:
// 0 members without the bedload integral on the batch's plain variant, 1 members with
// it, 2 wide members without it on the row-block variant.
constexpr int kBatchSegments = 3;
struct BatchDims {
  int32_t seg_first[kBatchSegments];  // the segment's first flat tile
  int32_t seg_tiles[kBatchSegments];  // its live tile count
  int32_t members;         // finalize blocks
  int32_t inflow_members;  // members with point sources
};

// What one member's point-source step needs (device-driven or host-driven).
struct InflowArgs {
  float *h0, *h1;  // the LI ping-pong pair (the same buffer twice for HLLC)
  int64_t pitch;
  int count;
  const int *rows, *cols, *off;
  const double *kt, *kq;
  float *comp;
  double *vol;
  DevCtl *ctl;
  double area;
};

struct FinArgs {
  DevCtl *ctl;
  Partials part;
  const double *sync;
  double area, n_cells, dx;
  int32_t nblk;
  int32_t pad;
};

struct StageCurve {
  int32_t n = 0;
  double *t = nullptr;      // host copies; evaluated per step on the host or device
  double *level = nullptr;
  double *d_t = nullptr;    // device copies for device-driven runs
  double *d_level = nullptr;
};

struct Inflows {
  int32_t count = 0;
  int32_t *d_rows = nullptr, *d_cols = nullptr, *d_off = nullptr;
  double *d_kt = nullptr, *d_kq = nullptr;
  float *d_comp = nullptr;      // per-entry Kahan debt (inflow.py:145-151)
  double *d_vol = nullptr;      // [count] requested volume this step (f64 of f32 depth)
  int32_t n_knots = 0;
};

}  // namespace rbs

// The opaque handle.
struct rbs_state {
  int32_t ny = 0, nx = 0;
  double dx = 0.0;
  int64_t pitch = 0;         // elements per row
  uint32_t armed = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t bc[4] = {0, 0, 0, 0};
  int32_t li_variant = 0;
  int32_t hllc_variant = 0;
  // LI ping-pong: index `cur` holds the current state
  int cur = 0;
  float *h[2] = {nullptr, nullptr};
  float *qx[2] = {nullptr, nullptr};
  float *qy[2] = {nullptr, nullptr};
  float *z = nullptr, *n = nullptr, *eta = nullptr;
  // optional
  float *h_comp = nullptr, *rain = nullptr, *infil = nullptr;
  float *hu = nullptr, *hv = nullptr, *h1 = nullptr, *hu1 = nullptr, *hv1 = nullptr;
  float *fh_x = nullptr, *fh_y = nullptr;   // HLLC limited boundary mass fluxes (scratch)
  float *d50 = nullptr, *qs_int_x = nullptr, *qs_comp_x = nullptr, *qs_int_y = nullptr,
        *qs_comp_y = nullptr, *z0 = nullptr, *dz_lo = nullptr, *dz_hi = nullptr;
  double *dz_cum = nullptr, *dz_unapplied = nullptr;
  double *sed_part = nullptr;  // bed-change totals: 5 per block + the 5 results
  // sub-grid channels: geometry, and the face components (ping-pong like qx / qy)
  float *ch_w = nullptr, *ch_d = nullptr, *ch_n = nullptr;
  float *qx_ch[2] = {nullptr, nullptr}, *qx_fp[2] = {nullptr, nullptr};
  float *qy_ch[2] = {nullptr, nullptr}, *qy_fp[2] = {nullptr, nullptr};
  int32_t chan_variant = 0;
  // dense device staging for field transfers (one block of rows at a time)
  void *xfer = nullptr;
  size_t xfer_cap = 0;
  rbs_release *resv_rec = nullptr;  // reservoir releases: record, scratch, block partials
  // reductions
  rbs::Partials part{};
  int32_t part_cap = 0;
  // the first level of a two-level fold of many block partials (kFoldBlocks entries)
  rbs::Partials fold{};
  // per LI tile: which buffers and compensations hold only +0 (the strip kernels
  // skip dry tiles with it); part_cap words, cleared by quiet_reset
  uint32_t *quiet = nullptr;
  // per HLLC block: bit 0 when stage 1 left the block's U1 cells all +0, bit 1 when
  // stage 2 left their U^n momentum all +0 (the strips skip dry blocks with it)
  uint32_t *hquiet = nullptr;
  rbs::DevCtl *ctl = nullptr;        // device
  rbs::DevCtl *ctl_host = nullptr;   // pinned host mirror for status reads
  double *d_sync = nullptr;
  int32_t sync_cap = 0;
  rbs::StageCurve stage[4];
  rbs::Inflows inflow;
  // captured device-driven step graphs, one per starting ping-pong parity
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_key = -1;     // scheme * 2 + limit of the captured graphs
  int graph_chunk = 16;   // steps per graph (even: a full graph keeps the parity)
  int speed_scheme = -1;  // scheme whose CFL speed ctl->speed_max holds (-1: stale)
  uint64_t bytes = 0;
  double rain_field_sum = 0.0;
};

namespace rbs {

inline size_t cells_bytes(const rbs_state *s) { return (size_t)s->ny * s->pitch * sizeof(float); }
inline size_t xf_rows(const rbs_state *s) { return (size_t)s->ny; }
inline size_t yf_rows(const rbs_state *s) { return (size_t)s->ny + 1; }

// Launch helpers implemented in the kernel translation units.
// `ev_start` / `ev_stop`, when given, are recorded by the kernel's own dispatch (its
// start and end), not by marker packets around it.
int li_launch_step(rbs_state *s, bool device_scalars, double dt, double rain, int limit,
                   double rain_scale, hipEvent_t ev_start = nullptr,
                   hipEvent_t ev_stop = nullptr);
int reduce_speed_launch(rbs_state *s, int scheme, int mode);
int reduce_partials_needed(const rbs_state *s);
int finalize_launch(rbs_state *s, int mode, int scheme);
struct LiArgs;
int set_scalars_launch(rbs_state *s, double t, double dt, float rain, float rain_scale,
                       const float stage[4] = nullptr);
int hllc_launch_step(rbs_state *s, bool device_scalars, double dt, double rain,
                     double rain_scale, double t);
int inflow_launch(rbs_state *s, bool device_scalars, double t, double dt);
int exner_launch(rbs_state *s, double porosity);
int li_num_blocks(const rbs_state *s);
int hllc_num_blocks(const rbs_state *s);
int hllc_variant_count();
int hllc_default_variant();
void hllc_variant_shape(int v, int *bx, int *by);
int li_variant_count();
int li_default_variant();
int li_wide_batch_variant();  // the row blocks a batch gives its wide plain members
int li_default_batch_variant();  // a batch's members without the bedload integral
int li_wide_batch_min_nx();
int li_initial_variant(int nx);
int chan_variant_count();
int chan_default_variant();
void chan_variant_shape(int v, int *bx, int *by);
int chan_num_blocks(const rbs_state *s);
int li_default_sed_variant();
void li_variant_shape(int v, int *bx, int *by, int *family);
int ensure_partials(rbs_state *s, int nblk);
// Clear the dry-tile words (stream-ordered on `stream`): every tile computes its next
// two steps.  Needed whenever h, qx, qy or a Kahan debt is written outside the LI step,
// or the tiling changes.
int quiet_reset(rbs_state *s, hipStream_t stream);

int driven_setup(rbs_state *s, const rbs_run_params *p, double t_stop, int64_t max_steps,
                 int64_t *steps0);
int batch_finalize_launch(const FinArgs *fin, const BatchDims *dims, int grid, int32_t *status,
                          hipStream_t stream);
void inflow_fill_args(const rbs_state *s, InflowArgs *out);
int inflow_batch_launch(const InflowArgs *list, const BatchDims *dims, int grid,
                        hipStream_t stream);
int li_batch_launch(int variant, int seg, const LiArgs *members, const int *tile_member,
                    const int *tile_first, const BatchDims *dims, int grid, hipStream_t stream);
size_t li_args_size();
void li_fill_args(const rbs_state *s, int limit, LiArgs *out);

// finalize modes
enum { FIN_PRIME = 0, FIN_EXPLICIT = 1, FIN_DRIVEN = 2, FIN_PRIME_DRIVEN = 3 };

}  // namespace rbs
