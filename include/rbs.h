/*
 * rbs.h -- C-ABI of librbs.so, the MI355X (gfx950) solver time-step library.
 *
 * The reference (BoykoNeov/river-basin-sim) has no native boundary: its solver is
 * Python driving Warp kernels through the scheme seam
 *     compute_dt(state, alpha, dt_max) -> float
 *     step(state, dt, rain, rain_scale, t)            (solver/core/schemes.py:1-52)
 * over a `State` of device arrays (solver/core/state.py:38-226).  This header is
 * the C boundary a Python ctypes binding (or any FFI) binds in its place.  Each
 * declaration names the reference interface it replaces.
 *
 * Conventions
 *   - Every call returns int: 0 = ok, < 0 = error; rbs_last_error() gives the text
 *     (thread-local).  RBS_E_HIP codes carry the HIP error in the message.
 *   - Plain pointers and sizes only.  Host buffers are borrowed for the duration of
 *     the call.  Device memory is owned by the library.
 *   - A state handle is not thread-safe; distinct handles may be driven from
 *     distinct host threads.  All work of one handle is ordered on its own stream.
 *   - Fields cross the boundary dense and row-major (y, x) in the reference's
 *     layout (solver/core/grid.py:9-23): cells (ny, nx), x-faces (ny, nx+1),
 *     y-faces (ny+1, nx).  The device layout is private (row-padded).
 */
#ifndef RBS_H
#define RBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
#define RBS_OK 0
#define RBS_E_ARG -1      /* bad argument (shape, size, id, order)            */
#define RBS_E_STATE -2    /* call not valid in this state (e.g. field unarmed) */
#define RBS_E_HIP -3      /* a HIP runtime call failed                         */
#define RBS_E_NOMEM -4    /* device allocation failed / over budget            */
#define RBS_E_RANGE -5    /* non-positive or non-finite dt, bad time, overflow */

/* ---- field ids (solver/core/state.py:38-96, sediment.py:548-571) ----------- */
enum rbs_field {
  RBS_F_H = 0,        /* (ny, nx)   f32  depth                                */
  RBS_F_Z = 1,        /* (ny, nx)   f32  bed                                  */
  RBS_F_N = 2,        /* (ny, nx)   f32  Manning n                            */
  RBS_F_QX = 3,       /* (ny, nx+1) f32  LI x-face discharge                  */
  RBS_F_QY = 4,       /* (ny+1, nx) f32  LI y-face discharge                  */
  RBS_F_HU = 5,       /* (ny, nx)   f32  HLLC x-momentum                      */
  RBS_F_HV = 6,       /* (ny, nx)   f32  HLLC y-momentum                      */
  RBS_F_RAIN = 7,     /* (ny, nx)   f32  rain field, m/s                      */
  RBS_F_INFIL = 8,    /* (ny, nx)   f32  infiltration rate, m/s               */
  RBS_F_H_COMP = 9,   /* (ny, nx)   f32  Kahan debt of areal sources          */
  RBS_F_D50 = 10,     /* (ny, nx)   f32  grain size, m                        */
  RBS_F_QS_INT_X = 11,  /* (ny, nx+1) f32 transport integral                  */
  RBS_F_QS_COMP_X = 12, /* (ny, nx+1) f32 its Kahan debt                      */
  RBS_F_QS_INT_Y = 13,  /* (ny+1, nx) f32                                     */
  RBS_F_QS_COMP_Y = 14, /* (ny+1, nx) f32                                     */
  RBS_F_Z0 = 15,      /* (ny, nx)   f32  pristine bed the Exner rebuild uses  */
  RBS_F_DZ_CUM = 16,  /* (ny, nx)   f64  cumulative bed change                */
  RBS_F_DZ_UNAPPLIED = 17, /* (ny, nx) f64 bed change the bounds refused      */
  RBS_F_DZ_LO = 18,   /* (ny, nx)   f32  lower bound on cumulative change     */
  RBS_F_DZ_HI = 19,   /* (ny, nx)   f32  upper bound                          */
  RBS_F_ETA = 20,     /* (ny, nx)   f32  water surface (read-only): h + z, or
                         the storage curve when channels are armed  */
  /* sub-grid channels (channels.py:178-199) */
  RBS_F_CH_W = 21,    /* (ny, nx)   f32  channel width, m (0: no channel)    */
  RBS_F_CH_D = 22,    /* (ny, nx)   f32  bank-full depth below the bed, m    */
  RBS_F_CH_N = 23,    /* (ny, nx)   f32  channel Manning n                   */
  RBS_F_QX_CH = 24,   /* (ny, nx+1) f32  x-face discharge per channel width  */
  RBS_F_QX_FP = 25,   /* (ny, nx+1) f32  x-face discharge per floodplain width */
  RBS_F_QY_CH = 26,   /* (ny+1, nx) f32                                      */
  RBS_F_QY_FP = 27,   /* (ny+1, nx) f32                                      */
  RBS_F_COUNT = 28
};

/* ---- arming flags (state.py:98-171, hllc.py:123-152, sediment.py:658-709) -- */
#define RBS_ARM_COMP (1u << 0)   /* h_comp: Kahan areal sources             */
#define RBS_ARM_RAIN (1u << 1)   /* rain field                              */
#define RBS_ARM_INFIL (1u << 2)  /* infiltration field                      */
#define RBS_ARM_HLLC (1u << 3)   /* momentum hu, hv + HLLC stage buffers    */
#define RBS_ARM_SED (1u << 4)    /* d50, transport integrals, dz_cum ...    */
#define RBS_ARM_CHAN (1u << 5)   /* sub-grid channel geometry + split faces */

/* ---- edges and edge types (state.py:81-96) ------------------------------- */
#define RBS_EDGE_WEST 0
#define RBS_EDGE_EAST 1
#define RBS_EDGE_NORTH 2
#define RBS_EDGE_SOUTH 3
#define RBS_BC_CLOSED 0
#define RBS_BC_OPEN 1
#define RBS_BC_FIXED_STAGE 2

/* ---- schemes (schemes.py:38-52) ------------------------------------------- */
#define RBS_SCHEME_LI 0
#define RBS_SCHEME_HLLC 1

typedef struct rbs_state rbs_state;

/* Device ledger (volumes in m^3; float64; each *_c is its Kahan compensation).
 * Replaces the host float64 sums of MassLedger (massbalance.py:65-158) and the
 * per-cell loss_cum reduction State.loss_volume (state.py:173-177). */
typedef struct rbs_ledger {
  double volume;            /* dx^2 * sum f64(h) of the current state          */
  double rain_in, rain_in_c;     /* rain volume (device-driven runs only)      */
  double point_in, point_in_c;   /* point-source volume (device injection)     */
  double infil_out, infil_out_c; /* infiltration sink                          */
  double edge_out, edge_out_c;   /* open-edge drain (LI) / banked edge flux (HLLC) */
  double t;                 /* device simulated clock                          */
  double dt;                /* dt of the next device-driven step               */
  int64_t steps;            /* steps taken                                     */
  float speed_max;          /* LI: max h; HLLC: max |u|,|v| + sqrt(g h)        */
  int32_t active;           /* device-driven run: 1 while t < t_stop           */
  double release_out, release_out_c; /* reservoir delivery rounding, banked as a sink */
} rbs_ledger;

/* Forcing a device-driven run evaluates per step (run.py:351-388). */
typedef struct rbs_run_params {
  int32_t scheme;           /* RBS_SCHEME_*                                     */
  int32_t limit;            /* LI donor limiter (local_inertial.py:551)         */
  double alpha;             /* [run] cfl                                        */
  double dt_max;            /* [run] dt_max                                     */
  double rain_m_s;          /* uniform rain rate (ignored with a rain field)    */
  double rain_duration;     /* raining while t < rain_duration - EPS_T          */
  int32_t rain_is_field;    /* 1: rain field scaled 1/0 by the raining flag     */
  int32_t n_sync;           /* number of sync times                             */
  const double *sync_times; /* sorted ascending: outputs, events, end_time ... */
  double rain_field_sum;    /* sum f64(rain field) in m/s (run.py:276)          */
} rbs_run_params;

/* ---- lifecycle ---------------------------------------------------------------- */
/* Last error text of the calling thread ("" if none). */
const char *rbs_last_error(void);
/* Library version string. */
const char *rbs_version(void);
/* Number of visible devices; selects nothing. */
int rbs_device_count(int32_t *count);

/* State.from_bed (state.py:179-226): allocates h, z, n, qx, qy (zeroed) on the
 * current device with a private stream.  `arm` optionally arms fields at once. */
int rbs_state_create(int32_t ny, int32_t nx, double dx, uint32_t arm, rbs_state **out);
int rbs_state_destroy(rbs_state *st);
/* State.arm_source_compensation / set_infiltration / set_rain_field / arm_hllc /
 * arm_sediment / arm_channels (state.py:98-171, hllc.py:123, sediment.py:658,
 * channels.py:289-330).  Idempotent per flag except RBS_ARM_SED, which refuses a
 * second arming (sediment.py:680-684).  RBS_ARM_CHAN allocates the geometry (width
 * and depth 0, channel n a copy of n) and the face components (0); with it armed the
 * local-inertial step, its CFL reduction and the bedload integral take their
 * sub-grid forms (local_inertial.py:617-680, :487-500, sediment.py:276-377).  The
 * HLLC scheme ignores channels, as the reference's does. */
int rbs_arm(rbs_state *st, uint32_t arm);
/* Armed flags of the state. */
int rbs_armed(const rbs_state *st, uint32_t *arm);
/* Device bytes held by the state (footprint for admission control). */
int rbs_state_bytes(const rbs_state *st, uint64_t *bytes);
/* Device bytes a state of ny x nx armed with `arm` will hold, without creating it (the
 * admission check of a server): *fields is what rbs_state_create and rbs_arm allocate
 * for the fields, which rbs_state_bytes reports for a fresh state; *total adds the
 * block partials and 1 MiB for the small allocations a run adds later.  No reference
 * counterpart (run.py:69-91 estimates the same from field sizes). */
int rbs_state_footprint(int32_t ny, int32_t nx, uint32_t arm, uint64_t *fields,
                        uint64_t *total);

/* wp.array(...) / .assign(...) and .numpy() on State fields (state.py:200-226).
 * `bytes` must equal the field's dense size.  Uploading h, z or hu/hv marks the
 * cached CFL reduction stale; uploading h, qx, qy or a Kahan debt clears the strip
 * kernels' dry-tile flags (every tile computes its next steps). */
int rbs_upload(rbs_state *st, int32_t field, const void *host, size_t bytes);
int rbs_download(rbs_state *st, int32_t field, void *host, size_t bytes);

/* State.set_open_boundaries (state.py:113-143): edge types indexed by RBS_EDGE_*. */
int rbs_set_boundaries(rbs_state *st, const int32_t edge_types[4]);
/* stage curve of one fixed_stage edge, piecewise linear (boundaries.py:49-70). */
int rbs_set_stage_curve(rbs_state *st, int32_t edge, int32_t n, const double *t,
                        const double *level);

/* ---- the scheme seam (schemes.py:1-52) ------------------------------------------ */
/* local_inertial.compute_dt (local_inertial.py:515-544) and hllc.compute_dt
 * (hllc.py:972-990): one device max-reduction, dt formula in float64. */
int rbs_compute_dt(rbs_state *st, int32_t scheme, double alpha, double dt_max,
                   double *dt_out);
/* local_inertial.step (local_inertial.py:547-727) as one fused kernel. */
int rbs_li_step(rbs_state *st, double dt, double rain, int32_t limit, double rain_scale,
                double t);
/* hllc.step (hllc.py:1272-1373): two fused SSP-RK2 stage kernels. */
int rbs_hllc_step(rbs_state *st, double dt, double rain, double rain_scale, double t);

/* ---- point sources (processes/inflow.py:116-200) ----------------------------------- */
/* InflowInjector.__init__: sources at (rows[k], cols[k]) with piecewise-linear
 * hydrographs knots[offsets[k] .. offsets[k+1]) of (t, Q).  Duplicate cells are
 * rejected (inflow.py:127-139). */
int rbs_set_inflows(rbs_state *st, int32_t count, const int32_t *rows, const int32_t *cols,
                    const int32_t *offsets, const double *knot_t, const double *knot_q);
/* InflowInjector.apply: midpoint discharge over [t, t+dt], compensated add; writes
 * the float32-requested volume (m^3) the ledger banks. */
int rbs_inflow_apply(rbs_state *st, double t, double dt, double *volume_out);

/* ---- morphology (processes/morphology.py:223-307, sediment.py:383-441) ------------- */
/* exner_update + rebuild_bed + refresh eta + clear the transport integral (its
 * Kahan debt kept).  Bounds are RBS_F_DZ_LO / RBS_F_DZ_HI (default -inf/+inf). */
int rbs_exner_activate(rbs_state *st, double porosity);
/* The totals MorphologyProcess and SedimentLedger read from the bed-change fields
 * (morphology.py:286-307, massbalance.py:193-306), reduced on the device in a fixed
 * order in float64: totals[0..4] = {sum dz_cum, sum |dz_cum|, min dz_cum, max dz_cum,
 * sum dz_unapplied}. */
int rbs_sed_totals(rbs_state *st, double *totals);

/* ---- output frames (run.py:389-400, state.py:232-256, zarr_writer.py:117-154) ------ */
/* Fields of a frame: depth h, cell-centred u and v (0 where h < H_DRY; from the face
 * discharges, or the momentum of an HLLC state), and the float64 bed change narrowed
 * to float32.  u and v come together. */
#define RBS_FRAME_DEPTH (1u << 0)
#define RBS_FRAME_U (1u << 1)
#define RBS_FRAME_V (1u << 2)
#define RBS_FRAME_BED_CHANGE (1u << 3)
typedef struct rbs_frames rbs_frames;
/* A queue of `slots` frames of `st`: a device slot and a pinned host slot each, and a
 * copy stream of its own.  Destroy it before its state.  Captures come from one host
 * thread; waits and releases may come from another (a writer thread). */
int rbs_frames_create(rbs_state *st, uint32_t fields, int32_t slots, rbs_frames **out);
int rbs_frames_destroy(rbs_frames *q);
/* Enqueue a frame of the state as it stands after the work already enqueued on the
 * state's stream (e.g. when rbs_run_until returns): one kernel on that stream forms
 * the fields and the clock into a free slot, then the queue's copy stream moves them
 * to the host while the state's stream goes on.  RBS_E_STATE when every slot is held. */
int rbs_frames_capture(rbs_frames *q, int32_t *slot);
/* *ready = 1 once the slot's frame is on the host. */
int rbs_frames_ready(rbs_frames *q, int32_t slot, int32_t *ready);
/* Block until the slot's frame is on the host; the pointers (any may be NULL) receive
 * the frame's clock and its dense (ny, nx) float32 planes in pinned memory (NULL for
 * fields the queue does not carry), valid until the slot is released. */
int rbs_frames_wait(rbs_frames *q, int32_t slot, double *t, float **depth, float **u, float **v,
                    float **bed_change);
/* Hand the slot back for reuse. */
int rbs_frames_release(rbs_frames *q, int32_t slot);
/* Device bytes held by the queue (pinned host bytes are the dense equivalent). */
int rbs_frames_bytes(const rbs_frames *q, uint64_t *bytes);

/* ---- reservoirs (processes/reservoir.py:60-275, io/config.py:147-255) ------------- */
/* One dam's release rule and geometry.  Boxes are inclusive (row0, col0, row1, col1);
 * the gauge is the pool's lowest-bed cell, fixed when the operator is built. */
#define RBS_RULE_FIXED 1          /* a constant release_m3_s                         */
#define RBS_RULE_TARGET_STAGE 2   /* 0 at target_stage_m rising to release_max_m3_s at the crest */
typedef struct rbs_reservoir {
  int32_t pool[4];
  int32_t outlet[4];
  int32_t gauge_row, gauge_col;
  int32_t rule;
  int32_t pad;
  double release_m3_s;
  double target_stage_m;
  double crest_m;
  double release_max_m3_s;
} rbs_reservoir;
/* One activation (reservoir.py:133-147); stage is NaN when the gauge cell is dry. */
typedef struct rbs_release {
  double time, stage, discharge_m3_s, volume_m3;
} rbs_release;
/* ReservoirOperator.advance (reservoir.py:219-270) as one kernel on the state's stream:
 * read the gauge stage, evaluate the rule, move min(Q dt_slow, pool volume) from the
 * pool (in proportion to depth, each cell's exact float32 change summed in float64)
 * to the outlet box (evenly, the float32 rounding of each cell banked into the
 * ledger's release_out).  Fills *out; the state's CFL reduction is marked stale. */
int rbs_reservoir_release(rbs_state *st, const rbs_reservoir *r, double t, double dt_slow,
                          rbs_release *out);
/* rbs_reservoir_release with no device allocation, for callers that cannot allocate
 * (code under graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same kernels and results; the record, the hand-off
 * between launches and the block partials live in a 32 KiB device array the library
 * holds on each GPU until the process ends (a release needs 56 bytes plus 8 per pool row,
 * at most 2048 rows' worth, and RBS_E_RANGE is returned beyond 32 KiB).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (several host threads) race. */
int rbs_reservoir_release_noalloc(rbs_state *st, const rbs_reservoir *r, double t,
                                  double dt_slow, rbs_release *out);

/* ---- ledger and device-driven runs -------------------------------------------------- */
/* MassLedger inputs: device volume, sink totals, clock (massbalance.py:109-135). */
int rbs_ledger_read(rbs_state *st, rbs_ledger *out);
/* Reset the device clock to t and clear the device ledger's flow totals. */
int rbs_clock_reset(rbs_state *st, double t);
/* MultiRateScheduler.ticks + the run loop's per-tick forcing (scheduler.py:171-246,
 * run.py:369-388) on the device: steps are chosen by the device dt reduction and the
 * device fill_span against `p->sync_times`, until the device clock reaches t_stop
 * (a sync time) or max_steps were taken.  Steps are replayed from a captured HIP
 * graph; the host does not see individual steps.  *steps_taken may be NULL. */
int rbs_run_until(rbs_state *st, const rbs_run_params *p, double t_stop, int64_t max_steps,
                  int64_t *steps_taken);
/* Block until the state's stream is idle. */
int rbs_sync(rbs_state *st);

/* ---- batched ensemble (no reference counterpart: the reference runs one scenario per
 * process, run.py:140) ------------------------------------------------------------ */
typedef struct rbs_batch rbs_batch;
int rbs_batch_create(int32_t capacity, rbs_batch **out);
int rbs_batch_destroy(rbs_batch *b);
/* Add a state to the batch with its run parameters; returns its slot. */
int rbs_batch_add(rbs_batch *b, rbs_state *st, const rbs_run_params *p, double t_stop,
                  int32_t *slot);
/* Remove a finished (or any) state from the batch. */
int rbs_batch_remove(rbs_batch *b, int32_t slot);
/* Advance every member by up to `steps` device-driven steps in batched launches;
 * members whose clock reached their t_stop idle.  n_active receives the number of
 * members still running. */
int rbs_batch_advance(rbs_batch *b, int32_t steps, int32_t *n_active);
/* The same advance split in two, so the host can prepare the next requests while the
 * device steps: rbs_batch_launch enqueues the steps and returns; rbs_batch_poll sets
 * *done once they finished (never blocks); rbs_batch_collect waits for them and
 * updates member status.  While an advance is in flight, add / remove / set_variant
 * return RBS_E_STATE. */
int rbs_batch_launch(rbs_batch *b, int32_t steps);
int rbs_batch_poll(rbs_batch *b, int32_t *done);
int rbs_batch_collect(rbs_batch *b, int32_t *n_active);
/* Per-member status: 1 = still running, 0 = reached t_stop. */
int rbs_batch_member_active(rbs_batch *b, int32_t slot, int32_t *active);

/* ---- measurement and tuning (no reference counterpart) ------------------------------------- */
/* Run `reps` host-driven LI steps of `dt` (as rbs_li_step with limit on), timing each
 * fused step kernel with HIP events on the state's own stream, stamped by the kernel's
 * dispatch (hipExtLaunchKernelGGL's start and stop events); *avg_ms receives the mean
 * kernel duration.  Advances the state. */
int rbs_time_li_kernel(rbs_state *st, int32_t reps, double dt, double rain, double rain_scale,
                       double *avg_ms);
/* Select the LI kernel variant, 0 <= variant < rbs_li_variant_count() (a new state
 * starts on the library's default, or on the 49-row row blocks when it is at least
 * 2048 cells wide).  Returns RBS_E_ARG for an unknown variant.
 * Depths and faces are bitwise identical across variants (the ledger's volume sum
 * may differ in its last bits: its partials follow the tile shape). */
int rbs_set_li_variant(rbs_state *st, int32_t variant);
/* The LI kernel variant a state runs (its initial one, or the last set). */
int rbs_get_li_variant(const rbs_state *st, int32_t *variant);
int rbs_li_variant_count(int32_t *count);
/* Shape of a variant: the block's tile of cells (bx x by) and its family
 * (0 = LDS tile, 1 = register strips with a column per lane, 3 = register strips side
 * by side in row blocks whose stores are whole 128-B lines; 2 is not used); *is_default
 * = 1 for the library default. */
int rbs_li_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *family,
                        int32_t *is_default);
/* Tile variants of the LI step of a state armed with sub-grid channels (the LDS
 * kernel that carries the two face components), as for LI; rbs_time_li_kernel times
 * whichever applies to the state. */
int rbs_set_chan_variant(rbs_state *st, int32_t variant);
int rbs_chan_variant_count(int32_t *count);
int rbs_chan_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *is_default);
/* HLLC stage-kernel variants, as for LI: LDS tiles (bx x by cells per block) and
 * register strips (bx = 58 columns per wave times the waves of a block, by rows per
 * strip; the default).  Every variant writes the same bits. */
int rbs_set_hllc_variant(rbs_state *st, int32_t variant);
int rbs_hllc_variant_count(int32_t *count);
int rbs_hllc_variant_info(int32_t variant, int32_t *bx, int32_t *by, int32_t *is_default);
/* Run `reps` host-driven HLLC steps of `dt` without sources, timing each pair of stage
 * kernels with HIP events on the state's stream; *avg_ms receives the mean.  Advances
 * the state. */
int rbs_time_hllc_step(rbs_state *st, int32_t reps, double dt, double *avg_ms);
/* The LI kernel variants a batch launches: `variant` for members without the bedload
 * integral, `variant_sed` for members with it (unless set: the 49-row row blocks and
 * the 8-row strips); takes effect at the next advance.  Members without the bedload
 * integral that are at least 2048 cells wide run the 49-row row blocks whatever
 * `variant` says, as they would alone; the three groups launch side by side in the
 * batch's graph. */
int rbs_batch_set_variant(rbs_batch *b, int32_t variant, int32_t variant_sed);
int rbs_batch_get_variant(rbs_batch *b, int32_t *variant, int32_t *variant_sed);
/* Host-side cost of membership changes so far (table rebuilds and graph captures:
 * count and wall milliseconds), graph replays launched, and the device time of the
 * batched steps (HIP events around each advance). */
int rbs_batch_stats(rbs_batch *b, int64_t *rebuilds, double *rebuild_ms,
                    int64_t *graph_launches, double *gpu_ms);

#ifdef __cplusplus
}
#endif
#endif
