/*
 * rbs_oracle.h -- CPU restatement of river-basin-sim's solver time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every pass follows one Warp kernel of the reference (cited per function in
 * rbs_oracle.c) in the same float32 operation order, with the same in-place
 * update sequence, built with -ffp-contract=off so no FMA is formed.  Float64
 * appears exactly where the reference uses it: the loss ledger, the Exner
 * cumulative bed change, and the host-side dt formula.
 *
 * Fields are dense row-major (y, x) exactly as the reference lays them out
 * (grid.py:9-23): cells (ny, nx), x-faces (ny, nx+1), y-faces (ny+1, nx).
 */
#ifndef RBS_ORACLE_H
#define RBS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge indices used for every per-edge table in the oracle. */
enum { ORC_WEST = 0, ORC_EAST = 1, ORC_NORTH = 2, ORC_SOUTH = 3 };
/* Edge types (state.py:84-92). */
enum { ORC_CLOSED = 0, ORC_OPEN = 1, ORC_FIXED_STAGE = 2 };

typedef struct orc_state {
  int32_t ny, nx;
  double dx;           /* Python float dx; kernels see (float)dx */
  /* always present (state.py:179-226) */
  float *h, *z, *n, *qx, *qy, *eta, *beta;
  /* optional, NULL when unarmed (state.py:70-80) */
  float *h_comp;       /* (ny, nx) Kahan debt of areal sources */
  float *rain;         /* (ny, nx) rain field, m/s */
  float *infil;        /* (ny, nx) infiltration rate, m/s */
  double *loss_cum;    /* (ny, nx) banked sink depth, m */
  int32_t bc[4];       /* edge types, indexed by ORC_WEST.. */
  /* HLLC momentum + scratch (hllc.py:96-152); NULL unless armed */
  float *hu, *hv;
  float *uvel, *vvel;
  float *fx_h, *fx_mn_l, *fx_mn_r, *fx_mt;   /* (ny, nx+1) */
  float *fy_h, *fy_mn_l, *fy_mn_r, *fy_mt;   /* (ny+1, nx) */
  float *dh, *dhu, *dhv, *h1, *hu1, *hv1;    /* (ny, nx) */
  /* sediment (sediment.py:548-571); NULL unless armed */
  float *d50;
  float *qs_int_x, *qs_comp_x;   /* (ny, nx+1) */
  float *qs_int_y, *qs_comp_y;   /* (ny+1, nx) */
  /* sub-grid channels (channels.py:178-199); NULL unless armed */
  float *ch_w, *ch_d, *ch_n;      /* (ny, nx) width, bank-full depth, Manning n */
  float *qx_ch, *qx_fp;           /* (ny, nx+1) channel / floodplain components */
  float *qy_ch, *qy_fp;           /* (ny+1, nx) */
} orc_state;

/* The powers the solver takes (see rbs_oracle.c): x^float32(7/3), x^float32(4/3),
 * x^float32(1/3), x^(3/2). */
float orc_pow73(float x);
float orc_pow43(float x);
float orc_cbrt(float x);
float orc_pow15(float x);
/* The float32 polynomial ln x inside the powers (test support). */
double orc_ln_coarse(float x);
/* 1 (default): correctly rounded powers, bit-identical to the HIP kernels;
 * 0: the C library's powf, as the reference's Warp-CPU runs evaluated them. */
void orc_set_pow_exact(int on);
int orc_get_pow_exact(void);
/* Test support: how many float32 bit patterns in [lo, hi) give a power (kind 0..3:
 * x^float32(7/3), x^float32(4/3), x^float32(1/3), x^(3/2)) other than powl's with the
 * same exponent, rounded; libm != 0 counts the C library's powf instead. */
long orc_pow_sweep(int kind, uint32_t lo, uint32_t hi, int libm);
/* The first maxn of those bit patterns (mode 1, kind 0..2); returns how many. */
long orc_pow_misses(int kind, uint32_t lo, uint32_t hi, uint32_t *out, long maxn);

/* Threads used by every pass (OpenMP); 1 gives the serial order of Warp-CPU. */
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- local-inertial ------------------------------------------------------ */
float orc_li_reduce_hmax(const orc_state *s);
double orc_li_compute_dt(const orc_state *s, double alpha, double dt_max);
void orc_li_step(orc_state *s, double dt, double rain, int limit, double rain_scale);
/* A 1000-step style driver: n_steps fixed-dt steps (used by the CPU baseline). */
void orc_li_run_fixed(orc_state *s, double dt, int n_steps, double rain, double rain_scale);

/* ---- sub-grid channels (channels.py:59-138) ------------------------------ */
float orc_eta_subgrid(float h, float z, float w, float d, float dx);
float orc_column_depth(float h, float w, float d, float dx);
/* friction.py:35-57 manning_denominator_radius */
float orc_manning_denominator_radius(float q_prev, float h_flow, float radius, float n,
                                     float g, float dt);

/* ---- HLLC ------------------------------------------------------------------ */
float orc_hllc_reduce_wavespeed(const orc_state *s);
double orc_hllc_compute_dt(const orc_state *s, double alpha, double dt_max);
/* stages: prescribed surface per edge (only read for fixed_stage edges) */
void orc_hllc_step(orc_state *s, double dt, double rain, double rain_scale,
                   const float stages[4]);
/* single-face HLLC flux (hllc.py:191-256), for known-answer tests */
void orc_hllc_flux(float hl, float ul, float vtl, float hr, float ur, float vtr,
                   float out[3]);

/* ---- sediment (sediment.py) ---------------------------------------------------- */
float orc_face_capacity(float q, float h, float radius, float n, float d50);
float orc_shields_number(float q, float h, float radius, float n, float d50);
void orc_accumulate_transport(orc_state *s, double dt);
void orc_exner_update(int32_t ny, int32_t nx, const float *qs_int_x,
                      const float *qs_int_y, const float *dz_lo, const float *dz_hi,
                      double dx, double inv_one_minus_p, double *dz_cum,
                      double *dz_unapplied);
void orc_rebuild_bed(int32_t ny, int32_t nx, const float *z0, const double *dz_cum,
                     float *z);

/* ---- point sources (inflow.py:91-113) --------------------------------------- */
void orc_inject_point_sources_c(orc_state *s, int32_t count, const int32_t *rows,
                                const int32_t *cols, const float *add_depth,
                                float *comp);

/* ---- scheduler (scheduler.py:48-72) ---------------------------------------------- */
double orc_fill_span(double span, double dt_raw);

#ifdef __cplusplus
}
#endif
#endif
