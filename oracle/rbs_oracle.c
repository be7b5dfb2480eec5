/*
 * rbs_oracle.c -- CPU restatement of river-basin-sim's solver time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * restates one reference Warp kernel; the citation above it is the reference
 * file:line it follows.  Arithmetic is float32 in the reference's left-to-right
 * expression order; the build uses -ffp-contract=off (no FMA) and no fast-math,
 * so Kahan compensation survives and every rounding happens where Warp's does.
 *
 * Passes run in the reference's launch order and update fields in place exactly
 * as the reference does (one Warp launch == one loop nest here).  OpenMP
 * parallelises the outer loop of a pass; every output element is still written
 * by exactly one iteration, so results are independent of the thread count.
 */
#include "rbs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* grid.py:44, 49 -- the Python constants as a float32 kernel sees them. */
#define G_F32 ((float)9.81)
#define HDRY_F32 ((float)1.0e-3)
#define H_DRY_F64 1.0e-3
#define GRAVITY_F64 9.81

#define IX(i, j, w) ((size_t)(i) * (size_t)(w) + (size_t)(j))

static int g_threads = 1;

void orc_set_threads(int n) {
  g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
  omp_set_num_threads(g_threads);
#endif
}
int orc_get_threads(void) { return g_threads; }

/* Warp's scalar max/min (builtin.h): a > b ? a : b. */
static inline float fmx(float a, float b) { return a > b ? a : b; }
static inline float fmn(float a, float b) { return a < b ? a : b; }
static inline float fclamp(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }

/*
 * Powers of a positive float32 (wp.pow in friction.py:31, :57, sediment.py:168, :192).
 * The reference's exponents are Python floats handed to a float32 pow, so they are
 * float32(7/3), float32(4/3), float32(1/3) and 1.5.  Warp's wp.pow is a libm powf,
 * accurate to about an ulp, whose last bit differs between math libraries.
 *
 * Mode 1 (the default) is the correctly rounded float32 power x^e, e = float32(p/3) =
 * p/3 - d, by a fixed sequence of IEEE basic operations (fused multiply-adds only where
 * written as fmaf/fma, which are correctly rounded), so it is the same bits on every
 * IEEE machine; the HIP kernels evaluate the same sequence:
 *   y ~ x^(-1/3): a bit-pattern seed y0 times the series of (1 - r0)^(-1/3) to r0^3,
 *     r0 = 1 - x y0^3, in float32 (|1 - x y^3| < 8.6e-5); r = 1 - x y^3 exactly in
 *     float64;
 *   x^e = x^q y^2 (1 - r)^(-2/3) x^(-d), q = 3, 2, 1 for p = 7, 4, 1: the binomial
 *     series of (1 - r)^(-2/3) to r^3 and exp(-u) = 1 - u + u^2/2, u = d ln x, merged
 *     into one factor 1 + g; ln x from a float32 polynomial (3e-8 absolute, so u is good
 *     to 2.4e-15 relative);
 *   x^q y^2 (1 + g) in float64, rounded once.
 * Mode 0 calls the C library's powf with the float32 exponents, as the reference's
 * Warp-CPU runs did; its published round-off-level mass residuals reproduce with it.
 * tests/test_oracle_pow.py checks mode 1 against the exact powers and against powl.
 */
static inline float cbrt_seed(float x) {
  uint32_t i;
  memcpy(&i, &x, 4);
  uint32_t j = 0x54a2fa8cu - i / 3u;
  float y;
  memcpy(&y, &j, 4);
  float r = fmaf(-(x * y), y * y, 1.0f);
  float t = fmaf(fmaf(r, (float)(14.0 / 81.0), (float)(2.0 / 9.0)), r, (float)(1.0 / 3.0));
  return fmaf(y * r, t, y);
}
/* ln x, x = 2^k m normal with m in [sqrt(1/2), sqrt(2)), k shifted by kadj: the bit
 * pattern offset by that of sqrt(1/2) splits off k and m; ln(1 + t) = t - t^2/2 +
 * t^3 R(t), t = m - 1 (exact), R of degree 6 fitted to the remainder on that range;
 * float32, then k ln 2 in float64.  3e-8 absolute. */
static inline double ln_coarse(float x, int kadj) {
  uint32_t i;
  memcpy(&i, &x, 4);
  int32_t ix = (int32_t)(i - 0x3f3504f3u);
  int k = (ix >> 23) + kadj;
  uint32_t mb = ((uint32_t)ix & 0x007fffffu) + 0x3f3504f3u;
  float m;
  memcpy(&m, &mb, 4);
  float t = m - 1.0f;
  float r = 0.08672088310927706f;
  r = fmaf(r, t, -0.14379001534630725f);
  r = fmaf(r, t, 0.1497802208269855f);
  r = fmaf(r, t, -0.16564426496452503f);
  r = fmaf(r, t, 0.19954873155337902f);
  r = fmaf(r, t, -0.2500169873952298f);
  r = fmaf(r, t, 0.3333416597960291f);
  float s2 = fmaf(r, t, -0.5f);
  float l = fmaf(t * t, s2, t);
  return fma((double)k, 0.6931471805599453, (double)l);
}
/* x^e, e = float32(p/3), p in {1, 4, 7}, x >= 0; subnormal x is first scaled by 2^48
 * (a power of two divisible by three) and the power scaled back by 2^(-16p). */
static float pow_third(float x, int p) {
  const double d = ((double)p - 3.0 * (double)(float)((double)p / 3.0)) / 3.0;
  float xs = x;
  int kadj = 0, tiny = x < 0x1p-126f;
  if (tiny) {
    xs = x * 0x1p48f;
    kadj = -48;
  }
  double a = (double)xs, b = (double)cbrt_seed(xs);
  double b2 = b * b;                 /* exact */
  double r = fma(-(a * b), b2, 1.0); /* a*b exact */
  double s2 = r * fma(r * (40.0 / 81.0) + 5.0 / 9.0, r, 2.0 / 3.0);
  double u = d * ln_coarse(xs, kadj);
  double g = fma(-u, fma(-0.5, u, s2) + 1.0, s2);
  double A;
  if (p == 7)
    A = ((a * a) * a) * b2;
  else if (p == 4)
    A = (a * a) * b2;
  else
    A = a * b2;
  double pw = fma(A, g, A);
  if (tiny) pw *= p == 7 ? 0x1p-112 : p == 4 ? 0x1p-64 : 0x1p-16;
  return (float)pw;
}
static int g_pow_exact = 1;
void orc_set_pow_exact(int on) { g_pow_exact = on != 0; }
int orc_get_pow_exact(void) { return g_pow_exact; }
float orc_pow73(float x) { /* x^float32(7/3) */
  if (!g_pow_exact) return powf(x, (float)(7.0 / 3.0));
  return pow_third(x, 7);
}
float orc_cbrt(float x) { /* x^float32(1/3) */
  if (!g_pow_exact) return powf(x, (float)(1.0 / 3.0));
  return pow_third(x, 1);
}
float orc_pow43(float x) { /* x^float32(4/3) */
  if (!g_pow_exact) return powf(x, (float)(4.0 / 3.0));
  return pow_third(x, 4);
}
float orc_pow15(float x) { /* x^(3/2) = x * sqrt(x) */
  if (!g_pow_exact) return powf(x, 1.5f);
  double a = (double)x;
  return (float)(a * sqrt(a));
}
double orc_ln_coarse(float x) { return ln_coarse(x, 0); }

/* friction.py:18-31 manning_denominator, 1 + g dt n^2 |q| / h^float32(7/3): the
 * product left to right, one float32 division, one sum (both modes). */
static inline float manning_denominator(float q_prev, float h_flow, float n, float g,
                                        float dt) {
  return 1.0f + g * dt * n * n * fabsf(q_prev) / orc_pow73(h_flow);
}

/* Every float32 in [lo, hi) (bit patterns) checked against powl with the same float32
 * exponent, rounded to float32; returns the number that differ.  kind 0..3:
 * x^float32(7/3), x^float32(4/3), x^float32(1/3), x^(3/2).  libm != 0 checks the C
 * library's powf instead of mode 1.  Test support only. */
long orc_pow_sweep(int kind, uint32_t lo, uint32_t hi, int libm) {
  const float ex[4] = {(float)(7.0 / 3.0), (float)(4.0 / 3.0), (float)(1.0 / 3.0), 1.5f};
  long bad = 0;
  int procs = 1;
#ifdef _OPENMP
  procs = omp_get_num_procs(); /* a test sweep: every processor, whatever the passes use */
#endif
#pragma omp parallel for reduction(+ : bad) schedule(static) num_threads(procs)
  for (long long u = (long long)lo; u < (long long)hi; ++u) {
    uint32_t bits = (uint32_t)u;
    float x;
    memcpy(&x, &bits, 4);
    float want = (float)powl((long double)x, (long double)ex[kind]);
    float got;
    if (libm)
      got = powf(x, ex[kind]);
    else if (kind == 3) {
      double a = (double)x;
      got = (float)(a * sqrt(a));
    } else
      got = pow_third(x, kind == 0 ? 7 : kind == 1 ? 4 : 1);
    bad += got != want;
  }
  return bad;
}

/* The first maxn bit patterns in [lo, hi) that orc_pow_sweep counts (mode 1), written
 * to out in increasing order; returns how many were written.  Test support only. */
long orc_pow_misses(int kind, uint32_t lo, uint32_t hi, uint32_t *out, long maxn) {
  const float ex[3] = {(float)(7.0 / 3.0), (float)(4.0 / 3.0), (float)(1.0 / 3.0)};
  long n = 0;
  for (long long u = (long long)lo; u < (long long)hi && n < maxn; ++u) {
    uint32_t bits = (uint32_t)u;
    float x;
    memcpy(&x, &bits, 4);
    float want = (float)powl((long double)x, (long double)ex[kind]);
    if (pow_third(x, kind == 0 ? 7 : kind == 1 ? 4 : 1) != want) out[n++] = bits;
  }
  return n;
}

/* friction.py:35-57 manning_denominator_radius: D = 1 + g dt n^2 |q| / (h R^float32(4/3)) */
float orc_manning_denominator_radius(float q_prev, float h_flow, float radius, float n,
                                     float g, float dt) {
  return 1.0f + g * dt * n * n * fabsf(q_prev) / (h_flow * orc_pow43(radius));
}

/* channels.py:59-72 eta_subgrid: the storage curve h -> eta */
float orc_eta_subgrid(float h, float z, float w, float d, float dx) {
  float eta = h + z;
  if (w > 0.0f && d > 0.0f) {
    float h_bf = w * d / dx;
    if (h <= h_bf)
      eta = z - d + h * dx / w;
    else
      eta = z + (h - h_bf);
  }
  return eta;
}

/* channels.py:75-85 column_depth: eta - z_ch, the depth that sets the wave speed */
float orc_column_depth(float h, float w, float d, float dx) {
  float col = h;
  if (w > 0.0f && d > 0.0f) {
    float h_bf = w * d / dx;
    if (h <= h_bf)
      col = h * dx / w;
    else
      col = d + (h - h_bf);
  }
  return col;
}

/* local_inertial.py:286-312 _channel_flux, per unit channel width; the face section
 * from channels.py:99-131 (min width, higher invert, bank-full saturation, A/P) */
static inline float channel_flux(float q_prev, float eta_max, float slope, float w_face,
                                 float z_bank, float z_ch, float n_ch, float dt, float g) {
  if (w_face <= 0.0f) return 0.0f;
  float h_ch = fmn(eta_max, z_bank) - z_ch;
  if (h_ch < HDRY_F32) return 0.0f;
  float radius = w_face * h_ch / (w_face + 2.0f * h_ch);
  float num = q_prev - g * h_ch * dt * slope;
  float den = orc_manning_denominator_radius(q_prev, h_ch, radius, n_ch, g, dt);
  return num / den;
}

/* local_inertial.py:315-331 _floodplain_flux: the plain face update on the bank */
static inline float floodplain_flux(float q_prev, float eta_max, float slope, float z_bank,
                                    float n_fp, float dt, float g);

/* sources.py:66-82 kahan_add -> (total, comp) */
static inline void kahan_add(float *total, float *comp, float inc) {
  float y = inc - *comp;
  float t = *total + y;
  *comp = (t - *total) - y;
  *total = t;
}

static inline float floodplain_flux(float q_prev, float eta_max, float slope, float z_bank,
                                    float n_fp, float dt, float g) {
  float h_fp = eta_max - z_bank;
  if (h_fp < HDRY_F32) return 0.0f;
  float num = q_prev - g * h_fp * dt * slope;
  float den = manning_denominator(q_prev, h_fp, n_fp, g, dt);
  return num / den;
}

/* ========================================================================== */
/* local-inertial                                                              */
/* ========================================================================== */

/* local_inertial.py:61-69 compute_eta */
static void li_compute_eta(orc_state *s) {
  const int ny = s->ny, nx = s->nx;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) s->eta[IX(i, j, nx)] = s->h[IX(i, j, nx)] + s->z[IX(i, j, nx)];
}

/* local_inertial.py:274-283 compute_eta_channels */
static void li_compute_eta_channels(orc_state *s) {
  const int ny = s->ny, nx = s->nx;
  const float dx = (float)s->dx;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      const size_t k = IX(i, j, nx);
      s->eta[k] = orc_eta_subgrid(s->h[k], s->z[k], s->ch_w[k], s->ch_d[k], dx);
    }
}

/* local_inertial.py:334-381 update_qx_channels / :384-429 update_qy_channels: one
 * interior face between cells a (left/top) and b (right/bottom); writes both
 * components and the total per unit cell width. */
static inline void channel_face(const orc_state *s, size_t a, size_t b, float dx, float dt,
                                float g, float *q, float *q_ch, float *q_fp) {
  float eta_a = s->eta[a], eta_b = s->eta[b];
  float eta_max = fmx(eta_a, eta_b);
  float slope = (eta_b - eta_a) / dx;
  float z_bank = fmx(s->z[a], s->z[b]);
  float fp = floodplain_flux(*q_fp, eta_max, slope, z_bank, 0.5f * (s->n[a] + s->n[b]), dt, g);
  float w_face = fmn(s->ch_w[a], s->ch_w[b]);
  float z_ch = fmx(s->z[a] - s->ch_d[a], s->z[b] - s->ch_d[b]);
  float ch = channel_flux(*q_ch, eta_max, slope, w_face, z_bank, z_ch,
                          0.5f * (s->ch_n[a] + s->ch_n[b]), dt, g);
  *q_fp = fp;
  *q_ch = ch;
  float frac = w_face / dx;
  *q = fp * (1.0f - frac) + ch * frac;
}

static void li_update_faces_channels(orc_state *s, float dx, float dt, float g) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  if (nx > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
      for (int j = 1; j < nx; ++j) {
        const size_t f = IX(i, j, w);
        channel_face(s, IX(i, j - 1, nx), IX(i, j, nx), dx, dt, g, &s->qx[f], &s->qx_ch[f],
                     &s->qx_fp[f]);
      }
  }
  if (ny > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const size_t f = IX(i, j, nx);
        channel_face(s, IX(i - 1, j, nx), IX(i, j, nx), dx, dt, g, &s->qy[f], &s->qy_ch[f],
                     &s->qy_fp[f]);
      }
  }
}

/* local_inertial.py:432-484 limit_qx_channels / limit_qy_channels: the total's donor
 * factor scales the total and both components (a zero total is left alone). */
static void li_limit_channels(orc_state *s) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  if (nx > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
      for (int j = 1; j < nx; ++j) {
        const size_t f = IX(i, j, w);
        float q = s->qx[f];
        if (q == 0.0f) continue;
        float b = s->beta[IX(i, j - 1, nx)];
        if (q < 0.0f) b = s->beta[IX(i, j, nx)];
        s->qx[f] = q * b;
        s->qx_ch[f] = s->qx_ch[f] * b;
        s->qx_fp[f] = s->qx_fp[f] * b;
      }
  }
  if (ny > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const size_t f = IX(i, j, nx);
        float q = s->qy[f];
        if (q == 0.0f) continue;
        float b = s->beta[IX(i - 1, j, nx)];
        if (q < 0.0f) b = s->beta[IX(i, j, nx)];
        s->qy[f] = q * b;
        s->qy_ch[f] = s->qy_ch[f] * b;
        s->qy_fp[f] = s->qy_fp[f] * b;
      }
  }
}

/* local_inertial.py:50-58 face_h_flow + :72-99 update_qx (interior faces j=1..nx-1) */
static void li_update_qx(orc_state *s, float dx, float dt, float g) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i) {
    for (int j = 1; j < nx; ++j) {
      float eta_l = s->eta[IX(i, j - 1, nx)];
      float eta_r = s->eta[IX(i, j, nx)];
      float h_flow = fmx(eta_l, eta_r) - fmx(s->z[IX(i, j - 1, nx)], s->z[IX(i, j, nx)]);
      if (h_flow < HDRY_F32) {
        s->qx[IX(i, j, w)] = 0.0f;
        continue;
      }
      float q = s->qx[IX(i, j, w)];
      float n_face = 0.5f * (s->n[IX(i, j - 1, nx)] + s->n[IX(i, j, nx)]);
      float slope = (eta_r - eta_l) / dx;
      float num = q - g * h_flow * dt * slope;
      float den = manning_denominator(q, h_flow, n_face, g, dt);
      s->qx[IX(i, j, w)] = num / den;
    }
  }
}

/* local_inertial.py:102-129 update_qy (interior faces i=1..ny-1) */
static void li_update_qy(orc_state *s, float dx, float dt, float g) {
  const int ny = s->ny, nx = s->nx;
#pragma omp parallel for schedule(static)
  for (int i = 1; i < ny; ++i) {
    for (int j = 0; j < nx; ++j) {
      float eta_t = s->eta[IX(i - 1, j, nx)];
      float eta_b = s->eta[IX(i, j, nx)];
      float h_flow = fmx(eta_t, eta_b) - fmx(s->z[IX(i - 1, j, nx)], s->z[IX(i, j, nx)]);
      if (h_flow < HDRY_F32) {
        s->qy[IX(i, j, nx)] = 0.0f;
        continue;
      }
      float q = s->qy[IX(i, j, nx)];
      float n_face = 0.5f * (s->n[IX(i - 1, j, nx)] + s->n[IX(i, j, nx)]);
      float slope = (eta_b - eta_t) / dx;
      float num = q - g * h_flow * dt * slope;
      float den = manning_denominator(q, h_flow, n_face, g, dt);
      s->qy[IX(i, j, nx)] = num / den;
    }
  }
}

/* boundaries.py:73-91 apply_closed_bc (_zero_qx_edges, _zero_qy_edges) */
static void li_closed_bc(orc_state *s) {
  const int ny = s->ny, nx = s->nx;
  for (int i = 0; i < ny; ++i) {
    s->qx[IX(i, 0, nx + 1)] = 0.0f;
    s->qx[IX(i, nx, nx + 1)] = 0.0f;
  }
  for (int j = 0; j < nx; ++j) {
    s->qy[IX(0, j, nx)] = 0.0f;
    s->qy[IX(ny, j, nx)] = 0.0f;
  }
}

/* local_inertial.py:197-239 compute_outflow_beta */
static void li_outflow_beta(orc_state *s, float dx, float dt) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i) {
    for (int j = 0; j < nx; ++j) {
      float q_out = 0.0f;
      float ql = s->qx[IX(i, j, w)];
      if (ql < 0.0f) q_out = q_out - ql;
      float qr = s->qx[IX(i, j + 1, w)];
      if (qr > 0.0f) q_out = q_out + qr;
      float qt = s->qy[IX(i, j, nx)];
      if (qt < 0.0f) q_out = q_out - qt;
      float qb = s->qy[IX(i + 1, j, nx)];
      if (qb > 0.0f) q_out = q_out + qb;
      float out_depth = dt / dx * q_out;
      if (out_depth > 0.0f)
        s->beta[IX(i, j, nx)] = fclamp(s->h[IX(i, j, nx)] / out_depth, 0.0f, 1.0f);
      else
        s->beta[IX(i, j, nx)] = 1.0f;
    }
  }
}

/* local_inertial.py:242-251 limit_qx */
static void li_limit_qx(orc_state *s) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i) {
    for (int j = 1; j < nx; ++j) {
      float q = s->qx[IX(i, j, w)];
      if (q > 0.0f)
        s->qx[IX(i, j, w)] = q * s->beta[IX(i, j - 1, nx)];
      else if (q < 0.0f)
        s->qx[IX(i, j, w)] = q * s->beta[IX(i, j, nx)];
    }
  }
}

/* local_inertial.py:254-263 limit_qy */
static void li_limit_qy(orc_state *s) {
  const int ny = s->ny, nx = s->nx;
#pragma omp parallel for schedule(static)
  for (int i = 1; i < ny; ++i) {
    for (int j = 0; j < nx; ++j) {
      float q = s->qy[IX(i, j, nx)];
      if (q > 0.0f)
        s->qy[IX(i, j, nx)] = q * s->beta[IX(i - 1, j, nx)];
      else if (q < 0.0f)
        s->qy[IX(i, j, nx)] = q * s->beta[IX(i, j, nx)];
    }
  }
}

/* local_inertial.py:132-151 update_h */
static void li_update_h(orc_state *s, float dx, float dt, float rain) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i) {
    for (int j = 0; j < nx; ++j) {
      float net = s->qx[IX(i, j, w)] - s->qx[IX(i, j + 1, w)] + s->qy[IX(i, j, nx)] -
                  s->qy[IX(i + 1, j, nx)];
      s->h[IX(i, j, nx)] = s->h[IX(i, j, nx)] + dt / dx * net + rain * dt;
    }
  }
}

/* local_inertial.py:154-167 apply_rain_field (uncompensated) */
static void apply_rain_field_plain(orc_state *s, float dt, float scale) {
  const size_t cells = (size_t)s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) s->h[k] = s->h[k] + s->rain[k] * scale * dt;
}

/* sources.py:85-101 add_uniform_rain_c */
static void add_uniform_rain_c(orc_state *s, float rate, float dt) {
  const size_t cells = (size_t)s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) kahan_add(&s->h[k], &s->h_comp[k], rate * dt);
}

/* sources.py:104-121 add_rain_field_c */
static void add_rain_field_c(orc_state *s, float dt, float scale) {
  const size_t cells = (size_t)s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) kahan_add(&s->h[k], &s->h_comp[k], s->rain[k] * scale * dt);
}

/* local_inertial.py:170-194 apply_infiltration */
static void apply_infiltration(orc_state *s, float dt) {
  const size_t cells = (size_t)s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) {
    float avail = s->h[k];
    float inf = s->infil[k] * dt;
    if (inf > avail) inf = avail;
    if (inf < 0.0f) inf = 0.0f;
    float h_new = avail - inf;
    s->h[k] = h_new;
    s->loss_cum[k] = s->loss_cum[k] + ((double)avail - (double)h_new);
  }
}

/* boundaries.py:94-106 _drain */
static inline float drain(float h_edge, float q_out, float dxdt) {
  float removed = dxdt * q_out;
  if (removed > h_edge) removed = h_edge;
  if (removed < 0.0f) removed = 0.0f;
  return h_edge - removed;
}

/* boundaries.py:109-208 _open_out_{west,east,north,south} + apply_open_outflow */
static void li_open_outflow(orc_state *s, double dt) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float dxdt = (float)(dt / s->dx);
  if (s->bc[ORC_WEST] == ORC_OPEN && nx > 1)
    for (int i = 0; i < ny; ++i) {
      float avail = s->h[IX(i, 0, nx)];
      float h_new = drain(avail, fmx(-s->qx[IX(i, 1, w)], 0.0f), dxdt);
      s->h[IX(i, 0, nx)] = h_new;
      s->loss_cum[IX(i, 0, nx)] += ((double)avail - (double)h_new);
    }
  if (s->bc[ORC_EAST] == ORC_OPEN && nx > 1)
    for (int i = 0; i < ny; ++i) {
      float avail = s->h[IX(i, nx - 1, nx)];
      float h_new = drain(avail, fmx(s->qx[IX(i, nx - 1, w)], 0.0f), dxdt);
      s->h[IX(i, nx - 1, nx)] = h_new;
      s->loss_cum[IX(i, nx - 1, nx)] += ((double)avail - (double)h_new);
    }
  if (s->bc[ORC_NORTH] == ORC_OPEN && ny > 1)
    for (int j = 0; j < nx; ++j) {
      float avail = s->h[IX(0, j, nx)];
      float h_new = drain(avail, fmx(-s->qy[IX(1, j, nx)], 0.0f), dxdt);
      s->h[IX(0, j, nx)] = h_new;
      s->loss_cum[IX(0, j, nx)] += ((double)avail - (double)h_new);
    }
  if (s->bc[ORC_SOUTH] == ORC_OPEN && ny > 1)
    for (int j = 0; j < nx; ++j) {
      float avail = s->h[IX(ny - 1, j, nx)];
      float h_new = drain(avail, fmx(s->qy[IX(ny - 1, j, nx)], 0.0f), dxdt);
      s->h[IX(ny - 1, j, nx)] = h_new;
      s->loss_cum[IX(ny - 1, j, nx)] += ((double)avail - (double)h_new);
    }
}

/* local_inertial.py:503-512 reduce_hmax (h_max starts at 0 via zero_()) */
float orc_li_reduce_hmax(const orc_state *s) {
  const size_t cells = (size_t)s->ny * s->nx;
  float m = 0.0f;
  if (s->ch_w) { /* local_inertial.py:487-500 reduce_column_max */
    const float dx = (float)s->dx;
    for (size_t k = 0; k < cells; ++k)
      m = fmx(m, orc_column_depth(s->h[k], s->ch_w[k], s->ch_d[k], dx));
    return m;
  }
  for (size_t k = 0; k < cells; ++k) m = fmx(m, s->h[k]);
  return m;
}

/* local_inertial.py:515-544 compute_dt (f64 formula on the host) */
double orc_li_compute_dt(const orc_state *s, double alpha, double dt_max) {
  double h_max = (double)orc_li_reduce_hmax(s);
  if (h_max <= H_DRY_F64) return dt_max;
  double dt = alpha * s->dx / sqrt(GRAVITY_F64 * h_max);
  return dt < dt_max ? dt : dt_max;
}

/* local_inertial.py:547-727 step; sub-grid channels switch the eta, face and
 * limiter passes onto their channel forms (:617-680) */
void orc_li_step(orc_state *s, double dt, double rain, int limit, double rain_scale) {
  const float dxf = (float)s->dx, dtf = (float)dt, gf = G_F32;
  const int chan = s->ch_w != NULL;
  if (chan) {
    li_compute_eta_channels(s);
    li_update_faces_channels(s, dxf, dtf, gf);
  } else {
    li_compute_eta(s);
    if (s->nx > 1) li_update_qx(s, dxf, dtf, gf);
    if (s->ny > 1) li_update_qy(s, dxf, dtf, gf);
  }
  li_closed_bc(s);
  if (limit) {
    li_outflow_beta(s, dxf, dtf);
    if (chan) {
      li_limit_channels(s);
    } else {
      if (s->nx > 1) li_limit_qx(s);
      if (s->ny > 1) li_limit_qy(s);
    }
  }
  if (s->d50) orc_accumulate_transport(s, dt);
  const int compensated = s->h_comp != NULL;
  li_update_h(s, dxf, dtf, compensated ? 0.0f : (float)rain);
  if (compensated && s->rain == NULL) add_uniform_rain_c(s, (float)rain, dtf);
  if (s->rain != NULL) {
    if (compensated)
      add_rain_field_c(s, dtf, (float)rain_scale);
    else
      apply_rain_field_plain(s, dtf, (float)rain_scale);
  }
  if (s->infil != NULL) apply_infiltration(s, dtf);
  int any_open = 0;
  for (int e = 0; e < 4; ++e) any_open |= (s->bc[e] == ORC_OPEN);
  if (any_open) li_open_outflow(s, dt);
}

void orc_li_run_fixed(orc_state *s, double dt, int n_steps, double rain, double rain_scale) {
  for (int k = 0; k < n_steps; ++k) orc_li_step(s, dt, rain, 1, rain_scale);
}

/* ========================================================================== */
/* sediment                                                                    */
/* ========================================================================== */
#define SUBMERGED_SG ((float)1.65)
#define SHIELDS_CRITICAL ((float)0.047)
#define MPM_COEFFICIENT 8.0f

/* sediment.py:161-168 kinematic_shear */
static inline float kinematic_shear(float q, float h, float radius, float n) {
  return G_F32 * n * n * q * q / (h * h * orc_cbrt(radius));
}
/* sediment.py:171-176 shields_number */
float orc_shields_number(float q, float h, float radius, float n, float d50) {
  return kinematic_shear(q, h, radius, n) / (SUBMERGED_SG * G_F32 * d50);
}
/* sediment.py:179-194 mpm_capacity */
static inline float mpm_capacity(float theta, float d50) {
  if (theta <= SHIELDS_CRITICAL) return 0.0f;
  float excess = theta - SHIELDS_CRITICAL;
  return MPM_COEFFICIENT * orc_pow15(excess) *
         sqrtf(SUBMERGED_SG * G_F32 * d50 * d50 * d50);
}
/* sediment.py:197-207 face_capacity */
float orc_face_capacity(float q, float h, float radius, float n, float d50) {
  if (h < HDRY_F32 || d50 <= 0.0f || q == 0.0f) return 0.0f;
  float qs = mpm_capacity(orc_shields_number(q, h, radius, n, d50), d50);
  return q < 0.0f ? -qs : qs;
}

/* sediment.py:276-302 _combined_capacity: floodplain and channel components, each
 * from its own flow depth and hydraulic radius, mixed by the face's width fraction */
static inline float combined_capacity(float q_ch, float q_fp, float eta_max, float z_bank,
                                      float z_ch, float w_face, float n_fp, float n_ch,
                                      float d50, float dx) {
  float h_fp = eta_max - z_bank;
  float qs_fp = orc_face_capacity(q_fp, h_fp, h_fp, n_fp, d50);
  float qs_ch = 0.0f;
  if (w_face > 0.0f) {
    float h_ch = fmn(eta_max, z_bank) - z_ch;
    if (h_ch >= HDRY_F32)
      qs_ch = orc_face_capacity(q_ch, h_ch, w_face * h_ch / (w_face + 2.0f * h_ch), n_ch, d50);
  }
  float frac = w_face / dx;
  return qs_fp * (1.0f - frac) + qs_ch * frac;
}

/* sediment.py:305-377 accumulate_qs_x_channels / accumulate_qs_y_channels */
static void accumulate_transport_channels(orc_state *s, float dtf) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float dx = (float)s->dx;
  if (nx > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
      for (int j = 1; j < nx; ++j) {
        const size_t a = IX(i, j - 1, nx), b = IX(i, j, nx), f = IX(i, j, w);
        float qs = combined_capacity(
            s->qx_ch[f], s->qx_fp[f], fmx(s->eta[a], s->eta[b]), fmx(s->z[a], s->z[b]),
            fmx(s->z[a] - s->ch_d[a], s->z[b] - s->ch_d[b]), fmn(s->ch_w[a], s->ch_w[b]),
            0.5f * (s->n[a] + s->n[b]), 0.5f * (s->ch_n[a] + s->ch_n[b]),
            0.5f * (s->d50[a] + s->d50[b]), dx);
        kahan_add(&s->qs_int_x[f], &s->qs_comp_x[f], qs * dtf);
      }
  }
  if (ny > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const size_t a = IX(i - 1, j, nx), b = IX(i, j, nx), f = IX(i, j, nx);
        float qs = combined_capacity(
            s->qy_ch[f], s->qy_fp[f], fmx(s->eta[a], s->eta[b]), fmx(s->z[a], s->z[b]),
            fmx(s->z[a] - s->ch_d[a], s->z[b] - s->ch_d[b]), fmn(s->ch_w[a], s->ch_w[b]),
            0.5f * (s->n[a] + s->n[b]), 0.5f * (s->ch_n[a] + s->ch_n[b]),
            0.5f * (s->d50[a] + s->d50[b]), dx);
        kahan_add(&s->qs_int_y[f], &s->qs_comp_y[f], qs * dtf);
      }
  }
}

/* sediment.py:218-273 accumulate_qs_x / accumulate_qs_y; :444-488 dispatch */
void orc_accumulate_transport(orc_state *s, double dt) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float dtf = (float)dt;
  if (s->ch_w) {
    accumulate_transport_channels(s, dtf);
    return;
  }
  if (nx > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
      for (int j = 1; j < nx; ++j) {
        float eta_max = fmx(s->eta[IX(i, j - 1, nx)], s->eta[IX(i, j, nx)]);
        float z_bank = fmx(s->z[IX(i, j - 1, nx)], s->z[IX(i, j, nx)]);
        float h_face = eta_max - z_bank;
        float qs = orc_face_capacity(s->qx[IX(i, j, w)], h_face, h_face,
                                     0.5f * (s->n[IX(i, j - 1, nx)] + s->n[IX(i, j, nx)]),
                                     0.5f * (s->d50[IX(i, j - 1, nx)] + s->d50[IX(i, j, nx)]));
        kahan_add(&s->qs_int_x[IX(i, j, w)], &s->qs_comp_x[IX(i, j, w)], qs * dtf);
      }
  }
  if (ny > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        float eta_max = fmx(s->eta[IX(i - 1, j, nx)], s->eta[IX(i, j, nx)]);
        float z_bank = fmx(s->z[IX(i - 1, j, nx)], s->z[IX(i, j, nx)]);
        float h_face = eta_max - z_bank;
        float qs = orc_face_capacity(s->qy[IX(i, j, nx)], h_face, h_face,
                                     0.5f * (s->n[IX(i - 1, j, nx)] + s->n[IX(i, j, nx)]),
                                     0.5f * (s->d50[IX(i - 1, j, nx)] + s->d50[IX(i, j, nx)]));
        kahan_add(&s->qs_int_y[IX(i, j, nx)], &s->qs_comp_y[IX(i, j, nx)], qs * dtf);
      }
  }
}

/* sediment.py:383-426 exner_update (float64, widened before subtracting) */
void orc_exner_update(int32_t ny, int32_t nx, const float *qs_int_x, const float *qs_int_y,
                      const float *dz_lo, const float *dz_hi, double dx,
                      double inv_one_minus_p, double *dz_cum, double *dz_unapplied) {
  /* dx and 1/(1-p) reach the kernel as float32 parameters, widened inside.  The
   * narrowing is held in float variables: GCC 11 dropped a (double)(float) cast on
   * a const double captured by the OpenMP region below. */
  const float dx_f = (float)dx;
  const float iomp_f = (float)inv_one_minus_p;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      const double dx64 = (double)dx_f;
      const double iomp = (double)iomp_f;
      double div_x = (double)qs_int_x[IX(i, j + 1, nx + 1)] - (double)qs_int_x[IX(i, j, nx + 1)];
      double div_y = (double)qs_int_y[IX(i + 1, j, nx)] - (double)qs_int_y[IX(i, j, nx)];
      double want = dz_cum[IX(i, j, nx)] - iomp * (div_x + div_y) / dx64;
      double nw = want;
      double lo = (double)dz_lo[IX(i, j, nx)];
      double hi = (double)dz_hi[IX(i, j, nx)];
      if (nw < lo) nw = lo;
      if (nw > hi) nw = hi;
      if (nw != want) dz_unapplied[IX(i, j, nx)] = dz_unapplied[IX(i, j, nx)] + (want - nw);
      dz_cum[IX(i, j, nx)] = nw;
    }
}

/* sediment.py:429-441 rebuild_bed */
void orc_rebuild_bed(int32_t ny, int32_t nx, const float *z0, const double *dz_cum, float *z) {
  const size_t cells = (size_t)ny * nx;
  for (size_t k = 0; k < cells; ++k) z[k] = (float)((double)z0[k] + dz_cum[k]);
}

/* inflow.py:91-113 _inject_point_sources_c */
void orc_inject_point_sources_c(orc_state *s, int32_t count, const int32_t *rows,
                                const int32_t *cols, const float *add_depth, float *comp) {
  for (int k = 0; k < count; ++k)
    kahan_add(&s->h[IX(rows[k], cols[k], s->nx)], &comp[k], add_depth[k]);
}

/* scheduler.py:48-72 fill_span */
double orc_fill_span(double span, double dt_raw) {
  if (!isfinite(span)) return dt_raw;
  if (span <= dt_raw) return span;
  double n = ceil(span / dt_raw);
  double dt = span / n;
  if (dt > dt_raw) dt = span / (n + 1.0);
  return dt;
}

/* ========================================================================== */
/* HLLC                                                                        */
/* ========================================================================== */

/* hllc.py:158-165 _minmod */
static inline float minmod(float a, float b) {
  if (a * b <= 0.0f) return 0.0f;
  if (fabsf(a) < fabsf(b)) return a;
  return b;
}
/* hllc.py:168-171 _cget (edge-clamped read) */
static inline float cget(const float *a, int i, int j, int ny, int nx) {
  int ii = i < 0 ? 0 : (i > ny - 1 ? ny - 1 : i);
  int jj = j < 0 ? 0 : (j > nx - 1 ? nx - 1 : j);
  return a[IX(ii, jj, nx)];
}
/* hllc.py:174-188 _dryfactor */
static inline float dryfactor(float h0, float hm, float hp, float dry) {
  if (h0 <= dry || hm <= dry || hp <= dry) return 0.0f;
  return 1.0f;
}

/* hllc.py:191-256 _hllc */
static void hllc(float hl, float ul, float vtl, float hr, float ur, float vtr, float g, float dry,
                 float out[3]) {
  float cl = sqrtf(g * fmx(hl, 0.0f));
  float cr = sqrtf(g * fmx(hr, 0.0f));
  int dry_l = hl <= dry;
  int dry_r = hr <= dry;
  if (dry_l && dry_r) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  float sl, sr;
  if (dry_l) {
    sl = ur - 2.0f * cr;
    sr = ur + cr;
  } else if (dry_r) {
    sl = ul - cl;
    sr = ul + 2.0f * cl;
  } else {
    float hs = 0.5f * (cl + cr) + 0.25f * (ul - ur);
    float h_star = fmx(hs * hs / g, 0.0f);
    float c_star = sqrtf(g * h_star);
    float u_star = 0.5f * (ul + ur) + cl - cr;
    sl = fmn(ul - cl, u_star - c_star);
    sr = fmx(ur + cr, u_star + c_star);
  }
  float fh_l = hl * ul;
  float fmn_l = hl * ul * ul + 0.5f * g * hl * hl;
  float fh_r = hr * ur;
  float fmn_r = hr * ur * ur + 0.5f * g * hr * hr;
  float fh, fm;
  if (sl >= 0.0f) {
    fh = fh_l;
    fm = fmn_l;
  } else if (sr <= 0.0f) {
    fh = fh_r;
    fm = fmn_r;
  } else {
    float inv = 1.0f / (sr - sl);
    fh = (sr * fh_l - sl * fh_r + sl * sr * (hr - hl)) * inv;
    fm = (sr * fmn_l - sl * fmn_r + sl * sr * (hr * ur - hl * ul)) * inv;
  }
  float den_s = hr * (ur - sr) - hl * (ul - sl);
  float s_star = 0.5f * (ul + ur);
  if (fabsf(den_s) > 1.0e-12f) s_star = (sl * hr * (ur - sr) - sr * hl * (ul - sl)) / den_s;
  float vt = vtr;
  if (s_star >= 0.0f) vt = vtl;
  out[0] = fh;
  out[1] = fm;
  out[2] = fh * vt;
}

void orc_hllc_flux(float hl, float ul, float vtl, float hr, float ur, float vtr, float out[3]) {
  hllc(hl, ul, vtl, hr, ur, vtr, G_F32, HDRY_F32, out);
}

/* hllc.py:262-282 _primitives (eta goes to s->eta) */
static void hl_primitives(orc_state *s, const float *h, const float *hu, const float *hv) {
  const size_t cells = (size_t)s->ny * s->nx;
  const float dry = HDRY_F32;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) {
    float hh = h[k];
    if (hh > dry) {
      s->uvel[k] = hu[k] / hh;
      s->vvel[k] = hv[k] / hh;
    } else {
      s->uvel[k] = 0.0f;
      s->vvel[k] = 0.0f;
    }
    s->eta[k] = hh + s->z[k];
  }
}

/* hllc.py:285-354 _flux_x over all faces (ny, nx+1) */
static void hl_flux_x(orc_state *s, const float *h) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float g = G_F32, dry = HDRY_F32;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j <= nx; ++j) {
      float e0 = cget(s->eta, i, j - 2, ny, nx), e1 = cget(s->eta, i, j - 1, ny, nx);
      float e2 = cget(s->eta, i, j, ny, nx), e3 = cget(s->eta, i, j + 1, ny, nx);
      float z0 = cget(s->z, i, j - 2, ny, nx), z1 = cget(s->z, i, j - 1, ny, nx);
      float z2 = cget(s->z, i, j, ny, nx), z3 = cget(s->z, i, j + 1, ny, nx);
      float u0 = cget(s->uvel, i, j - 2, ny, nx), u1 = cget(s->uvel, i, j - 1, ny, nx);
      float u2 = cget(s->uvel, i, j, ny, nx), u3 = cget(s->uvel, i, j + 1, ny, nx);
      float v0 = cget(s->vvel, i, j - 2, ny, nx), v1 = cget(s->vvel, i, j - 1, ny, nx);
      float v2 = cget(s->vvel, i, j, ny, nx), v3 = cget(s->vvel, i, j + 1, ny, nx);
      float h0 = cget(h, i, j - 2, ny, nx), h1 = cget(h, i, j - 1, ny, nx);
      float h2 = cget(h, i, j, ny, nx), h3 = cget(h, i, j + 1, ny, nx);
      float wl = dryfactor(h1, h0, h2, dry);
      float wr = dryfactor(h2, h1, h3, dry);
      float eta_l = e1 + 0.5f * wl * minmod(e1 - e0, e2 - e1);
      float eta_r = e2 - 0.5f * wr * minmod(e2 - e1, e3 - e2);
      float zf_l = z1 + 0.5f * wl * minmod(z1 - z0, z2 - z1);
      float zf_r = z2 - 0.5f * wr * minmod(z2 - z1, z3 - z2);
      float u_l = u1 + 0.5f * wl * minmod(u1 - u0, u2 - u1);
      float u_r = u2 - 0.5f * wr * minmod(u2 - u1, u3 - u2);
      float v_l = v1 + 0.5f * wl * minmod(v1 - v0, v2 - v1);
      float v_r = v2 - 0.5f * wr * minmod(v2 - v1, v3 - v2);
      float h_l = fmx(eta_l - zf_l, 0.0f);
      float h_r = fmx(eta_r - zf_r, 0.0f);
      float z_star = fmx(zf_l, zf_r);
      float hl_star = fmx(eta_l - z_star, 0.0f);
      float hr_star = fmx(eta_r - z_star, 0.0f);
      float f[3];
      hllc(hl_star, u_l, v_l, hr_star, u_r, v_r, g, dry, f);
      s->fx_h[IX(i, j, w)] = f[0];
      s->fx_mn_l[IX(i, j, w)] = f[1] + 0.5f * g * (h_l * h_l - hl_star * hl_star);
      s->fx_mn_r[IX(i, j, w)] = f[1] + 0.5f * g * (h_r * h_r - hr_star * hr_star);
      s->fx_mt[IX(i, j, w)] = f[2];
    }
}

/* hllc.py:357-424 _flux_y over all faces (ny+1, nx) */
static void hl_flux_y(orc_state *s, const float *h) {
  const int ny = s->ny, nx = s->nx;
  const float g = G_F32, dry = HDRY_F32;
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= ny; ++i)
    for (int j = 0; j < nx; ++j) {
      float e0 = cget(s->eta, i - 2, j, ny, nx), e1 = cget(s->eta, i - 1, j, ny, nx);
      float e2 = cget(s->eta, i, j, ny, nx), e3 = cget(s->eta, i + 1, j, ny, nx);
      float z0 = cget(s->z, i - 2, j, ny, nx), z1 = cget(s->z, i - 1, j, ny, nx);
      float z2 = cget(s->z, i, j, ny, nx), z3 = cget(s->z, i + 1, j, ny, nx);
      float u0 = cget(s->uvel, i - 2, j, ny, nx), u1 = cget(s->uvel, i - 1, j, ny, nx);
      float u2 = cget(s->uvel, i, j, ny, nx), u3 = cget(s->uvel, i + 1, j, ny, nx);
      float v0 = cget(s->vvel, i - 2, j, ny, nx), v1 = cget(s->vvel, i - 1, j, ny, nx);
      float v2 = cget(s->vvel, i, j, ny, nx), v3 = cget(s->vvel, i + 1, j, ny, nx);
      float h0 = cget(h, i - 2, j, ny, nx), h1 = cget(h, i - 1, j, ny, nx);
      float h2 = cget(h, i, j, ny, nx), h3 = cget(h, i + 1, j, ny, nx);
      float wl = dryfactor(h1, h0, h2, dry);
      float wr = dryfactor(h2, h1, h3, dry);
      float eta_l = e1 + 0.5f * wl * minmod(e1 - e0, e2 - e1);
      float eta_r = e2 - 0.5f * wr * minmod(e2 - e1, e3 - e2);
      float zf_l = z1 + 0.5f * wl * minmod(z1 - z0, z2 - z1);
      float zf_r = z2 - 0.5f * wr * minmod(z2 - z1, z3 - z2);
      float vn_l = v1 + 0.5f * wl * minmod(v1 - v0, v2 - v1);
      float vn_r = v2 - 0.5f * wr * minmod(v2 - v1, v3 - v2);
      float ut_l = u1 + 0.5f * wl * minmod(u1 - u0, u2 - u1);
      float ut_r = u2 - 0.5f * wr * minmod(u2 - u1, u3 - u2);
      float h_l = fmx(eta_l - zf_l, 0.0f);
      float h_r = fmx(eta_r - zf_r, 0.0f);
      float z_star = fmx(zf_l, zf_r);
      float hl_star = fmx(eta_l - z_star, 0.0f);
      float hr_star = fmx(eta_r - z_star, 0.0f);
      float f[3];
      hllc(hl_star, vn_l, ut_l, hr_star, vn_r, ut_r, g, dry, f);
      s->fy_h[IX(i, j, nx)] = f[0];
      s->fy_mn_l[IX(i, j, nx)] = f[1] + 0.5f * g * (h_l * h_l - hl_star * hl_star);
      s->fy_mn_r[IX(i, j, nx)] = f[1] + 0.5f * g * (h_r * h_r - hr_star * hr_star);
      s->fy_mt[IX(i, j, nx)] = f[2];
    }
}

/* hllc.py:993-1109 _apply_boundaries: stage ghosts (:537-628), then walls (:442-521) */
static void hl_apply_boundaries(orc_state *s, const float *h, const float stages[4]) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float g = G_F32, dry = HDRY_F32;
  float f[3];
  if (s->bc[ORC_WEST] == ORC_FIXED_STAGE)
    for (int i = 0; i < ny; ++i) {
      float h_g = fmx(stages[ORC_WEST] - s->z[IX(i, 0, nx)], 0.0f);
      float u = s->uvel[IX(i, 0, nx)], v = s->vvel[IX(i, 0, nx)];
      hllc(h_g, u, v, h[IX(i, 0, nx)], u, v, g, dry, f);
      s->fx_h[IX(i, 0, w)] = f[0];
      s->fx_mn_r[IX(i, 0, w)] = f[1];
      s->fx_mt[IX(i, 0, w)] = f[2];
    }
  if (s->bc[ORC_EAST] == ORC_FIXED_STAGE)
    for (int i = 0; i < ny; ++i) {
      float u = s->uvel[IX(i, nx - 1, nx)], v = s->vvel[IX(i, nx - 1, nx)];
      float h_g = fmx(stages[ORC_EAST] - s->z[IX(i, nx - 1, nx)], 0.0f);
      hllc(h[IX(i, nx - 1, nx)], u, v, h_g, u, v, g, dry, f);
      s->fx_h[IX(i, nx, w)] = f[0];
      s->fx_mn_l[IX(i, nx, w)] = f[1];
      s->fx_mt[IX(i, nx, w)] = f[2];
    }
  if (s->bc[ORC_NORTH] == ORC_FIXED_STAGE)
    for (int j = 0; j < nx; ++j) {
      float h_g = fmx(stages[ORC_NORTH] - s->z[IX(0, j, nx)], 0.0f);
      float u = s->uvel[IX(0, j, nx)], v = s->vvel[IX(0, j, nx)];
      hllc(h_g, v, u, h[IX(0, j, nx)], v, u, g, dry, f);
      s->fy_h[IX(0, j, nx)] = f[0];
      s->fy_mn_r[IX(0, j, nx)] = f[1];
      s->fy_mt[IX(0, j, nx)] = f[2];
    }
  if (s->bc[ORC_SOUTH] == ORC_FIXED_STAGE)
    for (int j = 0; j < nx; ++j) {
      float u = s->uvel[IX(ny - 1, j, nx)], v = s->vvel[IX(ny - 1, j, nx)];
      float h_g = fmx(stages[ORC_SOUTH] - s->z[IX(ny - 1, j, nx)], 0.0f);
      hllc(h[IX(ny - 1, j, nx)], v, u, h_g, v, u, g, dry, f);
      s->fy_h[IX(ny, j, nx)] = f[0];
      s->fy_mn_l[IX(ny, j, nx)] = f[1];
      s->fy_mt[IX(ny, j, nx)] = f[2];
    }
  if (s->bc[ORC_WEST] == ORC_CLOSED)
    for (int i = 0; i < ny; ++i) {
      float h0 = h[IX(i, 0, nx)], u = s->uvel[IX(i, 0, nx)], v = s->vvel[IX(i, 0, nx)];
      hllc(h0, -u, v, h0, u, v, g, dry, f);
      s->fx_h[IX(i, 0, w)] = f[0];
      s->fx_mn_r[IX(i, 0, w)] = f[1];
      s->fx_mt[IX(i, 0, w)] = f[2];
    }
  if (s->bc[ORC_EAST] == ORC_CLOSED)
    for (int i = 0; i < ny; ++i) {
      float h0 = h[IX(i, nx - 1, nx)], u = s->uvel[IX(i, nx - 1, nx)];
      float v = s->vvel[IX(i, nx - 1, nx)];
      hllc(h0, u, v, h0, -u, v, g, dry, f);
      s->fx_h[IX(i, nx, w)] = f[0];
      s->fx_mn_l[IX(i, nx, w)] = f[1];
      s->fx_mt[IX(i, nx, w)] = f[2];
    }
  if (s->bc[ORC_NORTH] == ORC_CLOSED)
    for (int j = 0; j < nx; ++j) {
      float h0 = h[IX(0, j, nx)], u = s->uvel[IX(0, j, nx)], v = s->vvel[IX(0, j, nx)];
      hllc(h0, -v, u, h0, v, u, g, dry, f);
      s->fy_h[IX(0, j, nx)] = f[0];
      s->fy_mn_r[IX(0, j, nx)] = f[1];
      s->fy_mt[IX(0, j, nx)] = f[2];
    }
  if (s->bc[ORC_SOUTH] == ORC_CLOSED)
    for (int j = 0; j < nx; ++j) {
      float h0 = h[IX(ny - 1, j, nx)], u = s->uvel[IX(ny - 1, j, nx)];
      float v = s->vvel[IX(ny - 1, j, nx)];
      hllc(h0, v, u, h0, -v, u, g, dry, f);
      s->fy_h[IX(ny, j, nx)] = f[0];
      s->fy_mn_l[IX(ny, j, nx)] = f[1];
      s->fy_mt[IX(ny, j, nx)] = f[2];
    }
}

/* hllc.py:700-773 _mass_beta, _limit_fx, _limit_fy */
static void hl_mass_limit(orc_state *s, const float *h, float dx, float dt) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      float q_out = 0.0f;
      float fl = s->fx_h[IX(i, j, w)];
      if (fl < 0.0f) q_out = q_out - fl;
      float fr = s->fx_h[IX(i, j + 1, w)];
      if (fr > 0.0f) q_out = q_out + fr;
      float ft = s->fy_h[IX(i, j, nx)];
      if (ft < 0.0f) q_out = q_out - ft;
      float fb = s->fy_h[IX(i + 1, j, nx)];
      if (fb > 0.0f) q_out = q_out + fb;
      float out_depth = dt / dx * q_out;
      if (out_depth > 0.0f)
        s->beta[IX(i, j, nx)] = fclamp(h[IX(i, j, nx)] / out_depth, 0.0f, 1.0f);
      else
        s->beta[IX(i, j, nx)] = 1.0f;
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j <= nx; ++j) {
      float q = s->fx_h[IX(i, j, w)];
      if (q > 0.0f) {
        if (j >= 1) s->fx_h[IX(i, j, w)] = q * s->beta[IX(i, j - 1, nx)];
      } else if (q < 0.0f) {
        if (j <= nx - 1) s->fx_h[IX(i, j, w)] = q * s->beta[IX(i, j, nx)];
      }
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= ny; ++i)
    for (int j = 0; j < nx; ++j) {
      float q = s->fy_h[IX(i, j, nx)];
      if (q > 0.0f) {
        if (i >= 1) s->fy_h[IX(i, j, nx)] = q * s->beta[IX(i - 1, j, nx)];
      } else if (q < 0.0f) {
        if (i <= ny - 1) s->fy_h[IX(i, j, nx)] = q * s->beta[IX(i, j, nx)];
      }
    }
}

/* hllc.py:776-847 _accumulate */
static void hl_accumulate(orc_state *s, const float *h, float dx) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float g = G_F32, dry = HDRY_F32;
  const float *eta = s->eta, *z = s->z;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      float inv_dx = 1.0f / dx;
      s->dh[IX(i, j, nx)] = -((s->fx_h[IX(i, j + 1, w)] - s->fx_h[IX(i, j, w)]) +
                              (s->fy_h[IX(i + 1, j, nx)] - s->fy_h[IX(i, j, nx)])) *
                            inv_dx;
      float hg = cget(h, i, j, ny, nx);
      float wx = dryfactor(hg, cget(h, i, j - 1, ny, nx), cget(h, i, j + 1, ny, nx), dry);
      float wy = dryfactor(hg, cget(h, i - 1, j, ny, nx), cget(h, i + 1, j, ny, nx), dry);
      float sz_x = wx * minmod(cget(z, i, j, ny, nx) - cget(z, i, j - 1, ny, nx),
                               cget(z, i, j + 1, ny, nx) - cget(z, i, j, ny, nx));
      float se_x = wx * minmod(cget(eta, i, j, ny, nx) - cget(eta, i, j - 1, ny, nx),
                               cget(eta, i, j + 1, ny, nx) - cget(eta, i, j, ny, nx));
      float e = eta[IX(i, j, nx)], zc = z[IX(i, j, nx)];
      float hp_x = fmx((e + 0.5f * se_x) - (zc + 0.5f * sz_x), 0.0f);
      float hm_x = fmx((e - 0.5f * se_x) - (zc - 0.5f * sz_x), 0.0f);
      float src_x = -g * 0.5f * (hp_x + hm_x) * sz_x;
      s->dhu[IX(i, j, nx)] = (-((s->fx_mn_l[IX(i, j + 1, w)] - s->fx_mn_r[IX(i, j, w)]) +
                                (s->fy_mt[IX(i + 1, j, nx)] - s->fy_mt[IX(i, j, nx)])) +
                              src_x) *
                             inv_dx;
      float sz_y = wy * minmod(cget(z, i, j, ny, nx) - cget(z, i - 1, j, ny, nx),
                               cget(z, i + 1, j, ny, nx) - cget(z, i, j, ny, nx));
      float se_y = wy * minmod(cget(eta, i, j, ny, nx) - cget(eta, i - 1, j, ny, nx),
                               cget(eta, i + 1, j, ny, nx) - cget(eta, i, j, ny, nx));
      float hp_y = fmx((e + 0.5f * se_y) - (zc + 0.5f * sz_y), 0.0f);
      float hm_y = fmx((e - 0.5f * se_y) - (zc - 0.5f * sz_y), 0.0f);
      float src_y = -g * 0.5f * (hp_y + hm_y) * sz_y;
      s->dhv[IX(i, j, nx)] = (-((s->fy_mn_l[IX(i + 1, j, nx)] - s->fy_mn_r[IX(i, j, nx)]) +
                                (s->fx_mt[IX(i, j + 1, w)] - s->fx_mt[IX(i, j, w)])) +
                              src_y) *
                             inv_dx;
    }
}

/* hllc.py:1152-1269 _eval_L */
/* The bedload integral of an HLLC stage (a coupling the library adds; the reference runs
 * morphology on the local-inertial scheme only, config.py:451-461): every interior
 * face adds 0.5*dt times the MPM capacity of its limited mass flux at the LI face
 * depth max(eta) - max(z) of the stage state (sediment.py:161-273), so the two
 * stages average the transport rate as SSP-RK2 averages the flux. */
static void hl_transport(orc_state *s, float dt) {
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  const float hdt = 0.5f * dt;
  if (nx > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ny; ++i)
      for (int j = 1; j < nx; ++j) {
        const size_t a = IX(i, j - 1, nx), b = IX(i, j, nx), f = IX(i, j, w);
        float hf = fmx(s->eta[a], s->eta[b]) - fmx(s->z[a], s->z[b]);
        float qs = orc_face_capacity(s->fx_h[f], hf, hf, 0.5f * (s->n[a] + s->n[b]),
                                     0.5f * (s->d50[a] + s->d50[b]));
        kahan_add(&s->qs_int_x[f], &s->qs_comp_x[f], qs * hdt);
      }
  }
  if (ny > 1) {
#pragma omp parallel for schedule(static)
    for (int i = 1; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const size_t a = IX(i - 1, j, nx), b = IX(i, j, nx), f = IX(i, j, nx);
        float hf = fmx(s->eta[a], s->eta[b]) - fmx(s->z[a], s->z[b]);
        float qs = orc_face_capacity(s->fy_h[f], hf, hf, 0.5f * (s->n[a] + s->n[b]),
                                     0.5f * (s->d50[a] + s->d50[b]));
        kahan_add(&s->qs_int_y[f], &s->qs_comp_y[f], qs * hdt);
      }
  }
}

static void hl_eval_L(orc_state *s, const float *h, const float *hu, const float *hv,
                      float dt, const float stages[4]) {
  const float dxf = (float)s->dx;
  hl_primitives(s, h, hu, hv);
  hl_flux_x(s, h);
  hl_flux_y(s, h);
  hl_apply_boundaries(s, h, stages);
  hl_mass_limit(s, h, dxf, dt);
  if (s->d50) hl_transport(s, dt);
  hl_accumulate(s, h, dxf);
}

/* hllc.py:638-677 _bank_* and :1115-1149 _bank_boundary_flux */
static void hl_bank(orc_state *s, double wt) {
  if (s->loss_cum == NULL) return;
  const int ny = s->ny, nx = s->nx, w = nx + 1;
  for (int e = 0; e < 4; ++e) {
    int banked = s->bc[e] == ORC_OPEN || s->bc[e] == ORC_FIXED_STAGE;
    if (!banked) continue;
    if (e == ORC_WEST)
      for (int i = 0; i < ny; ++i)
        s->loss_cum[IX(i, 0, nx)] = s->loss_cum[IX(i, 0, nx)] - wt * (double)s->fx_h[IX(i, 0, w)];
    if (e == ORC_EAST)
      for (int i = 0; i < ny; ++i)
        s->loss_cum[IX(i, nx - 1, nx)] =
            s->loss_cum[IX(i, nx - 1, nx)] + wt * (double)s->fx_h[IX(i, nx, w)];
    if (e == ORC_NORTH)
      for (int j = 0; j < nx; ++j)
        s->loss_cum[IX(0, j, nx)] = s->loss_cum[IX(0, j, nx)] - wt * (double)s->fy_h[IX(0, j, nx)];
    if (e == ORC_SOUTH)
      for (int j = 0; j < nx; ++j)
        s->loss_cum[IX(ny - 1, j, nx)] =
            s->loss_cum[IX(ny - 1, j, nx)] + wt * (double)s->fy_h[IX(ny, j, nx)];
  }
}

/* hllc.py:948-966 _reduce_wavespeed */
float orc_hllc_reduce_wavespeed(const orc_state *s) {
  const size_t cells = (size_t)s->ny * s->nx;
  const float g = G_F32, dry = HDRY_F32;
  float m = 0.0f;
  for (size_t k = 0; k < cells; ++k) {
    float hh = s->h[k];
    float c = sqrtf(g * fmx(hh, 0.0f));
    float sp = c;
    if (hh > dry) {
      float au = fabsf(s->hu[k]) / hh;
      float av = fabsf(s->hv[k]) / hh;
      sp = fmx(au, av) + c;
    }
    m = fmx(m, sp);
  }
  return m;
}

/* hllc.py:972-990 compute_dt */
double orc_hllc_compute_dt(const orc_state *s, double alpha, double dt_max) {
  double smax = (double)orc_hllc_reduce_wavespeed(s);
  if (smax <= 0.0) return dt_max;
  double dt = alpha * s->dx / smax;
  return dt < dt_max ? dt : dt_max;
}

/* hllc.py:1272-1373 step */
void orc_hllc_step(orc_state *s, double dt, double rain, double rain_scale,
                   const float stages[4]) {
  const int ny = s->ny, nx = s->nx;
  const size_t cells = (size_t)ny * nx;
  const float dtf = (float)dt, g = G_F32, dry = HDRY_F32;
  const double wt = 0.5 * dt / s->dx;
  /* stage 1 */
  hl_eval_L(s, s->h, s->hu, s->hv, dtf, stages);
  hl_bank(s, wt);
  /* hllc.py:850-878 _rk_stage1 */
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) {
    float hh = fmx(s->h[k] + dtf * s->dh[k], 0.0f);
    s->h1[k] = hh;
    if (hh > dry) {
      s->hu1[k] = s->hu[k] + dtf * s->dhu[k];
      s->hv1[k] = s->hv[k] + dtf * s->dhv[k];
    } else {
      s->hu1[k] = 0.0f;
      s->hv1[k] = 0.0f;
    }
  }
  /* stage 2 */
  hl_eval_L(s, s->h1, s->hu1, s->hv1, dtf, stages);
  hl_bank(s, wt);
  /* hllc.py:881-909 _rk_stage2 */
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) {
    float hh = fmx(0.5f * (s->h[k] + s->h1[k] + dtf * s->dh[k]), 0.0f);
    s->h[k] = hh;
    if (hh > dry) {
      s->hu[k] = 0.5f * (s->hu[k] + s->hu1[k] + dtf * s->dhu[k]);
      s->hv[k] = 0.5f * (s->hv[k] + s->hv1[k] + dtf * s->dhv[k]);
    } else {
      s->hu[k] = 0.0f;
      s->hv[k] = 0.0f;
    }
  }
  /* hllc.py:912-939 _friction */
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < cells; ++k) {
    float hh = s->h[k];
    if (hh <= dry) {
      s->hu[k] = 0.0f;
      s->hv[k] = 0.0f;
      continue;
    }
    float mom = sqrtf(s->hu[k] * s->hu[k] + s->hv[k] * s->hv[k]);
    float denom = manning_denominator(mom, hh, s->n[k], g, dtf);
    s->hu[k] = s->hu[k] / denom;
    s->hv[k] = s->hv[k] / denom;
  }
  /* hllc.py:1350-1373 sources */
  const int compensated = s->h_comp != NULL;
  if (compensated && s->rain == NULL)
    add_uniform_rain_c(s, (float)rain, dtf);
  else if (rain != 0.0) {
    /* hllc.py:942-945 _add_uniform_rain */
    const float r = (float)rain;
    for (size_t k = 0; k < cells; ++k) s->h[k] = s->h[k] + r * dtf;
  }
  if (s->rain != NULL) {
    if (compensated)
      add_rain_field_c(s, dtf, (float)rain_scale);
    else
      apply_rain_field_plain(s, dtf, (float)rain_scale);
  }
  if (s->infil != NULL) apply_infiltration(s, dtf);
}
