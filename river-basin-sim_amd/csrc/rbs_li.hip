// rbs_li.hip -- fused local-inertial time step for gfx950.
//
// One launch performs the whole of local_inertial.step (local_inertial.py:547-727)
// for a tile of BY x BX cells:
//   eta = h + z                                    (:61-69)
//   interior face update + dry guard + friction    (:72-129, friction.py:18-31)
//   closed edge faces = 0                          (boundaries.py:73-91)
//   donor outflow limiter beta, limit faces        (:197-263)
//   bedload transport integral (sediment armed)    (sediment.py:218-273)
//   continuity + uniform rain                      (:132-151)
//   Kahan areal sources / plain rain field         (sources.py:85-121, :154-167)
//   capped infiltration                            (:170-194)
//   open-edge drains, west/east/north/south order  (boundaries.py:109-208)
// The reference's separate launches become phases separated by __syncthreads over
// LDS copies of a halo'd tile.  A tile needs eta/z/n two cells out, the old faces
// and h one cell out; faces in the halo are recomputed bit-identically by the
// neighbouring tile (pure functions of the same inputs, no FMA contraction).
// Fields ping-pong (read index cur, write 1-cur) so no tile reads a value another
// tile has already overwritten.
//
// Per block the kernel also emits fixed-order partials: max h (the next CFL dt),
// sum f64(h) (ledger volume), infiltration and edge-drain depth removed.
#include <hip/hip_ext.h>

#include <cfloat>

#include "rbs_device.hpp"

namespace rbs {

enum : int {
  LI_LIMIT = 1 << 0,
  LI_COMP = 1 << 1,
  LI_FIELD = 1 << 2,
  LI_INFIL = 1 << 3,
  LI_SED = 1 << 4,
  LI_OPEN_W = 1 << 5,
  LI_OPEN_E = 1 << 6,
  LI_OPEN_N = 1 << 7,
  LI_OPEN_S = 1 << 8,
  LI_CHAN = 1 << 9,
};

struct LiArgs {
  float *h[2];   // ping-pong pair; the control block's `cur` names the current one
  float *qx[2];
  float *qy[2];
  const float *z;
  const float *n;
  float *h_comp;
  const float *rain;
  const float *infil;
  const float *d50;
  float *qs_int_x, *qs_comp_x, *qs_int_y, *qs_comp_y;
  // sub-grid channels: geometry and the ping-pong face components
  const float *ch_w, *ch_d, *ch_n;
  float *qx_ch[2], *qx_fp[2], *qy_ch[2], *qy_fp[2];
  const DevCtl *ctl;
  Partials part;
  uint32_t *quiet;  // per tile: which of its buffers and compensations are all +0
  double dx64;
  int64_t pitch;
  int32_t ny, nx;
  int32_t tiles_x, n_tiles;
  int32_t flags;
};

// ---- per-cell arithmetic shared by both kernel families ----------------------------
// Both families call these with the same operands, so their results are identical
// bit for bit (the build disables FMA contraction).

// limit_outflow (local_inertial.py:197-239): the donor factor of one cell from its
// four updated faces (left, right, top, bottom) and its depth.
__device__ __forceinline__ float donor_beta(float ql, float qr, float qt, float qb, float h,
                                            float dtdx) {
  float q_out = 0.0f;
  if (ql < 0.0f) q_out = q_out - ql;
  if (qr > 0.0f) q_out = q_out + qr;
  if (qt < 0.0f) q_out = q_out - qt;
  if (qb > 0.0f) q_out = q_out + qb;
  const float out_depth = dtdx * q_out;
  // h >= out_depth > 0 gives h / out_depth >= 1 exactly (rounding is monotone), so
  // the clamp returns 1 and the division can be skipped; NaN still divides
  const bool divide = out_depth > 0.0f && !(h >= out_depth);
  return divide ? fclamp(h / out_depth, 0.0f, 1.0f) : 1.0f;
}

// limit_qx / limit_qy (:242-263): each face scaled by its donor's factor.
// bc = this cell, bl/br/bt/bb = left/right/top/bottom neighbours.
// The reference leaves a zero face untouched; here a face that is not positive is
// scaled by the downstream donor's factor, which is the same bits: donor_beta returns
// a finite factor in [0, 1], and +-0 times such a factor is +-0 (a NaN face stays
// NaN).  One compare and one select per face, no branch (at 4096^2 the branches cost
// 3.5 % of the step kernel).
__device__ __forceinline__ void limit_faces(float &qxL, float &qxR, float &qyT, float &qyB,
                                            float bc, float bl, float br, float bt, float bb) {
  qxL = qxL * (qxL > 0.0f ? bl : bc);
  qxR = qxR * (qxR > 0.0f ? bc : br);
  qyT = qyT * (qyT > 0.0f ? bt : bc);
  qyB = qyB * (qyB > 0.0f ? bc : bb);
}

struct CellSums {
  float m_h = 0.0f;  // h_max starts at 0 (local_inertial.py:527)
  double vol = 0.0, inf = 0.0, edge = 0.0;
};

struct StepScalars {
  float dt, dtdx, dxdt, rain, rain_scale, rain_cont;
};

template <typename CtlPtr>
__device__ __forceinline__ StepScalars step_scalars(const LiArgs &a, CtlPtr ctl) {
  StepScalars k;
  const double dt64 = ctl->dt;
  k.dt = (float)dt64;
  k.dtdx = k.dt / (float)a.dx64;
  k.dxdt = (float)(dt64 / a.dx64);  // boundaries.py:179 (f64 quotient)
  k.rain = ctl->rain;
  k.rain_scale = ctl->rain_scale;
  k.rain_cont = (a.flags & LI_COMP) ? 0.0f : k.rain;
  return k;
}

// update_h (:132-151), areal sources (sources.py:85-121, local_inertial.py:154-167),
// capped infiltration (:170-194) and the open-edge drains in the reference's launch
// order (boundaries.py:181-208) for cell (r, c) with its limited faces.  `comp`
// (Kahan debt), `rain_v` and `infil_v` are that cell's values, read by the caller;
// the updated debt is returned in `comp`.
__device__ __forceinline__ float cell_update(int flags, int r, int c, int ny, int nx,
                                             float h_old, float qxL, float qxR, float qyT,
                                             float qyB, const StepScalars &k, float &comp,
                                             float rain_v, float infil_v, CellSums &sum) {
  const float net = qxL - qxR + qyT - qyB;
  float h = h_old + k.dtdx * net + k.rain_cont * k.dt;
  if (flags & LI_COMP) {
    if (!(flags & LI_FIELD)) kahan_add(h, comp, k.rain * k.dt);  // add_uniform_rain_c
    else kahan_add(h, comp, rain_v * k.rain_scale * k.dt);        // add_rain_field_c
  } else if (flags & LI_FIELD) {
    h = h + rain_v * k.rain_scale * k.dt;                         // apply_rain_field
  }
  if (flags & LI_INFIL) {
    const float avail = h;
    float inf = infil_v * k.dt;
    if (inf > avail) inf = avail;
    if (inf < 0.0f) inf = 0.0f;
    h = avail - inf;
    sum.inf += (double)avail - (double)h;
  }
  if ((flags & LI_OPEN_W) && c == 0 && nx > 1) {
    const float av = h;
    h = drain(av, fmx(-qxR, 0.0f), k.dxdt);
    sum.edge += (double)av - (double)h;
  }
  if ((flags & LI_OPEN_E) && c == nx - 1 && nx > 1) {
    const float av = h;
    h = drain(av, fmx(qxL, 0.0f), k.dxdt);
    sum.edge += (double)av - (double)h;
  }
  if ((flags & LI_OPEN_N) && r == 0 && ny > 1) {
    const float av = h;
    h = drain(av, fmx(-qyB, 0.0f), k.dxdt);
    sum.edge += (double)av - (double)h;
  }
  if ((flags & LI_OPEN_S) && r == ny - 1 && ny > 1) {
    const float av = h;
    h = drain(av, fmx(qyT, 0.0f), k.dxdt);
    sum.edge += (double)av - (double)h;
  }
  sum.m_h = fmx(sum.m_h, h);
  sum.vol += (double)h;
  return h;
}

// accumulate_qs_x / accumulate_qs_y (sediment.py:218-273) for one face: the MPM
// capacity at the face's flow depth, Kahan-added (times dt) to the integral at `o`.
//
// Most faces of a run carry no bedload (dry, still, or under the threshold): the d50
// pair is read only when the face is wet and moving, and a zero increment onto a zero
// compensation -- kahan_add of +-0 leaves both the integral (never -0: it starts at
// +0) and the +0 compensation exactly as they are -- reads one value and writes
// nothing.  The bits are those of the unconditional update.  Returns the bits of the
// face's compensation after the update.
template <typename DPtr, typename FloatPtr>
__device__ __forceinline__ uint32_t accumulate_face(float q, float eta_a, float eta_b, float z_a,
                                                float z_b, float n_a, float n_b, DPtr d50,
                                                int64_t oa, int64_t ob, float dt,
                                                FloatPtr integ, FloatPtr debt, int64_t o) {
  const float h_face = fmx(eta_a, eta_b) - fmx(z_a, z_b);
  float qs = 0.0f;  // face_capacity's dry and still exits, taken before the d50 reads
  if (!(h_face < kHDry) && q != 0.0f)
    qs = face_capacity(q, h_face, h_face, 0.5f * (n_a + n_b), 0.5f * (d50[oa] + d50[ob]));
  const float inc = qs * dt;
  float sc = debt[o];
  if (inc == 0.0f && sc == 0.0f) return __float_as_uint(sc);
  float si = integ[o];
  kahan_add(si, sc, inc);
  integ[o] = si;
  debt[o] = sc;
  return __float_as_uint(sc);
}

template <int NT = kThreads>
__device__ __forceinline__ void write_partials(const LiArgs &a, int tile, CellSums &sum,
                                               double *s_red) {
  block_reduce4<NT>(sum.m_h, sum.vol, sum.inf, sum.edge, s_red);
  if (threadIdx.x == 0) {
    glob(a.part.speed_max)[tile] = sum.m_h;
    glob(a.part.vol)[tile] = sum.vol;
    glob(a.part.infil)[tile] = sum.inf;
    glob(a.part.edge)[tile] = sum.edge;
  }
}

// ---- family 1: LDS tile ---------------------------------------------------------------
// BY x BX cells per block; the halo'd inputs are staged in LDS and the reference's
// launches become phases separated by barriers.  A tile needs eta/z/n two cells
// out, the old faces and h one cell out; faces in the halo are recomputed
// bit-identically by the neighbouring tile.
template <int BX, int BY>
__device__ __forceinline__ void li_tile(const LiArgs &a, const int tile) {
  const DevCtl *ctl = a.ctl;
  if (ctl->active != 1) return;  // device-driven run already at its stop time
  const int cur = ctl->cur;
  const float *__restrict__ h_in = a.h[cur];
  float *__restrict__ h_out = a.h[cur ^ 1];
  const float *__restrict__ qx_in = a.qx[cur];
  float *__restrict__ qx_out = a.qx[cur ^ 1];
  const float *__restrict__ qy_in = a.qy[cur];
  float *__restrict__ qy_out = a.qy[cur ^ 1];

  constexpr int EW = BX + 4, EH = BY + 4;  // cells two out: eta, z, n
  constexpr int RW = BX + 2, RH = BY + 2;  // cells one out: h, beta
  constexpr int QXW = BX + 3;              // x-faces of the ring-1 cells
  constexpr int QYH = BY + 3;              // y-faces of the ring-1 cells
  __shared__ float s_eta[EH][EW];
  __shared__ float s_z[EH][EW];
  __shared__ float s_n[EH][EW];
  __shared__ float s_h[RH][RW];
  __shared__ float s_qx[RH][QXW];
  __shared__ float s_qy[QYH][RW];
  __shared__ float s_beta[RH][RW];
  __shared__ double s_red[4 * (kThreads / 64)];

  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int y0 = ty * BY, x0 = tx * BX;
  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int tid = threadIdx.x;
  const int flags = a.flags;
  const StepScalars k = step_scalars(a, ctl);
  const float dx = (float)a.dx64;
  const float rdx = 1.0f / dx;  // IEEE division: RN(1/dx)

  // dry tiles (the strip kernels' rule): the flag, then the depths one cell out
  const bool no_rain = k.rain == 0.0f && (!(flags & LI_FIELD) || k.rain_scale == 0.0f);
  if (no_rain && a.quiet[tile] != 0u) {  // block-uniform
    uint32_t wet = 0;
    for (int kk = tid; kk < RH * RW; kk += kThreads) {
      const int lr = kk / RW, lc = kk - (kk / RW) * RW;
      const int gi = clampi(y0 + lr - 1, 0, ny - 1), gj = clampi(x0 + lc - 1, 0, nx - 1);
      wet |= __float_as_uint(h_in[(int64_t)gi * P + gj]);
    }
    if (!__syncthreads_or(wet != 0u)) {
      if (tid == 0) {
        a.part.speed_max[tile] = 0.0f;
        a.part.vol[tile] = 0.0;
        a.part.infil[tile] = 0.0;
        a.part.edge[tile] = 0.0;
      }
      return;
    }
  }
  // bits of everything the thread read (own cells' h, qx, qy) or wrote (the same and
  // the debts) that was not +0
  uint32_t nz = 0;

  // ---- phase A: stage the halo'd tile in LDS (clamped reads stay in bounds) ----
#pragma unroll
  for (int k0 = 0; k0 < EH * EW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < EH * EW) {
      const int lr = kk / EW, lc = kk - (kk / EW) * EW;
      const int gi = clampi(y0 + lr - 2, 0, ny - 1), gj = clampi(x0 + lc - 2, 0, nx - 1);
      const int64_t o = (int64_t)gi * P + gj;
      const float hv = h_in[o], zv = a.z[o];
      if (lr >= 2 && lr < BY + 2 && lc >= 2 && lc < BX + 2 && y0 + lr - 2 < ny && x0 + lc - 2 < nx)
        nz |= __float_as_uint(hv);  // an own cell
      s_eta[lr][lc] = hv + zv;  // compute_eta (:61-69)
      s_z[lr][lc] = zv;
      s_n[lr][lc] = a.n[o];
      if (lr >= 1 && lr <= BY + 2 && lc >= 1 && lc <= BX + 2) s_h[lr - 1][lc - 1] = hv;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < RH * QXW) {
      const int lr = kk / QXW, lf = kk - (kk / QXW) * QXW;
      const int gi = clampi(y0 + lr - 1, 0, ny - 1), gf = clampi(x0 + lf - 1, 0, nx);
      const float q = qx_in[(int64_t)gi * P + gf];
      if (lr >= 1 && lr <= BY && lf >= 1 && lf <= BX && y0 + lr - 1 < ny && x0 + lf - 1 < nx)
        nz |= __float_as_uint(q);  // an own cell's left face
      s_qx[lr][lf] = q;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < QYH * RW) {
      const int lf = kk / RW, lc = kk - (kk / RW) * RW;
      const int gf = clampi(y0 + lf - 1, 0, ny), gj = clampi(x0 + lc - 1, 0, nx - 1);
      const float q = qy_in[(int64_t)gf * P + gj];
      if (lf >= 1 && lf <= BY && lc >= 1 && lc <= BX && y0 + lf - 1 < ny && x0 + lc - 1 < nx)
        nz |= __float_as_uint(q);  // an own cell's top face
      s_qy[lf][lc] = q;
    }
  }
  __syncthreads();

  // ---- phase B: face update of every face of the ring-1 cells -------------------
  // Boundary and out-of-domain faces become 0: the closed BC (boundaries.py:87-91)
  // and the reference's "interior faces only" launch ranges (:599-612).
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < RH * QXW) {
      const int lr = kk / QXW, lf = kk - (kk / QXW) * QXW;
      const int r = y0 + lr - 1, f = x0 + lf - 1;
      float q = 0.0f;
      if (r >= 0 && r < ny && f >= 1 && f <= nx - 1)
        q = li_face(s_qx[lr][lf], s_eta[lr + 1][lf], s_eta[lr + 1][lf + 1], s_z[lr + 1][lf],
                    s_z[lr + 1][lf + 1], s_n[lr + 1][lf], s_n[lr + 1][lf + 1], dx, rdx, k.dt);
      s_qx[lr][lf] = q;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < QYH * RW) {
      const int lf = kk / RW, lc = kk - (kk / RW) * RW;
      const int f = y0 + lf - 1, c = x0 + lc - 1;
      float q = 0.0f;
      if (c >= 0 && c < nx && f >= 1 && f <= ny - 1)
        q = li_face(s_qy[lf][lc], s_eta[lf][lc + 1], s_eta[lf + 1][lc + 1], s_z[lf][lc + 1],
                    s_z[lf + 1][lc + 1], s_n[lf][lc + 1], s_n[lf + 1][lc + 1], dx, rdx, k.dt);
      s_qy[lf][lc] = q;
    }
  }
  __syncthreads();

  // ---- phase C: donor outflow limiter beta of the ring-1 cells --------------------
  const bool limit = flags & LI_LIMIT;
  if (limit) {
#pragma unroll
    for (int k0 = 0; k0 < RH * RW; k0 += kThreads) {
      const int kk = k0 + tid;
      if (kk < RH * RW) {
        const int lr = kk / RW, lc = kk - (kk / RW) * RW;
        s_beta[lr][lc] = donor_beta(s_qx[lr][lc], s_qx[lr][lc + 1], s_qy[lr][lc],
                                    s_qy[lr + 1][lc], s_h[lr][lc], k.dtdx);
      }
    }
    __syncthreads();
  }

  // ---- phase D: limited faces, continuity, sources, sinks, edge drains ------------
  CellSums sum;
#pragma unroll
  for (int k0 = 0; k0 < BY * BX; k0 += kThreads) {
    const int kk = k0 + tid;
    const int lr = kk / BX, lc = kk - (kk / BX) * BX;
    const int r = y0 + lr, c = x0 + lc;
    if (kk < BY * BX && r < ny && c < nx) {
      float qxL = s_qx[lr + 1][lc + 1], qxR = s_qx[lr + 1][lc + 2];
      float qyT = s_qy[lr + 1][lc + 1], qyB = s_qy[lr + 2][lc + 1];
      if (limit)
        limit_faces(qxL, qxR, qyT, qyB, s_beta[lr + 1][lc + 1], s_beta[lr + 1][lc],
                    s_beta[lr + 1][lc + 2], s_beta[lr][lc + 1], s_beta[lr + 2][lc + 1]);
      const int64_t o = (int64_t)r * P + c;
      if (flags & LI_SED) {  // own faces only: left (c >= 1) and top (r >= 1)
        if (c >= 1)
          nz |= accumulate_face(qxL, s_eta[lr + 2][lc + 1], s_eta[lr + 2][lc + 2],
                                s_z[lr + 2][lc + 1], s_z[lr + 2][lc + 2], s_n[lr + 2][lc + 1],
                                s_n[lr + 2][lc + 2], a.d50, o - 1, o, k.dt, a.qs_int_x,
                                a.qs_comp_x, o);
        if (r >= 1)
          nz |= accumulate_face(qyT, s_eta[lr + 1][lc + 2], s_eta[lr + 2][lc + 2],
                                s_z[lr + 1][lc + 2], s_z[lr + 2][lc + 2], s_n[lr + 1][lc + 2],
                                s_n[lr + 2][lc + 2], a.d50, o - P, o, k.dt, a.qs_int_y,
                                a.qs_comp_y, o);
      }
      float comp = (flags & LI_COMP) ? a.h_comp[o] : 0.0f;
      const float rain_v = (flags & LI_FIELD) ? a.rain[o] : 0.0f;
      const float infil_v = (flags & LI_INFIL) ? a.infil[o] : 0.0f;
      const float h = cell_update(flags, r, c, ny, nx, s_h[lr + 1][lc + 1], qxL, qxR, qyT, qyB,
                                  k, comp, rain_v, infil_v, sum);
      if (flags & LI_COMP) a.h_comp[o] = comp;
      nz |= __float_as_uint(h) | __float_as_uint(qxL) | __float_as_uint(qyT) |
            __float_as_uint(comp);
      h_out[o] = h;
      qx_out[o] = qxL;
      if (c == nx - 1) qx_out[o + 1] = qxR;
      qy_out[o] = qyT;
      if (r == ny - 1) qy_out[o + P] = qyB;
    }
  }
  write_partials(a, tile, sum, s_red);
  const bool still = no_rain && !__syncthreads_or(nz != 0u);
  if (tid == 0) a.quiet[tile] = still ? 1u : 0u;
}

// ---- family 2: register strips ----------------------------------------------------------
// Each wave owns a strip of 64 columns (global columns x0-2 .. x0+61) and marches
// down H rows, keeping a rolling window of rows in registers: x-neighbours come from
// lane shuffles, y-neighbours from the previous iterations, so there is no LDS and
// no barrier.  Row R's inputs are ingested at iteration R (prefetched one iteration
// ahead); its x-faces and the y-faces above it are updated at once, the donor factor
// of row R-1 follows, and row R-2 is finished.  The two edge lanes on each side only
// feed their neighbours: lanes 2..61 write, 60 columns per wave.  A wave starts two
// rows above its first row (those faces and factors are recomputed bit-identically
// by the wave above; the first of them only contributes its inputs).  The block's
// waves are stacked vertically.
constexpr int kStripW = 60;
constexpr int kStripWaves = kThreads / 64;

// Element `k` of the row at `base` (a wave-uniform pointer) through a 32-bit byte
// offset, so the load or store pairs a scalar base with one lane register.
template <typename T>
__device__ __forceinline__ T &at(T *base, uint32_t k) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + k * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &at(const T *base, uint32_t k) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + k * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T __attribute__((address_space(1))) &at(gptr<T> base, uint32_t k) {
  using B = gptr<unsigned char>;
  return *(gptr<T>)((B)base + k * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T __attribute__((address_space(1))) &at(gptr<const T> base,
                                                                        uint32_t k) {
  using B = gptr<const unsigned char>;
  return *(gptr<const T>)((B)base + k * (uint32_t)sizeof(T));
}
// Streaming (non-temporal) access in the row blocks for the fields a step rewrites --
// h, qx, qy and the Kahan debt, read once and written once per step -- and ordinary
// access for the static inputs (bed, Manning n, rain and infiltration fields), which
// are read again every step and so can stay in the last-level cache between steps
// while the step's own traffic streams past them (at 4096^2: 179 -> 170-173 us;
// streaming the static inputs too made it slower).  The channel step's static inputs
// (seven fields) are more than that cache holds: its row blocks stream their sources
// and read the rest ordinarily, which measured 3 % faster there.  The strips keep
// ordinary access: their edge lines are shared with the next strip.  The values are
// the same.
template <typename R>
__device__ __forceinline__ float ld_nt(const R &ref) {
  return __builtin_nontemporal_load(&ref);
}
template <typename R>
__device__ __forceinline__ void st_nt(R &ref, float v) {
  __builtin_nontemporal_store(v, &ref);
}

// The neighbouring lane's value through a DPP wave-wide shift (one VALU move, no LDS
// crossbar): lane i takes lane i-1's (wave_shr:1) or lane i+1's (wave_shl:1) value,
// and the lane with no neighbour (lane 0, lane 63) reads 0 (bound_ctrl), so no copy of
// the old value is needed.  What lanes 0 and 63 compute from the 0 (a face, then a
// donor factor) reaches lanes 1 and 62 at most, which write nothing either.
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138 /* wave_shr:1 */,
                                                 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130 /* wave_shl:1 */,
                                                 0xF, 0xF, true));
}

// NW > 0 ("row blocks"): the block's NW waves sit side by side on the same H rows,
// NW * 60 columns whose byte span is a multiple of 128 B, and each finished row's
// outputs go through LDS so that the block stores them as whole 256-B lines.  A
// wave's own 60 columns start 8 B off a line, and their direct stores leave the two
// lines at each strip edge written in parts by two waves at different times, which
// costs the HBM far more than the bytes (the LI step's store pattern of 60-column
// strips takes about 11% longer than the same bytes stored in whole lines).  One barrier per row keeps the
// double-buffered LDS rows in step.
//
// Dry tiles (the 60-column strips here and the LDS tiles above follow the same rule).
// A tile whose depths, one cell out included, are all +0 at a step with no
// rain has every face dry (face depth max(z_a, z_b) - max(z_a, z_b) = 0 < kHDry), so
// its faces, depths, compensations and partials come out +0 whatever the old faces
// were.  Such a tile can skip the step if the output buffer already holds +0 in its
// cells and its in-place compensations are +0: then nothing it would write differs
// from what memory holds.  The per-tile flag `quiet` keeps that knowledge: set by a
// computed step that read and wrote only +0 (h, qx, qy of both buffers and the Kahan
// debts of the tile's cells), cleared by one that did not, left by a skipped step.
// Writers outside the step touch only the current buffer's depths (point inflows,
// reservoirs), which the skip reads itself, one cell out, before relying on the flag;
// anything else (uploads of h, q or debts, a change of tiling) clears the flags.

// The strips' LDS, one allocation per kernel whatever li_strip instances it inlines
// (a static __shared__ array in li_strip itself would be allocated per instance).
template <int NT>
__device__ __forceinline__ double *strip_lds_red() {
  __shared__ double s_red[4 * (NT / 64)];
  return s_red;
}
template <int SLOTS, int BWC>
__device__ __forceinline__ float (&strip_lds_rows())[SLOTS][4][BWC] {
  __shared__ float s_row[SLOTS][4][BWC];
  return s_row;
}

// LIM = 1: the limiter is on, CMP = 1: the Kahan debt is armed (flags known at compile
// time, so that the compiler does not rebuild their lane masks with vector instructions
// every row: at 4096^2 162.9 -> 161.4 us for the limiter, 160.7 -> 158.8 us for the
// debt); 0 reads them from the flags.
template <int H, bool SED, int NW = 0, int FL = 1, int LIM = 0, int CMP = 0>
__device__ __forceinline__ void li_strip(const LiArgs &a, const int tile) {
  // FL: finished rows stored per barrier, from a ring of 2 FL LDS rows (two rows per
  // barrier measured 3 % faster than one at 4096^2; three no better, four cost
  // occupancy)
  static_assert(FL >= 1 && FL <= 4, "one to four rows per barrier");
  constexpr bool ROWS = NW > 0;
  constexpr int BWC = ROWS ? kStripW * NW : 1;  // a row block's own columns
  static_assert(!ROWS || (BWC * 4) % 128 == 0, "row blocks must span whole 128-B lines");
  // every pointer as a global one (a batch reads them from memory)
  const gptr<const DevCtl> ctl = glob(a.ctl);
  if (ctl->active != 1) return;
  const int cur = ctl->cur;
  const gptr<const float> __restrict__ h_in = glob(a.h[cur]);
  const gptr<float> __restrict__ h_out = glob(a.h[cur ^ 1]);
  const gptr<const float> __restrict__ qx_in = glob(a.qx[cur]);
  const gptr<float> __restrict__ qx_out = glob(a.qx[cur ^ 1]);
  const gptr<const float> __restrict__ qy_in = glob(a.qy[cur]);
  const gptr<float> __restrict__ qy_out = glob(a.qy[cur ^ 1]);
  const gptr<const float> __restrict__ zf = glob(a.z);
  const gptr<const float> __restrict__ nf = glob(a.n);
  const gptr<float> __restrict__ hcomp = glob(a.h_comp);
  const gptr<const float> __restrict__ rainf = glob(a.rain);
  const gptr<const float> __restrict__ infilf = glob(a.infil);
  constexpr int NT = ROWS ? 64 * NW : kThreads;
  double *s_red = strip_lds_red<NT>();
  // row blocks: the finished row's h, qx, qy and Kahan compensation, double-buffered
  auto &s_row = strip_lds_rows<ROWS ? 2 * FL : 1, BWC>();

  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int flags = CMP ? (a.flags | LI_COMP) : a.flags;  // CMP = 1: Kahan debt armed
  const bool limit = LIM ? true : (flags & LI_LIMIT) != 0;
  const StepScalars k = step_scalars(a, ctl);
  const float dx = (float)a.dx64;
  const float rdx = 1.0f / dx;  // IEEE division: RN(1/dx)

  // the wave index, and so every row index below, is wave-uniform: row offsets live
  // in scalar registers and only the column offset is per lane
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int strip = ROWS ? tx * NW + wave : tx;
  const int c = strip * kStripW - 2 + lane;
  const int y0 = ROWS ? ty * H : (ty * kStripWaves + wave) * H;  // block-uniform for ROWS
  const int y_end = y0 + H < ny ? y0 + H : ny;
  // a row block's waves past the east edge keep only the block's barriers and stores
  const bool idle = ROWS && strip * kStripW >= nx;
  const int x0 = ROWS ? tx * BWC : 0;  // the row block's first column
  // clamped reads stay in bounds; unsigned offsets let a load pair a scalar row
  // base with a 32-bit lane offset
  const uint32_t cc = clampi(c, 0, nx - 1);
  const uint32_t cf = clampi(c, 0, nx);
  const uint32_t cu = (uint32_t)c;  // used only where 0 <= c < nx
  const bool col_ok = c >= 0 && c < nx;
  const bool xface_ok = c >= 1 && c <= nx - 1;  // interior x-face left of column c
  const bool own = lane >= 2 && lane < 2 + kStripW && col_ok;

  // dry tiles are skipped by the 60-column strips (the row blocks, the kernel of wide
  // rain-on-terrain grids, keep their registers and compute every tile)
  constexpr bool DRY = !ROWS;
  const gptr<uint32_t> quiet = glob(a.quiet);
  const bool no_rain = k.rain == 0.0f && (!(flags & LI_FIELD) || k.rain_scale == 0.0f);
  if (DRY && no_rain && quiet[tile] != 0u) {  // block-uniform
    uint32_t wet = 0;
    if (y0 < ny)  // the tile's rows and one row either side, 64 columns per wave
      for (int r = y0 - 1; r <= y_end; ++r)
        wet |= __float_as_uint(at(h_in + (int64_t)clampi(r, 0, ny - 1) * P, cc));
    if (!__syncthreads_or(wet != 0u)) {
      if (threadIdx.x == 0) {
        glob(a.part.speed_max)[tile] = 0.0f;
        glob(a.part.vol)[tile] = 0.0;
        glob(a.part.infil)[tile] = 0.0;
        glob(a.part.edge)[tile] = 0.0;
      }
      return;
    }
  }
  // the wave's lanes that read or wrote anything but +0 (h, qx, qy in and out, the
  // debts): ballots, so the mask stays in scalar registers
  uint64_t nz = 0;

  // row blocks: the finished row i (iteration it) sits in LDS row slot it % (2 FL);
  // after every FL-th row of the block (and its last), one barrier, then wave w stores
  // the pending rows' columns [64w, 64w + 64) (and [64(w + NW), ...) when the block
  // spans more lines than waves).  Every wave reaches the barriers at the same
  // iterations (the rows are the block's), and a slot is written again only 2 FL rows
  // later, after a barrier that its readers had to pass first.
  auto flush_rows = [&](int it, int i) {
    const int off = i - y0;
    if ((off % FL) != FL - 1 && i != y_end - 1) return;  // block-uniform
    __syncthreads();
#pragma unroll
    for (int back = FL - 1; back >= 0; --back) {
      if (back > off % FL) continue;  // a short last group
      const int slot = (it - back) % (2 * FL);  // it >= 4 for a live row
      const int64_t row = (int64_t)(i - back) * P;  // uniform
#pragma unroll
      for (int part = 0; part < (BWC + 63) / 64; part += (NW > 0 ? NW : 1)) {
        const int j = 64 * (part + wave) + lane;
        if (j < BWC && x0 + j < nx) {
          // an empty asm makes the column opaque, so the loop-invariant per-lane
          // pointers (four 64-bit values) are not hoisted into registers: each store
          // pairs the row's scalar base with this 32-bit offset
          uint32_t col = (uint32_t)(x0 + j);
          asm volatile("" : "+v"(col));
          st_nt(at(h_out + row, col), s_row[slot][0][j]);
          st_nt(at(qx_out + row, col), s_row[slot][1][j]);
          st_nt(at(qy_out + row, col), s_row[slot][2][j]);
          if (flags & LI_COMP) st_nt(at(hcomp + row, col), s_row[slot][3][j]);
        }
      }
    }
  };

  CellSums sum;
  if (ROWS && idle && y0 < ny) {
    for (int it = 1; it < H + 4; ++it) {
      const int i = y0 - 4 + it;
      if (i >= y0 && i < y_end) flush_rows(it, i);
    }
  } else if (y0 < ny) {  // wave-uniform (block-uniform for row blocks)
    // the window: row R-1 inputs, updated faces of rows R-1 and R-2, old depths of
    // rows R-1 and R-2, donor factors of rows R-2 and R-3
    float e1 = 0.0f, z1 = 0.0f, n1 = 0.0f;
    float fx1 = 0.0f, fx2 = 0.0f, fy1 = 0.0f, fy2 = 0.0f;
    float h1 = 0.0f, h2 = 0.0f, b2 = 1.0f, b3 = 1.0f;
    float e2 = 0.0f, z2 = 0.0f, n2 = 0.0f, e3 = 0.0f, z3 = 0.0f, n3 = 0.0f;  // SED only
    // prefetched row: h, z, n, old qx, old qy; and the next finished row's sources
    float ph, pz, pn, pqx, pqy;
    float pcomp = 0.0f, prain = 0.0f, pinfil = 0.0f;
    // The lane offsets are made opaque at every use (an empty asm), as for the row
    // blocks' stores: otherwise the compiler hoists one 64-bit per-lane pointer per
    // field out of the loop (16 registers for the eight fields read), spills one, and
    // the reload's wait for every load in flight stalls the next row's prefetch.
    auto ingest = [&](int R) {
      const int64_t o = (int64_t)clampi(R, 0, ny - 1) * P;  // uniform
      const int64_t oy = (int64_t)clampi(R, 0, ny) * P;
      uint32_t kc = cc, kf = cf;
      asm volatile("" : "+v"(kc), "+v"(kf));
      ph = ROWS ? ld_nt(at(h_in + o, kc)) : at(h_in + o, kc);
      pz = at(zf + o, kc);
      pn = at(nf + o, kc);
      pqx = ROWS ? ld_nt(at(qx_in + o, kf)) : at(qx_in + o, kf);
      pqy = ROWS ? ld_nt(at(qy_in + oy, kc)) : at(qy_in + oy, kc);
    };
    auto sources = [&](int i) {
      if (i >= y0 && i < y_end && own) {
        const int64_t o = (int64_t)i * P;  // uniform
        uint32_t ku = cu;
        asm volatile("" : "+v"(ku));
        if (flags & LI_COMP) pcomp = ROWS ? ld_nt(at(hcomp + o, ku)) : at(hcomp + o, ku);
        if (flags & LI_FIELD) prain = at(rainf + o, ku);
        if (flags & LI_INFIL) pinfil = at(infilf + o, ku);
      }
    };
    // row y0-2 only feeds the y-faces above row y0-1: take its inputs, no faces
    ingest(y0 - 2);
    e1 = ph + pz;
    z1 = pz;
    n1 = pn;
    h1 = ph;
    ingest(y0 - 1);
    // Four rows per pass: the window of the previous rows is renamed instead of
    // copied at every row, and the next row's loads are no longer waited for at the
    // loop's back edge (at 4096^2 the row blocks: 200 -> 179 us; by two 186, by six or
    // eight 184)
    _Pragma("unroll 4")
    for (int it = 1; it < H + 4; ++it) {
      const int R = y0 - 2 + it;
      const float h0 = ph, z0 = pz, n0 = pn, qxo = pqx, qyo = pqy;
      if (DRY && R >= y0 && R < y_end)  // wave-uniform
        nz |= __ballot(own && (__float_as_uint(h0) | __float_as_uint(qxo) |
                               __float_as_uint(qyo)) != 0u);
      float comp = pcomp;
      const float rain_v = prain, infil_v = pinfil;
      // the next rows' loads issue at a raised wave priority, so that they leave before
      // the other waves' arithmetic (175 against 177 us at 4096^2; 3, or the row blocks'
      // stores raised too, no better)
      __builtin_amdgcn_s_setprio(2);
      if (it + 1 < H + 4) ingest(R + 1);
      sources(R - 1);  // the row the next iteration finishes
      __builtin_amdgcn_s_setprio(0);

      // x-faces of row R (left face of each column) and y-faces above row R
      const float e0 = h0 + z0;  // compute_eta (:61-69)
      const float el = from_left(e0), zl = from_left(z0), nl = from_left(n0);
      const bool fx_valid = xface_ok && R >= 0 && R < ny;
      const bool fy_valid = col_ok && R >= 1 && R <= ny - 1;
      float fx0 = 0.0f, fy0 = 0.0f;
      if (fx_valid) fx0 = li_face(qxo, el, e0, zl, z0, nl, n0, dx, rdx, k.dt);
      if (fy_valid) fy0 = li_face(qyo, e1, e0, z1, z0, n1, n0, dx, rdx, k.dt);

      // donor factor of row R-1
      const float fx1r = from_right(fx1);
      const float b1 = limit ? donor_beta(fx1, fx1r, fy1, fy0, h1, k.dtdx) : 1.0f;

      // finish row i = R-2
      const int i = R - 2;
      const float fx2r = from_right(fx2);
      const float b2l = from_left(b2), b2r = from_right(b2);
      float e2l = 0.0f, z2l = 0.0f, n2l = 0.0f;
      if (SED) {
        e2l = from_left(e2);
        z2l = from_left(z2);
        n2l = from_left(n2);
      }
      const bool row_live = i >= y0 && i < y_end;  // block-uniform for ROWS
      uint32_t bits = 0;  // what the lane writes (and the debts it leaves)
      if (row_live && own) {
        float qxL = fx2, qxR = fx2r, qyT = fy2, qyB = fy1;
        if (limit) limit_faces(qxL, qxR, qyT, qyB, b2, b2l, b2r, b3, b1);
        const int64_t row = (int64_t)i * P;  // uniform
        const int64_t o = row + c;
        if (SED && (flags & LI_SED)) {
          if (c >= 1)
            bits |= accumulate_face(qxL, e2l, e2, z2l, z2, n2l, n2, glob(a.d50), o - 1, o,
                                    k.dt, glob(a.qs_int_x), glob(a.qs_comp_x), o);
          if (i >= 1)
            bits |= accumulate_face(qyT, e3, e2, z3, z2, n3, n2, glob(a.d50), o - P, o, k.dt,
                                    glob(a.qs_int_y), glob(a.qs_comp_y), o);
        }
        const float h = cell_update(flags, i, c, ny, nx, h2, qxL, qxR, qyT, qyB, k, comp,
                                    rain_v, infil_v, sum);
        bits |= __float_as_uint(h) | __float_as_uint(qxL) | __float_as_uint(qyT);
        if (flags & LI_COMP) bits |= __float_as_uint(comp);
        if (ROWS) {
          const int j = kStripW * wave + lane - 2;  // in [0, BWC)
          const int slot = it % (2 * FL);
          s_row[slot][0][j] = h;
          s_row[slot][1][j] = qxL;
          s_row[slot][2][j] = qyT;
          s_row[slot][3][j] = comp;
        } else {
          if (flags & LI_COMP) at(hcomp + row, cu) = comp;
          at(h_out + row, cu) = h;
          at(qx_out + row, cu) = qxL;
          at(qy_out + row, cu) = qyT;
        }
        if (c == nx - 1) at(qx_out + row, cu + 1) = qxR;
        if (i == ny - 1) at(qy_out + row + P, cu) = qyB;
      }
      if (DRY && row_live) nz |= __ballot(bits != 0u);  // wave-uniform
      if (ROWS && row_live) flush_rows(it, i);
      // slide the window
      if (SED) {
        e3 = e2; z3 = z2; n3 = n2;
        e2 = e1; z2 = z1; n2 = n1;
      }
      e1 = e0; z1 = z0; n1 = n0;
      fx2 = fx1; fx1 = fx0;
      fy2 = fy1; fy1 = fy0;
      h2 = h1; h1 = h0;
      b3 = b2; b2 = b1;
    }
  }
  write_partials<NT>(a, tile, sum, s_red);
  // the tile's flag for the next steps (with rain falling no tile is skipped: 0)
  if (DRY) {
    const bool still = no_rain && !__syncthreads_or(nz != 0);
    if (threadIdx.x == 0) quiet[tile] = still ? 1u : 0u;
  }
}

// ---- family 3: LDS tile of a state with sub-grid channels --------------------------
// The same phases as li_tile with the channel forms of three of them
// (local_inertial.py:617-680): eta through the storage curve, each face updated as a
// floodplain and a channel component recombined into the total, and the limiter
// scaling the total and both components by the total's donor factor.  Continuity,
// sources and sinks read the totals exactly as in the plain step; the bedload
// integral mixes the two components' capacities (sediment.py:276-377); the CFL
// partial is the maximum water-column depth (:487-500).
template <int BX, int BY>
__device__ __forceinline__ void li_chan_tile(const LiArgs &a, const int tile) {
  const DevCtl *ctl = a.ctl;
  if (ctl->active != 1) return;
  const int cur = ctl->cur;
  const float *__restrict__ h_in = a.h[cur];
  float *__restrict__ h_out = a.h[cur ^ 1];
  const float *__restrict__ qxc_in = a.qx_ch[cur];
  const float *__restrict__ qxf_in = a.qx_fp[cur];
  const float *__restrict__ qyc_in = a.qy_ch[cur];
  const float *__restrict__ qyf_in = a.qy_fp[cur];
  float *__restrict__ qx_out = a.qx[cur ^ 1];
  float *__restrict__ qy_out = a.qy[cur ^ 1];
  float *__restrict__ qxc_out = a.qx_ch[cur ^ 1];
  float *__restrict__ qxf_out = a.qx_fp[cur ^ 1];
  float *__restrict__ qyc_out = a.qy_ch[cur ^ 1];
  float *__restrict__ qyf_out = a.qy_fp[cur ^ 1];

  constexpr int EW = BX + 4, EH = BY + 4;  // cells two out
  constexpr int RW = BX + 2, RH = BY + 2;  // cells one out
  constexpr int QXW = BX + 3, QYH = BY + 3;
  __shared__ float s_eta[EH][EW], s_z[EH][EW], s_n[EH][EW];
  __shared__ float s_w[EH][EW], s_d[EH][EW], s_nc[EH][EW];
  __shared__ float s_h[RH][RW], s_beta[RH][RW];
  __shared__ float s_qx[RH][QXW], s_qxc[RH][QXW], s_qxf[RH][QXW];
  __shared__ float s_qy[QYH][RW], s_qyc[QYH][RW], s_qyf[QYH][RW];
  __shared__ double s_red[4 * (kThreads / 64)];

  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int y0 = ty * BY, x0 = tx * BX;
  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int tid = threadIdx.x;
  const int flags = a.flags;
  const StepScalars k = step_scalars(a, ctl);
  const float dx = (float)a.dx64;
  const float rdx = 1.0f / dx;  // IEEE division: RN(1/dx)
  auto cell = [&](int lr, int lc) {
    return ChanCell{s_eta[lr][lc], s_z[lr][lc], s_n[lr][lc], s_w[lr][lc], s_d[lr][lc],
                    s_nc[lr][lc]};
  };

  // ---- phase A: stage the halo'd tile; eta through the storage curve -------------
#pragma unroll
  for (int k0 = 0; k0 < EH * EW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < EH * EW) {
      const int lr = kk / EW, lc = kk - (kk / EW) * EW;
      const int gi = clampi(y0 + lr - 2, 0, ny - 1), gj = clampi(x0 + lc - 2, 0, nx - 1);
      const int64_t o = (int64_t)gi * P + gj;
      const float hv = h_in[o], zv = a.z[o], wv = a.ch_w[o], dv = a.ch_d[o];
      s_eta[lr][lc] = eta_subgrid_u(hv, zv, wv, dv, dx, rdx);  // compute_eta_channels
      s_z[lr][lc] = zv;
      s_n[lr][lc] = a.n[o];
      s_w[lr][lc] = wv;
      s_d[lr][lc] = dv;
      s_nc[lr][lc] = a.ch_n[o];
      if (lr >= 1 && lr <= BY + 2 && lc >= 1 && lc <= BX + 2) s_h[lr - 1][lc - 1] = hv;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < RH * QXW) {
      const int lr = kk / QXW, lf = kk - (kk / QXW) * QXW;
      const int gi = clampi(y0 + lr - 1, 0, ny - 1), gf = clampi(x0 + lf - 1, 0, nx);
      const int64_t o = (int64_t)gi * P + gf;
      s_qxc[lr][lf] = qxc_in[o];
      s_qxf[lr][lf] = qxf_in[o];
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < QYH * RW) {
      const int lf = kk / RW, lc = kk - (kk / RW) * RW;
      const int gf = clampi(y0 + lf - 1, 0, ny), gj = clampi(x0 + lc - 1, 0, nx - 1);
      const int64_t o = (int64_t)gf * P + gj;
      s_qyc[lf][lc] = qyc_in[o];
      s_qyf[lf][lc] = qyf_in[o];
    }
  }
  __syncthreads();

  // ---- phase B: two-component update of every face of the ring-1 cells ----------
  // Boundary and out-of-domain faces (and their components) are 0.
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < RH * QXW) {
      const int lr = kk / QXW, lf = kk - (kk / QXW) * QXW;
      const int r = y0 + lr - 1, f = x0 + lf - 1;
      float q = 0.0f, qc = 0.0f, qf = 0.0f;
      if (r >= 0 && r < ny && f >= 1 && f <= nx - 1) {
        qc = s_qxc[lr][lf];
        qf = s_qxf[lr][lf];
        q = chan_face(cell(lr + 1, lf), cell(lr + 1, lf + 1), qc, qf, dx, rdx, k.dt);
      }
      s_qx[lr][lf] = q;
      s_qxc[lr][lf] = qc;
      s_qxf[lr][lf] = qf;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int kk = k0 + tid;
    if (kk < QYH * RW) {
      const int lf = kk / RW, lc = kk - (kk / RW) * RW;
      const int f = y0 + lf - 1, c = x0 + lc - 1;
      float q = 0.0f, qc = 0.0f, qf = 0.0f;
      if (c >= 0 && c < nx && f >= 1 && f <= ny - 1) {
        qc = s_qyc[lf][lc];
        qf = s_qyf[lf][lc];
        q = chan_face(cell(lf, lc + 1), cell(lf + 1, lc + 1), qc, qf, dx, rdx, k.dt);
      }
      s_qy[lf][lc] = q;
      s_qyc[lf][lc] = qc;
      s_qyf[lf][lc] = qf;
    }
  }
  __syncthreads();

  // ---- phase C: donor factors from the totals ---------------------------------------
  const bool limit = flags & LI_LIMIT;
  if (limit) {
#pragma unroll
    for (int k0 = 0; k0 < RH * RW; k0 += kThreads) {
      const int kk = k0 + tid;
      if (kk < RH * RW) {
        const int lr = kk / RW, lc = kk - (kk / RW) * RW;
        s_beta[lr][lc] = donor_beta(s_qx[lr][lc], s_qx[lr][lc + 1], s_qy[lr][lc],
                                    s_qy[lr + 1][lc], s_h[lr][lc], k.dtdx);
      }
    }
    __syncthreads();
  }

  // ---- phase D: limited faces and components, bedload, continuity, sources ---------
  CellSums sum;
  float m_col = 0.0f;  // reduce_column_max starts at 0
#pragma unroll
  for (int k0 = 0; k0 < BY * BX; k0 += kThreads) {
    const int kk = k0 + tid;
    const int lr = kk / BX, lc = kk - (kk / BX) * BX;
    const int r = y0 + lr, c = x0 + lc;
    if (kk < BY * BX && r < ny && c < nx) {
      float qxL = s_qx[lr + 1][lc + 1], qxR = s_qx[lr + 1][lc + 2];
      float qyT = s_qy[lr + 1][lc + 1], qyB = s_qy[lr + 2][lc + 1];
      // components of this cell's own faces (left x-face, top y-face)
      float cxL = s_qxc[lr + 1][lc + 1], fxL = s_qxf[lr + 1][lc + 1];
      float cyT = s_qyc[lr + 1][lc + 1], fyT = s_qyf[lr + 1][lc + 1];
      if (limit) {
        const float bc = s_beta[lr + 1][lc + 1];
        limit_components(qxL, cxL, fxL, s_beta[lr + 1][lc], bc);
        limit_components(qyT, cyT, fyT, s_beta[lr][lc + 1], bc);
        limit_faces(qxL, qxR, qyT, qyB, bc, s_beta[lr + 1][lc], s_beta[lr + 1][lc + 2],
                    s_beta[lr][lc + 1], s_beta[lr + 2][lc + 1]);
      }
      const int64_t o = (int64_t)r * P + c;
      if (flags & LI_SED) {  // own faces only: left (c >= 1) and top (r >= 1)
        const ChanCell self = cell(lr + 2, lc + 2);
        if (c >= 1) {
          const float qs = combined_capacity(cxL, fxL, cell(lr + 2, lc + 1), self,
                                             0.5f * (a.d50[o - 1] + a.d50[o]), dx, rdx);
          float si = a.qs_int_x[o], sc = a.qs_comp_x[o];
          kahan_add(si, sc, qs * k.dt);
          a.qs_int_x[o] = si;
          a.qs_comp_x[o] = sc;
        }
        if (r >= 1) {
          const float qs = combined_capacity(cyT, fyT, cell(lr + 1, lc + 2), self,
                                             0.5f * (a.d50[o - P] + a.d50[o]), dx, rdx);
          float si = a.qs_int_y[o], sc = a.qs_comp_y[o];
          kahan_add(si, sc, qs * k.dt);
          a.qs_int_y[o] = si;
          a.qs_comp_y[o] = sc;
        }
      }
      float comp = (flags & LI_COMP) ? a.h_comp[o] : 0.0f;
      const float rain_v = (flags & LI_FIELD) ? a.rain[o] : 0.0f;
      const float infil_v = (flags & LI_INFIL) ? a.infil[o] : 0.0f;
      const float h = cell_update(flags, r, c, ny, nx, s_h[lr + 1][lc + 1], qxL, qxR, qyT, qyB,
                                  k, comp, rain_v, infil_v, sum);
      m_col = fmx(m_col, column_depth_u(h, s_w[lr + 2][lc + 2], s_d[lr + 2][lc + 2], dx, rdx));
      if (flags & LI_COMP) a.h_comp[o] = comp;
      h_out[o] = h;
      qx_out[o] = qxL;
      qxc_out[o] = cxL;
      qxf_out[o] = fxL;
      if (c == nx - 1) {  // the closed east edge face: total and components 0
        qx_out[o + 1] = qxR;
        qxc_out[o + 1] = 0.0f;
        qxf_out[o + 1] = 0.0f;
      }
      qy_out[o] = qyT;
      qyc_out[o] = cyT;
      qyf_out[o] = fyT;
      if (r == ny - 1) {
        qy_out[o + P] = qyB;
        qyc_out[o + P] = 0.0f;
        qyf_out[o + P] = 0.0f;
      }
    }
  }
  sum.m_h = m_col;
  write_partials(a, tile, sum, s_red);
}

// One state: one block per tile.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void li_step_kernel(LiArgs a) {
  li_tile<BX, BY>(a, xcd_tile(blockIdx.x, a.n_tiles));
}
// W: the waves per SIMD the register allocator aims for (the bedload form keeps
// its extra window registers at 4)
template <int H, bool SED, int W, int NW = 0, int FL = 1>
__global__ __launch_bounds__(NW > 0 ? 64 * NW : kThreads)
__attribute__((amdgpu_waves_per_eu(SED ? 4 : W))) void li_strip_kernel(LiArgs a) {
  const int tile = xcd_tile(blockIdx.x, a.n_tiles);
  if constexpr (NW > 0 && !SED) {  // row blocks: the limiter's and debt's flags as constants
    if (a.flags & LI_LIMIT) {
      if (a.flags & LI_COMP) li_strip<H, SED, NW, FL, 1, 1>(a, tile);
      else li_strip<H, SED, NW, FL, 1>(a, tile);
      return;
    }
  }
  li_strip<H, SED, NW, FL>(a, tile);
}

// A batch of states in one launch per tile segment (seg 0: members without the
// bedload integral, seg 1: with it, seg 2: wide members without it on the row
// blocks): the segment's flat tile space, mapped back to
// (member, tile) by a table; each member keeps its own clock, dt and ledger.  The
// grid is sized for the batch's capacity; blocks past the segment's live tile count
// return at once.
__device__ __forceinline__ bool batch_tile(const int *__restrict__ tile_member,
                                           const int *__restrict__ tile_first,
                                           const BatchDims *dims, int seg, int &m, int &tile) {
  const int base = dims->seg_first[seg];
  const int total = dims->seg_tiles[seg];
  if ((int)blockIdx.x >= total) return false;
  const int t = xcd_tile(blockIdx.x, total);
  // the member index is block-uniform: its arguments are read with scalar loads
  m = __builtin_amdgcn_readfirstlane(tile_member[base + t]);
  tile = t - tile_first[m];
  return true;
}

template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void li_batch_kernel(const LiArgs *__restrict__ members,
                                                            const int *__restrict__ tile_member,
                                                            const int *__restrict__ tile_first,
                                                            const BatchDims *dims, int seg) {
  int m, tile;
  if (batch_tile(tile_member, tile_first, dims, seg, m, tile)) li_tile<BX, BY>(members[m], tile);
}
template <int H, bool SED, int W, int NW = 0, int FL = 1>
__global__ __launch_bounds__(NW > 0 ? 64 * NW : kThreads)
__attribute__((amdgpu_waves_per_eu(SED ? 4 : W))) void li_strip_batch_kernel(
    const LiArgs *__restrict__ members, const int *__restrict__ tile_member,
    const int *__restrict__ tile_first, const BatchDims *dims, int seg) {
  int m, tile;
  if (batch_tile(tile_member, tile_first, dims, seg, m, tile)) {
    if constexpr (NW > 0 && !SED) {  // as li_strip_kernel
      if (members[m].flags & LI_LIMIT) {
        if (members[m].flags & LI_COMP) li_strip<H, SED, NW, FL, 1, 1>(members[m], tile);
        else li_strip<H, SED, NW, FL, 1>(members[m], tile);
        return;
      }
    }
    li_strip<H, SED, NW, FL>(members[m], tile);
  }
}

// ---- variants (the autotuner picks one per grid shape) -------------------------------
// bx x by is the block's tile of cells.  `kern_sed` / `batch_sed` are the forms used
// when a state carries the bedload integral (the LDS family handles it at run time).
using StepFn = void (*)(LiArgs);
using BatchFn = void (*)(const LiArgs *, const int *, const int *, const BatchDims *, int);
struct LiVariant {
  int bx, by;
  int family;  // 0 LDS tile, 1 register strips, 3 register strips in row blocks
               // with line-aligned stores (2: retired form, two columns per lane)
  StepFn kern, kern_sed;
  BatchFn batch, batch_sed;
  int threads;  // block size
};
#define RBS_LI_TILE(X, Y)                                                                 \
  {X, Y, 0, li_step_kernel<X, Y>, li_step_kernel<X, Y>, li_batch_kernel<X, Y>,            \
   li_batch_kernel<X, Y>, kThreads}
#define RBS_LI_STRIP(H, W)                                                                \
  {kStripW, kStripWaves * (H), 1, li_strip_kernel<H, false, W>, li_strip_kernel<H, true, W>,  \
   li_strip_batch_kernel<H, false, W>, li_strip_batch_kernel<H, true, W>, kThreads}
#define RBS_LI_ROWS(H, W, NW, FL)                                                         \
  {kStripW * (NW), H, 3, li_strip_kernel<H, false, W, NW, FL>,                            \
   li_strip_kernel<H, true, W, NW, FL>, li_strip_batch_kernel<H, false, W, NW, FL>,       \
   li_strip_batch_kernel<H, true, W, NW, FL>, 64 * (NW)}
static const LiVariant kLiVariants[] = {
    RBS_LI_TILE(64, 16),     RBS_LI_TILE(64, 8),      RBS_LI_TILE(64, 32),
    RBS_LI_TILE(128, 8),     RBS_LI_TILE(128, 16),    RBS_LI_TILE(32, 32),
    RBS_LI_STRIP(16, 6),     RBS_LI_STRIP(32, 6),     RBS_LI_STRIP(64, 6),
    RBS_LI_ROWS(16, 6, 8, 2), RBS_LI_ROWS(32, 6, 8, 2), RBS_LI_ROWS(49, 6, 8, 2),
    // 8-row strips: a short serial chain per wave for grids where a few wet tiles
    // among skipped dry ones set the step's time
    RBS_LI_STRIP(8, 6),
};
constexpr int kNumLiVariants = sizeof(kLiVariants) / sizeof(kLiVariants[0]);
constexpr int kDefaultLiVariant = 7;    // strips of 32 rows
constexpr int kRowsVariantForWide = 11;  // row blocks of 49 rows, two rows per barrier

// ---- family 4: register strips of a state with sub-grid channels ----------------------
// li_strip's rolling window with the channel forms: each row's cells carry the
// storage-curve surface and the section (z, n, w, d, n_ch); a face's two components
// travel beside its total through the window, and the finished row limits its own
// faces' components by the same donor factor as their totals.
// NW > 0: row blocks, as for li_strip -- the block's NW strips side by side, each
// finished row's eight outputs staged in LDS and stored as whole lines.
template <int H, bool SED, int NW = 0, int FL = 1>
__device__ __forceinline__ void li_chan_strip(const LiArgs &a, const int tile) {
  static_assert(FL >= 1 && FL <= 4, "one to four rows per barrier");
  constexpr bool ROWS = NW > 0;
  constexpr int BWC = ROWS ? kStripW * NW : 1;
  static_assert(!ROWS || (BWC * 4) % 128 == 0, "row blocks must span whole 128-B lines");
  const DevCtl *ctl = a.ctl;
  if (ctl->active != 1) return;
  const int cur = ctl->cur;
  const float *__restrict__ h_in = a.h[cur];
  float *__restrict__ h_out = a.h[cur ^ 1];
  const float *__restrict__ qxc_in = a.qx_ch[cur];
  const float *__restrict__ qxf_in = a.qx_fp[cur];
  const float *__restrict__ qyc_in = a.qy_ch[cur];
  const float *__restrict__ qyf_in = a.qy_fp[cur];
  float *__restrict__ qx_out = a.qx[cur ^ 1];
  float *__restrict__ qy_out = a.qy[cur ^ 1];
  float *__restrict__ qxc_out = a.qx_ch[cur ^ 1];
  float *__restrict__ qxf_out = a.qx_fp[cur ^ 1];
  float *__restrict__ qyc_out = a.qy_ch[cur ^ 1];
  float *__restrict__ qyf_out = a.qy_fp[cur ^ 1];
  constexpr int NT = ROWS ? 64 * NW : kThreads;
  __shared__ double s_red[4 * (NT / 64)];
  // row blocks: h, qx, qx_ch, qx_fp, qy, qy_ch, qy_fp, Kahan debt; double-buffered
  __shared__ float s_row[ROWS ? 2 * FL : 1][8][BWC];

  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int flags = a.flags;
  const bool limit = flags & LI_LIMIT;
  const StepScalars k = step_scalars(a, ctl);
  const float dx = (float)a.dx64;
  const float rdx = 1.0f / dx;  // IEEE division: RN(1/dx)

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int strip = ROWS ? tx * NW + wave : tx;
  const int c = strip * kStripW - 2 + lane;
  const int y0 = ROWS ? ty * H : (ty * kStripWaves + wave) * H;  // block-uniform for ROWS
  const int y_end = y0 + H < ny ? y0 + H : ny;
  const bool idle = ROWS && strip * kStripW >= nx;
  const int x0 = ROWS ? tx * BWC : 0;
  const uint32_t cc = clampi(c, 0, nx - 1);
  const uint32_t cf = clampi(c, 0, nx);
  const uint32_t cu = (uint32_t)c;  // used only where 0 <= c < nx
  const bool col_ok = c >= 0 && c < nx;
  const bool xface_ok = c >= 1 && c <= nx - 1;
  const bool own = lane >= 2 && lane < 2 + kStripW && col_ok;

  // the ring of 2 FL LDS rows and one barrier per FL rows, as in li_strip
  auto flush_rows = [&](int it, int i) {
    const int off = i - y0;
    if ((off % FL) != FL - 1 && i != y_end - 1) return;  // block-uniform
    __syncthreads();  // the block's rows: every wave reaches it at the same iterations
    float *const outs[8] = {h_out, qx_out, qxc_out, qxf_out, qy_out, qyc_out, qyf_out,
                            a.h_comp};
#pragma unroll
    for (int back = FL - 1; back >= 0; --back) {
      if (back > off % FL) continue;  // a short last group
      const int slot = (it - back) % (2 * FL);  // it >= 4 for a live row
      const int64_t row = (int64_t)(i - back) * P;  // uniform
#pragma unroll
      for (int part = 0; part < (BWC + 63) / 64; part += (NW > 0 ? NW : 1)) {
        const int j = 64 * (part + wave) + lane;
        if (j < BWC && x0 + j < nx) {
          uint32_t col = (uint32_t)(x0 + j);
          asm volatile("" : "+v"(col));  // keep the offsets 32-bit (no hoisted pointers)
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (q < 7 || (flags & LI_COMP)) st_nt(at(outs[q] + row, col), s_row[slot][q][j]);
        }
      }
    }
  };

  CellSums sum;
  float m_col = 0.0f;  // reduce_column_max starts at 0
  if (ROWS && idle && y0 < ny) {
    for (int it = 1; it < H + 4; ++it) {
      const int i = y0 - 4 + it;
      if (i >= y0 && i < y_end) flush_rows(it, i);
    }
  } else if (y0 < ny) {  // wave-uniform (block-uniform for row blocks)
    // window: the cells of rows R-1 (c1), R-2 and R-3 (c2, c3: bedload only; else
    // just c2's width and depth for the column depth), the totals and components of
    // the faces of rows R-1 and R-2, old depths of rows R-1 and R-2, donor factors of
    // rows R-2 and R-3
    ChanCell c1{}, c2{}, c3{};
    float fx1 = 0.0f, fx2 = 0.0f, fy1 = 0.0f, fy2 = 0.0f;
    float fxc1 = 0.0f, fxf1 = 0.0f, fxc2 = 0.0f, fxf2 = 0.0f;
    float fyc1 = 0.0f, fyf1 = 0.0f, fyc2 = 0.0f, fyf2 = 0.0f;
    float h1 = 0.0f, h2 = 0.0f, b2 = 1.0f, b3 = 1.0f;
    // prefetched row: h, z, n, w, d, n_ch, old x-face and y-face components
    float ph, pz, pn, pw, pd, pnc, pxc, pxf, pyc, pyf;
    float pcomp = 0.0f, prain = 0.0f, pinfil = 0.0f;
    auto ingest = [&](int R) {
      const int64_t o = (int64_t)clampi(R, 0, ny - 1) * P;  // uniform
      const int64_t oy = (int64_t)clampi(R, 0, ny) * P;
      ph = at(h_in + o, cc);
      pz = at(a.z + o, cc);
      pn = at(a.n + o, cc);
      pw = at(a.ch_w + o, cc);
      pd = at(a.ch_d + o, cc);
      pnc = at(a.ch_n + o, cc);
      pxc = at(qxc_in + o, cf);
      pxf = at(qxf_in + o, cf);
      pyc = at(qyc_in + oy, cc);
      pyf = at(qyf_in + oy, cc);
    };
    auto sources = [&](int i) {
      if (i >= y0 && i < y_end && own) {
        const int64_t o = (int64_t)i * P;
        if (flags & LI_COMP) pcomp = ROWS ? ld_nt(at(a.h_comp + o, cu)) : at(a.h_comp + o, cu);
        if (flags & LI_FIELD) prain = ROWS ? ld_nt(at(a.rain + o, cu)) : at(a.rain + o, cu);
        if (flags & LI_INFIL) pinfil = ROWS ? ld_nt(at(a.infil + o, cu)) : at(a.infil + o, cu);
      }
    };
    ingest(y0 - 2);
    c1 = ChanCell{eta_subgrid_u(ph, pz, pw, pd, dx, rdx), pz, pn, pw, pd, pnc};
    h1 = ph;
    ingest(y0 - 1);
    // two rows per pass: fewer spilled registers than four (the window renamed every
    // second row; at 4096^2 349 -> 329 us, four rows per pass spill 60 B a lane, three
    // 413 us)
    _Pragma("unroll 2")
    for (int it = 1; it < H + 4; ++it) {
      const int R = y0 - 2 + it;
      const ChanCell c0{eta_subgrid_u(ph, pz, pw, pd, dx, rdx), pz, pn, pw, pd, pnc};
      const float h0 = ph;
      float qxc = pxc, qxf = pxf, qyc = pyc, qyf = pyf;
      float comp = pcomp;
      const float rain_v = prain, infil_v = pinfil;
      if (it + 1 < H + 4) ingest(R + 1);
      sources(R - 1);

      // x-faces of row R and y-faces above row R, both components
      const ChanCell cl{from_left(c0.eta), from_left(c0.z), from_left(c0.n),
                        from_left(c0.w), from_left(c0.d), from_left(c0.nc)};
      const bool fx_valid = xface_ok && R >= 0 && R < ny;
      const bool fy_valid = col_ok && R >= 1 && R <= ny - 1;
      float fx0 = 0.0f, fy0 = 0.0f;
      if (fx_valid) fx0 = chan_face(cl, c0, qxc, qxf, dx, rdx, k.dt);
      else qxc = qxf = 0.0f;
      if (fy_valid) fy0 = chan_face(c1, c0, qyc, qyf, dx, rdx, k.dt);
      else qyc = qyf = 0.0f;

      // donor factor of row R-1 from the totals
      const float fx1r = from_right(fx1);
      const float b1 = limit ? donor_beta(fx1, fx1r, fy1, fy0, h1, k.dtdx) : 1.0f;

      // finish row i = R-2
      const int i = R - 2;
      const float fx2r = from_right(fx2);
      const float b2l = from_left(b2), b2r = from_right(b2);
      ChanCell c2l{};
      if (SED) c2l = ChanCell{from_left(c2.eta), from_left(c2.z), from_left(c2.n),
                              from_left(c2.w), from_left(c2.d), from_left(c2.nc)};
      const bool row_live = i >= y0 && i < y_end;  // block-uniform for ROWS
      if (row_live && own) {
        float qxL = fx2, qxR = fx2r, qyT = fy2, qyB = fy1;
        float cxL = fxc2, fxL = fxf2, cyT = fyc2, fyT = fyf2;
        if (limit) {
          limit_components(qxL, cxL, fxL, b2l, b2);
          limit_components(qyT, cyT, fyT, b3, b2);
          limit_faces(qxL, qxR, qyT, qyB, b2, b2l, b2r, b3, b1);
        }
        const int64_t row = (int64_t)i * P;  // uniform
        const int64_t o = row + c;
        if (SED && (flags & LI_SED)) {
          if (c >= 1) {
            const float qs = combined_capacity(cxL, fxL, c2l, c2,
                                               0.5f * (a.d50[o - 1] + a.d50[o]), dx, rdx);
            float si = a.qs_int_x[o], sc = a.qs_comp_x[o];
            kahan_add(si, sc, qs * k.dt);
            a.qs_int_x[o] = si;
            a.qs_comp_x[o] = sc;
          }
          if (i >= 1) {
            const float qs = combined_capacity(cyT, fyT, c3, c2,
                                               0.5f * (a.d50[o - P] + a.d50[o]), dx, rdx);
            float si = a.qs_int_y[o], sc = a.qs_comp_y[o];
            kahan_add(si, sc, qs * k.dt);
            a.qs_int_y[o] = si;
            a.qs_comp_y[o] = sc;
          }
        }
        const float h = cell_update(flags, i, c, ny, nx, h2, qxL, qxR, qyT, qyB, k, comp,
                                    rain_v, infil_v, sum);
        m_col = fmx(m_col, column_depth_u(h, c2.w, c2.d, dx, rdx));
        if (ROWS) {
          const int j = kStripW * wave + lane - 2;  // in [0, BWC)
          const float v[8] = {h, qxL, cxL, fxL, qyT, cyT, fyT, comp};
#pragma unroll
          for (int q = 0; q < 8; ++q) s_row[it % (2 * FL)][q][j] = v[q];
        } else {
          if (flags & LI_COMP) at(a.h_comp + row, cu) = comp;
          at(h_out + row, cu) = h;
          at(qx_out + row, cu) = qxL;
          at(qxc_out + row, cu) = cxL;
          at(qxf_out + row, cu) = fxL;
          at(qy_out + row, cu) = qyT;
          at(qyc_out + row, cu) = cyT;
          at(qyf_out + row, cu) = fyT;
        }
        if (c == nx - 1) {  // the closed east edge face: total and components 0
          at(qx_out + row, cu + 1) = qxR;
          at(qxc_out + row, cu + 1) = 0.0f;
          at(qxf_out + row, cu + 1) = 0.0f;
        }
        if (i == ny - 1) {
          at(qy_out + row + P, cu) = qyB;
          at(qyc_out + row + P, cu) = 0.0f;
          at(qyf_out + row + P, cu) = 0.0f;
        }
      }
      if (ROWS && row_live) flush_rows(it, i);
      // slide the window
      if (SED) c3 = c2;
      c2 = c1;
      c1 = c0;
      fx2 = fx1; fx1 = fx0;
      fy2 = fy1; fy1 = fy0;
      fxc2 = fxc1; fxc1 = qxc;
      fxf2 = fxf1; fxf1 = qxf;
      fyc2 = fyc1; fyc1 = qyc;
      fyf2 = fyf1; fyf1 = qyf;
      h2 = h1; h1 = h0;
      b3 = b2; b2 = b1;
    }
  }
  sum.m_h = m_col;
  write_partials<NT>(a, tile, sum, s_red);
}

template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void li_chan_kernel(LiArgs a) {
  li_chan_tile<BX, BY>(a, xcd_tile(blockIdx.x, a.n_tiles));
}
// W: the waves per SIMD the register allocator aims for; the bedload form keeps its
// two extra rows of cells at 3 (fewer registers would spill)
template <int H, bool SED, int W, int NW = 0, int FL = 1>
__global__ __launch_bounds__(NW > 0 ? 64 * NW : kThreads)
__attribute__((amdgpu_waves_per_eu(SED ? 3 : W))) void li_chan_strip_kernel(LiArgs a) {
  li_chan_strip<H, SED, NW, FL>(a, xcd_tile(blockIdx.x, a.n_tiles));
}
struct ChanVariant {
  int bx, by;
  StepFn kern, kern_sed;
  int threads;  // block size
};
#define RBS_CHAN_TILE(X, Y) {X, Y, li_chan_kernel<X, Y>, li_chan_kernel<X, Y>, kThreads}
#define RBS_CHAN_STRIP(H, W)                                                    \
  {kStripW, kStripWaves * (H), li_chan_strip_kernel<H, false, W>,               \
   li_chan_strip_kernel<H, true, W>, kThreads}
#define RBS_CHAN_ROWS(H, W, NW, FL)                                             \
  {kStripW * (NW), H, li_chan_strip_kernel<H, false, W, NW, FL>,                \
   li_chan_strip_kernel<H, true, W, NW, FL>, 64 * (NW)}
static const ChanVariant kChanVariants[] = {
    RBS_CHAN_TILE(64, 8), RBS_CHAN_TILE(32, 16), RBS_CHAN_TILE(32, 8), RBS_CHAN_TILE(16, 16),
    RBS_CHAN_STRIP(8, 4), RBS_CHAN_STRIP(16, 4), RBS_CHAN_STRIP(32, 4), RBS_CHAN_STRIP(4, 4),
    RBS_CHAN_ROWS(8, 4, 8, 2), RBS_CHAN_ROWS(16, 4, 8, 2), RBS_CHAN_ROWS(32, 4, 8, 2),
};
constexpr int kNumChanVariants = sizeof(kChanVariants) / sizeof(kChanVariants[0]);
constexpr int kDefaultChanVariant = 9;  // row blocks of 16 rows: fastest measured at 4096^2
static const ChanVariant &chan_variant_at(int v) {
  if (v < 0 || v >= kNumChanVariants) v = kDefaultChanVariant;
  return kChanVariants[v];
}
int chan_variant_count() { return kNumChanVariants; }
int chan_default_variant() { return kDefaultChanVariant; }
void chan_variant_shape(int v, int *bx, int *by) {
  *bx = chan_variant_at(v).bx;
  *by = chan_variant_at(v).by;
}
int chan_num_blocks(const rbs_state *s) {
  const ChanVariant &v = chan_variant_at(s->chan_variant);
  return ((s->nx + v.bx - 1) / v.bx) * ((s->ny + v.by - 1) / v.by);
}

static const LiVariant &variant_at(int v) {
  if (v < 0 || v >= kNumLiVariants) v = kDefaultLiVariant;
  return kLiVariants[v];
}

int li_variant_count() { return kNumLiVariants; }
int li_default_variant() { return kDefaultLiVariant; }
// A new state's variant: the row blocks of 49 rows on grids at least 2048 cells wide
// (where a ragged last block idles few of their waves; at 4096^2 they re-read the
// fewest halo rows, 205 us against 214 us for 16-row blocks and 222 us for the
// default, measured with 50-row blocks), the library default otherwise.  49 rows:
// at 4096^2 the 9 x 84 = 756 blocks of eight waves all fit on the GPU at once (six
// waves per SIMD, 768 blocks), three on most CUs, and each of them walks 52 rows
// instead of 50-row blocks' 53 (1.5 us less per step); 48 rows give 774 blocks, and
// the six left over run after the rest (199 us).
int li_initial_variant(int nx) { return nx >= 2048 ? kRowsVariantForWide : kDefaultLiVariant; }
int li_wide_batch_variant() { return kRowsVariantForWide; }
// a batch's members without the bedload integral: the row blocks too.  The 256
// scenarios of sides 259-2027 took 5.9 s batched on them against 6.8 s on 32-row
// strips (6.1 s with the row blocks from 1024 columns up, 5.95 s from 512): the
// narrow members idle most of a 480-column block's waves, but their time is small
// beside the large members', whose halo re-reads the tall row blocks cut.
int li_default_batch_variant() { return kRowsVariantForWide; }
int li_wide_batch_min_nx() { return 2048; }
// batches of bedload members: strips of 8 rows.  Strips beat the LDS tiles inside a
// batch, where the members' tiles fill the GPU whatever their shapes; the bedload
// members are alluvial reaches, mostly dry floodplain, whose few wet tiles set the
// step's time, and 8 rows halve a wet tile's serial chain against 16.
int li_default_sed_variant() { return 12; }
void li_variant_shape(int v, int *bx, int *by, int *family) {
  const LiVariant &x = variant_at(v);
  *bx = x.bx;
  *by = x.by;
  *family = x.family;
}

size_t li_args_size() { return sizeof(LiArgs); }

int li_batch_launch(int variant, int seg, const LiArgs *members, const int *tile_member,
                    const int *tile_first, const BatchDims *dims, int grid, hipStream_t stream) {
  const LiVariant &v = variant_at(variant);
  hipLaunchKernelGGL(seg == 1 ? v.batch_sed : v.batch, dim3(grid), dim3(v.threads), 0, stream,
                     members, tile_member, tile_first, dims, seg);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

int li_num_blocks(const rbs_state *s) {
  if (s->armed & RBS_ARM_CHAN) return chan_num_blocks(s);
  const LiVariant &v = variant_at(s->li_variant);
  const int tx = (s->nx + v.bx - 1) / v.bx, ty = (s->ny + v.by - 1) / v.by;
  return tx * ty;
}

int li_launch_step(rbs_state *s, bool, double, double, int limit, double, hipEvent_t ev_start,
                   hipEvent_t ev_stop) {
  const LiVariant &v = variant_at(s->li_variant);
  LiArgs a{};
  li_fill_args(s, limit, &a);
  if (a.n_tiles > s->part_cap) {
    set_error("partial buffer too small (%d < %d)", s->part_cap, a.n_tiles);
    return RBS_E_STATE;
  }
  StepFn fn = (a.flags & LI_SED) ? v.kern_sed : v.kern;
  int threads = v.threads;
  if (a.flags & LI_CHAN) {
    const ChanVariant &cv = chan_variant_at(s->chan_variant);
    fn = (a.flags & LI_SED) ? cv.kern_sed : cv.kern;
    threads = cv.threads;
  }
  if (ev_start || ev_stop)
    hipExtLaunchKernelGGL(fn, dim3(a.n_tiles), dim3(threads), 0, s->stream, ev_start, ev_stop, 0, a);
  else
    hipLaunchKernelGGL(fn, dim3(a.n_tiles), dim3(threads), 0, s->stream, a);
  RBS_HIP(hipGetLastError());
  s->cur ^= 1;
  return RBS_OK;
}

void li_fill_args(const rbs_state *s, int limit, LiArgs *out) {
  const LiVariant &v = variant_at(s->li_variant);
  LiArgs a{};
  for (int k = 0; k < 2; ++k) {
    a.h[k] = s->h[k];
    a.qx[k] = s->qx[k];
    a.qy[k] = s->qy[k];
  }
  a.z = s->z;
  a.n = s->n;
  a.h_comp = s->h_comp;
  a.rain = s->rain;
  a.infil = s->infil;
  a.d50 = s->d50;
  a.qs_int_x = s->qs_int_x;
  a.qs_comp_x = s->qs_comp_x;
  a.qs_int_y = s->qs_int_y;
  a.qs_comp_y = s->qs_comp_y;
  a.ch_w = s->ch_w;
  a.ch_d = s->ch_d;
  a.ch_n = s->ch_n;
  for (int k = 0; k < 2; ++k) {
    a.qx_ch[k] = s->qx_ch[k];
    a.qx_fp[k] = s->qx_fp[k];
    a.qy_ch[k] = s->qy_ch[k];
    a.qy_fp[k] = s->qy_fp[k];
  }
  a.ctl = s->ctl;
  a.part = s->part;
  a.quiet = s->quiet;
  a.dx64 = s->dx;
  a.pitch = s->pitch;
  a.ny = s->ny;
  a.nx = s->nx;
  a.tiles_x = (s->nx + v.bx - 1) / v.bx;
  if (s->armed & RBS_ARM_CHAN) {
    const ChanVariant &cv = chan_variant_at(s->chan_variant);
    a.tiles_x = (s->nx + cv.bx - 1) / cv.bx;
  }
  a.n_tiles = li_num_blocks(s);
  int f = 0;
  if (limit) f |= LI_LIMIT;
  if (s->armed & RBS_ARM_COMP) f |= LI_COMP;
  if (s->armed & RBS_ARM_RAIN) f |= LI_FIELD;
  if (s->armed & RBS_ARM_INFIL) f |= LI_INFIL;
  if (s->armed & RBS_ARM_SED) f |= LI_SED;
  if (s->armed & RBS_ARM_CHAN) f |= LI_CHAN;
  if (s->bc[RBS_EDGE_WEST] == RBS_BC_OPEN) f |= LI_OPEN_W;
  if (s->bc[RBS_EDGE_EAST] == RBS_BC_OPEN) f |= LI_OPEN_E;
  if (s->bc[RBS_EDGE_NORTH] == RBS_BC_OPEN) f |= LI_OPEN_N;
  if (s->bc[RBS_EDGE_SOUTH] == RBS_BC_OPEN) f |= LI_OPEN_S;
  a.flags = f;
  *out = a;
}

}  // namespace rbs
