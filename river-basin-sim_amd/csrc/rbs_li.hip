// rbs_li.hip -- fused local-inertial time step for gfx950.
//
// One launch performs the whole of local_inertial.step (local_inertial.py:547-727)
// for a tile of BY x BX cells:
//   eta = h + z                                    (:61-69)
//   interior face update + dry guard + friction    (:72-129, friction.py:104-117)
//   closed edge faces = 0                          (boundaries.py:73-91)
//   donor outflow limiter beta, limit faces        (:197-263)
//   bedload transport integral (sediment armed)    (sediment.py:218-273)
//   continuity + uniform rain                      (:132-151)
//   Kahan areal sources / plain rain field         (sources.py:293-329, :154-167)
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
  const DevCtl *ctl;
  Partials part;
  double dx64;
  int64_t pitch;
  int32_t ny, nx;
  int32_t tiles_x, n_tiles;
  int32_t flags;
};

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

  const double dt64 = ctl->dt;
  const float dt = (float)dt64;
  const float dx = (float)a.dx64;
  const float dtdx = dt / dx;
  const float rain = ctl->rain;
  const float rain_scale = ctl->rain_scale;

  // ---- phase A: stage the halo'd tile in LDS (clamped reads stay in bounds) ----
#pragma unroll
  for (int k0 = 0; k0 < EH * EW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < EH * EW) {
      const int lr = k / EW, lc = k - (k / EW) * EW;
      const int gi = clampi(y0 + lr - 2, 0, ny - 1), gj = clampi(x0 + lc - 2, 0, nx - 1);
      const int64_t o = (int64_t)gi * P + gj;
      const float hv = h_in[o], zv = a.z[o];
      s_eta[lr][lc] = hv + zv;  // compute_eta (:61-69)
      s_z[lr][lc] = zv;
      s_n[lr][lc] = a.n[o];
      if (lr >= 1 && lr <= BY + 2 && lc >= 1 && lc <= BX + 2) s_h[lr - 1][lc - 1] = hv;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < RH * QXW) {
      const int lr = k / QXW, lf = k - (k / QXW) * QXW;
      const int gi = clampi(y0 + lr - 1, 0, ny - 1), gf = clampi(x0 + lf - 1, 0, nx);
      s_qx[lr][lf] = qx_in[(int64_t)gi * P + gf];
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < QYH * RW) {
      const int lf = k / RW, lc = k - (k / RW) * RW;
      const int gf = clampi(y0 + lf - 1, 0, ny), gj = clampi(x0 + lc - 1, 0, nx - 1);
      s_qy[lf][lc] = qy_in[(int64_t)gf * P + gj];
    }
  }
  __syncthreads();

  // ---- phase B: face update of every face of the ring-1 cells -------------------
  // Boundary and out-of-domain faces become 0: the closed BC (boundaries.py:87-91)
  // and the reference's "interior faces only" launch ranges (:599-612).
#pragma unroll
  for (int k0 = 0; k0 < RH * QXW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < RH * QXW) {
      const int lr = k / QXW, lf = k - (k / QXW) * QXW;
      const int r = y0 + lr - 1, f = x0 + lf - 1;
      float q = 0.0f;
      if (r >= 0 && r < ny && f >= 1 && f <= nx - 1)
        q = li_face(s_qx[lr][lf], s_eta[lr + 1][lf], s_eta[lr + 1][lf + 1], s_z[lr + 1][lf],
                    s_z[lr + 1][lf + 1], s_n[lr + 1][lf], s_n[lr + 1][lf + 1], dx, dt);
      s_qx[lr][lf] = q;
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < QYH * RW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < QYH * RW) {
      const int lf = k / RW, lc = k - (k / RW) * RW;
      const int f = y0 + lf - 1, c = x0 + lc - 1;
      float q = 0.0f;
      if (c >= 0 && c < nx && f >= 1 && f <= ny - 1)
        q = li_face(s_qy[lf][lc], s_eta[lf][lc + 1], s_eta[lf + 1][lc + 1], s_z[lf][lc + 1],
                    s_z[lf + 1][lc + 1], s_n[lf][lc + 1], s_n[lf + 1][lc + 1], dx, dt);
      s_qy[lf][lc] = q;
    }
  }
  __syncthreads();

  // ---- phase C: donor outflow limiter beta of the ring-1 cells (:197-239) ---------
  const bool limit = flags & LI_LIMIT;
  if (limit) {
#pragma unroll
    for (int k0 = 0; k0 < RH * RW; k0 += kThreads) {
      const int k = k0 + tid;
      if (k < RH * RW) {
        const int lr = k / RW, lc = k - (k / RW) * RW;
        float q_out = 0.0f;
        const float ql = s_qx[lr][lc];
        if (ql < 0.0f) q_out = q_out - ql;
        const float qr = s_qx[lr][lc + 1];
        if (qr > 0.0f) q_out = q_out + qr;
        const float qt = s_qy[lr][lc];
        if (qt < 0.0f) q_out = q_out - qt;
        const float qb = s_qy[lr + 1][lc];
        if (qb > 0.0f) q_out = q_out + qb;
        const float out_depth = dtdx * q_out;
        s_beta[lr][lc] = out_depth > 0.0f ? fclamp(s_h[lr][lc] / out_depth, 0.0f, 1.0f) : 1.0f;
      }
    }
    __syncthreads();
  }

  // ---- phase D: limited faces, continuity, sources, sinks, edge drains ------------
  float m_h = 0.0f;  // h_max starts at 0 (local_inertial.py:527)
  double s_vol = 0.0, s_inf = 0.0, s_edge = 0.0;
  const float dxdt = (float)(dt64 / a.dx64);  // boundaries.py:179 (f64 quotient)
  const float rain_cont = (flags & LI_COMP) ? 0.0f : rain;
#pragma unroll
  for (int k0 = 0; k0 < BY * BX; k0 += kThreads) {
    const int k = k0 + tid;
    const int lr = k / BX, lc = k - (k / BX) * BX;
    const int r = y0 + lr, c = x0 + lc;
    if (k < BY * BX && r < ny && c < nx) {
      float qxL = s_qx[lr + 1][lc + 1], qxR = s_qx[lr + 1][lc + 2];
      float qyT = s_qy[lr + 1][lc + 1], qyB = s_qy[lr + 2][lc + 1];
      if (limit) {  // limit_qx / limit_qy (:242-263): scale by the donor's beta
        const float bc = s_beta[lr + 1][lc + 1];
        if (qxL > 0.0f) qxL = qxL * s_beta[lr + 1][lc];
        else if (qxL < 0.0f) qxL = qxL * bc;
        if (qxR > 0.0f) qxR = qxR * bc;
        else if (qxR < 0.0f) qxR = qxR * s_beta[lr + 1][lc + 2];
        if (qyT > 0.0f) qyT = qyT * s_beta[lr][lc + 1];
        else if (qyT < 0.0f) qyT = qyT * bc;
        if (qyB > 0.0f) qyB = qyB * bc;
        else if (qyB < 0.0f) qyB = qyB * s_beta[lr + 2][lc + 1];
      }
      const int64_t o = (int64_t)r * P + c;
      if (flags & LI_SED) {  // accumulate_qs_x/y (sediment.py:218-273), own faces only
        if (c >= 1) {
          const float eta_max = fmx(s_eta[lr + 2][lc + 1], s_eta[lr + 2][lc + 2]);
          const float z_bank = fmx(s_z[lr + 2][lc + 1], s_z[lr + 2][lc + 2]);
          const float h_face = eta_max - z_bank;
          const float qs = face_capacity(qxL, h_face, h_face,
                                         0.5f * (s_n[lr + 2][lc + 1] + s_n[lr + 2][lc + 2]),
                                         0.5f * (a.d50[o - 1] + a.d50[o]));
          float si = a.qs_int_x[o], sc = a.qs_comp_x[o];
          kahan_add(si, sc, qs * dt);
          a.qs_int_x[o] = si;
          a.qs_comp_x[o] = sc;
        }
        if (r >= 1) {
          const float eta_max = fmx(s_eta[lr + 1][lc + 2], s_eta[lr + 2][lc + 2]);
          const float z_bank = fmx(s_z[lr + 1][lc + 2], s_z[lr + 2][lc + 2]);
          const float h_face = eta_max - z_bank;
          const float qs = face_capacity(qyT, h_face, h_face,
                                         0.5f * (s_n[lr + 1][lc + 2] + s_n[lr + 2][lc + 2]),
                                         0.5f * (a.d50[o - P] + a.d50[o]));
          float si = a.qs_int_y[o], sc = a.qs_comp_y[o];
          kahan_add(si, sc, qs * dt);
          a.qs_int_y[o] = si;
          a.qs_comp_y[o] = sc;
        }
      }
      // update_h (:132-151)
      const float net = qxL - qxR + qyT - qyB;
      float h = s_h[lr + 1][lc + 1] + dtdx * net + rain_cont * dt;
      if (flags & LI_COMP) {
        float cp = a.h_comp[o];
        if (!(flags & LI_FIELD)) kahan_add(h, cp, rain * dt);          // add_uniform_rain_c
        else kahan_add(h, cp, a.rain[o] * rain_scale * dt);            // add_rain_field_c
        a.h_comp[o] = cp;
      } else if (flags & LI_FIELD) {
        h = h + a.rain[o] * rain_scale * dt;                           // apply_rain_field
      }
      if (flags & LI_INFIL) {  // apply_infiltration (:170-194)
        const float avail = h;
        float inf = a.infil[o] * dt;
        if (inf > avail) inf = avail;
        if (inf < 0.0f) inf = 0.0f;
        h = avail - inf;
        s_inf += (double)avail - (double)h;
      }
      // open edges in the reference's launch order (boundaries.py:181-208)
      if ((flags & LI_OPEN_W) && c == 0 && nx > 1) {
        const float av = h;
        h = drain(av, fmx(-qxR, 0.0f), dxdt);
        s_edge += (double)av - (double)h;
      }
      if ((flags & LI_OPEN_E) && c == nx - 1 && nx > 1) {
        const float av = h;
        h = drain(av, fmx(qxL, 0.0f), dxdt);
        s_edge += (double)av - (double)h;
      }
      if ((flags & LI_OPEN_N) && r == 0 && ny > 1) {
        const float av = h;
        h = drain(av, fmx(-qyB, 0.0f), dxdt);
        s_edge += (double)av - (double)h;
      }
      if ((flags & LI_OPEN_S) && r == ny - 1 && ny > 1) {
        const float av = h;
        h = drain(av, fmx(qyT, 0.0f), dxdt);
        s_edge += (double)av - (double)h;
      }
      h_out[o] = h;
      qx_out[o] = qxL;
      if (c == nx - 1) qx_out[o + 1] = qxR;
      qy_out[o] = qyT;
      if (r == ny - 1) qy_out[o + P] = qyB;
      m_h = fmx(m_h, h);
      s_vol += (double)h;
    }
  }

  block_reduce4(m_h, s_vol, s_inf, s_edge, s_red);
  if (tid == 0) {
    a.part.speed_max[tile] = m_h;
    a.part.vol[tile] = s_vol;
    a.part.infil[tile] = s_inf;
    a.part.edge[tile] = s_edge;
  }
}

// One state: one block per tile.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void li_step_kernel(LiArgs a) {
  li_tile<BX, BY>(a, xcd_tile(blockIdx.x, a.n_tiles));
}

// A batch of states in one launch: the flat tile space of all members, mapped back
// to (member, tile) by a table; each member keeps its own clock, dt and ledger.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void li_batch_kernel(const LiArgs *__restrict__ members,
                                                            const int *__restrict__ tile_member,
                                                            const int *__restrict__ tile_first,
                                                            int total) {
  const int t = xcd_tile(blockIdx.x, total);
  const int m = tile_member[t];
  const LiArgs a = members[m];
  li_tile<BX, BY>(a, t - tile_first[m]);
}

// ---- tile variants (the autotuner picks one per grid shape) ----------------------
struct LiVariant {
  int bx, by;
  void (*kern)(LiArgs);
  void (*batch)(const LiArgs *, const int *, const int *, int);
};
#define RBS_LI_VARIANT(X, Y) {X, Y, li_step_kernel<X, Y>, li_batch_kernel<X, Y>}
static const LiVariant kLiVariants[] = {
    RBS_LI_VARIANT(64, 16), RBS_LI_VARIANT(64, 8),   RBS_LI_VARIANT(64, 32),
    RBS_LI_VARIANT(128, 8), RBS_LI_VARIANT(128, 16), RBS_LI_VARIANT(32, 32),
};
constexpr int kNumLiVariants = sizeof(kLiVariants) / sizeof(kLiVariants[0]);

int li_variant_count() { return kNumLiVariants; }

size_t li_args_size() { return sizeof(LiArgs); }

int li_batch_launch(int variant, const LiArgs *members, const int *tile_member,
                    const int *tile_first, int total, hipStream_t stream) {
  if (variant < 0 || variant >= kNumLiVariants) variant = 0;
  hipLaunchKernelGGL(kLiVariants[variant].batch, dim3(total), dim3(kThreads), 0, stream, members,
                     tile_member, tile_first, total);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

static const LiVariant &pick(const rbs_state *s) {
  int v = s->li_variant;
  if (v < 0 || v >= kNumLiVariants) v = 0;
  return kLiVariants[v];
}

int li_num_blocks(const rbs_state *s) {
  const LiVariant &v = pick(s);
  const int tx = (s->nx + v.bx - 1) / v.bx, ty = (s->ny + v.by - 1) / v.by;
  return tx * ty;
}

void li_fill_args(const rbs_state *s, int limit, LiArgs *out);

int li_launch_step(rbs_state *s, bool, double, double, int limit, double) {
  const LiVariant &v = pick(s);
  LiArgs a{};
  li_fill_args(s, limit, &a);
  if (a.n_tiles > s->part_cap) {
    set_error("partial buffer too small (%d < %d)", s->part_cap, a.n_tiles);
    return RBS_E_STATE;
  }
  hipLaunchKernelGGL(v.kern, dim3(a.n_tiles), dim3(kThreads), 0, s->stream, a);
  RBS_HIP(hipGetLastError());
  s->cur ^= 1;
  return RBS_OK;
}

void li_fill_args(const rbs_state *s, int limit, LiArgs *out) {
  const LiVariant &v = pick(s);
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
  a.ctl = s->ctl;
  a.part = s->part;
  a.dx64 = s->dx;
  a.pitch = s->pitch;
  a.ny = s->ny;
  a.nx = s->nx;
  a.tiles_x = (s->nx + v.bx - 1) / v.bx;
  a.n_tiles = li_num_blocks(s);
  int f = 0;
  if (limit) f |= LI_LIMIT;
  if (s->armed & RBS_ARM_COMP) f |= LI_COMP;
  if (s->armed & RBS_ARM_RAIN) f |= LI_FIELD;
  if (s->armed & RBS_ARM_INFIL) f |= LI_INFIL;
  if (s->armed & RBS_ARM_SED) f |= LI_SED;
  if (s->bc[RBS_EDGE_WEST] == RBS_BC_OPEN) f |= LI_OPEN_W;
  if (s->bc[RBS_EDGE_EAST] == RBS_BC_OPEN) f |= LI_OPEN_E;
  if (s->bc[RBS_EDGE_NORTH] == RBS_BC_OPEN) f |= LI_OPEN_N;
  if (s->bc[RBS_EDGE_SOUTH] == RBS_BC_OPEN) f |= LI_OPEN_S;
  a.flags = f;
  *out = a;
}

}  // namespace rbs
