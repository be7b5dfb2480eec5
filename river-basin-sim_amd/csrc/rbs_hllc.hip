// rbs_hllc.hip -- well-balanced HLLC finite-volume step for gfx950.
//
// hllc.step (hllc.py:1272-1373) is SSP-RK2 over the spatial operator _eval_L
// (hllc.py:1152-1269).  Each RK stage is one fused kernel over BY x BX tiles:
//   primitives u, v, eta of the stage state           (_primitives :262-282)
//   MUSCL-minmod + dry drop + Audusse + HLLC per face (_flux_x/_flux_y :285-424)
//   per-edge ghosts: fixed_stage, then closed walls    (_apply_boundaries :993-1109)
//   donor beta on the mass flux, ghost inflow unscaled (_mass_beta/_limit_* :700-773)
//   boundary banking, 0.5*dt/dx per stage            (_bank_* :638-677, :1115-1149)
//   flux divergence + centred bed-slope source        (_accumulate :776-847)
//   stage 1: Euler predictor U1                       (_rk_stage1 :850-878)
//   stage 2: Heun corrector in place, then friction, sources and the CFL
//            wave-speed partials                       (:881-990, :1338-1373)
//   with the bedload integral armed: each stage adds 0.5*dt times the MPM capacity
//   of every interior face's limited mass flux (the LI law and face depth,
//   sediment.py:161-273), the RK2 average of the transport rate.  The reference
//   runs morphology on the local-inertial scheme only (config.py:451-461); this
//   coupling is the library's own addition.
// A tile stages h, eta, z, u, v three cells out in LDS (the 4-point MUSCL stencil
// of the ring-1 cells' faces), computes every face of the ring-1 cells (their
// mass fluxes feed the donor beta the tile's own faces need), then works on its
// own cells.  Stage 1 writes U1 to separate buffers; stage 2 reads U1 with a halo
// and U^n only at its own cell, so it may overwrite U^n in place.
#include "rbs_device.hpp"

namespace rbs {

struct HlArgs {
  const float *h_in, *hu_in, *hv_in;  // stage input (U^n or U1)
  float *h_out, *hu_out, *hv_out;     // stage output (U1, or U^n in place)
  const float *hn, *hun, *hvn;        // stage 2: U^n
  const float *z, *n;
  float *h_comp;
  const float *rain, *infil;
  const float *d50;
  float *qs_int_x, *qs_comp_x, *qs_int_y, *qs_comp_y;
  uint32_t *hquiet;  // per block: U1 (bit 0) and U^n momentum (bit 1) left all +0
  const DevCtl *ctl;
  Partials part;
  double dx64;
  int64_t pitch;
  int32_t ny, nx, tiles_x, n_tiles;
  int32_t bc[4];
  int32_t flags;  // HL_* below
};

enum : int { HL_COMP = 1, HL_FIELD = 2, HL_INFIL = 4, HL_LOSS = 8, HL_SED = 16 };

// hllc.py:158-165 _minmod
__device__ __forceinline__ float minmod(float a, float b) {
  if (a * b <= 0.0f) return 0.0f;
  if (fabsf(a) < fabsf(b)) return a;
  return b;
}

// hllc.py:174-188 _dryfactor
__device__ __forceinline__ float dryfactor(float h0, float hm, float hp) {
  return (h0 <= kHDry || hm <= kHDry || hp <= kHDry) ? 0.0f : 1.0f;
}

// hllc.py:191-256 _hllc: (mass, normal momentum, transverse momentum) flux.
__device__ __forceinline__ void hllc_flux(float hl, float ul, float vtl, float hr, float ur,
                                          float vtr, float &fh, float &fm, float &ft) {
  const float g = kG, dry = kHDry;
  const float cl = sqrtf(g * fmx(hl, 0.0f));
  const float cr = sqrtf(g * fmx(hr, 0.0f));
  const bool dry_l = hl <= dry, dry_r = hr <= dry;
  if (dry_l && dry_r) {
    fh = fm = ft = 0.0f;
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
    const float hs = 0.5f * (cl + cr) + 0.25f * (ul - ur);
    const float h_star = fmx(hs * hs / g, 0.0f);
    const float c_star = sqrtf(g * h_star);
    const float u_star = 0.5f * (ul + ur) + cl - cr;
    sl = fmn(ul - cl, u_star - c_star);
    sr = fmx(ur + cr, u_star + c_star);
  }
  const float fh_l = hl * ul;
  const float fmn_l = hl * ul * ul + 0.5f * g * hl * hl;
  const float fh_r = hr * ur;
  const float fmn_r = hr * ur * ur + 0.5f * g * hr * hr;
  if (sl >= 0.0f) {
    fh = fh_l;
    fm = fmn_l;
  } else if (sr <= 0.0f) {
    fh = fh_r;
    fm = fmn_r;
  } else {
    const float inv = 1.0f / (sr - sl);
    fh = (sr * fh_l - sl * fh_r + sl * sr * (hr - hl)) * inv;
    fm = (sr * fmn_l - sl * fmn_r + sl * sr * (hr * ur - hl * ul)) * inv;
  }
  const float den_s = hr * (ur - sr) - hl * (ul - sl);
  float s_star = 0.5f * (ul + ur);
  if (fabsf(den_s) > 1.0e-12f) s_star = (sl * hr * (ur - sr) - sr * hl * (ul - sl)) / den_s;
  ft = fh * (s_star >= 0.0f ? vtl : vtr);
}

// One face of _flux_x / _flux_y: 4-point stencil p0..p3 along the face normal
// (cells f-2, f-1 | f, f+1); un is the normal velocity, ut the transverse one.
__device__ __forceinline__ void muscl_face(const float e[4], const float zz[4], const float un[4],
                                           const float ut[4], const float hh[4], float &fh,
                                           float &fl, float &fr, float &ft) {
  const float g = kG;
  const float wl = dryfactor(hh[1], hh[0], hh[2]);
  const float wr = dryfactor(hh[2], hh[1], hh[3]);
  const float eta_l = e[1] + 0.5f * wl * minmod(e[1] - e[0], e[2] - e[1]);
  const float eta_r = e[2] - 0.5f * wr * minmod(e[2] - e[1], e[3] - e[2]);
  const float zf_l = zz[1] + 0.5f * wl * minmod(zz[1] - zz[0], zz[2] - zz[1]);
  const float zf_r = zz[2] - 0.5f * wr * minmod(zz[2] - zz[1], zz[3] - zz[2]);
  const float un_l = un[1] + 0.5f * wl * minmod(un[1] - un[0], un[2] - un[1]);
  const float un_r = un[2] - 0.5f * wr * minmod(un[2] - un[1], un[3] - un[2]);
  const float ut_l = ut[1] + 0.5f * wl * minmod(ut[1] - ut[0], ut[2] - ut[1]);
  const float ut_r = ut[2] - 0.5f * wr * minmod(ut[2] - ut[1], ut[3] - ut[2]);
  const float h_l = fmx(eta_l - zf_l, 0.0f);
  const float h_r = fmx(eta_r - zf_r, 0.0f);
  const float z_star = fmx(zf_l, zf_r);
  const float hl_star = fmx(eta_l - z_star, 0.0f);
  const float hr_star = fmx(eta_r - z_star, 0.0f);
  float f0, f1, f2;
  hllc_flux(hl_star, un_l, ut_l, hr_star, un_r, ut_r, f0, f1, f2);
  fh = f0;
  fl = f1 + 0.5f * g * (h_l * h_l - hl_star * hl_star);
  fr = f1 + 0.5f * g * (h_r * h_r - hr_star * hr_star);
  ft = f2;
}

template <int BX, int BY, int STAGE>
__global__ __launch_bounds__(kThreads) void hllc_stage_kernel(HlArgs a) {
  const DevCtl *ctl = a.ctl;
  if (!ctl->active) return;

  constexpr int EW = BX + 6, EH = BY + 6;  // cells three out
  constexpr int RW = BX + 2, RH = BY + 2;  // ring-1 cells
  constexpr int XW = BX + 3;               // x-faces of ring-1 cells (per row)
  constexpr int YH = BY + 3;               // y-face rows of ring-1 cells
  __shared__ float s_h[EH][EW], s_e[EH][EW], s_z[EH][EW], s_u[EH][EW], s_v[EH][EW];
  __shared__ float s_fxh[RH][XW], s_fxl[RH][XW], s_fxr[RH][XW], s_fxt[RH][XW];
  __shared__ float s_fyh[YH][RW], s_fyl[YH][RW], s_fyr[YH][RW], s_fyt[YH][RW];
  __shared__ float s_beta[RH][RW];
  __shared__ double s_red[4 * (kThreads / 64)];

  const int tile = xcd_tile(blockIdx.x, a.n_tiles);
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int y0 = ty * BY, x0 = tx * BX;
  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int tid = threadIdx.x;
  const double dt64 = ctl->dt;
  const float dt = (float)dt64;
  const float dx = (float)a.dx64;
  const float dtdx = dt / dx;

  // ---- stage the halo'd state; primitives (_primitives) ------------------------
#pragma unroll
  for (int k0 = 0; k0 < EH * EW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < EH * EW) {
      const int lr = k / EW, lc = k - (k / EW) * EW;
      const int gi = clampi(y0 + lr - 3, 0, ny - 1), gj = clampi(x0 + lc - 3, 0, nx - 1);
      const int64_t o = (int64_t)gi * P + gj;
      const float hh = a.h_in[o];
      const float zv = a.z[o];
      float uu = 0.0f, vv = 0.0f;
      if (hh > kHDry) {
        uu = a.hu_in[o] / hh;
        vv = a.hv_in[o] / hh;
      }
      s_h[lr][lc] = hh;
      s_e[lr][lc] = hh + zv;
      s_z[lr][lc] = zv;
      s_u[lr][lc] = uu;
      s_v[lr][lc] = vv;
    }
  }
  __syncthreads();

  const int bw = a.bc[RBS_EDGE_WEST], be = a.bc[RBS_EDGE_EAST];
  const int bn = a.bc[RBS_EDGE_NORTH], bs = a.bc[RBS_EDGE_SOUTH];
  // ---- x-faces of the ring-1 cells ----------------------------------------------
#pragma unroll
  for (int k0 = 0; k0 < RH * XW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < RH * XW) {
      const int lr = k / XW, lf = k - (k / XW) * XW;
      const int r = y0 + lr - 1, f = x0 + lf - 1;
      float fh = 0.0f, fl = 0.0f, fr = 0.0f, ft = 0.0f;
      if (r >= 0 && r < ny && f >= 0 && f <= nx) {
        const int er = lr + 2;  // R3 row of r; cells f-2..f+1 are R3 cols lf..lf+3
        if (f == 0 && bw == RBS_BC_FIXED_STAGE) {  // _stage_x_west
          const float h_g = fmx(ctl->stage[RBS_EDGE_WEST] - s_z[er][lf + 2], 0.0f);
          const float u = s_u[er][lf + 2], v = s_v[er][lf + 2];
          hllc_flux(h_g, u, v, s_h[er][lf + 2], u, v, fh, fr, ft);
          fl = fr;
        } else if (f == nx && be == RBS_BC_FIXED_STAGE) {  // _stage_x_east
          const float u = s_u[er][lf + 1], v = s_v[er][lf + 1];
          const float h_g = fmx(ctl->stage[RBS_EDGE_EAST] - s_z[er][lf + 1], 0.0f);
          hllc_flux(s_h[er][lf + 1], u, v, h_g, u, v, fh, fl, ft);
          fr = fl;
        } else if (f == 0 && bw == RBS_BC_CLOSED) {  // _wall_x_west
          const float h0 = s_h[er][lf + 2], u = s_u[er][lf + 2], v = s_v[er][lf + 2];
          hllc_flux(h0, -u, v, h0, u, v, fh, fr, ft);
          fl = fr;
        } else if (f == nx && be == RBS_BC_CLOSED) {  // _wall_x_east
          const float h0 = s_h[er][lf + 1], u = s_u[er][lf + 1], v = s_v[er][lf + 1];
          hllc_flux(h0, u, v, h0, -u, v, fh, fl, ft);
          fr = fl;
        } else {  // interior, or open edge (transmissive: clamped neighbours)
          float e[4], zz[4], un[4], ut[4], hh[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            e[m] = s_e[er][lf + m];
            zz[m] = s_z[er][lf + m];
            un[m] = s_u[er][lf + m];
            ut[m] = s_v[er][lf + m];
            hh[m] = s_h[er][lf + m];
          }
          muscl_face(e, zz, un, ut, hh, fh, fl, fr, ft);
        }
      }
      s_fxh[lr][lf] = fh;
      s_fxl[lr][lf] = fl;
      s_fxr[lr][lf] = fr;
      s_fxt[lr][lf] = ft;
    }
  }
  // ---- y-faces of the ring-1 cells (normal v, transverse u) -----------------------
#pragma unroll
  for (int k0 = 0; k0 < YH * RW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < YH * RW) {
      const int lf = k / RW, lc = k - (k / RW) * RW;
      const int f = y0 + lf - 1, c = x0 + lc - 1;
      float fh = 0.0f, fl = 0.0f, fr = 0.0f, ft = 0.0f;
      if (c >= 0 && c < nx && f >= 0 && f <= ny) {
        const int ec = lc + 2;  // R3 col of c; rows f-2..f+1 are R3 rows lf..lf+3
        if (f == 0 && bn == RBS_BC_FIXED_STAGE) {  // _stage_y_north
          const float h_g = fmx(ctl->stage[RBS_EDGE_NORTH] - s_z[lf + 2][ec], 0.0f);
          const float u = s_u[lf + 2][ec], v = s_v[lf + 2][ec];
          hllc_flux(h_g, v, u, s_h[lf + 2][ec], v, u, fh, fr, ft);
          fl = fr;
        } else if (f == ny && bs == RBS_BC_FIXED_STAGE) {  // _stage_y_south
          const float u = s_u[lf + 1][ec], v = s_v[lf + 1][ec];
          const float h_g = fmx(ctl->stage[RBS_EDGE_SOUTH] - s_z[lf + 1][ec], 0.0f);
          hllc_flux(s_h[lf + 1][ec], v, u, h_g, v, u, fh, fl, ft);
          fr = fl;
        } else if (f == 0 && bn == RBS_BC_CLOSED) {  // _wall_y_north
          const float h0 = s_h[lf + 2][ec], u = s_u[lf + 2][ec], v = s_v[lf + 2][ec];
          hllc_flux(h0, -v, u, h0, v, u, fh, fr, ft);
          fl = fr;
        } else if (f == ny && bs == RBS_BC_CLOSED) {  // _wall_y_south
          const float h0 = s_h[lf + 1][ec], u = s_u[lf + 1][ec], v = s_v[lf + 1][ec];
          hllc_flux(h0, v, u, h0, -v, u, fh, fl, ft);
          fr = fl;
        } else {
          float e[4], zz[4], un[4], ut[4], hh[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            e[m] = s_e[lf + m][ec];
            zz[m] = s_z[lf + m][ec];
            un[m] = s_v[lf + m][ec];
            ut[m] = s_u[lf + m][ec];
            hh[m] = s_h[lf + m][ec];
          }
          muscl_face(e, zz, un, ut, hh, fh, fl, fr, ft);
        }
      }
      s_fyh[lf][lc] = fh;
      s_fyl[lf][lc] = fl;
      s_fyr[lf][lc] = fr;
      s_fyt[lf][lc] = ft;
    }
  }
  __syncthreads();

  // ---- _mass_beta of the ring-1 cells ----------------------------------------------
#pragma unroll
  for (int k0 = 0; k0 < RH * RW; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < RH * RW) {
      const int lr = k / RW, lc = k - (k / RW) * RW;
      float q_out = 0.0f;
      const float fl = s_fxh[lr][lc];
      if (fl < 0.0f) q_out = q_out - fl;
      const float fr = s_fxh[lr][lc + 1];
      if (fr > 0.0f) q_out = q_out + fr;
      const float ft = s_fyh[lr][lc];
      if (ft < 0.0f) q_out = q_out - ft;
      const float fb = s_fyh[lr + 1][lc];
      if (fb > 0.0f) q_out = q_out + fb;
      const float od = dtdx * q_out;
      s_beta[lr][lc] = od > 0.0f ? fclamp(s_h[lr + 2][lc + 2] / od, 0.0f, 1.0f) : 1.0f;
    }
  }
  __syncthreads();

  // ---- _limit_fx / _limit_fy on the tile's faces, banking of the edge faces ---------
  double bank = 0.0;
  const double wt = 0.5 * dt64 / a.dx64;  // hllc.py:1320 (float64)
  const bool loss = a.flags & HL_LOSS;
#pragma unroll
  for (int k0 = 0; k0 < BY * (BX + 1); k0 += kThreads) {
    const int k = k0 + tid;
    if (k < BY * (BX + 1)) {
      const int lr = k / (BX + 1), lfo = k - (k / (BX + 1)) * (BX + 1);
      const int r = y0 + lr, f = x0 + lfo;
      if (r < ny && f <= nx) {
        const int lf = lfo + 1;  // ring-face index
        float q = s_fxh[lr + 1][lf];
        if (q > 0.0f) {
          if (f >= 1) q = q * s_beta[lr + 1][lf - 1];
        } else if (q < 0.0f) {
          if (f <= nx - 1) q = q * s_beta[lr + 1][lf];
        }
        s_fxh[lr + 1][lf] = q;
        if (loss) {
          if (f == 0 && (bw == RBS_BC_OPEN || bw == RBS_BC_FIXED_STAGE)) bank -= wt * (double)q;
          if (f == nx && (be == RBS_BC_OPEN || be == RBS_BC_FIXED_STAGE)) bank += wt * (double)q;
        }
      }
    }
  }
#pragma unroll
  for (int k0 = 0; k0 < (BY + 1) * BX; k0 += kThreads) {
    const int k = k0 + tid;
    if (k < (BY + 1) * BX) {
      const int lfo = k / BX, lc = k - (k / BX) * BX;
      const int f = y0 + lfo, c = x0 + lc;
      if (f <= ny && c < nx) {
        const int lf = lfo + 1;
        float q = s_fyh[lf][lc + 1];
        if (q > 0.0f) {
          if (f >= 1) q = q * s_beta[lf - 1][lc + 1];
        } else if (q < 0.0f) {
          if (f <= ny - 1) q = q * s_beta[lf][lc + 1];
        }
        s_fyh[lf][lc + 1] = q;
        if (loss) {
          if (f == 0 && (bn == RBS_BC_OPEN || bn == RBS_BC_FIXED_STAGE)) bank -= wt * (double)q;
          if (f == ny && (bs == RBS_BC_OPEN || bs == RBS_BC_FIXED_STAGE)) bank += wt * (double)q;
        }
      }
    }
  }
  __syncthreads();

  // ---- _accumulate + RK stage (+ friction, sources on stage 2) -----------------------
  const float inv_dx = 1.0f / dx;
  const float g = kG;
  float m_s = 0.0f;
  double s_vol = 0.0, s_inf = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < BY * BX; k0 += kThreads) {
    const int k = k0 + tid;
    const int lr = k / BX, lc = k - (k / BX) * BX;
    const int r = y0 + lr, c = x0 + lc;
    if (k < BY * BX && r < ny && c < nx) {
      const int er = lr + 3, ec = lc + 3;
      const int64_t o = (int64_t)r * P + c;
      if (a.flags & HL_SED) {  // own interior faces: left (c >= 1), top (r >= 1)
        const float hdt = 0.5f * dt;
        if (c >= 1) {
          const float hf = fmx(s_e[er][ec - 1], s_e[er][ec]) - fmx(s_z[er][ec - 1], s_z[er][ec]);
          const float qs = face_capacity(s_fxh[lr + 1][lc + 1], hf, hf, 0.5f * (a.n[o - 1] + a.n[o]),
                                         0.5f * (a.d50[o - 1] + a.d50[o]));
          float si = a.qs_int_x[o], sc = a.qs_comp_x[o];
          kahan_add(si, sc, qs * hdt);
          a.qs_int_x[o] = si;
          a.qs_comp_x[o] = sc;
        }
        if (r >= 1) {
          const float hf = fmx(s_e[er - 1][ec], s_e[er][ec]) - fmx(s_z[er - 1][ec], s_z[er][ec]);
          const float qs = face_capacity(s_fyh[lr + 1][lc + 1], hf, hf, 0.5f * (a.n[o - P] + a.n[o]),
                                         0.5f * (a.d50[o - P] + a.d50[o]));
          float si = a.qs_int_y[o], sc = a.qs_comp_y[o];
          kahan_add(si, sc, qs * hdt);
          a.qs_int_y[o] = si;
          a.qs_comp_y[o] = sc;
        }
      }
      const float dh = -((s_fxh[lr + 1][lc + 2] - s_fxh[lr + 1][lc + 1]) +
                         (s_fyh[lr + 2][lc + 1] - s_fyh[lr + 1][lc + 1])) * inv_dx;
      const float hg = s_h[er][ec];
      const float wx = dryfactor(hg, s_h[er][ec - 1], s_h[er][ec + 1]);
      const float wy = dryfactor(hg, s_h[er - 1][ec], s_h[er + 1][ec]);
      const float zc = s_z[er][ec], ec_ = s_e[er][ec];
      const float sz_x = wx * minmod(zc - s_z[er][ec - 1], s_z[er][ec + 1] - zc);
      const float se_x = wx * minmod(ec_ - s_e[er][ec - 1], s_e[er][ec + 1] - ec_);
      const float hp_x = fmx((ec_ + 0.5f * se_x) - (zc + 0.5f * sz_x), 0.0f);
      const float hm_x = fmx((ec_ - 0.5f * se_x) - (zc - 0.5f * sz_x), 0.0f);
      const float src_x = -g * 0.5f * (hp_x + hm_x) * sz_x;
      const float dhu = (-((s_fxl[lr + 1][lc + 2] - s_fxr[lr + 1][lc + 1]) +
                           (s_fyt[lr + 2][lc + 1] - s_fyt[lr + 1][lc + 1])) + src_x) * inv_dx;
      const float sz_y = wy * minmod(zc - s_z[er - 1][ec], s_z[er + 1][ec] - zc);
      const float se_y = wy * minmod(ec_ - s_e[er - 1][ec], s_e[er + 1][ec] - ec_);
      const float hp_y = fmx((ec_ + 0.5f * se_y) - (zc + 0.5f * sz_y), 0.0f);
      const float hm_y = fmx((ec_ - 0.5f * se_y) - (zc - 0.5f * sz_y), 0.0f);
      const float src_y = -g * 0.5f * (hp_y + hm_y) * sz_y;
      const float dhv = (-((s_fyl[lr + 2][lc + 1] - s_fyr[lr + 1][lc + 1]) +
                           (s_fxt[lr + 1][lc + 2] - s_fxt[lr + 1][lc + 1])) + src_y) * inv_dx;
      if (STAGE == 1) {  // _rk_stage1
        const float hh = fmx(hg + dt * dh, 0.0f);
        float hu1 = 0.0f, hv1 = 0.0f;
        if (hh > kHDry) {
          hu1 = a.hu_in[o] + dt * dhu;
          hv1 = a.hv_in[o] + dt * dhv;
        }
        a.h_out[o] = hh;
        a.hu_out[o] = hu1;
        a.hv_out[o] = hv1;
      } else {  // _rk_stage2, in place on U^n
        float hh = fmx(0.5f * (a.hn[o] + hg + dt * dh), 0.0f);
        float hu = 0.0f, hv = 0.0f;
        if (hh > kHDry) {
          hu = 0.5f * (a.hun[o] + a.hu_in[o] + dt * dhu);
          hv = 0.5f * (a.hvn[o] + a.hv_in[o] + dt * dhv);
        }
        // _friction (:912-939)
        if (hh <= kHDry) {
          hu = 0.0f;
          hv = 0.0f;
        } else {
          const float mom = sqrtf(hu * hu + hv * hv);
          const float den = manning_den(mom, hh, a.n[o], g, dt);
          hu = hu / den;
          hv = hv / den;
        }
        // sources (:1350-1373)
        const float rain = ctl->rain;
        if ((a.flags & HL_COMP) && !(a.flags & HL_FIELD)) {
          float cp = a.h_comp[o];
          kahan_add(hh, cp, rain * dt);
          a.h_comp[o] = cp;
        } else if (rain != 0.0f) {
          hh = hh + rain * dt;
        }
        if (a.flags & HL_FIELD) {
          const float inc = a.rain[o] * ctl->rain_scale * dt;
          if (a.flags & HL_COMP) {
            float cp = a.h_comp[o];
            kahan_add(hh, cp, inc);
            a.h_comp[o] = cp;
          } else {
            hh = hh + inc;
          }
        }
        if (a.flags & HL_INFIL) {
          const float avail = hh;
          float inf = a.infil[o] * dt;
          if (inf > avail) inf = avail;
          if (inf < 0.0f) inf = 0.0f;
          hh = avail - inf;
          s_inf += (double)avail - (double)hh;
        }
        a.h_out[o] = hh;
        a.hu_out[o] = hu;
        a.hv_out[o] = hv;
        // _reduce_wavespeed (:948-966) of the new state
        const float cc = sqrtf(g * fmx(hh, 0.0f));
        float sp = cc;
        if (hh > kHDry) sp = fmx(fabsf(hu) / hh, fabsf(hv) / hh) + cc;
        m_s = fmx(m_s, sp);
        s_vol += (double)hh;
      }
    }
  }
  double dummy = 0.0;
  block_reduce4(m_s, s_vol, s_inf, bank, s_red);
  (void)dummy;
  if (tid == 0) {
    if (STAGE == 1) {
      a.part.edge[tile] = bank;
    } else {
      a.part.speed_max[tile] = m_s;
      a.part.vol[tile] = s_vol;
      a.part.infil[tile] = s_inf;
      a.part.edge[tile] = a.part.edge[tile] + bank;
    }
  }
}

// ---- strips: both stages without LDS ---------------------------------------------
//
// A wave owns 64 consecutive columns and writes the middle kHlStripW = 58.  A cell's
// update needs its two limited x-faces, a limited face the donor factors of the cells
// either side, a donor factor that cell's two faces, and a face the 4-point MUSCL
// stencil: three columns each side.  x-neighbours come from DPP wave shifts.  The
// wave marches down H rows with a register window; iteration R
//   takes row R's state and primitives,
//   reconstructs row R-1 along y (rows R-2..R) and forms the y-face above it,
//   reconstructs row R-2 along x and forms its x-faces (the left face of each column),
//   computes row R-2's donor factor, limits row R-2's x-faces and the y-face above it,
//   and updates row R-3.
// Rows R0-3 .. R0+H+2 are read once per wave.  Each cell's reconstruction along a
// direction serves both of its faces, and its slopes give the centred bed-slope
// source: hllc.py evaluates the slopes per face, and the products are the same
// (w is 0 or 1, and 0.5 w m = 0.5 (w m)).
constexpr int kHlStripW = 58;

struct HlRecon {             // one cell along one direction
  float ep, zp, np, tp, hp;  // at the face after it: that face's left state
  float em, zm, nm, tm, hm;  // at the face before it: that face's right state
  float src;                 // the centred bed-slope term of _accumulate (:776-847)
};

// hllc.py:285-424 (_flux_x/_flux_y up to the Audusse step) for one cell: neighbours a
// before it and b after it along the direction; un is the normal velocity there.
__device__ __forceinline__ HlRecon hl_recon(float h, float e, float z, float un, float ut,
                                            float ha, float ea, float za, float una,
                                            float uta, float hb, float eb, float zb,
                                            float unb, float utb) {
  const float g = kG;
  const float w = dryfactor(h, ha, hb);
  const float me = minmod(e - ea, eb - e);
  const float mz = minmod(z - za, zb - z);
  const float mn = minmod(un - una, unb - un);
  const float mt = minmod(ut - uta, utb - ut);
  const float hw = 0.5f * w;
  HlRecon r;
  r.ep = e + hw * me;
  r.em = e - hw * me;
  r.zp = z + hw * mz;
  r.zm = z - hw * mz;
  r.np = un + hw * mn;
  r.nm = un - hw * mn;
  r.tp = ut + hw * mt;
  r.tm = ut - hw * mt;
  r.hp = fmx(r.ep - r.zp, 0.0f);
  r.hm = fmx(r.em - r.zm, 0.0f);
  const float sz = w * mz;
  r.src = -g * 0.5f * (r.hp + r.hm) * sz;
  return r;
}

// One interior (or transmissive) face from the reconstructed states either side:
// Audusse's hydrostatic reconstruction, the HLLC flux, the pressure corrections.
__device__ __forceinline__ void hl_face(float eta_l, float zf_l, float un_l, float ut_l,
                                        float h_l, float eta_r, float zf_r, float un_r,
                                        float ut_r, float h_r, float &fh, float &fl, float &fr,
                                        float &ft) {
  const float g = kG;
  const float z_star = fmx(zf_l, zf_r);
  const float hl_star = fmx(eta_l - z_star, 0.0f);
  const float hr_star = fmx(eta_r - z_star, 0.0f);
  float f1;
  hllc_flux(hl_star, un_l, ut_l, hr_star, un_r, ut_r, fh, f1, ft);
  fl = f1 + 0.5f * g * (h_l * h_l - hl_star * hl_star);
  fr = f1 + 0.5f * g * (h_r * h_r - hr_star * hr_star);
}

// A fixed-stage or closed edge face (_stage_* and _wall_*, hllc.py:993-1109) from the
// inside cell (h, z, normal un, transverse ut); `first` when that cell comes after the
// face (west and north edges).
__device__ __forceinline__ void hl_edge_face(int bc, bool first, float stage, float h, float z,
                                             float un, float ut, float &fh, float &fl,
                                             float &fr, float &ft) {
  if (bc == RBS_BC_FIXED_STAGE) {
    const float h_g = fmx(stage - z, 0.0f);
    if (first) {
      hllc_flux(h_g, un, ut, h, un, ut, fh, fr, ft);
      fl = fr;
    } else {
      hllc_flux(h, un, ut, h_g, un, ut, fh, fl, ft);
      fr = fl;
    }
  } else if (first) {
    hllc_flux(h, -un, ut, h, un, ut, fh, fr, ft);
    fl = fr;
  } else {
    hllc_flux(h, un, ut, h, -un, ut, fh, fl, ft);
    fr = fl;
  }
}

// The neighbouring lane's value through a DPP wave-wide shift: lane i takes lane i-1's
// (wave_shr:1) or lane i+1's (wave_shl:1); the lane with no neighbour keeps its own.
__device__ __forceinline__ float hl_left(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138 /* wave_shr:1 */,
                                                 0xF, 0xF, true));
}
__device__ __forceinline__ float hl_right(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130 /* wave_shl:1 */,
                                                 0xF, 0xF, true));
}

template <int STAGE, int H, int NW, bool SED>
__global__ __launch_bounds__(64 * NW) void hllc_strip_kernel(HlArgs a) {
  const DevCtl *ctl = a.ctl;
  if (!ctl->active) return;
  __shared__ double s_red[4 * NW];

  const int tile = xcd_tile(blockIdx.x, a.n_tiles);
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ny = a.ny, nx = a.nx;
  const int64_t P = a.pitch;
  const int strip = tx * NW + wave;
  const int c = strip * kHlStripW - 3 + lane;
  const int y0 = ty * H;
  const int y_end = y0 + H < ny ? y0 + H : ny;
  const bool live = strip * kHlStripW < nx && y0 < ny;  // wave-uniform
  const int cc = clampi(c, 0, nx - 1);
  const bool own = lane >= 3 && lane < 3 + kHlStripW && c >= 0 && c < nx;

  const double dt64 = ctl->dt;
  const float dt = (float)dt64;
  const float dx = (float)a.dx64;
  const float dtdx = dt / dx;
  const float inv_dx = 1.0f / dx;
  const float g = kG;
  const double wt = 0.5 * dt64 / a.dx64;  // hllc.py:1320 (float64)
  const int bw = a.bc[RBS_EDGE_WEST], be = a.bc[RBS_EDGE_EAST];
  const int bn = a.bc[RBS_EDGE_NORTH], bs = a.bc[RBS_EDGE_SOUTH];
  const bool loss = a.flags & HL_LOSS;
  const bool bank_w = loss && bw != RBS_BC_CLOSED, bank_e = loss && be != RBS_BC_CLOSED;
  const bool bank_n = loss && bn != RBS_BC_CLOSED, bank_s = loss && bs != RBS_BC_CLOSED;

  // Dry blocks.  A block whose stage inputs hold h = +0 over all its rows and
  // columns, three out (for stage 2 also U^n's depth at its own cells), forms only +0
  // faces, and its update writes +0 to every output: the depths and momenta of U1 in
  // stage 1, of U^n in stage 2 (no rain, no Kahan debt to fold).  When the block's flag
  // says that its outputs already hold +0 (set by the last launch that computed the
  // block and wrote only +0; uploads of depth or momentum and a change of tiling clear
  // the flags; point sources and reservoirs write depths, which the check reads), the
  // block writes nothing but its partials, which are then +0 too.  Not with the bedload
  // integral, nor near a fixed-stage edge.
  const uint32_t qbit = STAGE == 1 ? 1u : 2u;
  {
    const int bx0 = tx * NW * kHlStripW - 3, bx1 = bx0 + NW * kHlStripW + 6;
    const bool near_stage = (bw == RBS_BC_FIXED_STAGE && bx0 <= 0) ||
                            (be == RBS_BC_FIXED_STAGE && bx0 <= nx && nx < bx1) ||
                            (bn == RBS_BC_FIXED_STAGE && y0 <= 4) ||
                            (bs == RBS_BC_FIXED_STAGE && y_end >= ny - 4);
    const bool no_sources = STAGE == 1 || (ctl->rain == 0.0f && !(a.flags & HL_COMP) &&
                                           (!(a.flags & HL_FIELD) || ctl->rain_scale == 0.0f));
    if (!SED && no_sources && !near_stage && (a.hquiet[tile] & qbit) != 0u) {  // block-uniform
      uint32_t wet = 0;
      if (live) {  // eight loads in flight per round
        for (int R0 = y0 - 3; R0 <= y_end + 2; R0 += 8) {
          uint32_t w8[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int R = R0 + k <= y_end + 2 ? R0 + k : y_end + 2;
            w8[k] = __float_as_uint(a.h_in[(int64_t)clampi(R, 0, ny - 1) * P + cc]);
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) wet |= w8[k];
        }
        if (STAGE == 2 && own)
          for (int i0 = y0; i0 < y_end; i0 += 8) {
            uint32_t w8[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              const int i = i0 + k < y_end ? i0 + k : y_end - 1;
              w8[k] = __float_as_uint(a.hn[(int64_t)i * P + c]);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) wet |= w8[k];
          }
      }
      if (!__syncthreads_or(wet != 0u)) {
        if (threadIdx.x == 0) {
          if (STAGE == 1) {
            a.part.edge[tile] = 0.0;
          } else {
            a.part.speed_max[tile] = 0.0f;
            a.part.vol[tile] = 0.0;
            a.part.infil[tile] = 0.0;
            a.part.edge[tile] = a.part.edge[tile] + 0.0;
          }
        }
        return;
      }
    }
  }

  float m_s = 0.0f;
  double s_vol = 0.0, s_inf = 0.0, bank = 0.0;
  uint32_t out_bits = 0;  // what the block wrote to the outputs the flag vouches for
  if (live) {
    // the window (rows relative to the newest, R)
    float h1 = 0.0f, e1 = 0.0f, z1 = 0.0f, u1 = 0.0f, v1 = 0.0f, hu1 = 0.0f, hv1 = 0.0f;
    float h2 = 0.0f, e2 = 0.0f, z2 = 0.0f, u2 = 0.0f, v2 = 0.0f, hu2 = 0.0f, hv2 = 0.0f;
    float hu3 = 0.0f, hv3 = 0.0f, hg3 = 0.0f;
    // row R-2 reconstructed along y at the face below it (the next y-face's left state)
    float pe = 0.0f, pz = 0.0f, pn = 0.0f, pt = 0.0f, ph = 0.0f;
    float yh2 = 0.0f, yl2 = 0.0f, yr2 = 0.0f, yt2 = 0.0f;  // the y-face above row R-2
    float yh3 = 0.0f, yr3 = 0.0f, yt3 = 0.0f;  // the y-face above row R-3, mass flux limited
    float xh3 = 0.0f, xl3 = 0.0f, xr3 = 0.0f, xt3 = 0.0f;  // row R-3's x-faces, limited
    float b3 = 1.0f;                                       // row R-3's donor factor
    float sx3 = 0.0f, sy2 = 0.0f, sy3 = 0.0f;              // bed-slope sources
    float dy2 = 0.0f, dy3 = 0.0f, dx3 = 0.0f;              // face depths (bedload only)
    float qh, qhu, qhv, qz;                                // the prefetched row
    auto ingest = [&](int R) {
      const int64_t o = (int64_t)clampi(R, 0, ny - 1) * P + cc;
      qh = a.h_in[o];
      qhu = a.hu_in[o];
      qhv = a.hv_in[o];
      qz = a.z[o];
    };
    // Dry rows.  When rows R-4..R hold h = +0 in every lane, every face this
    // iteration forms has both sides dry: the flux is +0 in all four components, the
    // donor factor 1 and the limited fluxes +0, exactly as the full arithmetic gives
    // them.  Such an iteration skips the reconstructions and fluxes and keeps only the
    // values later iterations read (row R-1's reconstruction at its lower face and the
    // bed-slope sources, with the signs of zero the formulas give); the update of row
    // R-3 runs as always.  Not with the bedload integral, nor in a strip that forms a
    // fixed-stage edge face, which can carry water into a dry cell.
    uint32_t dry_hist = 0;
    const int strip_c0 = strip * kHlStripW - 3;
    const int r_first = y0 - 3, r_last = y_end + 2;
    const bool stage_edge = (bw == RBS_BC_FIXED_STAGE && strip_c0 <= 0) ||
                            (be == RBS_BC_FIXED_STAGE && strip_c0 <= nx && nx <= strip_c0 + 63) ||
                            (bn == RBS_BC_FIXED_STAGE && r_first <= 0) ||
                            (bs == RBS_BC_FIXED_STAGE && r_last >= ny - 1);
    ingest(r_first);
    for (int R = r_first; R <= r_last; ++R) {
      const int it = R - r_first;
      // row R: _primitives (:262-282)
      const float h0 = qh, hu0 = qhu, hv0 = qhv, z0 = qz;
      if (R < r_last) ingest(R + 1);
      float u0 = 0.0f, v0 = 0.0f;
      if (h0 > kHDry) {
        u0 = hu0 / h0;
        v0 = hv0 / h0;
      }
      const float e0 = h0 + z0;
      dry_hist = (dry_hist << 1) | (__ballot(__float_as_uint(h0) != 0u) == 0 ? 1u : 0u);
      const bool dry = !SED && !stage_edge && (dry_hist & 0x1Fu) == 0x1Fu;  // wave-uniform

      // row R-1 along y; the face between rows R-2 and R-1 (normal v, transverse u)
      float yh1 = 0.0f, yl1 = 0.0f, yr1 = 0.0f, yt1 = 0.0f, sy1 = 0.0f, dy1 = 0.0f;
      if (dry) {  // the face is +0; row R-1's values at its lower face, as hl_recon forms them
        const float hw = 0.5f * 0.0f;
        const float me = minmod(e1 - e2, e0 - e1), mz = minmod(z1 - z2, z0 - z1);
        pe = e1 + hw * me;
        pz = z1 + hw * mz;
        pn = v1 + hw * minmod(v1 - v2, v0 - v1);
        pt = u1 + hw * minmod(u1 - u2, u0 - u1);
        ph = fmx(pe - pz, 0.0f);
        const float hm = fmx((e1 - hw * me) - (z1 - hw * mz), 0.0f);
        const float g0 = kG;
        sy1 = -g0 * 0.5f * (ph + hm) * (0.0f * mz);
      } else if (it >= 2) {  // wave-uniform, as every branch on `it` and R below
        const HlRecon ry = hl_recon(h1, e1, z1, v1, u1, h2, e2, z2, v2, u2, h0, e0, z0, v0, u0);
        sy1 = ry.src;
        if (it >= 3) {
          const int f = R - 1;
          if (f == 0 && bn != RBS_BC_OPEN)
            hl_edge_face(bn, true, ctl->stage[RBS_EDGE_NORTH], h1, z1, v1, u1, yh1, yl1, yr1, yt1);
          else if (f == ny && bs != RBS_BC_OPEN)
            hl_edge_face(bs, false, ctl->stage[RBS_EDGE_SOUTH], h2, z2, v2, u2, yh1, yl1, yr1,
                         yt1);
          else
            hl_face(pe, pz, pn, pt, ph, ry.em, ry.zm, ry.nm, ry.tm, ry.hm, yh1, yl1, yr1, yt1);
          if (SED) dy1 = fmx(e2, e1) - fmx(z2, z1);
        }
        pe = ry.ep;
        pz = ry.zp;
        pn = ry.np;
        pt = ry.tp;
        ph = ry.hp;
      }

      // row R-2 along x; its x-faces
      float xh = 0.0f, xl = 0.0f, xr = 0.0f, xt = 0.0f, sx = 0.0f, dxf = 0.0f;
      float b2 = 1.0f, xq = 0.0f, yq = 0.0f;
      if (dry) {  // row R-2's faces are +0, its donor factor 1; its x bed-slope source
        const float zl_ = hl_left(z2), zr_ = hl_right(z2), el_ = hl_left(e2), er_ = hl_right(e2);
        const float hw = 0.5f * 0.0f;
        const float me = minmod(e2 - el_, er_ - e2), mz = minmod(z2 - zl_, zr_ - z2);
        const float hp = fmx((e2 + hw * me) - (z2 + hw * mz), 0.0f);
        const float hm = fmx((e2 - hw * me) - (z2 - hw * mz), 0.0f);
        const float g0 = kG;
        sx = -g0 * 0.5f * (hp + hm) * (0.0f * mz);
        yq = yh2;
      } else if (it >= 4) {
        const float hl_ = hl_left(h2), el_ = hl_left(e2), zl_ = hl_left(z2);
        const float ul_ = hl_left(u2), vl_ = hl_left(v2);
        const float hr_ = hl_right(h2), er_ = hl_right(e2), zr_ = hl_right(z2);
        const float ur_ = hl_right(u2), vr_ = hl_right(v2);
        const HlRecon rx = hl_recon(h2, e2, z2, u2, v2, hl_, el_, zl_, ul_, vl_, hr_, er_, zr_,
                                    ur_, vr_);
        sx = rx.src;
        const float le = hl_left(rx.ep), lz = hl_left(rx.zp), ln = hl_left(rx.np);
        const float lt = hl_left(rx.tp), lh = hl_left(rx.hp);
        if (c == 0 && bw != RBS_BC_OPEN)
          hl_edge_face(bw, true, ctl->stage[RBS_EDGE_WEST], h2, z2, u2, v2, xh, xl, xr, xt);
        else if (c == nx && be != RBS_BC_OPEN)  // this lane holds column nx-1 (clamped)
          hl_edge_face(be, false, ctl->stage[RBS_EDGE_EAST], h2, z2, u2, v2, xh, xl, xr, xt);
        else
          hl_face(le, lz, ln, lt, lh, rx.em, rx.zm, rx.nm, rx.tm, rx.hm, xh, xl, xr, xt);
        if (SED) dxf = fmx(el_, e2) - fmx(zl_, z2);

        // _mass_beta of row R-2 (:700-740): the faces in the order left, right, top, bottom
        const float xhr = hl_right(xh);
        float q_out = 0.0f;
        if (xh < 0.0f) q_out = q_out - xh;
        if (xhr > 0.0f) q_out = q_out + xhr;
        if (yh2 < 0.0f) q_out = q_out - yh2;
        if (yh1 > 0.0f) q_out = q_out + yh1;
        const float od = dtdx * q_out;
        const bool divide = od > 0.0f && !(h2 >= od);
        b2 = divide ? fclamp(h2 / od, 0.0f, 1.0f) : 1.0f;
        // _limit_fx on row R-2, _limit_fy on the face above it (:743-773)
        const float b2l = hl_left(b2);
        const int f = R - 2;
        xq = xh * (xh > 0.0f ? (c >= 1 ? b2l : 1.0f) : (c <= nx - 1 ? b2 : 1.0f));
        yq = yh2 * (yh2 > 0.0f ? (f >= 1 ? b3 : 1.0f) : (f <= ny - 1 ? b2 : 1.0f));
      }

      // row i = R-3: _accumulate + the RK stage (+ friction and sources on stage 2)
      const int i = R - 3;
      const float qR = hl_right(xh3), lR = hl_right(xl3), tR = hl_right(xt3);
      if (it >= 6 && i < y_end && own) {
        const int64_t o = (int64_t)i * P + c;
        if (SED) {  // the cell's own interior faces: left (c >= 1), top (i >= 1)
          const float hdt = 0.5f * dt;
          if (c >= 1) {
            const float qs = face_capacity(xh3, dx3, dx3, 0.5f * (a.n[o - 1] + a.n[o]),
                                           0.5f * (a.d50[o - 1] + a.d50[o]));
            float si = a.qs_int_x[o], sc = a.qs_comp_x[o];
            kahan_add(si, sc, qs * hdt);
            a.qs_int_x[o] = si;
            a.qs_comp_x[o] = sc;
          }
          if (i >= 1) {
            const float qs = face_capacity(yh3, dy3, dy3, 0.5f * (a.n[o - P] + a.n[o]),
                                           0.5f * (a.d50[o - P] + a.d50[o]));
            float si = a.qs_int_y[o], sc = a.qs_comp_y[o];
            kahan_add(si, sc, qs * hdt);
            a.qs_int_y[o] = si;
            a.qs_comp_y[o] = sc;
          }
        }
        if (c == 0 && bank_w) bank -= wt * (double)xh3;
        if (c == nx - 1 && bank_e) bank += wt * (double)qR;
        if (i == 0 && bank_n) bank -= wt * (double)yh3;
        if (i == ny - 1 && bank_s) bank += wt * (double)yq;
        const float dh = -((qR - xh3) + (yq - yh3)) * inv_dx;
        const float dhu = (-((lR - xr3) + (yt2 - yt3)) + sx3) * inv_dx;
        const float dhv = (-((yl2 - yr3) + (tR - xt3)) + sy3) * inv_dx;
        const float hg = hg3;
        if (STAGE == 1) {  // _rk_stage1
          const float hh = fmx(hg + dt * dh, 0.0f);
          float hu_1 = 0.0f, hv_1 = 0.0f;
          if (hh > kHDry) {
            hu_1 = hu3 + dt * dhu;
            hv_1 = hv3 + dt * dhv;
          }
          a.h_out[o] = hh;
          a.hu_out[o] = hu_1;
          a.hv_out[o] = hv_1;
          out_bits |= __float_as_uint(hh) | __float_as_uint(hu_1) | __float_as_uint(hv_1);
        } else {  // _rk_stage2, in place on U^n
          float hh = fmx(0.5f * (a.hn[o] + hg + dt * dh), 0.0f);
          float hu = 0.0f, hv = 0.0f;
          if (hh > kHDry) {
            hu = 0.5f * (a.hun[o] + hu3 + dt * dhu);
            hv = 0.5f * (a.hvn[o] + hv3 + dt * dhv);
          }
          // _friction (:912-939)
          if (hh <= kHDry) {
            hu = 0.0f;
            hv = 0.0f;
          } else {
            const float mom = sqrtf(hu * hu + hv * hv);
            const float den = manning_den(mom, hh, a.n[o], g, dt);
            hu = hu / den;
            hv = hv / den;
          }
          // sources (:1350-1373)
          const float rain = ctl->rain;
          if ((a.flags & HL_COMP) && !(a.flags & HL_FIELD)) {
            float cp = a.h_comp[o];
            kahan_add(hh, cp, rain * dt);
            a.h_comp[o] = cp;
          } else if (rain != 0.0f) {
            hh = hh + rain * dt;
          }
          if (a.flags & HL_FIELD) {
            const float inc = a.rain[o] * ctl->rain_scale * dt;
            if (a.flags & HL_COMP) {
              float cp = a.h_comp[o];
              kahan_add(hh, cp, inc);
              a.h_comp[o] = cp;
            } else {
              hh = hh + inc;
            }
          }
          if (a.flags & HL_INFIL) {
            const float avail = hh;
            float inf = a.infil[o] * dt;
            if (inf > avail) inf = avail;
            if (inf < 0.0f) inf = 0.0f;
            hh = avail - inf;
            s_inf += (double)avail - (double)hh;
          }
          a.h_out[o] = hh;
          a.hu_out[o] = hu;
          a.hv_out[o] = hv;
          out_bits |= __float_as_uint(hu) | __float_as_uint(hv);
          // _reduce_wavespeed (:948-966) of the new state
          const float cw = sqrtf(g * fmx(hh, 0.0f));
          float sp = cw;
          if (hh > kHDry) sp = fmx(fabsf(hu) / hh, fabsf(hv) / hh) + cw;
          m_s = fmx(m_s, sp);
          s_vol += (double)hh;
        }
      }

      // slide the window down one row
      hu3 = hu2;
      hv3 = hv2;
      hg3 = h2;
      h2 = h1;
      e2 = e1;
      z2 = z1;
      u2 = u1;
      v2 = v1;
      hu2 = hu1;
      hv2 = hv1;
      h1 = h0;
      e1 = e0;
      z1 = z0;
      u1 = u0;
      v1 = v0;
      hu1 = hu0;
      hv1 = hv0;
      yh3 = yq;
      yr3 = yr2;
      yt3 = yt2;
      yh2 = yh1;
      yl2 = yl1;
      yr2 = yr1;
      yt2 = yt1;
      xh3 = xq;
      xl3 = xl;
      xr3 = xr;
      xt3 = xt;
      b3 = b2;
      sx3 = sx;
      sy3 = sy2;
      sy2 = sy1;
      dx3 = dxf;
      dy3 = dy2;
      dy2 = dy1;
    }
  }
  const bool quiet_out = !__syncthreads_or(out_bits != 0u);
  block_reduce4<64 * NW>(m_s, s_vol, s_inf, bank, s_red);
  if (threadIdx.x == 0) {
    a.hquiet[tile] = (a.hquiet[tile] & ~qbit) | (quiet_out ? qbit : 0u);
    if (STAGE == 1) {
      a.part.edge[tile] = bank;
    } else {
      a.part.speed_max[tile] = m_s;
      a.part.vol[tile] = s_vol;
      a.part.infil[tile] = s_inf;
      a.part.edge[tile] = a.part.edge[tile] + bank;
    }
  }
}

// ---- tile variants: LDS per block (and so occupancy) against halo overhead ---------
struct HlVariant {
  int bx, by, threads;  // a block's columns and rows, its threads
  void (*stage1)(HlArgs);
  void (*stage2)(HlArgs);
  void (*stage1_sed)(HlArgs);  // the same with the bedload integral
  void (*stage2_sed)(HlArgs);
};
#define RBS_HL_TILE(X, Y)                                                                  \
  {X, Y, kThreads, hllc_stage_kernel<X, Y, 1>, hllc_stage_kernel<X, Y, 2>,                  \
   hllc_stage_kernel<X, Y, 1>, hllc_stage_kernel<X, Y, 2>}
#define RBS_HL_STRIP(NW, H)                                                                \
  {kHlStripW * NW, H, 64 * NW, hllc_strip_kernel<1, H, NW, false>,                          \
   hllc_strip_kernel<2, H, NW, false>, hllc_strip_kernel<1, H, NW, true>,                  \
   hllc_strip_kernel<2, H, NW, true>}
static const HlVariant kHlVariants[] = {
    RBS_HL_TILE(64, 16),  RBS_HL_TILE(32, 16),  RBS_HL_TILE(64, 8),   RBS_HL_TILE(32, 8),
    RBS_HL_TILE(16, 16),  RBS_HL_TILE(16, 8),   RBS_HL_TILE(8, 16),   RBS_HL_STRIP(4, 32),
    RBS_HL_STRIP(4, 64),  RBS_HL_STRIP(4, 128), RBS_HL_STRIP(2, 64),  RBS_HL_STRIP(1, 64),
    RBS_HL_STRIP(1, 32),  RBS_HL_STRIP(1, 16),  RBS_HL_STRIP(2, 32),  RBS_HL_STRIP(1, 128),
};
constexpr int kNumHlVariants = sizeof(kHlVariants) / sizeof(kHlVariants[0]);
// one 58-column strip per block, 16 rows: on the 8192^2 valley (mostly dry) 317 us per
// step against 390 us for 32 rows (the finer blocks skip more dry area) and 2388 us for
// the fastest LDS tile (16x16); computing every cell, 16 and 32 rows ran alike
constexpr int kDefaultHlVariant = 13;

int hllc_variant_count() { return kNumHlVariants; }
int hllc_default_variant() { return kDefaultHlVariant; }
void hllc_variant_shape(int v, int *bx, int *by) {
  if (v < 0 || v >= kNumHlVariants) v = kDefaultHlVariant;
  *bx = kHlVariants[v].bx;
  *by = kHlVariants[v].by;
}

static const HlVariant &hl_variant(const rbs_state *s) {
  int v = s->hllc_variant;
  if (v < 0 || v >= kNumHlVariants) v = kDefaultHlVariant;
  return kHlVariants[v];
}

int hllc_num_blocks(const rbs_state *s) {
  const HlVariant &v = hl_variant(s);
  return ((s->nx + v.bx - 1) / v.bx) * ((s->ny + v.by - 1) / v.by);
}

int hllc_launch_step(rbs_state *s, bool, double, double, double, double) {
  if (!(s->armed & RBS_ARM_HLLC)) {
    set_error("HLLC step on a state without momentum (arm RBS_ARM_HLLC)");
    return RBS_E_STATE;
  }
  HlArgs a{};
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
  a.hquiet = s->hquiet;
  a.ctl = s->ctl;
  a.part = s->part;
  a.dx64 = s->dx;
  a.pitch = s->pitch;
  a.ny = s->ny;
  a.nx = s->nx;
  const HlVariant &v = hl_variant(s);
  a.tiles_x = (s->nx + v.bx - 1) / v.bx;
  a.n_tiles = hllc_num_blocks(s);
  for (int e = 0; e < 4; ++e) a.bc[e] = s->bc[e];
  int f = 0;
  if (s->armed & RBS_ARM_COMP) f |= HL_COMP;
  if (s->armed & RBS_ARM_RAIN) f |= HL_FIELD;
  if (s->armed & RBS_ARM_INFIL) f |= HL_INFIL;
  if (s->armed & RBS_ARM_SED) f |= HL_SED;
  for (int e = 0; e < 4; ++e)
    if (s->bc[e] != RBS_BC_CLOSED) f |= HL_LOSS;
  a.flags = f;
  if (a.n_tiles > s->part_cap) {
    set_error("partial buffer too small");
    return RBS_E_STATE;
  }
  float *h = s->h[s->cur];
  // stage 1: U^n -> U1
  a.h_in = h;
  a.hu_in = s->hu;
  a.hv_in = s->hv;
  a.h_out = s->h1;
  a.hu_out = s->hu1;
  a.hv_out = s->hv1;
  a.hn = h;
  a.hun = s->hu;
  a.hvn = s->hv;
  const bool sed = f & HL_SED;
  hipLaunchKernelGGL(sed ? v.stage1_sed : v.stage1, dim3(a.n_tiles), dim3(v.threads), 0,
                     s->stream, a);
  RBS_HIP(hipGetLastError());
  // stage 2: U1 (halo) + U^n (own cell) -> U^{n+1} in place
  a.h_in = s->h1;
  a.hu_in = s->hu1;
  a.hv_in = s->hv1;
  a.h_out = h;
  a.hu_out = s->hu;
  a.hv_out = s->hv;
  hipLaunchKernelGGL(sed ? v.stage2_sed : v.stage2, dim3(a.n_tiles), dim3(v.threads), 0,
                     s->stream, a);
  RBS_HIP(hipGetLastError());
  return RBS_OK;
}

}  // namespace rbs
