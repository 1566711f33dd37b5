// Host-side check harness for the kernel math (test infrastructure only).
//
// Runs the same __host__ __device__ functions the gfx950 kernels use
// (swe_math.h, swe_adjoint_math.h, swe_common.h) in plain loops over a padded
// grid, so the CPU test suite can compare them with the oracle and with the
// oracle's autograd without a GPU.  Nothing in the package loads this file.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <random>
#include <vector>

#include "swe_adjoint_math.h"
#include "swe_common.h"
#include "swe_host.h"
#include "swe2d_hip.h"

using namespace swe;

namespace {


// One member on a padded (ny+4) x (nx+4) grid.
template <class T> struct Member {
  int ny, nx, P;  // P = padded row length
  std::vector<T> h, hu, hv, z;           // raw padded state and bed
  std::vector<GhostMap> gm;
  std::vector<T> ph, pu, pv, peta;       // primitives
  T level;
  const swe_config* cfg;
  SweConsts<T> k;

  int idx(int r, int c) const { return (r + 2) * P + (c + 2); }

  void fill(const T* U, const T* bed, T lvl) {
    P = nx + 4;
    int n = (ny + 4) * P;
    h.assign(n, 0); hu.assign(n, 0); hv.assign(n, 0); z.assign(n, 0); gm.resize(n);
    ph.assign(n, 0); pu.assign(n, 0); pv.assign(n, 0); peta.assign(n, 0);
    level = lvl;
    int64_t plane = (int64_t)ny * nx;
    for (int r = -2; r < ny + 2; ++r)
      for (int c = -2; c < nx + 2; ++c) {
        GhostMap g = ghost_map(r, c, ny, nx, cfg->boundary);
        int64_t s = (int64_t)g.r * nx + g.c, m = (int64_t)g.mr * nx + g.mc;
        T H = U[s], HU = U[s + plane], HV = U[s + 2 * plane];
        T zz = bed[s];
        T mirror = (g.side == LEFT || g.side == RIGHT) ? U[m + plane] : U[m + 2 * plane];
        apply_ghost(g, H, HU, HV, mirror, zz, lvl, k);
        int i = idx(r, c);
        h[i] = H; hu[i] = HU; hv[i] = HV; z[i] = zz; gm[i] = g;
        primitive_cell(H, HU, HV, zz, k, ph[i], pu[i], pv[i], peta[i]);
      }
  }
  Cell<T> cell(int r, int c, bool xdir) const {
    int i = idx(r, c);
    return xdir ? Cell<T>{ph[i], pu[i], pv[i], peta[i], z[i]} : Cell<T>{ph[i], pv[i], pu[i], peta[i], z[i]};
  }
};

// cell at offset t along the sweep line (xdir: row r fixed, else column fixed)
template <class T> Cell<T> at(const Member<T>& m, bool xdir, int line, int t) {
  return xdir ? m.cell(line, t, true) : m.cell(t, line, false);
}

template <class T>
void forward_L(const Member<T>& m, T* rhs, T& ax, T& ay, int64_t plane) {
  int ny = m.ny, nx = m.nx;
  bool wsr = m.cfg->water_slope_reset != 0;
  std::vector<T> out(3 * plane, T(0));
  for (int dir = 0; dir < 2; ++dir) {
    bool xdir = dir == 0;
    int lines = xdir ? ny : nx, len = xdir ? nx : ny;
    T dl = xdir ? m.k.dx : m.k.dy;
    for (int line = 0; line < lines; ++line) {
      std::vector<FaceFlux<T>> F(len + 1);
      for (int j = -1; j < len; ++j) {  // face between t = j and j + 1
        Cell<T> a = at(m, xdir, line, j - 1), b = at(m, xdir, line, j), c = at(m, xdir, line, j + 1),
                d = at(m, xdir, line, j + 2);
        Incr<T> ib = increments(a, b, c, m.k, wsr), ic = increments(b, c, d, m.k, wsr);
        F[j + 1] = face_flux(side_plus(b, ib, m.k), b.h, b.z, side_minus(c, ic, m.k), c.h, c.z, m.k);
        if (xdir) ax = tmax(ax, F[j + 1].speed);
        else ay = tmax(ay, F[j + 1].speed);
      }
      for (int t = 0; t < len; ++t) {
        const FaceFlux<T>& A = F[t];      // face t-1 | t
        const FaceFlux<T>& Bf = F[t + 1]; // face t | t+1
        T rh = (A.fh - Bf.fh) / dl;
        T rn = ((A.fn - A.sR) - (Bf.fn - Bf.sL)) / dl;
        T rt = (A.ft - Bf.ft) / dl;
        int64_t cell = xdir ? (int64_t)line * nx + t : (int64_t)t * nx + line;
        if (xdir) {
          out[cell] = rh;
          out[cell + plane] = rn;
          out[cell + 2 * plane] = rt;
        } else {
          out[cell] = out[cell] + rh;
          out[cell + plane] = out[cell + plane] + rt;
          out[cell + 2 * plane] = out[cell + 2 * plane] + rn;
        }
      }
    }
  }
  memcpy(rhs, out.data(), sizeof(T) * 3 * plane);
}

// L^T: adjoint of the tendency -> adjoint of the interior state and the level.
template <class T>
void reverse_L(const Member<T>& m, const T* lam_rhs, T* lam_U, T& lam_level, int64_t plane) {
  int ny = m.ny, nx = m.nx;
  bool wsr = m.cfg->water_slope_reset != 0;
  int n = (ny + 4) * m.P;
  std::vector<T> gh(n, 0), gu(n, 0), gv(n, 0), geta(n, 0);  // adjoints of primitives
  for (int dir = 0; dir < 2; ++dir) {
    bool xdir = dir == 0;
    int lines = xdir ? ny : nx, len = xdir ? nx : ny;
    T dl = xdir ? m.k.dx : m.k.dy;
    for (int line = 0; line < lines; ++line) {
      auto lam = [&](int t, int ch) -> T {  // adjoint of rhs channel ch (h, n, t) at t
        if (t < 0 || t >= len) return T(0);
        int64_t cell = xdir ? (int64_t)line * nx + t : (int64_t)t * nx + line;
        int plane_ch = ch == 0 ? 0 : (ch == 1 ? (xdir ? 1 : 2) : (xdir ? 2 : 1));
        return lam_rhs[cell + plane_ch * plane];
      };
      // side adjoints per cell t in [-2, len+1]: plus side (face t|t+1), minus side (face t-1|t)
      int L0 = -2, cnt = len + 4;
      std::vector<SideGrad<T>> gp(cnt, SideGrad<T>{0, 0, 0, 0, 0}), gmn(cnt, SideGrad<T>{0, 0, 0, 0, 0});
      for (int j = -1; j < len; ++j) {
        Cell<T> a = at(m, xdir, line, j - 1), b = at(m, xdir, line, j), c = at(m, xdir, line, j + 1),
                d = at(m, xdir, line, j + 2);
        Incr<T> ib = increments(a, b, c, m.k, wsr), ic = increments(b, c, d, m.k, wsr);
        T gfh = (lam(j + 1, 0) - lam(j, 0)) / dl;
        T gfn = (lam(j + 1, 1) - lam(j, 1)) / dl;
        T gft = (lam(j + 1, 2) - lam(j, 2)) / dl;
        T gsL = lam(j, 1) / dl;
        T gsR = -lam(j + 1, 1) / dl;
        SideGrad<T> gL, gR;
        face_flux_bwd(side_plus(b, ib, m.k), b.h, b.z, side_minus(c, ic, m.k), c.h, c.z, m.k, gfh, gfn, gft, gsL,
                      gsR, gL, gR);
        gp[j - L0] = gL;
        gmn[j + 1 - L0] = gR;
      }
      // per cell: side adjoints -> increments -> neighbours
      std::vector<PrimGrad<T>> acc(cnt, PrimGrad<T>{0, 0, 0, 0});
      for (int t = -1; t <= len; ++t) {
        Cell<T> a = at(m, xdir, line, t - 1), b = at(m, xdir, line, t), c = at(m, xdir, line, t + 1);
        bool reset;
        Incr<T> ib = increments_r(a, b, c, m.k, wsr, reset);
        CellGrad<T> cg = sides_bwd(b, ib, reset, gp[t - L0], gmn[t - L0], m.k);
        PrimGrad<T>& pc = acc[t - L0];
        pc.h += cg.h;
        pc.eta += cg.eta;
        pc.un += cg.un;
        pc.ut += cg.ut;
        increments_bwd(a, b, c, cg, m.k, acc[t - 1 - L0], acc[t - L0], acc[t + 1 - L0]);
      }
      for (int t = L0; t < len + 2; ++t) {
        int i = xdir ? m.idx(line, t) : m.idx(t, line);
        const PrimGrad<T>& g = acc[t - L0];
        gh[i] += g.h;
        geta[i] += g.eta;
        if (xdir) {
          gu[i] += g.un;
          gv[i] += g.ut;
        } else {
          gv[i] += g.un;
          gu[i] += g.ut;
        }
      }
    }
  }
  // primitives -> raw padded state, then ghosts -> interior
  for (int64_t q = 0; q < 3 * plane; ++q) lam_U[q] = 0;
  lam_level = 0;
  for (int r = -2; r < ny + 2; ++r)
    for (int c = -2; c < nx + 2; ++c) {
      int i = m.idx(r, c);
      T a, b, d;
      primitive_bwd(m.h[i], m.hu[i], m.hv[i], m.k, gh[i], gu[i], gv[i], geta[i], a, b, d);
      const GhostMap& g = m.gm[i];
      bool in_r = r >= 0 && r < ny, in_c = c >= 0 && c < nx;
      if (in_r && in_c) {
        int64_t s = (int64_t)r * nx + c;
        lam_U[s] += a;
        lam_U[s + plane] += b;
        lam_U[s + 2 * plane] += d;
      } else if (in_r || in_c) {
        int64_t s = (int64_t)g.r * nx + g.c, mm = (int64_t)g.mr * nx + g.mc;
        T eh, eu, ev, gmir, glev;
        ghost_bwd(g, m.h[m.idx(g.r, g.c)], m.z[i], m.level, a, b, d, m.k, eh, eu, ev, gmir, glev);
        lam_U[s] += eh;
        lam_U[s + plane] += eu;
        lam_U[s + 2 * plane] += ev;
        if (g.side == LEFT || g.side == RIGHT) lam_U[mm + plane] += gmir;
        else lam_U[mm + 2 * plane] += gmir;
        lam_level += glev;
      }
    }
}

template <class T> struct Run {
  const swe_config* cfg;
  int B, ny, nx, bed_b;
  double dx, dy;
  int64_t plane;
  SweConsts<T> k;
};

template <class T>
void spatial(const Run<T>& R, const T* U, const T* bed, const T* level, int lstride, T* rhs, T* axay) {
  T ax = 0, ay = 0;
  for (int b = 0; b < R.B; ++b) {
    Member<T> m;
    m.ny = R.ny;
    m.nx = R.nx;
    m.cfg = R.cfg;
    m.k = R.k;
    m.fill(U + (int64_t)b * 3 * R.plane, bed + (R.bed_b == 1 ? 0 : (int64_t)b * R.plane), level[b * lstride]);
    forward_L(m, rhs + (int64_t)b * 3 * R.plane, ax, ay, R.plane);
  }
  axay[0] = ax;
  axay[1] = ay;
}

template <class T>
void spatial_bwd(const Run<T>& R, const T* U, const T* bed, const T* level, int lstride, const T* lam_rhs, T* lam_U,
                 T* lam_level) {
  for (int b = 0; b < R.B; ++b) {
    Member<T> m;
    m.ny = R.ny;
    m.nx = R.nx;
    m.cfg = R.cfg;
    m.k = R.k;
    m.fill(U + (int64_t)b * 3 * R.plane, bed + (R.bed_b == 1 ? 0 : (int64_t)b * R.plane), level[b * lstride]);
    T gl;
    reverse_L(m, lam_rhs + (int64_t)b * 3 * R.plane, lam_U + (int64_t)b * 3 * R.plane, gl, R.plane);
    lam_level[b] = gl;
  }
}

// forward step with given dt; records k1, k2
template <class T>
void step_fwd(const Run<T>& R, const T* U, const T* bed, const T* rough, const T* level, int lstride, T dt, T* k1,
              T* k2, T* out) {
  int64_t N = (int64_t)R.B * 3 * R.plane;
  std::vector<T> P(N);
  T axay[2];
  spatial(R, U, bed, level, lstride, k1, axay);
  T half = T(0.5) * dt;
  for (int b = 0; b < R.B; ++b)
    for (int64_t c = 0; c < R.plane; ++c) {
      int64_t i = (int64_t)b * 3 * R.plane + c, p = R.plane;
      T ch = U[i] + half * k1[i], cu = U[i + p] + half * k1[i + p], cv = U[i + 2 * p] + half * k1[i + 2 * p];
      clean_and_limit(U[i], U[i + p], U[i + 2 * p], ch, cu, cv, R.k, R.cfg->positivity_limiter != 0);
      P[i] = ch;
      P[i + p] = cu;
      P[i + 2 * p] = cv;
    }
  spatial(R, P.data(), bed, level, lstride, k2, axay);
  for (int b = 0; b < R.B; ++b)
    for (int64_t c = 0; c < R.plane; ++c) {
      int64_t i = (int64_t)b * 3 * R.plane + c, p = R.plane;
      T nh = U[i] + dt * k2[i], nu = U[i + p] + dt * k2[i + p], nv = U[i + 2 * p] + dt * k2[i + 2 * p];
      clean_and_limit(U[i], U[i + p], U[i + 2 * p], nh, nu, nv, R.k, R.cfg->positivity_limiter != 0);
      if (R.cfg->apply_friction) {
        T coef = rough[(R.bed_b == 1 ? 0 : (int64_t)b * R.plane) + c];
        friction(nh, nu, nv, coef, dt, R.k, R.cfg->friction_model);
      }
      out[i] = nh;
      out[i + p] = nu;
      out[i + 2 * p] = nv;
    }
}

template <class T>
void step_bwd(const Run<T>& R, const T* U, const T* bed, const T* rough, const T* level, int lstride, T dt,
              const T* lam_out, T* lam_U, T* lam_level) {
  int64_t N = (int64_t)R.B * 3 * R.plane, p = R.plane;
  std::vector<T> k1(N), k2(N), out(N), P(N), lk2(N), lP(N), lk1(N), lU(N, T(0)), tmp(N);
  step_fwd(R, U, bed, rough, level, lstride, dt, k1.data(), k2.data(), out.data());
  T half = T(0.5) * dt;
  bool pos = R.cfg->positivity_limiter != 0;
  for (int b = 0; b < R.B; ++b)
    for (int64_t c = 0; c < p; ++c) {
      int64_t i = (int64_t)b * 3 * p + c;
      T ch = U[i] + half * k1[i], cu = U[i + p] + half * k1[i + p], cv = U[i + 2 * p] + half * k1[i + 2 * p];
      clean_and_limit(U[i], U[i + p], U[i + 2 * p], ch, cu, cv, R.k, pos);
      P[i] = ch;
      P[i + p] = cu;
      P[i + 2 * p] = cv;
      // corrector + friction reverse
      T nh = U[i] + dt * k2[i], nu = U[i + p] + dt * k2[i + p], nv = U[i + 2 * p] + dt * k2[i + 2 * p];
      T vh = nh, vu = nu, vv = nv;
      clean_and_limit(U[i], U[i + p], U[i + 2 * p], vh, vu, vv, R.k, pos);
      T gh = lam_out[i], gu = lam_out[i + p], gv = lam_out[i + 2 * p];
      if (R.cfg->apply_friction) {
        T coef = rough[(R.bed_b == 1 ? 0 : (int64_t)b * p) + c];
        friction_bwd(vh, vu, vv, coef, dt, R.k, R.cfg->friction_model, gh, gu, gv, gh, gu, gv);
      }
      T go_h = 0, go_u = 0, go_v = 0, gc_h, gc_u, gc_v;
      clean_and_limit_bwd(U[i], U[i + p], U[i + 2 * p], nh, nu, nv, R.k, pos, gh, gu, gv, go_h, go_u, go_v, gc_h,
                          gc_u, gc_v);
      lU[i] += go_h + gc_h;
      lU[i + p] += go_u + gc_u;
      lU[i + 2 * p] += go_v + gc_v;
      lk2[i] = dt * gc_h;
      lk2[i + p] = dt * gc_u;
      lk2[i + 2 * p] = dt * gc_v;
    }
  std::vector<T> lev2(R.B), lev1(R.B);
  spatial_bwd(R, P.data(), bed, level, lstride, lk2.data(), lP.data(), lev2.data());
  for (int b = 0; b < R.B; ++b)
    for (int64_t c = 0; c < p; ++c) {
      int64_t i = (int64_t)b * 3 * p + c;
      T ch = U[i] + half * k1[i], cu = U[i + p] + half * k1[i + p], cv = U[i + 2 * p] + half * k1[i + 2 * p];
      T go_h = 0, go_u = 0, go_v = 0, gc_h, gc_u, gc_v;
      clean_and_limit_bwd(U[i], U[i + p], U[i + 2 * p], ch, cu, cv, R.k, pos, lP[i], lP[i + p], lP[i + 2 * p], go_h,
                          go_u, go_v, gc_h, gc_u, gc_v);
      lU[i] += go_h + gc_h;
      lU[i + p] += go_u + gc_u;
      lU[i + 2 * p] += go_v + gc_v;
      lk1[i] = half * gc_h;
      lk1[i + p] = half * gc_u;
      lk1[i + 2 * p] = half * gc_v;
    }
  spatial_bwd(R, U, bed, level, lstride, lk1.data(), tmp.data(), lev1.data());
  for (int64_t i = 0; i < N; ++i) lam_U[i] = lU[i] + tmp[i];
  for (int b = 0; b < R.B; ++b) lam_level[b] = lev1[b] + lev2[b];
}

template <class T> Run<T> make_run(const swe_config* cfg, int B, int ny, int nx, int bed_b, double dx, double dy) {
  Run<T> R;
  R.cfg = cfg;
  R.B = B;
  R.ny = ny;
  R.nx = nx;
  R.bed_b = bed_b;
  R.dx = dx;
  R.dy = dy;
  R.plane = (int64_t)ny * nx;
  R.k = host_consts<T>(cfg, dx, dy);
  return R;
}

}  // namespace

extern "C" {

#define HC_DEFINE(T, SUF)                                                                                        \
  void hc_spatial_operator_##SUF(const swe_config* cfg, int B, int ny, int nx, int bed_b, double dx, double dy,  \
                                 const T* U, const T* bed, const T* level, int lstride, T* rhs, T* axay) {      \
    Run<T> R = make_run<T>(cfg, B, ny, nx, bed_b, dx, dy);                                                      \
    spatial(R, U, bed, level, lstride, rhs, axay);                                                              \
  }                                                                                                              \
  void hc_spatial_operator_bwd_##SUF(const swe_config* cfg, int B, int ny, int nx, int bed_b, double dx,        \
                                     double dy, const T* U, const T* bed, const T* level, int lstride,          \
                                     const T* lam_rhs, T* lam_U, T* lam_level) {                                 \
    Run<T> R = make_run<T>(cfg, B, ny, nx, bed_b, dx, dy);                                                      \
    spatial_bwd(R, U, bed, level, lstride, lam_rhs, lam_U, lam_level);                                          \
  }                                                                                                              \
  void hc_step_fwd_##SUF(const swe_config* cfg, int B, int ny, int nx, int bed_b, double dx, double dy,         \
                         const T* U, const T* bed, const T* rough, const T* level, int lstride, T dt, T* k1,    \
                         T* k2, T* out) {                                                                        \
    Run<T> R = make_run<T>(cfg, B, ny, nx, bed_b, dx, dy);                                                      \
    step_fwd(R, U, bed, rough, level, lstride, dt, k1, k2, out);                                                \
  }                                                                                                              \
  void hc_step_bwd_##SUF(const swe_config* cfg, int B, int ny, int nx, int bed_b, double dx, double dy,         \
                         const T* U, const T* bed, const T* rough, const T* level, int lstride, T dt,           \
                         const T* lam_out, T* lam_U, T* lam_level) {                                            \
    Run<T> R = make_run<T>(cfg, B, ny, nx, bed_b, dx, dy);                                                      \
    step_bwd(R, U, bed, rough, level, lstride, dt, lam_out, lam_U, lam_level);                                  \
  }

HC_DEFINE(double, f64)
HC_DEFINE(float, f32)

}  // extern "C"

// per-face values of one sweep (member 0): out[(line * (len + 1) + j + 1) * 6 + q] for faces j = -1..len-1,
// q = fh, fn, ft, sL, sR, speed; xdir = 1 for x faces
extern "C" void hc_faces_f32(const swe_config* cfg, int ny, int nx, double dx, double dy, const float* U,
                             const float* bed, float level, int xdir, float* out) {
  Run<float> R = make_run<float>(cfg, 1, ny, nx, 1, dx, dy);
  Member<float> m;
  m.ny = ny;
  m.nx = nx;
  m.cfg = cfg;
  m.k = R.k;
  m.fill(U, bed, level);
  bool xd = xdir != 0;
  int lines = xd ? ny : nx, len = xd ? nx : ny;
  for (int line = 0; line < lines; ++line)
    for (int j = -1; j < len; ++j) {
      Cell<float> a = at(m, xd, line, j - 1), b = at(m, xd, line, j), c = at(m, xd, line, j + 1),
                  d = at(m, xd, line, j + 2);
      Incr<float> ib = increments(a, b, c, m.k, cfg->water_slope_reset != 0),
                  ic = increments(b, c, d, m.k, cfg->water_slope_reset != 0);
      FaceFlux<float> F = face_flux(side_plus(b, ib, m.k), b.h, b.z, side_minus(c, ic, m.k), c.h, c.z, m.k);
      float* o = out + ((int64_t)line * (len + 1) + j + 1) * 6;
      o[0] = F.fh; o[1] = F.fn; o[2] = F.ft; o[3] = F.sL; o[4] = F.sR; o[5] = F.speed;
    }
}

// div_rn_from / div_by (swe_math.h) against IEEE division on n random operand
// pairs: numerators log-uniform over [2^-100, 2^40], divisors over [2^-30, 2^30],
// both signs, one pair in 64 with a zero numerator.  The reciprocal estimate is
// taken exact and one ulp either side (the hardware estimate is within one ulp).
// mismatch[0..2]: div_rn_from with the three estimates; mismatch[3..5]: div_by.
extern "C" void hc_div_check(long long n, unsigned seed, long long* mismatch) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> unit(1.0, 2.0), ea(-100.0, 40.0), eb(-30.0, 30.0);
  for (int i = 0; i < 6; ++i) mismatch[i] = 0;
  for (long long i = 0; i < n; ++i) {
    float a = (float)ldexp(unit(rng), (int)ea(rng));
    float b = (float)ldexp(unit(rng), (int)eb(rng));
    if (rng() & 1) a = -a;
    if (rng() & 1) b = -b;
    if ((rng() & 63) == 0) a = 0.0f;
    const float want = a / b;
    const float r = 1.0f / b;
    const float est[3] = {r, nextafterf(r, 0.0f), nextafterf(r, 2.0f * r)};
    for (int j = 0; j < 3; ++j) {
      float got = div_rn_from(a, b, est[j]);
      if (memcmp(&got, &want, 4) != 0 && !(got == 0.0f && want == 0.0f)) ++mismatch[j];
      Divisor<float> d{b, est[j]};
      float got2 = div_by(a, d);
      if (memcmp(&got2, &want, 4) != 0 && !(got2 == 0.0f && want == 0.0f)) ++mismatch[3 + j];
    }
  }
}

// sqrt_rn_from (swe_math.h) against IEEE sqrt on n random arguments log-uniform
// over [2^-96, 2^100], from the exact root and from one ulp either side of it
// (the hardware estimate is within one ulp); mismatch[0..2] per estimate.
extern "C" void hc_sqrt_check(long long n, unsigned seed, long long* mismatch) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> unit(1.0, 2.0), ex(-96.0, 100.0);
  for (int i = 0; i < 3; ++i) mismatch[i] = 0;
  for (long long i = 0; i < n; ++i) {
    float x = (float)ldexp(unit(rng), (int)ex(rng));
    const float want = sqrtf(x);
    const float est[3] = {want, nextafterf(want, 0.0f), nextafterf(want, 2.0f * want)};
    for (int j = 0; j < 3; ++j) {
      float got = sqrt_rn_from(x, est[j]);
      if (memcmp(&got, &want, 4) != 0) ++mismatch[j];
    }
  }
}

// pow_third_from (swe_math.h) against h^third rounded to nearest from a double
// evaluation, for n arguments log-uniform over [1e-6, 1e4] (the friction's
// depths); the log2 and cube-root estimates are taken exact and moved by
// +-1e-6 relative (a bound for the hardware log2 / exp2).  out[0] = arguments
// off by one ulp, out[1] = off by more, out[2] = the same for powf.
extern "C" void hc_pow_check(long long n, unsigned seed, long long* out) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> lx(log(1e-6), log(1e4)), pert(-1e-6, 1e-6);
  const float third = (float)(1.0 / 3.0);
  out[0] = out[1] = out[2] = 0;
  for (long long i = 0; i < n; ++i) {
    const float x = (float)exp(lx(rng));
    const float want = (float)pow((double)x, (double)third);
    const float l2 = (float)(log2((double)x) * (1.0 + pert(rng)));
    const float y0 = (float)(cbrt((double)x) * (1.0 + pert(rng)));
    const float got = pow_third_from(x, third, y0, l2);
    int32_t a, b;
    memcpy(&a, &got, 4);
    memcpy(&b, &want, 4);
    const int32_t d = a > b ? a - b : b - a;
    if (d == 1) ++out[0];
    if (d > 1) ++out[1];
    const float pf = powf(x, third);
    if (memcmp(&pf, &want, 4) != 0) ++out[2];
  }
}
