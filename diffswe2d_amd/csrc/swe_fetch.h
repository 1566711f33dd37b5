// Ghost-aware fetch of a stencil cell, shared by the forward and adjoint kernels.
#pragma once

#include "swe_common.h"

namespace swe {

template <class T> __device__ inline T shup(T v) { return __shfl_up(v, 1, kWave); }
template <class T> __device__ inline T shdn(T v) { return __shfl_down(v, 1, kWave); }

template <class T> __device__ inline Cell<T> shup(const Cell<T>& c) {
  return Cell<T>{shup(c.h), shup(c.un), shup(c.ut), shup(c.eta), shup(c.z)};
}
template <class T> __device__ inline Cell<T> shdn(const Cell<T>& c) {
  return Cell<T>{shdn(c.h), shdn(c.un), shdn(c.ut), shdn(c.eta), shdn(c.z)};
}
template <class T> __device__ inline Side<T> shdn(const Side<T>& s) {
  return Side<T>{shdn(s.hlin), shdn(s.eta), shdn(s.z), shdn(s.un), shdn(s.ut)};
}

// eta = h + z is formed where a Cell is built (the value primitive_cell gives),
// so the row pipelines carry one register less per row.
template <class T> struct Prim {
  T h, u, v, z;
  T hr, qu, qv;  // raw (ghost-applied) state the primitives came from
  __device__ Cell<T> x() const { return Cell<T>{h, u, v, h + z, z}; }
  __device__ Cell<T> y() const { return Cell<T>{h, v, u, h + z, z}; }
};

// state of interior cell (r, c) as stage STAGE sees it
template <class T, int STAGE, class Args>
__device__ inline void load_state(const Args& a, int64_t base, int r, int c, T half_dt, T& h, T& hu,
                                  T& hv) {
  int64_t i = base + (int64_t)r * a.f.nx + c;
  int64_t P = a.f.plane;
  h = a.U[i];
  hu = a.U[i + P];
  hv = a.U[i + 2 * P];
  if (STAGE == 2) {
    T ch = h + half_dt * a.k1[i];
    T cu = hu + half_dt * a.k1[i + P];
    T cv = hv + half_dt * a.k1[i + 2 * P];
    clean_and_limit(h, hu, hv, ch, cu, cv, a.k, a.positivity != 0);
    h = ch;
    hu = cu;
    hv = cv;
  }
}

template <class T, int STAGE, class Args>
__device__ inline Prim<T> fetch(const Args& a, int b, int r, int c, T half_dt) {
  const Fields<T>& f = a.f;
  GhostMap g = ghost_map(r, c, f.ny, f.nx, f.bc);
  int64_t base = (int64_t)b * 3 * f.plane;
  T h, hu, hv;
  load_state<T, STAGE, Args>(a, base, g.r, g.c, half_dt, h, hu, hv);
  T z = f.bed[(int64_t)b * f.bed_bstride + (int64_t)g.r * f.nx + g.c];
  if (g.kind >= 0) {
    T mirror = T(0);
    if (g.kind == BC_WALL) {
      T mh, mu, mv;
      load_state<T, STAGE, Args>(a, base, g.mr, g.mc, half_dt, mh, mu, mv);
      mirror = (g.side == LEFT || g.side == RIGHT) ? mu : mv;
    }
    T level = (g.kind == BC_WATER_LEVEL || g.kind == BC_ABSORBING) ? f.level[b * f.level_stride] : T(0);
    apply_ghost(g, h, hu, hv, mirror, z, level, a.k);
  }
  Prim<T> p;
  T eta;
  primitive_cell(h, hu, hv, z, a.k, p.h, p.u, p.v, eta);
  p.z = z;
  p.hr = h;
  p.qu = hu;
  p.qv = hv;
  return p;
}

}  // namespace swe
