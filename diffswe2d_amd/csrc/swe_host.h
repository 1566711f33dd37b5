// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include "swe_math.h"
#include "../../include/swe2d_hip.h"

namespace swe {

// Fold the configuration the way the reference's Python does (double
// arithmetic, one cast to T): floors from riemann.py:6-17, reconstruction.py:
// 12-15 and friction.py:4-9, 0.5 * gravity, 1/3.
template <class T> inline SweConsts<T> host_consts(const swe_config* c, double dx, double dy) {
  const bool f32 = sizeof(T) == 4;
  const double prec = f32 ? 1e-6 : 1e-10;
  auto mx = [](double a, double b) { return a > b ? a : b; };
  SweConsts<T> k;
  k.g = (T)c->gravity;
  k.half_g = (T)(0.5 * c->gravity);
  k.dry = (T)c->dry_depth;
  k.eps = (T)c->epsilon;
  k.floor = (T)mx(mx(c->epsilon, c->dry_depth), prec);
  k.hll_floor = (T)mx(c->epsilon, prec);
  double ff = mx(c->epsilon, prec);
  k.fric_floor = (T)ff;
  k.fric_floor2 = (T)(ff * ff);
  k.theta = (T)c->limiter_theta;
  k.cfl = (T)c->cfl;
  k.dx = (T)dx;
  k.dy = (T)dy;
  k.ddx = Divisor<T>{k.dx, (T)(1.0 / (double)k.dx)};
  k.ddy = Divisor<T>{k.dy, (T)(1.0 / (double)k.dy)};
  k.cval = (T)c->constant_value;
  k.manning_g = (T)9.81;
  k.third = (T)(1.0 / 3.0);
  return k;
}

}  // namespace swe
