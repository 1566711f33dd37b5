// Scalar math of the shallow-water time step, shared by every kernel.
//
// Each function restates one piece of the reference with the same operations
// in the same order, so that FP32 results round like the reference's eager
// PyTorch ops (build with -ffp-contract=off; IEEE division and square root).
// Constants that the reference folds in Python double arithmetic arrive here
// already folded and cast once (see SweConsts, filled on the host).
//
// Min/max/clamp helpers follow torch semantics, including NaN propagation,
// so a NaN state still produces a NaN wave speed and a non-finite dt.
#pragma once

#include <hip/hip_runtime.h>

#define SWE_HD __host__ __device__ inline

namespace swe {

// ---------------------------------------------------------------- torch-like helpers
// torch.maximum / minimum / clamp propagate NaN.  IEEE 754-2019 maximum and
// minimum do the same in one gfx950 instruction (v_maximum3_f32 /
// v_minimum3_f32) where a compare-and-select would take two or three.  The
// two forms agree on every value except the sign of an exact zero result
// (maximum(-0, +0) is +0), which no later operation of the step can see: zeros
// only meet comparisons, products and sums, never a division or sign test.
template <class T> SWE_HD T tmax(T a, T b) { return __builtin_elementwise_maximum(a, b); }
template <class T> SWE_HD T tmin(T a, T b) { return __builtin_elementwise_minimum(a, b); }
template <class T> SWE_HD T clamp_lo(T x, T lo) { return __builtin_elementwise_maximum(x, lo); }
template <class T> SWE_HD T clamp_hi(T x, T hi) { return __builtin_elementwise_minimum(x, hi); }
template <class T> SWE_HD T tabs(T x) { return __builtin_fabs(x); }

// Median of three for lo <= hi: max(lo, min(x, hi)).  FP32 device code issues
// gfx950's single v_med3_f32 (min and max in one instruction); the value is
// the same as the two-instruction form for every ordered input, up to the sign
// of an exact zero (unobservable, as above).  Used where an input cannot be a
// NaN short of a non-finite state, which fails the step's dt check anyway.
template <class T> SWE_HD T med3(T x, T lo, T hi) { return clamp_hi(clamp_lo(x, lo), hi); }
template <> SWE_HD float med3<float>(float x, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
  return clamp_hi(clamp_lo(x, lo), hi);
#endif
}
template <class T> SWE_HD T clamp2(T x, T lo, T hi) { return med3(x, lo, hi); }

// ---------------------------------------------------------------- division
// FP32 a / b rounded like IEEE division.  The sequence is the one the compiler
// emits for '/' on gfx950 -- a reciprocal estimate, one Newton step on it, then
// two remainder corrections of the quotient, all fused multiply-adds -- without
// the operand range scaling (v_div_scale / v_div_fmas) and the special-value
// fix-up (v_div_fixup) around it.  Those act only when b is denormal or beyond
// 2^126, when the quotient leaves the normal range, or when |a| < 2^-103; the
// step's depths, momenta and fluxes reach none of these short of vanishing
// values, where the result is still within an ulp (and a -0 numerator gives
// +0).  8 instructions instead of 11.  FP64 keeps the plain division.
template <class T> SWE_HD T div_rn(T a, T b) { return a / b; }

SWE_HD float rcp_estimate(float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(b);
#else
  return 1.0f / b;
#endif
}

// the Newton step on a reciprocal estimate r within an ulp of 1 / b
SWE_HD float rcp_refine(float b, float r) { return __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r); }

// the quotient and its two remainder corrections from a refined reciprocal r
SWE_HD float div_rn_refined(float a, float b, float r) {
  float q = a * r;
  q = __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
}

// the whole sequence from a reciprocal estimate r
SWE_HD float div_rn_from(float a, float b, float r) { return div_rn_refined(a, b, rcp_refine(b, r)); }

template <> SWE_HD float div_rn<float>(float a, float b) { return div_rn_from(a, b, rcp_estimate(b)); }

// a / b for a divisor fixed for the whole run (dx, dy): its reciprocal comes
// rounded from the host, so only the two remainder corrections remain
// (5 instructions), with the same rounding as div_rn.
template <class T> struct Divisor {
  T b, r;
};

template <class T> SWE_HD T div_by(T a, const Divisor<T>& d) { return a / d.b; }

template <> SWE_HD float div_by<float>(float a, const Divisor<float>& d) {
  float q = a * d.r;
  q = __builtin_fmaf(__builtin_fmaf(-d.b, q, a), d.r, q);
  return __builtin_fmaf(__builtin_fmaf(-d.b, q, a), d.r, q);
}

// FP32 square root rounded like IEEE sqrt for x >= 2^-96 (and for +-0, +inf,
// NaN): the compiler's sequence -- the hardware estimate s, then s - 1 ulp or
// s + 1 ulp where the fused residual x - s' * s says so -- without the scaling
// it wraps around arguments below 2^-96.  The step's square roots are of g * h
// and u^2 + v^2 clamped at 1e-6 and 1e-12, far above.  9 instructions instead
// of 16.  FP64 keeps the plain square root.
template <class T> SWE_HD T sqrt_rn(T x) { return sqrt(x); }

// the pick from an estimate s within an ulp of sqrt(x)
SWE_HD float sqrt_rn_from(float x, float s) {
  float lo = __builtin_bit_cast(float, __builtin_bit_cast(int, s) - 1);
  float hi = __builtin_bit_cast(float, __builtin_bit_cast(int, s) + 1);
  float r_lo = __builtin_fmaf(-lo, s, x);
  float r_hi = __builtin_fmaf(-hi, s, x);
  float out = r_lo <= 0.0f ? lo : s;
  return r_hi > 0.0f ? hi : out;
}

template <> SWE_HD float sqrt_rn<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return sqrt_rn_from(x, __builtin_amdgcn_sqrtf(x));
#else
  return sqrtf(x);
#endif
}

// Constants of one run, folded on the host in double then cast to T once.
template <class T> struct SweConsts {
  T g;            // float(gravity)                      riemann.py:36
  T half_g;       // 0.5 * gravity                       riemann.py:69-70, reconstruction.py:101-106
  T dry;          // dry_depth
  T eps;          // epsilon
  T floor;        // max(eps, dry, prec)                 riemann.py:6-10, reconstruction.py:12-15
  T hll_floor;    // max(eps, prec)                      riemann.py:13-17
  T fric_floor;   // max(eps, prec)                      friction.py:4-9 (dry_depth = 0 form)
  T fric_floor2;  // max(eps, prec)^2 (double product)   friction.py:30-33
  T theta;        // limiter theta                       numerics.py:16
  T cfl;
  T dx, dy;       // float(dx), float(dy)
  Divisor<T> ddx, ddy;  // the same with their rounded reciprocals
  T cval;         // constant-boundary value
  T manning_g;    // 9.81 (hard-coded in friction.py:66)
  T third;        // 1.0 / 3.0                           friction.py:66
};

// ---------------------------------------------------------------- arithmetic policies
// RefRound: the reference's rounding -- IEEE division and square root, Manning's
// h^third correctly rounded -- so results match the reference's eager FP32 ops.
// FastRound (FP32 forward runs, an explicit opt-in; FP64 ignores it): division
// from the reciprocal estimate with one remainder correction, the hardware square
// root, h^third from the log2 / exp2 estimates, and face velocities taken from the
// reconstruction instead of being re-divided from the face momenta (the
// reference's (h_cn * u) / h_cn).  Each is within a few ulps of the reference's
// value; a run's distance from the reference is then a tolerance, not zero.
struct RefRound {};
struct FastRound {};

template <class A> inline constexpr bool is_fast = false;
template <> inline constexpr bool is_fast<FastRound> = true;

template <class A, class T> SWE_HD T adiv(T a, T b) { return div_rn(a, b); }

// a1 / b and a2 / b, each rounded as adiv rounds it, from one reciprocal of b
// (the compiler does not always share it when the quotients sit in separate
// branches)
template <class A, class T> SWE_HD void adiv2(T a1, T a2, T b, T& q1, T& q2) {
  q1 = adiv<A>(a1, b);
  q2 = adiv<A>(a2, b);
}
template <class A, class T> SWE_HD T adiv_by(T a, const Divisor<T>& d) { return div_by(a, d); }
template <class A, class T> SWE_HD T asqrt(T x) { return sqrt_rn(x); }

template <> SWE_HD float adiv<FastRound, float>(float a, float b) {
  const float r = rcp_estimate(b);
  const float q = a * r;
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
}
template <> SWE_HD float adiv_by<FastRound, float>(float a, const Divisor<float>& d) {
  const float q = a * d.r;
  return __builtin_fmaf(__builtin_fmaf(-d.b, q, a), d.r, q);
}
template <> SWE_HD void adiv2<RefRound, float>(float a1, float a2, float b, float& q1, float& q2) {
  const float r = rcp_refine(b, rcp_estimate(b));
  q1 = div_rn_refined(a1, b, r);
  q2 = div_rn_refined(a2, b, r);
}
template <> SWE_HD void adiv2<FastRound, float>(float a1, float a2, float b, float& q1, float& q2) {
  q1 = adiv<FastRound>(a1, b);
  q2 = adiv<FastRound>(a2, b);
}
template <> SWE_HD float asqrt<FastRound, float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

// ---------------------------------------------------------------- limiter (numerics.py:4-27)
// minmod of (theta*backward, centred, theta*forward): the smallest magnitude when
// all three share a strict sign, else 0 -- equal to sign(a)*min(|a|,|b|,|c|).
// All three are > 0 exactly when their minimum is (and < 0 when their maximum
// is); a NaN makes both tests false, as it does for the three separate tests.
// With lo <= hi that is 0 clamped to [lo, hi]: one median instruction in FP32
// (a NaN argument, which the three tests map to 0, only arises from a
// non-finite state).
template <class T> SWE_HD T minmod3(T a, T b, T c) {
  T lo = tmin(a, tmin(b, c));
  T hi = tmax(a, tmax(b, c));
  if constexpr (sizeof(T) == 4) return med3(T(0), lo, hi);
  return lo > T(0) ? lo : (hi < T(0) ? hi : T(0));
}

template <class T> SWE_HD T limited_increment(T qm, T q, T qp, T theta) {
  T back = q - qm;
  T centred = T(0.5) * (qp - qm);
  T fwd = qp - q;
  return minmod3(theta * back, centred, theta * fwd);
}

// ---------------------------------------------------------------- cells
// Primitive state of a (possibly ghost) cell as the reconstruction sees it
// (riemann.py:20-38 via reconstruction.py:27-28): depth clamped at 0, velocity 0
// where dry.  un/ut are the velocity normal/tangential to the sweep direction.
template <class T> struct Cell {
  T h, un, ut, eta, z;
};

template <class T, class A = RefRound>
SWE_HD void primitive_cell(T h_raw, T hu, T hv, T z, const SweConsts<T>& k, T& h, T& u, T& v, T& eta) {
  h = clamp_lo(h_raw, T(0));
  bool wet = h > k.dry;
  T safe = clamp_lo(h, k.floor);
  u = T(0);
  v = T(0);
  if (wet) adiv2<A>(hu, hv, safe, u, v);
  eta = h + z;
}

// Limited increments of one cell along the sweep direction, after the
// water-slope reset (reconstruction.py:26-42), plus the velocity-factor
// quotient 0.5*dh/max(h, floor) (reconstruction.py:62-73).
template <class T> struct Incr {
  T dh, deta, dz, dun, dut, q;
};

template <class T, class A = RefRound>
SWE_HD Incr<T> increments(const Cell<T>& m, const Cell<T>& c, const Cell<T>& p, const SweConsts<T>& k,
                          bool water_slope_reset) {
  Incr<T> r;
  r.dh = limited_increment(m.h, c.h, p.h, k.theta);
  r.deta = limited_increment(m.eta, c.eta, p.eta, k.theta);
  r.dz = limited_increment(m.z, c.z, p.z, k.theta);
  r.dun = limited_increment(m.un, c.un, p.un, k.theta);
  r.dut = limited_increment(m.ut, c.ut, p.ut, k.theta);
  if (water_slope_reset) {
    T base = c.eta - c.h;
    T half = T(0.5) * (r.deta - r.dh);
    T bed_minus = base - half;
    T bed_plus = base + half;
    bool reset = (r.deta * r.deta > r.dz * r.dz) && ((bed_minus > m.eta) || (bed_plus > p.eta));
    if (reset) r.deta = r.dh + r.dz;
  }
  r.q = adiv<A>(T(0.5) * r.dh, clamp_lo(c.h, k.floor));
  return r;
}

// increments() of a cell whose stencil (m, c, p) is all dry with depth exactly 0
// (h = u = v = 0, eta = z): the same values without the arithmetic on zeros --
// dh, dun, dut and q are +0, deta = dz (eta = z, so the water-slope reset never
// fires).
template <class T> SWE_HD Incr<T> increments_dry(const Cell<T>& m, const Cell<T>& c, const Cell<T>& p,
                                                 const SweConsts<T>& k) {
  Incr<T> r;
  r.dz = limited_increment(m.z, c.z, p.z, k.theta);
  r.deta = r.dz;
  r.dh = r.dun = r.dut = r.q = T(0);
  return r;
}

// Reconstructed values of a cell on one of its two faces
// (reconstruction.py:50-57, 71-96).
template <class T> struct Side {
  T hlin, eta, z, un, ut;
};

// cell as the LEFT state of its +1 face
template <class T> SWE_HD Side<T> side_plus(const Cell<T>& c, const Incr<T>& d, const SweConsts<T>& k) {
  T fac = c.h > k.dry ? clamp2(T(1) - d.q, T(-10), T(10)) : T(1);
  Side<T> s;
  s.hlin = c.h + T(0.5) * d.dh;
  s.eta = c.eta + T(0.5) * d.deta;
  s.z = c.z + T(0.5) * d.dz;
  T hf = T(0.5) * fac;
  s.un = c.un + hf * d.dun;
  s.ut = c.ut + hf * d.dut;
  return s;
}

// cell as the RIGHT state of its -1 face
template <class T> SWE_HD Side<T> side_minus(const Cell<T>& c, const Incr<T>& d, const SweConsts<T>& k) {
  T fac = c.h > k.dry ? clamp2(T(1) + d.q, T(-10), T(10)) : T(1);
  Side<T> s;
  s.hlin = c.h - T(0.5) * d.dh;
  s.eta = c.eta - T(0.5) * d.deta;
  s.z = c.z - T(0.5) * d.dz;
  T hf = T(0.5) * fac;
  s.un = c.un - hf * d.dun;
  s.ut = c.ut - hf * d.dut;
  return s;
}

// ---------------------------------------------------------------- face flux
// Hydrostatic face states (reconstruction.py:59-69, 98-106) and the two-wave HLL
// flux with dry-front speeds (riemann.py:41-106).  fn is the normal-momentum
// flux; the cell on the left of the face sees fn - sL, the cell on the right
// fn - sR (solver.py:503-509).
template <class T> struct FaceFlux {
  T fh, fn, ft, sL, sR, speed;
};

template <class T, class A = RefRound>
SWE_HD FaceFlux<T> face_flux(const Side<T>& L, T hL_cell, T zL_cell, const Side<T>& R, T hR_cell, T zR_cell,
                             const SweConsts<T>& k) {
  T z_face = tmin(tmax(L.z, R.z), tmin(L.eta, R.eta));
  T hcL = clamp_lo(tmin(L.eta - z_face, L.hlin), T(0));
  T hcR = clamp_lo(tmin(R.eta - z_face, R.hlin), T(0));

  FaceFlux<T> f;
  f.sL = (k.half_g * (hL_cell + hcL)) * (zL_cell - z_face);
  f.sR = (k.half_g * (hcR + hR_cell)) * (zR_cell - z_face);

  // primitive() of the face states
  T qnL = hcL * L.un, qtL = hcL * L.ut;
  T qnR = hcR * R.un, qtR = hcR * R.ut;
  T hL = clamp_lo(hcL, T(0)), hR = clamp_lo(hcR, T(0));
  bool wetL = hL > k.dry, wetR = hR > k.dry;
  T sfL = clamp_lo(hL, k.floor), sfR = clamp_lo(hR, k.floor);
  T nL = T(0), tL = T(0), nR = T(0), tR = T(0), cL = T(0), cR = T(0);
  if constexpr (is_fast<A> && sizeof(T) == 4) {
    // the face velocity is the reconstructed one: (h_cn * u) / h_cn without the re-division
    nL = wetL ? L.un : T(0);
    tL = wetL ? L.ut : T(0);
    nR = wetR ? R.un : T(0);
    tR = wetR ? R.ut : T(0);
    cL = wetL ? asqrt<A>(k.g * sfL) : T(0);
    cR = wetR ? asqrt<A>(k.g * sfR) : T(0);
  } else {
    // where(wet, q / sf, 0) and where(wet, sqrt(g sf), 0), one branch per side
    if (wetL) {
      adiv2<RefRound>(qnL, qtL, sfL, nL, tL);
      cL = sqrt_rn(k.g * sfL);
    }
    if (wetR) {
      adiv2<RefRound>(qnR, qtR, sfR, nR, tR);
      cR = sqrt_rn(k.g * sfR);
    }
  }

  T um = (T(0.5) * (nL + nR) + cL) - cR;
  T cm = T(0.5) * (cL + cR) + T(0.25) * (nL - nR);
  bool dryL = hL <= k.dry;
  bool dryR = hR <= k.dry;
  T SL = dryL ? nR - T(2) * cR : tmin(nL - cL, um - cm);
  T SR = dryR ? nL + T(2) * cL : tmax(nR + cR, um + cm);

  T mL = hL * nL, mR = hR * nR;
  T pL = hL * (nL * nL + k.half_g * hL);
  T pR = hR * (nR * nR + k.half_g * hR);
  T den = clamp_lo(SR - SL, k.hll_floor);
  T slsr = SL * SR;
  T m_star = adiv<A>((SR * mL - SL * mR) + slsr * (hR - hL), den);
  T p_star = adiv<A>((SR * pL - SL * pR) + slsr * (hR * nR - hL * nL), den);

  T fh, fn;
  if (SL >= T(0)) {
    fh = mL;
    fn = pL;
  } else if (SR <= T(0)) {
    fh = mR;
    fn = pR;
  } else {
    fh = m_star;
    fn = p_star;
  }
  T ft = (fh > T(0) ? tL : tR) * fh;
  T speed = tmax(tabs(SL), tabs(SR));
  if (dryL && dryR) {
    fh = T(0);
    fn = T(0);
    ft = T(0);
    speed = T(0);
  }
  f.fh = fh;
  f.fn = fn;
  f.ft = ft;
  f.speed = speed;
  return f;
}

// ---------------------------------------------------------------- positivity and friction
// Positivity limiter and dry-cell momentum removal (solver.py:547-570).
// theta = min(oh / max(-dh, eps), 1) where dh < 0.  When oh >= max(-dh, eps)
// the quotient is at least 1 and rounds to at least 1 (rounding is monotonic),
// so theta is 1 without the division; only a cell losing more than its depth
// divides.  Most wavefronts then skip the division altogether.  (The fast
// arithmetic's division can fall an ulp short of 1 there; it too gets 1.)
template <class T, class A = RefRound>
SWE_HD void clean_and_limit(T oh, T ou, T ov, T& ch, T& cu, T& cv, const SweConsts<T>& k, bool positivity) {
  if (positivity) {
    T dh = ch - oh, du = cu - ou, dv = cv - ov;
    T theta = T(1);
    if (dh < T(0)) {
      const T d = clamp_lo(-dh, k.eps);
      if (!(oh >= d)) theta = clamp_hi(adiv<A>(oh, d), T(1));
    }
    ch = oh + theta * dh;
    cu = ou + theta * du;
    cv = ov + theta * dv;
  }
  T h = clamp_lo(ch, T(0));
  bool wet = h > k.dry;
  ch = h;
  cu = wet ? cu : T(0);
  cv = wet ? cv : T(0);
}

// Semi-implicit Manning friction (friction.py:12-38, 62-66, 102-131).
// h^third with third = float(1/3) = 1/3 + 9.93e-9 (friction.py:66).  FP32: for
// a positive normal h (the friction floor keeps h >= 1e-6) the cube root is
// taken from a hardware log2 / exp2 estimate by two Newton steps on y^3 = h in
// FP64, then multiplied by h^(third - 1/3) = exp(t) ~ 1 + t + t^2 / 2 with
// t = (third - 1/3) ln h in FP64; the result is h^third rounded to nearest
// but for arguments within ~1e-14 of a rounding boundary, which land one ulp
// away.  This replaces a generic powf of ~160 instructions by ~30.  Other
// arguments go to powf; FP64 keeps pow.
SWE_HD float cbrt_estimate(float x, float& log2x) {
#if defined(__HIP_DEVICE_COMPILE__)
  log2x = __builtin_amdgcn_logf(x);
  return __builtin_amdgcn_exp2f(log2x * (1.0f / 3.0f));
#else
  log2x = log2f(x);
  return exp2f(log2x * (1.0f / 3.0f));
#endif
}

SWE_HD float pow_third_from(float x, float third, float y0, float log2x) {
  const double X = x;
  double Y = y0;
  const double r = (double)rcp_estimate(3.0f * y0 * y0);  // Newton slope; its error only scales the step
  Y = __builtin_fma(-__builtin_fma(Y * Y, Y, -X), r, Y);
  Y = __builtin_fma(-__builtin_fma(Y * Y, Y, -X), r, Y);
  const double t = ((double)third - 1.0 / 3.0) * ((double)log2x * 0.6931471805599453);
  return (float)(Y * __builtin_fma(t, __builtin_fma(t, 0.5, 1.0), 1.0));
}

template <class T> SWE_HD T pow_third(T x, T third);

template <> SWE_HD float pow_third<float>(float x, float third) {
  if (!(x >= 1.17549435e-38f && x <= 3.40282347e+38f)) return powf(x, third);
  float log2x;
  const float y0 = cbrt_estimate(x, log2x);
  return pow_third_from(x, third, y0, log2x);
}

template <> SWE_HD double pow_third<double>(double x, double third) { return pow(x, third); }

// h^third for the friction coefficient under policy A: FastRound takes the
// exp2(third * log2 h) estimate directly (FP32; within a few ulps).
template <class A, class T> SWE_HD T apow_third(T x, T third) { return pow_third(x, third); }
template <> SWE_HD float apow_third<FastRound, float>(float x, float third) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * third);
#else
  return exp2f(log2f(x) * third);
#endif
}

template <class T> SWE_HD T log_t(T x);
template <> SWE_HD float log_t<float>(float x) { return logf(x); }
template <> SWE_HD double log_t<double>(double x) { return log(x); }

// model 0: Manning n; model 1: roughness length z0 (BG_Flood smart friction, friction.py:41-59)
template <class T, class A = RefRound>
SWE_HD void friction(T& h_io, T& hu, T& hv, T coef, T dt, const SweConsts<T>& k, int model) {
  T h = clamp_lo(h_io, T(0));
  bool wet = h > k.dry;
  T safe = clamp_lo(h, k.floor);
  T u = adiv<A>(wet ? hu : T(0), safe);
  T v = adiv<A>(wet ? hv : T(0), safe);
  T s2 = u * u + v * v;
  bool moving = wet && (s2 > T(0));
  T speed = moving ? asqrt<A>(clamp_lo(s2, k.fric_floor2)) : T(0);
  T cf;
  if (model == 0) {
    T sh = clamp_lo(h, k.fric_floor);
    T n = clamp_lo(coef, T(0));
    cf = adiv<A>(k.manning_g * (n * n), apow_third<A>(sh, k.third));
  } else {
    T fl = k.fric_floor;
    T ratio = adiv<A>(clamp_lo(h, fl), clamp_lo(coef, fl));
    T shallow = adiv<A>(T(1), clamp_lo(T(0.46) * ratio, fl));
    T r = clamp_lo(ratio, fl);
    T lt = T(2.5) * ((log_t(r) - T(1)) + adiv<A>(T(1), r) * T(1.359));  // scalar / tensor = reciprocal * scalar
    T deep = adiv<A>(T(1), clamp_lo(lt * lt, fl));
    cf = ratio < T(2.718281828459045) ? shallow : deep;
  }
  T factor = adiv<A>(T(1), T(1) + adiv<A>((cf * speed) * dt, safe));
  h_io = h;
  hu = wet ? hu * factor : T(0);
  hv = wet ? hv * factor : T(0);
}

}  // namespace swe
