// Reverse-mode derivatives of the scalar math in swe_math.h.
//
// Each *_bwd function recomputes its forward from the same inputs and returns
// the adjoints of those inputs.  Branch and tie rules follow PyTorch autograd
// on the reference's ops, so gradients agree with the reference's own
// autograd (src/*.py via torch):
//   minimum/maximum   ties split 0.5 / 0.5
//   clamp             gradient passes on the inclusive side (x >= lo, x <= hi)
//   where             the selected branch only
//   sign              zero derivative (minmod's magnitude is |a|-based)
// The z components are the bed's adjoints; kernels that do not differentiate
// the bed leave them unused.
#pragma once

#include "swe_math.h"

namespace swe {

// a / b inside a gradient expression.  FP32 device code multiplies by the
// hardware reciprocal (v_rcp_f32, about 1 ulp) instead of the correctly rounded
// division sequence; divisions that recompute forward values use the forward's
// div_rn, so every branch of the reverse is taken exactly as in the forward.
// FP64 and host code divide.
template <class T> SWE_HD T gdiv(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
#endif
  return a / b;
}

// 1 / b for a denominator that several gradient terms share: the hardware
// reciprocal in FP32 device code (about 1 ulp), a division otherwise.
template <class T> SWE_HD T grcp(T b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(b);
#endif
  return T(1) / b;
}

// a / d for a divisor fixed for the whole run (dx, dy): FP32 multiplies by its
// rounded reciprocal, FP64 divides.
template <class T> SWE_HD T gdiv_by(T a, const Divisor<T>& d) {
  if constexpr (sizeof(T) == 4) return a * d.r;
  return a / d.b;
}

// weight of a in minimum(a, b) / maximum(a, b).  The two weights of one
// comparison are 1 and 0 or 0.5 and 0.5, so the second is 1 minus the first
// and g * w2 = g - g * w1 exactly; the reverse computes one and subtracts.
template <class T> SWE_HD T wmin(T a, T b) { return a < b ? T(1) : (a == b ? T(0.5) : T(0)); }
template <class T> SWE_HD T wmax(T a, T b) { return a > b ? T(1) : (a == b ? T(0.5) : T(0)); }

// ---------------------------------------------------------------- limiter
// adjoint of limited_increment(qm, q, qp): adds to gqm, gq, gqp
template <class T>
SWE_HD void limited_increment_bwd(T qm, T q, T qp, T theta, T g, T& gqm, T& gq, T& gqp) {
  T a = theta * (q - qm);
  T b = T(0.5) * (qp - qm);
  T c = theta * (qp - q);
  bool pos = tmin(a, tmin(b, c)) > T(0);  // all three > 0 (the forward's branch test)
  bool neg = tmax(a, tmax(b, c)) < T(0);  // all three < 0
  if (!(pos || neg) || g == T(0)) return;
  T aa = tabs(a), ab = tabs(b), ac = tabs(c);
  T mbc = tmin(ab, ac);
  // sign(x)^2 = 1 on the same-sign branch, so the gradient is the min weight
  T wa = wmin(aa, mbc);
  T wbc = T(1) - wa;          // wmin(mbc, aa)
  T wb = wbc * wmin(ab, ac);
  T wc = wbc - wb;            // wbc * wmin(ac, ab)
  T ga = g * wa, gb = g * wb, gc = g * wc;
  gq += theta * ga - theta * gc;
  gqm += -theta * ga - T(0.5) * gb;
  gqp += T(0.5) * gb + theta * gc;
}

// ---------------------------------------------------------------- face side values
// Adjoint of the reconstructed values one face takes from one cell.
template <class T> struct SideGrad {
  T hlin, eta, un, ut, hcell;
  T z, zcell;  // the face value of the bed and the cell bed (source term)
};

// Adjoints of one cell's increments, from both its faces (plus = its +1 face,
// minus = its -1 face); returns the adjoint of the cell's own primitive values
// and of its limited increments (after the water-slope reset and the velocity
// factor are reversed).
template <class T> struct CellGrad {
  T h, eta, un, ut;        // direct adjoints of the cell's primitive values
  T dh, deta, dun, dut;    // adjoints of the raw limited increments
  T z, dz;                 // the same for the bed
};

template <class T>
SWE_HD CellGrad<T> sides_bwd(const Cell<T>& c, const Incr<T>& d, bool reset, const SideGrad<T>& gp,
                             const SideGrad<T>& gm, const SweConsts<T>& k) {
  CellGrad<T> r;
  r.h = gp.hlin + gm.hlin + gp.hcell + gm.hcell;
  r.z = gp.z + gm.z + gp.zcell + gm.zcell;
  T gdz = T(0.5) * gp.z - T(0.5) * gm.z;
  r.eta = gp.eta + gm.eta;
  r.un = gp.un + gm.un;
  r.ut = gp.ut + gm.ut;
  T gdh = T(0.5) * gp.hlin - T(0.5) * gm.hlin;
  T gdeta = T(0.5) * gp.eta - T(0.5) * gm.eta;  // adjoint of the post-reset deta
  bool wet = c.h > k.dry;
  T facL_in = T(1) - d.q, facR_in = T(1) + d.q;
  T facL = wet ? clamp2(facL_in, T(-10), T(10)) : T(1);
  T facR = wet ? clamp2(facR_in, T(-10), T(10)) : T(1);
  r.dun = (T(0.5) * facL) * gp.un - (T(0.5) * facR) * gm.un;
  r.dut = (T(0.5) * facL) * gp.ut - (T(0.5) * facR) * gm.ut;
  T gfacL = T(0.5) * (gp.un * d.dun + gp.ut * d.dut);
  T gfacR = -T(0.5) * (gm.un * d.dun + gm.ut * d.dut);
  T gq = T(0);
  if (wet && facL_in >= T(-10) && facL_in <= T(10)) gq -= gfacL;
  if (wet && facR_in >= T(-10) && facR_in <= T(10)) gq += gfacR;
  // q = (0.5 dh) / max(h, floor)
  T s = clamp_lo(c.h, k.floor);
  T gq_s = gq * grcp(s);
  gdh += T(0.5) * gq_s;
  T gs = -gq_s * d.q;  // -gq * (0.5 dh) / s^2
  if (c.h >= k.floor) r.h += gs;
  // water-slope reset: deta' = reset ? dh + dz : deta
  if (reset) {
    gdh += gdeta;
    gdz += gdeta;
    r.deta = T(0);
  } else {
    r.deta = gdeta;
  }
  r.dh = gdh;
  r.dz = gdz;
  return r;
}

// Forward increments that also report whether the water-slope reset fired.
template <class T>
SWE_HD Incr<T> increments_r(const Cell<T>& m, const Cell<T>& c, const Cell<T>& p, const SweConsts<T>& k,
                            bool water_slope_reset, bool& reset) {
  Incr<T> r;
  r.dh = limited_increment(m.h, c.h, p.h, k.theta);
  r.deta = limited_increment(m.eta, c.eta, p.eta, k.theta);
  r.dz = limited_increment(m.z, c.z, p.z, k.theta);
  r.dun = limited_increment(m.un, c.un, p.un, k.theta);
  r.dut = limited_increment(m.ut, c.ut, p.ut, k.theta);
  reset = false;
  if (water_slope_reset) {
    T base = c.eta - c.h;
    T half = T(0.5) * (r.deta - r.dh);
    T bed_minus = base - half;
    T bed_plus = base + half;
    reset = (r.deta * r.deta > r.dz * r.dz) && ((bed_minus > m.eta) || (bed_plus > p.eta));
    if (reset) r.deta = r.dh + r.dz;
  }
  r.q = div_rn(T(0.5) * r.dh, clamp_lo(c.h, k.floor));
  return r;
}

// Spread the adjoints of a cell's raw increments onto the primitive values of
// the cell and its two neighbours (gm: cell - 1, gc: cell, gp: cell + 1).
template <class T> struct PrimGrad {
  T h, eta, un, ut, z;
};

template <class T>
SWE_HD void increments_bwd(const Cell<T>& m, const Cell<T>& c, const Cell<T>& p, const CellGrad<T>& g,
                           const SweConsts<T>& k, PrimGrad<T>& gm, PrimGrad<T>& gc, PrimGrad<T>& gp) {
  limited_increment_bwd(m.h, c.h, p.h, k.theta, g.dh, gm.h, gc.h, gp.h);
  limited_increment_bwd(m.eta, c.eta, p.eta, k.theta, g.deta, gm.eta, gc.eta, gp.eta);
  limited_increment_bwd(m.un, c.un, p.un, k.theta, g.dun, gm.un, gc.un, gp.un);
  limited_increment_bwd(m.ut, c.ut, p.ut, k.theta, g.dut, gm.ut, gc.ut, gp.ut);
  limited_increment_bwd(m.z, c.z, p.z, k.theta, g.dz, gm.z, gc.z, gp.z);
}

// ---------------------------------------------------------------- primitive variables
// adjoint of primitive_cell: (h, u, v, eta) -> (h_raw, hu, hv)
template <class T>
SWE_HD void primitive_bwd(T h_raw, T hu, T hv, const SweConsts<T>& k, T gh, T gu, T gv, T geta, T& g_hraw, T& g_hu,
                          T& g_hv) {
  T h = clamp_lo(h_raw, T(0));
  bool wet = h > k.dry;
  T safe = clamp_lo(h, k.floor);
  gh += geta;  // eta = h + z
  T r = grcp(safe);
  g_hu = wet ? gu * r : T(0);
  g_hv = wet ? gv * r : T(0);
  if (wet) {
    T gs = -(g_hu * hu + g_hv * hv) * r;  // -(gu hu + gv hv) / safe^2
    if (h >= k.floor) gh += gs;
  }
  g_hraw = h_raw >= T(0) ? gh : T(0);
}

// ---------------------------------------------------------------- face flux
// Adjoint of face_flux.  g* are the adjoints of (fh, fn, ft, sL, sR).
template <class T>
SWE_HD void face_flux_bwd(const Side<T>& L, T hL_cell, T zL_cell, const Side<T>& R, T hR_cell, T zR_cell,
                          const SweConsts<T>& k, T gfh, T gfn, T gft, T gsL, T gsR, SideGrad<T>& gL,
                          SideGrad<T>& gR) {
  // ---- forward recompute
  T zmax = tmax(L.z, R.z);
  T emin = tmin(L.eta, R.eta);
  T z_face = tmin(zmax, emin);
  T aL = L.eta - z_face, aR = R.eta - z_face;
  T mLh = tmin(aL, L.hlin), mRh = tmin(aR, R.hlin);
  T hcL = clamp_lo(mLh, T(0)), hcR = clamp_lo(mRh, T(0));
  T AL = k.half_g * (hL_cell + hcL), BL = zL_cell - z_face;
  T AR = k.half_g * (hcR + hR_cell), BR = zR_cell - z_face;
  T qnL = hcL * L.un, qtL = hcL * L.ut;
  T qnR = hcR * R.un, qtR = hcR * R.ut;
  T hL = clamp_lo(hcL, T(0)), hR = clamp_lo(hcR, T(0));
  bool wetL = hL > k.dry, wetR = hR > k.dry;
  T sfL = clamp_lo(hL, k.floor), sfR = clamp_lo(hR, k.floor);
  // the forward's face_flux forms: the same values
  T nL = T(0), tL = T(0), nR = T(0), tR = T(0), cL = T(0), cR = T(0);
  if (wetL) {
    adiv2<RefRound>(qnL, qtL, sfL, nL, tL);
    cL = sqrt_rn(k.g * sfL);
  }
  if (wetR) {
    adiv2<RefRound>(qnR, qtR, sfR, nR, tR);
    cR = sqrt_rn(k.g * sfR);
  }
  T um = (T(0.5) * (nL + nR) + cL) - cR;
  T cm = T(0.5) * (cL + cR) + T(0.25) * (nL - nR);
  bool dryL = hL <= k.dry, dryR = hR <= k.dry;
  T a1 = nL - cL, a2 = um - cm;
  T b1 = nR + cR, b2 = um + cm;
  T SL = dryL ? nR - T(2) * cR : tmin(a1, a2);
  T SR = dryR ? nL + T(2) * cL : tmax(b1, b2);
  T mL = hL * nL, mR = hR * nR;
  T iL = nL * nL + k.half_g * hL, iR = nR * nR + k.half_g * hR;
  T pL = hL * iL, pR = hR * iR;
  T dS = SR - SL;
  T den = clamp_lo(dS, k.hll_floor);
  T slsr = SL * SR;
  T dh_ = hR - hL;
  T dmn = hR * nR - hL * nL;
  T num_m = (SR * mL - SL * mR) + slsr * dh_;
  T num_p = (SR * pL - SL * pR) + slsr * dmn;
  T fh = SL >= T(0) ? mL : (SR <= T(0) ? mR : div_rn(num_m, den));
  bool both = dryL && dryR;

  // ---- reverse
  T g_hcL = T(0), g_hcR = T(0), g_zf = T(0);
  T g_hLc = T(0), g_hRc = T(0);
  T g_Lun = T(0), g_Lut = T(0), g_Run = T(0), g_Rut = T(0);
  if (!both) {
    T gnL = 0, gnR = 0, gtL = 0, gtR = 0, gcL = 0, gcR = 0, ghL = 0, ghR = 0;
    // ft = (fh > 0 ? tL : tR) * fh
    T sel = fh > T(0) ? tL : tR;
    T gfh_t = gfh + gft * sel;
    if (fh > T(0)) gtL += gft * fh;
    else gtR += gft * fh;
    T gmL = 0, gmR = 0, gpL = 0, gpR = 0, gms = 0, gps = 0;
    if (SL >= T(0)) {
      gmL = gfh_t;
      gpL = gfn;
    } else if (SR <= T(0)) {
      gmR = gfh_t;
      gpR = gfn;
    } else {
      gms = gfh_t;
      gps = gfn;
    }
    T gSL = 0, gSR = 0;
    if (gms != T(0) || gps != T(0)) {
      T rd = grcp(den);
      T gnum_m = gms * rd, gnum_p = gps * rd;
      T gden = -(gnum_m * num_m + gnum_p * num_p) * rd;  // -(gms num_m + gps num_p) / den^2
      if (dS >= k.hll_floor) {
        gSR += gden;
        gSL -= gden;
      }
      gSR += gnum_m * mL + gnum_p * pL;
      gmL += gnum_m * SR;
      gpL += gnum_p * SR;
      gSL -= gnum_m * mR + gnum_p * pR;
      gmR -= gnum_m * SL;
      gpR -= gnum_p * SL;
      T gslsr = gnum_m * dh_ + gnum_p * dmn;
      ghR += gnum_m * slsr;
      ghL -= gnum_m * slsr;
      T gdmn = gnum_p * slsr;
      ghR += gdmn * nR;
      gnR += gdmn * hR;
      ghL -= gdmn * nL;
      gnL -= gdmn * hL;
      gSL += gslsr * SR;
      gSR += gslsr * SL;
    }
    // p = h * (n^2 + half_g h), m = h * n
    ghL += gpL * iL + gmL * nL;
    T giL = gpL * hL;
    gnL += T(2) * nL * giL + gmL * hL;
    ghL += k.half_g * giL;
    ghR += gpR * iR + gmR * nR;
    T giR = gpR * hR;
    gnR += T(2) * nR * giR + gmR * hR;
    ghR += k.half_g * giR;
    // wave speeds
    T gum = 0, gcm = 0;
    if (dryL) {
      gnR += gSL;
      gcR -= T(2) * gSL;
    } else {
      T t1 = gSL * wmin(a1, a2), t2 = gSL - t1;
      gnL += t1;
      gcL -= t1;
      gum += t2;
      gcm -= t2;
    }
    if (dryR) {
      gnL += gSR;
      gcL += T(2) * gSR;
    } else {
      T t1 = gSR * wmax(b1, b2), t2 = gSR - t1;
      gnR += t1;
      gcR += t1;
      gum += t2;
      gcm += t2;
    }
    gnL += T(0.5) * gum + T(0.25) * gcm;
    gnR += T(0.5) * gum - T(0.25) * gcm;
    gcL += gum + T(0.5) * gcm;
    gcR += -gum + T(0.5) * gcm;
    // primitive of the face states
    T gsfL = 0, gsfR = 0;
    T gqnL = 0, gqtL = 0, gqnR = 0, gqtR = 0;
    if (wetL) {
      T r = grcp(sfL);
      gsfL += gdiv(gcL * k.g, T(2) * cL);
      gqnL = gnL * r;
      gqtL = gtL * r;
      gsfL -= (gqnL * qnL + gqtL * qtL) * r;  // (gn qn + gt qt) / sf^2
    }
    if (wetR) {
      T r = grcp(sfR);
      gsfR += gdiv(gcR * k.g, T(2) * cR);
      gqnR = gnR * r;
      gqtR = gtR * r;
      gsfR -= (gqnR * qnR + gqtR * qtR) * r;
    }
    if (hL >= k.floor) ghL += gsfL;
    if (hR >= k.floor) ghR += gsfR;
    if (hcL >= T(0)) g_hcL += ghL;
    if (hcR >= T(0)) g_hcR += ghR;
    g_hcL += gqnL * L.un + gqtL * L.ut;
    g_Lun += gqnL * hcL;
    g_Lut += gqtL * hcL;
    g_hcR += gqnR * R.un + gqtR * R.ut;
    g_Run += gqnR * hcR;
    g_Rut += gqtR * hcR;
  }
  // topographic source terms (never zeroed by the both-dry rule)
  T gAL = gsL * BL, gBL = gsL * AL;
  T gAR = gsR * BR, gBR = gsR * AR;
  g_hLc += k.half_g * gAL;
  g_hcL += k.half_g * gAL;
  g_hcR += k.half_g * gAR;
  g_hRc += k.half_g * gAR;
  g_zf -= gBL + gBR;
  // hydrostatic depths
  T gmLh = mLh >= T(0) ? g_hcL : T(0);
  T gmRh = mRh >= T(0) ? g_hcR : T(0);
  T gaL = gmLh * wmin(aL, L.hlin), gLh = gmLh - gaL;
  T gaR = gmRh * wmin(aR, R.hlin), gRh = gmRh - gaR;
  T gLeta = gaL, gReta = gaR;
  g_zf -= gaL + gaR;
  T gemin = g_zf * wmin(emin, zmax);
  T gLe = gemin * wmin(L.eta, R.eta);
  gLeta += gLe;
  gReta += gemin - gLe;
  T gzmax = g_zf - gemin;
  gL.z = gzmax * wmax(L.z, R.z);
  gR.z = gzmax - gL.z;
  gL.zcell = gBL;  // s_L = half_g (h_L + h_cL) (z_L - z_face)
  gR.zcell = gBR;
  gL.hlin = gLh;
  gL.eta = gLeta;
  gL.un = g_Lun;
  gL.ut = g_Lut;
  gL.hcell = g_hLc;
  gR.hlin = gRh;
  gR.eta = gReta;
  gR.un = g_Run;
  gR.ut = g_Rut;
  gR.hcell = g_hRc;
}

// ---------------------------------------------------------------- positivity and friction
// Adjoint of clean_and_limit(old, cand) -> out.  Adds to g_old, writes g_cand.
template <class T>
SWE_HD void clean_and_limit_bwd(T oh, T ou, T ov, T ch, T cu, T cv, const SweConsts<T>& k, bool positivity, T gh,
                                T gu, T gv, T& g_oh, T& g_ou, T& g_ov, T& g_ch, T& g_cu, T& g_cv) {
  T dh = ch - oh, du = cu - ou, dv = cv - ov;
  T theta = T(1);
  T s = T(0), q = T(0);
  if (positivity && dh < T(0)) {
    s = clamp_lo(-dh, k.eps);
    q = div_rn(oh, s);
    theta = clamp_hi(q, T(1));
  }
  T c2h = positivity ? oh + theta * dh : ch;
  T c2u = positivity ? ou + theta * du : cu;
  T c2v = positivity ? ov + theta * dv : cv;
  T h = clamp_lo(c2h, T(0));
  bool wet = h > k.dry;
  T g2h = c2h >= T(0) ? gh : T(0);
  T g2u = wet ? gu : T(0);
  T g2v = wet ? gv : T(0);
  (void)c2u;
  (void)c2v;
  if (!positivity) {
    g_ch = g2h;
    g_cu = g2u;
    g_cv = g2v;
    return;
  }
  // c2 = old + theta * delta
  g_oh += g2h;
  g_ou += g2u;
  g_ov += g2v;
  T gdh = theta * g2h, gdu = theta * g2u, gdv = theta * g2v;
  if (dh < T(0)) {
    T gtheta = g2h * dh + g2u * du + g2v * dv;
    T gq = q <= T(1) ? gtheta : T(0);
    T gq_s = gq * grcp(s);
    g_oh += gq_s;
    T gs = -gq_s * q;  // -gq oh / s^2
    if (-dh >= k.eps) gdh -= gs;  // s = max(-dh, eps)
  }
  // delta = cand - old
  g_ch = gdh;
  g_cu = gdu;
  g_cv = gdv;
  g_oh -= gdh;
  g_ou -= gdu;
  g_ov -= gdv;
}

// Adjoint of friction() (Manning, model 0) with respect to the state.
template <class T>
SWE_HD void friction_bwd(T h_in, T hu, T hv, T coef, T dt, const SweConsts<T>& k, int model, T gh_out, T gu_out,
                         T gv_out, T& g_h, T& g_hu, T& g_hv) {
  T h = clamp_lo(h_in, T(0));
  bool wet = h > k.dry;
  T safe = clamp_lo(h, k.floor);
  T um = wet ? hu : T(0), vm = wet ? hv : T(0);
  T u = div_rn(um, safe), v = div_rn(vm, safe);
  T s2 = u * u + v * v;
  bool moving = wet && (s2 > T(0));
  T ss2 = clamp_lo(s2, k.fric_floor2);
  T speed = moving ? sqrt_rn(ss2) : T(0);
  T sh = clamp_lo(h, k.fric_floor);
  T cf, gcf_dh = T(0);  // d cf / d h (through the clamp of h)
  if (model == 0) {
    T n = clamp_lo(coef, T(0));
    T num = k.manning_g * (n * n);
    T p = pow_third(sh, k.third);
    cf = div_rn(num, p);
    // d/dsh [num / sh^(1/3)] = -num / p^2 * (1/3) sh^(1/3 - 1), with sh^(1/3 - 1) = p / sh
    T dp = gdiv(k.third * p, sh);
    if (h >= k.fric_floor) gcf_dh = -(num / (p * p)) * dp;
  } else {
    T fl = k.fric_floor;
    T z0 = clamp_lo(coef, fl);
    T ratio = div_rn(clamp_lo(h, fl), z0);
    T dratio_dh = h >= fl ? T(1) / z0 : T(0);
    T x = T(0.46) * ratio;
    T shallow = div_rn(T(1), clamp_lo(x, fl));
    T r = clamp_lo(ratio, fl);
    T lt = T(2.5) * ((log_t(r) - T(1)) + div_rn(T(1), r) * T(1.359));
    T lt2 = lt * lt;
    T deep = div_rn(T(1), clamp_lo(lt2, fl));
    if (ratio < T(2.718281828459045)) {
      cf = shallow;
      T ds = x >= fl ? -(shallow * shallow) * T(0.46) : T(0);
      gcf_dh = ds * dratio_dh;
    } else {
      cf = deep;
      T dlt = T(2.5) * (T(1) / r - T(1.359) / (r * r));
      T dr = ratio >= fl ? T(1) : T(0);
      T dd = lt2 >= fl ? -(deep * deep) * T(2) * lt * dlt * dr : T(0);
      gcf_dh = dd * dratio_dh;
    }
  }
  T a = cf * speed;
  T b = a * dt;
  T c = div_rn(b, safe);
  T den = T(1) + c;
  T f = div_rn(T(1), den);
  T gf = wet ? (gu_out * hu + gv_out * hv) : T(0);
  T ghu = wet ? gu_out * f : T(0);
  T ghv = wet ? gv_out * f : T(0);
  T gh = gh_out;
  T gden = -gf * f * f;
  T rs = grcp(safe);
  T gb = gden * rs;
  T gsafe = -gb * c;  // -gden b / safe^2
  T ga = gb * dt;
  T gcf = ga * speed;
  T gspeed = ga * cf;
  gh += gcf * gcf_dh;
  if (moving) {
    T gss2 = gdiv(gspeed, T(2) * speed);
    T gs2 = s2 >= k.fric_floor2 ? gss2 : T(0);
    T gu = T(2) * u * gs2, gv = T(2) * v * gs2;
    T ghu_s = gu * rs, ghv_s = gv * rs;
    if (wet) {
      ghu += ghu_s;
      ghv += ghv_s;
    }
    gsafe -= (ghu_s * um + ghv_s * vm) * rs;  // (gu um + gv vm) / safe^2
  }
  if (h >= k.floor) gh += gsafe;
  g_h = h_in >= T(0) ? gh : T(0);
  g_hu = ghu;
  g_hv = ghv;
}

}  // namespace swe
