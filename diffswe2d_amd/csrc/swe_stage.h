// Stage arguments and the per-wavefront body of the forward stage kernels,
// shared by the single-model kernels and the multi-scenario ensemble kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "swe_common.h"
#include "swe_fetch.h"

namespace swe {

// accumulator bank: [3 banks][2 (x, y)][kSlots] raw bits of non-negative speeds
template <class T> struct Bits;
template <> struct Bits<float> { using U = unsigned int; };
template <> struct Bits<double> { using U = unsigned long long; };

template <class T> __device__ inline typename Bits<T>::U to_bits(T v);
template <> __device__ inline unsigned int to_bits<float>(float v) { return __float_as_uint(v); }
template <> __device__ inline unsigned long long to_bits<double>(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
template <class T> __device__ inline T from_bits(typename Bits<T>::U b);
template <> __device__ inline float from_bits<float>(unsigned int b) { return __uint_as_float(b); }
template <> __device__ inline double from_bits<double>(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

template <class T> struct StageArgs {
  SweConsts<T> k;
  Fields<T> f;
  int strips, segs, S;
  int wsr, positivity, friction, friction_model;
  const T* U;            // [B,3,ny,nx] state at step start
  const T* k1;           // stage 2: first-stage tendency
  T* out;                // stage 1: tendency; stage 2: new state
  T* hmax;               // stage 2, optional running max of depth [B,ny,nx]
  T* k2_out;             // stage 2, optional record of the second-stage tendency
  const T* rain;         // optional rain rate field [B or 1, ny, nx]
  const T* infil;        // optional infiltration field
  int64_t rain_bstride, infil_bstride;
  // optional rainfall record (RainfallForcing, forcing.py:5-38): rates [nt, ny, nx]
  // at times [nt], interpolated in time by each stage at its own model time
  const T* rain_times;
  const T* rain_rates;
  int rain_nt;
  double time0;          // model time of the step (host-driven) or of step 0 (device clock)
  typename Bits<T>::U* acc;       // stage 1 writes, stage 2 reads: [2][kSlots]
  typename Bits<T>::U* acc_next;  // stage 2 zeroes for the following step
  Control* ctl;          // control block at the head of the workspace
  int clock;             // 1: device-driven run (duration / elapsed kept on the device)
  int parity;            // step parity inside a device-driven run
  int step_index;        // step index inside a device-driven run
  double max_dt;         // host-driven steps: explicit max_dt
  T* dt_series;          // optional per-step dt record
  T* dt_out;             // stage 2, optional: the step's dt, written whether or not it is usable
  int record_wave;       // stage 1: accumulate wave speeds
  int fast;              // FP32: FastRound arithmetic (see swe_math.h); forward runs only
  int fixed_dt;          // host-driven stage 2: dt = max_dt (the dt-frozen map), not min(CFL dt, max_dt)
};

// dt of the step from the accumulated maxima (solver.py:618-632):
// min(cfl * reciprocal(max(ax/dx + ay/dy, eps)), max_dt)
template <class T> __device__ inline T step_dt(const typename Bits<T>::U* acc, const SweConsts<T>& k, T max_dt) {
  typename Bits<T>::U bx = 0, by = 0;
  for (int i = 0; i < kSlots; ++i) {
    typename Bits<T>::U a = acc[i], b = acc[kSlots + i];
    bx = a > bx ? a : bx;
    by = b > by ? b : by;
  }
  T ax = from_bits<T>(bx), ay = from_bits<T>(by);
  T den = clamp_lo(ax / k.dx + ay / k.dy, k.eps);
  T stable = (T(1) / den) * k.cfl;
  return tmin(stable, max_dt);
}

template <class T> __device__ inline bool dt_ok(T dt) { return isfinite(dt) && dt > T(0); }

// whether a stage launch needs the source-term paths of stage_strip
template <class T> inline bool stage_has_sources(const StageArgs<T>& a) {
  return a.rain != nullptr || a.infil != nullptr || a.rain_rates != nullptr;
}

// Where the rainfall record is sampled at one model time (forcing.py:24-38):
// right = searchsorted(times, t) clamped to [1, nt-1]; weight from the bracketing
// pair; times before the first / after the last record hold the end value.
template <class T> struct RainAt {
  int left, right, hold;  // hold: -1 interpolate, else the record index to copy
  T w;
};

template <class T> __device__ inline RainAt<T> rain_at(const T* times, int nt, double time) {
  RainAt<T> r;
  T t = (T)time;  // torch.as_tensor(time, dtype=times.dtype)
  if (nt == 1) {
    r.left = r.right = r.hold = 0;
    r.w = T(0);
    return r;
  }
  int lo = 0, hi = nt;
  while (lo < hi) {  // first index with times[i] >= t
    int mid = (lo + hi) >> 1;
    if (times[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  r.right = lo < 1 ? 1 : (lo > nt - 1 ? nt - 1 : lo);
  r.left = r.right - 1;
  T width = clamp_lo(times[r.right] - times[r.left], T(1.0e-12));
  r.w = (t - times[r.left]) / width;
  r.hold = t >= times[nt - 1] ? nt - 1 : (t <= times[0] ? 0 : -1);
  return r;
}

template <class T> __device__ inline T rain_value(const T* rates, int64_t plane, const RainAt<T>& r, int64_t cell) {
  if (r.hold >= 0) return rates[(int64_t)r.hold * plane + cell];
  return (T(1) - r.w) * rates[(int64_t)r.left * plane + cell] + r.w * rates[(int64_t)r.right * plane + cell];
}

template <class T> __device__ inline T wave_max(T v) {
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = tmax(v, o);
  }
  return v;
}

// One wavefront's strip: columns [strip*60 - 2, strip*60 + 62), rows of segment
// seg, member b.  Writes the stage output for its 60 x S cells and returns the
// wave-local maxima of the x- and y-face speeds it evaluated.  SRC: the launch
// carries a source term (rain field, rainfall record, infiltration); K2: stage 2
// records k2.  Without them those paths are compiled out, which frees the scalar
// and vector registers they would hold across the row loop (3 % of a plain step).
template <class T, int STAGE, class A = RefRound, bool SRC = true, bool K2 = true>
__device__ inline void stage_strip(const StageArgs<T>& a, int strip, int seg, int b, T dt, T half_dt, T& ax_out,
                                   T& ay_out, bool rain_rec = false, RainAt<T> rec = RainAt<T>{0, 0, 0, T(0)}) {
  const int lane = threadIdx.x & (kWave - 1);
  const Fields<T>& f = a.f;
  const SweConsts<T>& k = a.k;
  // stage 1 runs faster with LDS-routed lane shifts, stage 2 with DPP moves
  constexpr bool kDpp = STAGE == 2;
  const int x0 = strip * kOut;
  const int y0 = seg * a.S;
  const int col = x0 - kHalo + lane;
  const int rows = min(a.S, f.ny - y0);
  const int64_t mbase = (int64_t)b * 3 * f.plane;

  Prim<T> P2{}, P1{};
  Side<T> carryL{};       // row r-1 as the top state of face (r-1 | r)
  T gfh = 0, gft = 0, gfn = 0;  // flux of face (o-1 | o) as seen by row o
  T ax = T(0), ay = T(0);
  const bool lane_out = lane >= kHalo && lane < kWave - kHalo && col < f.nx;
  const bool xface_ok = lane >= 1 && lane <= kWave - 3 && col >= -1 && col <= f.nx - 1;
  const bool ycol_ok = col >= 0 && col < f.nx;
  const RowBufs<T> q = row_bufs<T>(a, b);
  // no ghost column and buffers usable (wave-uniform): interior rows load through q
  const bool strip_in = q.usable && x0 - kHalo >= 0 && x0 - kHalo + kWave <= f.nx;
  const unsigned voff = (unsigned)col * (unsigned)sizeof(T);  // meaningful where col >= 0
  // the member's output as a buffer of the same size (the rarer stage-2 loads and
  // stores -- roughness, the k2 record -- keep global addresses: each resource
  // held across the loop costs four scalar registers, and the stage-2 loop is
  // short of them)
  const __amdgpu_buffer_rsrc_t out_buf = buffer_rsrc(a.out + mbase, 3ull * q.plane_bytes);

  // Rows whose 64 cells all have depth exactly 0 (after the ghost rules): every
  // face between two such cells carries zero flux and zero speed, and every such
  // cell has zero increments of h, u, v (see increments_dry), so those parts of the
  // sweep are skipped with the values they would produce (+0 fluxes, exactly).
  // dry1..dry3: rows r-1..r-3 of the iteration; wave-uniform.
  bool dry1 = false, dry2 = false, dry3 = false;
  // stage 2: the step-start state of rows r-1 and r-2 as loaded (an interior
  // output cell's loads carry no ghost rule), for the output row's update
  T u1h = T(0), u1u = T(0), u1v = T(0), u2h = T(0), u2u = T(0), u2v = T(0);
  for (int it = 0; it < rows + 4; ++it) {
    const int r = y0 - kHalo + it;
    // row r's loads go out first; the x sweep of row o = r - 2 does not need them
    CellLoads<T> raw = fetch_loads<T, STAGE>(a, b, r, col, strip_in, &q, voff);
    // stage 2: the output row's step-start state and roughness, needed at its update
    T oh = T(0), ou = T(0), ov = T(0), coef = T(0);
    if (STAGE == 2) {
      oh = u2h;
      ou = u2u;
      ov = u2v;
      u2h = u1h;
      u2u = u1u;
      u2v = u1v;
      u1h = raw.h;
      u1u = raw.hu;
      u1v = raw.hv;
      if (it >= 4 && lane_out && a.friction)
        coef = f.rough[(int64_t)b * f.rough_bstride + (int64_t)(r - 2) * f.nx + col];
    }
    T rx_h = T(0), rx_u = T(0), rx_v = T(0);
    if (it >= 4 && !dry2) {
      // x sweep of row o (its state is P2)
      Cell<T> cx = P2.x();
      Cell<T> cm = shup<kDpp>(cx), cp = shdn<kDpp>(cx);
      Incr<T> xi = increments<T, A>(cm, cx, cp, k, a.wsr != 0);
      Side<T> right_of_face = side_minus(cx, xi, k);
      Side<T> left_of_face = side_plus(cx, xi, k);
      Side<T> nxt = shdn<kDpp>(right_of_face);
      FaceFlux<T> F = face_flux<T, A>(left_of_face, cx.h, cx.z, nxt, cp.h, cp.z, k);
      if (xface_ok) ax = tmax(ax, F.speed);
      T pfh = shup<kDpp>(F.fh), pft = shup<kDpp>(F.ft), pfn = shup<kDpp>(F.fn - F.sR);
      rx_h = adiv_by<A>(pfh - F.fh, k.ddx);
      rx_u = adiv_by<A>(pfn - (F.fn - F.sL), k.ddx);
      rx_v = adiv_by<A>(pft - F.ft, k.ddx);
    }
    Prim<T> Pn = fetch_finish<T, STAGE, StageArgs<T>, A>(a, b, raw, half_dt);
    const bool dry0 = __all(Pn.h == T(0));
    if (it >= 2) {
      // y increments of row r-1 and its two face states
      Cell<T> c1 = P1.y();
      Incr<T> yi = (dry2 && dry1 && dry0) ? increments_dry(P2.y(), c1, Pn.y(), k)
                                          : increments<T, A>(P2.y(), c1, Pn.y(), k, a.wsr != 0);
      Side<T> below = side_minus(c1, yi, k);
      if (it >= 3) {
        // face (r-2 | r-1)
        FaceFlux<T> G{};
        if (!(dry3 && dry2 && dry1 && dry0)) G = face_flux<T, A>(carryL, P2.h, P2.z, below, P1.h, P1.z, k);
        if (ycol_ok && r - 2 >= -1 && r - 2 <= f.ny - 1) ay = tmax(ay, G.speed);
        if (it >= 4) {
          // output row o = r - 2
          const int o = r - 2;
          T ry_h = adiv_by<A>(gfh - G.fh, k.ddy);
          T ry_u = adiv_by<A>(gft - G.ft, k.ddy);
          T ry_v = adiv_by<A>(gfn - (G.fn - G.sL), k.ddy);
          T th = rx_h + ry_h, tu = rx_u + ry_u, tv = rx_v + ry_v;
          if (lane_out) {
            int64_t i = mbase + (int64_t)o * f.nx + col;
            int64_t cell = (int64_t)o * f.nx + col;
            if constexpr (SRC) {
              if (a.rain != nullptr) th = th + a.rain[(int64_t)b * a.rain_bstride + cell];
              else if (rain_rec) th = th + rain_value(a.rain_rates, f.plane, rec, cell);
              if (a.infil != nullptr) th = th - clamp_lo(a.infil[(int64_t)b * a.infil_bstride + cell], T(0));
            }
            const unsigned so = (unsigned)o * q.row_bytes;
            if (STAGE == 1) {
              if (strip_in) {
                buf_store(th, out_buf, voff, so);
                buf_store(tu, out_buf, voff, so + q.plane_bytes);
                buf_store(tv, out_buf, voff, so + 2 * q.plane_bytes);
              } else {
                a.out[i] = th;
                a.out[i + f.plane] = tu;
                a.out[i + 2 * f.plane] = tv;
              }
            } else {
              if constexpr (K2) {
                if (a.k2_out != nullptr) {
                  a.k2_out[i] = th;
                  a.k2_out[i + f.plane] = tu;
                  a.k2_out[i + 2 * f.plane] = tv;
                }
              }
              T nh = oh + dt * th, nu = ou + dt * tu, nv = ov + dt * tv;
              clean_and_limit<T, A>(oh, ou, ov, nh, nu, nv, k, a.positivity != 0);
              // friction leaves a dry cell as the clean-up left it (h >= 0, zero
              // momentum): skipped when no cell of the row is wet
              if (a.friction && __any(nh > k.dry)) swe::friction<T, A>(nh, nu, nv, coef, dt, k, a.friction_model);
              if (strip_in) {
                buf_store(nh, out_buf, voff, so);
                buf_store(nu, out_buf, voff, so + q.plane_bytes);
                buf_store(nv, out_buf, voff, so + 2 * q.plane_bytes);
              } else {
                a.out[i] = nh;
                a.out[i + f.plane] = nu;
                a.out[i + 2 * f.plane] = nv;
              }
              if (a.hmax != nullptr) {
                int64_t hi = (int64_t)b * f.plane + cell;
                a.hmax[hi] = tmax(a.hmax[hi], nh);
              }
            }
          }
        }
        gfh = G.fh;
        gft = G.ft;
        gfn = G.fn - G.sR;
      }
      carryL = side_plus(c1, yi, k);
    }
    P2 = P1;
    P1 = Pn;
    dry3 = dry2;
    dry2 = dry1;
    dry1 = dry0;
  }

  ax_out = ax;
  ay_out = ay;
}

}  // namespace swe
