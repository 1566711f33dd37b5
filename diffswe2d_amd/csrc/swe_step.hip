// Forward time step of the shallow-water solver on gfx950.
//
// One step (solver.py:572-663) is two launches:
//   stage<1>: k1 = L(U) and the x/y maximum wave speeds
//   stage<2>: dt from those maxima, predictor = clean(U, U + dt/2 k1) fetched on
//             the fly, k2 = L(predictor), U' = clean(U, U + dt k2), friction
// L() fuses ghost cells, x/y hydrostatic MUSCL reconstruction, HLL fluxes and the
// flux difference (solver.py:448-531).
//
// Work decomposition: one wavefront owns a strip of 64 columns (60 outputs plus
// a 2-column halo each side) and streams down S rows.  x-neighbours come from
// cross-lane shuffles, y-neighbours from a register pipeline of the last rows,
// so the kernels use no LDS and no barriers.  State is SoA planes h, hu, hv.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "swe_common.h"
#include "swe_fetch.h"
#include "../../include/swe2d_hip.h"

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
  typename Bits<T>::U* acc;       // stage 1 writes, stage 2 reads: [2][kSlots]
  typename Bits<T>::U* acc_next;  // stage 2 zeroes for the following step
  Control* ctl;          // control block at the head of the workspace
  int clock;             // 1: device-driven run (duration / elapsed kept on the device)
  int parity;            // step parity inside a device-driven run
  int step_index;        // step index inside a device-driven run
  double max_dt;         // host-driven steps: explicit max_dt
  T* dt_series;          // optional per-step dt record
  int record_wave;       // stage 1: accumulate wave speeds
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

template <class T> __device__ inline T wave_max(T v) {
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = tmax(v, o);
  }
  return v;
}

template <class T, int STAGE>
__global__ __launch_bounds__(256) void stage_kernel(StageArgs<T> a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int b = blockIdx.y;
  const int strip = wid % a.strips;
  const int seg = wid / a.strips;
  if (seg >= a.segs) return;  // uniform per wavefront
  const Fields<T>& f = a.f;
  const SweConsts<T>& k = a.k;

  // ---- step control
  T dt = T(0), half_dt = T(0);
  if (a.clock) {
    if (a.ctl->done) return;
    double remaining = a.ctl->duration - a.ctl->elapsed[a.parity];
    if (remaining <= a.ctl->tol) {
      if (STAGE == 2 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        a.ctl->done = 1;
        a.ctl->steps = a.step_index;
      }
      return;
    }
    if (STAGE == 2) {
      double cap = a.ctl->cap >= 0.0 ? (remaining < a.ctl->cap ? remaining : a.ctl->cap) : remaining;
      dt = step_dt<T>(a.acc, k, (T)cap);
    }
  } else if (STAGE == 2) {
    dt = step_dt<T>(a.acc, k, (T)a.max_dt);
  }
  if (STAGE == 2) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2 * kSlots) a.acc_next[threadIdx.x] = 0;
    bool ok = dt_ok(dt);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
      a.ctl->last_dt = (double)dt;
      if (!ok) {
        a.ctl->status = isfinite(dt) ? 2 : 1;
        a.ctl->bad_dt = (double)dt;
      }
      if (a.clock) {
        if (!ok) {
          a.ctl->done = 1;
          a.ctl->steps = a.step_index;
        } else {
          int p = a.parity;
          double e = a.ctl->elapsed[p] + (double)dt;
          if (e > a.ctl->duration && e - a.ctl->duration <= a.ctl->tol) e = a.ctl->duration;
          a.ctl->elapsed[1 - p] = e;
          a.ctl->steps = a.step_index + 1;
        }
      }
      if (a.dt_series != nullptr && ok) a.dt_series[a.step_index] = dt;
    }
    if (!ok) return;
    half_dt = T(0.5) * dt;
  }

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

  for (int it = 0; it < rows + 4; ++it) {
    const int r = y0 - kHalo + it;
    Prim<T> Pn = fetch<T, STAGE>(a, b, r, col, half_dt);
    if (it >= 2) {
      // y increments of row r-1 and its two face states
      Cell<T> c1 = P1.y();
      Incr<T> yi = increments(P2.y(), c1, Pn.y(), k, a.wsr != 0);
      Side<T> below = side_minus(c1, yi, k);
      if (it >= 3) {
        // face (r-2 | r-1)
        FaceFlux<T> G = face_flux(carryL, P2.h, P2.z, below, P1.h, P1.z, k);
        if (ycol_ok && r - 2 >= -1 && r - 2 <= f.ny - 1) ay = tmax(ay, G.speed);
        if (it >= 4) {
          // output row o = r - 2 (its state is P2)
          const int o = r - 2;
          T ry_h = (gfh - G.fh) / k.dy;
          T ry_u = (gft - G.ft) / k.dy;
          T ry_v = (gfn - (G.fn - G.sL)) / k.dy;
          // x sweep of row o
          Cell<T> cx = P2.x();
          Cell<T> cm = shup(cx), cp = shdn(cx);
          Incr<T> xi = increments(cm, cx, cp, k, a.wsr != 0);
          Side<T> right_of_face = side_minus(cx, xi, k);
          Side<T> left_of_face = side_plus(cx, xi, k);
          Side<T> nxt = shdn(right_of_face);
          FaceFlux<T> F = face_flux(left_of_face, cx.h, cx.z, nxt, cp.h, cp.z, k);
          if (xface_ok) ax = tmax(ax, F.speed);
          T pfh = shup(F.fh), pft = shup(F.ft), pfn = shup(F.fn - F.sR);
          T rx_h = (pfh - F.fh) / k.dx;
          T rx_u = (pfn - (F.fn - F.sL)) / k.dx;
          T rx_v = (pft - F.ft) / k.dx;
          T th = rx_h + ry_h, tu = rx_u + ry_u, tv = rx_v + ry_v;
          if (lane_out) {
            int64_t i = mbase + (int64_t)o * f.nx + col;
            int64_t cell = (int64_t)o * f.nx + col;
            if (a.rain != nullptr) th = th + a.rain[(int64_t)b * a.rain_bstride + cell];
            if (a.infil != nullptr) th = th - clamp_lo(a.infil[(int64_t)b * a.infil_bstride + cell], T(0));
            if (STAGE == 1) {
              a.out[i] = th;
              a.out[i + f.plane] = tu;
              a.out[i + 2 * f.plane] = tv;
            } else {
              if (a.k2_out != nullptr) {
                a.k2_out[i] = th;
                a.k2_out[i + f.plane] = tu;
                a.k2_out[i + 2 * f.plane] = tv;
              }
              T oh = a.U[i], ou = a.U[i + f.plane], ov = a.U[i + 2 * f.plane];
              T nh = oh + dt * th, nu = ou + dt * tu, nv = ov + dt * tv;
              clean_and_limit(oh, ou, ov, nh, nu, nv, k, a.positivity != 0);
              if (a.friction) {
                T coef = f.rough[(int64_t)b * f.rough_bstride + cell];
                swe::friction(nh, nu, nv, coef, dt, k, a.friction_model);
              }
              a.out[i] = nh;
              a.out[i + f.plane] = nu;
              a.out[i + 2 * f.plane] = nv;
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
  }

  if (STAGE == 1 && a.record_wave) {
    ax = wave_max(ax);
    ay = wave_max(ay);
    if (lane == 0) {
      int slot = (wid + b * 7) & (kSlots - 1);
      atomicMax(&a.acc[slot], to_bits<T>(ax));
      atomicMax(&a.acc[kSlots + slot], to_bits<T>(ay));
    }
  }
}

// ------------------------------------------------------------------ control kernels
__global__ void zero_words(uint32_t* p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void run_begin_kernel(Control* c, uint32_t* acc_words, int n_acc_words, double duration,
                                 double cap) {
  int t = threadIdx.x;
  for (int i = t; i < n_acc_words; i += blockDim.x) acc_words[i] = 0u;
  if (t == 0) {
    c->elapsed[0] = 0.0;
    c->elapsed[1] = 0.0;
    c->duration = duration;
    double tol = 1e-12 * fabs(duration);
    c->tol = tol > 1e-12 ? tol : 1e-12;
    c->cap = cap;
    c->bad_dt = 0.0;
    c->done = 0;
    c->steps = 0;
    c->status = 0;
    c->last_dt = 0.0;
  }
}

template <class T>
__global__ void gauge_kernel(const T* U, const T* bed, int64_t plane, int nx, int64_t bed_bstride,
                             const int32_t* gauges, int n, T* depth_out, T* level_out, int step_index,
                             const Control* ctl) {
  int g = threadIdx.x;
  if (ctl != nullptr && ctl->done && ctl->steps <= step_index) return;
  if (g >= n) return;
  int64_t cell = (int64_t)gauges[2 * g] * nx + gauges[2 * g + 1];
  T h = U[cell];
  depth_out[(int64_t)step_index * n + g] = h;
  if (level_out != nullptr) level_out[(int64_t)step_index * n + g] = h + bed[cell];
}

}  // namespace swe

// ===================================================================== host side
using namespace swe;

namespace {

thread_local int g_last_error = 0;
thread_local char g_last_msg[256] = {0};

int fail(int code, const char* msg) {
  g_last_error = code;
  strncpy(g_last_msg, msg, sizeof(g_last_msg) - 1);
  return code;
}

constexpr size_t kCtrlBytes = 256;
template <class T> constexpr size_t acc_bytes() { return 3 * 2 * kSlots * sizeof(typename Bits<T>::U); }
constexpr size_t kAccRegion = 3 * 2 * kSlots * 8;  // sized for 8-byte words whatever the dtype
inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

template <class T> size_t workspace_bytes(const swe_grid* g) {
  size_t state = (size_t)g->batch * 3 * (size_t)g->ny * (size_t)g->nx * sizeof(T);
  return kCtrlBytes + kAccRegion + align256(state);
}

template <class T> struct WsView {
  Control* ctl;
  typename Bits<T>::U* acc;  // [3][2][kSlots]
  T* k1;
};

template <class T> WsView<T> ws_view(void* ws) {
  char* p = (char*)ws;
  WsView<T> v;
  v.ctl = (Control*)p;
  v.acc = (typename Bits<T>::U*)(p + kCtrlBytes);
  v.k1 = (T*)(p + kCtrlBytes + kAccRegion);
  return v;
}

int check_grid(const swe_grid* g) {
  if (g == nullptr) return fail(SWE_ERR_ARG, "grid is null");
  if (g->batch <= 0 || g->ny <= 0 || g->nx <= 0) return fail(SWE_ERR_SHAPE, "batch, ny, nx must be positive");
  if (g->ny < 2 || g->nx < 2)
    return fail(SWE_ERR_SHAPE, "ng cannot exceed either interior grid dimension (ng = 2)");
  if (!(g->bed_batch == 1 || g->bed_batch == g->batch))
    return fail(SWE_ERR_SHAPE, "bed batch must be 1 or the state batch");
  if (!(g->rough_batch == 1 || g->rough_batch == g->batch))
    return fail(SWE_ERR_SHAPE, "roughness batch must be 1 or the state batch");
  if (!(g->dx > 0.0) || !(g->dy > 0.0)) return fail(SWE_ERR_ARG, "dx and dy must be positive");
  return 0;
}

int check_cfg(const swe_config* c) {
  if (c == nullptr) return fail(SWE_ERR_ARG, "config is null");
  for (int s = 0; s < 4; ++s)
    if (c->boundary[s] < 0 || c->boundary[s] > 5) return fail(SWE_ERR_BC, "unsupported boundary type code");
  if (!(c->gravity > 0.0) || !(c->epsilon > 0.0) || !(c->dry_depth >= 0.0))
    return fail(SWE_ERR_ARG, "Require gravity > 0, epsilon > 0, dry_depth >= 0.");
  if (c->friction_model < 0 || c->friction_model > 1) return fail(SWE_ERR_ARG, "unknown friction model");
  return 0;
}

template <class T> SweConsts<T> make_consts(const swe_config* c, const swe_grid* g) {
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
  k.dx = (T)g->dx;
  k.dy = (T)g->dy;
  k.cval = (T)c->constant_value;
  k.manning_g = (T)9.81;
  k.third = (T)(1.0 / 3.0);
  return k;
}

int pick_rows(const swe_grid* g, int rows_hint) {
  if (rows_hint > 0) return rows_hint;
  int strips = (g->nx + kOut - 1) / kOut;
  long long target = 4096;  // wavefronts wanted in flight
  long long per_col = (long long)strips * g->batch;
  int S = (int)((long long)g->ny * per_col / target);
  if (S < 8) S = 8;
  if (S > 64) S = 64;
  if (S > g->ny) S = g->ny;
  return S;
}

template <class T>
StageArgs<T> make_args(const swe_grid* g, const swe_config* c, const T* bed, const T* rough, const T* level,
                       int level_stride, int rows_hint) {
  StageArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.k = make_consts<T>(c, g);
  a.f.B = g->batch;
  a.f.ny = g->ny;
  a.f.nx = g->nx;
  a.f.plane = (int64_t)g->ny * g->nx;
  a.f.bed_bstride = g->bed_batch == 1 ? 0 : a.f.plane;
  a.f.rough_bstride = g->rough_batch == 1 ? 0 : a.f.plane;
  for (int s = 0; s < 4; ++s) a.f.bc[s] = c->boundary[s];
  a.f.bed = bed;
  a.f.rough = rough;
  a.f.level = level;
  a.f.level_stride = level_stride;
  a.S = pick_rows(g, rows_hint);
  a.strips = (g->nx + kOut - 1) / kOut;
  a.segs = (g->ny + a.S - 1) / a.S;
  a.wsr = c->water_slope_reset;
  a.positivity = c->positivity_limiter;
  a.friction = c->apply_friction;
  a.friction_model = c->friction_model;
  return a;
}

template <class T> dim3 stage_grid(const StageArgs<T>& a) {
  long long waves = (long long)a.strips * a.segs;
  return dim3((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock), (unsigned)a.f.B, 1);
}

template <class T, int STAGE> int launch_stage(const StageArgs<T>& a, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel<T, STAGE>), stage_grid(a), dim3(kWave * kWavesPerBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(e));
  return 0;
}

int check_level(const void* level, int level_stride, const swe_grid* g, const swe_config* c) {
  bool needs = false;
  for (int s = 0; s < 4; ++s) needs |= (c->boundary[s] == BC_WATER_LEVEL || c->boundary[s] == BC_ABSORBING);
  if (needs && level == nullptr) return fail(SWE_ERR_ARG, "boundary level pointer is required");
  if (level_stride != 0 && level_stride != 1) return fail(SWE_ERR_ARG, "level stride must be 0 or 1");
  (void)g;
  return 0;
}

template <class T>
int spatial_operator_impl(const swe_grid* g, const swe_config* c, const T* U, const T* bed, const T* level,
                          int level_stride, const T* rain, const T* infil, T* rhs, void* ws, int rows_hint,
                          void* stream) {
  int e;
  if ((e = check_grid(g)) || (e = check_cfg(c)) || (e = check_level(level, level_stride, g, c))) return e;
  if (!U || !bed || !rhs || !ws) return fail(SWE_ERR_ARG, "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  WsView<T> w = ws_view<T>(ws);
  auto* bank = w.acc + 2 * 2 * kSlots;  // bank 2: standalone operator calls
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(256), 0, s, (uint32_t*)bank,
                     (int64_t)(2 * kSlots * sizeof(typename Bits<T>::U) / 4));
  StageArgs<T> a = make_args<T>(g, c, bed, bed, level, level_stride, rows_hint);
  a.U = U;
  a.out = rhs;
  a.rain = rain;
  a.infil = infil;
  a.rain_bstride = a.infil_bstride = 0;
  a.acc = bank;
  a.record_wave = 1;
  return launch_stage<T, 1>(a, s);
}

template <class T>
int step_impl(const swe_grid* g, const swe_config* c, int stages, const T* U, const T* bed, const T* rough,
              const T* level, int level_stride, double max_dt, int parity, T* U_out, void* ws, int rows_hint,
              void* stream, T* k1_rec = nullptr, T* k2_rec = nullptr) {
  int e;
  if ((e = check_grid(g)) || (e = check_cfg(c)) || (e = check_level(level, level_stride, g, c))) return e;
  if (stages < 1 || stages > 3) return fail(SWE_ERR_ARG, "stages must be 1, 2 or 3");
  if (!U || !bed || !ws || ((stages & 2) && (!U_out || (c->apply_friction && !rough))))
    return fail(SWE_ERR_ARG, "null pointer argument");
  if (U == U_out) return fail(SWE_ERR_ARG, "U and U_out must not alias");
  hipStream_t s = (hipStream_t)stream;
  WsView<T> w = ws_view<T>(ws);
  int p = parity & 1;
  StageArgs<T> a = make_args<T>(g, c, bed, rough, level, level_stride, rows_hint);
  T* k1buf = k1_rec != nullptr ? k1_rec : w.k1;
  a.U = U;
  a.out = k1buf;
  a.acc = w.acc + p * 2 * kSlots;
  a.acc_next = w.acc + (1 - p) * 2 * kSlots;
  a.record_wave = 1;
  a.max_dt = max_dt;
  a.ctl = w.ctl;
  a.clock = 0;
  if ((stages & 1) && (e = launch_stage<T, 1>(a, s))) return e;
  if (stages & 2) {
    StageArgs<T> b2 = a;
    b2.k1 = k1buf;
    b2.k2_out = k2_rec;
    b2.out = U_out;
    b2.record_wave = 0;
    if ((e = launch_stage<T, 2>(b2, s))) return e;
  }
  return 0;
}

template <class T>
int advance_impl(const swe_grid* g, const swe_config* c, const T* bed, const T* rough, const T* level,
                 int level_stride, T* U_a, T* U_b, int first_step, int n_steps, T* hmax, T* dt_series,
                 int dt_capacity, const int32_t* gauges, int n_gauges, T* gauge_depth, T* gauge_level,
                 int gauge_capacity, void* ws, int rows_hint, void* stream) {
  int e;
  if ((e = check_grid(g)) || (e = check_cfg(c)) || (e = check_level(level, level_stride, g, c))) return e;
  if (!U_a || !U_b || !bed || !ws || (c->apply_friction && !rough)) return fail(SWE_ERR_ARG, "null pointer argument");
  if (first_step < 0 || n_steps < 0) return fail(SWE_ERR_ARG, "negative step index");
  if (dt_series && first_step + n_steps > dt_capacity) return fail(SWE_ERR_CAPACITY, "dt series capacity exceeded");
  if (n_gauges > 0 && (!gauges || !gauge_depth || first_step + n_steps > gauge_capacity))
    return fail(SWE_ERR_CAPACITY, "gauge series capacity exceeded");
  if (n_gauges > 1024) return fail(SWE_ERR_ARG, "at most 1024 gauges");
  hipStream_t s = (hipStream_t)stream;
  WsView<T> w = ws_view<T>(ws);
  StageArgs<T> a = make_args<T>(g, c, bed, rough, level, level_stride, rows_hint);
  a.ctl = w.ctl;
  a.clock = 1;
  a.hmax = hmax;
  a.dt_series = dt_series;
  for (int k = first_step; k < first_step + n_steps; ++k) {
    int p = k & 1;
    T* src = p == 0 ? U_a : U_b;
    T* dst = p == 0 ? U_b : U_a;
    StageArgs<T> s1 = a;
    s1.U = src;
    s1.out = w.k1;
    s1.acc = w.acc + p * 2 * kSlots;
    s1.acc_next = w.acc + (1 - p) * 2 * kSlots;
    s1.parity = p;
    s1.step_index = k;
    s1.record_wave = 1;
    if ((e = launch_stage<T, 1>(s1, s))) return e;
    StageArgs<T> s2 = s1;
    s2.k1 = w.k1;
    s2.out = dst;
    s2.record_wave = 0;
    if ((e = launch_stage<T, 2>(s2, s))) return e;
    if (n_gauges > 0) {
      hipLaunchKernelGGL((gauge_kernel<T>), dim3(1), dim3(1024), 0, s, (const T*)dst, bed, a.f.plane, g->nx,
                         a.f.bed_bstride, gauges, n_gauges, gauge_depth, gauge_level, k, (const Control*)w.ctl);
      hipError_t he = hipGetLastError();
      if (he != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(he));
    }
  }
  return 0;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int swe_abi_version(void) { return SWE_ABI_VERSION; }

const char* swe_strerror(int code) {
  switch (code) {
    case SWE_OK: return "ok";
    case SWE_ERR_ARG: return "invalid argument";
    case SWE_ERR_SHAPE: return "invalid shape";
    case SWE_ERR_BC: return "unsupported boundary type";
    case SWE_ERR_HIP: return "HIP runtime error";
    case SWE_ERR_DT_NONFINITE: return "The calculated timestep is non-finite.";
    case SWE_ERR_DT_NONPOSITIVE: return "Non-positive timestep";
    case SWE_ERR_CAPACITY: return "capacity exceeded";
    default: return "unknown error";
  }
}

const char* swe_last_error_message(void) { return g_last_msg; }

size_t swe_workspace_bytes(const swe_grid* g, int dtype_bytes) {
  if (check_grid(g)) return 0;
  return dtype_bytes == 8 ? workspace_bytes<double>(g) : workspace_bytes<float>(g);
}

size_t swe_control_bytes(void) { return sizeof(Control); }

int swe_workspace_init(void* ws, size_t bytes, void* stream) {
  if (!ws || bytes < kCtrlBytes + kAccRegion) return fail(SWE_ERR_ARG, "workspace too small");
  int64_t words = (int64_t)(kCtrlBytes + kAccRegion) / 4;
  hipLaunchKernelGGL(zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (uint32_t*)ws, words);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(SWE_ERR_HIP, hipGetErrorString(e));
}

int swe_spatial_operator_f32(const swe_grid* g, const swe_config* c, const float* U, const float* bed,
                             const float* level, int level_stride, const float* rain, const float* infil,
                             float* rhs, void* ws, int rows_hint, void* stream) {
  return spatial_operator_impl<float>(g, c, U, bed, level, level_stride, rain, infil, rhs, ws, rows_hint, stream);
}

int swe_spatial_operator_f64(const swe_grid* g, const swe_config* c, const double* U, const double* bed,
                             const double* level, int level_stride, const double* rain, const double* infil,
                             double* rhs, void* ws, int rows_hint, void* stream) {
  return spatial_operator_impl<double>(g, c, U, bed, level, level_stride, rain, infil, rhs, ws, rows_hint, stream);
}

int swe_step_f32(const swe_grid* g, const swe_config* c, const float* U, const float* bed, const float* rough,
                 const float* level, int level_stride, double max_dt, int parity, float* U_out, void* ws,
                 int rows_hint, void* stream) {
  return step_impl<float>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint,
                          stream);
}

int swe_step_f64(const swe_grid* g, const swe_config* c, const double* U, const double* bed, const double* rough,
                 const double* level, int level_stride, double max_dt, int parity, double* U_out, void* ws,
                 int rows_hint, void* stream) {
  return step_impl<double>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint,
                           stream);
}

int swe_step_record_f32(const swe_grid* g, const swe_config* c, const float* U, const float* bed,
                        const float* rough, const float* level, int level_stride, double max_dt, int parity,
                        float* U_out, float* k1_out, float* k2_out, void* ws, int rows_hint, void* stream) {
  if (!k1_out || !k2_out) return fail(SWE_ERR_ARG, "k1_out and k2_out are required");
  return step_impl<float>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint,
                          stream, k1_out, k2_out);
}

int swe_step_record_f64(const swe_grid* g, const swe_config* c, const double* U, const double* bed,
                        const double* rough, const double* level, int level_stride, double max_dt, int parity,
                        double* U_out, double* k1_out, double* k2_out, void* ws, int rows_hint, void* stream) {
  if (!k1_out || !k2_out) return fail(SWE_ERR_ARG, "k1_out and k2_out are required");
  return step_impl<double>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint,
                           stream, k1_out, k2_out);
}

int swe_step_stage_f32(const swe_grid* g, const swe_config* c, int stage, const float* U, const float* bed,
                       const float* rough, const float* level, int level_stride, double max_dt, int parity,
                       float* U_out, void* ws, int rows_hint, void* stream) {
  if (stage != 1 && stage != 2) return fail(SWE_ERR_ARG, "stage must be 1 or 2");
  return step_impl<float>(g, c, stage, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint,
                          stream);
}

int swe_run_begin(void* ws, double duration, double max_dt_cap, int dtype_bytes, void* stream) {
  if (!ws) return fail(SWE_ERR_ARG, "null workspace");
  if (!(duration >= 0.0)) return fail(SWE_ERR_ARG, "duration must be non-negative");
  size_t acc = dtype_bytes == 8 ? acc_bytes<double>() : acc_bytes<float>();
  hipLaunchKernelGGL(run_begin_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, (Control*)ws,
                     (uint32_t*)((char*)ws + kCtrlBytes), (int)(acc / 4), duration, max_dt_cap);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(SWE_ERR_HIP, hipGetErrorString(e));
}

int swe_advance_f32(const swe_grid* g, const swe_config* c, const float* bed, const float* rough,
                    const float* level, int level_stride, float* U_a, float* U_b, int first_step, int n_steps,
                    float* hmax, float* dt_series, int dt_capacity, const int32_t* gauges, int n_gauges,
                    float* gauge_depth, float* gauge_level, int gauge_capacity, void* ws, int rows_hint,
                    void* stream) {
  return advance_impl<float>(g, c, bed, rough, level, level_stride, U_a, U_b, first_step, n_steps, hmax,
                             dt_series, dt_capacity, gauges, n_gauges, gauge_depth, gauge_level, gauge_capacity,
                             ws, rows_hint, stream);
}

int swe_advance_f64(const swe_grid* g, const swe_config* c, const double* bed, const double* rough,
                    const double* level, int level_stride, double* U_a, double* U_b, int first_step,
                    int n_steps, double* hmax, double* dt_series, int dt_capacity, const int32_t* gauges,
                    int n_gauges, double* gauge_depth, double* gauge_level, int gauge_capacity, void* ws,
                    int rows_hint, void* stream) {
  return advance_impl<double>(g, c, bed, rough, level, level_stride, U_a, U_b, first_step, n_steps, hmax,
                              dt_series, dt_capacity, gauges, n_gauges, gauge_depth, gauge_level, gauge_capacity,
                              ws, rows_hint, stream);
}

}  // extern "C"
