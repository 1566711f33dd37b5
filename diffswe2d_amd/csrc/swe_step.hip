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

#include <algorithm>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <limits>
#include <mutex>

#include "swe_common.h"
#include "swe_host.h"
#include "swe_fetch.h"
#include "swe_stage.h"
#include "../../include/swe2d_hip.h"

namespace swe {

template <class T, int STAGE, class A, bool SRC, bool K2>
__global__ __launch_bounds__(256) void stage_kernel(StageArgs<T> a) {
  const int lane = threadIdx.x & (kWave - 1);
  // the wavefront's index, made wave-uniform for the compiler (it cannot see that
  // threadIdx.x >> 6 is): the strip's rows and columns then live in scalar registers
  const int wid = blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.y;
  const int strip = wid % a.strips;
  const int seg = wid / a.strips;
  if (seg >= a.segs) return;  // uniform per wavefront
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
    dt = a.fixed_dt ? (T)a.max_dt : step_dt<T>(a.acc, k, (T)a.max_dt);
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
      if (a.dt_out != nullptr) *a.dt_out = dt;
    }
    if (!ok) return;
    half_dt = T(0.5) * dt;
  }

  T ax, ay;
  RainAt<T> rec{0, 0, 0, T(0)};
  if constexpr (SRC) {
    if (a.rain_rates != nullptr) {
      // model_time = U.new_tensor(start + elapsed) (solver.py:751-757); stage 2 samples
      // the record at time + 0.5 dt in double (:634-637)
      double t = a.clock ? a.time0 + a.ctl->elapsed[a.parity] : a.time0;
      double ts = (double)(T)t;
      rec = rain_at<T>(a.rain_times, a.rain_nt, STAGE == 1 ? ts : ts + 0.5 * (double)dt);
    }
  }
  stage_strip<T, STAGE, A, SRC, K2>(a, strip, seg, b, dt, half_dt, ax, ay, a.rain_rates != nullptr, rec);
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
// Water volume sum h * dx * dy of every member, in FP64 and deterministic: each block
// sums a fixed stride of the member's depth plane, one wave per member adds the
// block partials in index order.
template <class T>
__global__ __launch_bounds__(256) void volume_partial_kernel(const T* __restrict__ U, int64_t plane, int nblk,
                                                             double* __restrict__ partial) {
  const int b = blockIdx.y;
  const T* h = U + (int64_t)b * 3 * plane;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < plane; i += (int64_t)nblk * 256) s += (double)h[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double wave_sum[4];
  if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[(int64_t)b * nblk + blockIdx.x] = (wave_sum[0] + wave_sum[1]) + (wave_sum[2] + wave_sum[3]);
}

__global__ __launch_bounds__(64) void volume_finish_kernel(const double* __restrict__ partial, int nblk,
                                                           double cell_area, double* __restrict__ out) {
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 64) s += partial[(int64_t)b * nblk + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) out[b] = s * cell_area;
}

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
    c->chunked = 0;
    c->chunk = 0;
    c->n_chunks = 0;
    c->levels = nullptr;
  }
}

// Arms a chunked run: every chunk lasts `duration`, chunk c uses levels[c].
template <class T>
__global__ void run_begin_chunked_kernel(Control* c, int n_chunks, const T* levels) {
  if (threadIdx.x != 0) return;
  c->chunked = 1;
  c->chunk = 0;
  c->n_chunks = n_chunks;
  c->levels = levels;
  c->level_f64 = (double)levels[0];
  c->level_f32 = (float)levels[0];
}

// After stage 2 of a chunked step: record the step's chunk; at a chunk end sample
// h + z at the gauges (member 0) from the step's output, then move to the next
// chunk (its level, elapsed time 0) or finish the run after the last one.
template <class T>
__global__ void chunk_clock_kernel(Control* c, const T* U_out, const T* bed, int nx, const int32_t* gauges,
                                   int n_gauges, T* obs, int32_t* chunk_series, int step_index, int parity) {
  __shared__ int complete;
  if (threadIdx.x == 0) {
    complete = 0;
    if (!c->done) {
      if (chunk_series != nullptr) chunk_series[step_index] = c->chunk;
      complete = (c->duration - c->elapsed[1 - parity]) <= c->tol;
    }
  }
  __syncthreads();
  if (!complete) return;
  const int chunk = c->chunk;
  for (int g = threadIdx.x; g < n_gauges; g += blockDim.x) {
    int64_t cell = (int64_t)gauges[2 * g] * nx + gauges[2 * g + 1];
    obs[(int64_t)chunk * n_gauges + g] = U_out[cell] + bed[cell];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int next = chunk + 1;
    c->chunk = next;
    if (next >= c->n_chunks) {
      c->done = 1;
    } else {
      c->elapsed[1 - parity] = 0.0;
      T lv = ((const T*)c->levels)[next];
      c->level_f64 = (double)lv;
      c->level_f32 = (float)lv;
    }
  }
}

// Dynamic boundary series (solver.py:391-415), evaluated before each step of a
// device-driven run: the level at the step's start time start + elapsed, in the
// run's dtype -- time rounded like U.new_tensor(current_time), the interval from
// searchsorted(times, time) clamped to [1, n - 1] (n = 1: both ends are sample 0,
// as times[-1] is), weight (time - t0) / clamp(t1 - t0, min = eps), level
// w0 + weight * (w1 - w0) -- written where the stage kernels read the level.
// Optionally records each step's level, interval index and weight.
template <class T>
__global__ void dynamic_level_kernel(Control* c, const T* times, const T* levels, int n, double start, T eps,
                                     int parity, int step_index, T* step_level, int32_t* step_idx, T* step_weight) {
  if (threadIdx.x != 0 || c->done) return;
  const T t = (T)(start + c->elapsed[parity]);
  int lo = 0, hi = n;
  while (lo < hi) {  // first index with times[i] >= t
    const int mid = (lo + hi) >> 1;
    if (times[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  const int i = n > 1 ? (lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo)) : 0;
  const int i0 = i > 0 ? i - 1 : n - 1;
  const T t0 = times[i0], t1 = times[i];
  const T w0 = levels[i0], w1 = levels[i];
  const T weight = (t - t0) / tmax(t1 - t0, eps);
  const T level = w0 + weight * (w1 - w0);
  if (sizeof(T) == 8) c->level_f64 = (double)level;
  else c->level_f32 = (float)level;
  if (step_level != nullptr) step_level[step_index] = level;
  if (step_idx != nullptr) step_idx[step_index] = i;
  if (step_weight != nullptr) step_weight[step_index] = weight;
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
  if (c->arithmetic != SWE_ARITH_REFERENCE && c->arithmetic != SWE_ARITH_FAST)
    return fail(SWE_ERR_ARG, "unknown arithmetic mode");
  if (c->fixed_dt != 0 && c->fixed_dt != 1) return fail(SWE_ERR_ARG, "fixed_dt must be 0 or 1");
  return 0;
}


int pick_rows(const swe_grid* g, int rows_hint) {
  if (rows_hint > 0) return rows_hint;
  int strips = (g->nx + kOut - 1) / kOut;
  long long target = 4096;  // wavefronts wanted in flight
  long long per_col = (long long)strips * g->batch;
  int S = (int)((long long)g->ny * per_col / target);
  // small grids: short strips, so that enough wavefronts run at once; each
  // wavefront's rows form one dependent chain, whose latency sets the step time.
  // One-row strips when even they give at most one wavefront per SIMD (1,024 on
  // the chip): the chain is then shortest (5 rows instead of 6) and nothing waits
  // for a SIMD; two-row strips otherwise.
  if (S < 2) S = (long long)g->ny * per_col <= 1024 ? 1 : 2;
  // large grids: 48 rows (measured on 8192^2: 48 rows 1.763 ms per step, 32 rows
  // 1.778, 64 rows 1.793, 128 rows 1.928; taller strips pay less halo but leave
  // fewer, longer wavefronts to balance the SIMDs)
  if (S > 48) S = 48;
  if (S > g->ny) S = g->ny;
  return S;
}

template <class T>
StageArgs<T> make_args(const swe_grid* g, const swe_config* c, const T* bed, const T* rough, const T* level,
                       int level_stride, int rows_hint) {
  StageArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.k = host_consts<T>(c, g->dx, g->dy);
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
  a.fast = c->arithmetic == SWE_ARITH_FAST;
  a.fixed_dt = c->fixed_dt;
  return a;
}

template <class T> dim3 stage_grid(const StageArgs<T>& a) {
  long long waves = (long long)a.strips * a.segs;
  return dim3((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock), (unsigned)a.f.B, 1);
}

template <class T, int STAGE, class A, bool SRC> void launch_stage_src(const StageArgs<T>& a, hipStream_t s) {
  const dim3 grid = stage_grid(a), block(kWave * kWavesPerBlock);
  if (STAGE == 2 && a.k2_out != nullptr)
    hipLaunchKernelGGL((stage_kernel<T, STAGE, A, SRC, STAGE == 2>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((stage_kernel<T, STAGE, A, SRC, false>), grid, block, 0, s, a);
}

template <class T, int STAGE, class A> void launch_stage_as(const StageArgs<T>& a, hipStream_t s) {
  if (stage_has_sources(a)) launch_stage_src<T, STAGE, A, true>(a, s);
  else launch_stage_src<T, STAGE, A, false>(a, s);
}

template <class T, int STAGE> int launch_stage(const StageArgs<T>& a, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (a.fast) launch_stage_as<T, STAGE, FastRound>(a, s);
    else launch_stage_as<T, STAGE, RefRound>(a, s);
  } else {
    launch_stage_as<T, STAGE, RefRound>(a, s);
  }
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

template <class T> int apply_forcing(StageArgs<T>& a, const swe_grid* g, const swe_forcing* fz) {
  if (fz == nullptr) return 0;
  if ((fz->rain_rates != nullptr) != (fz->rain_times != nullptr) || (fz->rain_rates && fz->rain_nt < 1))
    return fail(SWE_ERR_ARG, "rainfall record needs times, rates and rain_nt >= 1");
  if (fz->infiltration && fz->infil_bstride != 0 && fz->infil_bstride != (int64_t)g->ny * g->nx)
    return fail(SWE_ERR_ARG, "infiltration member stride must be 0 or ny * nx");
  a.rain_times = (const T*)fz->rain_times;
  a.rain_rates = (const T*)fz->rain_rates;
  a.rain_nt = fz->rain_nt;
  a.time0 = fz->time;
  a.infil = (const T*)fz->infiltration;
  a.infil_bstride = fz->infil_bstride;
  return 0;
}

template <class T>
int step_impl(const swe_grid* g, const swe_config* c, int stages, const T* U, const T* bed, const T* rough,
              const T* level, int level_stride, double max_dt, int parity, T* U_out, void* ws, int rows_hint,
              void* stream, T* k1_rec = nullptr, T* k2_rec = nullptr, const swe_forcing* fz = nullptr,
              T* dt_out = nullptr) {
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
  if ((e = apply_forcing<T>(a, g, fz))) return e;
  T* k1buf = k1_rec != nullptr ? k1_rec : w.k1;
  if (k1_rec != nullptr) a.fast = 0;  // a recorded step feeds the adjoint, which replays the reference rounding
  a.U = U;
  a.out = k1buf;
  a.acc = w.acc + p * 2 * kSlots;
  a.acc_next = w.acc + (1 - p) * 2 * kSlots;
  a.record_wave = 1;
  a.max_dt = max_dt;
  a.ctl = w.ctl;
  a.clock = 0;
  a.dt_out = dt_out;
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
                 int gauge_capacity, void* ws, int rows_hint, void* stream, T* tape_k1 = nullptr,
                 T* tape_k2 = nullptr, int tape_first = 0, const int32_t* chunk_gauges = nullptr,
                 int chunk_n_gauges = 0, T* chunk_obs = nullptr, int32_t* chunk_series = nullptr,
                 bool chunked = false, const swe_forcing* fz = nullptr, const swe_dynamic_level* dyn = nullptr) {
  int e;
  if ((e = check_grid(g)) || (e = check_cfg(c)) || (e = check_level(level, level_stride, g, c))) return e;
  if (!U_a || (!U_b && !tape_k1) || !bed || !ws || (c->apply_friction && !rough))
    return fail(SWE_ERR_ARG, "null pointer argument");
  if (tape_k1 && (!tape_k2 || first_step < tape_first)) return fail(SWE_ERR_ARG, "bad tape arguments");
  if (first_step < 0 || n_steps < 0) return fail(SWE_ERR_ARG, "negative step index");
  if (c->fixed_dt) return fail(SWE_ERR_ARG, "fixed_dt applies to host-driven steps; a device-driven run picks its dt");
  if (dt_series && first_step + n_steps > dt_capacity) return fail(SWE_ERR_CAPACITY, "dt series capacity exceeded");
  if (n_gauges > 0 && (!gauges || !gauge_depth || first_step + n_steps > gauge_capacity))
    return fail(SWE_ERR_CAPACITY, "gauge series capacity exceeded");
  if (n_gauges > 1024) return fail(SWE_ERR_ARG, "at most 1024 gauges");
  hipStream_t s = (hipStream_t)stream;
  WsView<T> w = ws_view<T>(ws);
  StageArgs<T> a = make_args<T>(g, c, bed, rough, level, level_stride, rows_hint);
  if ((e = apply_forcing<T>(a, g, fz))) return e;
  a.ctl = w.ctl;
  a.clock = 1;
  a.hmax = hmax;
  a.dt_series = dt_series;
  if (tape_k1 != nullptr || chunked) a.fast = 0;  // taped / chunked runs feed the adjoint
  if (dyn != nullptr && (dyn->step_level || dyn->step_index || dyn->step_weight) &&
      first_step + n_steps > dt_capacity)
    return fail(SWE_ERR_CAPACITY, "dynamic level record capacity exceeded");
  const size_t state = (size_t)g->batch * 3 * (size_t)g->ny * g->nx;
  for (int k = first_step; k < first_step + n_steps; ++k) {
    int p = k & 1;
    T* src = p == 0 ? U_a : U_b;
    T* dst = p == 0 ? U_b : U_a;
    T* k1 = w.k1;
    T* k2 = nullptr;
    if (tape_k1) {  // tape: U_a holds the states (slot j = step tape_first + j), k1/k2 per step
      size_t j = (size_t)(k - tape_first);
      src = U_a + j * state;
      dst = U_a + (j + 1) * state;
      k1 = tape_k1 + j * state;
      k2 = tape_k2 + j * state;
    }
    if (dyn != nullptr) {
      hipLaunchKernelGGL((dynamic_level_kernel<T>), dim3(1), dim3(64), 0, s, w.ctl, (const T*)dyn->times,
                         (const T*)dyn->levels, dyn->n, dyn->start_time, std::numeric_limits<T>::epsilon(), p, k,
                         (T*)dyn->step_level, dyn->step_index, (T*)dyn->step_weight);
      hipError_t he = hipGetLastError();
      if (he != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(he));
    }
    StageArgs<T> s1 = a;
    s1.U = src;
    s1.out = k1;
    s1.acc = w.acc + p * 2 * kSlots;
    s1.acc_next = w.acc + (1 - p) * 2 * kSlots;
    s1.parity = p;
    s1.step_index = k;
    s1.record_wave = 1;
    if ((e = launch_stage<T, 1>(s1, s))) return e;
    StageArgs<T> s2 = s1;
    s2.k1 = k1;
    s2.k2_out = k2;
    s2.out = dst;
    s2.record_wave = 0;
    if ((e = launch_stage<T, 2>(s2, s))) return e;
    if (chunked) {
      hipLaunchKernelGGL((chunk_clock_kernel<T>), dim3(1), dim3(256), 0, s, w.ctl, (const T*)dst, bed, g->nx,
                         chunk_gauges, chunk_n_gauges, chunk_obs, chunk_series, k, p);
      hipError_t he = hipGetLastError();
      if (he != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(he));
    }
    if (n_gauges > 0) {
      hipLaunchKernelGGL((gauge_kernel<T>), dim3(1), dim3(1024), 0, s, (const T*)dst, bed, a.f.plane, g->nx,
                         a.f.bed_bstride, gauges, n_gauges, gauge_depth, gauge_level, k, (const Control*)w.ctl);
      hipError_t he = hipGetLastError();
      if (he != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(he));
    }
  }
  return 0;
}

template <class T>
int advance_chunked(const swe_grid* g, const swe_config* c, const T* bed, const T* rough, T* U_a, T* U_b, T* tape_k1,
                    T* tape_k2, int tape_first, int first_step, int n_steps, T* dt_series, int32_t* chunk_series,
                    int capacity, const int32_t* gauges, int n_gauges, T* obs, void* ws, int rows_hint,
                    void* stream) {
  if (!ws || (n_gauges > 0 && (!gauges || !obs))) return fail(SWE_ERR_ARG, "null pointer argument");
  if (n_gauges < 0 || n_gauges > 1024) return fail(SWE_ERR_ARG, "0 to 1024 gauges");
  if ((dt_series || chunk_series) && first_step + n_steps > capacity)
    return fail(SWE_ERR_CAPACITY, "step series capacity exceeded");
  // the stage kernels read the chunk's level from the control block
  Control* ctl = (Control*)ws;
  const T* level = sizeof(T) == 8 ? (const T*)&ctl->level_f64 : (const T*)&ctl->level_f32;
  return advance_impl<T>(g, c, bed, rough, level, 0, U_a, U_b, first_step, n_steps, nullptr, dt_series, capacity,
                         nullptr, 0, nullptr, nullptr, 0, ws, rows_hint, stream, tape_k1, tape_k2, tape_first, gauges,
                         n_gauges, obs, chunk_series, true);
}

}  // namespace

// ------------------------------------------------------------------ step without allocation
// Scratch of swe_step_noalloc: the workspace layout of one step (control block,
// wave-speed banks, first-stage tendency) in a fixed array that each device holds
// from module load to the end of the process.  Host code never uses the array's
// name as an address: the device address comes from hipGetSymbolAddress, once
// per device, on the first call with that device current.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(hipStream_t stream, void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(SWE_ERR_ARG, "device index beyond the scratch table");
  if (stream != nullptr) {
    int sdev = -1;
    e = hipStreamGetDevice(stream, &sdev);
    if (e != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(e));
    if (sdev != dev) return fail(SWE_ERR_ARG, "the stream belongs to a device other than the current one");
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(SWE_ERR_HIP, e != hipSuccess ? hipGetErrorString(e) : "no address for the step scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

template <class T>
int step_noalloc_impl(const swe_grid* g, const swe_config* c, const T* U, const T* bed, const T* rough,
                      const T* level, int level_stride, double max_dt, T* U_out, T* dt_out, int rows_hint,
                      void* stream) {
  int e;
  if ((e = check_grid(g))) return e;
  if (!dt_out) return fail(SWE_ERR_ARG, "dt_out is required");
  const size_t need = workspace_bytes<T>(g);
  if (need > kNoAllocScratchBytes) {
    char msg[200];
    snprintf(msg, sizeof(msg),
             "batch * ny * nx = %lld cells need %zu scratch bytes; the fixed scratch holds %zu (%lld cells)",
             (long long)g->batch * g->ny * g->nx, need, kNoAllocScratchBytes,
             (long long)((kNoAllocScratchBytes - kCtrlBytes - kAccRegion) / (3 * sizeof(T))));
    return fail(SWE_ERR_CAPACITY, msg);
  }
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  if ((e = scratch_address(s, &ws))) return e;
  // the step accumulates its wave speeds into bank 0, which the previous call left
  // holding its own: clear it with a kernel (no memset node in a captured graph)
  WsView<T> w = ws_view<T>(ws);
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(256), 0, s, (uint32_t*)w.acc,
                     (int64_t)(2 * kSlots * sizeof(typename Bits<T>::U) / 4));
  hipError_t he = hipGetLastError();
  if (he != hipSuccess) return fail(SWE_ERR_HIP, hipGetErrorString(he));
  return step_impl<T>(g, c, 3, U, bed, rough, level, level_stride, max_dt, 0, U_out, ws, rows_hint, stream, nullptr,
                      nullptr, nullptr, dt_out);
}

template <class T>
int advance_dynamic(const swe_grid* g, const swe_config* c, const swe_dynamic_level* dyn, const swe_forcing* fz,
                    const T* bed, const T* rough, T* U_a, T* U_b, T* tape_k1, T* tape_k2, int tape_first,
                    int first_step, int n_steps, T* hmax, T* dt_series, int capacity, const int32_t* gauges,
                    int n_gauges, T* gauge_depth, T* gauge_level, void* ws, int rows_hint, void* stream) {
  if (!dyn || !dyn->times || !dyn->levels || dyn->n < 1) return fail(SWE_ERR_ARG, "dynamic series needs times, levels, n >= 1");
  if (!ws) return fail(SWE_ERR_ARG, "null workspace");
  if ((tape_k1 == nullptr) != (tape_k2 == nullptr)) return fail(SWE_ERR_ARG, "record both tapes or neither");
  // the stage kernels read the step's level from the control block
  Control* ctl = (Control*)ws;
  const T* level = sizeof(T) == 8 ? (const T*)&ctl->level_f64 : (const T*)&ctl->level_f32;
  return advance_impl<T>(g, c, bed, rough, level, 0, U_a, U_b, first_step, n_steps, hmax, dt_series, capacity,
                         gauges, n_gauges, gauge_depth, gauge_level, capacity, ws, rows_hint, stream, tape_k1, tape_k2,
                         tape_first, nullptr, 0, nullptr, nullptr, false, fz, dyn);
}

}  // namespace

// ===================================================================== C ABI
template <class T>
int water_volume_impl(const swe_grid* g, const T* U, double* volume, void* ws, void* stream) {
  if (check_grid(g)) return g_last_error;
  if (!U || !volume || !ws) return fail(SWE_ERR_ARG, "U, volume and ws are required");
  const int64_t plane = (int64_t)g->ny * g->nx;
  // partials borrow the stage-tendency region of the workspace (3 * plane * sizeof(T) bytes per member),
  // which holds nothing between steps
  const int64_t room = 3 * plane * (int64_t)sizeof(T) / 8;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>({1024, room, (plane + 255) / 256}));
  double* partial = (double*)ws_view<T>(ws).k1;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(volume_partial_kernel<T>, dim3(nblk, g->batch), dim3(256), 0, s, U, plane, nblk, partial);
  hipLaunchKernelGGL(volume_finish_kernel, dim3(g->batch), dim3(64), 0, s, partial, nblk, g->dx * g->dy, volume);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(SWE_ERR_HIP, hipGetErrorString(e));
}

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

int swe_step_forced_f32(const swe_grid* g, const swe_config* c, const swe_forcing* fz, const float* U,
                        const float* bed, const float* rough, const float* level, int level_stride, double max_dt,
                        int parity, float* U_out, float* k1_out, float* k2_out, void* ws, int rows_hint,
                        void* stream) {
  if ((k1_out == nullptr) != (k2_out == nullptr)) return fail(SWE_ERR_ARG, "record both k1 and k2 or neither");
  return step_impl<float>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint, stream,
                          k1_out, k2_out, fz);
}

int swe_step_forced_f64(const swe_grid* g, const swe_config* c, const swe_forcing* fz, const double* U,
                        const double* bed, const double* rough, const double* level, int level_stride, double max_dt,
                        int parity, double* U_out, double* k1_out, double* k2_out, void* ws, int rows_hint,
                        void* stream) {
  if ((k1_out == nullptr) != (k2_out == nullptr)) return fail(SWE_ERR_ARG, "record both k1 and k2 or neither");
  return step_impl<double>(g, c, 3, U, bed, rough, level, level_stride, max_dt, parity, U_out, ws, rows_hint, stream,
                           k1_out, k2_out, fz);
}

int swe_advance_forced_f32(const swe_grid* g, const swe_config* c, const swe_forcing* fz, const float* bed,
                           const float* rough, const float* level, int level_stride, float* U_a, float* U_b,
                           int first_step, int n_steps, float* hmax, float* dt_series, int dt_capacity,
                           const int32_t* gauges, int n_gauges, float* gauge_depth, float* gauge_level,
                           int gauge_capacity, void* ws, int rows_hint, void* stream) {
  return advance_impl<float>(g, c, bed, rough, level, level_stride, U_a, U_b, first_step, n_steps, hmax, dt_series,
                             dt_capacity, gauges, n_gauges, gauge_depth, gauge_level, gauge_capacity, ws, rows_hint,
                             stream, nullptr, nullptr, 0, nullptr, 0, nullptr, nullptr, false, fz);
}

int swe_advance_forced_f64(const swe_grid* g, const swe_config* c, const swe_forcing* fz, const double* bed,
                           const double* rough, const double* level, int level_stride, double* U_a, double* U_b,
                           int first_step, int n_steps, double* hmax, double* dt_series, int dt_capacity,
                           const int32_t* gauges, int n_gauges, double* gauge_depth, double* gauge_level,
                           int gauge_capacity, void* ws, int rows_hint, void* stream) {
  return advance_impl<double>(g, c, bed, rough, level, level_stride, U_a, U_b, first_step, n_steps, hmax, dt_series,
                              dt_capacity, gauges, n_gauges, gauge_depth, gauge_level, gauge_capacity, ws, rows_hint,
                              stream, nullptr, nullptr, 0, nullptr, 0, nullptr, nullptr, false, fz);
}

int swe_run_begin_chunked(void* ws, double chunk_seconds, int n_chunks, const void* levels, int dtype_bytes,
                          void* stream) {
  if (!levels || n_chunks <= 0) return fail(SWE_ERR_ARG, "chunked runs need levels and n_chunks > 0");
  if (!(chunk_seconds > 0.0)) return fail(SWE_ERR_ARG, "chunk duration must be positive");
  int e = swe_run_begin(ws, chunk_seconds, -1.0, dtype_bytes, stream);
  if (e) return e;
  if (dtype_bytes == 8)
    hipLaunchKernelGGL((run_begin_chunked_kernel<double>), dim3(1), dim3(64), 0, (hipStream_t)stream, (Control*)ws,
                       n_chunks, (const double*)levels);
  else
    hipLaunchKernelGGL((run_begin_chunked_kernel<float>), dim3(1), dim3(64), 0, (hipStream_t)stream, (Control*)ws,
                       n_chunks, (const float*)levels);
  hipError_t he = hipGetLastError();
  return he == hipSuccess ? 0 : fail(SWE_ERR_HIP, hipGetErrorString(he));
}

int swe_advance_chunked_f32(const swe_grid* g, const swe_config* c, const float* bed, const float* rough,
                            float* U_a, float* U_b, float* tape_k1, float* tape_k2, int tape_first, int first_step,
                            int n_steps, float* dt_series, int32_t* chunk_series, int capacity,
                            const int32_t* gauges, int n_gauges, float* obs, void* ws, int rows_hint, void* stream) {
  return advance_chunked<float>(g, c, bed, rough, U_a, U_b, tape_k1, tape_k2, tape_first, first_step, n_steps,
                                dt_series, chunk_series, capacity, gauges, n_gauges, obs, ws, rows_hint, stream);
}

int swe_advance_chunked_f64(const swe_grid* g, const swe_config* c, const double* bed, const double* rough,
                            double* U_a, double* U_b, double* tape_k1, double* tape_k2, int tape_first,
                            int first_step, int n_steps, double* dt_series, int32_t* chunk_series, int capacity,
                            const int32_t* gauges, int n_gauges, double* obs, void* ws, int rows_hint,
                            void* stream) {
  return advance_chunked<double>(g, c, bed, rough, U_a, U_b, tape_k1, tape_k2, tape_first, first_step, n_steps,
                                 dt_series, chunk_series, capacity, gauges, n_gauges, obs, ws, rows_hint, stream);
}

int swe_advance_tape_f32(const swe_grid* g, const swe_config* c, const float* bed, const float* rough,
                         const float* level, int level_stride, float* tape_U, float* tape_k1, float* tape_k2,
                         int tape_first, int first_step, int n_steps, float* hmax, float* dt_series, int dt_capacity,
                         void* ws, int rows_hint, void* stream) {
  if (!tape_U || !tape_k1 || !tape_k2) return fail(SWE_ERR_ARG, "tape pointers are required");
  return advance_impl<float>(g, c, bed, rough, level, level_stride, tape_U, nullptr, first_step, n_steps, hmax,
                             dt_series, dt_capacity, nullptr, 0, nullptr, nullptr, 0, ws, rows_hint, stream, tape_k1,
                             tape_k2, tape_first);
}

int swe_advance_tape_f64(const swe_grid* g, const swe_config* c, const double* bed, const double* rough,
                         const double* level, int level_stride, double* tape_U, double* tape_k1, double* tape_k2,
                         int tape_first, int first_step, int n_steps, double* hmax, double* dt_series,
                         int dt_capacity, void* ws, int rows_hint, void* stream) {
  if (!tape_U || !tape_k1 || !tape_k2) return fail(SWE_ERR_ARG, "tape pointers are required");
  return advance_impl<double>(g, c, bed, rough, level, level_stride, tape_U, nullptr, first_step, n_steps, hmax,
                              dt_series, dt_capacity, nullptr, 0, nullptr, nullptr, 0, ws, rows_hint, stream, tape_k1,
                              tape_k2, tape_first);
}

int swe_step_noalloc(const swe_grid* g, const swe_config* c, int dtype_bytes, const void* U, const void* bed,
                     const void* rough, const void* level, int level_stride, double max_dt, void* U_out,
                     void* dt_out, int rows_hint, void* stream) {
  if (dtype_bytes == 4)
    return step_noalloc_impl<float>(g, c, (const float*)U, (const float*)bed, (const float*)rough,
                                    (const float*)level, level_stride, max_dt, (float*)U_out, (float*)dt_out,
                                    rows_hint, stream);
  if (dtype_bytes == 8)
    return step_noalloc_impl<double>(g, c, (const double*)U, (const double*)bed, (const double*)rough,
                                     (const double*)level, level_stride, max_dt, (double*)U_out, (double*)dt_out,
                                     rows_hint, stream);
  return fail(SWE_ERR_ARG, "dtype_bytes must be 4 or 8");
}

int swe_advance_dynamic_f32(const swe_grid* g, const swe_config* c, const swe_dynamic_level* dyn,
                            const swe_forcing* fz, const float* bed, const float* rough, float* U_a, float* U_b,
                            float* tape_k1, float* tape_k2, int tape_first, int first_step, int n_steps, float* hmax,
                            float* dt_series, int capacity, const int32_t* gauges, int n_gauges, float* gauge_depth,
                            float* gauge_level, void* ws, int rows_hint, void* stream) {
  return advance_dynamic<float>(g, c, dyn, fz, bed, rough, U_a, U_b, tape_k1, tape_k2, tape_first, first_step, n_steps,
                                hmax, dt_series, capacity, gauges, n_gauges, gauge_depth, gauge_level, ws, rows_hint,
                                stream);
}

int swe_advance_dynamic_f64(const swe_grid* g, const swe_config* c, const swe_dynamic_level* dyn,
                            const swe_forcing* fz, const double* bed, const double* rough, double* U_a, double* U_b,
                            double* tape_k1, double* tape_k2, int tape_first, int first_step, int n_steps,
                            double* hmax, double* dt_series, int capacity, const int32_t* gauges, int n_gauges,
                            double* gauge_depth, double* gauge_level, void* ws, int rows_hint, void* stream) {
  return advance_dynamic<double>(g, c, dyn, fz, bed, rough, U_a, U_b, tape_k1, tape_k2, tape_first, first_step,
                                 n_steps, hmax, dt_series, capacity, gauges, n_gauges, gauge_depth, gauge_level, ws,
                                 rows_hint, stream);
}

int swe_water_volume_f32(const swe_grid* g, const float* U, double* volume, void* ws, void* stream) {
  return water_volume_impl<float>(g, U, volume, ws, stream);
}

int swe_water_volume_f64(const swe_grid* g, const double* U, double* volume, void* ws, void* stream) {
  return water_volume_impl<double>(g, U, volume, ws, stream);
}

}  // extern "C"
