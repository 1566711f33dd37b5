// Multi-scenario ensemble of forward runs on gfx950 (FP32).
//
// Each scenario has its own grid, bed, roughness, piecewise-constant boundary
// level per output chunk, gauges and its own adaptive dt.  All scenarios advance
// in lockstep "global steps"; one global step is three launches over every
// unfinished scenario:
//   ens_stage<1>  k1 and wave-speed maxima (one accumulator bank per scenario)
//   ens_stage<2>  dt of each scenario from its own maxima and the time left in
//                 its current chunk; predictor, k2, update, friction
//   ens_clock     per scenario: elapsed time, chunk boundary, gauge sample at
//                 each chunk end (h + z), clearing of the next accumulator bank
// A scenario's chunk logic is SWE2D.forward(U, t_end=chunk) called once per
// chunk (solver.py:719-779) with the level of that chunk (the inversion
// driver's run_model_chunk, bc_inversion_train.py:332-349).  Wavefronts find
// their scenario by binary search over the scenarios' wave offsets, so one
// launch covers uneven grids; the sequence of launches is graph-capturable.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "swe_common.h"
#include "swe_fetch.h"
#include "swe_host.h"
#include "swe_stage.h"
#include "../../include/swe2d_hip.h"

namespace swe {

struct EnsEntry {
  StageArgs<float> a;
  float *Ua, *Ub, *k1;
  const float* levels;
  int n_chunks;
  double chunk, tol;
  const int32_t* gauges;
  int n_gauges;
  float* gauge_out;
  long long wave_begin;
};

struct EnsClock {
  double elapsed;  // time advanced inside the current chunk
  int chunk;
  int done;
  int steps;
  int status;
  int final_parity;
  int pad;
  unsigned int acc[2][2 * kSlots];  // [step parity][x slots, y slots]
};

struct EnsHeader {
  int n;
  int n_done;
  long long total_waves;
};

__device__ inline int find_scenario(const EnsEntry* tab, int n, long long wid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tab[mid].wave_begin <= wid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int STAGE, class A>
__global__ __launch_bounds__(256) void ens_stage_kernel(const EnsEntry* tab, EnsClock* clk, const EnsHeader* hdr,
                                                        int parity) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long wid = (long long)blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  if (wid >= hdr->total_waves) return;
  int s = find_scenario(tab, hdr->n, wid);
  s = __builtin_amdgcn_readfirstlane(s);
  EnsClock& c = clk[s];
  if (c.done) return;
  const EnsEntry& e = tab[s];
  StageArgs<float> a = e.a;
  a.U = parity ? e.Ub : e.Ua;
  a.f.level = e.levels + c.chunk;
  a.f.level_stride = 0;
  const long long w = wid - e.wave_begin;
  const int strip = (int)(w % a.strips), seg = (int)(w / a.strips);
  float dt = 0.f, half = 0.f;
  if (STAGE == 1) {
    a.out = e.k1;
  } else {
    double remaining = e.chunk - c.elapsed;
    dt = step_dt<float>(c.acc[parity], a.k, (float)remaining);
    if (!dt_ok(dt)) return;
    half = 0.5f * dt;
    a.k1 = e.k1;
    a.out = parity ? e.Ua : e.Ub;
  }
  float ax, ay;
  stage_strip<float, STAGE, A, false, false>(a, strip, seg, 0, dt, half, ax, ay);  // no sources, no k2 record
  if (STAGE == 1) {
    ax = wave_max(ax);
    ay = wave_max(ay);
    if (lane == 0) {
      int slot = (int)(w & (kSlots - 1));
      atomicMax(&c.acc[parity][slot], to_bits<float>(ax));
      atomicMax(&c.acc[parity][kSlots + slot], to_bits<float>(ay));
    }
  }
}

__global__ void ens_clock_kernel(const EnsEntry* tab, EnsClock* clk, EnsHeader* hdr, int parity) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= hdr->n) return;
  EnsClock& c = clk[s];
  if (c.done) return;
  const EnsEntry& e = tab[s];
  double remaining = e.chunk - c.elapsed;
  float dt = step_dt<float>(c.acc[parity], e.a.k, (float)remaining);
  for (int i = 0; i < 2 * kSlots; ++i) c.acc[1 - parity][i] = 0u;
  if (!dt_ok(dt)) {
    c.status = isfinite(dt) ? 2 : 1;
    c.done = 1;
    c.final_parity = parity;
    atomicAdd(&hdr->n_done, 1);
    return;
  }
  c.steps += 1;
  double el = c.elapsed + (double)dt;
  if (el > e.chunk && el - e.chunk <= e.tol) el = e.chunk;
  if (e.chunk - el <= e.tol) {
    const float* Unew = parity ? e.Ua : e.Ub;
    const int nx = e.a.f.nx;
    for (int g = 0; g < e.n_gauges; ++g) {
      int64_t cell = (int64_t)e.gauges[2 * g] * nx + e.gauges[2 * g + 1];
      e.gauge_out[(int64_t)c.chunk * e.n_gauges + g] = Unew[cell] + e.a.f.bed[cell];
    }
    c.chunk += 1;
    c.elapsed = 0.0;
    if (c.chunk >= e.n_chunks) {
      c.done = 1;
      c.final_parity = 1 - parity;
      atomicAdd(&hdr->n_done, 1);
    }
  } else {
    c.elapsed = el;
  }
}

__global__ void ens_reset_kernel(EnsClock* clk, EnsHeader* hdr, int n, long long total_waves) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) {
    hdr->n = n;
    hdr->n_done = 0;
    hdr->total_waves = total_waves;
  }
  if (s >= n) return;
  EnsClock& c = clk[s];
  c.elapsed = 0.0;
  c.chunk = 0;
  c.done = 0;
  c.steps = 0;
  c.status = 0;
  c.final_parity = 0;
  for (int i = 0; i < 2 * kSlots; ++i) {
    c.acc[0][i] = 0u;
    c.acc[1][i] = 0u;
  }
}

}  // namespace swe

using namespace swe;

namespace {

thread_local char e_msg[256] = {0};
int efail(int code, const char* m) {
  strncpy(e_msg, m, sizeof(e_msg) - 1);
  return code;
}
inline size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

struct EnsLayout {
  size_t hdr, tab, clk, total;
};
EnsLayout ens_layout(int n) {
  EnsLayout L;
  L.hdr = 0;
  L.tab = 256;
  L.clk = L.tab + a256(sizeof(EnsEntry) * (size_t)n);
  L.total = L.clk + a256(sizeof(EnsClock) * (size_t)n);
  return L;
}

int pick_rows_ens(int ny, int nx, int hint) {
  if (hint > 0) return hint;
  (void)nx;
  int S = ny / 8;
  if (S < 8) S = 8;
  if (S > 32) S = 32;
  return S;
}

}  // namespace

extern "C" {

size_t swe_ensemble_workspace_bytes(int n) { return n > 0 ? ens_layout(n).total : 0; }

const char* swe_ensemble_last_error_message(void) { return e_msg; }

int swe_ensemble_init(const swe_config* cfg, int n, const swe_scenario* sc, void* ws, size_t bytes,
                      long long* total_waves, void* stream) {
  if (!cfg || !sc || !ws || n <= 0 || !total_waves) return efail(SWE_ERR_ARG, "null argument or empty ensemble");
  EnsLayout L = ens_layout(n);
  if (bytes < L.total) return efail(SWE_ERR_ARG, "ensemble workspace too small");
  for (int q = 0; q < 4; ++q)
    if (cfg->boundary[q] < 0 || cfg->boundary[q] > 5) return efail(SWE_ERR_BC, "bad boundary code");
  if (cfg->fixed_dt) return efail(SWE_ERR_ARG, "fixed_dt applies to host-driven steps; the ensemble picks its dt");
  EnsEntry* host = (EnsEntry*)malloc(sizeof(EnsEntry) * (size_t)n);
  if (!host) return efail(SWE_ERR_ARG, "host allocation failed");
  memset(host, 0, sizeof(EnsEntry) * (size_t)n);
  long long waves = 0;
  for (int i = 0; i < n; ++i) {
    const swe_scenario& s = sc[i];
    if (s.ny < 2 || s.nx < 2 || !(s.dx > 0) || !(s.dy > 0) || !s.U_a || !s.U_b || !s.k1 || !s.bed ||
        (cfg->apply_friction && !s.roughness) || !s.levels || s.n_chunks <= 0 || !(s.chunk_seconds > 0) ||
        (s.n_gauges > 0 && (!s.gauges || !s.gauge_out))) {
      free(host);
      return efail(SWE_ERR_ARG, "invalid scenario description");
    }
    EnsEntry& e = host[i];
    StageArgs<float>& a = e.a;
    a.k = host_consts<float>(cfg, s.dx, s.dy);
    a.f.B = 1;
    a.f.ny = s.ny;
    a.f.nx = s.nx;
    a.f.plane = (int64_t)s.ny * s.nx;
    a.f.bed_bstride = 0;
    a.f.rough_bstride = 0;
    for (int q = 0; q < 4; ++q) a.f.bc[q] = cfg->boundary[q];
    a.f.bed = s.bed;
    a.f.rough = s.roughness;
    a.S = pick_rows_ens(s.ny, s.nx, s.rows_hint);
    a.strips = (s.nx + kOut - 1) / kOut;
    a.segs = (s.ny + a.S - 1) / a.S;
    a.wsr = cfg->water_slope_reset;
    a.positivity = cfg->positivity_limiter;
    a.friction = cfg->apply_friction;
    a.friction_model = cfg->friction_model;
    e.Ua = s.U_a;
    e.Ub = s.U_b;
    e.k1 = s.k1;
    e.levels = s.levels;
    e.n_chunks = s.n_chunks;
    e.chunk = s.chunk_seconds;
    double tol = 1e-12 * fabs(s.chunk_seconds);
    e.tol = tol > 1e-12 ? tol : 1e-12;
    e.gauges = s.gauges;
    e.n_gauges = s.n_gauges;
    e.gauge_out = s.gauge_out;
    e.wave_begin = waves;
    waves += (long long)a.strips * a.segs;
  }
  hipStream_t st = (hipStream_t)stream;
  hipError_t err = hipMemcpyAsync((char*)ws + L.tab, host, sizeof(EnsEntry) * (size_t)n, hipMemcpyHostToDevice, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  free(host);
  if (err != hipSuccess) return efail(SWE_ERR_HIP, hipGetErrorString(err));
  hipLaunchKernelGGL(ens_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, st, (EnsClock*)((char*)ws + L.clk),
                     (EnsHeader*)ws, n, waves);
  err = hipGetLastError();
  if (err != hipSuccess) return efail(SWE_ERR_HIP, hipGetErrorString(err));
  *total_waves = waves;
  return 0;
}

int swe_ensemble_advance(void* ws, int n, long long total_waves, int first_step, int n_steps, int arithmetic,
                         void* stream) {
  if (!ws || n <= 0 || total_waves <= 0 || first_step < 0 || n_steps < 0) return efail(SWE_ERR_ARG, "bad argument");
  if (arithmetic != SWE_ARITH_REFERENCE && arithmetic != SWE_ARITH_FAST) return efail(SWE_ERR_ARG, "bad arithmetic");
  const bool fast = arithmetic == SWE_ARITH_FAST;
  EnsLayout L = ens_layout(n);
  const EnsEntry* tab = (const EnsEntry*)((char*)ws + L.tab);
  EnsClock* clk = (EnsClock*)((char*)ws + L.clk);
  EnsHeader* hdr = (EnsHeader*)ws;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid((unsigned)((total_waves + kWavesPerBlock - 1) / kWavesPerBlock)), block(kWave * kWavesPerBlock);
  for (int g = first_step; g < first_step + n_steps; ++g) {
    int p = g & 1;
    if (fast) {
      hipLaunchKernelGGL((ens_stage_kernel<1, FastRound>), grid, block, 0, st, tab, clk, (const EnsHeader*)hdr, p);
      hipLaunchKernelGGL((ens_stage_kernel<2, FastRound>), grid, block, 0, st, tab, clk, (const EnsHeader*)hdr, p);
    } else {
      hipLaunchKernelGGL((ens_stage_kernel<1, RefRound>), grid, block, 0, st, tab, clk, (const EnsHeader*)hdr, p);
      hipLaunchKernelGGL((ens_stage_kernel<2, RefRound>), grid, block, 0, st, tab, clk, (const EnsHeader*)hdr, p);
    }
    hipLaunchKernelGGL(ens_clock_kernel, dim3((n + 63) / 64), dim3(64), 0, st, tab, clk, hdr, p);
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return efail(SWE_ERR_HIP, hipGetErrorString(err));
  return 0;
}

size_t swe_ensemble_clock_offset(int n) { return n > 0 ? ens_layout(n).clk : 0; }
size_t swe_ensemble_clock_bytes(void) { return sizeof(EnsClock); }

}  // extern "C"
