// Probes: the VALU issue rate the stage kernels are measured against, the
// kernels' FP32 division, square root and friction power on the device next to
// the compiler's IEEE forms and powf, and their one-lane wavefront shifts.
//
// VALU issue-rate probe:
// Each lane runs 8 independent FP32 multiply/add chains (no memory traffic in
// the loop; 16 non-packed VALU instructions per iteration in the gfx950 ISA),
// so with every SIMD holding several wavefronts the launch time is set by how
// fast the CUs issue plain FP32 VALU instructions.
#include <hip/hip_runtime.h>

#include "../../include/swe2d_hip.h"
#include "swe_fetch.h"
#include "swe_host.h"
#include "swe_math.h"

namespace {

__global__ __launch_bounds__(256) void valu_probe_kernel(float* out, int iters, float x, float y) {
  float a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
  for (int i = 0; i < iters; ++i) {
    a0 = a0 * x + y;
    a1 = a1 * x + y;
    a2 = a2 * x + y;
    a3 = a3 * x + y;
    a4 = a4 * x + y;
    a5 = a5 * x + y;
    a6 = a6 * x + y;
    a7 = a7 * x + y;
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

// Mix probe: the arithmetic of one x-sweep cell of the stage kernels -- the
// primitives of a cell from its state (two divisions), the limited increments,
// both face states, the HLL face flux with its square roots and star divisions,
// the flux differences over dx -- on register values only, each iteration on
// new wet, subcritical states (so the same branches as the basin's ocean cells
// are taken) with the lane shifts of the kernels.  Its issue rate is the VALU
// ceiling of the stage kernels' own instruction mix (transcendental estimates,
// FMAs, compares, selects, DPP moves); SQ_INSTS_VALU per launch gives the count.
__global__ __launch_bounds__(256) void valu_mix_probe_kernel(float* out, int iters, swe::SweConsts<float> k) {
  using namespace swe;
  const int lane = threadIdx.x & 63;
  float acc = 0.0f;
  for (int i = 0; i < iters; ++i) {
    const float h_raw = 1.0f + 0.01f * (float)(lane & 7) + 1.0e-3f * (float)(i & 15);
    const float hu = 0.05f + 1.0e-3f * (float)((lane + i) & 31);
    const float hv = -0.02f + 5.0e-4f * (float)((lane * 3 + i) & 31);
    const float z = -1.0f - 0.01f * (float)((lane * 5 + i) & 15);
    float h, u, v, eta;
    primitive_cell(h_raw, hu, hv, z, k, h, u, v, eta);
    Cell<float> cx{h, u, v, eta, z};
    Cell<float> cm = shup(cx), cp = shdn(cx);
    Incr<float> xi = increments(cm, cx, cp, k, true);
    Side<float> right_of_face = side_minus(cx, xi, k);
    Side<float> left_of_face = side_plus(cx, xi, k);
    Side<float> nxt = shdn(right_of_face);
    FaceFlux<float> F = face_flux(left_of_face, cx.h, cx.z, nxt, cp.h, cp.z, k);
    float pfh = shup(F.fh), pft = shup(F.ft), pfn = shup(F.fn - F.sR);
    acc += div_by(pfh - F.fh, k.ddx) + div_by(pfn - (F.fn - F.sL), k.ddx) + div_by(pft - F.ft, k.ddx) +
           F.speed;
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void division_probe_kernel(const float* a, const float* b, int64_t n, float* q_rn,
                                                             float* q_by, float* q_ieee) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  q_rn[i] = swe::div_rn(x, y);
  q_by[i] = swe::div_by(x, swe::Divisor<float>{y, (float)(1.0 / (double)y)});
  q_ieee[i] = x / y;
}

__global__ __launch_bounds__(256) void sqrt_probe_kernel(const float* x, int64_t n, float* s_rn, float* s_ieee) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s_rn[i] = swe::sqrt_rn(x[i]);
  s_ieee[i] = sqrtf(x[i]);
}

__global__ __launch_bounds__(256) void pow_probe_kernel(const float* x, int64_t n, float third, float* y_fast,
                                                        float* y_powf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y_fast[i] = swe::pow_third(x[i], third);
  y_powf[i] = powf(x[i], third);
}

__global__ __launch_bounds__(256) void lane_shift_probe_kernel(const double* x, double* up64, double* dn64,
                                                               float* up32, float* dn32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double v = x[i];
  up64[i] = swe::shup(v);
  dn64[i] = swe::shdn(v);
  up32[i] = swe::shup((float)v);
  dn32[i] = swe::shdn((float)v);
}

}  // namespace

extern "C" int swe_lane_shift_probe(const double* x, long long n, double* up64, double* dn64, float* up32,
                                    float* dn32, void* stream) {
  if (!x || !up64 || !dn64 || !up32 || !dn32 || n <= 0 || n % 256 != 0) return SWE_ERR_ARG;
  hipLaunchKernelGGL(lane_shift_probe_kernel, dim3((unsigned)(n / 256)), dim3(256), 0, (hipStream_t)stream, x, up64,
                     dn64, up32, dn32);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}

extern "C" int swe_pow_probe(const float* x, long long n, float third, float* y_fast, float* y_powf, void* stream) {
  if (!x || !y_fast || !y_powf || n <= 0) return SWE_ERR_ARG;
  hipLaunchKernelGGL(pow_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x,
                     (int64_t)n, third, y_fast, y_powf);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}

extern "C" int swe_sqrt_probe(const float* x, long long n, float* s_rn, float* s_ieee, void* stream) {
  if (!x || !s_rn || !s_ieee || n <= 0) return SWE_ERR_ARG;
  hipLaunchKernelGGL(sqrt_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x,
                     (int64_t)n, s_rn, s_ieee);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}

extern "C" int swe_division_probe(const float* a, const float* b, long long n, float* q_rn, float* q_by,
                                  float* q_ieee, void* stream) {
  if (!a || !b || !q_rn || !q_by || !q_ieee || n <= 0) return SWE_ERR_ARG;
  hipLaunchKernelGGL(division_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a,
                     b, (int64_t)n, q_rn, q_by, q_ieee);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}

extern "C" int swe_valu_mix_probe(float* out, int blocks, int iters, void* stream) {
  if (!out || blocks <= 0 || iters <= 0) return SWE_ERR_ARG;
  swe_config c{};
  c.gravity = 9.81;
  c.dry_depth = 1e-6;
  c.epsilon = 1e-12;
  c.cfl = 0.5;
  c.limiter_theta = 1.3;
  const swe::SweConsts<float> k = swe::host_consts<float>(&c, 100.0, 100.0);
  hipLaunchKernelGGL(valu_mix_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, k);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}

extern "C" int swe_valu_probe(float* out, int blocks, int iters, void* stream) {
  if (!out || blocks <= 0 || iters <= 0) return SWE_ERR_ARG;
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, 0.999f,
                     0.001f);
  return hipGetLastError() == hipSuccess ? 0 : SWE_ERR_HIP;
}
