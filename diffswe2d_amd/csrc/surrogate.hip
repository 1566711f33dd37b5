// BF16 MFMA kernels of the 1-D convolution surrogate (gfx950).
//
// Every convolution (forward, input gradient, weight gradient) is a GEMM in
// "NT" form, C[m][n] = sum_k A[m][k] * B[n][k], with both operands bf16 and
// K contiguous, accumulated in FP32 by gfx950's v_mfma_f32_16x16x32_bf16.  A
// 256-thread block computes a 64 x 64 tile of C: per 32-wide K step it stages
// the 64 x 32 operand slices in LDS, and each of its four wavefronts reads its
// 32 x 32 share as a 2 x 2 grid of 16 x 16 MFMA tiles (each A and B slice read
// from HBM once per block instead of once per wavefront).  Long-K products
// (weight gradients over batch x time) are split over blockIdx.z into FP32
// partials that a second kernel sums in a fixed order, so results are
// deterministic.
//
// Data layouts: activations are channel-major [C][B*L] (bf16); im2col writes
// the "column" operand [B*L][Cin*Kw] (or its transpose) with 'same' zero
// padding and dilation.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include "../../include/swe2d_hip.h"

namespace sur {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float x) {
  // round to nearest even (NaN stays NaN)
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ inline bf16x4 load4(const uint16_t* p, bool ok) {
  if (!ok) return bf16x4{0, 0, 0, 0};
  return *reinterpret_cast<const bf16x4*>(p);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 64;              // C tile of a block, in M and in N
constexpr int kKStep = 32;             // K per v_mfma_f32_16x16x32_bf16
constexpr int kLdsRow = kKStep + 8;    // bf16 per LDS row: 80 bytes, so the 16 rows one
                                       // MFMA operand read touches start in distinct banks

// C = A * B^T on gfx950's 16x16x32 BF16 MFMA.  Per K step the block stages the
// 64 x 32 slices of A and B in LDS (one coalesced 16-byte load per thread and
// operand) and each of its four wavefronts reads its 32 x 32 share as 2 x 2
// MFMA tiles: lane l holds row (l % 16) of a tile and K elements 8 (l / 16) ..
// 8 (l / 16) + 7 of the step, the same K mapping for A and B.
__global__ __launch_bounds__(256) void gemm_nt_bf16_kernel(int M, int N, int K, const uint16_t* __restrict__ A,
                                                           long lda, const uint16_t* __restrict__ B, long ldb,
                                                           float* __restrict__ C, long ldc, long split_stride,
                                                           int k_per_split) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kLdsRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kLdsRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = blockIdx.y * kTile, bn = blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int kb = blockIdx.z * k_per_split;
  const int ke = min(K, kb + k_per_split);
  // staging: thread t moves row t / 4, K elements 8 (t % 4) .. 8 (t % 4) + 7 of each slice
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  const int am = bm + sr, bnr = bn + sr;
  const int r = lane & 15, g = lane >> 4;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k = kb; k < ke; k += kKStep) {
    // K and every split boundary are multiples of 16, so an 8-element chunk is
    // either wholly before ke or wholly past it
    const int kk = k + sk;
    const bool ina = kk < ke && am < M, inb = kk < ke && bnr < N;
    const bf16x4 a_lo = load4(A + (long)am * lda + kk, ina), a_hi = load4(A + (long)am * lda + kk + 4, ina);
    const bf16x4 b_lo = load4(B + (long)bnr * ldb + kk, inb), b_hi = load4(B + (long)bnr * ldb + kk + 4, inb);
    __syncthreads();  // the previous step's operand reads are done
    *reinterpret_cast<bf16x4*>(&As[sr * kLdsRow + sk]) = a_lo;
    *reinterpret_cast<bf16x4*>(&As[sr * kLdsRow + sk + 4]) = a_hi;
    *reinterpret_cast<bf16x4*>(&Bs[sr * kLdsRow + sk]) = b_lo;
    *reinterpret_cast<bf16x4*>(&Bs[sr * kLdsRow + sk + 4]) = b_hi;
    __syncthreads();
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&As[(wm + r) * kLdsRow + 8 * g]);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&As[(wm + 16 + r) * kLdsRow + 8 * g]);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(&Bs[(wn + r) * kLdsRow + 8 * g]);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(&Bs[(wn + 16 + r) * kLdsRow + 8 * g]);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
  }
  // D layout of a 16 x 16 MFMA tile: lane l holds rows 4 (l / 16) .. +3 of column l % 16
  float* Cs = C + (long)blockIdx.z * split_stride;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        int m = bm + wm + 16 * i + 4 * g + q, n = bn + wn + 16 * j + r;
        if (m < M && n < N) Cs[(long)m * ldc + n] = acc[i][j][q];
      }
}

// sum of split-K partials in split order: C[m][n] = sum_s P[s][m][n]
__global__ void reduce_splits_kernel(const float* P, long split_stride, int splits, float* C, long count) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int k = 0; k < splits; ++k) s += P[(long)k * split_stride + i];
  C[i] = s;
}

// im2col of channel-major X [Cin][B*L] -> col [B*L][Kpad] (transpose = 0) or
// col [Kpad][B*L] (transpose = 1), with col index ci*Kw + k and tap offset
// (k - (Kw-1)/2) * dil; entries beyond Cin*Kw (padding of K to 16) are 0.
__global__ void im2col_kernel(const uint16_t* X, int Cin, int Bn, int L, int Kw, int dil, int Kpad, int transpose,
                              uint16_t* col) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long BL = (long)Bn * L;
  long total = BL * Kpad;
  if (idx >= total) return;
  // consecutive threads take consecutive output elements: along K for [B*L][Kpad],
  // along B*L for the transpose (whose reads of X are then consecutive too)
  int kc = transpose ? (int)(idx / BL) : (int)(idx % Kpad);
  long n = transpose ? idx % BL : idx / Kpad;
  int b = (int)(n / L), t = (int)(n % L);
  uint16_t v = 0;
  if (kc < Cin * Kw) {
    int ci = kc / Kw, k = kc % Kw;
    int tt = t + (k - (Kw - 1) / 2) * dil;
    if (tt >= 0 && tt < L) v = X[(long)ci * Bn * L + (long)b * L + tt];
  }
  if (transpose) col[(long)kc * Bn * L + n] = v;
  else col[n * Kpad + kc] = v;
}

// col2im (adjoint of im2col): dX[ci][b*L + t] = sum_k dcol[ci*Kw + k][b*L + t - off_k]
__global__ void col2im_kernel(const float* dcol, int Cin, int Bn, int L, int Kw, int dil, float* dX) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)Cin * Bn * L;
  if (idx >= total) return;
  int ci = (int)(idx / ((long)Bn * L));
  long n = idx % ((long)Bn * L);
  int b = (int)(n / L), t = (int)(n % L);
  float s = 0.f;
  for (int k = 0; k < Kw; ++k) {
    int ts = t - (k - (Kw - 1) / 2) * dil;
    if (ts >= 0 && ts < L) s += dcol[((long)ci * Kw + k) * Bn * L + (long)b * L + ts];
  }
  dX[idx] = s;
}

// epilogue: y = act(C + bias[m]) -> fp32 (pre-activation kept for backward) and bf16
__global__ void bias_act_kernel(const float* C, const float* bias, int M, long N, int relu, float* pre,
                                uint16_t* out_bf16, float* out_f32) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  int m = (int)(i / N);
  float v = C[i] + (bias ? bias[m] : 0.f);
  if (pre) pre[i] = v;
  float y = relu ? (v > 0.f ? v : 0.f) : v;
  if (out_bf16) out_bf16[i] = f2bf(y);
  if (out_f32) out_f32[i] = y;
}

// backward through the activation: g_pre = g_out * relu'(pre); also bf16 copy and
// the transpose [N][M] the input-gradient GEMM wants
__global__ void act_bwd_kernel(const float* g_out, const float* pre, int M, long N, int relu, float* g_pre,
                               uint16_t* g_bf16, uint16_t* g_bf16_t, int ldt) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  float g = g_out[i];
  if (relu && !(pre[i] > 0.f)) g = 0.f;
  if (g_pre) g_pre[i] = g;
  uint16_t b = f2bf(g);
  if (g_bf16) g_bf16[i] = b;
  if (g_bf16_t) {
    int m = (int)(i / N);
    long n = i % N;
    g_bf16_t[n * ldt + m] = b;
  }
}

// bias gradient: db[m] = sum_n g[m][n], one block per m, fixed reduction order
__global__ __launch_bounds__(256) void bias_grad_kernel(const float* g, long N, float* db) {
  __shared__ float part[256];
  int m = blockIdx.x;
  float s = 0.f;
  for (long n = threadIdx.x; n < N; n += 256) s += g[(long)m * N + n];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[m] = part[0];
}

// float -> bf16 with an optional transpose of a [R][Cc] matrix into [Cc][Rpad]
// (rows beyond R padded with zeros)
__global__ void cast_bf16_kernel(const float* x, long count, uint16_t* y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < count) y[i] = f2bf(x[i]);
}

__global__ void cast_pad_rows_kernel(const float* x, int R, int Cc, int Cpad, uint16_t* y) {
  // y[r][c] for c < Cpad: x[r][c] when c < Cc else 0
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)R * Cpad) return;
  int r = (int)(i / Cpad), c = (int)(i % Cpad);
  y[i] = c < Cc ? f2bf(x[(long)r * Cc + c]) : (uint16_t)0;
}

__global__ void transpose_bf16_kernel(const float* x, int R, int Cc, int Rpad, uint16_t* y) {
  // y[c][r] = bf16(x[r][c]) for r < R, 0 for R <= r < Rpad
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)Cc * Rpad) return;
  int c = (int)(i / Rpad), r = (int)(i % Rpad);
  y[i] = r < R ? f2bf(x[(long)r * Cc + c]) : (uint16_t)0;
}

}  // namespace sur

using namespace sur;

namespace {
thread_local char s_msg[256] = {0};
int sfail(int code, const char* m) {
  strncpy(s_msg, m, sizeof(s_msg) - 1);
  return code;
}
inline dim3 g1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }
int last(void) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : sfail(SWE_ERR_HIP, hipGetErrorString(e));
}
}  // namespace

extern "C" {

const char* sur_last_error_message(void) { return s_msg; }

int sur_gemm_nt_bf16(int M, int N, int K, const void* A, long lda, const void* B, long ldb, float* C, long ldc,
                     int splits, float* partials, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return sfail(SWE_ERR_ARG, "bad gemm arguments");
  if (K % 16 || lda % 4 || ldb % 4 || lda < K || ldb < K || ldc < N)
    return sfail(SWE_ERR_ARG, "K must be a multiple of 16 and leading dimensions multiples of 4");
  if (((uintptr_t)A & 7) || ((uintptr_t)B & 7)) return sfail(SWE_ERR_ARG, "operands must be 8-byte aligned");
  if (splits < 1) splits = 1;
  int kps = ((K + splits - 1) / splits + 15) / 16 * 16;
  splits = (K + kps - 1) / kps;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((N + 63) / 64, (M + 63) / 64, splits);
  if (splits == 1) {
    hipLaunchKernelGGL(gemm_nt_bf16_kernel, grid, dim3(256), 0, s, M, N, K, (const uint16_t*)A, lda,
                       (const uint16_t*)B, ldb, C, ldc, 0L, kps);
  } else {
    if (!partials || ldc != N) return sfail(SWE_ERR_ARG, "split-K needs partials and ldc == N");
    long stride = (long)M * N;
    hipLaunchKernelGGL(gemm_nt_bf16_kernel, grid, dim3(256), 0, s, M, N, K, (const uint16_t*)A, lda,
                       (const uint16_t*)B, ldb, partials, (long)N, stride, kps);
    hipLaunchKernelGGL(reduce_splits_kernel, g1(stride), dim3(256), 0, s, (const float*)partials, stride, splits, C,
                       stride);
  }
  return last();
}

int sur_im2col(const void* X, int Cin, int Bn, int L, int Kw, int dil, int Kpad, int transpose, void* col,
               void* stream) {
  if (!X || !col || Kpad < Cin * Kw) return sfail(SWE_ERR_ARG, "bad im2col arguments");
  long total = (long)Bn * L * Kpad;
  hipLaunchKernelGGL(im2col_kernel, g1(total), dim3(256), 0, (hipStream_t)stream, (const uint16_t*)X, Cin, Bn, L,
                     Kw, dil, Kpad, transpose, (uint16_t*)col);
  return last();
}

int sur_col2im(const float* dcol, int Cin, int Bn, int L, int Kw, int dil, float* dX, void* stream) {
  if (!dcol || !dX) return sfail(SWE_ERR_ARG, "bad col2im arguments");
  hipLaunchKernelGGL(col2im_kernel, g1((long)Cin * Bn * L), dim3(256), 0, (hipStream_t)stream, dcol, Cin, Bn, L, Kw,
                     dil, dX);
  return last();
}

int sur_bias_act(const float* C, const float* bias, int M, long N, int relu, float* pre, void* out_bf16,
                 float* out_f32, void* stream) {
  hipLaunchKernelGGL(bias_act_kernel, g1((long)M * N), dim3(256), 0, (hipStream_t)stream, C, bias, M, N, relu, pre,
                     (uint16_t*)out_bf16, out_f32);
  return last();
}

int sur_act_bwd(const float* g_out, const float* pre, int M, long N, int relu, float* g_pre, void* g_bf16,
                void* g_bf16_t, int ldt, void* stream) {
  if (g_bf16_t && ldt < M) return sfail(SWE_ERR_ARG, "transposed gradient needs ldt >= M");
  hipLaunchKernelGGL(act_bwd_kernel, g1((long)M * N), dim3(256), 0, (hipStream_t)stream, g_out, pre, M, N, relu,
                     g_pre, (uint16_t*)g_bf16, (uint16_t*)g_bf16_t, ldt);
  return last();
}

int sur_bias_grad(const float* g, int M, long N, float* db, void* stream) {
  hipLaunchKernelGGL(bias_grad_kernel, dim3(M), dim3(256), 0, (hipStream_t)stream, g, N, db);
  return last();
}

int sur_cast_bf16(const float* x, long count, void* y, void* stream) {
  hipLaunchKernelGGL(cast_bf16_kernel, g1(count), dim3(256), 0, (hipStream_t)stream, x, count, (uint16_t*)y);
  return last();
}

int sur_cast_pad_rows(const float* x, int R, int Cc, int Cpad, void* y, void* stream) {
  hipLaunchKernelGGL(cast_pad_rows_kernel, g1((long)R * Cpad), dim3(256), 0, (hipStream_t)stream, x, R, Cc, Cpad,
                     (uint16_t*)y);
  return last();
}

int sur_transpose_bf16(const float* x, int R, int Cc, int Rpad, void* y, void* stream) {
  hipLaunchKernelGGL(transpose_bf16_kernel, g1((long)Cc * Rpad), dim3(256), 0, (hipStream_t)stream, x, R, Cc, Rpad,
                     (uint16_t*)y);
  return last();
}

}  // extern "C"
