// Learned motion-proposal network: forward, backward and Adam on gfx950.
//
// Network (per shot row r of a scan; scans are ragged runs of rows):
//   h1 = GELU(X W1^T + b1)      X [M x 13], W1 [256 x 13]
//   h2 = GELU(h1 W2^T + b2)     W2 [256 x 256]
//   u  = h2 W3^T + b3           W3 [6 x 256]
//   y[r, o] = bc[o] + sum_{i<6, t<5} Wc[o, i, t] u[r + t - 2, i]
//            (1D convolution along shot order, zero padding inside each scan)
// Matrix products run on v_mfma_f32_16x16x16_bf16: FP32 operands rounded to
// bf16 (round to nearest even), FP32 accumulation.  Reductions over rows
// (bias and weight gradients) use fixed-order block partials, so results are
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/disorder.h"
#include "kernels.hpp"
#include "transform.hpp"

namespace dis {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 13, kHid = 256, kOut = 6, kK = 5;

// parameter offsets in the flat FP32 buffer
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kHid * kIn;
constexpr int oW2 = oB1 + kHid;
constexpr int oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid;
constexpr int oB3 = oW3 + kOut * kHid;
constexpr int oWc = oB3 + kOut;
constexpr int oBc = oWc + kOut * kOut * kK;
constexpr int kParams = oBc + kOut;

__device__ __forceinline__ short bf16_rne(float f) {
  const unsigned int u = __float_as_uint(f);
  return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// C[i][j] = sum_k A(i,k) B(k,j) (+ bias[j]) with A(i,k) = A[i*sa0 + k*sa1],
// B(k,j) = B[k*sb0 + j*sb1], C row-major with ldc.  EPI: 0 plain store,
// 1 bias + store Z and GELU(Z) (H), 2 bias only.  Each wave computes one
// 16 x 16 tile; a workgroup of 4 waves covers 32 x 32.
template <int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(const float* __restrict__ A, long long sa0, long long sa1,
                                                  const float* __restrict__ B, long long sb0, long long sb1,
                                                  int M, int N, int K, const float* __restrict__ bias,
                                                  float* __restrict__ C, float* __restrict__ H, int ldc) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int i0 = blockIdx.y * 32 + (wv >> 1) * 16;
  const int j0 = blockIdx.x * 32 + (wv & 1) * 16;
  f32x4 d = {0.f, 0.f, 0.f, 0.f};
  const int ia = i0 + row;  // A row of this lane
  const int jb = j0 + row;  // B column of this lane
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * kq + j;
      const float av = (ia < M && k < K) ? A[ia * sa0 + k * sa1] : 0.f;
      const float bv = (jb < N && k < K) ? B[k * sb0 + jb * sb1] : 0.f;
      a[j] = bf16_rne(av);
      b[j] = bf16_rne(bv);
    }
    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
  }
  // lane holds D[4 kq + j][row]
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = i0 + 4 * kq + j, c = j0 + row;
    if (i < M && c < N) {
      float z = d[j];
      if (EPI >= 1) z += bias[c];
      C[(long long)i * ldc + c] = z;
      if (EPI == 1) H[(long long)i * ldc + c] = gelu(z);
    }
  }
}

// rows -> scan index table is given as scan_of[r]; offsets scan_off[b].
__global__ void conv_fwd_kernel(const float* __restrict__ U, const float* __restrict__ P, const int* __restrict__ scan_of,
                                const int* __restrict__ scan_off, int M, float* __restrict__ Y) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * kOut) return;
  const int r = e / kOut, o = e - r * kOut;
  const int b = scan_of[r];
  const int lo = scan_off[b], hi = scan_off[b + 1];
  float acc = P[oBc + o];
  for (int t = 0; t < kK; ++t) {
    const int rr = r + t - kK / 2;
    if (rr < lo || rr >= hi) continue;
    for (int i = 0; i < kOut; ++i) acc += P[oWc + (o * kOut + i) * kK + t] * U[rr * kOut + i];
  }
  Y[e] = acc;
}

// dU[r, i] = sum_{o, t} Wc[o, i, t] dY[r - t + 2, o] (inside the scan)
__global__ void conv_bwd_data_kernel(const float* __restrict__ dY, const float* __restrict__ P,
                                     const int* __restrict__ scan_of, const int* __restrict__ scan_off, int M,
                                     float* __restrict__ dU) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * kOut) return;
  const int r = e / kOut, i = e - r * kOut;
  const int b = scan_of[r];
  const int lo = scan_off[b], hi = scan_off[b + 1];
  float acc = 0.f;
  for (int t = 0; t < kK; ++t) {
    const int rr = r - t + kK / 2;
    if (rr < lo || rr >= hi) continue;
    for (int o = 0; o < kOut; ++o) acc += P[oWc + (o * kOut + i) * kK + t] * dY[rr * kOut + o];
  }
  dU[e] = acc;
}

constexpr int kRedBlocksNN = 256;

// Partial sums over rows, fixed order: grid (kRedBlocksNN), each block sums a
// contiguous row range.  Quantities: conv weights dWc[o,i,t] (180) and dbc (6).
__global__ __launch_bounds__(256) void conv_bwd_param_kernel(const float* __restrict__ dY, const float* __restrict__ U,
                                                            const int* __restrict__ scan_of,
                                                            const int* __restrict__ scan_off, int M,
                                                            float* __restrict__ partial) {
  constexpr int NQ = kOut * kOut * kK + kOut;
  const int rows_per = (M + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * rows_per, r1 = min(M, r0 + rows_per);
  for (int q = threadIdx.x; q < NQ; q += blockDim.x) {
    float acc = 0.f;
    if (q < kOut * kOut * kK) {
      const int o = q / (kOut * kK), rem = q - o * kOut * kK, i = rem / kK, t = rem - i * kK;
      for (int r = r0; r < r1; ++r) {
        const int b = scan_of[r];
        const int rr = r + t - kK / 2;
        if (rr < scan_off[b] || rr >= scan_off[b + 1]) continue;
        acc += dY[r * kOut + o] * U[rr * kOut + i];
      }
    } else {
      const int o = q - kOut * kOut * kK;
      for (int r = r0; r < r1; ++r) acc += dY[r * kOut + o];
    }
    partial[(long long)blockIdx.x * NQ + q] = acc;
  }
}

// column sums of Z [M x N] (bias gradients), block partials over rows
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ Z, int M, int N,
                                                    float* __restrict__ partial) {
  const int rows_per = (M + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * rows_per, r1 = min(M, r0 + rows_per);
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    float acc = 0.f;
    for (int r = r0; r < r1; ++r) acc += Z[(long long)r * N + c];
    partial[(long long)blockIdx.x * N + c] = acc;
  }
}

__global__ void sum_partials_kernel(const float* __restrict__ partial, int nb, int nq, float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  float acc = 0.f;
  for (int b = 0; b < nb; ++b) acc += partial[(long long)b * nq + q];
  out[q] = acc;
}

// dZ = dH * GELU'(Z)
__global__ void gelu_bwd_kernel(const float* __restrict__ dH, const float* __restrict__ Z, long long n,
                                float* __restrict__ dZ) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    dZ[e] = dH[e] * gelu_grad(Z[e]);
}

// dY = 2 (Y - T) / (M * 6); per-block partial squared errors
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ Y, const float* __restrict__ Tg, int n,
                                                 float scale, float* __restrict__ dY, float* __restrict__ partial) {
  __shared__ float sh[256];
  float acc = 0.f;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const float r = Y[e] - Tg[e];
    dY[e] = 2.0f * r * scale;
    acc += r * r;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <int EPI>
void gemm(const float* A, long long sa0, long long sa1, const float* B, long long sb0, long long sb1, int M, int N,
          int K, const float* bias, float* C, float* H, int ldc, hipStream_t s) {
  const dim3 grid((N + 31) / 32, (M + 31) / 32);
  hipLaunchKernelGGL(gemm_kernel<EPI>, grid, dim3(256), 0, s, A, sa0, sa1, B, sb0, sb1, M, N, K, bias, C, H, ldc);
}

struct Act {  // activations of one forward pass, row-major [M x width]
  float *Z1, *H1, *Z2, *H2, *U, *Y;
};

size_t act_floats(int M) { return (size_t)M * (4 * kHid + 2 * kOut); }

Act carve(float* base, int M) {
  Act a;
  a.Z1 = base;
  a.H1 = a.Z1 + (size_t)M * kHid;
  a.Z2 = a.H1 + (size_t)M * kHid;
  a.H2 = a.Z2 + (size_t)M * kHid;
  a.U = a.H2 + (size_t)M * kHid;
  a.Y = a.U + (size_t)M * kOut;
  return a;
}

void forward(const float* P, const float* X, int M, const int* scan_of, const int* scan_off, const Act& a,
             hipStream_t s) {
  gemm<1>(X, kIn, 1, P + oW1, 1, kIn, M, kHid, kIn, P + oB1, a.Z1, a.H1, kHid, s);
  gemm<1>(a.H1, kHid, 1, P + oW2, 1, kHid, M, kHid, kHid, P + oB2, a.Z2, a.H2, kHid, s);
  gemm<2>(a.H2, kHid, 1, P + oW3, 1, kHid, M, kOut, kHid, P + oB3, a.U, nullptr, kOut, s);
  hipLaunchKernelGGL(conv_fwd_kernel, dim3((M * kOut + 255) / 256), dim3(256), 0, s, a.U, P, scan_of, scan_off, M,
                     a.Y);
}

// Gradients of all parameters (G, flat like P) and of the input (dX, or
// null) from dY.  Work buffers: dU [M x 6], dH [M x 256], dZ [M x 256],
// partial [kRedBlocksNN x 256].
void backward(const float* P, const float* X, int M, const int* scan_of, const int* scan_off, const Act& a,
              const float* dY, float* G, float* dX, float* dU, float* dH, float* dZ, float* partial,
              hipStream_t s) {
  const int nb = std::max(1, std::min(kRedBlocksNN, M));
  constexpr int NQc = kOut * kOut * kK + kOut;
  // conv
  hipLaunchKernelGGL(conv_bwd_param_kernel, dim3(nb), dim3(256), 0, s, dY, a.U, scan_of, scan_off, M, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, nb, NQc, G + oWc);
  hipLaunchKernelGGL(conv_bwd_data_kernel, dim3((M * kOut + 255) / 256), dim3(256), 0, s, dY, P, scan_of, scan_off,
                     M, dU);
  // layer 3: u = h2 W3^T + b3 ; dW3 = dU^T h2 [6 x 256], db3 = colsum dU, dH2 = dU W3 [M x 256]
  gemm<0>(dU, 1, kOut, a.H2, kHid, 1, kOut, kHid, M, nullptr, G + oW3, nullptr, kHid, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(nb), dim3(256), 0, s, dU, M, kOut, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, nb, kOut, G + oB3);
  gemm<0>(dU, kOut, 1, P + oW3, kHid, 1, M, kHid, kOut, nullptr, dH, nullptr, kHid, s);
  // layer 2
  const long long nh = (long long)M * kHid;
  const unsigned ge = (unsigned)std::min<long long>((nh + 255) / 256, 8192);
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(ge), dim3(256), 0, s, dH, a.Z2, nh, dZ);
  gemm<0>(dZ, 1, kHid, a.H1, kHid, 1, kHid, kHid, M, nullptr, G + oW2, nullptr, kHid, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(nb), dim3(256), 0, s, dZ, M, kHid, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, nb, kHid, G + oB2);
  gemm<0>(dZ, kHid, 1, P + oW2, kHid, 1, M, kHid, kHid, nullptr, dH, nullptr, kHid, s);
  // layer 1
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(ge), dim3(256), 0, s, dH, a.Z1, nh, dZ);
  gemm<0>(dZ, 1, kHid, X, kIn, 1, kHid, kIn, M, nullptr, G + oW1, nullptr, kIn, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(nb), dim3(256), 0, s, dZ, M, kHid, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, nb, kHid, G + oB1);
  if (dX) gemm<0>(dZ, kHid, 1, P + oW1, kIn, 1, M, kIn, kHid, nullptr, dX, nullptr, kIn, s);
}

}  // namespace

}  // namespace dis

using namespace dis;

namespace {

// scan_of per row and the scan offsets, uploaded once per call
int upload_scans(dis_ctx* c, const int32_t* scan_off, int nscan, int M, const int** d_of, const int** d_off) {
  if (nscan <= 0 || scan_off[0] != 0 || scan_off[nscan] != M) return ctx_fail(c, DIS_E_INVALID, "bad scan offsets");
  std::vector<int> of(std::max(1, M));
  for (int b = 0; b < nscan; ++b) {
    if (scan_off[b + 1] < scan_off[b]) return ctx_fail(c, DIS_E_INVALID, "scan offsets not monotone");
    for (int r = scan_off[b]; r < scan_off[b + 1]; ++r) of[r] = b;
  }
  Staging st;
  const size_t a = st.add(of.data(), sizeof(int) * of.size());
  const size_t b = st.add(scan_off, sizeof(int32_t) * (nscan + 1));
  const unsigned char* d = nullptr;
  int rc = ctx_upload_bytes(c, st, &d);
  if (rc) return rc;
  *d_of = (const int*)(d + a);
  *d_off = (const int*)(d + b);
  return DIS_OK;
}

}  // namespace

extern "C" {

int dis_mlp_param_count(void) { return kParams; }

int dis_mlp_forward(dis_ctx* c, const float* params, const float* X, int32_t M, const int32_t* scan_off, int32_t nscan,
                    float* Y) {
  if (!c) return DIS_E_INVALID;
  if (!params || !X || !Y || M <= 0 || !scan_off) return ctx_fail(c, DIS_E_INVALID, "dis_mlp_forward: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  const int *d_of = nullptr, *d_off = nullptr;
  int st = upload_scans(c, scan_off, nscan, M, &d_of, &d_off);
  if (st) return st;
  void* ws = nullptr;
  if ((st = ctx_workspace(c, sizeof(float) * act_floats(M), &ws))) return st;
  Act a = carve((float*)ws, M);
  a.Y = Y;
  hipStream_t s = ctx_stream(c);
  forward(params, X, M, d_of, d_off, a, s);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_mlp_forward failed");
  return DIS_OK;
}

int dis_mlp_backward(dis_ctx* c, const float* params, const float* X, int32_t M, const int32_t* scan_off,
                     int32_t nscan, const float* dY, float* grads, float* dX) {
  if (!c) return DIS_E_INVALID;
  if (!params || !X || !dY || !grads || M <= 0 || !scan_off)
    return ctx_fail(c, DIS_E_INVALID, "dis_mlp_backward: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  const int *d_of = nullptr, *d_off = nullptr;
  int st = upload_scans(c, scan_off, nscan, M, &d_of, &d_off);
  if (st) return st;
  const size_t na = act_floats(M), nw = (size_t)M * (2 * kHid + kOut) + (size_t)kRedBlocksNN * kHid;
  void* ws = nullptr;
  if ((st = ctx_workspace(c, sizeof(float) * (na + nw), &ws))) return st;
  Act a = carve((float*)ws, M);
  float* dU = (float*)ws + na;
  float* dH = dU + (size_t)M * kOut;
  float* dZ = dH + (size_t)M * kHid;
  float* partial = dZ + (size_t)M * kHid;
  hipStream_t s = ctx_stream(c);
  forward(params, X, M, d_of, d_off, a, s);
  backward(params, X, M, d_of, d_off, a, dY, grads, dX, dU, dH, dZ, partial, s);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_mlp_backward failed");
  return DIS_OK;
}

int dis_mlp_train_step(dis_ctx* c, float* params, float* adam_m, float* adam_v, int32_t step, float lr,
                       const float* X, const float* target, int32_t M, const int32_t* scan_off, int32_t nscan,
                       float* loss) {
  if (!c) return DIS_E_INVALID;
  if (!params || !adam_m || !adam_v || !X || !target || M <= 0 || !scan_off || step < 1)
    return ctx_fail(c, DIS_E_INVALID, "dis_mlp_train_step: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  const int *d_of = nullptr, *d_off = nullptr;
  int st = upload_scans(c, scan_off, nscan, M, &d_of, &d_off);
  if (st) return st;
  const size_t na = act_floats(M);
  const size_t nw = (size_t)M * (2 * kHid + 2 * kOut) + (size_t)kRedBlocksNN * kHid + kParams + 256;
  void* ws = nullptr;
  if ((st = ctx_workspace(c, sizeof(float) * (na + nw), &ws))) return st;
  Act a = carve((float*)ws, M);
  float* dY = (float*)ws + na;
  float* dU = dY + (size_t)M * kOut;
  float* dH = dU + (size_t)M * kOut;
  float* dZ = dH + (size_t)M * kHid;
  float* partial = dZ + (size_t)M * kHid;
  float* G = partial + (size_t)kRedBlocksNN * kHid;
  float* lpart = G + kParams;
  hipStream_t s = ctx_stream(c);
  forward(params, X, M, d_of, d_off, a, s);
  const int n = M * kOut;
  const int lb = 64;
  hipLaunchKernelGGL(mse_kernel, dim3(lb), dim3(256), 0, s, a.Y, target, n, 1.0f / (float)n, dY, lpart);
  backward(params, X, M, d_of, d_off, a, dY, G, nullptr, dU, dH, dZ, partial, s);
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, params, G, adam_m, adam_v, kParams, lr,
                     b1, b2, 1e-8f, c1, c2);
  if (loss) {
    float lp[lb];
    if (hipMemcpyAsync(lp, lpart, sizeof(lp), hipMemcpyDeviceToHost, s) != hipSuccess)
      return ctx_fail(c, DIS_E_HIP, "loss copy failed");
    if (hipStreamSynchronize(s) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "dis_mlp_train_step failed");
    double t = 0;
    for (int i = 0; i < lb; ++i) t += lp[i];
    *loss = (float)(t / n);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_mlp_train_step failed");
  return DIS_OK;
}

}  // extern "C"
