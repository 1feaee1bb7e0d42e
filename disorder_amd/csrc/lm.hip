// Device kernels of the LM motion step (Reconstruction/LMsolver.m:52-141):
// per-state residual energy, gradient and Gauss-Newton Hessian reductions,
// and the damped 6x6 normal-equation solves, all in FP64.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lm.hpp"

namespace dis {

namespace {

constexpr int kLmBlocks = 64;  // partial blocks per state

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Hessian pairs ha = [diag(1..6), nchoosek(1:6,2)] (LMsolver.m:26)
__constant__ int kHi[21] = {0, 1, 2, 3, 4, 5, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4};
__constant__ int kHj[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};

// grid (kLmBlocks, nlist): block b of listed state sl sums a slice of the
// state's (sample, readout plane in [i3lo, i3lo + i3n), coil) elements.
__global__ __launch_bounds__(kThreads) void lm_reduce_kernel(LmReduceArgs a) {
  const int sl = blockIdx.y;
  const int s = a.slist[sl];
  const int o0 = a.off[s], o1 = a.off[s + 1];
  const int nec = o1 - o0;
  const long long tot = (long long)nec * a.i3n * a.C;
  double acc[kLmSums];
#pragma unroll
  for (int q = 0; q < kLmSums; ++q) acc[q] = 0.0;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < tot;
       idx += (long long)gridDim.x * kThreads) {
    const long long j = o0 + idx % nec;
    const long long rest = idx / nec;  // (i3 - i3lo) + i3n c
    const long long c = rest / a.i3n;
    const long long i3 = a.i3lo + (rest - c * a.i3n);
    const long long yi = j + (long long)a.Ns * (i3 + (long long)a.n3 * c);
    const float2 Y = a.Yr[yi], d = a.y[yi];
    const float2 r = make_float2(Y.x - d.x, Y.y - d.y);
    acc[0] += (double)r.x * r.x + (double)r.y * r.y;
    float2 J[6];
#pragma unroll
    for (int g = 0; g < 6; ++g) J[g] = a.J[g][yi];
#pragma unroll
    for (int g = 0; g < 6; ++g) acc[1 + g] += (double)J[g].x * r.x + (double)J[g].y * r.y;  // Re(J conj r)
#pragma unroll
    for (int h = 0; h < 21; ++h) {
      const float2 u = J[kHi[h]], v = J[kHj[h]];
      acc[7 + h] += (double)u.x * v.x + (double)u.y * v.y;  // Re(J_i conj J_j)
    }
  }
  __shared__ double sh[kThreads / 64][kLmSums];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kLmSums; ++q) {
    const double v = wsum(acc[q]);
    if (lane == 0) sh[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kLmSums) {
    double t = 0;
    for (int i = 0; i < kThreads / 64; ++i) t += sh[i][threadIdx.x];
    a.partial[((long long)sl * gridDim.x + blockIdx.x) * kLmSums + threadIdx.x] = t;
  }
}

__global__ void lm_finalize_kernel(const double* partial, int nb, int nlist, double* out) {
  const int sl = blockIdx.x;
  const int q = threadIdx.x;
  if (q < kLmSums && sl < nlist) {
    double t = 0;
    for (int b = 0; b < nb; ++b) t += partial[((long long)sl * nb + b) * kLmSums + q];
    out[(long long)sl * kLmSums + q] = t;
  }
}

// Damped 6x6 solve per listed state (LMsolver.m:132-141):
//   MH_ii = (1 + w) H_ii, MH_ij = H_ij,  step = -(winit / w) MH \ g
// Gaussian elimination with partial pivoting in FP64, one thread per state.
__global__ void lm_solve_kernel(const double* sums, const double* w, int nlist, double winit, double* step) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nlist) return;
  const double* S = sums + (long long)sl * kLmSums;
  double M[6][7];
  for (int h = 0; h < 21; ++h) {
    const int i = kHi[h], j = kHj[h];
    const double v = S[7 + h];
    if (i == j) {
      M[i][i] = (1.0 + w[sl]) * v;
    } else {
      M[i][j] = v;
      M[j][i] = v;
    }
  }
  for (int i = 0; i < 6; ++i) M[i][6] = S[1 + i];
  for (int k = 0; k < 6; ++k) {
    int piv = k;
    double best = fabs(M[k][k]);
    for (int i = k + 1; i < 6; ++i)
      if (fabs(M[i][k]) > best) {
        best = fabs(M[i][k]);
        piv = i;
      }
    if (piv != k)
      for (int j = 0; j < 7; ++j) {
        const double tmp = M[k][j];
        M[k][j] = M[piv][j];
        M[piv][j] = tmp;
      }
    const double d = M[k][k];
    for (int i = k + 1; i < 6; ++i) {
      const double f = M[i][k] / d;
      for (int j = k; j < 7; ++j) M[i][j] -= f * M[k][j];
    }
  }
  double x[6];
  for (int i = 5; i >= 0; --i) {
    double v = M[i][6];
    for (int j = i + 1; j < 6; ++j) v -= M[i][j] * x[j];
    x[i] = v / M[i][i];
  }
  const double sc = -(winit / w[sl]);
  for (int i = 0; i < 6; ++i) step[(long long)sl * 6 + i] = sc * x[i];
}

}  // namespace

hipError_t launch_lm_reduce(const LmReduceArgs& a, int nlist, double* partial_out, double* sums_out,
                            hipStream_t st) {
  (void)partial_out;
  if (a.i3lo < 0 || a.i3n <= 0 || a.i3lo + a.i3n > a.n3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lm_reduce_kernel, dim3(kLmBlocks, nlist), dim3(kThreads), 0, st, a);
  hipLaunchKernelGGL(lm_finalize_kernel, dim3(nlist), dim3(64), 0, st, a.partial, kLmBlocks, nlist, sums_out);
  return hipGetLastError();
}

hipError_t launch_lm_solve(const double* sums, const double* w, int nlist, double winit, double* step,
                           hipStream_t st) {
  hipLaunchKernelGGL(lm_solve_kernel, dim3((nlist + 63) / 64), dim3(64), 0, st, sums, w, nlist, winit, step);
  return hipGetLastError();
}

int lm_partial_count() { return kLmBlocks; }

}  // namespace dis
