// Device kernels of the LM motion step (Reconstruction/LMsolver.m:52-141):
// per-state residual energy, gradient and Gauss-Newton Hessian reductions
// and the damped 6x6 normal-equation solves, in FP64; and the fractional
// finite-difference readout weighting (E.Fd, :88-91), in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>

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
    const float2 Y = a.Yr[yi], d = a.y ? a.y[yi] : make_float2(0.f, 0.f);
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

// grid (tiles of P samples, nlist): the tile's P consecutive samples of one
// listed state, every coil in turn, staged in LDS with the samples fastest
// (coalesced loads along the compact rows) and transformed along the readout
// (Reconstruction/LMsolver.m:88-91 with fftGPU(., 3), Methods/fftGPU.m:28-45)
__global__ __launch_bounds__(kThreads) void readout_filter_kernel(ReadoutFilterArgs a, int P) {
  constexpr int kE = kMaxTile / kThreads;
  __shared__ float2 lds[kMaxTile];
  __shared__ float2 tw[1024];
  const int tid = threadIdx.x;
  const int s = a.slist[blockIdx.y];
  const int o0 = a.off[s], o1 = a.off[s + 1];
  const int j0 = o0 + blockIdx.x * P;
  if (j0 >= o1) return;  // uniform: the whole tile lies past the state's samples
  const int n = a.nF, tot = P * n;
  const int Pv = min(P, o1 - j0);
  for (int k = tid; k < n; k += kThreads) tw[k] = a.tw[k];
  float en[kE];
#pragma unroll
  for (int i = 0; i < kE; ++i) en[i] = 0.f;
  for (int c = 0; c < a.C; ++c) {
    const long long base = (long long)j0 + (long long)a.Ns * (a.lo + (long long)a.n3 * c);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        const int k = e / P, p = e - k * P;
        float2 v = make_float2(0.f, 0.f);
        if (p < Pv) {
          const long long idx = base + p + (long long)a.Ns * k;
          v = a.Yr[idx];
          if (a.y) {
            const float2 d = a.y[idx];
            v = make_float2(v.x - d.x, v.y - d.y);
          }
        }
        lds[e] = v;
      }
    }
    __syncthreads();
    fft_tile<false>(lds, tw, a.plan, P, tid);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        const int k = e / P, p = e - k * P;
        if (p < Pv) {
          const float f = a.Fd[j0 + p + (long long)a.Ns * k];
          const float2 v = make_float2(lds[e].x * f, lds[e].y * f);
          a.Yr[base + p + (long long)a.Ns * k] = v;
          en[i] += v.x * v.x + v.y * v.y;
        }
      }
    }
    __syncthreads();
  }
  if (a.epart) {
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        const int k = e / P, p = e - k * P;
        if (p < Pv) a.epart[j0 + p + (long long)a.Ns * (a.lo + k)] = en[i];
      }
    }
  }
}

}  // namespace

hipError_t launch_readout_filter(ReadoutFilterArgs a, int nlist, int maxnec, hipStream_t st) {
  if (a.nF <= 0 || a.nF > 1024 || a.lo < 0 || a.lo + a.nF > a.n3 || nlist <= 0) return hipErrorInvalidValue;
  int P = 1;
  while (2 * P * a.nF <= kMaxTile && 2 * P <= 64) P *= 2;
  const int tiles = std::max(1, (maxnec + P - 1) / P);
  hipLaunchKernelGGL(readout_filter_kernel, dim3(tiles, nlist), dim3(kThreads), 0, st, a, P);
  return hipGetLastError();
}

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
