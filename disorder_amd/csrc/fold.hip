// SENSE folding inside the coil stage (Methods/fold.m, ifold.m;
// Reconstruction/encode.m:100-104, decode.m:73-76, encodedecode.m:100-104,
// 167-170): the reconstruction grid n1 x n2 is larger than the acquired
// phase-encode grid ny1 x ny2, and after the coil multiply the image is
// aliased onto the acquired grid before the Fourier rows are taken.
//
// fold.m (A empty, fl = 0) sums, for every acquired position i, the samples
// r with (r - o) mod ny = i, o = ceil((n - ny) / 2); ifold.m is its adjoint,
// the periodic replication out[r] = in[(r - o) mod ny].  Per state the
// stage runs as a few batched kernels over the whole state (coils and
// readout planes together):
//   forward  [dim-1 translation] -> coil multiply + fold (both axes) ->
//            FFT along dim 1 (ny1) -> DFT along dim 2 at the state's samples
//   adjoint  DFT^H along dim 2 from the state's rows (per-state CSR by k1) ->
//            inverse FFT along dim 1 -> unfold + conj(S) coil sum ->
//            [dim-1 translation]
// with the same CoilArgs contract (U/V state strides, state lists, weights,
// energy partials) as the fused stage.  Scratch: CoilArgs::fwork.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "transform.hpp"

namespace dis {
namespace {

constexpr int kFoldBlocks = 4096;

inline int blocks_for(long long n) { return (int)std::max<long long>(1, std::min<long long>(kFoldBlocks, (n + kThreads - 1) / kThreads)); }

__device__ __forceinline__ int state_of(const CoilArgs& a, int sl) { return a.slist ? a.slist[sl] : a.state0 + sl; }

// Columns of length n along dim 1 ([n x H], contiguous pencils), P per tile:
// out = FFT^-1(phase(c) FFT(in)) / n  (translation along dim 1, in place allowed).
__global__ __launch_bounds__(kThreads) void translate1_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                               int n, long long H, int P, FftPlan pl,
                                                               const float2* __restrict__ gtw, const double* coef,
                                                               double invN1, const int* slist, int state0, int sl) {
  __shared__ float2 buf[kMaxTile];
  __shared__ float2 tw[1024];
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += kThreads) tw[k] = gtw[k];
  const double cf = coef[slist ? slist[sl] : state0 + sl] * invN1;
  const long long h0 = (long long)blockIdx.x * P;
  const int Pv = (int)(H - h0 < P ? H - h0 : P);
  const int tot = P * n;
  for (int e = tid; e < tot; e += kThreads) buf[e] = e < Pv * n ? in[h0 * n + e] : make_float2(0.f, 0.f);
  __syncthreads();
  fft_tile<true>(buf, tw, pl, P, tid);
  for (int e = tid; e < tot; e += kThreads) buf[e] = cswap(cmul(buf[e], phase_cycles(cf, freq_index(e % n, n))));
  __syncthreads();
  fft_tile<true>(buf, tw, pl, P, tid);
  const float inv = 1.0f / (float)n;
  for (int e = tid; e < Pv * n; e += kThreads) out[h0 * n + e] = cscale(cswap(buf[e]), inv);
}

// In-place FFT (forward, or inverse without 1/n) of length n along dim 1 of [n x H].
template <bool INV>
__global__ __launch_bounds__(kThreads) void fft1_kernel(float2* __restrict__ z, int n, long long H, int P, FftPlan pl,
                                                         const float2* __restrict__ gtw) {
  __shared__ float2 buf[kMaxTile];
  __shared__ float2 tw[1024];
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += kThreads) tw[k] = gtw[k];
  const long long h0 = (long long)blockIdx.x * P;
  const int Pv = (int)(H - h0 < P ? H - h0 : P);
  const int tot = P * n;
  for (int e = tid; e < tot; e += kThreads) {
    const float2 v = e < Pv * n ? z[h0 * n + e] : make_float2(0.f, 0.f);
    buf[e] = INV ? cswap(v) : v;
  }
  __syncthreads();
  fft_tile<true>(buf, tw, pl, P, tid);
  for (int e = tid; e < Pv * n; e += kThreads) z[h0 * n + e] = INV ? cswap(buf[e]) : buf[e];
}

struct FoldGeom {
  int n1, n2, n3, C;
  int ny1, ny2, o1, o2;
};

// Z[i1 + ny1 (i2 + ny2 (i3 + n3 c))] = sum over r1 = i1 + o1 (mod ny1),
// r2 = i2 + o2 (mod ny2) of S_c[r] X[r]   (coil multiply, then fold.m on
// dims 1 and 2)
__global__ void fold_mul_kernel(const float2* __restrict__ X, const float2* __restrict__ S, FoldGeom f,
                                float2* __restrict__ Z) {
  const long long nz = (long long)f.ny1 * f.ny2 * f.n3 * f.C;
  const long long nplane = (long long)f.n1 * f.n2;
  const long long nvox = nplane * f.n3;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nz; e += (long long)gridDim.x * blockDim.x) {
    const int i1 = (int)(e % f.ny1);
    long long t = e / f.ny1;
    const int i2 = (int)(t % f.ny2);
    t /= f.ny2;
    const int i3 = (int)(t % f.n3);
    const int c = (int)(t / f.n3);
    const float2* Xp = X + i3 * nplane;
    const float2* Sp = S + (long long)c * nvox + i3 * nplane;
    float2 acc = make_float2(0.f, 0.f);
    for (int r2 = (i2 + f.o2) % f.ny2; r2 < f.n2; r2 += f.ny2)
      for (int r1 = (i1 + f.o1) % f.ny1; r1 < f.n1; r1 += f.ny1) {
        const long long g = r1 + (long long)f.n1 * r2;
        acc = cadd(acc, cmul(Sp[g], Xp[g]));
      }
    Z[e] = acc;
  }
}

// V[r] = sum_c conj(S_c[r]) Z[(r1 - o1) mod ny1, (r2 - o2) mod ny2, i3, c]  (ifold.m, decode.m:82-84)
__global__ void unfold_mul_kernel(const float2* __restrict__ Z, const float2* __restrict__ S, FoldGeom f,
                                  float2* __restrict__ V) {
  const long long nplane = (long long)f.n1 * f.n2;
  const long long nvox = nplane * f.n3;
  const long long zc = (long long)f.ny1 * f.ny2 * f.n3;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nvox; e += (long long)gridDim.x * blockDim.x) {
    const int r1 = (int)(e % f.n1);
    long long t = e / f.n1;
    const int r2 = (int)(t % f.n2);
    const int i3 = (int)(t / f.n2);
    int i1 = (r1 - f.o1) % f.ny1;
    if (i1 < 0) i1 += f.ny1;
    int i2 = (r2 - f.o2) % f.ny2;
    if (i2 < 0) i2 += f.ny2;
    const long long zi = i1 + (long long)f.ny1 * (i2 + (long long)f.ny2 * i3);
    float2 acc = make_float2(0.f, 0.f);
    for (int c = 0; c < f.C; ++c) acc = cadd(acc, cmulc(Z[zi + zc * c], S[e + nvox * c]));
    V[e] = acc;
  }
}

// Forward rows: per (sample of the state, readout plane) the DFT along dim 2
// of the dim-1 spectra at row k1, for every coil, scaled 1/sqrt(ny1 ny2).
// ENCODE writes y; ENERGY the per-(sample, plane) residual energy over the
// coils; E^H E keeps We * E x for the adjoint in Yc [nec x n3 x C].
template <int MODE>
__global__ __launch_bounds__(kThreads) void rows_fwd_kernel(CoilArgs a, FoldGeom f, int sl, const float2* __restrict__ Z,
                                                            float2* __restrict__ Yc, int maxnec) {
  __shared__ float2 tw[1024];
  for (int k = threadIdx.x; k < f.ny2; k += kThreads) tw[k] = a.fy2tw[k];
  __syncthreads();
  const int s = state_of(a, sl);
  const int o0 = a.off[s], nec = a.off[s + 1] - o0;
  const long long tot = (long long)nec * f.n3;
  const float inv_sqrt = rsqrtf((float)f.ny1 * (float)f.ny2);
  const long long zc = (long long)f.ny1 * f.ny2 * f.n3;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int jl = (int)(e % nec);
    const int i3 = (int)(e / nec);
    const int j = o0 + jl;
    const int k1 = a.k1[j], k2 = a.k2[j];
    const float w = a.We ? a.We[j] : 1.0f;
    float esum = 0.f;
    for (int c = 0; c < f.C; ++c) {
      const float2* Zr = Z + k1 + (long long)f.ny1 * f.ny2 * i3 + zc * c;
      float2 acc = make_float2(0.f, 0.f);
      int idx = 0;
      for (int i2 = 0; i2 < f.ny2; ++i2) {
        acc = cadd(acc, cmul(tw[idx], Zr[(long long)f.ny1 * i2]));
        idx += k2;
        if (idx >= f.ny2) idx -= f.ny2;
      }
      const float2 Y = cscale(acc, inv_sqrt);
      const long long yi = j + (long long)a.Ns * (i3 + (long long)f.n3 * c);
      if (MODE == COIL_ENCODE) a.yout[yi] = Y;
      if (MODE == COIL_ENERGY) {
        const float2 r = csub(Y, a.yin[yi]);
        esum += r.x * r.x + r.y * r.y;
      }
      if (MODE == COIL_EHE) Yc[jl + (long long)maxnec * (i3 + (long long)f.n3 * c)] = cscale(Y, w);
    }
    if (MODE == COIL_ENERGY) a.epart[j + (long long)a.Ns * i3] = esum;
  }
}

// Adjoint rows: Z[i1, i2, i3, c] = sum over the state's samples in row i1 of
// conj(W2^{k2 i2}) y_j / sqrt(ny1 ny2)  (DECODE: y = We .* yin; E^H E: Yc).
template <int MODE>
__global__ __launch_bounds__(kThreads) void rows_adj_kernel(CoilArgs a, FoldGeom f, int sl, const float2* __restrict__ Yc,
                                                            int maxnec, float2* __restrict__ Z) {
  __shared__ float2 tw[1024];
  for (int k = threadIdx.x; k < f.ny2; k += kThreads) tw[k] = a.fy2tw[k];
  __syncthreads();
  const int s = state_of(a, sl);
  const int o0 = a.off[s];
  const int* rp = a.rowptr + (long long)s * (f.n1 + 1);
  const long long tot = (long long)f.ny1 * f.ny2 * f.n3;
  const float inv_sqrt = rsqrtf((float)f.ny1 * (float)f.ny2);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int i1 = (int)(e % f.ny1);
    long long t = e / f.ny1;
    const int i2 = (int)(t % f.ny2);
    const int i3 = (int)(t / f.ny2);
    const int q0 = rp[i1], q1 = rp[i1 + 1];
    for (int c = 0; c < f.C; ++c) {
      float2 acc = make_float2(0.f, 0.f);
      for (int q = q0; q < q1; ++q) {
        const int ent = a.csr[o0 + q];
        const int k2 = ent >> 16, jl = ent & 0xffff;
        float2 yv;
        if (MODE == COIL_DECODE) {
          const int j = o0 + jl;
          yv = cscale(a.yin[j + (long long)a.Ns * (i3 + (long long)f.n3 * c)], a.We ? a.We[j] : 1.0f);
        } else {
          yv = Yc[jl + (long long)maxnec * (i3 + (long long)f.n3 * c)];
        }
        acc = cadd(acc, cmulc(yv, tw[(int)(((long long)k2 * i2) % f.ny2)]));
      }
      Z[e + tot * c] = cscale(acc, inv_sqrt);
    }
  }
}

int tile_pencils(int n) { return std::max(1, kMaxTile / n); }

// Strided FFT along dim 2 (length n2 = pl.n) of Z [L x n2 x H] (L = ny1
// contiguous positions), P positions per tile, then the weighted energy
// sum W[k1 + L k2] |X|^2 of the tile into part[blockIdx.x] (FP64, fixed order).
__global__ __launch_bounds__(kThreads) void fft2_energy_kernel(const float2* __restrict__ z, long long L, long long H,
                                                                int P, long long tiles_lo, FftPlan pl,
                                                                const float2* __restrict__ gtw,
                                                                const float* __restrict__ W, double* __restrict__ part) {
  __shared__ float2 buf[kMaxTile];
  __shared__ float2 tw[1024];
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int n = pl.n;
  for (int k = tid; k < n; k += kThreads) tw[k] = gtw[k];
  const long long hi0 = blockIdx.x / tiles_lo;
  const long long lo0 = (blockIdx.x - hi0 * tiles_lo) * (long long)P;
  const int Pv = (int)(L - lo0 < P ? L - lo0 : P);
  const int tot = P * n;
  for (int e = tid; e < tot; e += kThreads) {
    const int j = e / P, p = e - j * P;
    buf[e] = p < Pv ? z[(lo0 + p) + L * ((long long)j + (long long)n * hi0)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_tile<false>(buf, tw, pl, P, tid);
  double acc = 0.0;
  for (int e = tid; e < tot; e += kThreads) {
    const int j = e / P, p = e - j * P;
    if (p >= Pv) continue;
    const float2 v = buf[e];
    acc += (double)W[(lo0 + p) + L * j] * ((double)v.x * v.x + (double)v.y * v.y);
  }
  red[tid] = acc;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = red[0];
}

__global__ void sum_partials_kernel(const double* __restrict__ part, long long n, double* __restrict__ out) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += kThreads) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace

// traceEstimation.m:61-79 with SENSE folding: sum over coils of
// sum_k W(k1, k2) |F2(fold(S_c x)) / sqrt(ny1 ny2)|^2, the readout in image
// space.  Scratch from context buffer slot 8.
int fold_fidelity_trace(dis_ctx* c, const int32_t n[3], const float2* S, int C, int ny1, int ny2, const float* W,
                        const float2* x, double* out) {
  FftPlan p1, p2;
  if (ny1 <= 0 || ny2 <= 0 || ny1 > n[0] || ny2 > n[1] || ny1 > 1024 || ny2 > 1024 || make_plan(ny1, &p1) ||
      make_plan(ny2, &p2))
    return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_fidelity_trace_folded: unsupported acquired grid");
  const float2 *tw1 = nullptr, *tw2 = nullptr;
  int rc = ctx_twiddles(c, ny1, &tw1);
  if (!rc) rc = ctx_twiddles(c, ny2, &tw2);
  if (rc) return rc;
  FoldGeom f;
  f.n1 = n[0];
  f.n2 = n[1];
  f.n3 = n[2];
  f.C = C;
  f.ny1 = ny1;
  f.ny2 = ny2;
  f.o1 = (n[0] - ny1 + 1) / 2;
  f.o2 = (n[1] - ny2 + 1) / 2;
  const long long nz = (long long)ny1 * ny2 * n[2] * C;
  const long long L = ny1, H = (long long)n[2] * C;
  const int P2 = (int)std::min<long long>(tile_pencils(ny2), L);
  const long long tiles_lo = (L + P2 - 1) / P2;
  const long long ntiles = tiles_lo * H;
  void* buf = nullptr;
  rc = ctx_buffer(c, 8, sizeof(float2) * nz + sizeof(double) * (ntiles + 1) + 256, &buf);
  if (rc) return rc;
  float2* Z = (float2*)buf;
  double* part = (double*)((char*)buf + ((sizeof(float2) * nz + 255) & ~size_t(255)));
  double* tot = part + ntiles;
  hipStream_t st = ctx_stream(c);
  hipLaunchKernelGGL(fold_mul_kernel, dim3(blocks_for(nz)), dim3(kThreads), 0, st, x, S, f, Z);
  const long long Hy = (long long)ny2 * n[2] * C;
  const int Py = tile_pencils(ny1);
  hipLaunchKernelGGL(fft1_kernel<false>, dim3((unsigned)((Hy + Py - 1) / Py)), dim3(kThreads), 0, st, Z, ny1, Hy, Py,
                     p1, tw1);
  hipLaunchKernelGGL(fft2_energy_kernel, dim3((unsigned)ntiles), dim3(kThreads), 0, st, Z, L, H, P2, tiles_lo, p2,
                     tw2, W, part);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kThreads), 0, st, part, ntiles, tot);
  double h = 0.0;
  if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&h, tot, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_fidelity_trace_folded failed");
  *out = h / ((double)ny1 * (double)ny2);
  return DIS_OK;
}

size_t coil_fold_work_bytes(int n1, int n2, int n3, int C, int ny1, int ny2, int maxnec) {
  const size_t x = (size_t)n1 * n2 * n3;
  const size_t z = (size_t)ny1 * ny2 * n3 * C;
  const size_t yc = (size_t)std::max(1, maxnec) * n3 * C;
  return sizeof(float2) * (x + z + yc) + 512;
}

hipError_t launch_coil_fold(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  // U/V indexed by state number need the state on the host: only lists with
  // list-relative strides, or consecutive states
  if (a.fy1 <= 0 || a.fy2 <= 0 || a.fy1 > a.n1 || a.fy2 > a.n2 || a.fy1 > 1024 || a.fy2 > 1024 || !a.fwork ||
      !a.fy1tw || !a.fy2tw || a.n1 > 1024 || (a.slist && (!a.us_rel || (a.V && !a.vs_rel))))
    return hipErrorInvalidValue;
  FoldGeom f;
  f.n1 = a.n1;
  f.n2 = a.n2;
  f.n3 = a.n3;
  f.C = a.C;
  f.ny1 = a.fy1;
  f.ny2 = a.fy2;
  f.o1 = (a.n1 - a.fy1 + 1) / 2;  // FOV(2) = ceil((NX - NY) / 2), fold.m:42-43
  f.o2 = (a.n2 - a.fy2 + 1) / 2;
  const long long nv = (long long)a.n1 * a.n2 * a.n3;
  const long long nz = (long long)f.ny1 * f.ny2 * f.n3 * f.C;
  float2* X = a.fwork;
  float2* Z = X + nv;
  float2* Yc = Z + nz;
  const bool fwd = mode != COIL_DECODE;
  const bool adj = mode == COIL_DECODE || mode == COIL_EHE;
  const long long H1 = (long long)a.n2 * a.n3;
  const int P1 = tile_pencils(a.n1);
  const int Py = tile_pencils(f.ny1);
  const long long Hy = (long long)f.ny2 * f.n3 * f.C;
  const long long nrow = (long long)a.fmaxnec * f.n3;
  for (int sl = 0; sl < nstates; ++sl) {
    if (fwd) {
      // U indexed like the fused stage: by list position (us_rel) or by state
      // (the state number itself is only known on the device for lists)
      const float2* Us = a.U + (long long)(a.us_rel ? sl : a.state0 + sl) * a.us;
      const float2* Xs = Us;
      if (a.tfwd) {
        hipLaunchKernelGGL(translate1_kernel, dim3((unsigned)((H1 + P1 - 1) / P1)), dim3(kThreads), 0, st, Us, X,
                           a.n1, H1, P1, a.plan1, a.tw1, a.tfwd, a.invN1, a.slist, a.state0, sl);
        Xs = X;
      }
      hipLaunchKernelGGL(fold_mul_kernel, dim3(blocks_for(nz)), dim3(kThreads), 0, st, Xs, a.Smaps, f, Z);
      hipLaunchKernelGGL(fft1_kernel<false>, dim3((unsigned)((Hy + Py - 1) / Py)), dim3(kThreads), 0, st, Z, f.ny1, Hy,
                         Py, a.fy1pl, a.fy1tw);
      if (mode == COIL_ENCODE)
        hipLaunchKernelGGL(rows_fwd_kernel<COIL_ENCODE>, dim3(blocks_for(nrow)), dim3(kThreads), 0, st, a, f, sl, Z, Yc,
                           a.fmaxnec);
      if (mode == COIL_ENERGY)
        hipLaunchKernelGGL(rows_fwd_kernel<COIL_ENERGY>, dim3(blocks_for(nrow)), dim3(kThreads), 0, st, a, f, sl, Z, Yc,
                           a.fmaxnec);
      if (mode == COIL_EHE)
        hipLaunchKernelGGL(rows_fwd_kernel<COIL_EHE>, dim3(blocks_for(nrow)), dim3(kThreads), 0, st, a, f, sl, Z, Yc,
                           a.fmaxnec);
    }
    if (adj) {
      const long long nzy = (long long)f.ny1 * f.ny2 * f.n3;
      if (mode == COIL_DECODE)
        hipLaunchKernelGGL(rows_adj_kernel<COIL_DECODE>, dim3(blocks_for(nzy)), dim3(kThreads), 0, st, a, f, sl, Yc,
                           a.fmaxnec, Z);
      else
        hipLaunchKernelGGL(rows_adj_kernel<COIL_EHE>, dim3(blocks_for(nzy)), dim3(kThreads), 0, st, a, f, sl, Yc,
                           a.fmaxnec, Z);
      hipLaunchKernelGGL(fft1_kernel<true>, dim3((unsigned)((Hy + Py - 1) / Py)), dim3(kThreads), 0, st, Z, f.ny1, Hy,
                         Py, a.fy1pl, a.fy1tw);
      float2* Vs = a.V + (long long)(a.vs_rel ? sl : a.state0 + sl) * a.vs;
      hipLaunchKernelGGL(unfold_mul_kernel, dim3(blocks_for(nv)), dim3(kThreads), 0, st, Z, a.Smaps, f, Vs);
      if (a.tadj)
        hipLaunchKernelGGL(translate1_kernel, dim3((unsigned)((H1 + P1 - 1) / P1)), dim3(kThreads), 0, st, Vs, Vs,
                           a.n1, H1, P1, a.plan1, a.tw1, a.tadj, a.invN1, a.slist, a.state0, sl);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace dis
