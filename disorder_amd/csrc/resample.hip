// Pyramid resampling (Methods/resampling.m:45-89 with fo = 0, mirror = 0;
// Methods/downsampleOperators.m:34-60) and the per-level data plumbing.
//
// resample_kernel: along one axis of an [L x n x H] array (L = product of the
// dimensions before the axis, H = product of those after it, batch
// included), each pencil is transformed (FFT of length n_in), its spectrum is
// cropped or zero-padded about DC, and transformed back at length n_out:
//   out = (1/n_in) IFFT_{n_out}( crop(FFT_{n_in}(x)) )
// which is resampling.m's fft/N_in, fftshift, crop/pad centred at
// ceil((Nmax+1)/2), ifftshift, ifft*N_out.  The kept signed frequencies are
// f in [-ceil((Nmin-1)/2), floor((Nmin-1)/2)], Nmin = min(n_in, n_out).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/disorder.h"
#include "kernels.hpp"
#include "transform.hpp"

namespace dis {

namespace {

struct ResampleArgs {
  const float2* in;
  float2* out;
  long long L, H;  // pencil index space: i_lo < L, i_hi < H
  int nin, nout;
  int P;           // pencils per tile
  long long tiles_lo;  // strided: tiles along i_lo per i_hi
  FftPlan pin, pout;
  const float2* twin;
  const float2* twout;
};

// signed frequency of FFT-order index k of a length-n transform
__device__ __forceinline__ int signed_freq(int k, int n) { return k < (n + 1) / 2 ? k : k - n; }

template <bool CONTIG>
__global__ __launch_bounds__(kThreads) void resample_kernel(ResampleArgs a) {
  __shared__ float2 bin[kMaxTile];
  __shared__ float2 bout[kMaxTile];
  __shared__ float2 twi[1024];
  __shared__ float2 two[1024];
  const int tid = threadIdx.x;
  for (int k = tid; k < a.nin; k += kThreads) twi[k] = a.twin[k];
  for (int k = tid; k < a.nout; k += kThreads) two[k] = a.twout[k];
  const int P = a.P;
  // tile origin
  long long hi0, lo0;
  int Pv;
  if (CONTIG) {  // L == 1: P consecutive i_hi
    hi0 = (long long)blockIdx.x * P;
    lo0 = 0;
    Pv = (int)(a.H - hi0 < P ? a.H - hi0 : P);
  } else {
    hi0 = blockIdx.x / a.tiles_lo;
    lo0 = (blockIdx.x - hi0 * a.tiles_lo) * (long long)P;
    Pv = (int)(a.L - lo0 < P ? a.L - lo0 : P);
  }
  auto in_addr = [&](int p, int j) -> long long {
    return CONTIG ? (long long)j + (long long)a.nin * (hi0 + p) : (lo0 + p) + a.L * ((long long)j + (long long)a.nin * hi0);
  };
  auto out_addr = [&](int p, int j) -> long long {
    return CONTIG ? (long long)j + (long long)a.nout * (hi0 + p)
                  : (lo0 + p) + a.L * ((long long)j + (long long)a.nout * hi0);
  };
  const int totin = P * a.nin, totout = P * a.nout;
  for (int e = tid; e < totin; e += kThreads) {
    int p, j;
    if (CONTIG) { p = e / a.nin; j = e - p * a.nin; } else { j = e / P; p = e - j * P; }
    bin[e] = p < Pv ? a.in[in_addr(p, j)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_tile<CONTIG>(bin, twi, a.pin, P, tid);
  const int nmin = a.nin < a.nout ? a.nin : a.nout;
  const int flo = -(nmin / 2), fhi = (nmin - 1) / 2;  // -ceil((nmin-1)/2) == -(nmin/2)
  for (int e = tid; e < totout; e += kThreads) {
    int p, q;
    if (CONTIG) { p = e / a.nout; q = e - p * a.nout; } else { q = e / P; p = e - q * P; }
    const int f = signed_freq(q, a.nout);
    float2 v = make_float2(0.f, 0.f);
    if (f >= flo && f <= fhi) {
      const int j = f < 0 ? f + a.nin : f;
      v = bin[CONTIG ? p * a.nin + j : j * P + p];
    }
    bout[e] = cswap(v);  // inverse transform by swapping real and imaginary parts
  }
  __syncthreads();
  fft_tile<CONTIG>(bout, two, a.pout, P, tid);
  const float scale = 1.0f / (float)a.nin;
  for (int e = tid; e < totout; e += kThreads) {
    int p, q;
    if (CONTIG) { p = e / a.nout; q = e - p * a.nout; } else { q = e / P; p = e - q * P; }
    if (p < Pv) a.out[out_addr(p, q)] = cscale(cswap(bout[e]), scale);
  }
}

// Radial Tukey window of buildFilter(N, 'tukeyIso', [], gpu, gibbsRing)
// (Methods/buildFilter.m:69-92, ifftshifted): |k| = sqrt(sum_d (2 pi f_d / N_d)^2)
// over the signed frequencies f_d of the DC-first grid; 1 below pi (1 - g),
// a raised-cosine taper to 0 at pi, 0 beyond.
__device__ __forceinline__ float tukey_iso(int f1, int f2, int f3, int n1, int n2, int n3, float g) {
  const float k1 = 6.28318530717958647692f * (float)f1 / (float)n1;
  const float k2 = 6.28318530717958647692f * (float)f2 / (float)n2;
  const float k3 = 6.28318530717958647692f * (float)f3 / (float)n3;
  const float kk = sqrtf(k1 * k1 + k2 * k2 + k3 * k3);
  const float pi = 3.14159265358979323846f;
  const float alpha = 1.0f - g;
  if (kk >= pi) return 0.f;
  if (kk >= pi * alpha) return 0.5f * (1.0f + cosf(pi * ((kk - pi * alpha) / ((1.0f - alpha) * pi))));
  return 1.f;
}

// buildFilter(N, 'CubicBSpline') (Methods/buildFilter.m:102-106): per
// dimension 6 sinc(f/N)^4 / (4 + 2 cos(2 pi f/N)), sinc(t) = sin(pi t)/(pi t).
__device__ __forceinline__ float bspline1(int f, int n) {
  const float t = (float)f / (float)n;
  const float s = f == 0 ? 1.0f : sinpif(t) / (3.14159265358979323846f * t);
  const float s2 = s * s;
  return 6.0f * s2 * s2 / (4.0f + 2.0f * cospif(2.0f * t));
}

struct RowFilterArgs {
  float2* y;       // [Ns x n3 x C], filtered in place
  long long Ns, C;
  int n3, P;
  long long tiles_lo;
  FftPlan pl;
  const float2* tw;
  const int2* k12;  // level FFT-order indices of every row
  int n1, n2;
  float g;
};

// filtering(y, H) (Methods/filtering.m:40-67) for compact rows whose phase
// encodes are already in k-space: per row, FFT along the readout, times the
// window at (k1, k2, k3), inverse FFT.  Tiles of P consecutive rows, strided
// LDS layout; every tile reads all its pencils before writing them.
__global__ __launch_bounds__(kThreads) void tukey_rows_kernel(RowFilterArgs a) {
  __shared__ float2 buf[kMaxTile];
  __shared__ float2 tw[1024];
  const int tid = threadIdx.x;
  for (int k = tid; k < a.n3; k += kThreads) tw[k] = a.tw[k];
  const int P = a.P;
  const long long c = blockIdx.x / a.tiles_lo;
  const long long r0 = (blockIdx.x - c * a.tiles_lo) * (long long)P;
  const int Pv = (int)(a.Ns - r0 < P ? a.Ns - r0 : P);
  const int tot = P * a.n3;
  for (int e = tid; e < tot; e += kThreads) {
    const int j = e / P, p = e - j * P;
    buf[e] = p < Pv ? a.y[(r0 + p) + a.Ns * ((long long)j + (long long)a.n3 * c)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_tile<false>(buf, tw, a.pl, P, tid);
  for (int e = tid; e < tot; e += kThreads) {
    const int j = e / P, p = e - j * P;
    float h = 0.f;
    if (p < Pv) {
      const int2 k = a.k12[r0 + p];
      const int f1 = signed_freq(k.x, a.n1), f2 = signed_freq(k.y, a.n2), f3 = signed_freq(j, a.n3);
      h = a.g <= 1.0f ? tukey_iso(f1, f2, f3, a.n1, a.n2, a.n3, a.g)
                      : bspline1(f1, a.n1) * bspline1(f2, a.n2) * bspline1(f3, a.n3);
    }
    buf[e] = cswap(cscale(buf[e], h));
  }
  __syncthreads();
  fft_tile<false>(buf, tw, a.pl, P, tid);
  const float inv = 1.0f / (float)a.n3;
  for (int e = tid; e < tot; e += kThreads) {
    const int j = e / P, p = e - j * P;
    if (p < Pv) a.y[(r0 + p) + a.Ns * ((long long)j + (long long)a.n3 * c)] = cscale(cswap(buf[e]), inv);
  }
}

// out[r + Ns_out (i3 + n3 c)] = scale * in[rows[r] + Ns_in (i3 + n3 c)]
__global__ void gather_rows_kernel(const float2* __restrict__ in, long long Ns_in, const int* __restrict__ rows,
                                   long long Ns_out, long long ncol, float scale, float2* __restrict__ out) {
  const long long tot = Ns_out * ncol;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long col = e / Ns_out, r = e - col * Ns_out;
    out[e] = cscale(in[rows[r] + Ns_in * col], scale);
  }
}

// Se = 1 / (sum_c |S_c|^2 + eps) per voxel (solveXT.m:231-232)
__global__ void coil_precond_kernel(const float2* __restrict__ S, long long nv, int C, float eps,
                                    float* __restrict__ Se) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 z = S[v + nv * c];
      s += z.x * z.x + z.y * z.y;
    }
    Se[v] = 1.0f / (s + eps);
  }
}

}  // namespace

// One axis of the resampling: x [L x nin x H] -> out [L x nout x H].
int resample_axis(dis_ctx* c, const float2* x, float2* out, long long L, int nin, int nout, long long H) {
  ResampleArgs a;
  a.in = x;
  a.out = out;
  a.L = L;
  a.H = H;
  a.nin = nin;
  a.nout = nout;
  int st = make_plan(nin, &a.pin);
  if (!st) st = make_plan(nout, &a.pout);
  if (st) return ctx_fail(c, DIS_E_UNSUPPORTED, "resample: length with a prime factor above 7 or above 1024");
  const float2* tw = nullptr;
  if ((st = ctx_twiddles(c, nin, &tw))) return st;
  a.twin = tw;
  if ((st = ctx_twiddles(c, nout, &tw))) return st;
  a.twout = tw;
  const int nmax = std::max(nin, nout);
  int P = std::max(1, kMaxTile / nmax);
  const bool contig = (L == 1);
  long long tiles;
  if (contig) {
    P = (int)std::min<long long>(P, std::max<long long>(1, H));
    tiles = (H + P - 1) / P;
  } else {
    P = (int)std::min<long long>(P, L);
    a.tiles_lo = (L + P - 1) / P;
    tiles = a.tiles_lo * H;
  }
  a.P = P;
  if (contig) a.tiles_lo = 1;
  if (tiles > 0x7fffffffLL) return ctx_fail(c, DIS_E_UNSUPPORTED, "resample: too many tiles");
  hipStream_t s = ctx_stream(c);
  if (contig)
    hipLaunchKernelGGL(resample_kernel<true>, dim3((unsigned)tiles), dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(resample_kernel<false>, dim3((unsigned)tiles), dim3(kThreads), 0, s, a);
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "resample launch failed");
  return DIS_OK;
}

}  // namespace dis

using namespace dis;

extern "C" {

int dis_resample(dis_ctx* c, const int32_t nin[3], const int32_t nout[3], int64_t batch, const void* x, void* out) {
  if (!c) return DIS_E_INVALID;
  if (!nin || !nout || !x || !out || batch <= 0) return ctx_fail(c, DIS_E_INVALID, "dis_resample: invalid argument");
  for (int d = 0; d < 3; ++d)
    if (nin[d] <= 0 || nout[d] <= 0) return ctx_fail(c, DIS_E_INVALID, "dis_resample: bad size");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  // axes in order 0, 1, 2; intermediate shapes in two workspace halves
  long long cur[3] = {nin[0], nin[1], nin[2]};
  int todo = 0;
  for (int d = 0; d < 3; ++d) todo += (nin[d] != nout[d]);
  const long long vin = (long long)nin[0] * nin[1] * nin[2];
  const long long vout = (long long)nout[0] * nout[1] * nout[2];
  if (todo == 0) {
    if (hipMemcpyAsync(out, x, sizeof(float2) * vin * batch, hipMemcpyDeviceToDevice, ctx_stream(c)) != hipSuccess)
      return ctx_fail(c, DIS_E_HIP, "dis_resample: copy failed");
    return hipStreamSynchronize(ctx_stream(c)) == hipSuccess ? DIS_OK : ctx_fail(c, DIS_E_HIP, "sync failed");
  }
  // largest intermediate volume
  long long vmax = 0;
  {
    long long sh[3] = {nin[0], nin[1], nin[2]};
    for (int d = 0; d < 3; ++d) {
      sh[d] = nout[d];
      vmax = std::max(vmax, sh[0] * sh[1] * sh[2]);
    }
  }
  void* ws = nullptr;
  int st = todo > 1 ? ctx_workspace(c, sizeof(float2) * vmax * batch * (todo > 2 ? 2 : 1), &ws) : DIS_OK;
  if (st) return st;
  float2* bufs[2] = {(float2*)ws, (float2*)ws + vmax * batch};
  const float2* src = (const float2*)x;
  int done = 0;
  for (int d = 0; d < 3; ++d) {
    if (cur[d] == nout[d]) continue;
    ++done;
    float2* dst = (done == todo) ? (float2*)out : bufs[(done - 1) & 1];
    long long L = 1, H = batch;
    for (int e = 0; e < d; ++e) L *= cur[e];
    for (int e = d + 1; e < 3; ++e) H *= cur[e];
    st = resample_axis(c, src, dst, L, (int)cur[d], nout[d], H);
    if (st) return st;
    cur[d] = nout[d];
    src = dst;
  }
  (void)vout;
  if (hipStreamSynchronize(ctx_stream(c)) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "dis_resample failed");
  return DIS_OK;
}

int dis_gather_samples(dis_ctx* c, const void* y, int32_t Ns_in, int64_t ncol, const int32_t* rows, int32_t Ns_out,
                       float scale, void* out) {
  if (!c) return DIS_E_INVALID;
  if (!y || !out || Ns_in < 0 || Ns_out < 0 || ncol < 0 || (Ns_out > 0 && !rows))
    return ctx_fail(c, DIS_E_INVALID, "dis_gather_samples: invalid argument");
  for (int r = 0; r < Ns_out; ++r)
    if (rows[r] < 0 || rows[r] >= Ns_in) return ctx_fail(c, DIS_E_INVALID, "dis_gather_samples: row out of range");
  if (Ns_out == 0 || ncol == 0) return DIS_OK;
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  Staging stg;
  stg.add(rows, sizeof(int32_t) * Ns_out);
  const unsigned char* d = nullptr;
  int st = ctx_upload_bytes(c, stg, &d);
  if (st) return st;
  const long long tot = (long long)Ns_out * ncol;
  const unsigned nb = (unsigned)std::min<long long>((tot + 255) / 256, 8192);
  hipLaunchKernelGGL(gather_rows_kernel, dim3(nb), dim3(256), 0, ctx_stream(c), (const float2*)y, (long long)Ns_in,
                     (const int*)d, (long long)Ns_out, (long long)ncol, scale, (float2*)out);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx_stream(c)) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_gather_samples failed");
  return DIS_OK;
}

int dis_filter_rows(dis_ctx* c, void* y, int64_t Ns, int32_t n3, int32_t C, const int32_t* k1, const int32_t* k2,
                    int32_t n1, int32_t n2, float gibbs) {
  if (!c) return DIS_E_INVALID;
  if (!y || !k1 || !k2 || Ns < 0 || n3 <= 0 || C <= 0 || n1 <= 0 || n2 <= 0 || !(gibbs > 0.f))
    return ctx_fail(c, DIS_E_INVALID, "dis_filter_rows: invalid argument");
  if (Ns == 0) return DIS_OK;
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  for (long long r = 0; r < Ns; ++r)
    if (k1[r] < 0 || k1[r] >= n1 || k2[r] < 0 || k2[r] >= n2)
      return ctx_fail(c, DIS_E_INVALID, "dis_filter_rows: sample index outside the grid");
  RowFilterArgs a{};
  if (make_plan(n3, &a.pl)) return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_filter_rows: unsupported readout length");
  int st = ctx_twiddles(c, n3, &a.tw);
  if (st) return st;
  std::vector<int2> h((size_t)Ns);
  for (long long r = 0; r < Ns; ++r) h[r] = make_int2(k1[r], k2[r]);
  void* dk = nullptr;
  st = ctx_buffer(c, 5, sizeof(int2) * (size_t)Ns, &dk);
  if (st) return st;
  hipStream_t s = ctx_stream(c);
  if (hipMemcpyAsync(dk, h.data(), sizeof(int2) * (size_t)Ns, hipMemcpyHostToDevice, s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_filter_rows: upload failed");
  a.y = (float2*)y;
  a.Ns = Ns;
  a.C = C;
  a.n3 = n3;
  a.k12 = (const int2*)dk;
  a.n1 = n1;
  a.n2 = n2;
  a.g = gibbs;
  a.P = (int)std::min<long long>(std::max(1, kMaxTile / n3), Ns);
  a.tiles_lo = (Ns + a.P - 1) / a.P;
  const long long tiles = a.tiles_lo * C;
  if (tiles > 0x7fffffffLL) return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_filter_rows: too many tiles");
  hipLaunchKernelGGL(tukey_rows_kernel, dim3((unsigned)tiles), dim3(kThreads), 0, s, a);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_filter_rows failed");
  return DIS_OK;
}

int dis_coil_precond(dis_ctx* c, const void* S, int64_t nv, int32_t C, float eps, float* Se) {
  if (!c) return DIS_E_INVALID;
  if (!S || !Se || nv <= 0 || C <= 0) return ctx_fail(c, DIS_E_INVALID, "dis_coil_precond: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  const unsigned nb = (unsigned)std::min<long long>((nv + 255) / 256, 16384);
  hipLaunchKernelGGL(coil_precond_kernel, dim3(nb), dim3(256), 0, ctx_stream(c), (const float2*)S, (long long)nv, C,
                     eps, Se);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx_stream(c)) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_coil_precond failed");
  return DIS_OK;
}

}  // extern "C"
