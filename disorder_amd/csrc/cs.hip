// Shearlet-regularised final reconstruction (Reconstruction/CSsolver.m,
// the shearlet term of Reconstruction/regularize.m:27-50, the trace
// estimates of Reconstruction/traceEstimation.m).
//
// The 3-D shearlets of buildShearlet(N, 1, ., 1, 'kos') are products of two
// 2-D filters on coordinate planes (combineShearlets): band w is
// S_w(k1,k2,k3) = A_w(plane a) B_w(plane b) with planes (k1,k2), (k1,k3) and
// (k3,k2).  The planes (11 filters each, a few MB) are all that lives on the
// device; every pass evaluates the band it needs from them on the fly, so no
// [N x 50] filter bank is ever stored or read.
//
// fft_axis_kernel: one axis of a 3-D (inverse) FFT over [L x n x H] with a
// fused load op (x conj(S_w), x We_w, x S_c) and store op (plain, += S_w v,
// IRLS weights (|v| + eps)^(p-2), weighted energies, += scale v).  Tiles of P
// pencils are read whole into LDS before they are written, so the passes run
// in place.
// rsh_kernel: the same passes on the register path for lengths 16 R1
// (two-level FFT, sixteen lanes per pencil, one LDS exchange per transform),
// plus the fused middle pass of a band along the last axis: inverse FFT,
// times the weights, forward FFT, without a round trip through HBM.
//   regulariser  (la p/2) IFFT3( sum_w S_w FFT3( We_w IFFT3( conj(S_w) FFT3 x )))
//   = 3 + 5 NW + 3 passes; weights and energies 3 + 3 NW passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/disorder.h"
#include "kernels.hpp"
#include "regfft.hpp"
#include "transform.hpp"

struct dis_shearlet {
  int n[3];
  int nf;      // filters per plane
  int nw;      // bands
  float2* planes[3];
  int4* bands_h;  // host copy: (plane a, filter a, plane b, filter b)
  float* We;      // [nw x nvox] IRLS weights
  float2* Xf;     // FFT3(x)
  float2* t;      // band work volume
  float2* Yf;     // band accumulator
  double* part;   // energy partials
  double* red;    // reduced energies [max(nw, 64)]
  long long nvox;
  long long npart;
  bool weights_set;
};

namespace dis {

typedef dis_shearlet Shearlet;

enum LoadOp { LD_PLAIN = 0, LD_CONJ_BAND = 1, LD_MUL_W = 2, LD_MUL_COIL = 3 };
enum StoreOp { ST_PLAIN = 0, ST_ACC_BAND = 1, ST_WEIGHTS = 2, ST_ENERGY = 3, ST_ACC_SCALED = 4, ST_ENERGY_GRID = 5 };
enum PassKind { PK_FWD = 0, PK_INV = 1, PK_MID = 2 };  // MID: inverse, x W, forward

namespace {

struct AxisArgs {
  const float2* in;
  float2* out;
  long long L, H;
  int n, P;
  long long tiles_lo;
  FftPlan pl;
  const float2* tw;
  float scale;
  int n1, n2, n3;
  const float2* pl0;
  const float2* pl1;
  const float2* pl2;
  int4 band;
  const float* W;  // weights (band volume) or grid weights
  float* Wout;
  float eps, pexp;
  const float2* coil;
  double* part;
  int axis;
};

__device__ __forceinline__ float2 plane_value(const AxisArgs& a, int p, int i, int k1, int k2, int k3) {
  if (p == 0) return a.pl0[k1 + (long long)a.n1 * (k2 + (long long)a.n2 * i)];
  if (p == 1) return a.pl1[k1 + (long long)a.n1 * (k3 + (long long)a.n3 * i)];
  return a.pl2[k3 + (long long)a.n3 * (k2 + (long long)a.n2 * i)];
}

__device__ __forceinline__ float2 band_value(const AxisArgs& a, long long v) {
  const int k1 = (int)(v % a.n1);
  const long long t = v / a.n1;
  const int k2 = (int)(t % a.n2);
  const int k3 = (int)(t / a.n2);
  return cmul(plane_value(a, a.band.x, a.band.y, k1, k2, k3), plane_value(a, a.band.z, a.band.w, k1, k2, k3));
}

template <int LOP>
__device__ __forceinline__ float2 load_op(const AxisArgs& a, long long g) {
  float2 v = a.in[g];
  if (LOP == LD_CONJ_BAND) v = cmulc(v, band_value(a, g));
  if (LOP == LD_MUL_W) v = cscale(v, a.W[g]);
  if (LOP == LD_MUL_COIL) v = cmul(v, a.coil[g]);
  return v;
}

template <int SOP>
__device__ __forceinline__ void store_op(const AxisArgs& a, long long g, float2 v, double& acc) {
  if (SOP == ST_PLAIN) a.out[g] = v;
  if (SOP == ST_ACC_BAND) a.out[g] = cadd(a.out[g], cmul(band_value(a, g), v));
  if (SOP == ST_ACC_SCALED) a.out[g] = cadd(a.out[g], v);
  if (SOP == ST_WEIGHTS) a.Wout[g] = powf(sqrtf(v.x * v.x + v.y * v.y) + a.eps, a.pexp);
  if (SOP == ST_ENERGY) acc += (double)a.W[g] * ((double)v.x * v.x + (double)v.y * v.y);
  if (SOP == ST_ENERGY_GRID) acc += (double)a.W[g % ((long long)a.n1 * a.n2)] * ((double)v.x * v.x + (double)v.y * v.y);
}

template <int SOP>
__device__ __forceinline__ void block_energy(const AxisArgs& a, double acc, double* red, int tid) {
  if (SOP == ST_ENERGY || SOP == ST_ENERGY_GRID) {
    red[tid] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.part[blockIdx.x] = red[0];
  }
}

// Register path, N = 16 R1.  CONTIG (axis 0): pencil q = 16 blockIdx + p,
// lanes 16 p + t.  Strided: sixteen consecutive positions lo of the
// contiguous index space per tile, lanes p + 16 t.
template <int R1, bool CONTIG, int KIND, int LOP, int SOP>
__global__ __launch_bounds__(kThreads) void rsh_kernel(AxisArgs a) {
  constexpr int N = R1 * 16;
  constexpr int XB = CONTIG ? RegFftSize<R1>::XB : RegFftSize<R1>::XB_STRIDED;
  __shared__ fvec2 xraw[16 * XB];
  __shared__ fvec2 traw[N];
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  const int p = CONTIG ? (tid >> 4) : (tid & 15);
  const int t = CONTIG ? (tid & 15) : (tid >> 4);
  lds_float2* tw = (lds_float2*)traw;
  lds_float2* xb = (lds_float2*)xraw + p * XB;
  for (int k = tid; k < N; k += kThreads) st2(tw + k, a.tw[k]);
  __syncthreads();
  long long hi0, lo0;
  int Pv;
  if (CONTIG) {
    hi0 = (long long)blockIdx.x * 16;
    lo0 = 0;
    Pv = (int)(a.H - hi0 < 16 ? a.H - hi0 : 16);
  } else {
    hi0 = blockIdx.x / a.tiles_lo;
    lo0 = (blockIdx.x - hi0 * a.tiles_lo) * 16LL;
    Pv = (int)(a.L - lo0 < 16 ? a.L - lo0 : 16);
  }
  const bool live = p < Pv;
  auto addr = [&](int j) -> long long {
    return CONTIG ? (long long)j + (long long)N * (hi0 + p) : (lo0 + p) + a.L * ((long long)j + (long long)N * hi0);
  };
  // this lane's fixed coordinates; the pencil index j supplies the third
  int ca = 0, cb = 0;
  if (a.axis == 0) {
    const int q = (int)(hi0 + p);
    ca = q % a.n2;   // k2
    cb = q / a.n2;   // k3
  } else if (a.axis == 1) {
    ca = (int)(lo0 + p);  // k1
    cb = (int)hi0;        // k3
  } else {
    const int lo = (int)(lo0 + p);
    ca = lo % a.n1;  // k1
    cb = lo / a.n1;  // k2
  }
  auto kc = [&](int j, int& k1, int& k2, int& k3) {
    if (a.axis == 0) { k1 = j; k2 = ca; k3 = cb; }
    else if (a.axis == 1) { k1 = ca; k2 = j; k3 = cb; }
    else { k1 = ca; k2 = cb; k3 = j; }
  };
  auto pv = [&](int pl, int i, int k1, int k2, int k3) -> float2 {
    if (pl == 0) return a.pl0[k1 + a.n1 * (k2 + a.n2 * i)];
    if (pl == 1) return a.pl1[k1 + a.n1 * (k3 + a.n3 * i)];
    return a.pl2[k3 + a.n3 * (k2 + a.n2 * i)];
  };
  auto bandv = [&](int j) -> float2 {
    int k1, k2, k3;
    kc(j, k1, k2, k3);
    return cmul(pv(a.band.x, a.band.y, k1, k2, k3), pv(a.band.z, a.band.w, k1, k2, k3));
  };
  auto ld = [&](int j) -> float2 {
    const long long g = addr(j);
    float2 x = a.in[g];
    if (LOP == LD_CONJ_BAND) x = cmulc(x, bandv(j));
    if (LOP == LD_MUL_W) x = cscale(x, a.W[g]);
    if (LOP == LD_MUL_COIL) x = cmul(x, a.coil[g]);
    return x;
  };
  auto st = [&](int j, float2 y, double& acc) {
    const long long g = addr(j);
    if (SOP == ST_PLAIN) a.out[g] = y;
    if (SOP == ST_ACC_BAND) a.out[g] = cadd(a.out[g], cmul(bandv(j), y));
    if (SOP == ST_ACC_SCALED) a.out[g] = cadd(a.out[g], y);
    if (SOP == ST_WEIGHTS) a.Wout[g] = powf(sqrtf(y.x * y.x + y.y * y.y) + a.eps, a.pexp);
    if (SOP == ST_ENERGY) acc += (double)a.W[g] * ((double)y.x * y.x + (double)y.y * y.y);
    if (SOP == ST_ENERGY_GRID) {
      int k1, k2, k3;
      kc(j, k1, k2, k3);
      acc += (double)a.W[k1 + a.n1 * k2] * ((double)y.x * y.x + (double)y.y * y.y);
    }
  };
  float2 v[16];
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    float2 x = live ? ld(t + 16 * r) : make_float2(0.f, 0.f);
    v[r] = KIND == PK_FWD ? x : cswap(x);
  }
  reg_fft<R1, 16>(v, t, xb, tw);
  double acc = 0.0;
  if (KIND == PK_MID) {
    const float inv = 1.0f / (float)N;
    if (t < R1) {
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) {
        const float w = live ? a.W[addr(t + R1 * k2)] : 0.f;
        v[k2] = cscale(cswap(v[k2]), inv * w);
      }
    }
    reg_fft<16, R1>(v, t, xb, tw);
    if (live) {
#pragma unroll
      for (int r = 0; r < R1; ++r) st(t + 16 * r, cscale(v[r], a.scale), acc);
    }
  } else if (t < R1 && live) {
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
      float2 y = KIND == PK_INV ? cswap(v[k2]) : v[k2];
      st(t + R1 * k2, cscale(y, a.scale), acc);
    }
  }
  block_energy<SOP>(a, acc, red, tid);
}

template <bool CONTIG, bool INV, int LOP, int SOP>
__global__ __launch_bounds__(kThreads) void fft_axis_kernel(AxisArgs a) {
  __shared__ float2 buf[kMaxTile];
  __shared__ float2 tw[1024];
  __shared__ double red[kThreads];
  const int tid = threadIdx.x;
  for (int k = tid; k < a.n; k += kThreads) tw[k] = a.tw[k];
  const int P = a.P;
  long long hi0, lo0;
  int Pv;
  if (CONTIG) {
    hi0 = (long long)blockIdx.x * P;
    lo0 = 0;
    Pv = (int)(a.H - hi0 < P ? a.H - hi0 : P);
  } else {
    hi0 = blockIdx.x / a.tiles_lo;
    lo0 = (blockIdx.x - hi0 * a.tiles_lo) * (long long)P;
    Pv = (int)(a.L - lo0 < P ? a.L - lo0 : P);
  }
  auto addr = [&](int p, int j) -> long long {
    return CONTIG ? (long long)j + (long long)a.n * (hi0 + p) : (lo0 + p) + a.L * ((long long)j + (long long)a.n * hi0);
  };
  const int tot = P * a.n;
  for (int e = tid; e < tot; e += kThreads) {
    int p, j;
    if (CONTIG) { p = e / a.n; j = e - p * a.n; } else { j = e / P; p = e - j * P; }
    float2 v = make_float2(0.f, 0.f);
    if (p < Pv) {
      const long long g = addr(p, j);
      v = a.in[g];
      if (LOP == LD_CONJ_BAND) v = cmulc(v, band_value(a, g));
      if (LOP == LD_MUL_W) v = cscale(v, a.W[g]);
      if (LOP == LD_MUL_COIL) v = cmul(v, a.coil[g]);
    }
    buf[e] = INV ? cswap(v) : v;
  }
  __syncthreads();
  fft_tile<CONTIG>(buf, tw, a.pl, P, tid);
  double acc = 0.0;
  for (int e = tid; e < tot; e += kThreads) {
    int p, j;
    if (CONTIG) { p = e / a.n; j = e - p * a.n; } else { j = e / P; p = e - j * P; }
    if (p >= Pv) continue;
    float2 v = buf[e];
    if (INV) v = cswap(v);
    v = cscale(v, a.scale);
    const long long g = addr(p, j);
    if (SOP == ST_PLAIN) a.out[g] = v;
    if (SOP == ST_ACC_BAND) a.out[g] = cadd(a.out[g], cmul(band_value(a, g), v));
    if (SOP == ST_ACC_SCALED) a.out[g] = cadd(a.out[g], v);
    if (SOP == ST_WEIGHTS) a.Wout[g] = powf(sqrtf(v.x * v.x + v.y * v.y) + a.eps, a.pexp);
    if (SOP == ST_ENERGY) acc += (double)a.W[g] * ((double)v.x * v.x + (double)v.y * v.y);
    if (SOP == ST_ENERGY_GRID) acc += (double)a.W[g % ((long long)a.n1 * a.n2)] * ((double)v.x * v.x + (double)v.y * v.y);
  }
  if (SOP == ST_ENERGY || SOP == ST_ENERGY_GRID) {
    red[tid] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) a.part[blockIdx.x] = red[0];
  }
}

// fixed-order sum of partials into out[0]
__global__ void reduce_partials_kernel(const double* __restrict__ part, long long n, double* __restrict__ out) {
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

__global__ void zero_kernel(float2* __restrict__ x, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = make_float2(0.f, 0.f);
}

struct Pass {
  const float2* in;
  float2* out;
  int axis;
  bool inv;
  bool mid;  // inverse, x W, forward in one register-path pass
  int lop, sop;
  float scale;
  int4 band;
  const float* W;
  float* Wout;
  const float2* coil;
  float eps, pexp;
};

double pass_bytes(const Pass& ps) {
  // algorithmic HBM bytes per voxel: the input, the load operand, the weights of a fused pass, the store
  double bpv = 8.0 + (ps.lop == LD_MUL_W ? 4.0 : ps.lop == LD_MUL_COIL ? 8.0 : 0.0) + (ps.mid ? 4.0 : 0.0);
  bpv += ps.sop == ST_PLAIN ? 8.0 : (ps.sop == ST_ACC_BAND || ps.sop == ST_ACC_SCALED) ? 16.0
         : ps.sop == ST_WEIGHTS ? 4.0 : ps.sop == ST_ENERGY ? 4.0 : 0.0;
  return bpv;
}

template <int R1, bool CONTIG>
int launch_rsh(const Pass& ps, AxisArgs& a, long long tiles, hipStream_t s) {
  const int kind = ps.mid ? PK_MID : ps.inv ? PK_INV : PK_FWD;
  const int key = kind | (ps.lop << 2) | (ps.sop << 4);
#define DIS_RK(K, LO, SO) \
  case (K | (LO << 2) | (SO << 4)): \
    hipLaunchKernelGGL((rsh_kernel<R1, CONTIG, K, LO, SO>), dim3((unsigned)tiles), dim3(kThreads), 0, s, a); \
    return DIS_OK;
  if (CONTIG) {
    switch (key) {
      DIS_RK(PK_INV, LD_CONJ_BAND, ST_PLAIN)
      DIS_RK(PK_FWD, LD_PLAIN, ST_ACC_BAND)
      DIS_RK(PK_FWD, LD_PLAIN, ST_PLAIN)
      DIS_RK(PK_INV, LD_PLAIN, ST_PLAIN)
      DIS_RK(PK_FWD, LD_MUL_COIL, ST_PLAIN)
      DIS_RK(PK_FWD, LD_MUL_W, ST_PLAIN)
      default: return DIS_E_UNSUPPORTED;
    }
  } else {
    switch (key) {
      DIS_RK(PK_INV, LD_PLAIN, ST_PLAIN)
      DIS_RK(PK_FWD, LD_PLAIN, ST_PLAIN)
      DIS_RK(PK_FWD, LD_PLAIN, ST_ENERGY_GRID)
      DIS_RK(PK_MID, LD_PLAIN, ST_PLAIN)
      DIS_RK(PK_INV, LD_PLAIN, ST_WEIGHTS)
      DIS_RK(PK_INV, LD_PLAIN, ST_ENERGY)
      DIS_RK(PK_INV, LD_PLAIN, ST_ACC_SCALED)
      DIS_RK(PK_FWD, LD_PLAIN, ST_ACC_BAND)
      default: return DIS_E_UNSUPPORTED;
    }
  }
#undef DIS_RK
}

int run_rsh(dis_ctx* c, Shearlet* sh, const Pass& ps, AxisArgs& a, int R1, bool contig, long long* ntiles_out) {
  long long tiles;
  if (contig) {
    tiles = (a.H + 15) / 16;
    a.tiles_lo = 1;
  } else {
    a.tiles_lo = (a.L + 15) / 16;
    tiles = a.tiles_lo * a.H;
  }
  a.P = 16;
  if (tiles > sh->npart) return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: partial buffer too small");
  if (ntiles_out) *ntiles_out = tiles;
  hipStream_t s = ctx_stream(c);
  const int prof = prof_begin(c, DIS_K_SHEAR);
  int st = DIS_E_UNSUPPORTED;
#define DIS_RR(R)                                                                     \
  case R:                                                                             \
    st = contig ? launch_rsh<R, true>(ps, a, tiles, s) : launch_rsh<R, false>(ps, a, tiles, s); \
    break;
  switch (R1) {
    DIS_RR(2)
    DIS_RR(4)
    DIS_RR(3)
    DIS_RR(6)
    DIS_RR(8)
    DIS_RR(9)
    DIS_RR(10)
    DIS_RR(12)
    DIS_RR(14)
    DIS_RR(16)
    default: break;
  }
#undef DIS_RR
  prof_end(c, prof, pass_bytes(ps) * (double)sh->nvox);
  if (st) return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: unsupported register pass combination");
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "shearlet register pass launch failed");
  return DIS_OK;
}

#define DIS_AX(C, I, LO, SO) hipLaunchKernelGGL((fft_axis_kernel<C, I, LO, SO>), dim3((unsigned)tiles), dim3(kThreads), 0, s, a)

int run_pass(dis_ctx* c, Shearlet* sh, const Pass& ps, long long* ntiles_out) {
  AxisArgs a{};
  const int* n = sh->n;
  const int ax = ps.axis;
  a.in = ps.in;
  a.out = ps.out;
  a.n = n[ax];
  a.L = ax == 0 ? 1 : ax == 1 ? n[0] : (long long)n[0] * n[1];
  a.H = ax == 0 ? (long long)n[1] * n[2] : ax == 1 ? n[2] : 1;
  if (make_plan(a.n, &a.pl)) return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: length above 1024");
  int st = ctx_twiddles(c, a.n, &a.tw);
  if (st) return st;
  a.scale = ps.scale;
  a.n1 = n[0];
  a.n2 = n[1];
  a.n3 = n[2];
  a.pl0 = sh->planes[0];
  a.pl1 = sh->planes[1];
  a.pl2 = sh->planes[2];
  a.axis = ax;
  a.band = ps.band;
  a.W = ps.W;
  a.Wout = ps.Wout;
  a.coil = ps.coil;
  a.eps = ps.eps;
  a.pexp = ps.pexp;
  a.part = sh->part;
  const bool contig = ax == 0;
  const int R1 = reg_radix(a.n);
  if (R1) return run_rsh(c, sh, ps, a, R1, contig, ntiles_out);
  if (ps.mid) return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: fused pass needs a register-path length");
  int P = std::max(1, kMaxTile / a.n);
  long long tiles;
  if (contig) {
    P = (int)std::min<long long>(P, std::max<long long>(1, a.H));
    tiles = (a.H + P - 1) / P;
    a.tiles_lo = 1;
  } else {
    P = (int)std::min<long long>(P, a.L);
    a.tiles_lo = (a.L + P - 1) / P;
    tiles = a.tiles_lo * a.H;
  }
  a.P = P;
  if (tiles > sh->npart) return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: partial buffer too small");
  if (ntiles_out) *ntiles_out = tiles;
  hipStream_t s = ctx_stream(c);
  const int key = (ps.inv ? 1 : 0) | (ps.lop << 1) | (ps.sop << 3);
  const double bpv = pass_bytes(ps);
  const int prof = prof_begin(c, DIS_K_SHEAR);
  // the (direction, load, store) combinations the solver uses
#define DIS_AXC(I, LO, SO) \
  do {                     \
    if (contig)            \
      DIS_AX(true, I, LO, SO); \
    else                   \
      DIS_AX(false, I, LO, SO); \
  } while (0)
  switch (key) {
    case (0 | (LD_PLAIN << 1) | (ST_PLAIN << 3)): DIS_AXC(false, LD_PLAIN, ST_PLAIN); break;
    case (1 | (LD_PLAIN << 1) | (ST_PLAIN << 3)): DIS_AXC(true, LD_PLAIN, ST_PLAIN); break;
    case (1 | (LD_CONJ_BAND << 1) | (ST_PLAIN << 3)): DIS_AXC(true, LD_CONJ_BAND, ST_PLAIN); break;
    case (0 | (LD_MUL_W << 1) | (ST_PLAIN << 3)): DIS_AXC(false, LD_MUL_W, ST_PLAIN); break;
    case (0 | (LD_PLAIN << 1) | (ST_ACC_BAND << 3)): DIS_AXC(false, LD_PLAIN, ST_ACC_BAND); break;
    case (1 | (LD_PLAIN << 1) | (ST_WEIGHTS << 3)): DIS_AXC(true, LD_PLAIN, ST_WEIGHTS); break;
    case (1 | (LD_PLAIN << 1) | (ST_ENERGY << 3)): DIS_AXC(true, LD_PLAIN, ST_ENERGY); break;
    case (1 | (LD_PLAIN << 1) | (ST_ACC_SCALED << 3)): DIS_AXC(true, LD_PLAIN, ST_ACC_SCALED); break;
    case (0 | (LD_MUL_COIL << 1) | (ST_PLAIN << 3)): DIS_AXC(false, LD_MUL_COIL, ST_PLAIN); break;
    case (0 | (LD_PLAIN << 1) | (ST_ENERGY_GRID << 3)): DIS_AXC(false, LD_PLAIN, ST_ENERGY_GRID); break;
    default: return ctx_fail(c, DIS_E_UNSUPPORTED, "shearlet: unsupported pass combination");
  }
#undef DIS_AXC
  prof_end(c, prof, bpv * (double)sh->nvox);
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "shearlet pass launch failed");
  return DIS_OK;
}

Pass mk(const float2* in, float2* out, int axis, bool inv, int lop = LD_PLAIN, int sop = ST_PLAIN, float scale = 1.f) {
  Pass p{};
  p.in = in;
  p.out = out;
  p.axis = axis;
  p.inv = inv;
  p.lop = lop;
  p.sop = sop;
  p.scale = scale;
  return p;
}

// X^ = FFT3(x) into sh->Xf
int forward3(dis_ctx* c, Shearlet* sh, const float2* x) {
  int st = run_pass(c, sh, mk(x, sh->Xf, 0, false), nullptr);
  if (!st) st = run_pass(c, sh, mk(sh->Xf, sh->Xf, 1, false), nullptr);
  if (!st) st = run_pass(c, sh, mk(sh->Xf, sh->Xf, 2, false), nullptr);
  return st;
}

// t = IFFT3(conj(S_w) X^) with the last axis finished by `last` (store op)
int band_coefficients(dis_ctx* c, Shearlet* sh, int w, Pass last, long long* ntiles) {
  const int4 b = sh->bands_h[w];
  Pass p0 = mk(sh->Xf, sh->t, 0, true, LD_CONJ_BAND, ST_PLAIN, 1.f / sh->n[0]);
  p0.band = b;
  int st = run_pass(c, sh, p0, nullptr);
  if (!st) st = run_pass(c, sh, mk(sh->t, sh->t, 1, true, LD_PLAIN, ST_PLAIN, 1.f / sh->n[1]), nullptr);
  if (st) return st;
  last.in = sh->t;
  last.axis = 2;
  last.inv = true;
  last.lop = LD_PLAIN;
  last.scale = 1.f / sh->n[2];
  last.band = b;
  return run_pass(c, sh, last, ntiles);
}

int reduce_into(dis_ctx* c, Shearlet* sh, long long ntiles, double* dst) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kThreads), 0, ctx_stream(c), sh->part, ntiles, dst);
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "shearlet reduce launch failed");
  return DIS_OK;
}

}  // namespace

int shearlet_weights(dis_ctx* c, Shearlet* sh, const float2* x, double eps, double p) {
  int st = forward3(c, sh, x);
  for (int w = 0; w < sh->nw && !st; ++w) {
    Pass last{};
    last.out = sh->t;
    last.sop = ST_WEIGHTS;
    last.Wout = sh->We + (long long)w * sh->nvox;
    last.eps = (float)eps;
    last.pexp = (float)(p - 2.0);
    st = band_coefficients(c, sh, w, last, nullptr);
  }
  if (!st) sh->weights_set = true;
  return st;
}

int shearlet_apply(dis_ctx* c, Shearlet* sh, const float2* x, double scale, float2* out) {
  if (!sh->weights_set) return ctx_fail(c, DIS_E_INVALID, "shearlet: weights not computed");
  hipStream_t s = ctx_stream(c);
  int st = forward3(c, sh, x);
  if (st) return st;
  hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, s, sh->Yf, sh->nvox);
  for (int w = 0; w < sh->nw && !st; ++w) {
    const float* Ww = sh->We + (long long)w * sh->nvox;
    if (reg_radix(sh->n[2])) {
      // IFFT axis 0 (x conj(S_w)), IFFT axis 1, [IFFT axis 2, x We_w, FFT axis 2], FFT axis 1
      Pass p0 = mk(sh->Xf, sh->t, 0, true, LD_CONJ_BAND, ST_PLAIN, 1.f / sh->n[0]);
      p0.band = sh->bands_h[w];
      st = run_pass(c, sh, p0, nullptr);
      if (!st) st = run_pass(c, sh, mk(sh->t, sh->t, 1, true, LD_PLAIN, ST_PLAIN, 1.f / sh->n[1]), nullptr);
      Pass pm = mk(sh->t, sh->t, 2, true, LD_PLAIN, ST_PLAIN, 1.f);
      pm.mid = true;
      pm.W = Ww;
      if (!st) st = run_pass(c, sh, pm, nullptr);
      if (!st) st = run_pass(c, sh, mk(sh->t, sh->t, 1, false), nullptr);
      if (st) return st;
      Pass f0 = mk(sh->t, sh->Yf, 0, false, LD_PLAIN, ST_ACC_BAND, 1.f);
      f0.band = sh->bands_h[w];
      st = run_pass(c, sh, f0, nullptr);
      continue;
    }
    Pass last{};
    last.out = sh->t;
    last.sop = ST_PLAIN;
    st = band_coefficients(c, sh, w, last, nullptr);
    if (st) return st;
    Pass f0 = mk(sh->t, sh->t, 0, false, LD_MUL_W, ST_PLAIN, 1.f);
    f0.W = Ww;
    st = run_pass(c, sh, f0, nullptr);
    if (!st) st = run_pass(c, sh, mk(sh->t, sh->t, 1, false), nullptr);
    if (st) return st;
    Pass f2 = mk(sh->t, sh->Yf, 2, false, LD_PLAIN, ST_ACC_BAND, 1.f);
    f2.band = sh->bands_h[w];
    st = run_pass(c, sh, f2, nullptr);
  }
  if (st) return st;
  st = run_pass(c, sh, mk(sh->Yf, sh->Yf, 0, true, LD_PLAIN, ST_PLAIN, 1.f / sh->n[0]), nullptr);
  if (!st) st = run_pass(c, sh, mk(sh->Yf, sh->Yf, 1, true, LD_PLAIN, ST_PLAIN, 1.f / sh->n[1]), nullptr);
  if (!st) st = run_pass(c, sh, mk(sh->Yf, out, 2, true, LD_PLAIN, ST_ACC_SCALED, (float)(scale / sh->n[2])), nullptr);
  return st;
}

// sum_w sum_v We_w |IFFT3(conj(S_w) FFT3 x)|^2 (per band into red[w]) -> host total
int shearlet_energy(dis_ctx* c, Shearlet* sh, const float2* x, double* total) {
  if (!sh->weights_set) return ctx_fail(c, DIS_E_INVALID, "shearlet: weights not computed");
  int st = forward3(c, sh, x);
  for (int w = 0; w < sh->nw && !st; ++w) {
    Pass last{};
    last.out = sh->t;
    last.sop = ST_ENERGY;
    last.W = sh->We + (long long)w * sh->nvox;
    long long nt = 0;
    st = band_coefficients(c, sh, w, last, &nt);
    if (!st) st = reduce_into(c, sh, nt, sh->red + w);
  }
  if (st) return st;
  std::vector<double> h(sh->nw);
  hipStream_t s = ctx_stream(c);
  if (hipMemcpyAsync(h.data(), sh->red, sizeof(double) * sh->nw, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "shearlet energy readback failed");
  double t = 0;
  for (double v : h) t += v;
  *total = t;
  return DIS_OK;
}

// sum_c sum_k Wgrid(k1,k2) |F2(S_c x)/sqrt(n1 n2)|^2
int fidelity_trace(dis_ctx* c, Shearlet* sh, const float2* Smaps, int C, const float* Wgrid, const float2* x,
                   double* total) {
  int st = DIS_OK;
  const float s2 = 1.0f / std::sqrt((float)sh->n[0] * (float)sh->n[1]);
  for (int cc = 0; cc < C && !st; ++cc) {
    Pass p0 = mk(x, sh->t, 0, false, LD_MUL_COIL, ST_PLAIN, 1.f);
    p0.coil = Smaps + (long long)cc * sh->nvox;
    st = run_pass(c, sh, p0, nullptr);
    if (st) return st;
    Pass p1 = mk(sh->t, sh->t, 1, false, LD_PLAIN, ST_ENERGY_GRID, s2);
    p1.W = Wgrid;
    long long nt = 0;
    st = run_pass(c, sh, p1, &nt);
    if (!st) st = reduce_into(c, sh, nt, sh->red + cc);
  }
  if (st) return st;
  std::vector<double> h(C);
  hipStream_t s = ctx_stream(c);
  if (hipMemcpyAsync(h.data(), sh->red, sizeof(double) * C, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "fidelity trace readback failed");
  double t = 0;
  for (double v : h) t += v;
  *total = t;
  return DIS_OK;
}

}  // namespace dis

using namespace dis;

extern "C" {

int dis_shearlet_create(dis_ctx* c, const int32_t n[3], const void* p12, const void* p13, const void* p32,
                        int32_t nf, const int32_t* bands, int32_t nw, dis_shearlet** out) {
  if (!c) return DIS_E_INVALID;
  if (!n || !p12 || !p13 || !p32 || !bands || !out || nf <= 0 || nw <= 0 || nw > 4096)
    return ctx_fail(c, DIS_E_INVALID, "dis_shearlet_create: invalid argument");
  for (int d = 0; d < 3; ++d) {
    FftPlan pl;
    if (n[d] <= 0 || make_plan(n[d], &pl)) return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_shearlet_create: unsupported size");
  }
  for (int w = 0; w < nw; ++w)
    for (int q = 0; q < 2; ++q) {
      const int pl = bands[4 * w + 2 * q], fi = bands[4 * w + 2 * q + 1];
      if (pl < 0 || pl > 2 || fi < 0 || fi >= nf) return ctx_fail(c, DIS_E_INVALID, "dis_shearlet_create: bad band table");
    }
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  Shearlet* sh = new Shearlet{};
  for (int d = 0; d < 3; ++d) sh->n[d] = n[d];
  sh->nf = nf;
  sh->nw = nw;
  sh->nvox = (long long)n[0] * n[1] * n[2];
  long long maxt = 0;
  for (int ax = 0; ax < 3; ++ax) {
    const long long L = ax == 0 ? 1 : ax == 1 ? n[0] : (long long)n[0] * n[1];
    const long long H = ax == 0 ? (long long)n[1] * n[2] : ax == 1 ? n[2] : 1;
    for (int P0 : {std::max(1, kMaxTile / n[ax]), 16}) {  // LDS-engine and register-path tilings
      long long t = ax == 0 ? (H + std::min<long long>(P0, H) - 1) / std::min<long long>(P0, H)
                            : ((L + std::min<long long>(P0, L) - 1) / std::min<long long>(P0, L)) * H;
      maxt = std::max(maxt, t);
    }
  }
  sh->npart = maxt;
  const size_t ps[3] = {(size_t)n[0] * n[1] * nf, (size_t)n[0] * n[2] * nf, (size_t)n[2] * n[1] * nf};
  const void* src[3] = {p12, p13, p32};
  bool ok = true;
  for (int d = 0; d < 3 && ok; ++d) {
    ok = hipMalloc(&sh->planes[d], sizeof(float2) * ps[d]) == hipSuccess &&
         hipMemcpy(sh->planes[d], src[d], sizeof(float2) * ps[d], hipMemcpyHostToDevice) == hipSuccess;
  }
  ok = ok && hipMalloc(&sh->We, sizeof(float) * sh->nvox * nw) == hipSuccess;
  ok = ok && hipMalloc(&sh->Xf, sizeof(float2) * sh->nvox) == hipSuccess;
  ok = ok && hipMalloc(&sh->t, sizeof(float2) * sh->nvox) == hipSuccess;
  ok = ok && hipMalloc(&sh->Yf, sizeof(float2) * sh->nvox) == hipSuccess;
  ok = ok && hipMalloc(&sh->part, sizeof(double) * sh->npart) == hipSuccess;
  ok = ok && hipMalloc(&sh->red, sizeof(double) * std::max(nw, 4096)) == hipSuccess;
  sh->bands_h = new int4[nw];
  for (int w = 0; w < nw; ++w)
    sh->bands_h[w] = make_int4(bands[4 * w], bands[4 * w + 1], bands[4 * w + 2], bands[4 * w + 3]);
  if (!ok) {
    (void)dis_shearlet_destroy(sh);
    return ctx_fail(c, DIS_E_NOMEM, "dis_shearlet_create: device allocation failed");
  }
  *out = sh;
  return DIS_OK;
}

int dis_shearlet_destroy(dis_shearlet* sh) {
  if (!sh) return DIS_OK;
  for (int d = 0; d < 3; ++d)
    if (sh->planes[d]) (void)hipFree(sh->planes[d]);
  if (sh->We) (void)hipFree(sh->We);
  if (sh->Xf) (void)hipFree(sh->Xf);
  if (sh->t) (void)hipFree(sh->t);
  if (sh->Yf) (void)hipFree(sh->Yf);
  if (sh->part) (void)hipFree(sh->part);
  if (sh->red) (void)hipFree(sh->red);
  delete[] sh->bands_h;
  delete sh;
  return DIS_OK;
}

int dis_shearlet_weights(dis_ctx* c, dis_shearlet* sh, const void* x, double epsil, double p) {
  if (!c) return DIS_E_INVALID;
  if (!sh || !x) return ctx_fail(c, DIS_E_INVALID, "dis_shearlet_weights: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  int st = shearlet_weights(c, sh, (const float2*)x, epsil, p);
  if (st) return st;
  return hipStreamSynchronize(ctx_stream(c)) == hipSuccess ? DIS_OK : ctx_fail(c, DIS_E_HIP, "sync failed");
}

int dis_shearlet_regularize(dis_ctx* c, dis_shearlet* sh, const void* x, double scale, void* out) {
  if (!c) return DIS_E_INVALID;
  if (!sh || !x || !out) return ctx_fail(c, DIS_E_INVALID, "dis_shearlet_regularize: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  int st = shearlet_apply(c, sh, (const float2*)x, scale, (float2*)out);
  if (st) return st;
  return hipStreamSynchronize(ctx_stream(c)) == hipSuccess ? DIS_OK : ctx_fail(c, DIS_E_HIP, "sync failed");
}

int dis_shearlet_energy(dis_ctx* c, dis_shearlet* sh, const void* x, double* out) {
  if (!c) return DIS_E_INVALID;
  if (!sh || !x || !out) return ctx_fail(c, DIS_E_INVALID, "dis_shearlet_energy: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  return shearlet_energy(c, sh, (const float2*)x, out);
}

int dis_fidelity_trace_folded(dis_ctx* c, const int32_t n[3], const void* S, int32_t C, int32_t ny1, int32_t ny2,
                              const float* Wgrid, const void* x, double* out) {
  if (!c) return DIS_E_INVALID;
  if (!n || !S || !Wgrid || !x || !out || C <= 0 || C > 4096 || n[0] <= 0 || n[1] <= 0 || n[2] <= 0)
    return ctx_fail(c, DIS_E_INVALID, "dis_fidelity_trace_folded: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  return fold_fidelity_trace(c, n, (const float2*)S, C, ny1, ny2, Wgrid, (const float2*)x, out);
}

int dis_fidelity_trace(dis_ctx* c, dis_shearlet* sh, const void* S, int32_t C, const float* Wgrid, const void* x,
                       double* out) {
  if (!c) return DIS_E_INVALID;
  if (!sh || !S || !Wgrid || !x || !out || C <= 0 || C > 4096)
    return ctx_fail(c, DIS_E_INVALID, "dis_fidelity_trace: invalid argument");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  return fidelity_trace(c, sh, (const float2*)S, C, Wgrid, (const float2*)x, out);
}

}  // extern "C"
