// Fused pencil passes of the sinc rigid transform on gfx950.
//
// Each workgroup owns a tile of P pencils along one axis, stages it in LDS,
// runs FFT -> per-pencil phase ramp -> inverse FFT, and writes it back.  The
// three modes share the tile machinery:
//   MAP   : one state per grid.y, in -> out (in place allowed)
//   BCAST : one input volume, all states looped inside the workgroup; the
//           forward FFT of the input is done once and kept in registers
//   ACCUM : all states looped inside the workgroup and summed in the
//           frequency domain; one inverse FFT per tile (state sum of T^H)
//   FLIP  : the >90 degree flip of Methods/flipping.m along one axis
#include "kernels.hpp"
#include "regfft.hpp"

namespace dis {

constexpr int kE = kMaxTile / kThreads;  // elements per thread per tile
constexpr int kMaxPencils = 512;          // tile pencils (host caps P)

struct TileInfo {
  long long base;
  long long jstride;
  int Pv;   // valid pencils in this tile
  int t_a;  // CONTIG: first pencil index q0; strided: first contiguous index i1_0
  int t_b;  // strided: the other coordinate (i3 for axis 1, i2 for axis 2)
};

__device__ __forceinline__ TileInfo tile_info(const PassArgs& a, int t) {
  TileInfo ti;
  if (a.axis == 0) {
    const int np = a.n2 * a.n3;
    const int q0 = t * a.P;
    ti.Pv = min(a.P, np - q0);
    ti.base = (long long)q0 * a.n1;
    ti.jstride = 1;
    ti.t_a = q0;
    ti.t_b = 0;
  } else {
    const int line = t / a.tiles_per_line;
    const int i10 = (t - line * a.tiles_per_line) * a.P;
    ti.Pv = min(a.P, a.n1 - i10);
    ti.t_a = i10;
    ti.t_b = line;
    if (a.axis == 1) {
      ti.base = (long long)i10 + (long long)a.n1 * a.n2 * line;
      ti.jstride = a.n1;
    } else {
      ti.base = (long long)i10 + (long long)a.n1 * line;
      ti.jstride = (long long)a.n1 * a.n2;
    }
  }
  return ti;
}

// Workgroups are dealt to the eight XCDs round-robin, so without a remap an
// XCD's L2 sees every eighth tile (on strided axes every eighth 128-byte
// column segment).  With the tile count a multiple of 8, XCD x (blocks
// b = x mod 8) instead takes the contiguous range of tiles
// [x ntiles/8, (x+1) ntiles/8).
__device__ __forceinline__ int xcd_tile(const PassArgs& a, int b) {
  const int n = (int)gridDim.x;
  if (n & 7) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

// coordinate of pencil p of the tile along axis `src`
__device__ __forceinline__ int src_coord(const PassArgs& a, const TileInfo& ti, int p) {
  if (a.axis == 0) {
    const int q = ti.t_a + p;
    const int i2 = q % a.n2;
    return a.src == 1 ? i2 : q / a.n2;
  }
  if (a.src == 0) return ti.t_a + p;
  return ti.t_b;  // axis 1 -> src 2 (i3), axis 2 -> src 1 (i2)
}

template <bool CONTIG>
__device__ __forceinline__ void elem_pj(int e, int n, int P, int& p, int& j) {
  if (CONTIG) { p = e / n; j = e - p * n; }
  else { j = e / P; p = e - j * P; }
}

template <bool CONTIG>
__device__ __forceinline__ long long gaddr(const TileInfo& ti, int e, int n, int P) {
  int p, j;
  elem_pj<CONTIG>(e, n, P, p, j);
  return CONTIG ? ti.base + e : ti.base + p + (long long)j * ti.jstride;
}

template <bool CONTIG>
__device__ __forceinline__ void load_tile(float2* lds, const float2* __restrict__ g, const TileInfo& ti, int n,
                                          int P, int tid) {
  const int tot = P * n;
#pragma unroll
  for (int i = 0; i < kE; ++i) {
    const int e = tid + i * kThreads;
    if (e < tot) {
      int p, j;
      elem_pj<CONTIG>(e, n, P, p, j);
      lds[e] = (p < ti.Pv) ? g[gaddr<CONTIG>(ti, e, n, P)] : make_float2(0.f, 0.f);
    }
  }
}

// per-pencil phase coefficient in cycles per frequency index
__device__ __forceinline__ void pencil_coefs(double* cP, const PassArgs& a, const TileInfo& ti, int s, int tid) {
  const double A = a.coefA ? a.coefA[s] : 0.0;
  const double B = a.coefB ? a.coefB[s] : 0.0;
  for (int p = tid; p < a.P; p += kThreads) {
    double r = 0.0;
    if (a.src >= 0) r = a.rscale * (double)(src_coord(a, ti, p) + 1 - a.rcent);
    cP[p] = (A * r + B) * a.invN;
  }
}

template <bool CONTIG, int MODE>
__global__ __launch_bounds__(kThreads) void pass_kernel(PassArgs a) {
  __shared__ float2 lds[kMaxTile];
  __shared__ float2 tw[1024];
  __shared__ double cP[kMaxPencils];
  const int tid = threadIdx.x;
  const int n = a.plan.n;
  const int P = a.P;
  const int tot = P * n;
  const float scale = 1.0f / (float)n;
  for (int k = tid; k < n; k += kThreads) tw[k] = a.tw[k];
  const TileInfo ti = tile_info(a, xcd_tile(a, blockIdx.x));

  if (MODE == MODE_MAP) {
    const int s = a.states ? a.states[blockIdx.y] : (int)blockIdx.y;
    pencil_coefs(cP, a, ti, s, tid);
    load_tile<CONTIG>(lds, a.in + a.in_sstride * s, ti, n, P, tid);
    __syncthreads();
    fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        int p, j;
        elem_pj<CONTIG>(e, n, P, p, j);
        lds[e] = cswap(cmul(lds[e], phase_cycles(cP[p], freq_index(j, n))));
      }
    }
    __syncthreads();
    fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
    float2* o = a.out + a.out_sstride * s;
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        int p, j;
        elem_pj<CONTIG>(e, n, P, p, j);
        if (p < ti.Pv) o[gaddr<CONTIG>(ti, e, n, P)] = cscale(cswap(lds[e]), scale);
      }
    }
  } else if (MODE == MODE_BCAST) {
    load_tile<CONTIG>(lds, a.in, ti, n, P, tid);
    __syncthreads();
    fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
    float2 spec[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      spec[i] = (e < tot) ? lds[e] : make_float2(0.f, 0.f);
    }
    for (int s = 0; s < a.S; ++s) {
      pencil_coefs(cP, a, ti, s, tid);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          lds[e] = cswap(cmul(spec[i], phase_cycles(cP[p], freq_index(j, n))));
        }
      }
      __syncthreads();
      fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
      float2* o = a.out + a.out_sstride * s;
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          if (p < ti.Pv) o[gaddr<CONTIG>(ti, e, n, P)] = cscale(cswap(lds[e]), scale);
        }
      }
      __syncthreads();
    }
  } else if (MODE == MODE_ACCUM) {
    float2 acc[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i) acc[i] = make_float2(0.f, 0.f);
    for (int s = 0; s < a.S; ++s) {
      pencil_coefs(cP, a, ti, s, tid);
      load_tile<CONTIG>(lds, a.in + a.in_sstride * s, ti, n, P, tid);
      __syncthreads();
      fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          acc[i] = cadd(acc[i], cmul(lds[e], phase_cycles(cP[p], freq_index(j, n))));
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) lds[e] = cswap(acc[i]);
    }
    __syncthreads();
    fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        int p, j;
        elem_pj<CONTIG>(e, n, P, p, j);
        if (p < ti.Pv) {
          const long long g = gaddr<CONTIG>(ti, e, n, P);
          float2 v = cscale(cswap(lds[e]), scale);
          if (a.accumulate) v = cadd(v, a.out[g]);
          a.out[g] = v;
        }
      }
    }
  } else if (MODE == MODE_DUAL) {
    const int s = a.states ? a.states[blockIdx.y] : (int)blockIdx.y;
    pencil_coefs(cP, a, ti, s, tid);
    __shared__ double cB[kMaxPencils];
    {
      const double A2 = a.coefA2 ? a.coefA2[s] : 0.0;
      const double B2 = a.coefB2 ? a.coefB2[s] : 0.0;
      for (int pp = tid; pp < P; pp += kThreads) {
        double r = 0.0;
        if (a.src >= 0) r = a.rscale * (double)(src_coord(a, ti, pp) + 1 - a.rcent);
        cB[pp] = (A2 * r + B2) * 2.0 * M_PI * a.invN;  // gamma * 2 pi / Nfull: beta = i cB m
      }
    }
    float2 x1[kE];
    load_tile<CONTIG>(lds, a.in + a.in_sstride * s, ti, n, P, tid);
    __syncthreads();
    fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      x1[i] = (e < tot) ? lds[e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 x2[kE];
    if (a.in2) {
      load_tile<CONTIG>(lds, a.in2 + a.in2_sstride * s, ti, n, P, tid);
      __syncthreads();
      fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        x2[i] = (e < tot) ? lds[e] : make_float2(0.f, 0.f);
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int i = 0; i < kE; ++i) x2[i] = make_float2(0.f, 0.f);
    }
    for (int which = 0; which < 2; ++which) {
      float2* dst = which == 0 ? a.out : a.out2;
      if (!dst) continue;
      const long long dss = which == 0 ? a.out_sstride : a.out2_sstride;
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          const int m = freq_index(j, n);
          const float2 ph = phase_cycles(cP[p], m);
          float2 val = x1[i];
          if (which == 1) {
            const float bm = (float)(cB[p] * (double)m);
            val = cadd(make_float2(-bm * val.y, bm * val.x), x2[i]);  // i beta X1 + X2
          }
          lds[e] = cswap(cmul(val, ph));
        }
      }
      __syncthreads();
      fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
      float2* o = dst + dss * s;
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          if (p < ti.Pv) o[gaddr<CONTIG>(ti, e, n, P)] = cscale(cswap(lds[e]), scale);
        }
      }
      __syncthreads();
    }
  } else if (MODE == MODE_FLIP) {
    // Methods/flipping.m:11-22 along this axis: shift by h1=-c+1/2 with
    // c = n/2+1, reverse, shift by -h1.  Odd n gives integer shifts, which
    // Methods/shifting.m:29-33 applies as circular index shifts.
    const int s = a.states ? a.states[blockIdx.y] : (int)blockIdx.y;
    const float2* in = a.in + a.in_sstride * s;
    float2* o = a.out + a.out_sstride * s;
    load_tile<CONTIG>(lds, in, ti, n, P, tid);
    __syncthreads();
    float2 v[kE];
    if (n & 1) {
      const int h1 = -(n + 1) / 2;  // -c + 1/2 with c = n/2 + 1
#pragma unroll
      for (int i = 0; i < kE; ++i) {
        const int e = tid + i * kThreads;
        if (e < tot) {
          int p, j;
          elem_pj<CONTIG>(e, n, P, p, j);
          // x3[j] = x2[(j+h1) mod n], x2[i] = x1[n-1-i], x1[i] = x[(i-h1) mod n]
          const int i2 = ((j + h1) % n + n) % n;
          const int i1 = n - 1 - i2;
          const int i0 = ((i1 - h1) % n + n) % n;
          v[i] = CONTIG ? lds[p * n + i0] : lds[p + P * i0];
        }
      }
    } else {
      const double h1 = -(double)(n / 2 + 1) + 0.5;
      for (int pass = 0; pass < 2; ++pass) {
        const double c = (pass == 0 ? -h1 : h1) / (double)n;  // exp(-i h k), k = 2 pi m / n
        fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
#pragma unroll
        for (int i = 0; i < kE; ++i) {
          const int e = tid + i * kThreads;
          if (e < tot) {
            int p, j;
            elem_pj<CONTIG>(e, n, P, p, j);
            lds[e] = cswap(cmul(lds[e], phase_cycles(c, freq_index(j, n))));
          }
        }
        __syncthreads();
        fft_tile<CONTIG>(lds, tw, a.plan, P, tid);
        if (pass == 0) {
          // back to image space (swap + scale) and reverse along the pencil
#pragma unroll
          for (int i = 0; i < kE; ++i) {
            const int e = tid + i * kThreads;
            if (e < tot) {
              int p, j;
              elem_pj<CONTIG>(e, n, P, p, j);
              const int jr = n - 1 - j;
              v[i] = cscale(cswap(CONTIG ? lds[p * n + jr] : lds[p + P * jr]), scale);
            }
          }
          __syncthreads();
#pragma unroll
          for (int i = 0; i < kE; ++i) {
            const int e = tid + i * kThreads;
            if (e < tot) lds[e] = v[i];
          }
          __syncthreads();
        } else {
#pragma unroll
          for (int i = 0; i < kE; ++i) {
            const int e = tid + i * kThreads;
            if (e < tot) v[i] = cscale(cswap(lds[e]), scale);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kE; ++i) {
      const int e = tid + i * kThreads;
      if (e < tot) {
        int p, j;
        elem_pj<CONTIG>(e, n, P, p, j);
        if (p < ti.Pv) o[gaddr<CONTIG>(ti, e, n, P)] = v[i];
      }
    }
  }
}

hipError_t launch_pass_generic(const PassArgs& a, int mode, int grid_states, hipStream_t st) {
  const int n = a.plan.n;
  int ntiles;
  if (a.axis == 0) ntiles = (a.n2 * a.n3 + a.P - 1) / a.P;
  else ntiles = a.tiles_per_line * (a.axis == 1 ? a.n3 : a.n2);
  (void)n;
  dim3 grid(ntiles, (mode == MODE_MAP || mode == MODE_FLIP || mode == MODE_DUAL) ? grid_states : 1);
  const bool contig = a.axis == 0;
#define DIS_LAUNCH(C, M) hipLaunchKernelGGL((pass_kernel<C, M>), grid, dim3(kThreads), 0, st, a)
  switch (mode) {
    case MODE_MAP: if (contig) DIS_LAUNCH(true, MODE_MAP); else DIS_LAUNCH(false, MODE_MAP); break;
    case MODE_BCAST: if (contig) DIS_LAUNCH(true, MODE_BCAST); else DIS_LAUNCH(false, MODE_BCAST); break;
    case MODE_ACCUM: if (contig) DIS_LAUNCH(true, MODE_ACCUM); else DIS_LAUNCH(false, MODE_ACCUM); break;
    case MODE_FLIP: if (contig) DIS_LAUNCH(true, MODE_FLIP); else DIS_LAUNCH(false, MODE_FLIP); break;
    case MODE_DUAL: if (contig) DIS_LAUNCH(true, MODE_DUAL); else DIS_LAUNCH(false, MODE_DUAL); break;
    default: return hipErrorInvalidValue;
  }
#undef DIS_LAUNCH
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// Register path for N = 16 * R1: sixteen lanes per pencil, sixteen pencils
// per workgroup, data in registers from the global load to the global store,
// one LDS exchange per transform (regfft.hpp).
template <int R1, bool CONTIG>
__device__ __forceinline__ void rload(float2 (&v)[16], const float2* __restrict__ g, const TileInfo& ti, int p,
                                      int t) {
  constexpr int N = R1 * 16;
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    const int j = t + 16 * r;
    v[r] = (p < ti.Pv) ? g[CONTIG ? ti.base + (long long)p * N + j : ti.base + p + (long long)j * ti.jstride]
                       : make_float2(0.f, 0.f);
  }
}

template <int R1, bool CONTIG>
__device__ __forceinline__ void rstore(const float2 (&v)[16], float2* __restrict__ g, const TileInfo& ti, int p,
                                       int t, float scale, bool acc) {
  constexpr int N = R1 * 16;
  if (p >= ti.Pv) return;
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    const int j = t + 16 * r;
    const long long ad = CONTIG ? ti.base + (long long)p * N + j : ti.base + p + (long long)j * ti.jstride;
    float2 o = cscale(cswap(v[r]), scale);
    if (acc) o = cadd(o, g[ad]);
    g[ad] = o;
  }
}

__device__ __forceinline__ double lane_coef(const PassArgs& a, const TileInfo& ti, int s, int p) {
  const double A = a.coefA ? a.coefA[s] : 0.0;
  const double B = a.coefB ? a.coefB[s] : 0.0;
  double r = 0.0;
  if (a.src >= 0) r = a.rscale * (double)(src_coord(a, ti, p) + 1 - a.rcent);
  return (A * r + B) * a.invN;
}

template <int R1, bool CONTIG, int MODE>
__global__ __launch_bounds__(kThreads) void rpass_kernel(PassArgs a) {
  constexpr int N = R1 * 16;
  constexpr int XB = CONTIG ? RegFftSize<R1>::XB : RegFftSize<R1>::XB_STRIDED;
  __shared__ fvec2 xraw[16 * XB];
  __shared__ fvec2 traw[N];
  const int tid = threadIdx.x;
  const int p = CONTIG ? (tid >> 4) : (tid & 15);
  const int t = CONTIG ? (tid & 15) : (tid >> 4);
  lds_float2* tw = (lds_float2*)traw;
  lds_float2* xb = (lds_float2*)xraw + p * XB;
  for (int k = tid; k < N; k += kThreads) st2(tw + k, a.tw[k]);
  __syncthreads();
  const TileInfo ti = tile_info(a, xcd_tile(a, blockIdx.x));
  const float scale = 1.0f / (float)N;
  float2 v[16];
  if (MODE == MODE_MAP) {
    const int s = a.states ? a.states[blockIdx.y] : (int)blockIdx.y;
    const double c = lane_coef(a, ti, s, p);
    rload<R1, CONTIG>(v, a.in + a.in_sstride * s, ti, p, t);
    reg_fft<R1, 16>(v, t, xb, tw);
    if (t < R1) {
      float2 ph[16];
      ramp16<R1>(ph, c, t);
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = cswap(cmul(v[k], ph[k]));
    }
    reg_fft<16, R1>(v, t, xb, tw);
    rstore<R1, CONTIG>(v, a.out + a.out_sstride * s, ti, p, t, scale, false);
  } else if (MODE == MODE_BCAST) {
    rload<R1, CONTIG>(v, a.in, ti, p, t);
    reg_fft<R1, 16>(v, t, xb, tw);
    float2 spec[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) spec[k] = v[k];
    for (int s = 0; s < a.S; ++s) {
      const double c = lane_coef(a, ti, s, p);
      if (t < R1) {
        float2 ph[16];
        ramp16<R1>(ph, c, t);
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = cswap(cmul(spec[k], ph[k]));
      }
      reg_fft<16, R1>(v, t, xb, tw);
      rstore<R1, CONTIG>(v, a.out + a.out_sstride * s, ti, p, t, scale, false);
    }
  } else if (MODE == MODE_DUAL) {
    const int s = a.states ? a.states[blockIdx.y] : (int)blockIdx.y;
    const double c = lane_coef(a, ti, s, p);
    double gam = 0.0;
    {
      const double A2 = a.coefA2 ? a.coefA2[s] : 0.0;
      const double B2 = a.coefB2 ? a.coefB2[s] : 0.0;
      double r = 0.0;
      if (a.src >= 0) r = a.rscale * (double)(src_coord(a, ti, p) + 1 - a.rcent);
      gam = (A2 * r + B2) * 2.0 * M_PI * a.invN;  // beta = i gam m
    }
    float2 x2[16];
    rload<R1, CONTIG>(v, a.in + a.in_sstride * s, ti, p, t);
    reg_fft<R1, 16>(v, t, xb, tw);
    if (a.in2) {
      rload<R1, CONTIG>(x2, a.in2 + a.in2_sstride * s, ti, p, t);
      reg_fft<R1, 16>(x2, t, xb, tw);
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) x2[k] = make_float2(0.f, 0.f);
    }
    float2 ph[16];
    if (t < R1) ramp16<R1>(ph, c, t);
    if (a.out) {
      float2 w[16];
      if (t < R1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = cswap(cmul(v[k], ph[k]));
      }
      reg_fft<16, R1>(w, t, xb, tw);
      rstore<R1, CONTIG>(w, a.out + a.out_sstride * s, ti, p, t, scale, false);
    }
    if (t < R1) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int j = t + R1 * k;
        const int m = j < N / 2 ? j : j - N;
        const float bm = (float)(gam * (double)m);
        const float2 val = cadd(make_float2(-bm * v[k].y, bm * v[k].x), x2[k]);  // i beta X1 + X2
        x2[k] = cswap(cmul(val, ph[k]));
      }
    }
    reg_fft<16, R1>(x2, t, xb, tw);
    rstore<R1, CONTIG>(x2, a.out2 + a.out2_sstride * s, ti, p, t, scale, false);
  } else if (MODE == MODE_ACCUM) {
    float2 acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = make_float2(0.f, 0.f);
    for (int s = 0; s < a.S; ++s) {
      const double c = lane_coef(a, ti, s, p);
      rload<R1, CONTIG>(v, a.in + a.in_sstride * s, ti, p, t);
      reg_fft<R1, 16>(v, t, xb, tw);
      if (t < R1) {
        float2 ph[16];
        ramp16<R1>(ph, c, t);
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] = cadd(acc[k], cmul(v[k], ph[k]));
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = cswap(acc[k]);
    reg_fft<16, R1>(v, t, xb, tw);
    rstore<R1, CONTIG>(v, a.out, ti, p, t, scale, a.accumulate != 0);
  }
}

// R1 of the register path for length n (0 when not available)
int reg_radix(int n) {
  switch (n) {
    case 256: return 16;
    case 224: return 14;
    case 192: return 12;
    case 160: return 10;
    case 144: return 9;
    case 128: return 8;
    case 96: return 6;
    case 64: return 4;
    case 48: return 3;
    case 32: return 2;
    default: return 0;
  }
}

template <int R1>
hipError_t launch_rpass(PassArgs a, int mode, int grid_states, hipStream_t st) {
  a.P = 16;
  a.tiles_per_line = (a.n1 + 15) / 16;
  int ntiles;
  if (a.axis == 0) ntiles = (a.n2 * a.n3 + 15) / 16;
  else ntiles = a.tiles_per_line * (a.axis == 1 ? a.n3 : a.n2);
  dim3 grid(ntiles, (mode == MODE_MAP || mode == MODE_DUAL) ? grid_states : 1);
  const bool contig = a.axis == 0;
#define DIS_RL(C, M) hipLaunchKernelGGL((rpass_kernel<R1, C, M>), grid, dim3(kThreads), 0, st, a)
  switch (mode) {
    case MODE_MAP: if (contig) DIS_RL(true, MODE_MAP); else DIS_RL(false, MODE_MAP); break;
    case MODE_BCAST: if (contig) DIS_RL(true, MODE_BCAST); else DIS_RL(false, MODE_BCAST); break;
    case MODE_ACCUM: if (contig) DIS_RL(true, MODE_ACCUM); else DIS_RL(false, MODE_ACCUM); break;
    case MODE_DUAL: if (contig) DIS_RL(true, MODE_DUAL); else DIS_RL(false, MODE_DUAL); break;
    default: return hipErrorInvalidValue;
  }
#undef DIS_RL
  return hipGetLastError();
}

hipError_t launch_pass(const PassArgs& a, int mode, int grid_states, hipStream_t st) {
  const int n = a.plan.n;
  const int R1 = reg_radix(n);
  if (R1 && mode != MODE_FLIP && (a.axis == 0 || a.n1 >= 16) && !a.force_generic) {
    switch (R1) {
      case 16: return launch_rpass<16>(a, mode, grid_states, st);
      case 14: return launch_rpass<14>(a, mode, grid_states, st);
      case 12: return launch_rpass<12>(a, mode, grid_states, st);
      case 10: return launch_rpass<10>(a, mode, grid_states, st);
      case 9: return launch_rpass<9>(a, mode, grid_states, st);
      case 8: return launch_rpass<8>(a, mode, grid_states, st);
      case 6: return launch_rpass<6>(a, mode, grid_states, st);
      case 4: return launch_rpass<4>(a, mode, grid_states, st);
      case 3: return launch_rpass<3>(a, mode, grid_states, st);
      case 2: return launch_rpass<2>(a, mode, grid_states, st);
    }
  }
  return launch_pass_generic(a, mode, grid_states, st);
}

__global__ void twiddle_kernel(float2* tw, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    tw[k] = make_float2((float)c, (float)s);
  }
}

hipError_t launch_twiddles(float2* tw, int n, hipStream_t st) {
  hipLaunchKernelGGL(twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tw, n);
  return hipGetLastError();
}

__global__ void bcast_kernel(const float2* __restrict__ x, float2* __restrict__ out, long long nvox, int S) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nvox) {
    const float2 v = x[i];
    for (int s = 0; s < S; ++s) out[i + nvox * s] = v;
  }
}

hipError_t launch_bcast(const float2* x, float2* out, long long nvox, int S, hipStream_t st) {
  hipLaunchKernelGGL(bcast_kernel, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, x, out, nvox, S);
  return hipGetLastError();
}

__global__ void sum_states_kernel(const float2* __restrict__ in, long long sstride, float2* __restrict__ out,
                                  long long nvox, int S, int accumulate) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nvox) {
    float2 acc = accumulate ? out[i] : make_float2(0.f, 0.f);
    for (int s = 0; s < S; ++s) acc = cadd(acc, in[i + sstride * s]);
    out[i] = acc;
  }
}

hipError_t launch_sum_states(const float2* in, long long sstride, float2* out, long long nvox, int S, int accumulate,
                             hipStream_t st) {
  hipLaunchKernelGGL(sum_states_kernel, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, in, sstride, out,
                     nvox, S, accumulate);
  return hipGetLastError();
}

// per-sample energy: sum over planes i3lo..i3hi-1 (in fixed order) of the
// per-(sample, plane) coil sums, weighted by the readout mask Mbe (or 1).
__global__ void energy_reduce_kernel(const float* __restrict__ epart, int Ns, int i3lo, int i3hi,
                                     const float* __restrict__ Mbe, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < Ns) {
    double acc = 0.0;
    for (int i3 = i3lo; i3 < i3hi; ++i3) {
      const double w = Mbe ? (double)Mbe[i3] : 1.0;
      acc += w * (double)epart[j + (long long)Ns * i3];
    }
    out[j] = acc;
  }
}

hipError_t launch_energy_reduce(const float* epart, int Ns, int n3, const float* Mbe, int i3lo, int i3hi, double* out,
                                hipStream_t st) {
  if (i3lo < 0 || i3hi > n3 || i3lo > i3hi) return hipErrorInvalidValue;
  hipLaunchKernelGGL(energy_reduce_kernel, dim3((Ns + 255) / 256), dim3(256), 0, st, epart, Ns, i3lo, i3hi, Mbe, out);
  return hipGetLastError();
}

}  // namespace dis
