// Stockham mixed-radix FFT engine for batches of pencils staged in LDS.
//
// A workgroup of kThreads threads holds a tile of P pencils of length n
// (P*n <= kMaxTile complex values) in LDS and transforms all of them at once.
// Stages use radices {16, 8, 4, 2, 3, 5, 7} with compile-time unrolled
// butterflies, and any other prime factor of n through a runtime-radix stage
// (each output element is its butterfly's R-term sum); the stage loop is
// runtime.  Any n <= 1024 is supported.  The inverse transform is the
// forward transform applied to real/imaginary-swapped data (no separate
// inverse kernels), unnormalised.
//
// Two LDS layouts:
//   CONTIG  : element (p, j) at p*n + j   (pencils along the contiguous axis;
//             butterflies are mapped with the in-pencil index fastest)
//   strided : element (p, j) at p + P*j   (pencils along a strided axis; the
//             tile holds P consecutive positions of the contiguous axis and
//             butterflies are mapped with the pencil index fastest)
// Both keep consecutive lanes on consecutive LDS addresses in every stage.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dis {

constexpr int kThreads = 256;
constexpr int kMaxTile = 4096;
constexpr int kMaxStages = 12;

// LDS-qualified 2-float vector: keeps ds_* (b64) instructions inside the
// out-of-line stages, which receive their buffers as LDS pointers.
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) fvec2 lds_float2;
__device__ __forceinline__ float2 ld2(const lds_float2* p) {
  const fvec2 t = *p;
  return make_float2(t.x, t.y);
}
__device__ __forceinline__ void st2(lds_float2* p, float2 v) {
  fvec2 t;
  t.x = v.x;
  t.y = v.y;
  *p = t;
}

struct FftPlan {
  int n;
  int nst;
  int radix[kMaxStages];
};

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cswap(float2 a) { return make_float2(a.y, a.x); }
// multiply by -i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// ------------------------------------------------------------ small DFTs
// Forward DFT of length R in place: V[k] = sum_r v[r] exp(-2 pi i r k / R).
template <int R> struct Dft;

template <> struct Dft<2> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct Dft<4> {
  static __device__ __forceinline__ void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

// 4-point DFT on strided registers (used to compose 8 and 16)
__device__ __forceinline__ void dft4s(float2& a, float2& b, float2& c, float2& d) {
  float2 t0 = cadd(a, c), t1 = csub(a, c);
  float2 t2 = cadd(b, d), t3 = cmul_mi(csub(b, d));
  a = cadd(t0, t2);
  c = csub(t0, t2);
  b = cadd(t1, t3);
  d = csub(t1, t3);
}

#define DIS_SQRT1_2 0.70710678118654752440f
#define DIS_C16_1 0.92387953251128675613f  // cos(pi/8)
#define DIS_S16_1 0.38268343236508977173f  // sin(pi/8)

// multiply by W_N^k = exp(-2 pi i k / N) for the few constants used below
__device__ __forceinline__ float2 w8_1(float2 a) {  // (1 - i)/sqrt2
  return make_float2((a.x + a.y) * DIS_SQRT1_2, (a.y - a.x) * DIS_SQRT1_2);
}
__device__ __forceinline__ float2 w8_3(float2 a) {  // (-1 - i)/sqrt2
  return make_float2((a.y - a.x) * DIS_SQRT1_2, -(a.x + a.y) * DIS_SQRT1_2);
}
__device__ __forceinline__ float2 wconst(float2 a, float c, float s) {  // a * (c - i s)
  return make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
}

template <> struct Dft<8> {
  // r = 2a + b, k = c + 4d:  W8^{rk} = W4^{ac} W8^{bc} W2^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 y00 = v[0], y01 = v[2], y02 = v[4], y03 = v[6];
    float2 y10 = v[1], y11 = v[3], y12 = v[5], y13 = v[7];
    dft4s(y00, y01, y02, y03);
    dft4s(y10, y11, y12, y13);
    y11 = w8_1(y11);
    y12 = cmul_mi(y12);
    y13 = w8_3(y13);
    v[0] = cadd(y00, y10); v[4] = csub(y00, y10);
    v[1] = cadd(y01, y11); v[5] = csub(y01, y11);
    v[2] = cadd(y02, y12); v[6] = csub(y02, y12);
    v[3] = cadd(y03, y13); v[7] = csub(y03, y13);
  }
};

template <> struct Dft<16> {
  // r = 4a + b, k = c + 4d:  W16^{rk} = W4^{ac} W16^{bc} W4^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 y[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      y[b][0] = v[b]; y[b][1] = v[b + 4]; y[b][2] = v[b + 8]; y[b][3] = v[b + 12];
      dft4s(y[b][0], y[b][1], y[b][2], y[b][3]);
    }
    // twiddles W16^{bc}
    y[1][1] = wconst(y[1][1], DIS_C16_1, DIS_S16_1);
    y[1][2] = w8_1(y[1][2]);
    y[1][3] = wconst(y[1][3], DIS_S16_1, DIS_C16_1);
    y[2][1] = w8_1(y[2][1]);
    y[2][2] = cmul_mi(y[2][2]);
    y[2][3] = w8_3(y[2][3]);
    y[3][1] = wconst(y[3][1], DIS_S16_1, DIS_C16_1);
    y[3][2] = w8_3(y[3][2]);
    y[3][3] = wconst(y[3][3], -DIS_C16_1, -DIS_S16_1);  // W16^9
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float2 a0 = y[0][c], a1 = y[1][c], a2 = y[2][c], a3 = y[3][c];
      dft4s(a0, a1, a2, a3);
      v[c] = a0; v[c + 4] = a1; v[c + 8] = a2; v[c + 12] = a3;
    }
  }
};

// Odd prime radix: V[k] = v0 + sum_{r=1}^{h} (v_r+v_{R-r}) cos(2 pi rk/R)
//                         - i sum (v_r - v_{R-r}) sin(2 pi rk/R)
template <int R> struct OddConst;
template <> struct OddConst<3> {
  static __device__ __forceinline__ float c(int i) { return -0.5f; }
  static __device__ __forceinline__ float s(int i) { return 0.86602540378443864676f; }
};
template <> struct OddConst<5> {
  // cos/sin(2 pi q/5), q = 1, 2
  static __device__ __forceinline__ float c(int q) { return q == 1 ? 0.30901699437494742410f : -0.80901699437494742410f; }
  static __device__ __forceinline__ float s(int q) { return q == 1 ? 0.95105651629515357212f : 0.58778525229247312917f; }
};
template <> struct OddConst<7> {
  // cos/sin(2 pi q/7), q = 1, 2, 3
  static __device__ __forceinline__ float c(int q) {
    return q == 1 ? 0.62348980185873353053f : (q == 2 ? -0.22252093395631440429f : -0.90096886790241912624f);
  }
  static __device__ __forceinline__ float s(int q) {
    return q == 1 ? 0.78183148246802980871f : (q == 2 ? 0.97492791218182360702f : 0.43388373911755812048f);
  }
};

template <int R> struct DftOdd {
  static __device__ __forceinline__ void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 sp[H], sm[H];
    float2 v0 = v[0];
    float2 tot = v0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      sp[r - 1] = cadd(v[r], v[R - r]);
      sm[r - 1] = csub(v[r], v[R - r]);
      tot = cadd(tot, sp[r - 1]);
    }
    float2 out[R];
    out[0] = tot;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float2 a = v0, b = make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int q = (r * k) % R;  // angle index in [1, R-1]
        const int qq = q <= H ? q : R - q;
        const float c = OddConst<R>::c(qq);
        const float s = q <= H ? OddConst<R>::s(qq) : -OddConst<R>::s(qq);
        a.x += sp[r - 1].x * c; a.y += sp[r - 1].y * c;
        b.x += sm[r - 1].x * s; b.y += sm[r - 1].y * s;
      }
      // V[k] = a - i b ; V[R-k] = a + i b
      out[k] = make_float2(a.x + b.y, a.y - b.x);
      out[R - k] = make_float2(a.x - b.y, a.y + b.x);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};
template <> struct Dft<3> { static __device__ __forceinline__ void run(float2* v) { DftOdd<3>::run(v); } };
template <> struct Dft<5> { static __device__ __forceinline__ void run(float2* v) { DftOdd<5>::run(v); } };
template <> struct Dft<7> { static __device__ __forceinline__ void run(float2* v) { DftOdd<7>::run(v); } };

// ------------------------------------------------------------ Stockham stage
// One radix-R stage of the expand-form Stockham algorithm over all P pencils
// of the tile:  for butterfly b of a pencil (k = b mod Ns)
//   v[r] = in[b + r n/R] * W_n^{r k n/(Ns R)};  V = DFT_R(v);
//   out[(b/Ns) Ns R + k + r Ns] = V[r]
// In place: every thread loads all its butterflies, barrier, then stores.
// Kept out of line: inlining all radices into the stage loop lets the
// compiler hoist per-radix state and triples register pressure.
template <int R, bool CONTIG>
__device__ __noinline__ void stockham_stage(lds_float2* buf, const lds_float2* tw, int n, int Ns, int P, int tid) {
  constexpr int MAXB = (kMaxTile / R + kThreads - 1) / kThreads;
  const int nb = n / R;
  const int NB = P * nb;
  const int twstep = n / (Ns * R);
  float2 v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int q = tid + i * kThreads;
    dst[i] = -1;
    if (q < NB) {
      int p, b;
      if (CONTIG) { p = q / nb; b = q - p * nb; }
      else { b = q / P; p = q - b * P; }
      const int k = b % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = b + r * nb;
        v[i][r] = ld2(CONTIG ? buf + (p * n + j) : buf + (p + P * j));
      }
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], ld2(tw + r * k * twstep));
      }
      Dft<R>::run(v[i]);
      const int jd = (b / Ns) * Ns * R + k;
      dst[i] = CONTIG ? (p * n + jd) : (p + P * jd);
    }
  }
  __syncthreads();
  const int sd = CONTIG ? Ns : Ns * P;
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    if (dst[i] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) st2(buf + (dst[i] + r * sd), v[i][r]);
    }
  }
  __syncthreads();
}

// Radix-R Stockham stage for any R (the prime factors above 7): output r of
// butterfly b (k = b mod Ns) is
//   sum_q in[b + q n/R] W_n^{q k n/(Ns R)} W_R^{q r} = sum_q in[b + q n/R] W_n^{q m},
//   m = k n/(Ns R) + r n/R < n,
// written to (b/Ns) Ns R + k + r Ns: the same stage as stockham_stage with
// the R-point DFT evaluated directly (R^2 / R = R complex MACs per element).
// Element u of the tile is (p, r, b) with b fastest along the lanes, so the
// loads of consecutive lanes are consecutive in LDS.
template <bool CONTIG>
__device__ __noinline__ void generic_stage(lds_float2* buf, const lds_float2* tw, int n, int R, int Ns, int P,
                                           int tid) {
  constexpr int kE = kMaxTile / kThreads;
  const int nb = n / R;
  const int tot = P * n;
  const int twstep = n / (Ns * R);
  float2 out[kE];
  int dst[kE];
#pragma unroll
  for (int i = 0; i < kE; ++i) {
    const int u = tid + i * kThreads;
    dst[i] = -1;
    out[i] = make_float2(0.f, 0.f);
    if (u < tot) {
      int p, rem;
      if (CONTIG) { p = u / n; rem = u - p * n; }
      else { rem = u / P; p = u - rem * P; }
      const int r = rem / nb, b = rem - r * nb;
      const int k = b % Ns;
      const int m = k * twstep + r * nb;
      float2 acc = make_float2(0.f, 0.f);
      int e = 0;
      for (int q = 0; q < R; ++q) {
        const int j = b + q * nb;
        const float2 x = ld2(CONTIG ? buf + (p * n + j) : buf + (p + P * j));
        const float2 w = ld2(tw + e);
        acc = make_float2(acc.x + x.x * w.x - x.y * w.y, acc.y + x.x * w.y + x.y * w.x);
        e += m;
        if (e >= n) e -= n;
      }
      const int jd = (b / Ns) * Ns * R + k + r * Ns;
      dst[i] = CONTIG ? (p * n + jd) : (p + P * jd);
      out[i] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kE; ++i)
    if (dst[i] >= 0) st2(buf + dst[i], out[i]);
  __syncthreads();
}

// Forward unnormalised FFT of every pencil of the tile.  Caller must have
// synchronised after filling the tile; returns synchronised.
template <bool CONTIG>
__device__ void fft_tile(float2* gbuf, const float2* gtw, const FftPlan& pl, int P, int tid) {
  lds_float2* buf = (lds_float2*)gbuf;
  const lds_float2* tw = (const lds_float2*)gtw;
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix[s];
    switch (R) {
      case 16: stockham_stage<16, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 8: stockham_stage<8, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 4: stockham_stage<4, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 2: stockham_stage<2, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 3: stockham_stage<3, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 5: stockham_stage<5, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      case 7: stockham_stage<7, CONTIG>(buf, tw, pl.n, Ns, P, tid); break;
      default: generic_stage<CONTIG>(buf, tw, pl.n, R, Ns, P, tid); break;
    }
    Ns *= R;
  }
}

// Spectral index of FFT-order position j: j for j < ceil(n/2), else j - n.
__device__ __forceinline__ int freq_index(int j, int n) { return j < (n + 1) / 2 ? j : j - n; }

// exp(2 pi i c m) with the argument reduced in double before the float sincos.
__device__ __forceinline__ float2 phase_cycles(double c, int m) {
  double f = c * (double)m;
  f -= rint(f);
  float s, co;
  sincospif((float)(2.0 * f), &s, &co);
  return make_float2(co, s);
}

}  // namespace dis
