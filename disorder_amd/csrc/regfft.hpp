// Register-resident two-level FFT for pencils of length N = R1 * 16.
//
// Sixteen lanes own one pencil.  Decimation in time with n = t + 16 r and
// k = k1 + R1 k2:
//   stage 1 (lane t < 16):  Y_t[k1] = DFT_R1_r( x[t + 16 r] ) * W_N^{t k1}
//   exchange through LDS (one write, one read per element)
//   stage 2 (lane k1 < R1): X[k1 + R1 k2] = DFT_16_t( Y_t[k1] )
// The forward output (lane k1 holds X[k1 + R1 k2]) is exactly the stage-1
// input pattern of the swapped factorisation (16, R1), so the inverse of a
// fused pass runs reg_fft<16, R1> on the same registers and returns the data
// to the original input layout (lane t holds y[t + 16 r]) with no extra
// exchange.  For R1 = 16 both factorisations are the same.
#pragma once
#include "fft_engine.hpp"

namespace dis {

#define DIS_SQRT3_2 0.86602540378443864676f

template <> struct Dft<12> {
  // r = 4a + b, k = c + 3d:  W12^{rk} = W3^{ac} W12^{bc} W4^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 y[4][3];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float2 t3[3] = {v[b], v[b + 4], v[b + 8]};
      DftOdd<3>::run(t3);
      y[b][0] = t3[0];
      y[b][1] = t3[1];
      y[b][2] = t3[2];
    }
    y[1][1] = wconst(y[1][1], DIS_SQRT3_2, 0.5f);   // W12^1
    y[1][2] = wconst(y[1][2], 0.5f, DIS_SQRT3_2);   // W12^2
    y[2][1] = wconst(y[2][1], 0.5f, DIS_SQRT3_2);   // W12^2
    y[2][2] = wconst(y[2][2], -0.5f, DIS_SQRT3_2);  // W12^4
    y[3][1] = cmul_mi(y[3][1]);                     // W12^3 = -i
    y[3][2] = make_float2(-y[3][2].x, -y[3][2].y);  // W12^6 = -1
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float2 a0 = y[0][c], a1 = y[1][c], a2 = y[2][c], a3 = y[3][c];
      dft4s(a0, a1, a2, a3);
      v[c] = a0;
      v[c + 3] = a1;
      v[c + 6] = a2;
      v[c + 9] = a3;
    }
  }
};

template <> struct Dft<6> {
  // r = 2a + b, k = c + 3d:  W6^{rk} = W3^{ac} W6^{bc} W2^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 e[3] = {v[0], v[2], v[4]};
    float2 o[3] = {v[1], v[3], v[5]};
    DftOdd<3>::run(e);
    DftOdd<3>::run(o);
    o[1] = wconst(o[1], 0.5f, DIS_SQRT3_2);   // W6^1
    o[2] = wconst(o[2], -0.5f, DIS_SQRT3_2);  // W6^2
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] = cadd(e[c], o[c]);
      v[c + 3] = csub(e[c], o[c]);
    }
  }
};

template <> struct Dft<9> {
  // r = 3a + b, k = c + 3d:  W9^{rk} = W3^{ac} W9^{bc} W3^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 y[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float2 t3[3] = {v[b], v[b + 3], v[b + 6]};
      DftOdd<3>::run(t3);
      y[b][0] = t3[0];
      y[b][1] = t3[1];
      y[b][2] = t3[2];
    }
    y[1][1] = wconst(y[1][1], 0.76604444311897803520f, 0.64278760968653932632f);   // W9^1
    y[1][2] = wconst(y[1][2], 0.17364817766693034885f, 0.98480775301220805936f);   // W9^2
    y[2][1] = wconst(y[2][1], 0.17364817766693034885f, 0.98480775301220805936f);   // W9^2
    y[2][2] = wconst(y[2][2], -0.93969262078590838405f, 0.34202014332566873304f);  // W9^4
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float2 t3[3] = {y[0][c], y[1][c], y[2][c]};
      DftOdd<3>::run(t3);
      v[c] = t3[0];
      v[c + 3] = t3[1];
      v[c + 6] = t3[2];
    }
  }
};

template <> struct Dft<10> {
  // r = 2a + b, k = c + 5d:  W10^{rk} = W5^{ac} W10^{bc} W2^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 e[5], o[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
      e[a] = v[2 * a];
      o[a] = v[2 * a + 1];
    }
    DftOdd<5>::run(e);
    DftOdd<5>::run(o);
    o[1] = wconst(o[1], 0.80901699437494745f, 0.58778525229247314f);  // W10^1
    o[2] = wconst(o[2], 0.30901699437494745f, 0.95105651629515353f);  // W10^2
    o[3] = wconst(o[3], -0.30901699437494734f, 0.95105651629515364f);  // W10^3
    o[4] = wconst(o[4], -0.80901699437494734f, 0.58778525229247325f);  // W10^4
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      v[c] = cadd(e[c], o[c]);
      v[c + 5] = csub(e[c], o[c]);
    }
  }
};

template <> struct Dft<14> {
  // r = 2a + b, k = c + 7d:  W14^{rk} = W7^{ac} W14^{bc} W2^{bd}
  static __device__ __forceinline__ void run(float2* v) {
    float2 e[7], o[7];
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      e[a] = v[2 * a];
      o[a] = v[2 * a + 1];
    }
    DftOdd<7>::run(e);
    DftOdd<7>::run(o);
    o[1] = wconst(o[1], 0.90096886790241915f, 0.43388373911755812f);  // W14^1
    o[2] = wconst(o[2], 0.62348980185873359f, 0.7818314824680298f);  // W14^2
    o[3] = wconst(o[3], 0.22252093395631445f, 0.97492791218182362f);  // W14^3
    o[4] = wconst(o[4], -0.22252093395631434f, 0.97492791218182362f);  // W14^4
    o[5] = wconst(o[5], -0.62348980185873348f, 0.78183148246802991f);  // W14^5
    o[6] = wconst(o[6], -0.90096886790241903f, 0.43388373911755823f);  // W14^6
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      v[c] = cadd(e[c], o[c]);
      v[c + 7] = csub(e[c], o[c]);
    }
  }
};

// Exchange-buffer size (elements) per pencil for both factorisations,
// padded so that the lanes of a half-wavefront (32 lanes, LDS b64 accesses
// are served per half) hit distinct banks:
//   CONTIG (lane = 16 p + t): two pencils of 16 consecutive t -> stride = 16 mod 32
//   strided (lane = p + 16 t): sixteen pencils, two t values -> stride = 2 mod 32 (x odd)
template <int R1> struct RegFftSize {
  static constexpr int N = R1 * 16;
  static constexpr int BASE = (R1 * 17 > 16 * (R1 + 1)) ? R1 * 17 : 16 * (R1 + 1);
  static constexpr int pad_to(int base, int residue) { return base + (((residue - base % 32) % 32) + 32) % 32; }
  static constexpr int XB = pad_to(BASE, 16);
  static constexpr int XB_STRIDED = pad_to(BASE, 18);
};

// Orders a wavefront's LDS accesses around an exchange among its own lanes.
// The LDS executes one wavefront's instructions in issue order, so only the
// compiler has to be kept from moving accesses across this point (the
// wavefront-scope fences emit no instructions).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One two-level FFT of a pencil: RA x RB = N, lanes t < RB enter with
// v[r] = x[t + RB r] (r < RA); lanes k1 < RA leave with v[k2] = X[k1 + RA k2]
// (k2 < RB).  xb: this pencil's exchange region; tw: W_N^i, i < N.  Contains
// two workgroup barriers (every lane of the workgroup must call it), or with
// WAVE two wavefront-order points: WAVE requires the pencil's lanes to sit in
// one wavefront (lane = 16 p + t) and xb to be touched by no other pencil.
template <int RA, int RB, bool WAVE = false>
__device__ __forceinline__ void reg_fft(float2 (&v)[16], int t, lds_float2* xb, const lds_float2* tw) {
  constexpr int LD = RB + 1;  // padded exchange row
  if (t < RB) {
    Dft<RA>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < RA; ++k1) v[k1] = cmul(v[k1], ld2(tw + t * k1));
#pragma unroll
    for (int k1 = 0; k1 < RA; ++k1) st2(xb + k1 * LD + t, v[k1]);
  }
  if (WAVE) wave_lds_sync(); else __syncthreads();
  if (t < RA) {
#pragma unroll
    for (int tt = 0; tt < RB; ++tt) v[tt] = ld2(xb + t * LD + tt);
    Dft<RB>::run(v);
  }
  if (WAVE) wave_lds_sync(); else __syncthreads();
}

// Phase ramp exp(2 pi i c m) over the forward output of lane k1 (< R1):
// element k2 has frequency index m = k1 + R1 k2 (k2 < 8) or k1 + R1 k2 - N
// (k2 >= 8, FFT order with the Nyquist bin negative).  Two anchors from
// double-reduced sincos and a 7-step complex recurrence each.
template <int R1>
__device__ __forceinline__ void ramp16(float2 (&ph)[16], double c, int k1) {
  constexpr int N = R1 * 16;
  const float2 a0 = phase_cycles(c, k1);
  const float2 a1 = phase_cycles(c, k1 - N / 2);
  const float2 w = phase_cycles(c, R1);
  ph[0] = a0;
  ph[8] = a1;
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    ph[k] = cmul(ph[k - 1], w);
    ph[k + 8] = cmul(ph[k + 7], w);
  }
}

// Phase ramp over the stage-1 input layout of reg_fft<R1,16> (lane t holds
// element t + 16 r, r < R1): m = t + 16 r below N/2, t + 16 r - N at or
// above it.  Even R1: rows r < R1/2 lie below N/2.  Odd R1: rows
// r < (R1-1)/2 lie below, row (R1-1)/2 straddles N/2 = 16 (R1-1)/2 + 8
// (lanes t >= 8 at or above it), the rest lie above.
template <int R1>
__device__ __forceinline__ void ramp_in(float2 (&ph)[16], double c, int t) {
  constexpr int N = R1 * 16;
  constexpr int H = R1 / 2;
  const float2 w = phase_cycles(c, 16);
  if constexpr (R1 % 2 == 0) {
    const float2 a0 = phase_cycles(c, t);
    const float2 a1 = phase_cycles(c, t - N / 2);
    ph[0] = a0;
    ph[H] = a1;
#pragma unroll
    for (int k = 1; k < H; ++k) {
      ph[k] = cmul(ph[k - 1], w);
      ph[k + H] = cmul(ph[k + H - 1], w);
    }
  } else {
    const float2 a0 = phase_cycles(c, t);
    const float2 a1 = phase_cycles(c, t + 16 * (H + 1) - N);
    ph[0] = a0;
#pragma unroll
    for (int k = 1; k <= H; ++k) ph[k] = cmul(ph[k - 1], w);
    ph[H + 1] = a1;
#pragma unroll
    for (int k = H + 2; k < R1; ++k) ph[k] = cmul(ph[k - 1], w);
    if (t >= 8) ph[H] = cmulc(a1, w);  // m = t + 16 H - N
  }
}

}  // namespace dis
