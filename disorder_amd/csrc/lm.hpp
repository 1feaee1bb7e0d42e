// LM reductions and solves (FP64) and the host pipeline of the transform
// gradient.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dis {

// 1 energy + 6 gradient + 21 Hessian sums per state
constexpr int kLmSums = 28;

struct LmReduceArgs {
  const float2* Yr;      // [Ns, n3, C] encoded transformed volume E_s(T_s x)
  const float2* y;       // [Ns, n3, C] data, or null when Yr already holds the residual
  const float2* J[6];    // [Ns, n3, C] encoded gradient volumes E_s(dT_s x / dT_g)
  const int* off;        // [S+1]
  const int* slist;      // [nlist] states reduced
  int Ns, n3, C;
  int i3lo, i3n;         // readout planes summed (E.nF of LMsolver.m:33,68,98)
  double* partial;       // [nlist][lm_partial_count()][kLmSums]
};

// Fractional finite-difference weighting of LMsolver (E.Fd): per sample j
// of the listed states and coil c, the readout planes lo .. lo + nF - 1 of
// Yr - y (y optional) go through the unnormalised nF-point DFT and are
// multiplied by Fd[j + Ns k] (k the DFT index); the result replaces those
// planes of Yr.  epart (optional) receives per (sample, plane lo + k) the
// coil sum of the squared magnitudes.
struct ReadoutFilterArgs {
  float2* Yr;            // [Ns, n3, C] in/out
  const float2* y;       // [Ns, n3, C] or null
  const float* Fd;       // [Ns, nF]
  float* epart;          // [Ns, n3] or null
  const int* off;        // [S+1]
  const int* slist;      // [nlist] states filtered
  int Ns, n3, C;
  int lo, nF;
  const float2* tw;      // twiddles of length nF
  FftPlan plan;          // FFT plan of length nF
};

hipError_t launch_readout_filter(ReadoutFilterArgs a, int nlist, int maxnec, hipStream_t st);

hipError_t launch_lm_reduce(const LmReduceArgs& a, int nlist, double* partial_out, double* sums_out,
                            hipStream_t st);
hipError_t launch_lm_solve(const double* sums, const double* w, int nlist, double winit, double* step,
                           hipStream_t st);
int lm_partial_count();

}  // namespace dis
