// LM reductions and solves (FP64) and the host pipeline of the transform
// gradient.
#pragma once
#include <hip/hip_runtime.h>

namespace dis {

// 1 energy + 6 gradient + 21 Hessian sums per state
constexpr int kLmSums = 28;

struct LmReduceArgs {
  const float2* Yr;      // [Ns, n3, C] encoded transformed volume E_s(T_s x)
  const float2* y;       // [Ns, n3, C] data
  const float2* J[6];    // [Ns, n3, C] encoded gradient volumes E_s(dT_s x / dT_g)
  const int* off;        // [S+1]
  const int* slist;      // [nlist] states reduced
  int Ns, n3, C;
  int i3lo, i3n;         // readout planes summed (E.nF of LMsolver.m:33,68,98)
  double* partial;       // [nlist][lm_partial_count()][kLmSums]
};

hipError_t launch_lm_reduce(const LmReduceArgs& a, int nlist, double* partial_out, double* sums_out,
                            hipStream_t st);
hipError_t launch_lm_solve(const double* sums, const double* w, int nlist, double winit, double* step,
                           hipStream_t st);
int lm_partial_count();

}  // namespace dis
