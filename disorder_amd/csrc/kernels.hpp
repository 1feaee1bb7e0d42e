// Kernel argument blocks and host-side launchers shared by the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_engine.hpp"
#include "lanetab.hpp"

namespace dis {

// One fused 1D pass along `axis` of every pencil of a (n1,n2,n3) column-major
// volume:  y = F^-1 diag(phi) F x  with
//   phi(j) = exp(i (A_s r_src + B_s) k_axis(j)),
//   k_axis(j) = 2 pi m(j) / Nfull_axis   (m = FFT-order frequency index),
//   r_src(i) = rscale (i + 1 - rcent)    (full-resolution voxel units).
// A_s, B_s are per motion state.  This is one shear (src >= 0) or one
// translation (src < 0) of the sinc rigid transform.
struct PassArgs {
  const float2* in;
  float2* out;
  long long in_sstride;   // state stride of `in` in elements (0 = broadcast)
  long long out_sstride;  // state stride of `out`
  int n1, n2, n3;
  int axis;
  int src;
  int P;                  // pencils per tile
  int tiles_per_line;     // strided axes: tiles along the contiguous axis
  const float2* tw;       // twiddles of length n (device)
  FftPlan plan;
  const double* coefA;    // [S] or null
  const double* coefB;    // [S] or null
  double rscale;
  int rcent;
  double invN;            // 1 / Nfull_axis
  int S;                  // states looped inside a workgroup (BCAST / ACCUM)
  const int* states;      // MAP: optional list of state indices for grid.y
  int accumulate;         // ACCUM: add into existing `out`
  int force_generic;      // use the LDS-staged engine even where the register path exists
  // DUAL mode (derivative chains of the transform):
  //   X1 = F in, X2 = F in2 (0 when in2 is null),
  //   out  = F^-1(phi X1)              (skipped when out is null)
  //   out2 = F^-1(phi (beta X1 + X2))  with beta = i (A2_s r_src + B2_s) k_axis
  const float2* in2;
  float2* out2;
  long long in2_sstride, out2_sstride;
  const double* coefA2;
  const double* coefB2;
};

enum PassMode { MODE_MAP = 0, MODE_BCAST = 1, MODE_ACCUM = 2, MODE_FLIP = 3, MODE_DUAL = 4 };

// Launch a pass of the given mode; grid_states = gridDim.y for MAP/FLIP.
hipError_t launch_pass(const PassArgs& a, int mode, int grid_states, hipStream_t st);
hipError_t launch_pass_generic(const PassArgs& a, int mode, int grid_states, hipStream_t st);
int reg_radix(int n);

// twiddle table tw[k] = exp(-2 pi i k / n), k < n, computed in double.
hipError_t launch_twiddles(float2* tw, int n, hipStream_t st);

// ---------------------------------------------------------------- coil stage
// One workgroup per (motion state, readout plane).  Forward part: coil
// multiply, FFT along dim 1, and the DFT along dim 2 evaluated at the
// state's samples; adjoint part: the transpose of it with conj(S) and the
// coil sum.  Optional fused dim-1 translation on either side.
struct CoilArgs {
  const float2* U;        // transformed volumes, state stride us (0 = broadcast)
  long long us;
  float2* V;              // adjoint output volumes, state stride vs
  long long vs;
  const float2* Smaps;    // [n1,n2,n3,C]
  int n1, n2, n3, C;
  const int* k1;          // [Ns] FFT-order row index along dim 1
  const int* k2;          // [Ns] along dim 2
  const int* off;         // [S+1] state offsets
  const float* We;        // [Ns] or null
  int Ns;
  const float2* yin;      // [Ns, n3, C] data (DECODE, ENERGY) or null
  float2* yout;           // [Ns, n3, C] (ENCODE) or null
  float* epart;           // [Ns, n3] per-(sample, plane) energy (ENERGY) or null
  const double* tfwd;     // [S] dim-1 translation coefficient before the coil stage, or null
  const double* tadj;     // [S] dim-1 translation coefficient after the adjoint, or null
  double invN1;           // 1/Nfull_1 for the translation phase
  const float2* tw1;      // twiddles n1
  const float2* tw2;      // twiddles n2
  FftPlan plan1;
  int G;                  // pencils (dim-2 positions) per tile
  int state0;             // first state handled by this launch
  int vs_rel;             // V indexed relative to state0
  int us_rel;             // U indexed relative to state0
  int force_generic;      // use the LDS-staged engine even where the register path exists
  const int* rowptr;      // [S][n1+1] per-state row pointers of the samples sorted by k1
  const int* csr;         // [Ns] (k2 << 16) | (sample index within its state), sorted by k1 per state
  const int* slist;       // optional [nstates] list of the states handled (U/V indexed by list position)
  const int* sord;        // optional [Ns] register path: sample handled by each slot, per chunk of kCoilRegChunk
  // optional register-path adjoint lane table (lanetab.hpp): one block of
  // kThreads int4 per state and chunk of kCoilRegChunk row-sorted samples,
  // lane L taking chunk samples cs L .. cs L + cs - 1 (cs = ceil(samples in
  // the chunk / kThreads)); adjoff [S+1] gives each state's first block
  const int* adjtab;
  const int* adjoff;
  int* queue;             // persistent launch: 8 per-XCD work counters (zeroed before the launch), else null
  // SENSE folding: acquired phase-encode grid fy1 x fy2 smaller than n1 x n2
  // (0 = none); k1/k2 then index the acquired grid
  int fy1, fy2;
  int fmaxnec;            // largest state (samples) handled
  float2* fwork;          // scratch of coil_fold_work_bytes
  FftPlan fy1pl;          // FFT plan and twiddles of length fy1; twiddles of fy2
  const float2* fy1tw;
  const float2* fy2tw;
  // readout planes handled: i3lo .. i3lo + i3n - 1 (i3n = 0: every plane).
  // Kernels that do not restrict the planes compute all of them; callers
  // read only the range.
  int i3lo, i3n;
};

enum CoilMode { COIL_ENCODE = 1, COIL_DECODE = 2, COIL_EHE = 3, COIL_ENERGY = 4 };

// Samples per chunk of the register-path coil stage (3 per thread).
constexpr int kCoilRegChunk = 768;
static_assert(kAdjLanes == kThreads && kAdjMaxSamples == kCoilRegChunk, "adjoint lane table shape");

hipError_t launch_coil(const CoilArgs& a, int mode, int nstates, hipStream_t st);
// Persistent register-path coil stage: resident workgroups pull (state,
// readout-plane) items from per-XCD queues (stealing from the others when
// their own runs dry) until every item is done.  a.queue must point to 8
// ints; they are zeroed by a kernel on the stream first.  Returns
// hipErrorInvalidValue where the register path does not apply.
hipError_t launch_coil_persistent(const CoilArgs& a, int mode, int nstates, hipStream_t st);
bool coil_queue_available(int n1, int n2);
// true when a plan folded to fy1 x fy2 runs its coil stage on the register
// path (n1 a register length that fy1 divides), false when it needs the fold
// stage
bool coil_fold_register(int n1, int n2, int fy1);
// Samples per chunk of the coil-stage kernel that launch_coil picks for a
// (n1, n2) plane: a state with more samples is processed in several chunks.
int coil_chunk(int n1, int n2);
hipError_t launch_coil_generic(const CoilArgs& a, int mode, int nstates, hipStream_t st);
// The coil stage with SENSE folding (CoilArgs::fy1 > 0): batched kernels
// per state; fwork must hold coil_fold_work_bytes(...).  launch_coil and
// launch_coil_persistent dispatch here when folding is on.
hipError_t launch_coil_fold(const CoilArgs& a, int mode, int nstates, hipStream_t st);
size_t coil_fold_work_bytes(int n1, int n2, int n3, int C, int ny1, int ny2, int maxnec);

// ---------------------------------------------------------------- utilities
hipError_t launch_bcast(const float2* x, float2* out, long long nvox, int S, hipStream_t st);
hipError_t launch_sum_states(const float2* in, long long sstride, float2* out, long long nvox, int S,
                             int accumulate, hipStream_t st);
// per-sample energy over readout planes [i3lo, i3hi), weighted by Mbe (or 1)
hipError_t launch_energy_reduce(const float* epart, int Ns, int n3, const float* Mbe, int i3lo, int i3hi,
                                double* out, hipStream_t st);

}  // namespace dis
