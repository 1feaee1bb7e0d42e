// Host orchestration of the sinc rigid transform and of the encoding operator.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <vector>

#include "../../include/disorder.h"
#include "kernels.hpp"

namespace dis {

// Tuned launch shape of the pencil passes along one axis of one grid.
struct PassTune {
  int valid = 0;
  int generic = 0;   // LDS Stockham engine even where the register path exists
  int P = 0;         // pencils per workgroup tile (generic engine)
  int maxradix = 16; // largest power-of-two radix of the generic plan
};

// Level grid with its FFT plans and twiddle tables.
struct Grid {
  int n[3];
  int Nfull[3];
  long long nvox;
  FftPlan plan[3];
  const float2* tw[3];
  PassTune tune[3];
};

// Device-side view of a plan (sampling + coil maps).
struct OpView {
  dis_geom g;
  const float2* S;
  const int* k1;
  const int* k2;
  const int* off;
  const float* We;
  const int* rowptr;  // per-state CSR of the samples by k1
  const int* csr;
  const int* sord;  // register coil stage: sample of each slot (or null)
  const int* adjtab = nullptr;  // register coil stage: adjoint lane table (or null)
  const int* adjoff = nullptr;  // its first block per state
  const int* hoff;  // host copy of the offsets
  int fold[2];      // acquired phase-encode grid when SENSE folding is on, else {0, 0}
  int roi[2];       // readout planes [roi[0], roi[1]) of LMsolver (E.nF), or {0, 0} for all
  const float* Fd = nullptr;  // [Ns, fdn] fractional finite-difference weights of LMsolver (E.Fd), or null
  int fdn = 0;
};

// Fills the SENSE-folding fields of a coil-stage argument block (no-op
// without folding): plan and twiddles of the acquired grid, scratch for
// states of up to maxnec samples (context buffer slot 6).
int fold_setup(dis_ctx* c, const OpView& v, int maxnec, CoilArgs* a);
// Fidelity trace sample of traceEstimation.m:61-79 with SENSE folding onto
// the acquired ny1 x ny2 grid (weights W [ny1 x ny2] on that grid).
int fold_fidelity_trace(dis_ctx* c, const int32_t n[3], const float2* S, int C, int ny1, int ny2, const float* W,
                        const float2* x, double* out);

struct EnergyOut {
  const float* Mbe = nullptr;  // host [n3] or null
  double* per_sample = nullptr;  // host [Ns] or null
  double* total = nullptr;       // host or null
};

// Collects small host arrays for one upload; offsets are 16-byte aligned.
struct Staging {
  std::vector<unsigned char> bytes;
  size_t add(const void* p, size_t nbytes) {
    size_t o = (bytes.size() + 15) & ~size_t(15);
    bytes.resize(o + nbytes);
    if (nbytes) std::memcpy(bytes.data() + o, p, nbytes);
    return o;
  }
};

// context helpers implemented in api.hip
int ctx_fail(dis_ctx* c, int code, const char* msg);
hipStream_t ctx_stream(dis_ctx* c);
hipError_t ctx_set_device(dis_ctx* c);
int ctx_twiddles(dis_ctx* c, int n, const float2** out);
int ctx_workspace(dis_ctx* c, size_t bytes, void** out);
size_t ctx_workspace_limit(dis_ctx* c);
// Named device buffers that outlive one call (slot 0 is the op workspace).
int ctx_buffer(dis_ctx* c, int slot, size_t bytes, void** out);
int ctx_upload_bytes(dis_ctx* c, const Staging& st, const unsigned char** dptr);
// Upload tape (0 off, 1 record, 2 replay from the start) and its release.
void ctx_tape_mode(dis_ctx* c, int mode);
void ctx_tape_free(dis_ctx* c);
// Switch event profiling on (1) / off (0) / leave it (-1); returns the old state.
bool ctx_profiling(dis_ctx* c, int on);
int ctx_device(dis_ctx* c);
// Records, once per device, the device address of the allocation-free
// Gram's static scratch (coilcomp.hip); called with the device current.
int coil_gram_scratch_init(int device);
// Coil-stage schedule: 0 auto (work queue for uneven shots), 1 static grid,
// 2 persistent work queue.
int ctx_coil_schedule(dis_ctx* c);
long long ctx_cache_budget(dis_ctx* c);

// Event timing around a launch when profiling is enabled (no-ops otherwise).
int prof_begin(dis_ctx* c, int kclass);  // returns a slot or -1
void prof_end(dis_ctx* c, int slot, double bytes);
// Inside a group, launches of one class issued back to back with nothing
// else on the stream between them share one event pair (the first launch's
// start, the last one's end) and count as that many launches; a class change
// or the end of the outermost group records the end event.  Two events per
// run of same-class launches instead of two per launch: a transform of 64
// states in two-state batches records a handful of events, not a thousand.
struct ProfGroup {
  explicit ProfGroup(dis_ctx* c);
  ~ProfGroup();
  ProfGroup(const ProfGroup&) = delete;
  ProfGroup& operator=(const ProfGroup&) = delete;
  dis_ctx* c;
};

int make_plan(int n, FftPlan* pl, int maxradix = 16);
// Tuned pass shapes per (n1, n2, n3, axis), kept by the context.
bool ctx_tune_get(dis_ctx* c, const int n[3], int axis, PassTune* t);
void ctx_tune_set(dis_ctx* c, const int n[3], int axis, const PassTune& t);
// Time the candidate pass shapes of every axis of a grid on S states and
// keep the fastest (register path; LDS engine with max radix 4/8/16 and
// 1..Pmax pencils per tile).  out[3*axis + {0,1,2}] = generic, P, maxradix;
// ms_out[3] (or null) the chosen shapes' times.
int autotune_grid(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], int S, int32_t* out, double* ms_out);
int make_grid(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], Grid* g);

// T: host [Sld x 6] column-major; the states transformed are [0, S).
// Forward: U[s] = T_s x for s < S.  x_sstride = 0 broadcasts one volume.
// skip_t1 leaves out the final dim-1 translation (fused into the coil stage).
int rigid_forward(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* x,
                  long long x_sstride, float2* U, bool skip_t1);
// Inverse: out = sum_s T_s^H V[s] (sumR) or out[s] = T_s^H V[s].  `work`
// (S writable volumes, may equal V) holds the intermediate states when sumR.  skip_t1 leaves out the
// first dim-1 translation (fused into the coil stage).  accumulate adds into
// out (sumR only).
int rigid_inverse(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* V,
                  long long v_sstride, float2* work, float2* out, bool sumR, bool skip_t1, bool accumulate);

// E, E^H, E^H E and the residual energy through one coil-stage mode.
int op_apply(dis_ctx* c, const OpView& v, const double* T, const float2* x, const float2* yin, float2* yout,
             float2* xout, int mode, EnergyOut* eo);
// The k-space scratch [Ns x n3 x C] in which op_apply's E^H E leaves E x
// (its forward half), or null where E^H E runs as one fused or folded
// stage without it.
float2* ehe_kspace(dis_ctx* c, const OpView& v);

// Tx[s] = T_s x and G[g][s] = d(T_s x)/dT_g (g < 6) for states [0,S) of a
// [Sld x 6] motion array.  Buffers: Tx [S], G [6 S], X0 [3 S], Y [2 S] volumes.
int rigid_gradient(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* x, float2* Tx,
                   float2* G, float2* X0, float2* Y);
// One LMsolver call over the motion states (T [NM x 6] column-major, w, cT in/out).
int run_lm(dis_ctx* c, const OpView& v, double* T, double* w, unsigned char* cT, double winit, const double tL[2],
           const float2* x, const float2* y, double* En_out, int* n_trials);

// Raw per-state LM sums [NM x 28]: sum |r|^2, sum Re(J_g conj r) (6),
// sum Re(J_i conj J_j) over the 21 pairs of LMsolver.m:26.
int lm_partials(dis_ctx* c, const OpView& v, const double* T, const float2* x, const float2* y, double* sums);

// Coil compression: Gram matrix P (device, complex128 column-major C x C)
// and the MFMA apply out(v,k) = sum_c A(k,c) in(v,c) with the FP32 real
// operand Bm [2C][2Cc] on the device.
int coil_gram(dis_ctx* c, const float2* S, long long nv, int C, double2* P_dev_out);
int coil_apply(dis_ctx* c, const float* Bm_dev, int Cc, int C, const float2* in, long long nv, long long in_cs,
               float2* out, long long out_cs, int terms);

// Batched CG over independent problems (x0 = 0, nIt iterations of
// 'RelativeResidual2Error' with tol 0); iterations 2..nIt are replayed from
// a captured hipGraph when use_graph.  res [nvol x nIt] (or null) receives
// sqrt(sum|conj(z) r| / bnorm2) after each iteration; ms[3] (or null) the
// setup, first (recording) iteration and remaining iterations times.
int run_cg_batched(dis_ctx* c, int nvol, const OpView* views, const double* const* T, const float2* const* y,
                   const float* const* Se, float2* const* x, int nIt, bool use_graph, double* res, double* ms);

// Shearlet regulariser (cs.hip): out += scale * R_sh(x), R_sh the IRLS-weighted
// shearlet term of regularize.m:27-50 with the weights last computed.
int shearlet_apply(dis_ctx* c, dis_shearlet* sh, const float2* x, double scale, float2* out);

// sh (or null): adds (la_sh p / 2) R_sh to the CG operator (CSsolver.m:76);
// 'Energy' stopping is not available with it.
int run_cg(dis_ctx* c, const OpView& v, const double* T, const float2* y, const float* Se, const float* Ma,
           double lam, float2* x, bool x_init, int nIt, int tolType, double tol, double* En, int* nEn, int* nn,
           dis_shearlet* sh = nullptr, double sh_scale = 0.0);

}  // namespace dis
