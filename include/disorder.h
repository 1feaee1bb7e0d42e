/* C-ABI of the MI355X-native DISORDER aligned-reconstruction hot path.
 *
 * Every entry point returns an int status (DIS_OK = 0, negative on error);
 * the message of the last error of a context is dis_last_error(ctx).
 * Nothing unwinds across this boundary.
 *
 * Data conventions (match MATLAB's mxGPUArray / mxGetComplexSingles layout,
 * so a MEX gateway passes pointers through without copies or transposes):
 *   - complex data is interleaved single precision (re, im) pairs;
 *   - arrays are column-major; a volume is [n1, n2, n3] with dims 1-2 the
 *     phase-encode plane and dim 3 the readout;
 *   - a stack of per-state volumes is [n1, n2, n3, S] (MATLAB keeps the state
 *     on dim 5 with singleton dim 4: same memory);
 *   - compact k-space data y is [Ns, n3, C] (MATLAB [Ns, 1, n3, C]); samples
 *     of all states are concatenated in state order;
 *   - motion T is a HOST array [S x 6] column-major (T[s + S*g]), g = 0..2
 *     translations in full-resolution voxels, g = 3..5 rotations in radians
 *     (MATLAB E.Tr [1,1,1,1,S,6]);
 *   - every array argument named x, y, out, S_maps, G is a DEVICE pointer on
 *     the context's GPU; every other pointer is a host pointer.
 * A context is not reentrant: use one thread per context.  All calls are
 * synchronous on return.
 */
#ifndef DISORDER_H
#define DISORDER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIS_OK 0
#define DIS_E_INVALID (-1)     /* bad argument or shape */
#define DIS_E_UNSUPPORTED (-2) /* feature of the reference not supported here */
#define DIS_E_NOMEM (-3)       /* device allocation failed or workspace limit too small */
#define DIS_E_HIP (-4)         /* HIP runtime error */
#define DIS_E_NODEVICE (-5)    /* no usable GPU */

typedef struct dis_ctx dis_ctx;
typedef struct dis_plan dis_plan;
typedef struct dis_shearlet dis_shearlet;

/* Library version (major*10000 + minor*100 + patch). */
int dis_version(void);
/* Static text of a status code. */
const char* dis_status_string(int status);

/* One HIP device, one stream, a growable workspace. */
int dis_ctx_create(int device, dis_ctx** out);
int dis_ctx_destroy(dis_ctx* ctx);
const char* dis_last_error(const dis_ctx* ctx);
int dis_sync(dis_ctx* ctx);
/* Upper bound on the workspace in bytes (0 = no bound).  Larger problems are
 * processed in batches of motion states that fit. */
int dis_set_workspace_limit(dis_ctx* ctx, size_t bytes);

/* sincRigidTransform with precomputeFactorsSincRigidTransform(kG,rkG,T,di,0,1,1)
 * (replaces Methods/sincRigidTransform.m:1 and
 * Methods/precomputeFactorsSincRigidTransform.m:1 as called from
 * Reconstruction/encode.m:67-69 and decode.m:104-106).
 *   n      : level grid [n1,n2,n3] of x;  Nfull : full-resolution grid
 *            (generateTransformGrids(Nfull,gpu,n,ceil((n+1)/2),1),
 *            Methods/downsampleOperators.m:63)
 *   di = 1 : x is [n] (x_per_state = 0) or [n,S] (= 1); out is [n,S]
 *   di = 0 : x is [n,S]; out is [n] when sumR, else [n,S]
 * Rotations beyond 90 degrees take the flip path of Methods/flipping.m. */
int dis_rigid_transform(dis_ctx* ctx, const int32_t n[3], const int32_t Nfull[3], const double* T, int32_t S,
                        int32_t di, int32_t sumR, int32_t x_per_state, const void* x, void* out);

/* Encoding operator description (the E/EH structs of Reconstruction/solveXT.m
 * after Methods/buildHarmonicSampling.m):
 *   n        : level grid of x, also the phase-encode grid of y unless
 *              SENSE folding is set (dis_plan_set_fold: y then lives on the
 *              smaller acquired grid)
 *   Nfull    : full-resolution grid of the motion parameters
 *   ncoils   : C (E.Sf dim 4)
 *   nstates  : segments (E.NSe); the first nmotion (E.NMs) are rigidly
 *              transformed, the rest (outer segments, encode.m:76-79) are not
 *   nsamples : Ns = sum of segment sizes */
typedef struct {
  int32_t n[3];
  int32_t Nfull[3];
  int32_t ncoils;
  int32_t nstates;
  int32_t nmotion;
  int32_t nsamples;
} dis_geom;

/* S_maps: DEVICE [n1,n2,n3,C] coil sensitivities (E.Sf), borrowed: must
 * outlive the plan.  k1, k2: 0-based FFT-order phase-encode indices of every
 * sample (rows of E.Fs{1}, E.Fs{2}); state_off: [nstates+1] segment offsets
 * (E.nSt/E.nEc); We: [Ns] sample weights (EH.We) or NULL. */
int dis_plan_create(dis_ctx* ctx, const dis_geom* g, const void* S_maps, const int32_t* k1, const int32_t* k2,
                    const int32_t* state_off, const float* We, dis_plan** out);
int dis_plan_destroy(dis_plan* plan);
/* Replace EH.We ([Ns] host, or NULL for none). */
int dis_plan_set_weights(dis_plan* plan, const float* We);
/* SENSE folding of the phase encodes (E.Uf{1:2}, EH.Ub{1:2}; Methods/fold.m,
 * ifold.m; Reconstruction/encode.m:100-104, decode.m:73-76): the data were
 * acquired on an ny1 x ny2 phase-encode grid smaller than the plan's n1 x n2
 * reconstruction grid, so E aliases S_c T_s x onto ny1 x ny2 (fold.m with
 * FOV(2) = ceil((n - ny) / 2)) before the Fourier rows of size ny, and E^H
 * replicates back (ifold.m).  k1/k2 of the plan must index the ny grid.
 * ny = n turns folding off.  Every operator of the plan (encode, decode,
 * encodedecode, energy, CG, LM) then runs folded. */
int dis_plan_set_fold(dis_plan* plan, int32_t ny1, int32_t ny2);

/* Readout region of interest of the motion estimation (E.nF of
 * Reconstruction/LMsolver.m:33,68,98, set per level by solveXT.m:345 from
 * FatShiftDir and redFOV): dis_lm and dis_lm_partials sum residuals,
 * gradients and Hessians over readout planes lo .. hi-1 (0-based; MATLAB
 * E.nF = lo+1:hi) only, still dividing by numel(y) of the whole data
 * (LMsolver.m:32).  lo = hi = 0 (or the full range) restores every plane.
 * Only a contiguous range is supported (solveXT builds E.nF as one). */
int dis_plan_set_readout_roi(dis_plan* plan, int32_t lo, int32_t hi);

/* Fractional finite-difference motion estimation (E.Fd of
 * Reconstruction/LMsolver.m:88-91,101-104,186-189; built per level by
 * Reconstruction/solveXT.m:353-354 and handed to LMsolver at :506 when
 * parXT.fractionOrder ~= 0): with weights set, dis_lm and dis_lm_partials
 * pass the residual of every state, its six Jacobians and every trial
 * residual through the unnormalised nF-point DFT along the readout planes
 * of the ROI (fftGPU along dim 3) and multiply them by Fd before the
 * energies, gradients and Hessians are summed.  Fd: host [nsamples x nF]
 * float, sample fastest, in the plan's sample order (the filter of
 * buildFilter(..., 'FractionalFiniteDiscreteIso', ...) at each sample's
 * phase-encode position, as buildHarmonicSampling samples it); nF must equal
 * the readout ROI's length (n3 without one) when LM runs.  NULL clears.
 * Encode, decode, E^H E, energies and CG are unaffected.  Not available on a
 * SENSE-folded plan (the reference's buildHarmonicSampling cannot reshape
 * an image-grid filter onto a smaller acquired grid either). */
int dis_plan_set_lm_filter(dis_plan* plan, const float* Fd, int32_t nF);

/* y = E x  (Reconstruction/encode.m:1).  T: host [nmotion x 6]. */
int dis_encode(dis_plan* plan, const double* T, const void* x, void* y);
/* x = E^H (We .* y)  (Reconstruction/decode.m:1). */
int dis_decode(dis_plan* plan, const double* T, const void* y, void* x);
/* out = E^H We E x  (Reconstruction/encodedecode.m:1). */
int dis_encodedecode(dis_plan* plan, const double* T, const void* x, void* out);
/* Reconstruction/computeEnergy.m:34-57 (no = 2): residual E x - y with
 * readout weights Mbe ([n3] host or NULL, EH.Mbe).  per_sample ([Ns] host or
 * NULL) receives sum over readout and coils per sample (avdi); total (host or
 * NULL) the sum over all samples.  Accumulated in FP64. */
int dis_energy(dis_plan* plan, const double* T, const void* x, const void* y, const float* Mbe,
               double* per_sample, double* total);

/* Preconditioned CG on the normal equations (replaces
 * Reconstruction/CGsolver.m:1 with A.Se, C.Ma and a Tikhonov R.Ti.la):
 *   b = Ma.*decode(y);  r = b - Ma.*(E^H E x + lambda x);  z = Se.*r; ...
 * Se, Ma: DEVICE [n1,n2,n3] real preconditioner (A.Se) and mask (C.Ma), or
 * NULL.  x: DEVICE [n1,n2,n3], initial guess when x_init (else zeroed), the
 * solution on return.  tolType: 1 'RelativeResidual2Error', 6 'Energy',
 * 7 'None' (Reconstruction/stopCondition.m), or 8, the rule of
 * Simulation/Xsolver.m:53-75: stop once max|alpha p|^2 < tol or
 * sqrt|sum conj(z) r| < 1e-10 (nIt then caps the iterations).  Inner products, the step
 * scalars and the residual norm are FP64.  En (host, 2 doubles per odd
 * iteration: fidelity, regulariser) receives the 'Energy' history, nEn its
 * column count, nn the iterations run (0 on the early exit of
 * CGsolver.m:70-75). */
int dis_cg(dis_plan* plan, const double* T, const void* y, const float* Se, const float* Ma, double lambda, void* x,
           int32_t x_init, int32_t nIt, int32_t tolType, double tol, double* En, int32_t* nEn, int32_t* nn);

/* Several independent reconstructions in one batched CG (replaces nvol calls
 * of Reconstruction/CGsolver.m:1 with A.Se, no mask, no regulariser,
 * tolType 'RelativeResidual2Error', tol 0 and x0 = 0, i.e. exactly nIt
 * iterations, as solveXT.m:431 runs them with motion).  plans[i] (all on
 * ctx), T[i] (host [nmotion_i x 6]), y[i], Se[i] (DEVICE, Se or NULL) and
 * x[i] (DEVICE output) describe volume i; sizes may differ.  Iteration 1 runs
 * eagerly and records every coefficient upload; with use_graph, iterations
 * 2..nIt of all volumes are one captured hipGraph replayed nIt-1 times (no
 * host work between iterations), otherwise they are launched eagerly from the
 * recording.  A volume whose CGsolver.m:101 break condition (|rs| < 1e-6)
 * is met keeps its iterate from then on.  res (host [nvol x nIt] or NULL)
 * receives sqrt(sum|conj(z) r| / sum|conj(Se b) b|) after each iteration;
 * ms (host [3] or NULL): setup, first iteration, remaining iterations. */
int dis_cg_batched(dis_ctx* ctx, int32_t nvol, dis_plan* const* plans, const double* const* T, const void* const* y,
                   const float* const* Se, void* const* x, int32_t nIt, int32_t use_graph, double* res, double* ms);

/* Autotuner (no counterpart in the reference, which has no kernels to tune):
 * for every axis of the grid n (motion grid Nfull, or n when NULL) times the
 * fused pencil pass on S states for each candidate shape -- the register
 * path where it exists, and the LDS Stockham engine with largest radix 4, 8
 * or 16 and 1..Pmax pencils per workgroup tile (LDS tile = P x n) -- and
 * keeps the fastest in the context; every later call on that grid uses it.
 * choice (host [9] or NULL): per axis {generic, pencils per tile, radix};
 * ms (host [3] or NULL): the chosen pass times. */
int dis_autotune(dis_ctx* ctx, const int32_t n[3], const int32_t Nfull[3], int32_t S, int32_t* choice, double* ms);
/* Set the pass shape of one axis of grid n by hand (generic: LDS engine;
 * P: pencils per tile, a power of two, 0 = largest that fits; maxradix 4, 8
 * or 16). */
int dis_set_pass_shape(dis_ctx* ctx, const int32_t n[3], int32_t axis, int32_t generic, int32_t P, int32_t maxradix);
/* Schedule of the coil stage of encode/decode/encodedecode/energy: 0 auto
 * (default: one workgroup per (shot, readout-plane) item, which measures
 * faster than the work queue also for uneven shots), 1 always one
 * workgroup per item, 2 the persistent launch fed by per-XCD
 * (shot, readout-plane) work queues.  Results are identical; only the
 * order in which items run changes. */
int dis_set_coil_schedule(dis_ctx* ctx, int32_t mode);
/* Bytes of state volumes a rigid transform keeps in flight (default 208
 * MiB; 0 = every state in every launch).  The passes of a transform (of
 * encode, decode, E^H E, the CG and LM trial transforms) run state batch by
 * state batch, each batch as many states as the budget holds, every pass of
 * a batch before the next batch starts, so a batch's intermediate volumes
 * stay in the last-level cache (MALL) between passes instead of streaming
 * through HBM.  Results are unchanged except for the order in which the
 * adjoint adds the states' contributions (one fixed order per budget). */
int dis_set_cache_budget(dis_ctx* ctx, int64_t bytes);
/* Forget every tuned shape (back to the built-in choice). */
int dis_autotune_clear(dis_ctx* ctx);

/* sincRigidTransformGradient (Methods/sincRigidTransformGradient.m:1) with
 * the factors of precomputeFactorsSincRigidTransform(kG,rkG,T,1,1,1,1)
 * (Reconstruction/LMsolver.m:60-66,96): x is one volume [n]; Tx (or NULL)
 * receives T_s x [n,S]; G receives dT_s x / dT_g as [n,S,6] (g fastest
 * after the state: G[:, s, g]).  Rotations beyond 90 degrees flip the
 * volume as the forward transform does; the flip itself is not
 * differentiated (as in the reference). */
int dis_rigid_transform_grad(dis_ctx* ctx, const int32_t n[3], const int32_t Nfull[3], const double* T, int32_t S,
                             const void* x, void* Tx, void* G);

/* Per-state partial sums of LMsolver (Reconstruction/LMsolver.m:52-122) for
 * the motion states of the plan: sums [nmotion x 28] (host, FP64, row per
 * state): sum_j |E_s T_s x - y|^2, sum Re(J_g conj r) (g = 1..6), and
 * sum Re(J_i conj J_j) over ha = [diag(1..6), nchoosek(1:6,2)] (:26), with
 * J_g = E_s dT_s x / dT_g; sums over the state's samples, readout and coils. */
int dis_lm_partials(dis_plan* plan, const double* T, const void* x, const void* y, double* sums);

/* One LMsolver call (replaces Reconstruction/LMsolver.m:1) on the motion
 * states: T [nmotion x 6] (host, in/out), w [nmotion] damping (E.w, in/out),
 * cT [nmotion] convergence flags (E.cT, in/out), winit (E.winit), tL[2]
 * (E.tL: translation voxels, rotation degrees).  Damped 6x6 normal equations
 * are solved in FP64 on the device.  En [nmotion] receives min(EnPrev, En)
 * (E.En), n_trials (or NULL) the number of trial-state evaluations.  EH.We is
 * not applied (as in the reference's LM encodes). */
int dis_lm(dis_plan* plan, double* T, double* w, uint8_t* cT, double winit, const double* tL, const void* x,
           const void* y, double* En, int32_t* n_trials);

/* ---- Coil-array compression (replaces Methods/compressCoils.m:1 with
 * useFold = 0, useInvMot = 0; the unfolded Gram branch :74-90, taken when the
 * coil maps and the data share their spatial size as in
 * Reconstruction/solveXT.m:61).  Arrays are [nv x C] complex64 on the device
 * with coil stride nv (the coil dimension last, as S [n1,n2,n3,C] and y). */

/* Gram matrix P(i,j) = sum_v Shat_i(v) conj(S_j(v)), Shat = S/(sum_c|S_c|^2 +
 * 1e-9) (compressCoils.m:78-87).  P: host complex128 (re, im pairs),
 * column-major C x C.  Split-bf16 MFMA products (three bf16 products per
 * FP32 product) accumulated in FP32 over chunks of 4096 voxels, the chunk
 * partials summed in FP64.  C <= 64. */
int dis_coil_gram(dis_ctx* ctx, const void* S, int64_t nv, int32_t C, double* P);

/* dis_coil_gram that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The block
 * partials live in a static per-GPU device array of 256 x 64 x 64 complex128
 * (16 MiB, held until the process ends); C <= 64 (larger counts return
 * DIS_E_UNSUPPORTED), any nv.  P_dev: DEVICE complex128 column-major C x C
 * output.  Stream-ordered: returns once the kernels are enqueued on the
 * context's stream (dis_sync waits).  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (other contexts, other
 * threads) race; calls on one context are ordered by its stream. */
int dis_coil_gram_noalloc(dis_ctx* ctx, const void* S, int64_t nv, int32_t C, void* P_dev);

/* [U,D] = svd(P) and the component counts (compressCoils.m:93-111) on the
 * host in FP64.  P is Hermitian PSD, so U holds its eigenvectors (column
 * phases arbitrary, as for any SVD) and D [C] the |eigenvalues| descending.
 * Ccount[n] = first k with cumsum(D)/sum(D) >= perc[n] when perc[n] < 1, else
 * perc[n].  U: host complex128 column-major C x C (may be NULL); D may be NULL. */
int dis_coil_basis(const double* P, int32_t C, const double* perc, int32_t nperc, int32_t* Ccount, double* D,
                   double* U);

/* out(v,k) = sum_c A(k,c) in(v,c) (aplGPU(A,S,4), compressCoils.m:117-118):
 * A host complex128 column-major Cc x C; in [nv x C], out [nv x Cc] on the
 * device.  BF16 MFMA with split operands accumulated in FP32: terms = 3
 * (hi/lo split, three products) or 6 (three-way split, six products). */
int dis_coil_apply(dis_ctx* ctx, const double* A, int32_t Cc, int32_t C, const void* in, int64_t nv, void* out,
                   int32_t terms);

/* The whole compressCoils (compressCoils.m:75-118): Gram, basis, then
 * S_out = A S and y_out = A y with A = U(:,1:max(Ccount))^H.  S_out and y_out
 * hold max(Ccount) <= C coil planes of nvS / nvy elements.  y may be NULL. */
int dis_compress_coils(dis_ctx* ctx, const void* S, int64_t nvS, int32_t C, const double* perc, int32_t nperc,
                       const void* y, int64_t nvy, void* S_out, void* y_out, int32_t terms, int32_t* Ccount,
                       double* D, double* U);

/* ---- Pyramid levels (Methods/downsampleOperators.m:34-60). */

/* Fourier resampling (replaces Methods/resampling.m:1 with fo = 0, mirror = 0)
 * of `batch` volumes [nin] -> [nout] (device complex64, column-major, volume
 * stride prod(n)): per axis fft/nin, crop or zero-pad the spectrum about DC
 * keeping frequencies [-ceil((m-1)/2), floor((m-1)/2)], m = min(nin, nout),
 * then ifft*nout.  Sizes <= 1024 (any factors). */
int dis_resample(dis_ctx* ctx, const int32_t nin[3], const int32_t nout[3], int64_t batch, const void* x, void* out);

/* Compact k-space rows of a pyramid level: out[r + Ns_out*col] =
 * scale * y[rows[r] + Ns_in*col] for col < ncol (y [Ns_in x ncol] device
 * complex64, rows host).  With the readout resampled by dis_resample this is
 * downsampleOperators.m:48-51 on the sampled rows. */
int dis_gather_samples(dis_ctx* ctx, const void* y, int32_t Ns_in, int64_t ncol, const int32_t* rows, int32_t Ns_out,
                       float scale, void* out);

/* CG preconditioner Se = 1/(sum_c |S_c|^2 + eps) per voxel
 * (Reconstruction/solveXT.m:231-232); S [nv x C] device, Se [nv] device float. */
int dis_coil_precond(dis_ctx* ctx, const void* S, int64_t nv, int32_t C, float eps, float* Se);
/* Gibbs-ringing filter of the level data (Methods/downsampleOperators.m:54-57,
 * Reconstruction/solveXT.m:237-238): filtering(y, buildFilter(NY, 'tukeyIso',
 * [], gpu, gibbs)) for compact rows y [Ns x n3 x C] (DEVICE, in place) whose
 * phase encodes are in k-space: per row an FFT along the readout, the radial
 * Tukey window at (k1[r], k2[r], k3) of the n1 x n2 x n3 grid (host int32
 * FFT-order indices), and the inverse FFT.  0 < gibbs <= 1; gibbs > 1
 * selects the cubic B-spline filter buildFilter(N, 'CubicBSpline') that
 * downsampleOperators.m:54-55 uses at coarse levels otherwise. */
int dis_filter_rows(dis_ctx* ctx, void* y, int64_t Ns, int32_t n3, int32_t C, const int32_t* k1, const int32_t* k2,
                    int32_t n1, int32_t n2, float gibbs);

/* ---- Learned motion proposal (no counterpart in the reference; proposes
 * per-shot motion updates that Reconstruction/LMsolver.m then accepts or
 * rejects).  Network per shot row: MLP 13 -> 256 -> 256 -> 6 with GELU,
 * then a 1D convolution along shot order (kernel 5, 6 -> 6 channels, zero
 * padding inside each scan).  Parameters are one flat FP32 device buffer of
 * dis_mlp_param_count() floats: W1 [256x13], b1, W2 [256x256], b2, W3 [6x256],
 * b3, Wc [6x6x5] (out, in, tap), bc.  Matrix products are BF16 MFMA with FP32
 * accumulation.  X [M x 13], Y/target [M x 6] are row-major FP32 device
 * arrays; scan_off [nscan+1] (host) delimits the scans (runs of rows in shot
 * order). */
int dis_mlp_param_count(void);
int dis_mlp_forward(dis_ctx* ctx, const float* params, const float* X, int32_t M, const int32_t* scan_off,
                    int32_t nscan, float* Y);
/* Gradients of sum(dY .* Y) with respect to every parameter (grads, flat like
 * params) and to X (dX [M x 13], or NULL). */
int dis_mlp_backward(dis_ctx* ctx, const float* params, const float* X, int32_t M, const int32_t* scan_off,
                     int32_t nscan, const float* dY, float* grads, float* dX);
/* One Adam step (beta 0.9/0.999, eps 1e-8, bias-corrected, step >= 1) on the
 * mean squared error between Y and target; loss (host, or NULL) receives it. */
int dis_mlp_train_step(dis_ctx* ctx, float* params, float* adam_m, float* adam_v, int32_t step, float lr,
                       const float* X, const float* target, int32_t M, const int32_t* scan_off, int32_t nscan,
                       float* loss);

/* ---- Shearlet-regularised final reconstruction (Reconstruction/CSsolver.m) ----
 * The 3-D shearlet system of Shearlets/buildShearlet.m (CSsolver.m:52-54:
 * one scale, K = 1, 'kos' directional and 'maxflat2' scaling filters) is
 * passed as its 2-D plane filters (combineShearlets, buildShearlet.m:94-106):
 * p12 [n1 x n2 x nf], p13 [n1 x n3 x nf], p32 [n3 x n2 x nf] complex64 host
 * arrays in FFT order (DC first), and bands [nw x 4] = (plane a, filter a,
 * plane b, filter b) with planes 0 = (k1,k2), 1 = (k1,k3), 2 = (k3,k2):
 * band w is S_w(k1,k2,k3) = plane_a[filter a] * plane_b[filter b].  The
 * handle owns the IRLS weights (nw float volumes) and three work volumes. */
int dis_shearlet_create(dis_ctx* ctx, const int32_t n[3], const void* p12, const void* p13, const void* p32,
                        int32_t nf, const int32_t* bands, int32_t nw, dis_shearlet** out);
int dis_shearlet_destroy(dis_shearlet* sh);
/* R.Sh.We = (|IFFT3(conj(S_w) FFT3(x))| + epsil)^(p - 2) for every band
 * (CSsolver.m:55-59 computeShearlet, :84-96). */
int dis_shearlet_weights(dis_ctx* ctx, dis_shearlet* sh, const void* x, double epsil, double p);
/* out += scale * IFFT3(sum_w S_w FFT3(We_w IFFT3(conj(S_w) FFT3(x)))): the
 * shearlet term y{5} of Reconstruction/regularize.m:27-50 with scale =
 * R.Sh.la * R.Sh.p / 2.  x, out: DEVICE [n1 x n2 x n3] complex64. */
int dis_shearlet_regularize(dis_ctx* ctx, dis_shearlet* sh, const void* x, double scale, void* out);
/* sum_w || sqrt(We_w) IFFT3(conj(S_w) FFT3(x)) ||^2 (one probe of the
 * regulariser trace, Reconstruction/traceEstimation.m:82-97); host result. */
int dis_shearlet_energy(dis_ctx* ctx, dis_shearlet* sh, const void* x, double* out);
/* sum_c sum_k Wgrid(k1,k2) |F2(S_c x)(k) / sqrt(n1 n2)|^2 (one probe of the
 * fidelity trace, traceEstimation.m:61-79 with EH.We laid on the k-space
 * grid, DC first): S DEVICE [n1 x n2 x n3 x C], Wgrid DEVICE float
 * [n1 x n2], x DEVICE complex64 volume; uses sh for its work buffers. */
int dis_fidelity_trace(dis_ctx* ctx, dis_shearlet* sh, const void* S, int32_t C, const float* Wgrid, const void* x,
                       double* out);
/* The same fidelity trace sample with SENSE folding (traceEstimation.m:61-79
 * with E.Uf): sum_c sum_k Wgrid(k1,k2) |F2(fold(S_c x))/sqrt(ny1 ny2)|^2 on the
 * acquired ny1 x ny2 grid (Wgrid: DEVICE float [ny1 x ny2]); S [n1,n2,n3,C]
 * and x [n1,n2,n3] on the device; out: host double. */
int dis_fidelity_trace_folded(dis_ctx* ctx, const int32_t n[3], const void* S, int32_t C, int32_t ny1, int32_t ny2,
                              const float* Wgrid, const void* x, double* out);
/* Reconstruction/CGsolver.m:1 with the shearlet term (CSsolver.m:76):
 * dis_cg with R x = lambda x + (la_sh p / 2) R_sh x.  tolType 'Energy' is
 * not supported with the shearlet term. */
int dis_cg_cs(dis_plan* plan, const double* T, const void* y, const float* Se, const float* Ma, double lambda,
              dis_shearlet* sh, double sh_scale, void* x, int32_t x_init, int32_t nIt, int32_t tolType, double tol,
              int32_t* nn);

/* Streaming-copy probe (no counterpart in the reference: a measurement of
 * the GPU's sustainable HBM rate that the HBM-bound passes are compared
 * with).  Copies nbytes (a multiple of 32 KiB) from DEVICE in to DEVICE
 * out reps times after one warm-up copy; ms receives the average time of one
 * copy.  variant: 0 grid-stride 8-byte lanes, 1 16-byte lanes, 2 four
 * 16-byte loads in flight per lane, 3 = 2 with nontemporal accesses, 4 the
 * pencil passes' tile shape (32 KiB per workgroup, 16 8-byte loads per
 * lane, then the stores), 5 = 4 with each workgroup looping over tiles and
 * the next tile's loads issued before the current tile's stores, 6 and 7
 * the strided register passes' access patterns on floor(nbytes / V) whole
 * 256 x 192 x 256 complex64 volumes (V = 100,663,296 bytes; nbytes >= V):
 * per workgroup the pencils of 16 consecutive dim-1 positions, lane
 * (p, t) element t + 16 r, along the readout axis at one dim-2 position (6,
 * elements n1 n2 apart) or along dim 2 within one readout plane (7, n1
 * apart), XCD-contiguous tile ranges.  blocks: grid size (0 = 8 per CU;
 * variants 4, 6 and 7: one per tile). */
int dis_stream_probe(dis_ctx* ctx, const void* in, void* out, int64_t nbytes, int32_t variant, int32_t blocks,
                     int32_t reps, double* ms);

/* Kernel classes for dis_profile_read. */
#define DIS_K_PASS 0   /* fused shear/translation pencil pass, one state per grid row */
#define DIS_K_BCAST 1  /* first forward pass, one input volume to all states */
#define DIS_K_ACCUM 2  /* last inverse pass with the state sum */
#define DIS_K_COIL 3   /* coil stage (coil multiply, dim-1 FFT, sampled dim-2 DFT, adjoint) */
#define DIS_K_FLIP 4   /* >90 degree flip pass */
#define DIS_K_GRAM 5   /* coil-compression Gram matrix */
#define DIS_K_GEMM 6   /* coil-compression apply (MFMA) */
#define DIS_K_SHEAR 7  /* shearlet FFT-axis pass (CS reconstruction) */
#define DIS_K_COUNT 8

/* Per-kernel-class timing with HIP events on the context's stream (no
 * counterpart in the reference, which only prints tic/toc totals,
 * Reconstruction/solveXT.m:184,263).  Events bracket every launch, except
 * that the passes of one rigid transform issued back to back share one event
 * pair per run of same-class launches (ms_total then also holds the few
 * microseconds between those launches).  Enabling resets the counters.
 * bytes_total is the algorithmic HBM traffic of the launches (each pass
 * reads and writes every state volume once). */
int dis_profile_enable(dis_ctx* ctx, int32_t on);
int dis_profile_read(dis_ctx* ctx, int32_t kernel_class, int64_t* launches, double* ms_total, double* bytes_total);

#ifdef __cplusplus
}
#endif
#endif /* DISORDER_H */
