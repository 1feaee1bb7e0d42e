// MATLAB MEX gateway over the C-ABI of include/disorder.h.
//
// Build inside MATLAB (R2018a or later, Parallel Computing Toolbox):
//   mex -R2018a -I<repo>/include disorder_mex.cpp -L<repo>/disorder_amd -ldisorder_hip
// It is compile-only outside MATLAB: neither mex.h nor gpu/mxGPUArray.h
// ships with ROCm.
//
// Usage (every array argument is a gpuArray of complex single unless noted):
//   h  = disorder_mex('ctx_create', device)                 -> uint64 handle
//   disorder_mex('ctx_destroy', h)
//   p  = disorder_mex('plan_create', h, n, Nfull, nmotion, S, k1, k2, off, We)
//         n, Nfull: 1x3 double; S: [n1 n2 n3 C] gpuArray (kept alive by the
//         caller for the plan's lifetime); k1, k2: int32 0-based FFT-order
//         indices (E.Fs rows); off: int32 [nstates+1]; We: single or []
//   disorder_mex('plan_destroy', p)
//   y  = disorder_mex('encode', p, T, x)                    -- encode.m
//   x  = disorder_mex('decode', p, T, y)                    -- decode.m
//   z  = disorder_mex('encodedecode', p, T, x)              -- encodedecode.m
//   [e, es] = disorder_mex('energy', p, T, x, y, Mbe)       -- computeEnergy.m
//   [x, En, nn] = disorder_mex('cg', p, T, y, Se, Ma, lambda, x0, nIt, tolType, tol)
//   [T, w, cT, En] = disorder_mex('lm', p, T, w, cT, winit, tL, x, y)
//   xo = disorder_mex('transform', h, n, Nfull, T, di, sumR, x)
//   [Tx, G] = disorder_mex('gradient', h, n, Nfull, T, x)
//   xo = disorder_mex('resample', h, Nres, x)
// T is a double [S x 6] host matrix (E.Tr reshaped to [S 6]).
// Errors raise disorder:hip with dis_last_error's text.
#include <cstdint>
#include <cstring>
#include <string>

#include "gpu/mxGPUArray.h"
#include "mex.h"

#include "disorder.h"

namespace {

bool g_init = false;

void fail_status(dis_ctx* ctx, int st) {
  const char* msg = ctx ? dis_last_error(ctx) : dis_status_string(st);
  mexErrMsgIdAndTxt("disorder:hip", "%s (status %d)", msg && *msg ? msg : dis_status_string(st), st);
}

void check(dis_ctx* ctx, int st) {
  if (st != DIS_OK) fail_status(ctx, st);
}

template <class T>
T* handle(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("disorder:arg", "expected a handle");
  return reinterpret_cast<T*>(static_cast<uintptr_t>(*static_cast<uint64_t*>(mxGetData(a))));
}

mxArray* make_handle(const void* p) {
  mxArray* a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *static_cast<uint64_t*>(mxGetData(a)) = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p));
  return a;
}

// Read-only complex single gpuArray (the caller keeps the mxArray alive).
const mxGPUArray* gpu_in(const mxArray* a) {
  if (!mxIsGPUArray(a)) mexErrMsgIdAndTxt("disorder:arg", "expected a gpuArray");
  const mxGPUArray* g = mxGPUCreateFromMxArray(a);
  if (mxGPUGetClassID(g) != mxSINGLE_CLASS || mxGPUGetComplexity(g) != mxCOMPLEX)
    mexErrMsgIdAndTxt("disorder:arg", "expected complex single data");
  return g;
}

mxGPUArray* gpu_out(mwSize nd, const mwSize* dims) {
  return mxGPUCreateGPUArray(nd, dims, mxSINGLE_CLASS, mxCOMPLEX, MX_GPU_DO_NOT_INITIALIZE);
}

const void* dptr(const mxGPUArray* g) { return mxGPUGetDataReadOnly(g); }

void read3(const mxArray* a, int32_t out[3]) {
  if (mxGetNumberOfElements(a) != 3) mexErrMsgIdAndTxt("disorder:arg", "expected a 1x3 size");
  const double* v = mxGetPr(a);
  for (int i = 0; i < 3; ++i) out[i] = static_cast<int32_t>(v[i]);
}

// T [S x 6] double host matrix, column-major as the ABI expects
const double* motion(const mxArray* a, int32_t* S) {
  if (!mxIsDouble(a) || mxGetN(a) != 6) mexErrMsgIdAndTxt("disorder:arg", "T must be a double [S x 6] matrix");
  *S = static_cast<int32_t>(mxGetM(a));
  return mxGetPr(a);
}

struct PlanBox {  // plan plus the facts the gateway needs for output shapes
  dis_plan* plan;
  dis_ctx* ctx;
  int32_t n[3];
  int32_t C, nmotion, Ns;
};

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (!g_init) {
    mxInitGPU();
    mexLock();
    g_init = true;
  }
  if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("disorder:arg", "first argument is the command");
  char cmdbuf[64];
  mxGetString(prhs[0], cmdbuf, sizeof(cmdbuf));
  const std::string cmd(cmdbuf);

  if (cmd == "ctx_create") {
    dis_ctx* c = nullptr;
    const int dev = nrhs > 1 ? static_cast<int>(mxGetScalar(prhs[1])) : 0;
    check(nullptr, dis_ctx_create(dev, &c));
    plhs[0] = make_handle(c);
    return;
  }
  if (cmd == "ctx_destroy") {
    check(nullptr, dis_ctx_destroy(handle<dis_ctx>(prhs[1])));
    return;
  }
  if (cmd == "plan_create") {
    dis_ctx* c = handle<dis_ctx>(prhs[1]);
    dis_geom g{};
    read3(prhs[2], g.n);
    read3(prhs[3], g.Nfull);
    g.nmotion = static_cast<int32_t>(mxGetScalar(prhs[4]));
    const mxGPUArray* S = gpu_in(prhs[5]);
    const mwSize* sd = mxGPUGetDimensions(S);
    g.ncoils = mxGPUGetNumberOfDimensions(S) > 3 ? static_cast<int32_t>(sd[3]) : 1;
    const mxArray *k1 = prhs[6], *k2 = prhs[7], *off = prhs[8];
    if (!mxIsInt32(k1) || !mxIsInt32(k2) || !mxIsInt32(off)) mexErrMsgIdAndTxt("disorder:arg", "k1, k2, off: int32");
    g.nsamples = static_cast<int32_t>(mxGetNumberOfElements(k1));
    g.nstates = static_cast<int32_t>(mxGetNumberOfElements(off)) - 1;
    const float* We = (nrhs > 9 && !mxIsEmpty(prhs[9])) ? static_cast<const float*>(mxGetData(prhs[9])) : nullptr;
    PlanBox* b = new PlanBox{};
    b->ctx = c;
    std::memcpy(b->n, g.n, sizeof(b->n));
    b->C = g.ncoils;
    b->nmotion = g.nmotion;
    b->Ns = g.nsamples;
    const int st = dis_plan_create(c, &g, dptr(S), static_cast<const int32_t*>(mxGetData(k1)),
                                   static_cast<const int32_t*>(mxGetData(k2)),
                                   static_cast<const int32_t*>(mxGetData(off)), We, &b->plan);
    mxGPUDestroyGPUArray(S);
    if (st != DIS_OK) {
      delete b;
      fail_status(c, st);
    }
    plhs[0] = make_handle(b);
    return;
  }
  if (cmd == "plan_set_fold") {  // E.Uf{1:2}.NY: disorder_mex('plan_set_fold', plan, [NY1 NY2])
    PlanBox* b = handle<PlanBox>(prhs[1]);
    if (mxGetNumberOfElements(prhs[2]) < 2) mexErrMsgIdAndTxt("disorder:arg", "expected [NY1 NY2]");
    const double* v = mxGetPr(prhs[2]);
    check(b->ctx, dis_plan_set_fold(b->plan, static_cast<int32_t>(v[0]), static_cast<int32_t>(v[1])));
    return;
  }
  if (cmd == "plan_set_readout_roi") {  // E.nF: disorder_mex('plan_set_readout_roi', plan, E.nF)
    PlanBox* b = handle<PlanBox>(prhs[1]);
    int32_t lo = 0, hi = 0;
    const size_t m = mxGetNumberOfElements(prhs[2]);
    if (m > 0) {  // E.nF is one contiguous 1-based range (solveXT.m:345)
      const double* v = mxGetPr(prhs[2]);
      lo = static_cast<int32_t>(v[0]) - 1;
      hi = static_cast<int32_t>(v[m - 1]);
      if (hi - lo != static_cast<int32_t>(m)) mexErrMsgIdAndTxt("disorder:arg", "E.nF must be a contiguous range");
    }
    check(b->ctx, dis_plan_set_readout_roi(b->plan, lo, hi));
    return;
  }
  if (cmd == "plan_set_lm_filter") {  // E.Fd: disorder_mex('plan_set_lm_filter', plan, single(GRes))
    PlanBox* b = handle<PlanBox>(prhs[1]);
    if (mxIsEmpty(prhs[2])) {
      check(b->ctx, dis_plan_set_lm_filter(b->plan, nullptr, 0));
      return;
    }
    if (!mxIsSingle(prhs[2]) || mxIsComplex(prhs[2]) || static_cast<int64_t>(mxGetM(prhs[2])) != b->Ns)
      mexErrMsgIdAndTxt("disorder:arg", "E.Fd must be real single [nsamples x 1 x nF]");
    const int32_t nF = static_cast<int32_t>(mxGetNumberOfElements(prhs[2]) / mxGetM(prhs[2]));
    check(b->ctx, dis_plan_set_lm_filter(b->plan, static_cast<const float*>(mxGetData(prhs[2])), nF));
    return;
  }
  if (cmd == "plan_destroy") {
    PlanBox* b = handle<PlanBox>(prhs[1]);
    check(b->ctx, dis_plan_destroy(b->plan));
    delete b;
    return;
  }
  if (cmd == "encode" || cmd == "decode" || cmd == "encodedecode") {
    PlanBox* b = handle<PlanBox>(prhs[1]);
    int32_t S = 0;
    const double* T = motion(prhs[2], &S);
    const mxGPUArray* in = gpu_in(prhs[3]);
    mxGPUArray* out;
    int st;
    if (cmd == "encode") {
      const mwSize d[3] = {(mwSize)b->Ns, (mwSize)b->n[2], (mwSize)b->C};
      out = gpu_out(3, d);
      st = dis_encode(b->plan, T, dptr(in), mxGPUGetData(out));
    } else {
      const mwSize d[3] = {(mwSize)b->n[0], (mwSize)b->n[1], (mwSize)b->n[2]};
      out = gpu_out(3, d);
      st = cmd == "decode" ? dis_decode(b->plan, T, dptr(in), mxGPUGetData(out))
                           : dis_encodedecode(b->plan, T, dptr(in), mxGPUGetData(out));
    }
    mxGPUDestroyGPUArray(in);
    if (st != DIS_OK) {
      mxGPUDestroyGPUArray(out);
      fail_status(b->ctx, st);
    }
    plhs[0] = mxGPUCreateMxArrayOnGPU(out);
    mxGPUDestroyGPUArray(out);
    return;
  }
  if (cmd == "energy") {
    PlanBox* b = handle<PlanBox>(prhs[1]);
    int32_t S = 0;
    const double* T = motion(prhs[2], &S);
    const mxGPUArray* x = gpu_in(prhs[3]);
    const mxGPUArray* y = gpu_in(prhs[4]);
    const float* Mbe = (nrhs > 5 && !mxIsEmpty(prhs[5])) ? static_cast<const float*>(mxGetData(prhs[5])) : nullptr;
    plhs[0] = mxCreateDoubleScalar(0);
    mxArray* ps = mxCreateDoubleMatrix(b->Ns, 1, mxREAL);
    const int st = dis_energy(b->plan, T, dptr(x), dptr(y), Mbe, mxGetPr(ps), mxGetPr(plhs[0]));
    mxGPUDestroyGPUArray(x);
    mxGPUDestroyGPUArray(y);
    check(b->ctx, st);
    if (nlhs > 1) plhs[1] = ps; else mxDestroyArray(ps);
    return;
  }
  if (cmd == "cg") {
    PlanBox* b = handle<PlanBox>(prhs[1]);
    int32_t S = 0;
    const double* T = motion(prhs[2], &S);
    const mxGPUArray* y = gpu_in(prhs[3]);
    const float* Se = nullptr;  // real single gpuArrays or []
    const float* Ma = nullptr;
    const mxGPUArray* gSe = nullptr;
    const mxGPUArray* gMa = nullptr;
    if (!mxIsEmpty(prhs[4])) {
      gSe = mxGPUCreateFromMxArray(prhs[4]);
      Se = static_cast<const float*>(mxGPUGetDataReadOnly(gSe));
    }
    if (!mxIsEmpty(prhs[5])) {
      gMa = mxGPUCreateFromMxArray(prhs[5]);
      Ma = static_cast<const float*>(mxGPUGetDataReadOnly(gMa));
    }
    const double lambda = mxGetScalar(prhs[6]);
    const mwSize d[3] = {(mwSize)b->n[0], (mwSize)b->n[1], (mwSize)b->n[2]};
    mxGPUArray* x = nullptr;
    int32_t x_init = 0;
    if (!mxIsEmpty(prhs[7])) {
      x = mxGPUCopyFromMxArray(prhs[7]);
      x_init = 1;
    } else {
      x = gpu_out(3, d);
    }
    const int32_t nIt = static_cast<int32_t>(mxGetScalar(prhs[8]));
    char tt[32];
    mxGetString(prhs[9], tt, sizeof(tt));
    const int32_t tolType = !std::strcmp(tt, "RelativeResidual2Error") ? 1 : !std::strcmp(tt, "Energy") ? 6 : 7;
    const double tol = mxGetScalar(prhs[10]);
    mxArray* En = mxCreateDoubleMatrix(2, nIt / 2 + 2, mxREAL);
    int32_t nEn = 0, nn = 0;
    const int st = dis_cg(b->plan, T, dptr(y), Se, Ma, lambda, mxGPUGetData(x), x_init, nIt, tolType, tol,
                          mxGetPr(En), &nEn, &nn);
    mxGPUDestroyGPUArray(y);
    if (gSe) mxGPUDestroyGPUArray(gSe);
    if (gMa) mxGPUDestroyGPUArray(gMa);
    if (st != DIS_OK) {
      mxGPUDestroyGPUArray(x);
      fail_status(b->ctx, st);
    }
    mxSetN(En, nEn);
    plhs[0] = mxGPUCreateMxArrayOnGPU(x);
    mxGPUDestroyGPUArray(x);
    if (nlhs > 1) plhs[1] = En; else mxDestroyArray(En);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(nn);
    return;
  }
  if (cmd == "lm") {
    PlanBox* b = handle<PlanBox>(prhs[1]);
    mxArray* T = mxDuplicateArray(prhs[2]);
    mxArray* w = mxDuplicateArray(prhs[3]);   // double [S x 1]
    mxArray* cT = mxDuplicateArray(prhs[4]);  // uint8 [S x 1]
    const double winit = mxGetScalar(prhs[5]);
    const double* tL = mxGetPr(prhs[6]);
    const mxGPUArray* x = gpu_in(prhs[7]);
    const mxGPUArray* y = gpu_in(prhs[8]);
    mxArray* En = mxCreateDoubleMatrix(b->nmotion, 1, mxREAL);
    int32_t ntr = 0;
    const int st = dis_lm(b->plan, mxGetPr(T), mxGetPr(w), static_cast<uint8_t*>(mxGetData(cT)), winit, tL, dptr(x),
                          dptr(y), mxGetPr(En), &ntr);
    mxGPUDestroyGPUArray(x);
    mxGPUDestroyGPUArray(y);
    check(b->ctx, st);
    plhs[0] = T;
    if (nlhs > 1) plhs[1] = w; else mxDestroyArray(w);
    if (nlhs > 2) plhs[2] = cT; else mxDestroyArray(cT);
    if (nlhs > 3) plhs[3] = En; else mxDestroyArray(En);
    return;
  }
  if (cmd == "transform") {
    dis_ctx* c = handle<dis_ctx>(prhs[1]);
    int32_t n[3], Nf[3], S = 0;
    read3(prhs[2], n);
    read3(prhs[3], Nf);
    const double* T = motion(prhs[4], &S);
    const int32_t di = static_cast<int32_t>(mxGetScalar(prhs[5]));
    const int32_t sumR = static_cast<int32_t>(mxGetScalar(prhs[6]));
    const mxGPUArray* x = gpu_in(prhs[7]);
    const int32_t per_state = mxGPUGetNumberOfDimensions(x) > 3 ? 1 : 0;
    const mwSize dS[4] = {(mwSize)n[0], (mwSize)n[1], (mwSize)n[2], (mwSize)S};
    mxGPUArray* out = gpu_out((di == 0 && sumR) ? 3 : 4, dS);
    const int st = dis_rigid_transform(c, n, Nf, T, S, di, sumR, per_state, dptr(x), mxGPUGetData(out));
    mxGPUDestroyGPUArray(x);
    if (st != DIS_OK) {
      mxGPUDestroyGPUArray(out);
      fail_status(c, st);
    }
    plhs[0] = mxGPUCreateMxArrayOnGPU(out);
    mxGPUDestroyGPUArray(out);
    return;
  }
  if (cmd == "gradient") {
    dis_ctx* c = handle<dis_ctx>(prhs[1]);
    int32_t n[3], Nf[3], S = 0;
    read3(prhs[2], n);
    read3(prhs[3], Nf);
    const double* T = motion(prhs[4], &S);
    const mxGPUArray* x = gpu_in(prhs[5]);
    const mwSize dT[4] = {(mwSize)n[0], (mwSize)n[1], (mwSize)n[2], (mwSize)S};
    const mwSize dG[5] = {(mwSize)n[0], (mwSize)n[1], (mwSize)n[2], (mwSize)S, 6};
    mxGPUArray* Tx = gpu_out(4, dT);
    mxGPUArray* G = gpu_out(5, dG);
    const int st = dis_rigid_transform_grad(c, n, Nf, T, S, dptr(x), mxGPUGetData(Tx), mxGPUGetData(G));
    mxGPUDestroyGPUArray(x);
    if (st != DIS_OK) {
      mxGPUDestroyGPUArray(Tx);
      mxGPUDestroyGPUArray(G);
      fail_status(c, st);
    }
    plhs[0] = mxGPUCreateMxArrayOnGPU(Tx);
    if (nlhs > 1) plhs[1] = mxGPUCreateMxArrayOnGPU(G);
    mxGPUDestroyGPUArray(Tx);
    mxGPUDestroyGPUArray(G);
    return;
  }
  if (cmd == "resample") {
    dis_ctx* c = handle<dis_ctx>(prhs[1]);
    int32_t nout[3];
    read3(prhs[2], nout);
    const mxGPUArray* x = gpu_in(prhs[3]);
    const mwSize* d = mxGPUGetDimensions(x);
    const mwSize nd = mxGPUGetNumberOfDimensions(x);
    int32_t nin[3] = {(int32_t)d[0], nd > 1 ? (int32_t)d[1] : 1, nd > 2 ? (int32_t)d[2] : 1};
    int64_t batch = 1;
    for (mwSize i = 3; i < nd; ++i) batch *= (int64_t)d[i];
    const mwSize od[4] = {(mwSize)nout[0], (mwSize)nout[1], (mwSize)nout[2], (mwSize)batch};
    mxGPUArray* out = gpu_out(4, od);
    const int st = dis_resample(c, nin, nout, batch, dptr(x), mxGPUGetData(out));
    mxGPUDestroyGPUArray(x);
    if (st != DIS_OK) {
      mxGPUDestroyGPUArray(out);
      fail_status(c, st);
    }
    plhs[0] = mxGPUCreateMxArrayOnGPU(out);
    mxGPUDestroyGPUArray(out);
    return;
  }
  mexErrMsgIdAndTxt("disorder:arg", "unknown command '%s'", cmdbuf);
}
