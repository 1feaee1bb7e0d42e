// Host orchestration: the pass sequence of the sinc rigid transform
// (Methods/sincRigidTransform.m:38-114) and the batching of motion states
// through the coil stage (Reconstruction/encode.m, decode.m, encodedecode.m).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lm.hpp"
#include "lm_host.hpp"
#include "transform.hpp"

namespace dis {

namespace {

constexpr int kPer0[3] = {0, 2, 1};  // per(1,:) - 1
constexpr int kPer1[3] = {1, 0, 2};  // per(2,:) - 1

#define DCHK(c, expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) return ctx_fail((c), DIS_E_HIP, hipGetErrorString(e_)); \
  } while (0)

double wrap_to_pi(double a) {  // Mapping Toolbox wrapToPi
  if (a < -M_PI || a > M_PI) {
    double b = a + M_PI;
    const bool pos = b > 0;
    b = b - std::floor(b / (2 * M_PI)) * (2 * M_PI);
    if (b == 0 && pos) b = 2 * M_PI;
    a = b - M_PI;
  }
  return a;
}

double wrap_to_pi_half(double a) {  // Methods/wrapToPiHalf.m:9-12
  a = a - std::floor(a / M_PI) * M_PI;
  if (a > M_PI / 2) a -= M_PI;
  return a;
}

// Per-state factors of Methods/precomputeFactorsSincRigidTransform.m:45-60.
struct StateParam {
  double tan2[3], sn[3], t[3];
  bool flip[3];
};

StateParam state_param(const double* T, int Sld, int s) {
  StateParam p;
  for (int d = 0; d < 3; ++d) p.t[d] = T[s + (size_t)Sld * d];
  for (int m = 0; m < 3; ++m) {
    double th = wrap_to_pi(T[s + (size_t)Sld * (3 + m)]);
    p.flip[m] = std::fabs(th) > M_PI / 2;
    th = wrap_to_pi_half(th);
    p.tan2[m] = std::tan(th / 2);
    p.sn[m] = std::sin(th);
  }
  return p;
}

// One step of the pass sequence: a fused shear/translation pass (flip_m < 0)
// or the flip of rotation flip_m along `axis`.
struct Step {
  int axis, src;
  int flip_m;
  size_t offA, offB;  // staging offsets of the [S] coefficient arrays
  bool hasA, hasB;
  size_t offList;     // flip: staging offset of the flagged state list
  int nlist;
};

// Build the pass sequence of the forward (di=1) or inverse (di=0) transform
// for states [0,S) of T (rows s0..s0+S-1 of a [Sld x 6] array).
std::vector<Step> build_steps(const double* T, int S, int Sld, int di, bool skip_t1, Staging& st) {
  std::vector<StateParam> sp(S);
  for (int s = 0; s < S; ++s) sp[s] = state_param(T, Sld, s);
  const double sgn_tan = di == 1 ? 1.0 : -1.0;  // (-1)^(di-1)
  const double sgn_sin = di == 1 ? -1.0 : 1.0;  // (-1)^di
  const double sgn_t = di == 1 ? -1.0 : 1.0;    // (-1)^di
  std::vector<double> buf(S);
  auto shear = [&](int axis, int src, int m, bool tan_kind, int tdim) {
    Step stp{};
    stp.axis = axis;
    stp.src = src;
    stp.flip_m = -1;
    stp.hasA = m >= 0;
    stp.hasB = tdim >= 0;
    if (stp.hasA) {
      for (int s = 0; s < S; ++s) buf[s] = tan_kind ? sgn_tan * sp[s].tan2[m] : sgn_sin * sp[s].sn[m];
      stp.offA = st.add(buf.data(), sizeof(double) * S);
    }
    if (stp.hasB) {
      for (int s = 0; s < S; ++s) buf[s] = sgn_t * sp[s].t[tdim];
      stp.offB = st.add(buf.data(), sizeof(double) * S);
    }
    return stp;
  };
  auto flips = [&](int m, std::vector<Step>& out) {
    std::vector<int> lst;
    for (int s = 0; s < S; ++s)
      if (sp[s].flip[m]) lst.push_back(s);
    if (lst.empty()) return;
    const size_t o = st.add(lst.data(), sizeof(int) * lst.size());
    for (int ax : {kPer0[m], kPer1[m]}) {
      Step f{};
      f.axis = ax;
      f.src = -1;
      f.flip_m = m;
      f.offList = o;
      f.nlist = (int)lst.size();
      out.push_back(f);
    }
  };
  std::vector<Step> steps;
  if (di == 1) {
    for (int m = 0; m < 3; ++m) {
      const int a = kPer0[m], b = kPer1[m];
      flips(m, steps);
      steps.push_back(shear(a, b, m, true, -1));
      steps.push_back(shear(b, a, m, false, -1));
      // the dim-2 translation joins the last shear along dim 2 (:53, :66-71)
      steps.push_back(shear(a, b, m, true, m == 2 ? 1 : -1));
    }
    steps.push_back(shear(2, -1, -1, false, 2));
    if (!skip_t1) steps.push_back(shear(0, -1, -1, false, 0));
  } else {
    if (!skip_t1) steps.push_back(shear(0, -1, -1, false, 0));
    steps.push_back(shear(2, -1, -1, false, 2));
    for (int m = 2; m >= 0; --m) {
      const int a = kPer0[m], b = kPer1[m];
      steps.push_back(shear(a, b, m, true, m == 2 ? 1 : -1));
      steps.push_back(shear(b, a, m, false, -1));
      steps.push_back(shear(a, b, m, true, -1));
      flips(m, steps);
    }
  }
  return steps;
}

PassArgs pass_args(const Grid& gr, const Step& stp, const unsigned char* dbase, int S) {
  PassArgs a{};
  a.n1 = gr.n[0];
  a.n2 = gr.n[1];
  a.n3 = gr.n[2];
  a.axis = stp.axis;
  a.src = stp.src;
  const int n = gr.n[stp.axis];
  int P = 1;
  while (P * 2 * n <= kMaxTile && P * 2 <= 512) P *= 2;
  const PassTune& tu = gr.tune[stp.axis];
  if (tu.valid && tu.P > 0 && tu.P < P) P = tu.P;
  if (stp.axis != 0 && P > gr.n[0]) P = gr.n[0];
  a.P = P;
  a.force_generic = tu.valid ? tu.generic : 0;
  a.tiles_per_line = (gr.n[0] + P - 1) / P;
  a.tw = gr.tw[stp.axis];
  a.plan = gr.plan[stp.axis];
  a.coefA = stp.hasA ? (const double*)(dbase + stp.offA) : nullptr;
  a.coefB = stp.hasB ? (const double*)(dbase + stp.offB) : nullptr;
  if (stp.src >= 0) {
    a.rscale = (double)gr.Nfull[stp.src] / (double)gr.n[stp.src];
    a.rcent = (gr.n[stp.src] + 2) / 2;  // ceil((n+1)/2)
  }
  a.invN = 1.0 / (double)gr.Nfull[stp.axis];
  a.S = S;
  a.states = stp.flip_m >= 0 ? (const int*)(dbase + stp.offList) : nullptr;
  return a;
}

// States per batch of a transform: as many state volumes as the cache budget
// holds (at least one).  With several batches, every pass of a batch runs
// before the next batch starts, so the batch's intermediate volumes are read
// back from the last-level cache (MALL, 256 MB on MI355X) instead of HBM
// (E^H E at 256x192x256, 64 states: two states per batch, the 18 in-place
// passes 43.5 -> 38.9 ms).
int state_batch(dis_ctx* c, long long nvox, int S) {
  const long long budget = ctx_cache_budget(c);
  if (budget <= 0) return S;
  const long long per = std::max(1ll, budget / (8ll * nvox));
  return (int)std::min<long long>(S, per);
}

int run_steps(dis_ctx* c, const Grid& gr, const std::vector<Step>& steps, const unsigned char* dbase, int S,
              const float2* in, long long in_ss, float2* W, bool bcast_first, float2* final_out, int final_mode,
              bool accumulate) {
  hipStream_t stream = ctx_stream(c);
  const long long nv = gr.nvox;
  const double V = 8.0 * (double)nv;  // bytes of one complex64 volume
  // flips address states through lists: a sequence with flips runs unbatched
  bool flips = false;
  for (const Step& stp : steps) flips = flips || stp.flip_m >= 0;
  const int B = flips ? S : state_batch(c, nv, S);
  const size_t ns = steps.size();
  // the broadcast first pass and the state-summing last pass stay single
  // launches over every state: batching them would re-read the input or
  // re-accumulate the output per batch, and the flat summing pass keeps the
  // order in which the states' contributions are added
  const bool flat_first = B < S && bcast_first;
  const bool flat_last = B < S && final_mode == MODE_ACCUM;
  // one launch of step i for states [b0, b0 + nb)
  auto step = [&](size_t i, int b0, int nb) -> int {
    const Step& stp = steps[i];
    PassArgs a = pass_args(gr, stp, dbase, S);
    if (a.coefA) a.coefA += b0;
    if (a.coefB) a.coefB += b0;
    a.S = nb;
    const bool first = i == 0, last = i + 1 == ns;
    a.in = first ? in + b0 * in_ss : W + b0 * nv;
    a.in_sstride = first ? in_ss : nv;
    a.out = W + b0 * nv;
    a.out_sstride = nv;
    if (stp.flip_m >= 0) {
      a.in = W;
      a.in_sstride = nv;
      a.out = W;
      const int ps = prof_begin(c, DIS_K_FLIP);
      DCHK(c, launch_pass(a, MODE_FLIP, stp.nlist, stream));
      prof_end(c, ps, 2.0 * V * stp.nlist);
    } else if (first && bcast_first) {
      const int ps = prof_begin(c, DIS_K_BCAST);
      DCHK(c, launch_pass(a, MODE_BCAST, 1, stream));
      prof_end(c, ps, V * (1.0 + nb));
    } else if (last && final_mode == MODE_ACCUM) {
      a.out = final_out;
      a.out_sstride = 0;
      const bool acc = accumulate || b0 > 0;  // later batches add to the earlier ones' sum
      a.accumulate = acc ? 1 : 0;
      const int ps = prof_begin(c, DIS_K_ACCUM);
      DCHK(c, launch_pass(a, MODE_ACCUM, 1, stream));
      prof_end(c, ps, V * (nb + (acc ? 2.0 : 1.0)));
    } else {
      if (last && final_out) a.out = final_out + b0 * nv;
      const int ps = prof_begin(c, DIS_K_PASS);
      DCHK(c, launch_pass(a, MODE_MAP, nb, stream));
      prof_end(c, ps, 2.0 * V * nb);
    }
    return DIS_OK;
  };
  int rc = DIS_OK;
  ProfGroup group(c);
  if (flat_first && (rc = step(0, 0, S))) return rc;
  for (int b0 = 0; b0 < S; b0 += B) {
    const int nb = std::min(B, S - b0);
    for (size_t i = flat_first ? 1 : 0; i < ns - (flat_last ? 1 : 0); ++i)
      if ((rc = step(i, b0, nb))) return rc;
  }
  if (flat_last && (rc = step(ns - 1, 0, S))) return rc;
  return DIS_OK;
}

}  // namespace

int make_plan(int n, FftPlan* pl, int maxradix) {
  if (n <= 0 || n > 1024) return DIS_E_UNSUPPORTED;
  const int lr = maxradix >= 16 ? 4 : maxradix >= 8 ? 3 : maxradix >= 4 ? 2 : 1;  // log2 of the largest radix
  pl->n = n;
  pl->nst = 0;
  int m = n;
  for (int r : {7, 5, 3}) {
    while (m % r == 0) {
      if (pl->nst >= kMaxStages) return DIS_E_UNSUPPORTED;
      pl->radix[pl->nst++] = r;
      m /= r;
    }
  }
  int a = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++a;
  }
  // any other prime factor takes a runtime-radix stage (last)
  std::vector<int> big;
  for (int p = 11; m > 1; p += 2) {
    if ((long long)p * p > m) p = m;  // m itself is prime
    while (m % p == 0) {
      big.push_back(p);
      m /= p;
    }
  }
  std::vector<int> two;
  while (a >= lr) {
    two.push_back(1 << lr);
    a -= lr;
  }
  if (a > 0) two.push_back(1 << a);
  // powers of two first (the largest radix leads)
  std::vector<int> all(two.begin(), two.end());
  for (int s = 0; s < pl->nst; ++s) all.push_back(pl->radix[s]);
  for (int r : big) all.push_back(r);
  if ((int)all.size() > kMaxStages) return DIS_E_UNSUPPORTED;
  pl->nst = (int)all.size();
  for (int s = 0; s < pl->nst; ++s) pl->radix[s] = all[s];
  return DIS_OK;
}

int make_grid(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], Grid* g) {
  for (int d = 0; d < 3; ++d) {
    if (n[d] <= 0 || Nfull[d] <= 0) return ctx_fail(c, DIS_E_INVALID, "grid sizes must be positive");
    g->n[d] = n[d];
    g->Nfull[d] = Nfull[d];
    g->tune[d] = PassTune{};
    PassTune t;
    const int nn[3] = {n[0], n[1], n[2]};
    if (ctx_tune_get(c, nn, d, &t)) g->tune[d] = t;
    if (make_plan(n[d], &g->plan[d], g->tune[d].valid ? g->tune[d].maxradix : 16) != DIS_OK)
      return ctx_fail(c, DIS_E_UNSUPPORTED, "grid sizes must be at most 1024");
    int st = ctx_twiddles(c, n[d], &g->tw[d]);
    if (st) return st;
  }
  g->nvox = (long long)n[0] * n[1] * n[2];
  return DIS_OK;
}

int rigid_forward(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* x, long long x_sstride,
                  float2* U, bool skip_t1) {
  Staging st;
  std::vector<Step> steps = build_steps(T, S, Sld, 1, skip_t1, st);
  const unsigned char* d = nullptr;
  int rc = ctx_upload_bytes(c, st, &d);
  if (rc) return rc;
  hipStream_t stream = ctx_stream(c);
  const bool flip0 = !steps.empty() && steps[0].flip_m == 0;
  bool bcast = x_sstride == 0 && !flip0;
  const float2* in = x;
  long long in_ss = x_sstride;
  if (flip0) {  // materialise the states before flipping some of them
    if (x_sstride == 0) {
      DCHK(c, launch_bcast(x, U, gr.nvox, S, stream));
    } else if (x != U) {
      DCHK(c, hipMemcpyAsync(U, x, sizeof(float2) * gr.nvox * S, hipMemcpyDeviceToDevice, stream));
    }
    in = U;
    in_ss = gr.nvox;
  }
  return run_steps(c, gr, steps, d, S, in, in_ss, U, bcast, nullptr, MODE_MAP, false);
}

int rigid_inverse(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* V, long long v_sstride,
                  float2* work, float2* out, bool sumR, bool skip_t1, bool accumulate) {
  Staging st;
  std::vector<Step> steps = build_steps(T, S, Sld, 0, skip_t1, st);
  const unsigned char* d = nullptr;
  int rc = ctx_upload_bytes(c, st, &d);
  if (rc) return rc;
  hipStream_t stream = ctx_stream(c);
  const bool flip_last = steps.back().flip_m >= 0;
  float2* W = sumR ? work : out;
  if (sumR && !flip_last) return run_steps(c, gr, steps, d, S, V, v_sstride, W, false, out, MODE_ACCUM, accumulate);
  rc = run_steps(c, gr, steps, d, S, V, v_sstride, W, false, nullptr, MODE_MAP, false);
  if (rc) return rc;
  if (sumR) DCHK(c, launch_sum_states(W, gr.nvox, out, gr.nvox, S, accumulate ? 1 : 0, stream));
  return DIS_OK;
}

int fold_setup(dis_ctx* c, const OpView& v, int maxnec, CoilArgs* a) {
  if (v.fold[0] <= 0) return DIS_OK;
  const dis_geom& g = v.g;
  if (make_plan(v.fold[0], &a->fy1pl)) return ctx_fail(c, DIS_E_UNSUPPORTED, "SENSE folding: unsupported acquired length");
  int rc = ctx_twiddles(c, v.fold[0], &a->fy1tw);
  if (!rc) rc = ctx_twiddles(c, v.fold[1], &a->fy2tw);
  if (rc) return rc;
  void* w = nullptr;
  rc = ctx_buffer(c, 6, coil_fold_work_bytes(g.n[0], g.n[1], g.n[2], g.ncoils, v.fold[0], v.fold[1], maxnec), &w);
  if (rc) return rc;
  a->fwork = (float2*)w;
  a->fy1 = v.fold[0];
  a->fy2 = v.fold[1];
  a->fmaxnec = std::max(1, maxnec);
  return DIS_OK;
}

float2* ehe_kspace(dis_ctx* c, const OpView& v) {
  const dis_geom& g = v.g;
  if (!(v.fold[0] == 0 || coil_fold_register(g.n[0], g.n[1], v.fold[0]))) return nullptr;
  void* yb = nullptr;
  if (ctx_buffer(c, 7, sizeof(float2) * (size_t)std::max(1, g.nsamples) * g.n[2] * g.ncoils, &yb)) return nullptr;
  return (float2*)yb;
}

int op_apply(dis_ctx* c, const OpView& v, const double* T, const float2* x, const float2* yin, float2* yout,
             float2* xout, int mode, EnergyOut* eo) {
  const dis_geom& g = v.g;
  Grid gr;
  int rc = make_grid(c, g.n, g.Nfull, &gr);
  if (rc) return rc;
  hipStream_t stream = ctx_stream(c);
  const int S = g.nstates;
  bool anyT = false;
  for (int i = 0; i < 6 * g.nmotion; ++i)
    if (T[i] != 0.0) anyT = true;
  const int NM = anyT ? g.nmotion : 0;  // encode.m:61 -- no transform at all when T == 0
  const bool fwd = mode != COIL_DECODE;
  const bool adj = mode == COIL_DECODE || mode == COIL_EHE;
  const long long nv = gr.nvox;
  const size_t vol = sizeof(float2) * nv;
  int maxnec = 0;
  for (int s = 0; s < S; ++s) maxnec = std::max(maxnec, v.hoff[s + 1] - v.hoff[s]);
  // E^H E runs in place (U -> U) only when one workgroup pass covers all
  // coils and samples: with several chunks the adjoint of the first would
  // overwrite the U that the forward of the next still reads
  const bool inplace = !(mode == COIL_EHE && (g.ncoils > 8 || maxnec > coil_chunk(g.n[0], g.n[1])));
  const int vols_per_state = inplace ? 1 : 2;

  // workspace: [Sb (x2) volumes][energy partials Ns*n3 floats][per-sample doubles Ns]
  const size_t ns = (size_t)std::max(1, g.nsamples);
  const size_t epbytes = mode == COIL_ENERGY ? ((sizeof(float) * ns * g.n[2] + 255) & ~size_t(255)) : 0;
  const size_t esbytes = mode == COIL_ENERGY ? sizeof(double) * ns : 0;
  size_t lim = ctx_workspace_limit(c);
  int Sb = std::max(1, S);
  if (lim) {
    if (lim < vol * vols_per_state + epbytes + esbytes) return ctx_fail(c, DIS_E_NOMEM, "workspace limit below one state");
    Sb = (int)std::min<size_t>((size_t)Sb, (lim - epbytes - esbytes) / (vol * vols_per_state));
  }
  void* ws = nullptr;
  rc = ctx_workspace(c, vol * Sb * vols_per_state + epbytes + esbytes, &ws);
  if (rc) return rc;
  float2* U = (float2*)ws;
  float2* Vb = inplace ? U : U + nv * Sb;
  float* epart = (float*)((char*)ws + vol * Sb * vols_per_state);
  double* esamp = (double*)((char*)epart + epbytes);

  const int G = std::max(1, std::min(g.n[1], kMaxTile / g.n[0]));
  bool first_out = true;
  // the persistent work-queue launch on the register path runs only when
  // forced (schedule 2): with the coil stage split into its forward and
  // adjoint halves the static grid is faster also for uneven shots
  // (E^H E at 256x192x256, 64 shots with 10 % truncated: 32.4 ms static
  // against 60.8 ms persistent)
  int* queue = nullptr;
  {
    const bool want = ctx_coil_schedule(c) == 2 && NM > 0;
    if (want && coil_queue_available(g.n[0], g.n[1])) {
      void* qb = nullptr;
      rc = ctx_buffer(c, 4, 64, &qb);
      if (rc) return rc;
      queue = (int*)qb;
    }
  }
  CoilArgs fold_a{};
  rc = fold_setup(c, v, maxnec, &fold_a);
  if (rc) return rc;
  auto coil_args = [&](int s0, const float2* Usrc, long long us) {
    CoilArgs a = fold_a;
    a.U = Usrc;
    a.us = us;
    a.us_rel = 1;
    a.V = Vb;
    a.vs = nv;
    a.vs_rel = 1;
    a.Smaps = v.S;
    a.n1 = g.n[0];
    a.n2 = g.n[1];
    a.n3 = g.n[2];
    a.C = g.ncoils;
    a.k1 = v.k1;
    a.k2 = v.k2;
    a.off = v.off;
    a.We = v.We;
    a.rowptr = v.rowptr;
    a.csr = v.csr;
    a.sord = v.sord;
    a.adjtab = v.adjtab;
    a.adjoff = v.adjoff;
    a.Ns = g.nsamples;
    a.yin = yin;
    a.yout = yout;
    a.epart = epart;
    a.invN1 = 1.0 / (double)g.Nfull[0];
    a.tw1 = gr.tw[0];
    a.tw2 = gr.tw[1];
    a.plan1 = gr.plan[0];
    a.G = G;
    a.state0 = s0;
    return a;
  };

  // E^H E runs the coil stage as its two halves, the forward into a k-space
  // scratch [Ns x n3 x C] and the adjoint from it: at 256 x 192 x 256, 8
  // coils, 64 shots the two launches take 13.9 + 15.7 ms against 63.7 ms for
  // the fused on-chip kernel, for 2 x 805 MB of extra traffic (every plan on
  // the register path, including plans folded along dim 2 only)
  const bool split = mode == COIL_EHE && (v.fold[0] == 0 || coil_fold_register(g.n[0], g.n[1], v.fold[0]));
  float2* yscr = nullptr;
  if (split) {
    void* yb = nullptr;
    rc = ctx_buffer(c, 7, sizeof(float2) * (size_t)std::max(1, g.nsamples) * g.n[2] * g.ncoils, &yb);
    if (rc) return rc;
    yscr = (float2*)yb;
  }
  auto run_coil = [&](const CoilArgs& a0, int sb, bool prof) -> int {
    const int nmodes = split ? 2 : 1;
    for (int h = 0; h < nmodes; ++h) {
      CoilArgs a = a0;
      int m = mode;
      if (split) {
        m = h == 0 ? COIL_ENCODE : COIL_DECODE;
        if (h == 0) {
          a.yout = yscr;
          a.tadj = nullptr;
        } else {
          a.yin = yscr;
          a.tfwd = nullptr;
        }
      }
      const int ps = prof ? prof_begin(c, DIS_K_COIL) : -1;
      if (queue && a.U != x) {
        a.queue = queue;
        DCHK(c, launch_coil_persistent(a, m, sb, stream));
      } else {
        DCHK(c, launch_coil(a, m, sb, stream));
      }
      // pass model of the coil stage per state: (6C+2) volumes (E^H E), (3C+1) forward or adjoint only
      const double cv = m == COIL_EHE ? 6.0 * g.ncoils + 2 : 3.0 * g.ncoils + 1;
      prof_end(c, ps, cv * 8.0 * (double)nv * sb);
    }
    return DIS_OK;
  };

  // motion states, in batches of Sb
  for (int s0 = 0; s0 < NM; s0 += Sb) {
    const int sb = std::min(Sb, NM - s0);
    if (fwd) {
      rc = rigid_forward(c, gr, T + s0, sb, g.nmotion, x, 0, U, true);
      if (rc) return rc;
    }
    Staging st;
    std::vector<double> tf(sb), ta(sb);
    for (int s = 0; s < sb; ++s) {
      tf[s] = -T[s0 + s];  // forward e1 along dim 1: (-1)^1 * t1
      ta[s] = T[s0 + s];   // inverse: (-1)^0 * t1
    }
    const size_t otf = st.add(tf.data(), sizeof(double) * sb);
    const size_t ota = st.add(ta.data(), sizeof(double) * sb);
    const unsigned char* d = nullptr;
    rc = ctx_upload_bytes(c, st, &d);  // stream-ordered after the transform kernels
    if (rc) return rc;
    CoilArgs a = coil_args(s0, U, nv);
    a.tfwd = fwd ? (const double*)(d + otf) - s0 : nullptr;
    a.tadj = adj ? (const double*)(d + ota) - s0 : nullptr;
    rc = run_coil(a, sb, true);
    if (rc) return rc;
    if (adj) {
      rc = rigid_inverse(c, gr, T + s0, sb, g.nmotion, Vb, nv, Vb, xout, true, true, !first_out);
      if (rc) return rc;
      first_out = false;
    }
  }
  // segments without a transform: the coil stage reads x directly
  for (int s0 = NM; s0 < S; s0 += Sb) {
    const int sb = std::min(Sb, S - s0);
    CoilArgs a = coil_args(s0, x, 0);
    rc = run_coil(a, sb, false);
    if (rc) return rc;
    if (adj) {
      DCHK(c, launch_sum_states(Vb, nv, xout, nv, sb, first_out ? 0 : 1, stream));
      first_out = false;
    }
  }
  if (mode == COIL_ENERGY && eo) {
    const int Ns = g.nsamples;
    const unsigned char* d = nullptr;
    if (eo->Mbe) {
      Staging st;
      st.add(eo->Mbe, sizeof(float) * g.n[2]);
      rc = ctx_upload_bytes(c, st, &d);
      if (rc) return rc;
    }
    std::vector<double> h(ns, 0.0);
    if (Ns > 0) {
      DCHK(c, launch_energy_reduce(epart, Ns, g.n[2], eo->Mbe ? (const float*)d : nullptr, 0, g.n[2], esamp, stream));
      DCHK(c, hipMemcpyAsync(h.data(), esamp, sizeof(double) * Ns, hipMemcpyDeviceToHost, stream));
    }
    DCHK(c, hipStreamSynchronize(stream));
    if (eo->per_sample) std::copy(h.begin(), h.begin() + Ns, eo->per_sample);
    if (eo->total) {
      double t = 0;
      for (int j = 0; j < Ns; ++j) t += h[j];
      *eo->total = t;
    }
  }
  return DIS_OK;
}

}  // namespace dis

// ===========================================================================
// Transform gradient (Methods/sincRigidTransformGradient.m:28-147) and the LM
// step (Reconstruction/LMsolver.m:43-225).
// ===========================================================================
namespace dis {

namespace {

// Per-state coefficient arrays of the gradient chain, staged for one upload.
struct GradCoefs {
  size_t tanA[3], sinA[3], tc[3], mcos[3], tB[3], minus1;
  size_t flist[3];
  int nflip[3];
};

GradCoefs stage_grad(const double* T, int S, int Sld, Staging& st) {
  GradCoefs gc{};
  std::vector<double> b(S);
  std::vector<StateParam> sp(S);
  for (int s = 0; s < S; ++s) sp[s] = state_param(T, Sld, s);
  for (int m = 0; m < 3; ++m) {
    for (int s = 0; s < S; ++s) b[s] = sp[s].tan2[m];  // (-1)^(di-1) tan, di = 1
    gc.tanA[m] = st.add(b.data(), sizeof(double) * S);
    for (int s = 0; s < S; ++s) b[s] = -sp[s].sn[m];  // (-1)^di sin
    gc.sinA[m] = st.add(b.data(), sizeof(double) * S);
    for (int s = 0; s < S; ++s) b[s] = 0.5 * (1.0 + sp[s].tan2[m] * sp[s].tan2[m]);  // d tan(th/2)/d th
    gc.tc[m] = st.add(b.data(), sizeof(double) * S);
    for (int s = 0; s < S; ++s) {
      // cos of the wrapped angle: tan(th/2) = t2  ->  cos th = (1 - t2^2) / (1 + t2^2)
      const double t2 = sp[s].tan2[m];
      b[s] = -(1.0 - t2 * t2) / (1.0 + t2 * t2);
    }
    gc.mcos[m] = st.add(b.data(), sizeof(double) * S);
    for (int s = 0; s < S; ++s) b[s] = -sp[s].t[m];  // (-1)^di t
    gc.tB[m] = st.add(b.data(), sizeof(double) * S);
    std::vector<int> lst;
    for (int s = 0; s < S; ++s)
      if (sp[s].flip[m]) lst.push_back(s);
    gc.nflip[m] = (int)lst.size();
    gc.flist[m] = lst.empty() ? 0 : st.add(lst.data(), sizeof(int) * lst.size());
  }
  for (int s = 0; s < S; ++s) b[s] = -1.0;  // etg1 = -i k
  gc.minus1 = st.add(b.data(), sizeof(double) * S);
  return gc;
}

}  // namespace

int rigid_gradient(dis_ctx* c, const Grid& gr, const double* T, int S, int Sld, const float2* x, float2* Tx,
                   float2* G, float2* X0, float2* Y) {
  Staging st;
  GradCoefs gc = stage_grad(T, S, Sld, st);
  const unsigned char* d = nullptr;
  int rc = ctx_upload_bytes(c, st, &d);
  if (rc) return rc;
  hipStream_t stream = ctx_stream(c);
  const long long nv = gr.nvox;
  auto dp = [&](size_t off) { return (const double*)(d + off); };
  auto base = [&](int axis, int src) {
    Step stp{};
    stp.axis = axis;
    stp.src = src;
    stp.flip_m = -1;
    PassArgs a = pass_args(gr, stp, d, S);
    return a;
  };
  auto map = [&](int axis, int src, const double* A, const double* B, const float2* in, long long in_ss,
                 float2* out) -> int {
    PassArgs a = base(axis, src);
    a.coefA = A;
    a.coefB = B;
    a.in = in;
    a.in_sstride = in_ss;
    a.out = out;
    a.out_sstride = nv;
    const int ps = prof_begin(c, DIS_K_PASS);
    DCHK(c, launch_pass(a, MODE_MAP, S, stream));
    prof_end(c, ps, 16.0 * nv * S);
    return DIS_OK;
  };
  auto dual = [&](int axis, int src, const double* A, const double* B, const double* A2, const double* B2,
                  const float2* in, long long in_ss, const float2* in2, float2* out, float2* out2) -> int {
    PassArgs a = base(axis, src);
    a.coefA = A;
    a.coefB = B;
    a.coefA2 = A2;
    a.coefB2 = B2;
    a.in = in;
    a.in_sstride = in_ss;
    a.in2 = in2;
    a.in2_sstride = nv;
    a.out = out;
    a.out_sstride = nv;
    a.out2 = out2;
    a.out2_sstride = nv;
    const int ps = prof_begin(c, DIS_K_PASS);
    DCHK(c, launch_pass(a, MODE_DUAL, S, stream));
    prof_end(c, ps, 8.0 * nv * S * ((in2 ? 2 : 1) + (out ? 1 : 0) + 1));
    return DIS_OK;
  };
  auto flips = [&](int m, float2* buf) -> int {
    if (!gc.nflip[m]) return DIS_OK;
    for (int ax : {kPer0[m], kPer1[m]}) {
      Step f{};
      f.axis = ax;
      f.src = -1;
      f.flip_m = m;
      PassArgs a = pass_args(gr, f, d, S);
      a.states = (const int*)(d + gc.flist[m]);
      a.in = buf;
      a.in_sstride = nv;
      a.out = buf;
      a.out_sstride = nv;
      DCHK(c, launch_pass(a, MODE_FLIP, gc.nflip[m], stream));
    }
    return DIS_OK;
  };
  auto rotation = [&](int m, const float2* in, long long in_ss, float2* out, bool with_t2) -> int {
    const int a = kPer0[m], b = kPer1[m];
    int r = map(a, b, dp(gc.tanA[m]), nullptr, in, in_ss, out);
    if (!r) r = map(b, a, dp(gc.sinA[m]), nullptr, out, nv, out);
    if (!r) r = map(a, b, dp(gc.tanA[m]), with_t2 ? dp(gc.tB[1]) : nullptr, out, nv, out);
    return r;
  };
  auto translate_rest = [&](float2* buf) -> int {  // t3 then t1 (t2 is fused into the last shear)
    int r = map(2, -1, nullptr, dp(gc.tB[2]), buf, nv, buf);
    if (!r) r = map(0, -1, nullptr, dp(gc.tB[0]), buf, nv, buf);
    return r;
  };
  float2* X0m[3] = {X0, X0 + nv * S, X0 + 2 * nv * S};
  float2* Y1 = Y;
  float2* Y2 = Y + nv * S;
  // inputs of the three rotations: X0[0] = flip(x) (x itself when no state flips), X0[1], X0[2]
  const float2* in0 = x;
  long long in0_ss = 0;
  if (gc.nflip[0]) {
    DCHK(c, launch_bcast(x, X0m[0], nv, S, stream));
    if ((rc = flips(0, X0m[0]))) return rc;
    in0 = X0m[0];
    in0_ss = nv;
  }
  if ((rc = rotation(0, in0, in0_ss, X0m[1], false))) return rc;
  if ((rc = flips(1, X0m[1]))) return rc;
  if ((rc = rotation(1, X0m[1], nv, X0m[2], false))) return rc;
  if ((rc = flips(2, X0m[2]))) return rc;
  if ((rc = rotation(2, X0m[2], nv, Tx, true))) return rc;
  if ((rc = translate_rest(Tx))) return rc;
  // translation derivatives: d(T x)/dt_d = F_d^-1 (-i k_d) F_d (T x)   (:29-36)
  for (int dd = 0; dd < 3; ++dd)
    if ((rc = dual(dd, -1, nullptr, nullptr, nullptr, dp(gc.minus1), Tx, nv, nullptr, nullptr, G + dd * nv * S)))
      return rc;
  // rotation derivatives: product rule through the three shears (:39-52), then the
  // later rotations and the translation
  for (int m = 0; m < 3; ++m) {
    const int a = kPer0[m], b = kPer1[m];
    float2* Gm = G + (3 + m) * nv * S;
    const float2* xin = m == 0 ? in0 : X0m[m];
    const long long xin_ss = m == 0 ? in0_ss : nv;
    if ((rc = dual(a, b, dp(gc.tanA[m]), nullptr, dp(gc.tc[m]), nullptr, xin, xin_ss, nullptr, Y1, Y2))) return rc;
    if ((rc = dual(b, a, dp(gc.sinA[m]), nullptr, dp(gc.mcos[m]), nullptr, Y1, nv, Y2, Y1, Y2))) return rc;
    if (m < 2) {
      if ((rc = dual(a, b, dp(gc.tanA[m]), nullptr, dp(gc.tc[m]), nullptr, Y1, nv, Y2, nullptr, Gm))) return rc;
      for (int mm = m + 1; mm < 3; ++mm) {
        if ((rc = flips(mm, Gm))) return rc;
        if ((rc = rotation(mm, Gm, nv, Gm, mm == 2))) return rc;
      }
    } else {
      // the dim-2 translation joins the last product-rule pass (k-space along dim 2)
      if ((rc = dual(a, b, dp(gc.tanA[m]), dp(gc.tB[1]), dp(gc.tc[m]), nullptr, Y1, nv, Y2, nullptr, Gm))) return rc;
    }
    if ((rc = translate_rest(Gm))) return rc;
  }
  return DIS_OK;
}

int run_lm_impl(dis_ctx* c, const OpView& v, double* T, double* w, unsigned char* cT, double winit,
                const double tL[2], const float2* x, const float2* y, double* En_out, int* n_trials,
                double* sums_only);

int run_lm(dis_ctx* c, const OpView& v, double* T, double* w, unsigned char* cT, double winit, const double tL[2],
           const float2* x, const float2* y, double* En_out, int* n_trials) {
  return run_lm_impl(c, v, T, w, cT, winit, tL, x, y, En_out, n_trials, nullptr);
}

int lm_partials(dis_ctx* c, const OpView& v, const double* T, const float2* x, const float2* y, double* sums) {
  std::vector<double> Tc(T, T + 6 * (size_t)v.g.nmotion), w(v.g.nmotion, 1.0), En(v.g.nmotion);
  std::vector<unsigned char> cT(v.g.nmotion, 0);
  const double tL[2] = {0, 0};
  return run_lm_impl(c, v, Tc.data(), w.data(), cT.data(), 1.0, tL, x, y, En.data(), nullptr, sums);
}

// One call of LMsolver for the motion states of the plan.  With sums_only,
// stops after the partial derivatives and returns the raw per-state sums.
int run_lm_impl(dis_ctx* c, const OpView& v, double* T, double* w, unsigned char* cT, double winit,
                const double tL[2], const float2* x, const float2* y, double* En_out, int* n_trials,
                double* sums_only) {
  const dis_geom& g = v.g;
  Grid gr;
  int rc = make_grid(c, g.n, g.Nfull, &gr);
  if (rc) return rc;
  hipStream_t stream = ctx_stream(c);
  const int NM = g.nmotion;
  const long long nv = gr.nvox;
  const long long ny = (long long)g.nsamples * g.n[2] * g.ncoils;
  const double NElY = (double)ny;  // numel(y) before the readout crop (LMsolver.m:32-33)
  // readout ROI E.nF (LMsolver.m:33,68,98): sums over planes [ro_lo, ro_hi)
  const int ro_lo = v.roi[1] > 0 ? v.roi[0] : 0;
  const int ro_hi = v.roi[1] > 0 ? v.roi[1] : g.n[2];
  // fractional finite-difference weights E.Fd (LMsolver.m:88-91,101-104,186-189):
  // residuals, Jacobians and trial residuals are weighted along the ROI's
  // readout DFT before their sums
  const bool filt = v.Fd != nullptr;
  FftPlan fd_plan{};
  const float2* fd_tw = nullptr;
  if (filt) {
    if (v.fold[0] > 0) return ctx_fail(c, DIS_E_UNSUPPORTED, "LMsolver: fractional finite-difference weights on a folded plan");
    if (v.fdn != ro_hi - ro_lo)
      return ctx_fail(c, DIS_E_INVALID, "LMsolver: the fractional finite-difference weights do not match the readout ROI");
    if (make_plan(v.fdn, &fd_plan) != DIS_OK) return ctx_fail(c, DIS_E_UNSUPPORTED, "LMsolver: readout ROI above 1024 planes");
    rc = ctx_twiddles(c, v.fdn, &fd_tw);
    if (rc) return rc;
  }
  {
    const int mc = lm_check_motion(NM, T, w);
    if (mc) return ctx_fail(c, DIS_E_INVALID, lm_status_text(mc));
  }
  std::vector<int> est;
  for (int s = 0; s < NM; ++s)
    if (!cT[s]) est.push_back(s);
  std::vector<double> EnPrev(NM, 0.0), sums((size_t)NM * kLmSums, 0.0);
  if (n_trials) *n_trials = 0;
  if (est.empty()) {
    for (int s = 0; s < NM; ++s) En_out[s] = 0.0;
    return DIS_OK;
  }
  // y-space buffers (slot 2): Yr, J[6]; partials; per-sample energies
  const int nb = lm_partial_count();
  const size_t ybytes = sizeof(float2) * ny;
  const size_t pbytes = sizeof(double) * (size_t)NM * nb * kLmSums;
  const size_t sbytes = sizeof(double) * (size_t)NM * kLmSums;
  const size_t ebytes = sizeof(float) * (size_t)std::max(1, g.nsamples) * g.n[2];
  const size_t esbytes = sizeof(double) * (size_t)std::max(1, g.nsamples);
  const size_t stepbytes = sizeof(double) * (size_t)NM * 8;
  void* yb = nullptr;
  rc = ctx_buffer(c, 2, 7 * ybytes + pbytes + sbytes + ebytes + esbytes + stepbytes + 4096, &yb);
  if (rc) return rc;
  float2* Yr = (float2*)yb;
  float2* J[6];
  for (int gi = 0; gi < 6; ++gi) J[gi] = Yr + (gi + 1) * ny;
  double* partial = (double*)((char*)yb + 7 * ybytes);
  double* dsums = (double*)((char*)partial + pbytes);
  float* epart = (float*)((char*)dsums + sbytes);
  double* esamp = (double*)((char*)epart + ebytes);
  double* dstep = (double*)((char*)esamp + esbytes);
  // volume workspace (slot 0): per batch state Tx, G[6], X0[3], Y[2] = 12 volumes
  const size_t vol = sizeof(float2) * nv;
  int Sb = (int)est.size();
  const size_t lim = ctx_workspace_limit(c);
  if (lim) {
    if (lim < 12 * vol) return ctx_fail(c, DIS_E_NOMEM, "workspace limit below the 12 volumes of one LM state");
    Sb = (int)std::min<size_t>((size_t)Sb, lim / (12 * vol));
  }
  void* ws = nullptr;
  rc = ctx_workspace(c, 12 * vol * (size_t)Sb, &ws);
  if (rc) return rc;
  float2* Tx = (float2*)ws;
  float2* G = Tx + nv * Sb;
  float2* X0 = G + 6 * nv * Sb;
  float2* Yv = X0 + 3 * nv * Sb;
  const int G16 = std::max(1, std::min(g.n[1], kMaxTile / g.n[0]));
  CoilArgs fold_a{};
  {
    int maxnec = 0;
    for (int s = 0; s < g.nstates; ++s) maxnec = std::max(maxnec, v.hoff[s + 1] - v.hoff[s]);
    rc = fold_setup(c, v, maxnec, &fold_a);
    if (rc) return rc;
  }
  int maxnec_all = 0;
  for (int s = 0; s < g.nstates; ++s) maxnec_all = std::max(maxnec_all, v.hoff[s + 1] - v.hoff[s]);
  auto readout_filter = [&](float2* Y, const float2* ysub, float* ep, const int* dlist) {
    ReadoutFilterArgs fa{};
    fa.Yr = Y;
    fa.y = ysub;
    fa.Fd = v.Fd;
    fa.epart = ep;
    fa.off = v.off;
    fa.slist = dlist;
    fa.Ns = g.nsamples;
    fa.n3 = g.n[2];
    fa.C = g.ncoils;
    fa.lo = ro_lo;
    fa.nF = ro_hi - ro_lo;
    fa.tw = fd_tw;
    fa.plan = fd_plan;
    return fa;
  };
  auto coil_args = [&](const int* dlist, const float2* U, int mode) {
    CoilArgs a = fold_a;
    a.U = U;
    a.us = nv;
    a.us_rel = 1;
    a.V = nullptr;
    a.Smaps = v.S;
    a.n1 = g.n[0];
    a.n2 = g.n[1];
    a.n3 = g.n[2];
    a.C = g.ncoils;
    a.k1 = v.k1;
    a.k2 = v.k2;
    a.off = v.off;
    a.We = nullptr;  // LMsolver encodes without EH.We (LMsolver.m:72,99)
    a.rowptr = v.rowptr;
    a.csr = v.csr;
    a.sord = v.sord;
    a.adjtab = v.adjtab;
    a.adjoff = v.adjoff;
    a.Ns = g.nsamples;
    a.yin = y;
    a.yout = nullptr;
    a.epart = epart;
    a.invN1 = 1.0 / (double)g.Nfull[0];
    a.tw1 = gr.tw[0];
    a.tw2 = gr.tw[1];
    a.plan1 = gr.plan[0];
    a.G = G16;
    a.slist = dlist;
    a.i3lo = ro_lo;
    a.i3n = ro_hi - ro_lo;
    (void)mode;
    return a;
  };
  // ---- partial derivatives of the estimated states (LMsolver.m:52-122)
  for (size_t b0 = 0; b0 < est.size(); b0 += Sb) {
    const int sb = (int)std::min<size_t>(Sb, est.size() - b0);
    std::vector<double> Tb((size_t)sb * 6);
    std::vector<int> lst(sb);
    for (int i = 0; i < sb; ++i) {
      lst[i] = est[b0 + i];
      for (int gi = 0; gi < 6; ++gi) Tb[i + (size_t)sb * gi] = T[lst[i] + (size_t)NM * gi];
    }
    rc = rigid_gradient(c, gr, Tb.data(), sb, sb, x, Tx, G, X0, Yv);
    if (rc) return rc;
    Staging st;
    const size_t ol = st.add(lst.data(), sizeof(int) * sb);
    const unsigned char* d = nullptr;
    rc = ctx_upload_bytes(c, st, &d);
    if (rc) return rc;
    const int* dlist = (const int*)(d + ol);
    CoilArgs a = coil_args(dlist, Tx, COIL_ENCODE);
    a.yout = Yr;
    DCHK(c, launch_coil(a, COIL_ENCODE, sb, stream));
    for (int gi = 0; gi < 6; ++gi) {
      CoilArgs ag = coil_args(dlist, G + (long long)gi * nv * sb, COIL_ENCODE);
      ag.yout = J[gi];
      DCHK(c, launch_coil(ag, COIL_ENCODE, sb, stream));
    }
    if (filt) {  // residual and Jacobians through fftGPU(., 3) .* E.Fd (LMsolver.m:88-91, 101-104)
      for (int q = 0; q < 7; ++q) {
        ReadoutFilterArgs fa = readout_filter(q == 0 ? Yr : J[q - 1], q == 0 ? y : nullptr, nullptr, dlist);
        DCHK(c, launch_readout_filter(fa, sb, maxnec_all, stream));
      }
    }
    LmReduceArgs ra{};
    ra.Yr = Yr;
    ra.y = filt ? nullptr : y;
    for (int gi = 0; gi < 6; ++gi) ra.J[gi] = J[gi];
    ra.off = v.off;
    ra.slist = dlist;
    ra.Ns = g.nsamples;
    ra.n3 = g.n[2];
    ra.C = g.ncoils;
    ra.i3lo = ro_lo;
    ra.i3n = ro_hi - ro_lo;
    ra.partial = partial;
    DCHK(c, launch_lm_reduce(ra, sb, partial, dsums, stream));
    std::vector<double> h((size_t)sb * kLmSums);
    DCHK(c, hipMemcpyAsync(h.data(), dsums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    DCHK(c, hipStreamSynchronize(stream));
    for (int i = 0; i < sb; ++i)
      for (int q = 0; q < kLmSums; ++q) sums[(size_t)lst[i] * kLmSums + q] = h[(size_t)i * kLmSums + q];
  }
  if (sums_only) {
    std::copy(sums.begin(), sums.end(), sums_only);
    return DIS_OK;
  }
  LmUpdate u;
  u.NM = NM;
  u.est = est;
  u.EnPrev = EnPrev;
  for (int s : est) u.EnPrev[s] = sums[(size_t)s * kLmSums] / NElY;
  // damped 6x6 solves (FP64, on the device) for the listed states
  auto solve = [&](const std::vector<int>& sol, std::vector<double>& steps) -> int {
    const int ns = (int)sol.size();
    std::vector<double> hs((size_t)ns * kLmSums), hw(ns);
    for (int i = 0; i < ns; ++i) {
      for (int q = 0; q < kLmSums; ++q) hs[(size_t)i * kLmSums + q] = sums[(size_t)sol[i] * kLmSums + q];
      hw[i] = w[sol[i]];
    }
    Staging st;
    const size_t os = st.add(hs.data(), sizeof(double) * hs.size());
    const size_t ow = st.add(hw.data(), sizeof(double) * ns);
    const unsigned char* d = nullptr;
    int r = ctx_upload_bytes(c, st, &d);
    if (r) return r;
    DCHK(c, launch_lm_solve((const double*)(d + os), (const double*)(d + ow), ns, winit, dstep, stream));
    DCHK(c, hipMemcpyAsync(steps.data(), dstep, sizeof(double) * (size_t)ns * 6, hipMemcpyDeviceToHost, stream));
    DCHK(c, hipStreamSynchronize(stream));
    return DIS_OK;
  };
  // trial energies at Tup of the listed states
  auto trial = [&](const std::vector<int>& todo, const std::vector<double>& Tup, std::vector<double>& Et) -> int {
    for (size_t b0 = 0; b0 < todo.size(); b0 += Sb) {
      const int sb = (int)std::min<size_t>(Sb, todo.size() - b0);
      std::vector<double> Tb((size_t)sb * 6);
      std::vector<int> lst(sb);
      for (int i = 0; i < sb; ++i) {
        lst[i] = todo[b0 + i];
        for (int gi = 0; gi < 6; ++gi) Tb[i + (size_t)sb * gi] = Tup[lst[i] + (size_t)NM * gi];
      }
      int r = rigid_forward(c, gr, Tb.data(), sb, sb, x, 0, Tx, false);
      if (r) return r;
      Staging st;
      const size_t ol = st.add(lst.data(), sizeof(int) * sb);
      const unsigned char* d = nullptr;
      r = ctx_upload_bytes(c, st, &d);
      if (r) return r;
      if (filt) {  // encode, then the weighted residual's energies (LMsolver.m:186-189)
        CoilArgs a = coil_args((const int*)(d + ol), Tx, COIL_ENCODE);
        a.yout = Yr;
        DCHK(c, launch_coil(a, COIL_ENCODE, sb, stream));
        ReadoutFilterArgs fa = readout_filter(Yr, y, epart, (const int*)(d + ol));
        DCHK(c, launch_readout_filter(fa, sb, maxnec_all, stream));
      } else {
        CoilArgs a = coil_args((const int*)(d + ol), Tx, COIL_ENERGY);
        DCHK(c, launch_coil(a, COIL_ENERGY, sb, stream));
      }
      DCHK(c, launch_energy_reduce(epart, g.nsamples, g.n[2], nullptr, ro_lo, ro_hi, esamp, stream));
      std::vector<double> he(std::max(1, g.nsamples));
      DCHK(c, hipMemcpyAsync(he.data(), esamp, sizeof(double) * g.nsamples, hipMemcpyDeviceToHost, stream));
      DCHK(c, hipStreamSynchronize(stream));
      for (int s : lst) {
        double e = 0;
        for (int j = v.hoff[s]; j < v.hoff[s + 1]; ++j) e += he[j];
        Et[s] = e / NElY;
      }
    }
    return DIS_OK;
  };
  int cb_rc = DIS_OK;
  const int lst = lm_update(u, T, w, cT, tL, solve, trial, n_trials, &cb_rc);
  if (lst == LM_CALLBACK) return cb_rc;
  if (lst) return ctx_fail(c, DIS_E_INVALID, lm_status_text(lst));
  for (int s = 0; s < NM; ++s) En_out[s] = std::min(u.EnPrev[s], u.En[s]);
  return DIS_OK;
}


int autotune_grid(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], int S, int32_t* out, double* ms_out) {
  hipStream_t stream = ctx_stream(c);
  const long long nv = (long long)n[0] * n[1] * n[2];
  void* ws = nullptr;
  int rc = ctx_workspace(c, sizeof(float2) * nv * S * 2, &ws);
  if (rc) return rc;
  DCHK(c, hipMemsetAsync(ws, 0, sizeof(float2) * nv * S * 2, stream));
  float2* in = (float2*)ws;
  float2* outv = in + nv * S;
  hipEvent_t e0, e1;
  DCHK(c, hipEventCreate(&e0));
  DCHK(c, hipEventCreate(&e1));
  const bool prof = ctx_profiling(c, 0);
  const int nn[3] = {n[0], n[1], n[2]};
  for (int axis = 0; axis < 3; ++axis) {
    const int len = n[axis];
    int Pmax = 1;
    while (Pmax * 2 * len <= kMaxTile && Pmax * 2 <= 512) Pmax *= 2;
    if (axis != 0) Pmax = std::min(Pmax, n[0]);
    std::vector<PassTune> cands;
    if (reg_radix(len) && (axis == 0 || n[0] >= 16)) {
      PassTune t;
      t.valid = 1;
      cands.push_back(t);  // register path
    }
    for (int mr : {4, 8, 16})
      for (int P = Pmax; P >= 1; P /= 2) {
        PassTune t;
        t.valid = 1;
        t.generic = 1;
        t.P = P;
        t.maxradix = mr;
        cands.push_back(t);
        if (P <= Pmax / 16) break;  // tiles below 1/16 of the largest are never faster
      }
    double best = 1e30;
    PassTune bt;
    for (const PassTune& t : cands) {
      ctx_tune_set(c, nn, axis, t);
      Grid gr;
      rc = make_grid(c, n, Nfull, &gr);
      if (rc) break;
      Step stp{};
      stp.axis = axis;
      stp.src = (axis + 1) % 3;
      stp.flip_m = -1;
      PassArgs a = pass_args(gr, stp, nullptr, S);
      a.in = in;
      a.in_sstride = nv;
      a.out = outv;
      a.out_sstride = nv;
      DCHK(c, launch_pass(a, MODE_MAP, S, stream));  // warm-up
      DCHK(c, hipEventRecord(e0, stream));
      const int reps = 3;
      for (int r = 0; r < reps; ++r) DCHK(c, launch_pass(a, MODE_MAP, S, stream));
      DCHK(c, hipEventRecord(e1, stream));
      DCHK(c, hipEventSynchronize(e1));
      float ms = 0;
      DCHK(c, hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best) {
        best = ms / reps;
        bt = t;
      }
    }
    if (rc) break;
    ctx_tune_set(c, nn, axis, bt);
    if (out) {
      out[3 * axis] = bt.generic;
      out[3 * axis + 1] = bt.generic ? bt.P : 16;
      out[3 * axis + 2] = bt.generic ? bt.maxradix : reg_radix(len);
    }
    if (ms_out) ms_out[axis] = best;
  }
  ctx_profiling(c, prof ? 1 : 0);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace dis
