// Host loop of the preconditioned CG of Reconstruction/CGsolver.m:45-109.
// Vectors stay on the device; the FP64 scalars are computed by device
// kernels and read back once per iteration for the stopping tests
// (stopCondition.m: 'RelativeResidual2Error' = 1, 'Energy' = 6, 'None' = 7).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "solvers.hpp"
#include "transform.hpp"

namespace dis {

#define CCHK(c, expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return ctx_fail((c), DIS_E_HIP, hipGetErrorString(e_)); \
  } while (0)

int run_cg(dis_ctx* c, const OpView& v, const double* T, const float2* y, const float* Se, const float* Ma,
           double lam, float2* x, bool x_init, int nIt, int tolType, double tol, double* En, int* nEn, int* nn,
           dis_shearlet* sh, double sh_scale) {
  // tolType 8: Simulation/Xsolver.m:53-75 -- stop when max |alpha p|^2 < tol, or when sqrt|rs| < 1e-10,
  // or after nIt iterations (the caller turns Xsolver's operation budget into nIt)
  if (tolType != 1 && tolType != 6 && tolType != 7 && tolType != 8)
    return ctx_fail(c, DIS_E_UNSUPPORTED,
                    "CG tolType must be RelativeResidual2Error(1), Energy(6), None(7) or MaxUpdate(8)");
  if (sh && tolType == 6) return ctx_fail(c, DIS_E_UNSUPPORTED, "CG with the shearlet term: tolType 'Energy' not supported");
  hipStream_t st = ctx_stream(c);
  const long long n = (long long)v.g.n[0] * v.g.n[1] * v.g.n[2];
  const size_t vb = sizeof(float2) * n;
  void* base = nullptr;
  const size_t scal_off = 4 * vb;
  const size_t part_off = scal_off + 256;
  const size_t bytes = part_off + 2 * sizeof(double2) * kCgPartials;
  int rc = ctx_buffer(c, 1, bytes, &base);
  if (rc) return rc;
  float2* r = (float2*)base;
  float2* z = r + n;
  float2* p = z + n;
  float2* q = p + n;
  CgScalars* sc = (CgScalars*)((char*)base + scal_off);
  double2* part = (double2*)((char*)base + part_off);
  double2* part2 = part + kCgPartials;
  const float flam = (float)lam;
  auto energy = [&](double* out2) -> int {
    EnergyOut eo;
    double fid = 0;
    eo.total = &fid;
    int r2 = op_apply(c, v, T, x, y, nullptr, nullptr, COIL_ENERGY, &eo);
    if (r2) return r2;
    double reg = 0;
    if (lam != 0) {
      CCHK(c, cg_dot(x, x, n, part, &sc->dot, st));
      double2 h;
      CCHK(c, hipMemcpyAsync(&h, &sc->dot, sizeof(h), hipMemcpyDeviceToHost, st));
      CCHK(c, hipStreamSynchronize(st));
      reg = std::fabs(lam * h.x);
    }
    out2[0] = fid;
    out2[1] = reg;
    return DIS_OK;
  };

  // 'Energy' (CGsolver.m:87, computeEnergy of every odd iterate): E x of the
  // iterate is kept in k-space, E x_{k+1} = E x_k + g_k E p_k with E p_k the
  // forward half that E^H E leaves in its k-space scratch, so the fidelity
  // energy sum |E x - y|^2 is a reduction instead of another forward encode
  // (operators whose E^H E keeps no k-space scratch re-encode)
  float2* Ep = tolType == 6 ? ehe_kspace(c, v) : nullptr;
  float2* Ex = nullptr;
  const long long ny = (long long)v.g.nsamples * v.g.n[2] * v.g.ncoils;
  if (Ep && ny > 0) {
    void* eb = nullptr;
    rc = ctx_buffer(c, 9, sizeof(float2) * ny, &eb);
    if (rc) return rc;
    Ex = (float2*)eb;
  } else {
    Ep = nullptr;
  }
  auto kenergy = [&](double* out2, double fid) -> int {
    double reg = 0;
    if (lam != 0) {
      CCHK(c, cg_dot(x, x, n, part, &sc->dot, st));
      double2 h;
      CCHK(c, hipMemcpyAsync(&h, &sc->dot, sizeof(h), hipMemcpyDeviceToHost, st));
      CCHK(c, hipStreamSynchronize(st));
      reg = std::fabs(lam * h.x);
    }
    out2[0] = fid;
    out2[1] = reg;
    return DIS_OK;
  };

  // b = C(decode(y)) (CGsolver.m:46) into r
  rc = op_apply(c, v, T, nullptr, y, nullptr, r, COIL_DECODE, nullptr);
  if (rc) return rc;
  if (Ma) CCHK(c, cg_mask(r, Ma, n, st));
  const bool need_b = tolType == 1;
  if (x_init) {
    // EHE = C(encodedecode(x) + R x) (:55)
    rc = op_apply(c, v, T, x, nullptr, nullptr, q, COIL_EHE, nullptr);
    if (!rc && sh) rc = shearlet_apply(c, sh, x, sh_scale, q);
    if (rc) return rc;
    if (Ex) CCHK(c, hipMemcpyAsync(Ex, Ep, sizeof(float2) * ny, hipMemcpyDeviceToDevice, st));  // E x_0
    CCHK(c, cg_constrain_dot(q, x, Ma, flam, n, part, &sc->dot, st));
    CCHK(c, cg_residual(r, q, r, z, p, Se, n, part, need_b ? part2 : nullptr, &sc->rsold, &sc->bnorm2, st));
  } else {
    CCHK(c, hipMemsetAsync(x, 0, vb, st));
    if (Ex) CCHK(c, hipMemsetAsync(Ex, 0, sizeof(float2) * ny, st));
    CCHK(c, cg_residual(r, nullptr, r, z, p, Se, n, part, need_b ? part2 : nullptr, &sc->rsold, &sc->bnorm2, st));
  }
  CgScalars h;
  CCHK(c, hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
  CCHK(c, hipStreamSynchronize(st));
  // early termination (:70-75); Xsolver only stops on sqrt|rs| < 1e-10
  if (std::hypot(h.rsold.x, h.rsold.y) < (tolType == 8 ? 1e-20 : 1e-6)) {
    if (nn) *nn = 0;
    if (nEn) *nEn = 0;
    return DIS_OK;
  }
  const double bnorm2 = h.bnorm2.x;
  double err = INFINITY;
  int it = 0;
  std::vector<double> Ehist;  // [fid, reg] of odd iterations
  for (it = 1; it <= nIt; ++it) {
    rc = op_apply(c, v, T, p, nullptr, nullptr, q, COIL_EHE, nullptr);
    if (!rc && sh) rc = shearlet_apply(c, sh, p, sh_scale, q);
    if (rc) return rc;
    CCHK(c, cg_constrain_dot(q, p, Ma, flam, n, part, &sc->pq, st));
    CCHK(c, cg_step(sc, st));
    CCHK(c, cg_update(x, r, z, p, q, Se, sc, n, part, part2, st));
    if (Ex) CCHK(c, cg_kspace_update(Ex, Ep, y, sc, ny, it % 2 == 1, part, st));
    if (tolType == 8) CCHK(c, cg_maxabs2(p, n, part, &sc->dot, st));  // before p changes
    CCHK(c, cg_direction(p, z, sc, n, st));
    if (tolType == 6 && it % 2 == 1 && !Ex) {
      double e2[2];
      rc = energy(e2);
      if (rc) return rc;
      Ehist.push_back(e2[0]);
      Ehist.push_back(e2[1]);
    }
    CCHK(c, hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
    CCHK(c, hipStreamSynchronize(st));
    if (tolType == 6 && it % 2 == 1 && Ex) {
      double e2[2];
      rc = kenergy(e2, h.fid.x);
      if (rc) return rc;
      Ehist.push_back(e2[0]);
      Ehist.push_back(e2[1]);
    }
    if (tolType == 8) {
      const double g2 = h.g.x * h.g.x + h.g.y * h.g.y;
      if (g2 * h.dot.x < tol) break;                 // max |alpha p|^2 < toler
      if (std::sqrt(std::hypot(h.rs.x, h.rs.y)) < 1e-10) break;
      continue;
    }
    if (tolType == 1) {
      err = std::sqrt(h.zr_abs.x / bnorm2);
    } else if (tolType == 6 && it % 2 == 1 && it > 1) {
      const size_t k = Ehist.size();
      const double e0 = Ehist[k - 4] + Ehist[k - 3];
      const double e1 = Ehist[k - 2] + Ehist[k - 1];
      err = (e0 - e1) / e0;
    }
    if (err < tol || std::hypot(h.rs.x, h.rs.y) < 1e-6) break;
  }
  if (it > nIt) it = nIt;
  if (nn) *nn = it;
  const int ne = (int)(Ehist.size() / 2);
  if (nEn) *nEn = ne;
  if (En)
    for (size_t k = 0; k < Ehist.size(); ++k) En[k] = Ehist[k];
  return DIS_OK;
}

}  // namespace dis

namespace dis {

// Per-volume CG state of the batched solver (device).
struct CgVol {
  float2 *r, *z, *p, *q;
  CgScalars* sc;
  double2 *part, *part2;
  double* hist;
  int* cnt;
  long long n;
  bool active;
  void* base;
};

namespace {

// Everything the batched solver holds on the device or in the context,
// released on every return path (also after a failed capture).
struct BatchedScope {
  dis_ctx* c;
  hipStream_t st;
  std::vector<CgVol> vs;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  int prof_restore = -1;
  BatchedScope(dis_ctx* c_, int nvol) : c(c_), st(ctx_stream(c_)), vs(nvol) {}
  ~BatchedScope() {
    (void)hipStreamSynchronize(st);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    for (auto& v : vs)
      if (v.base) (void)hipFree(v.base);
    ctx_tape_free(c);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (prof_restore >= 0) ctx_profiling(c, prof_restore);
  }
};

}  // namespace

int run_cg_batched(dis_ctx* c, int nvol, const OpView* views, const double* const* T, const float2* const* y,
                   const float* const* Se, float2* const* x, int nIt, bool use_graph, double* res, double* ms) {
  BatchedScope sc(c, nvol);
  hipStream_t st = sc.st;
  std::vector<CgVol>& vs = sc.vs;
  for (auto& e : sc.ev) CCHK(c, hipEventCreate(&e));
  hipEvent_t e0 = sc.ev[0], e1 = sc.ev[1], e1b = sc.ev[2], e2 = sc.ev[3];
  CCHK(c, hipEventRecord(e0, st));
  // setup: b = decode(y), x = 0, r = b, z = Se r, p = z (CGsolver.m:46-68)
  for (int i = 0; i < nvol; ++i) {
    CgVol& v = vs[i];
    v.n = (long long)views[i].g.n[0] * views[i].g.n[1] * views[i].g.n[2];
    const size_t vb = sizeof(float2) * v.n;
    const size_t scal = 4 * vb, part = scal + 256, hist = part + 2 * sizeof(double2) * kCgPartials;
    const size_t cnt = hist + sizeof(double) * (nIt + 1), bytes = cnt + 256;
    if (hipMalloc(&v.base, bytes) != hipSuccess) {
      v.base = nullptr;
      (void)hipGetLastError();
      return ctx_fail(c, DIS_E_NOMEM, "batched CG: state allocation failed");
    }
    v.r = (float2*)v.base;
    v.z = v.r + v.n;
    v.p = v.z + v.n;
    v.q = v.p + v.n;
    v.sc = (CgScalars*)((char*)v.base + scal);
    v.part = (double2*)((char*)v.base + part);
    v.part2 = v.part + kCgPartials;
    v.hist = (double*)((char*)v.base + hist);
    v.cnt = (int*)((char*)v.base + cnt);
    int rc = op_apply(c, views[i], T[i], nullptr, y[i], nullptr, v.r, COIL_DECODE, nullptr);
    if (rc) return rc;
    CCHK(c, hipMemsetAsync(x[i], 0, vb, st));
    CCHK(c, hipMemsetAsync(v.cnt, 0, sizeof(int), st));
    CCHK(c, cg_residual(v.r, nullptr, v.r, v.z, v.p, Se ? Se[i] : nullptr, v.n, v.part, v.part2, &v.sc->rsold,
                        &v.sc->bnorm2, st));
  }
  std::vector<CgScalars> hs(nvol);
  for (int i = 0; i < nvol; ++i) CCHK(c, hipMemcpyAsync(&hs[i], vs[i].sc, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
  CCHK(c, hipStreamSynchronize(st));
  int nact = 0;
  for (int i = 0; i < nvol; ++i) {
    vs[i].active = std::hypot(hs[i].rsold.x, hs[i].rsold.y) >= 1e-6;  // early termination (:70-75)
    nact += vs[i].active;
  }
  CCHK(c, hipEventRecord(e1, st));
  // one iteration of every active volume (CGsolver.m:80-97)
  auto iteration = [&]() -> int {
    for (int i = 0; i < nvol; ++i) {
      CgVol& v = vs[i];
      if (!v.active) continue;
      int rc = op_apply(c, views[i], T[i], v.p, nullptr, nullptr, v.q, COIL_EHE, nullptr);
      if (rc) return rc;
      CCHK(c, cg_constrain_dot(v.q, v.p, nullptr, 0.f, v.n, v.part, &v.sc->pq, st));
      CCHK(c, cg_step_guarded(v.sc, st));
      CCHK(c, cg_update(x[i], v.r, v.z, v.p, v.q, Se ? Se[i] : nullptr, v.sc, v.n, v.part, v.part2, st));
      CCHK(c, cg_direction(v.p, v.z, v.sc, v.n, st));
      CCHK(c, cg_record(v.sc, v.hist, v.cnt, st));
    }
    return DIS_OK;
  };
  if (nact > 0 && nIt > 0) {
    // iteration 1 records every coefficient upload on the tape
    ctx_tape_mode(c, 1);
    int rc = iteration();
    if (rc) return rc;
    CCHK(c, hipStreamSynchronize(st));
    CCHK(c, hipEventRecord(e1b, st));
    if (nIt > 1) {
      sc.prof_restore = ctx_profiling(c, 0) ? 1 : 0;
      ctx_tape_mode(c, 2);
      if (use_graph) {
        CCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = iteration();
        const hipError_t ec = hipStreamEndCapture(st, &sc.g);
        if (rc) return rc;
        if (ec != hipSuccess) return ctx_fail(c, DIS_E_HIP, "batched CG: graph capture failed");
        if (hipGraphInstantiate(&sc.ge, sc.g, nullptr, nullptr, 0) != hipSuccess)
          return ctx_fail(c, DIS_E_HIP, "batched CG: graph instantiation failed");
        for (int k = 2; k <= nIt; ++k)
          if (hipGraphLaunch(sc.ge, st) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "batched CG: graph launch failed");
      } else {
        for (int k = 2; k <= nIt; ++k) {
          ctx_tape_mode(c, 2);
          rc = iteration();
          if (rc) return rc;
        }
      }
    }
  } else {
    CCHK(c, hipEventRecord(e1b, st));
  }
  CCHK(c, hipEventRecord(e2, st));
  CCHK(c, hipStreamSynchronize(st));
  if (ms) {
    float a = 0, b = 0, d = 0;
    CCHK(c, hipEventElapsedTime(&a, e0, e1));
    CCHK(c, hipEventElapsedTime(&b, e1, e1b));
    CCHK(c, hipEventElapsedTime(&d, e1b, e2));
    ms[0] = a;
    ms[1] = b;
    ms[2] = d;
  }
  if (res) {
    for (int i = 0; i < nvol; ++i) {
      std::vector<double> h(nIt, 0.0);
      int cnt = 0;
      CCHK(c, hipMemcpy(&cnt, vs[i].cnt, sizeof(int), hipMemcpyDeviceToHost));
      if (cnt > 0) CCHK(c, hipMemcpy(h.data(), vs[i].hist, sizeof(double) * std::min(cnt, nIt), hipMemcpyDeviceToHost));
      const double bn = hs[i].bnorm2.x;
      for (int k = 0; k < nIt; ++k) res[(size_t)i * nIt + k] = (k < cnt && bn > 0) ? std::sqrt(h[k] / bn) : 0.0;
    }
  }
  return DIS_OK;
}

}  // namespace dis
