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
           double lam, float2* x, bool x_init, int nIt, int tolType, double tol, double* En, int* nEn, int* nn) {
  if (tolType != 1 && tolType != 6 && tolType != 7)
    return ctx_fail(c, DIS_E_UNSUPPORTED, "CG tolType must be RelativeResidual2Error(1), Energy(6) or None(7)");
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

  // b = C(decode(y)) (CGsolver.m:46) into r
  rc = op_apply(c, v, T, nullptr, y, nullptr, r, COIL_DECODE, nullptr);
  if (rc) return rc;
  if (Ma) CCHK(c, cg_mask(r, Ma, n, st));
  const bool need_b = tolType == 1;
  if (x_init) {
    // EHE = C(encodedecode(x) + R x) (:55)
    rc = op_apply(c, v, T, x, nullptr, nullptr, q, COIL_EHE, nullptr);
    if (rc) return rc;
    CCHK(c, cg_constrain_dot(q, x, Ma, flam, n, part, &sc->dot, st));
    CCHK(c, cg_residual(r, q, r, z, p, Se, n, part, need_b ? part2 : nullptr, &sc->rsold, &sc->bnorm2, st));
  } else {
    CCHK(c, hipMemsetAsync(x, 0, vb, st));
    CCHK(c, cg_residual(r, nullptr, r, z, p, Se, n, part, need_b ? part2 : nullptr, &sc->rsold, &sc->bnorm2, st));
  }
  CgScalars h;
  CCHK(c, hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
  CCHK(c, hipStreamSynchronize(st));
  if (std::hypot(h.rsold.x, h.rsold.y) < 1e-6) {  // early termination (:70-75)
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
    if (rc) return rc;
    CCHK(c, cg_constrain_dot(q, p, Ma, flam, n, part, &sc->pq, st));
    CCHK(c, cg_step(sc, st));
    CCHK(c, cg_update(x, r, z, p, q, Se, sc, n, part, part2, st));
    CCHK(c, cg_direction(p, z, sc, n, st));
    if (tolType == 6 && it % 2 == 1) {
      double e2[2];
      rc = energy(e2);
      if (rc) return rc;
      Ehist.push_back(e2[0]);
      Ehist.push_back(e2[1]);
    }
    CCHK(c, hipMemcpyAsync(&h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
    CCHK(c, hipStreamSynchronize(st));
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
