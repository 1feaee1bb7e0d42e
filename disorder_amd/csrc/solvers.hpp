// Device scalars and launchers of the CG vector kernels (FP64 reductions).
#pragma once
#include <hip/hip_runtime.h>

namespace dis {

// CG scalars kept on the device between kernels (complex double).
struct CgScalars {
  double2 rsold;   // sum conj(z) r of the previous iteration
  double2 pq;      // sum conj(p) EHE p
  double2 g;       // step rsold / pq
  double2 rs;      // sum conj(z) r
  double2 d;       // rs / (rsold + eps)
  double2 zr_abs;  // sum |conj(z) r|  (RelativeResidual2Error)
  double2 bnorm2;  // sum |conj(P b) b|
  double2 dot;     // scratch result
  double2 fid;     // sum |E x - y|^2 from the k-space iterate (x: the fidelity energy)
};

constexpr int kCgPartials = 1024;

hipError_t cg_dot(const float2* a, const float2* b, long long n, double2* partial, double2* out, hipStream_t st);
// Ex += g Ep with the step g of the x update (sc->g, in FP32 like x); with
// energy, sc->fid.x = sum |Ex - y|^2 in FP64 (fixed-order reduction).
hipError_t cg_kspace_update(float2* Ex, const float2* Ep, const float2* y, CgScalars* sc, long long n, bool energy,
                            double2* partial, hipStream_t st);
hipError_t cg_mask(float2* v, const float* Ma, long long n, hipStream_t st);
// out->x = max_i |a_i|^2
hipError_t cg_maxabs2(const float2* a, long long n, double2* partial, double2* out, hipStream_t st);
hipError_t cg_constrain_dot(float2* q, const float2* p, const float* Ma, float lam, long long n, double2* partial,
                            double2* out, hipStream_t st);
hipError_t cg_residual(const float2* b, const float2* q, float2* r, float2* z, float2* p, const float* Se, long long n,
                       double2* partial, double2* partial_b, double2* rs_out, double2* bn_out, hipStream_t st);
hipError_t cg_step(CgScalars* sc, hipStream_t st);
hipError_t cg_step_guarded(CgScalars* sc, hipStream_t st);
// hist[*cnt] = zr_abs; ++*cnt  (device-side residual history)
hipError_t cg_record(const CgScalars* sc, double* hist, int* cnt, hipStream_t st);
hipError_t cg_update(float2* x, float2* r, float2* z, const float2* p, const float2* q, const float* Se, CgScalars* sc,
                     long long n, double2* partial, double2* partial_abs, hipStream_t st);
hipError_t cg_direction(float2* p, const float2* z, const CgScalars* sc, long long n, hipStream_t st);

}  // namespace dis
