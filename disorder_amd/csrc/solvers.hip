// Vector kernels of the preconditioned CG of Reconstruction/CGsolver.m:64-104
// and the deterministic FP64 reductions behind its scalars.
//
// Every inner product accumulates FP32 complex products in FP64: each thread
// sums a grid-stride slice, a wavefront reduces by shuffles, LDS combines the
// waves, and a single-block kernel adds the per-block partials in a fixed
// order (no atomics, so results are bit-reproducible).
#include "kernels.hpp"
#include "solvers.hpp"

namespace dis {

namespace {

constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block reduction of (a, b) into partial[blockIdx.x]
__device__ __forceinline__ void block_store(double a, double b, double2* partial) {
  __shared__ double sa[kThreads / 64], sb[kThreads / 64];
  __syncthreads();  // the previous call's reader is done with sa/sb
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int i = 0; i < kThreads / 64; ++i) {
      ta += sa[i];
      tb += sb[i];
    }
    partial[blockIdx.x] = make_double2(ta, tb);
  }
}

__global__ __launch_bounds__(kThreads) void finalize_kernel(const double2* partial, int nb, double2* out) {
  __shared__ double2 s[kThreads];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    a += partial[i].x;
    b += partial[i].y;
  }
  s[threadIdx.x] = make_double2(a, b);
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int i = 0; i < kThreads; ++i) {
      ta += s[i].x;
      tb += s[i].y;
    }
    *out = make_double2(ta, tb);
  }
}

// sum conj(a) * b
__global__ __launch_bounds__(kThreads) void dot_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                                       long long n, double2* partial) {
  double re = 0, im = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 u = a[i], v = b[i];
    re += (double)u.x * v.x + (double)u.y * v.y;
    im += (double)u.x * v.y - (double)u.y * v.x;
  }
  block_store(re, im, partial);
}

// per-block max |a|^2 into partial[].x (FP64)
__global__ __launch_bounds__(kThreads) void maxabs2_kernel(const float2* __restrict__ a, long long n,
                                                           double2* partial) {
  __shared__ double s[kThreads];
  double m = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 u = a[i];
    m = fmax(m, (double)u.x * u.x + (double)u.y * u.y);
  }
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s[0], 0.0);
}

__global__ __launch_bounds__(kThreads) void finalize_max_kernel(const double2* partial, int nb, double2* out) {
  __shared__ double s[kThreads];
  double m = 0;
  for (int i = threadIdx.x; i < nb; i += kThreads) m = fmax(m, partial[i].x);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = make_double2(s[0], 0.0);
}

// q = Ma .* (q + lam p); partial of sum conj(p) q
__global__ __launch_bounds__(kThreads) void constrain_dot_kernel(float2* __restrict__ q, const float2* __restrict__ p,
                                                                 const float* __restrict__ Ma, float lam, long long n,
                                                                 double2* partial) {
  double re = 0, im = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 u = p[i];
    float2 v = q[i];
    v.x += lam * u.x;
    v.y += lam * u.y;
    if (Ma) {
      const float m = Ma[i];
      v.x *= m;
      v.y *= m;
    }
    q[i] = v;
    re += (double)u.x * v.x + (double)u.y * v.y;
    im += (double)u.x * v.y - (double)u.y * v.x;
  }
  block_store(re, im, partial);
}

// r = b - q (b may alias r); z = Se .* r; partials of sum conj(z) r and sum |conj(Se b) b|
__global__ __launch_bounds__(kThreads) void residual_kernel(const float2* __restrict__ b, const float2* __restrict__ q,
                                                            float2* __restrict__ r, float2* __restrict__ z,
                                                            float2* __restrict__ p, const float* __restrict__ Se,
                                                            long long n, double2* partial, double2* partial_b) {
  double re = 0, im = 0, bb = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 bv = b[i];
    float2 rv = bv;
    if (q) {
      const float2 qv = q[i];
      rv.x -= qv.x;
      rv.y -= qv.y;
    }
    const float s = Se ? Se[i] : 1.0f;
    const float2 zv = make_float2(s * rv.x, s * rv.y);
    r[i] = rv;
    z[i] = zv;
    p[i] = zv;
    re += (double)zv.x * rv.x + (double)zv.y * rv.y;
    im += (double)zv.x * rv.y - (double)zv.y * rv.x;
    // |conj(Se b) b| = |Se| |b|^2
    bb += fabs((double)s) * ((double)bv.x * bv.x + (double)bv.y * bv.y);
  }
  block_store(re, im, partial);
  if (partial_b) block_store(bb, 0.0, partial_b);
}

// x += g p; r -= g q; z = Se .* r; partial of sum conj(z) r and sum |conj(z) r|
__global__ __launch_bounds__(kThreads) void update_kernel(float2* __restrict__ x, float2* __restrict__ r,
                                                          float2* __restrict__ z, const float2* __restrict__ p,
                                                          const float2* __restrict__ q, const float* __restrict__ Se,
                                                          const CgScalars* sc, long long n, double2* partial,
                                                          double2* partial_abs) {
  const float gx = (float)sc->g.x, gy = (float)sc->g.y;
  double re = 0, im = 0, ab = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 pv = p[i], qv = q[i];
    float2 xv = x[i], rv = r[i];
    xv.x += gx * pv.x - gy * pv.y;
    xv.y += gx * pv.y + gy * pv.x;
    rv.x -= gx * qv.x - gy * qv.y;
    rv.y -= gx * qv.y + gy * qv.x;
    const float s = Se ? Se[i] : 1.0f;
    const float2 zv = make_float2(s * rv.x, s * rv.y);
    x[i] = xv;
    r[i] = rv;
    z[i] = zv;
    const double pr = (double)zv.x * rv.x + (double)zv.y * rv.y;
    const double pi = (double)zv.x * rv.y - (double)zv.y * rv.x;
    re += pr;
    im += pi;
    ab += sqrt(pr * pr + pi * pi);
  }
  block_store(re, im, partial);
  block_store(ab, 0.0, partial_abs);
}

// Ex += g Ep (the same FP32 step as x += g p); partial of sum |Ex - y|^2
template <bool ENERGY>
__global__ __launch_bounds__(kThreads) void kspace_update_kernel(float2* __restrict__ Ex, const float2* __restrict__ Ep,
                                                                 const float2* __restrict__ y, const CgScalars* sc,
                                                                 long long n, double2* partial) {
  const float gx = (float)sc->g.x, gy = (float)sc->g.y;
  double e = 0;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 pv = Ep[i];
    float2 xv = Ex[i];
    xv.x += gx * pv.x - gy * pv.y;
    xv.y += gx * pv.y + gy * pv.x;
    Ex[i] = xv;
    if (ENERGY) {
      const float2 yv = y[i];
      const double dx = (double)xv.x - yv.x, dy = (double)xv.y - yv.y;
      e += dx * dx + dy * dy;
    }
  }
  if (ENERGY) block_store(e, 0.0, partial);
}

// p = z + d p
__global__ __launch_bounds__(kThreads) void direction_kernel(float2* __restrict__ p, const float2* __restrict__ z,
                                                             const CgScalars* sc, long long n) {
  const float dx = (float)sc->d.x, dy = (float)sc->d.y;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float2 pv = p[i], zv = z[i];
    p[i] = make_float2(zv.x + dx * pv.x - dy * pv.y, zv.y + dx * pv.y + dy * pv.x);
  }
}

__global__ void mask_kernel(float2* __restrict__ v, const float* __restrict__ Ma, long long n) {
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
    const float m = Ma[i];
    float2 a = v[i];
    a.x *= m;
    a.y *= m;
    v[i] = a;
  }
}

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double den = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}

// g = rsold / pq   (CGsolver.m:82)
__global__ void step_kernel(CgScalars* sc) { sc->g = cdiv(sc->rsold, sc->pq); }

// Fixed-count iterations (no host test between them): once |rsold| < 1e-6
// (the break of CGsolver.m:101) the step is zero, so x and r stay where the
// loop would have stopped.
__global__ void step_guarded_kernel(CgScalars* sc) {
  const double2 rs = sc->rsold;
  const bool stop = hypot(rs.x, rs.y) < 1e-6 || (sc->pq.x == 0.0 && sc->pq.y == 0.0);
  sc->g = stop ? make_double2(0.0, 0.0) : cdiv(rs, sc->pq);
}

__global__ void record_kernel(const CgScalars* sc, double* hist, int* cnt) {
  const int k = *cnt;
  hist[k] = sc->zr_abs.x;
  *cnt = k + 1;
}

// rs from the reduction; d = rs / (rsold + eps); rsold = rs  (CGsolver.m:94-97)
__global__ void beta_kernel(CgScalars* sc) {
  const double2 rs = sc->rs;
  const double2 den = make_double2(sc->rsold.x + 2.220446049250313e-16, sc->rsold.y);
  sc->d = cdiv(rs, den);
  sc->rsold = rs;
}

int grid_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  return (int)(b < kRedBlocks ? b : kRedBlocks);
}

}  // namespace

hipError_t cg_dot(const float2* a, const float2* b, long long n, double2* partial, double2* out, hipStream_t st) {
  const int nb = grid_for(n);
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(kThreads), 0, st, a, b, n, partial);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, out);
  return hipGetLastError();
}

hipError_t cg_maxabs2(const float2* a, long long n, double2* partial, double2* out, hipStream_t st) {
  const int nb = grid_for(n);
  hipLaunchKernelGGL(maxabs2_kernel, dim3(nb), dim3(kThreads), 0, st, a, n, partial);
  hipLaunchKernelGGL(finalize_max_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, out);
  return hipGetLastError();
}

hipError_t cg_mask(float2* v, const float* Ma, long long n, hipStream_t st) {
  hipLaunchKernelGGL(mask_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, v, Ma, n);
  return hipGetLastError();
}

hipError_t cg_constrain_dot(float2* q, const float2* p, const float* Ma, float lam, long long n, double2* partial,
                            double2* out, hipStream_t st) {
  const int nb = grid_for(n);
  hipLaunchKernelGGL(constrain_dot_kernel, dim3(nb), dim3(kThreads), 0, st, q, p, Ma, lam, n, partial);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, out);
  return hipGetLastError();
}

hipError_t cg_residual(const float2* b, const float2* q, float2* r, float2* z, float2* p, const float* Se, long long n,
                       double2* partial, double2* partial_b, double2* rs_out, double2* bn_out, hipStream_t st) {
  const int nb = grid_for(n);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(kThreads), 0, st, b, q, r, z, p, Se, n, partial, partial_b);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, rs_out);
  if (partial_b) hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial_b, nb, bn_out);
  return hipGetLastError();
}

hipError_t cg_step(CgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(step_kernel, dim3(1), dim3(1), 0, st, sc);
  return hipGetLastError();
}

hipError_t cg_step_guarded(CgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(step_guarded_kernel, dim3(1), dim3(1), 0, st, sc);
  return hipGetLastError();
}

hipError_t cg_record(const CgScalars* sc, double* hist, int* cnt, hipStream_t st) {
  hipLaunchKernelGGL(record_kernel, dim3(1), dim3(1), 0, st, sc, hist, cnt);
  return hipGetLastError();
}

hipError_t cg_update(float2* x, float2* r, float2* z, const float2* p, const float2* q, const float* Se, CgScalars* sc,
                     long long n, double2* partial, double2* partial_abs, hipStream_t st) {
  const int nb = grid_for(n);
  hipLaunchKernelGGL(update_kernel, dim3(nb), dim3(kThreads), 0, st, x, r, z, p, q, Se, sc, n, partial, partial_abs);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, &sc->rs);
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial_abs, nb, &sc->zr_abs);
  hipLaunchKernelGGL(beta_kernel, dim3(1), dim3(1), 0, st, sc);
  return hipGetLastError();
}

hipError_t cg_kspace_update(float2* Ex, const float2* Ep, const float2* y, CgScalars* sc, long long n, bool energy,
                            double2* partial, hipStream_t st) {
  const int nb = grid_for(n);
  if (energy) {
    hipLaunchKernelGGL(kspace_update_kernel<true>, dim3(nb), dim3(kThreads), 0, st, Ex, Ep, y, sc, n, partial);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st, partial, nb, &sc->fid);
  } else {
    hipLaunchKernelGGL(kspace_update_kernel<false>, dim3(nb), dim3(kThreads), 0, st, Ex, Ep, y, sc, n, partial);
  }
  return hipGetLastError();
}

hipError_t cg_direction(float2* p, const float2* z, const CgScalars* sc, long long n, hipStream_t st) {
  hipLaunchKernelGGL(direction_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, p, z, sc, n);
  return hipGetLastError();
}

}  // namespace dis
