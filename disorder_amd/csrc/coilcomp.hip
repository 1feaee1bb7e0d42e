// Coil-array compression (Methods/compressCoils.m:75-118) on gfx950.
//
//  * gram_kernel: P(i,j) = sum_v Shat_i(v) conj(S_j(v)),
//    Shat = S / (sum_c |S_c|^2 + 1e-9)  (compressCoils.m:78-87; normm is the
//    squared norm, Methods/normm.m:15).  FP32 products, FP64 accumulation per
//    block, fixed-order FP64 reduction of the block partials (deterministic).
//  * host: [U,D] = svd(P) (:93).  P is Hermitian positive semi-definite, so
//    its SVD is its eigendecomposition with |eigenvalues| sorted descending;
//    computed by cyclic complex Jacobi in FP64.  Component counts by energy
//    fraction or fixed count (:107-111).
//  * apply_kernel: out(v, k) = sum_c A(k, c) in(v, c), A = U(:,1:Cc)^H
//    (:115-118, aplGPU along dim 4), as a real GEMM [2Cc x 2C] . [2C x nv] on
//    the matrix cores (v_mfma_f32_16x16x16_bf16).  FP32 operands are split
//    into bf16 terms and the products accumulate in FP32:
//      TERMS = 3: a = h + l, ab ~ hh + hl + lh          (~2^-16 relative)
//      TERMS = 6: a = h + m + l, ab ~ hh + hm + mh + hl + lh + mm (~FP32)
//    Real and imaginary parts interleave along K (k = 2c + {re, im}) and
//    along N (n = 2k' + {re, im}).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

#include "../../include/disorder.h"
#include "kernels.hpp"
#include "transform.hpp"

namespace dis {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxCoils = 64;

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
  const unsigned int u = __float_as_uint(f);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// ------------------------------------------------------------------ Gram
constexpr int kGramTile = 32;     // voxels staged per step
constexpr int kGramLd = 66;       // padded coil stride of the staging tiles
constexpr int kGramFlush = 8;     // FP32 steps (of kGramTile voxels) per FP64 flush

// One block accumulates a contiguous chunk of voxels.  Thread (ti, tj) owns a
// 4x4 block of P (256 threads cover 64 x 64 coils).
__global__ __launch_bounds__(256) void gram_kernel(const float2* __restrict__ S, long long nv, int C,
                                                  long long chunk, double2* __restrict__ partial) {
  __shared__ float2 sh[kGramTile][kGramLd];  // Shat
  __shared__ float2 ss[kGramTile][kGramLd];  // S
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  const long long v0 = (long long)blockIdx.x * chunk;
  const long long v1 = v0 + chunk < nv ? v0 + chunk : nv;
  double2 accd[4][4];
  float2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accd[a][b] = make_double2(0, 0);
      acc[a][b] = make_float2(0, 0);
    }
  int steps = 0;
  for (long long vb = v0; vb < v1; vb += kGramTile) {
    // stage kGramTile voxels x 64 coils (coalesced along voxels per coil)
    for (int e = tid; e < kGramTile * kMaxCoils; e += 256) {
      const int vv = e % kGramTile, cc = e / kGramTile;
      const long long v = vb + vv;
      ss[vv][cc] = (v < v1 && cc < C) ? S[v + nv * cc] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    {
      // squared coil norm of voxel vv by 8 lanes, 8 coils each
      const int vv = tid >> 3, part = tid & 7;
      float nrm = 0.f;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float2 s = ss[vv][8 * part + i];
        nrm += s.x * s.x + s.y * s.y;
      }
      nrm += __shfl_xor(nrm, 1, 64);
      nrm += __shfl_xor(nrm, 2, 64);
      nrm += __shfl_xor(nrm, 4, 64);
      const float inv = 1.0f / (nrm + 1e-9f);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float2 s = ss[vv][8 * part + i];
        sh[vv][8 * part + i] = make_float2(s.x * inv, s.y * inv);
      }
    }
    __syncthreads();
    for (int vv = 0; vv < kGramTile; ++vv) {
      float2 hi[4], sj[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) hi[a] = sh[vv][4 * ti + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) sj[b] = ss[vv][4 * tj + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cadd(acc[a][b], cmulc(hi[a], sj[b]));
    }
    __syncthreads();
    if (++steps == kGramFlush) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          accd[a][b].x += acc[a][b].x;
          accd[a][b].y += acc[a][b].y;
          acc[a][b] = make_float2(0, 0);
        }
      steps = 0;
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accd[a][b].x += acc[a][b].x;
      accd[a][b].y += acc[a][b].y;
      const int i = 4 * ti + a, j = 4 * tj + b;
      partial[(long long)blockIdx.x * 4096 + i * 64 + j] = accd[a][b];
    }
}

// P entry e = i * 64 + j summed over the block partials in a fixed order:
// four strided partial sums per entry, then combined in order.
__global__ __launch_bounds__(256) void gram_finalize_kernel(const double2* __restrict__ partial, int nb,
                                                           int C, double2* __restrict__ out) {
  __shared__ double2 red[4][64];
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  double x = 0, y = 0;
  for (int b = q; b < nb; b += 4) {
    const double2 p = partial[(long long)b * 4096 + e];
    x += p.x;
    y += p.y;
  }
  red[q][threadIdx.x & 63] = make_double2(x, y);
  __syncthreads();
  if (q == 0) {
    double2 t = red[0][threadIdx.x];
    for (int r = 1; r < 4; ++r) {
      t.x += red[r][threadIdx.x].x;
      t.y += red[r][threadIdx.x].y;
    }
    const int i = e >> 6, j = e & 63;
    if (i < C && j < C) out[i + (long long)C * j] = t;  // column-major C x C
  }
}

// ------------------------------------------------------------------ apply
template <int TERMS>
__device__ __forceinline__ void split_terms(float a, short (&t)[3]);

// ------------------------------------------------------------ Gram on MFMA
// The Gram matrix as one real GEMM over the voxels, G = X Y^T with
// X rows 2c + {re, im} of Shat and Y rows 2c + {re, im} of S (2C <= 128 rows,
// K = voxels): P_re(i,j) = G[2i][2j] + G[2i+1][2j+1],
// P_im(i,j) = G[2i+1][2j] - G[2i][2j+1].  A workgroup takes a chunk of
// voxels in tiles of kGmKt: the 256 threads stage a tile (coalesced along
// voxels per coil), form Shat with the voxel's squared coil norm, and store
// X and Y split into bf16 terms in LDS; wave w then accumulates output rows
// [32 w, 32 w + 32) x all 128 columns (16 tiles of 16 x 16) with
// v_mfma_f32_16x16x16_bf16 (hi.hi + hi.lo + lo.hi; with TERMS = 6 also
// mid terms) in FP32.  Each workgroup writes its FP32 128 x 128 partial; a
// fixed-order FP64 sum over the partials gives P (deterministic).
// Fragment layouts as in apply_kernel: lane l holds A[l&15][4(l>>4)+j],
// B[4(l>>4)+j][l&15], D[4(l>>4)+j][l&15].
constexpr int kGmKt = 32;             // voxels per staged tile (two MFMA K steps)
constexpr int kGmLd = kGmKt + 4;      // padded row stride of the bf16 tiles
constexpr int kGmRows = 2 * kMaxCoils;
constexpr int kGmChunk = 4096;        // voxels per workgroup (FP32 accumulation length)
constexpr int kGmBlocksMax = 8192;

template <int TERMS>
__global__ __launch_bounds__(256) void gram_mfma_kernel(const float2* __restrict__ S, long long nv, int C,
                                                       long long chunk, float* __restrict__ partial) {
  constexpr int NT = TERMS == 6 ? 3 : 2;  // bf16 terms per value
  __shared__ __align__(16) short xs[NT][kGmRows][kGmLd];
  __shared__ __align__(16) short ys[NT][kGmRows][kGmLd];
  __shared__ float nrm[8][kGmKt];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int i16 = lane & 15, q = lane >> 4;
  const long long v0 = (long long)blockIdx.x * chunk;
  const long long v1 = v0 + chunk < nv ? v0 + chunk : nv;
  f32x4 acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int R = 2 * C;
  const bool wave_live = 32 * w < R;  // rows of this wave hold coils
  // staging role: voxel tv, coils 8 tg .. 8 tg + 7
  const int tv = tid & (kGmKt - 1), tg = tid / kGmKt;
  // the next tile's maps are loaded while this tile's MFMAs run
  auto load_tile = [&](long long vb, float2 (&s)[8]) {
    const long long v = vb + tv;
    const bool vok = v < v1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = 8 * tg + k;
      s[k] = (vok && c < C) ? S[v + nv * c] : make_float2(0.f, 0.f);
    }
  };
  float2 s[8], sn[8];
  load_tile(v0, s);
  for (long long vb = v0; vb < v1; vb += kGmKt) {
    float part = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) part += s[k].x * s[k].x + s[k].y * s[k].y;
    nrm[tg][tv] = part;
    __syncthreads();
    float tot = 0.f;
#pragma unroll
    for (int g = 0; g < 8; ++g) tot += nrm[g][tv];
    const float inv = 1.0f / (tot + 1e-9f);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int r = 2 * (8 * tg + k);
      const float vals[4] = {s[k].x * inv, s[k].y * inv, s[k].x, s[k].y};
      short t[4][3];
#pragma unroll
      for (int e = 0; e < 4; ++e) split_terms<TERMS>(vals[e], t[e]);
#pragma unroll
      for (int h = 0; h < NT; ++h) {
        xs[h][r][tv] = t[0][h];
        xs[h][r + 1][tv] = t[1][h];
        ys[h][r][tv] = t[2][h];
        ys[h][r + 1][tv] = t[3][h];
      }
    }
    __syncthreads();
    load_tile(vb + kGmKt, sn);  // past v1 the loads are predicated off
    if (wave_live) {
#pragma unroll
      for (int ks = 0; ks < kGmKt / 16; ++ks) {
        const int k0 = 16 * ks + 4 * q;
        bf16x4 A[2][NT];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
          for (int h = 0; h < NT; ++h) A[mt][h] = *(const bf16x4*)&xs[h][32 * w + 16 * mt + i16][k0];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
          if (16 * nt >= R) break;  // uniform: columns beyond the coils are zero
          bf16x4 B[NT];
#pragma unroll
          for (int h = 0; h < NT; ++h) B[h] = *(const bf16x4*)&ys[h][16 * nt + i16][k0];
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            f32x4 d = acc[mt][nt];
            // smallest terms first
            if (TERMS == 6) {
              d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][1], B[1], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][0], B[2], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][2], B[0], d, 0, 0, 0);
            }
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][0], B[1], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][1], B[0], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[mt][0], B[0], d, 0, 0, 0);
            acc[mt][nt] = d;
          }
        }
      }
    }
    __syncthreads();  // the next tile overwrites the staging buffers
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = sn[k];
  }
  // partial[blk][row][col], row = 32 w + 16 mt + 4 q + j, col = 16 nt + i16
  float* out = partial + (long long)blockIdx.x * (kGmRows * kGmRows);
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(32 * w + 16 * mt + 4 * q + j) * kGmRows + 16 * nt + i16] = acc[mt][nt][j];
}

// P(i, j) from the FP32 block partials of G, summed in FP64 in a fixed
// order: four strided partial sums per entry, then combined in order.
__global__ __launch_bounds__(256) void gram_mfma_finalize_kernel(const float* __restrict__ partial, int nb, int C,
                                                                double2* __restrict__ out) {
  __shared__ double2 red[4][64];
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);  // e = i * 64 + j
  const int qq = threadIdx.x >> 6;
  const int i = e >> 6, j = e & 63;
  double x = 0, y = 0;
  if (i < C && j < C) {
    for (int b = qq; b < nb; b += 4) {
      const float* g = partial + (long long)b * (kGmRows * kGmRows);
      x += (double)g[(2 * i) * kGmRows + 2 * j] + (double)g[(2 * i + 1) * kGmRows + 2 * j + 1];
      y += (double)g[(2 * i + 1) * kGmRows + 2 * j] - (double)g[(2 * i) * kGmRows + 2 * j + 1];
    }
  }
  red[qq][threadIdx.x & 63] = make_double2(x, y);
  __syncthreads();
  if (qq == 0 && i < C && j < C) {
    double2 t = red[0][threadIdx.x];
    for (int r = 1; r < 4; ++r) {
      t.x += red[r][threadIdx.x].x;
      t.y += red[r][threadIdx.x].y;
    }
    out[i + (long long)C * j] = t;  // column-major C x C
  }
}

template <int TERMS>
__device__ __forceinline__ void split_terms(float a, short (&t)[3]) {
  const unsigned short h = f2bf(a);
  const float r1 = a - bf2f(h);
  const unsigned short m = f2bf(r1);
  t[0] = (short)h;
  t[1] = (short)m;
  if (TERMS == 6) {
    t[2] = (short)f2bf(r1 - bf2f(m));
  } else {
    t[2] = 0;
  }
}

// D = Aop . Bop with Aop[m][k] = Bm[k][m] (the compression matrix, real
// output column m of the current 16-column tile, real input index k) and
// Bop[k][n] = in(voxel n, k) (16 voxels of a tile).  The constant operand
// stays in registers for the whole kernel; each wave walks voxel tiles with
// all sixteen loads of a tile in flight.  Operand layouts of
// v_mfma_f32_16x16x16_bf16: lane l holds A[l&15][4(l>>4)+j], B[4(l>>4)+j][l&15]
// and D[4(l>>4)+j][l&15], j < 4.  With k = 2c + {re, im}, lane (row, kq)
// needs coils 8 ks + 2 kq + {0, 1} of voxel `row` (two float2 loads per
// K step, 16 lanes reading 128 contiguous bytes per coil), and returns
// output coils 8 mt + 2 kq + {0, 1} of voxel `row` as float2 pairs.
// Cc > 8 runs one sweep over the voxels per 16-column tile.
constexpr int kApplyMaxKs = kMaxCoils / 8;  // K steps of 16 real inputs

// KS = ceil(C / 8) K steps.  Loads are branch-free: out-of-range voxels and
// coils read a clamped in-range address and are zeroed by a select.
template <int TERMS, int KS>
__global__ __launch_bounds__(256) void apply_kernel(const float2* __restrict__ in, long long nv, long long in_cs,
                                                   int C, const float* __restrict__ Bm, int Cc,
                                                   float2* __restrict__ out, long long out_cs) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 15, kq = lane >> 4;
  const long long ntiles = (nv + 15) >> 4;
  const long long w0 = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  const int K = 2 * C, Nn = 2 * Cc;
  for (int mt = 0; mt * 16 < Nn; ++mt) {
    // constant operand: rows m = 16 mt + row, columns k = 16 ks + 4 kq + j
    bf16x4 a0[KS], a1[KS], a2[KS];
    const int m = 16 * mt + row;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 16 * ks + 4 * kq + j;
        const float av = (k < K && m < Nn) ? Bm[(long long)k * Nn + m] : 0.f;
        short t[3];
        split_terms<TERMS>(av, t);
        a0[ks][j] = t[0];
        a1[ks][j] = t[1];
        a2[ks][j] = t[2];
      }
    const int oc0 = 8 * mt + 2 * kq;  // output coils of this lane
    for (long long t = w0; t < ntiles; t += wstride) {
      const long long v = t * 16 + row;
      const bool vok = v < nv;
      const float2* __restrict__ v0 = in + (vok ? v : nv - 1);  // coil 0
      const float2* __restrict__ src = v0 + in_cs * (2 * kq);
      float2 x[KS][2];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const bool cok = 8 * ks + 2 * kq + h < C;
          const float2 val = *(cok ? src + in_cs * (8 * ks + h) : v0);
          x[ks][h] = (vok && cok) ? val : make_float2(0.f, 0.f);
        }
      f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const float xv[4] = {x[ks][0].x, x[ks][0].y, x[ks][1].x, x[ks][1].y};
        bf16x4 b0, b1, b2;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          short tt[3];
          split_terms<TERMS>(xv[j], tt);
          b0[j] = tt[0];
          b1[j] = tt[1];
          b2[j] = tt[2];
        }
        // smallest terms first
        if (TERMS == 6) {
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[ks], b1, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0[ks], b2, d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[ks], b0, d, 0, 0, 0);
        }
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0[ks], b1, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[ks], b0, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0[ks], b0, d, 0, 0, 0);
      }
      if (vok) {
        if (oc0 < Cc) out[v + out_cs * oc0] = make_float2(d[0], d[1]);
        if (oc0 + 1 < Cc) out[v + out_cs * (oc0 + 1)] = make_float2(d[2], d[3]);
      }
    }
  }
}

template <int TERMS, int KS>
void launch_apply_ks(int ks, dim3 grid, hipStream_t s, const float2* in, long long nv, long long in_cs, int C,
                     const float* Bm, int Cc, float2* out, long long out_cs) {
  if constexpr (KS > 1) {
    if (ks < KS) {
      launch_apply_ks<TERMS, KS - 1>(ks, grid, s, in, nv, in_cs, C, Bm, Cc, out, out_cs);
      return;
    }
  }
  // one resident wave per tile stream: blocks = CUs x resident blocks per CU
  static int resident = 0;
  if (!resident) {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, apply_kernel<TERMS, KS>, 256, 0) == hipSuccess)
      resident = std::max(1, ncu * per_cu);
    else
      resident = 1024;
  }
  const unsigned nb = (unsigned)std::min<long long>(grid.x, resident);
  hipLaunchKernelGGL((apply_kernel<TERMS, KS>), dim3(nb), dim3(256), 0, s, in, nv, in_cs, C, Bm, Cc, out, out_cs);
}

// ------------------------------------------------------------------ host eig
// Cyclic Jacobi for a Hermitian matrix A (column-major n x n, FP64).  On
// return the diagonal of A holds the eigenvalues and V the eigenvectors
// (columns).  Each rotation first makes A(p,q) real by a phase on column q,
// then zeroes it by a real plane rotation.
void hermitian_jacobi(std::vector<std::complex<double>>& A, int n, std::vector<std::complex<double>>& V) {
  typedef std::complex<double> cd;
  auto at = [&](int i, int j) -> cd& { return A[i + (size_t)n * j]; };
  V.assign((size_t)n * n, cd(0, 0));
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  double fro = 0;
  for (auto& z : A) fro += std::norm(z);
  fro = std::sqrt(fro);
  if (fro == 0) return;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < q; ++p) off += std::norm(at(p, q));
    if (std::sqrt(2 * off) <= 1e-15 * fro) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cd apq = at(p, q);
        const double g = std::abs(apq);
        if (g <= 1e-300) continue;
        const cd ph = apq / g;  // A(p,q) = g e^{i phi}
        const double app = at(p, p).real(), aqq = at(q, q).real();
        const double tau = (aqq - app) / (2 * g);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
        const double cs = 1 / std::sqrt(1 + t * t), sn = t * cs;
        // G = diag(1, e^{-i phi}) R, R = [[c, s], [-s, c]] on (p, q)
        const cd vpp = cs, vqp = -sn * std::conj(ph), vpq = sn, vqq = cs * std::conj(ph);
        for (int k = 0; k < n; ++k) {  // A <- A G
          const cd x = at(k, p), y = at(k, q);
          at(k, p) = x * vpp + y * vqp;
          at(k, q) = x * vpq + y * vqq;
        }
        for (int k = 0; k < n; ++k) {  // A <- G^H A
          const cd x = at(p, k), y = at(q, k);
          at(p, k) = std::conj(vpp) * x + std::conj(vqp) * y;
          at(q, k) = std::conj(vpq) * x + std::conj(vqq) * y;
        }
        at(p, q) = 0;
        at(q, p) = 0;
        for (int k = 0; k < n; ++k) {  // V <- V G
          cd& x = V[k + (size_t)n * p];
          cd& y = V[k + (size_t)n * q];
          const cd xv = x, yv = y;
          x = xv * vpp + yv * vqp;
          y = xv * vpq + yv * vqq;
        }
      }
  }
}

}  // namespace

// ---- allocation-free Gram (dis_coil_gram_noalloc)
// Block partials of the Gram kernel in a file-scope device array: each GPU
// holds its own copy (kGramStaticBlocks x 64 x 64 complex128 = 16 MiB) until
// the process ends.  Host code never passes `scratch` itself (in host code
// the name is only a stand-in); it uses the address hipGetSymbolAddress
// returned for the device, recorded once per device at context creation.
constexpr int kGramStaticBlocks = 256;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[(size_t)kGramStaticBlocks * 4096];
static double2* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int coil_gram_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return DIS_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return DIS_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return DIS_E_HIP;
  }
  g_scratch_dev[device] = (double2*)p;
  return DIS_OK;
}

int coil_gram(dis_ctx* c, const float2* S, long long nv, int C, double2* P_dev_out) {
  // split-bf16 MFMA GEMM over voxel chunks of about kGmChunk, FP32 partials
  const long long nb0 = (nv + kGmChunk - 1) / kGmChunk;
  const int nb = (int)std::max<long long>(1, std::min<long long>(nb0, kGmBlocksMax));
  const long long chunk = ((nv + nb - 1) / nb + kGmKt - 1) / kGmKt * kGmKt;
  void* ws = nullptr;
  int st = ctx_workspace(c, sizeof(float) * kGmRows * kGmRows * (size_t)nb, &ws);
  if (st) return st;
  hipStream_t s = ctx_stream(c);
  const int ps = prof_begin(c, DIS_K_GRAM);
  hipLaunchKernelGGL(gram_mfma_kernel<3>, dim3(nb), dim3(256), 0, s, S, nv, C, chunk, (float*)ws);
  hipLaunchKernelGGL(gram_mfma_finalize_kernel, dim3(64), dim3(256), 0, s, (const float*)ws, nb, C, P_dev_out);
  prof_end(c, ps, 8.0 * (double)nv * C);
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "coil Gram launch failed");
  return DIS_OK;
}

int coil_apply(dis_ctx* c, const float* Bm_dev, int Cc, int C, const float2* in, long long nv, long long in_cs,
               float2* out, long long out_cs, int terms) {
  if (nv <= 0) return DIS_OK;
  const long long ntiles = (nv + 15) / 16;
  const dim3 grid((unsigned)std::min<long long>((ntiles + 3) / 4, 1 << 20));
  const int ks = (C + 7) / 8;
  hipStream_t s = ctx_stream(c);
  const int ps = prof_begin(c, DIS_K_GEMM);
  if (terms == 6)
    launch_apply_ks<6, kApplyMaxKs>(ks, grid, s, in, nv, in_cs, C, Bm_dev, Cc, out, out_cs);
  else
    launch_apply_ks<3, kApplyMaxKs>(ks, grid, s, in, nv, in_cs, C, Bm_dev, Cc, out, out_cs);
  prof_end(c, ps, 8.0 * (double)nv * (C + Cc));
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "coil apply launch failed");
  return DIS_OK;
}

}  // namespace dis

using namespace dis;

namespace {

// FP32 real operand of A [Cc x C] (column-major complex128): Bm[k][n] with
// k = 2c + {0: re, 1: im} of the input and n = 2k' + {0: re, 1: im} of the output.
std::vector<float> real_operand(const std::complex<double>* A, int Cc, int C) {
  const int Nn = 2 * Cc;
  std::vector<float> Bm((size_t)2 * C * Nn);
  for (int c = 0; c < C; ++c)
    for (int k = 0; k < Cc; ++k) {
      const std::complex<double> a = A[k + (size_t)Cc * c];
      Bm[(size_t)(2 * c) * Nn + 2 * k] = (float)a.real();
      Bm[(size_t)(2 * c + 1) * Nn + 2 * k] = (float)-a.imag();
      Bm[(size_t)(2 * c) * Nn + 2 * k + 1] = (float)a.imag();
      Bm[(size_t)(2 * c + 1) * Nn + 2 * k + 1] = (float)a.real();
    }
  return Bm;
}

int upload_operand(dis_ctx* c, const std::vector<float>& Bm, const float** out) {
  Staging stg;
  stg.add(Bm.data(), Bm.size() * sizeof(float));
  const unsigned char* d = nullptr;
  int st = ctx_upload_bytes(c, stg, &d);
  if (st) return st;
  *out = (const float*)d;
  return DIS_OK;
}

}  // namespace

extern "C" {

int dis_coil_gram(dis_ctx* c, const void* S, int64_t nv, int32_t C, double* P) {
  if (!c) return DIS_E_INVALID;
  if (!S || !P || nv <= 0 || C <= 0) return ctx_fail(c, DIS_E_INVALID, "dis_coil_gram: invalid argument");
  if (C > kMaxCoils) return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_coil_gram: more than 64 coils");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  void* buf = nullptr;
  int st = ctx_buffer(c, 3, sizeof(double2) * C * C, &buf);
  if (st) return st;
  st = coil_gram(c, (const float2*)S, nv, C, (double2*)buf);
  if (st) return st;
  if (hipMemcpyAsync(P, buf, sizeof(double2) * C * C, hipMemcpyDeviceToHost, ctx_stream(c)) != hipSuccess ||
      hipStreamSynchronize(ctx_stream(c)) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_coil_gram: copy of P failed");
  return DIS_OK;
}

int dis_coil_gram_noalloc(dis_ctx* c, const void* S, int64_t nv, int32_t C, void* P_dev) {
  if (!c) return DIS_E_INVALID;
  if (!S || !P_dev || nv <= 0 || C <= 0) return ctx_fail(c, DIS_E_INVALID, "dis_coil_gram_noalloc: invalid argument");
  if (C > kMaxCoils)
    return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_coil_gram_noalloc: more than 64 coils (the static scratch holds 64 x 64)");
  const int dev = ctx_device(c);
  double2* ws = (dev >= 0 && dev < kMaxDevices) ? g_scratch_dev[dev] : nullptr;
  if (!ws) return ctx_fail(c, DIS_E_HIP, "dis_coil_gram_noalloc: static scratch not set up for this device");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  // the same kernels as dis_coil_gram with at most kGramStaticBlocks blocks;
  // every block writes its whole 64 x 64 partial, so nothing is cleared
  const long long nb0 = (nv + 4095) / 4096;
  const int nb = (int)std::max<long long>(1, std::min<long long>(nb0, kGramStaticBlocks));
  const long long chunk = ((nv + nb - 1) / nb + kGramTile - 1) / kGramTile * kGramTile;
  hipStream_t s = ctx_stream(c);
  hipLaunchKernelGGL(gram_kernel, dim3(nb), dim3(256), 0, s, (const float2*)S, (long long)nv, (int)C, chunk, ws);
  hipLaunchKernelGGL(gram_finalize_kernel, dim3(64), dim3(256), 0, s, (const double2*)ws, nb, (int)C,
                     (double2*)P_dev);
  if (hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "dis_coil_gram_noalloc: launch failed");
  return DIS_OK;
}

int dis_coil_basis(const double* P, int32_t C, const double* perc, int32_t nperc, int32_t* Ccount, double* D,
                   double* U) {
  if (!P || C <= 0 || (nperc > 0 && (!perc || !Ccount))) return DIS_E_INVALID;
  typedef std::complex<double> cd;
  std::vector<cd> A((size_t)C * C), V;
  for (size_t e = 0; e < A.size(); ++e) A[e] = cd(P[2 * e], P[2 * e + 1]);
  // symmetrise the FP64 rounding of the Gram sums
  for (int j = 0; j < C; ++j)
    for (int i = 0; i <= j; ++i) {
      const cd a = 0.5 * (A[i + (size_t)C * j] + std::conj(A[j + (size_t)C * i]));
      A[i + (size_t)C * j] = a;
      A[j + (size_t)C * i] = std::conj(a);
    }
  hermitian_jacobi(A, C, V);
  std::vector<int> ord(C);
  for (int i = 0; i < C; ++i) ord[i] = i;
  std::vector<double> lam(C);
  for (int i = 0; i < C; ++i) lam[i] = std::fabs(A[i + (size_t)C * i].real());
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return lam[a] > lam[b]; });
  double tot = 0;
  for (int i = 0; i < C; ++i) tot += lam[i];
  for (int n = 0; n < nperc; ++n) {
    if (perc[n] < 1) {
      double cs = 0;
      int k = C;
      for (int i = 0; i < C; ++i) {
        cs += lam[ord[i]];
        if (tot > 0 && cs / tot >= perc[n]) {
          k = i + 1;
          break;
        }
      }
      Ccount[n] = k;
    } else {
      Ccount[n] = (int)std::min<double>(perc[n], C);
    }
  }
  if (D)
    for (int i = 0; i < C; ++i) D[i] = lam[ord[i]];
  if (U)
    for (int j = 0; j < C; ++j)
      for (int i = 0; i < C; ++i) {
        const cd u = V[i + (size_t)C * ord[j]];
        U[2 * (i + (size_t)C * j)] = u.real();
        U[2 * (i + (size_t)C * j) + 1] = u.imag();
      }
  return DIS_OK;
}

int dis_coil_apply(dis_ctx* c, const double* A, int32_t Cc, int32_t C, const void* in, int64_t nv, void* out,
                   int32_t terms) {
  if (!c) return DIS_E_INVALID;
  if (!A || !in || !out || Cc <= 0 || C <= 0 || nv < 0 || (terms != 3 && terms != 6))
    return ctx_fail(c, DIS_E_INVALID, "dis_coil_apply: invalid argument");
  if (C > kMaxCoils || Cc > kMaxCoils) return ctx_fail(c, DIS_E_UNSUPPORTED, "dis_coil_apply: more than 64 coils");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  const std::vector<float> Bm = real_operand((const std::complex<double>*)A, Cc, C);
  const float* Bd = nullptr;
  int st = upload_operand(c, Bm, &Bd);
  if (st) return st;
  st = coil_apply(c, Bd, Cc, C, (const float2*)in, nv, nv, (float2*)out, nv, terms);
  if (st) return st;
  if (hipStreamSynchronize(ctx_stream(c)) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "dis_coil_apply failed");
  return DIS_OK;
}

int dis_compress_coils(dis_ctx* c, const void* S, int64_t nvS, int32_t C, const double* perc, int32_t nperc,
                       const void* y, int64_t nvy, void* S_out, void* y_out, int32_t terms, int32_t* Ccount,
                       double* D, double* U) {
  if (!c) return DIS_E_INVALID;
  if (!S || !S_out || nvS <= 0 || C <= 0 || nperc <= 0 || !perc || !Ccount || (y && (!y_out || nvy < 0)))
    return ctx_fail(c, DIS_E_INVALID, "dis_compress_coils: invalid argument");
  std::vector<double> P((size_t)2 * C * C), Uh((size_t)2 * C * C);
  int st = dis_coil_gram(c, S, nvS, C, P.data());
  if (st) return st;
  st = dis_coil_basis(P.data(), C, perc, nperc, Ccount, D, Uh.data());
  if (st) return ctx_fail(c, st, "dis_compress_coils: eigendecomposition failed");
  int Cc = 0;
  for (int n = 0; n < nperc; ++n) Cc = std::max(Cc, (int)Ccount[n]);
  // A = U(:, 1:Cc)^H, column-major Cc x C
  std::vector<double> A((size_t)2 * Cc * C);
  for (int k = 0; k < Cc; ++k)
    for (int i = 0; i < C; ++i) {
      A[2 * (k + (size_t)Cc * i)] = Uh[2 * (i + (size_t)C * k)];
      A[2 * (k + (size_t)Cc * i) + 1] = -Uh[2 * (i + (size_t)C * k) + 1];
    }
  if (U) std::copy(Uh.begin(), Uh.end(), U);
  st = dis_coil_apply(c, A.data(), Cc, C, S, nvS, S_out, terms);
  if (st) return st;
  if (y && nvy > 0) st = dis_coil_apply(c, A.data(), Cc, C, y, nvy, y_out, terms);
  return st;
}

}  // extern "C"
