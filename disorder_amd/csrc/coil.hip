// Coil stage of E, E^H and E^H E on gfx950.
//
// One workgroup per (motion state s, readout plane i3).  The sampled k-space
// value of sample j (phase-encode position (k1_j, k2_j)) for coil c is
//   Y_c[j, i3] = 1/sqrt(n1 n2) sum_{i2} W2^{k2_j i2} sum_{i1} W1^{k1_j i1} S_c u
// (the unitary row-DFT form of Reconstruction/encode.m:120-136).  The plane is
// walked in tiles of G dim-2 rows: each tile is multiplied by S_c, FFT'd along
// dim 1 in LDS, and every thread accumulates the samples it owns (registers)
// with the dim-2 DFT weight W2^{k2 i2}.  The adjoint (Reconstruction/
// decode.m:43-84) stages the chunk's weighted sample values in LDS; each
// tile element (row i2, spectrum column k1) then gathers the samples of its
// k1 row from the state's CSR in a fixed order (no atomics, so sums are
// reproducible run to run), the tile is inverse-FFT'd along dim 1,
// multiplied by conj(S_c) and the coils are summed in registers.  E^H E
// runs the forward half into a k-space scratch [Ns x n3 x C] and the adjoint
// half from it, as two launches; the one-launch form (MODE = COIL_EHE, both
// halves in one workgroup, the k-space kept on chip) is kept in the source
// but measured slower (its registers are live across both halves).
#include <algorithm>

#include "kernels.hpp"
#include "regfft.hpp"

namespace dis {

constexpr int kMaxJ = 4;   // samples per thread per chunk
constexpr int kCC = 8;     // coils per chunk
constexpr int kTE = kMaxTile / kThreads;

template <int MODE>
__global__ __launch_bounds__(kThreads) void coil_kernel(CoilArgs a) {
  constexpr bool FWD = MODE == COIL_ENCODE || MODE == COIL_EHE || MODE == COIL_ENERGY;
  constexpr bool ADJ = MODE == COIL_DECODE || MODE == COIL_EHE;
  __shared__ float2 lds[kMaxTile];
  __shared__ float2 tw1[1024];
  __shared__ float2 tw2[1024];
  __shared__ float2 wst[ADJ ? kMaxJ * kThreads : 1];  // adjoint: chunk's sample values of the current coil
  const int tid = threadIdx.x;
  const int n1 = a.n1, n2 = a.n2;
  const int i3 = a.i3lo + blockIdx.x;
  const int sl = blockIdx.y;
  const int s = a.slist ? a.slist[sl] : a.state0 + sl;
  const int G = a.G;
  const long long nplane = (long long)n1 * n2;
  const long long nvox = nplane * a.n3;
  const float inv_sqrt = rsqrtf((float)n1 * (float)n2);
  const float inv_n1 = 1.0f / (float)n1;
  for (int k = tid; k < n1; k += kThreads) tw1[k] = a.tw1[k];
  for (int k = tid; k < n2; k += kThreads) tw2[k] = a.tw2[k];
  __syncthreads();

  const int o0 = a.off[s], o1 = a.off[s + 1];
  const float2* Up = FWD ? a.U + (a.us_rel ? sl : s) * a.us + i3 * nplane : nullptr;
  float2* Vp = ADJ ? a.V + (long long)(a.vs_rel ? sl : s) * a.vs + i3 * nplane : nullptr;
  const int ngroups = (n2 + G - 1) / G;
  const double cf = a.tfwd ? a.tfwd[s] * a.invN1 : 0.0;
  const double ca = a.tadj ? a.tadj[s] * a.invN1 : 0.0;

  bool first_out = true;
  for (int sc0 = o0; sc0 < o1 || (sc0 == o0 && ADJ); sc0 += kMaxJ * kThreads) {
    const int sc1 = min(o1, sc0 + kMaxJ * kThreads);
    int jk1[kMaxJ], jk2[kMaxJ];
    float jw[kMaxJ];
#pragma unroll
    for (int i = 0; i < kMaxJ; ++i) {
      const int j = sc0 + tid + i * kThreads;
      if (j < sc1) {
        jk1[i] = a.k1[j];
        jk2[i] = a.k2[j];
        jw[i] = a.We ? a.We[j] : 1.0f;
      } else {
        jk1[i] = -1;
        jk2[i] = 0;
        jw[i] = 0.f;
      }
    }
    float esum[kMaxJ];
#pragma unroll
    for (int i = 0; i < kMaxJ; ++i) esum[i] = 0.f;

    for (int c0 = 0; c0 < a.C; c0 += kCC) {
      const int cc = min(kCC, a.C - c0);
      float2 W[kMaxJ][kCC];
#pragma unroll
      for (int i = 0; i < kMaxJ; ++i)
#pragma unroll
        for (int c = 0; c < kCC; ++c) W[i][c] = make_float2(0.f, 0.f);

      if (FWD) {
        for (int g = 0; g < ngroups; ++g) {
          const int r0 = g * G;
          const int Gv = min(G, n2 - r0);
          const int tot = G * n1;
          const int totv = Gv * n1;
          float2 ureg[kTE];
#pragma unroll
          for (int i = 0; i < kTE; ++i) {
            const int e = tid + i * kThreads;
            if (e < tot) lds[e] = e < totv ? Up[(long long)r0 * n1 + e] : make_float2(0.f, 0.f);
          }
          __syncthreads();
          if (a.tfwd) {  // last translation of T (along dim 1) before the coil multiply
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) {
                const int j = e % n1;
                lds[e] = cswap(cmul(lds[e], phase_cycles(cf, freq_index(j, n1))));
              }
            }
            __syncthreads();
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              ureg[i] = (e < tot) ? cscale(cswap(lds[e]), inv_n1) : make_float2(0.f, 0.f);
            }
          } else {
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              ureg[i] = (e < tot) ? lds[e] : make_float2(0.f, 0.f);
            }
          }
          __syncthreads();
          for (int c = 0; c < cc; ++c) {
            const float2* Sp = a.Smaps + (long long)(c0 + c) * nvox + i3 * nplane + (long long)r0 * n1;
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) lds[e] = e < totv ? cmul(ureg[i], Sp[e]) : make_float2(0.f, 0.f);
            }
            __syncthreads();
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
#pragma unroll
            for (int i = 0; i < kMaxJ; ++i) {
              if (jk1[i] >= 0) {
                int idx = (int)(((long long)jk2[i] * r0) % n2);
                float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                for (int cq = 0; cq < kCC; ++cq)
                  if (cq == c) acc = W[i][cq];
                for (int p = 0; p < Gv; ++p) {
                  acc = cadd(acc, cmul(tw2[idx], lds[p * n1 + jk1[i]]));
                  idx += jk2[i];
                  if (idx >= n2) idx -= n2;
                }
                // W[i][c] with runtime c: write through a small switch to keep registers
#pragma unroll
                for (int cq = 0; cq < kCC; ++cq)
                  if (cq == c) W[i][cq] = acc;
              }
            }
            __syncthreads();
          }
        }
        // samples of this chunk are complete
#pragma unroll
        for (int i = 0; i < kMaxJ; ++i) {
          const int j = sc0 + tid + i * kThreads;
          if (jk1[i] >= 0) {
#pragma unroll
            for (int c = 0; c < kCC; ++c) {
              if (c < cc) {
                const float2 Y = cscale(W[i][c], inv_sqrt);
                const long long yi = j + (long long)a.Ns * (i3 + (long long)a.n3 * (c0 + c));
                if (MODE == COIL_ENCODE) a.yout[yi] = Y;
                if (MODE == COIL_ENERGY) {
                  const float2 r = csub(Y, a.yin[yi]);
                  esum[i] += r.x * r.x + r.y * r.y;
                }
                if (MODE == COIL_EHE) W[i][c] = cscale(Y, jw[i]);
              }
            }
          }
        }
      }
      if (MODE == COIL_DECODE) {
#pragma unroll
        for (int i = 0; i < kMaxJ; ++i) {
          const int j = sc0 + tid + i * kThreads;
          if (jk1[i] >= 0) {
#pragma unroll
            for (int c = 0; c < kCC; ++c)
              if (c < cc) W[i][c] = cscale(a.yin[j + (long long)a.Ns * (i3 + (long long)a.n3 * (c0 + c))], jw[i]);
          }
        }
      }

      if (ADJ) {
        for (int g = 0; g < ngroups; ++g) {
          const int r0 = g * G;
          const int Gv = min(G, n2 - r0);
          const int tot = G * n1;
          const int totv = Gv * n1;
          float2 o[kTE];
#pragma unroll
          for (int i = 0; i < kTE; ++i) o[i] = make_float2(0.f, 0.f);
          for (int c = 0; c < cc; ++c) {
            // stage this coil's sample values by slot (j - sc0)
#pragma unroll
            for (int i = 0; i < kMaxJ; ++i) {
              float2 w = make_float2(0.f, 0.f);
#pragma unroll
              for (int cq = 0; cq < kCC; ++cq)
                if (cq == c) w = W[i][cq];
              wst[tid + i * kThreads] = w;
            }
            __syncthreads();
            // element (p, k1) of the tile sums conj(W2^{k2 (r0 + p)}) w_j over
            // the samples j of row k1 in this chunk, in CSR order; stored
            // real/imag-swapped for the inverse FFT
            const int* rp = a.rowptr + (long long)s * (n1 + 1);
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) {
                const int p = e / n1, k1 = e - p * n1;
                float2 v = make_float2(0.f, 0.f);
                if (p < Gv) {
                  const int r = r0 + p;
                  for (int q = rp[k1]; q < rp[k1 + 1]; ++q) {
                    const int ent = a.csr[o0 + q];
                    const int j = o0 + (ent & 0xffff);
                    if (j < sc0 || j >= sc1) continue;
                    const int k2 = ent >> 16;
                    v = cadd(v, cmulc(wst[j - sc0], tw2[(k2 * r) % n2]));
                  }
                }
                lds[e] = make_float2(v.y, v.x);
              }
            }
            __syncthreads();
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
            const float2* Sp = a.Smaps + (long long)(c0 + c) * nvox + i3 * nplane + (long long)r0 * n1;
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < totv) o[i] = cadd(o[i], cscale(cmulc(cswap(lds[e]), Sp[e]), inv_sqrt));
            }
            __syncthreads();
          }
          if (a.tadj) {  // first translation of T^H (along dim 1) after the coil sum
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) lds[e] = o[i];
            }
            __syncthreads();
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) {
                const int j = e % n1;
                lds[e] = cswap(cmul(lds[e], phase_cycles(ca, freq_index(j, n1))));
              }
            }
            __syncthreads();
            fft_tile<true>(lds, tw1, a.plan1, G, tid);
#pragma unroll
            for (int i = 0; i < kTE; ++i) {
              const int e = tid + i * kThreads;
              if (e < tot) o[i] = cscale(cswap(lds[e]), inv_n1);
            }
            __syncthreads();
          }
          float2* Vt = Vp + (long long)r0 * n1;
#pragma unroll
          for (int i = 0; i < kTE; ++i) {
            const int e = tid + i * kThreads;
            if (e < totv) Vt[e] = first_out ? o[i] : cadd(Vt[e], o[i]);
          }
        }
        first_out = false;
      }
    }
    if (MODE == COIL_ENERGY) {
#pragma unroll
      for (int i = 0; i < kMaxJ; ++i) {
        const int j = sc0 + tid + i * kThreads;
        if (jk1[i] >= 0) a.epart[j + (long long)a.Ns * i3] = esum[i];
      }
    }
    if (sc1 >= o1) break;
  }
}

hipError_t launch_coil_generic(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  if (a.fy1 > 0) return launch_coil_fold(a, mode, nstates, st);
  dim3 grid(a.i3n ? a.i3n : a.n3, nstates);
  switch (mode) {
    case COIL_ENCODE: hipLaunchKernelGGL(coil_kernel<COIL_ENCODE>, grid, dim3(kThreads), 0, st, a); break;
    case COIL_DECODE: hipLaunchKernelGGL(coil_kernel<COIL_DECODE>, grid, dim3(kThreads), 0, st, a); break;
    case COIL_EHE: hipLaunchKernelGGL(coil_kernel<COIL_EHE>, grid, dim3(kThreads), 0, st, a); break;
    case COIL_ENERGY: hipLaunchKernelGGL(coil_kernel<COIL_ENERGY>, grid, dim3(kThreads), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// Register path for n1 = 16 * R1.  Sixteen dim-2 rows per tile, sixteen
// lanes per row (regfft.hpp).  Workgroups are numbered so that the S states
// of one readout plane run back to back on one XCD: they share that plane's
// coil maps through the XCD's L2.  The adjoint adds each sample into the 16
// rows of a tile with a geometric recurrence over the rows; the lanes own
// either one phase-encode row each (samples of a row summed in CSR order)
// or, in adjoint-only launches of plans with a lane table (lanetab.hpp),
// equal runs of row-sorted samples, so that no wave waits on its most-sampled
// row.
constexpr int kRJ = 3;  // samples per thread per chunk on the register path
static_assert(kRJ * kThreads == kCoilRegChunk, "slot tables are built per chunk of kCoilRegChunk samples");
constexpr int kHornerBatch = 16;  // rows of a full group whose LDS reads may be in flight together

// Next (state, plane) item for a persistent workgroup whose home queue is
// `home`: items of queue k are planes i3 = k (mod 8), all states of a plane
// consecutive.  Returns k * per_queue + local, or -1 when every queue is dry.
__device__ __forceinline__ int queue_next(int* queue, int home, int per_queue) {
  for (int q = 0; q < 8; ++q) {
    const int k = (home + q) & 7;
    const int v = atomicAdd(queue + k, 1);
    if (v < per_queue) return k * per_queue + v;
  }
  return -1;
}

// BAL: adjoint-only instance driven by the plan's lane table (a.adjtab set,
// no fold along dim 1); kept a separate instance so that each kernel holds
// one adjoint loop.
template <int R1, int MODE, bool PERSIST = false, bool BAL = false>
__global__ __launch_bounds__(kThreads, 2) void rcoil_kernel(CoilArgs a, int nstates) {
  static_assert(!BAL || MODE == COIL_DECODE, "the lane-table adjoint is an adjoint-only launch");
  constexpr bool FWD = MODE == COIL_ENCODE || MODE == COIL_EHE || MODE == COIL_ENERGY;
  constexpr bool ADJ = MODE == COIL_DECODE || MODE == COIL_EHE;
  constexpr int N = R1 * 16;
  constexpr int XB = RegFftSize<R1>::XB;
  // lane-table adjoint: tile rows kThreads elements apart, and the lanes'
  // carry slots (one column per lane) in a region of their own at the same
  // stride, a multiple of 64 elements (512 bytes) so that a segment's 16
  // stores pair into ds_write2st64 (n1 = 224 with stride 480: 82 single
  // stores, 10 % slower).  The carries live from the scatter to the deferred
  // adds and the FFT exchange only after them, so the two share the region
  // and the exchange needs no workgroup barrier
  static_assert(!BAL || N <= kThreads, "the lane table covers n1 <= kThreads");
  constexpr int NS = BAL ? kThreads : N;
  constexpr int XCB = 16 * XB > 16 * kThreads ? 16 * XB : 16 * kThreads;
  // forward: the 16-row tile is stored column-major (spectrum column k1's
  // 16 rows contiguous, columns ZS = 18 elements apart, 16-byte aligned), so
  // a sample's Horner chain reads its column with 16-byte loads
  constexpr int ZS = 18;
  constexpr int XZF = FWD ? N * ZS : 0;
  constexpr int XZ0 = (16 * XB > 16 * NS) ? 16 * XB : 16 * NS;
  // forward-only launches: the FFT exchange has an LDS region of its own
  // behind the tile (80 KiB at N = 256, two workgroups per CU), so a
  // pencil's exchange is ordered within its wavefront only, and a wave's
  // next FFT runs while other waves still read the tile; one workgroup
  // barrier before each tile store and one after it (not in persistent
  // launches, whose queue slot would take the LDS past two workgroups per CU)
  constexpr bool SEPX = (MODE == COIL_ENCODE || MODE == COIL_ENERGY) && !PERSIST;
  constexpr int XZ = SEPX ? XZF + 16 * XB : BAL ? 16 * NS + XCB : (XZ0 > XZF ? XZ0 : XZF);
  constexpr int kEntCap = 2048;
  // the 16-row tile and the FFT exchange: aliased, or the exchange behind
  // the tile (SEPX), or behind the rows in the carries' region (BAL)
  __shared__ fvec2 xzraw[XZ];
  // row-owner adjoint: the sample values of the current coil and the
  // state's CSR by k1
  constexpr int kWcElems = kRJ * kThreads;
  constexpr int kCsrInts = N + 1 + kEntCap;
  constexpr int kRowOwnerElems = kWcElems + (kCsrInts + 1) / 2;
  __shared__ fvec2 adjraw[kRowOwnerElems];
  fvec2* const wcraw = adjraw;
  int* const csrraw = (int*)(adjraw + kWcElems);
  __shared__ fvec2 t1raw[N];
  __shared__ fvec2 t2raw[1024];
  const int tid = threadIdx.x;
  const int p = tid >> 4, t = tid & 15;
  lds_float2* tw1 = (lds_float2*)t1raw;
  lds_float2* tw2 = (lds_float2*)t2raw;
  lds_float2* xb = (lds_float2*)xzraw + (SEPX ? XZF : BAL ? 16 * NS : 0) + p * XB;
  lds_float2* Zt = (lds_float2*)xzraw;  // [16 rows][N] natural order (row stride NS)
  lds_float2* Cr = (lds_float2*)xzraw + 16 * NS;  // lane-table carries [16][kThreads] (BAL)
  const int n2 = a.n2;
  // dim-2 DFT of the sampled rows: length n2, or under SENSE folding along
  // dim 2 (fy1 == n1) the acquired length m2 = fy2 with the image rows
  // aliased at offset o2 = ceil((n2 - m2) / 2) (Methods/fold.m): the m2-point
  // DFT of the alias sum at k2 is sum_r W_m2^{k2 (r - o2)} v[r], so the same
  // row recurrences run with m2-point twiddles started at -k2 o2
  const bool fold2 = a.fy1 > 0;
  const int m2 = fold2 ? a.fy2 : n2;
  const int o2 = fold2 ? (n2 - m2 + 1) / 2 : 0;
  const int o2p = (m2 - o2 % m2) % m2;  // row shift equal to -o2 modulo m2, so k2 (r0 + o2p) >= 0
  // folding along dim 1 by an integer factor q1 = N / m1: the m1-point DFT
  // of the alias sum at k1 is P(k1) FFT_N(v)[q1 k1] with P(k1) = W_m1^{-k1 o1}
  // = exp(2 pi i k1 o1 / m1), so samples gather spectrum column q1 k1 and
  // take the phase; the adjoint publishes conj(P) y into column q1 k1 and
  // the inverse FFT replicates it periodically (Methods/ifold.m)
  const int m1 = fold2 ? a.fy1 : N;
  const int q1 = N / m1;
  const int o1 = (N - m1 + 1) / 2;
  auto fold1_phase = [&](int k1) -> float2 {
    if (q1 == 1) return make_float2(1.f, 0.f);
    float sn, cs;
    sincospif(2.0f * (float)((k1 * o1) % m1) / (float)m1, &sn, &cs);
    return make_float2(cs, sn);
  };
  const float2* tws = fold2 ? a.fy2tw : a.tw2;
  for (int k = tid; k < N; k += kThreads) st2(tw1 + k, a.tw1[k]);
  for (int k = tid; k < m2; k += kThreads) st2(tw2 + k, tws[k]);
  __syncthreads();

  // XCD-aware items: XCD x takes planes i3lo + x, i3lo + x + 8, ... with all
  // states of a plane consecutive (static: one item per workgroup from
  // blockIdx; persistent: items pulled from the per-XCD queues).
  const int n3e = a.i3n ? a.i3n : a.n3;
  const int per_queue = ((n3e + 7) / 8) * nstates;
  auto item_body = [&](const int sl, const int i3) {
  const int s = a.slist ? a.slist[sl] : a.state0 + sl;
  const long long nplane = (long long)N * n2;
  const long long nvox = nplane * a.n3;
  const float inv_sqrt = rsqrtf((float)m1 * (float)m2);
  const float inv_n = 1.0f / (float)N;
  const int o0 = a.off[s], o1 = a.off[s + 1];
  const float2* Up = FWD ? a.U + (a.us_rel ? sl : s) * a.us + i3 * nplane : nullptr;
  float2* Vp = ADJ ? a.V + (long long)(a.vs_rel ? sl : s) * a.vs + i3 * nplane : nullptr;
  const int ngroups = (n2 + 15) / 16;
  const double cf = a.tfwd ? a.tfwd[s] * a.invN1 : 0.0;
  const double ca = a.tadj ? a.tadj[s] * a.invN1 : 0.0;

  // adjoint-only launches with a lane table: every lane takes cs <= 3
  // row-sorted samples (one chunk per state)
  constexpr bool bal = BAL;
  int binfo = 0;
  bool first_out = true;
  for (int sc0 = o0; sc0 < o1 || (sc0 == o0 && ADJ); sc0 += kRJ * kThreads) {
    const int sc1 = min(o1, sc0 + kRJ * kThreads);
    int jk1[kRJ], jk2[kRJ], jstep[kRJ], jg[kRJ];
    float jw[kRJ];
    int bent[kRJ];
    if (bal) {
      const int ch = (sc0 - o0) / (kRJ * kThreads);  // this chunk's table block
      const int4 tb = ((const int4*)a.adjtab)[((long long)a.adjoff[s] + ch) * kThreads + tid];
      bent[0] = tb.x;
      bent[1] = tb.y;
      bent[2] = tb.z;
      binfo = tb.w;
    }
#pragma unroll
    for (int i = 0; i < kRJ; ++i) {
      if (bal) {
        // lane-table entry: row (k1), k2 and the sample within its state
        const bool ok = (binfo >> (kAdjValidShift + i)) & 1;
        const int j = ok ? o0 + ((bent[i] >> kAdjSampleShift) & 0x3fff) : -1;
        jg[i] = j;
        jk1[i] = ok ? (bent[i] >> kAdjRowShift) & 0xff : -1;
        jk2[i] = ok ? bent[i] & 1023 : 0;
        jw[i] = ok ? (a.We ? a.We[j] : 1.0f) : 0.f;
      } else {
        // slot -> sample: the plan deals each chunk's samples to slots so
        // that consecutive lanes read distinct LDS banks in the row gathers
        const int q = sc0 + tid + i * kThreads;
        const bool ok = q < sc1;
        const int j = ok ? (a.sord ? a.sord[q] : q) : -1;
        jg[i] = j;
        jk1[i] = ok ? a.k1[j] : -1;
        jk2[i] = ok ? a.k2[j] : 0;
        jw[i] = ok ? (a.We ? a.We[j] : 1.0f) : 0.f;
      }
      jstep[i] = (16 * jk2[i]) % m2;  // twiddle-index advance per 16-row group
    }
    float esum[kRJ];
#pragma unroll
    for (int i = 0; i < kRJ; ++i) esum[i] = 0.f;

    for (int c0 = 0; c0 < a.C; c0 += kCC) {
      const int cc = min(kCC, a.C - c0);
      float2 W[kRJ][kCC];
#pragma unroll
      for (int i = 0; i < kRJ; ++i)
#pragma unroll
        for (int c = 0; c < kCC; ++c) W[i][c] = make_float2(0.f, 0.f);

      if (FWD) {
        int ib[kRJ];  // W2 index of row r0 for each sample: (k2 (r0 - o2)) mod m2
#pragma unroll
        for (int i = 0; i < kRJ; ++i) ib[i] = (jk2[i] * o2p) % m2;
        for (int g = 0; g < ngroups; ++g) {
          const int r0 = g * 16;
          const int Gv = min(16, n2 - r0);
          const bool row_ok = p < Gv;
          const long long rowoff = (long long)(r0 + (row_ok ? p : 0)) * N;
          float2 u[16];
#pragma unroll
          for (int r = 0; r < R1; ++r) {
            const float2 ld = Up[rowoff + t + 16 * r];
            u[r] = row_ok ? ld : make_float2(0.f, 0.f);
          }
          if (a.tfwd) {  // last translation of T along dim 1
            reg_fft<R1, 16, SEPX>(u, t, xb, tw1);
            if (t < R1) {
              float2 ph[16];
              ramp16<R1>(ph, cf, t);
#pragma unroll
              for (int k = 0; k < 16; ++k) u[k] = cswap(cmul(u[k], ph[k]));
            }
            reg_fft<16, R1, SEPX>(u, t, xb, tw1);
#pragma unroll
            for (int r = 0; r < R1; ++r) u[r] = cscale(cswap(u[r]), inv_n);
          }
          // W[i][0] is the current coil's accumulator: the row of eight is
          // rotated by one after every coil (eight rotations per group, the
          // identity), so the sums take a static register index instead of a
          // select over all eight
          // encode launches: unrolled twice, half the rotation's moves (11.8 ->
          // 11.5 ms at 256x192x256); the energy instance unrolled so ran 2.5x
          // slower, and full unrolling overflows the instruction cache
          constexpr int kCoilUnroll = MODE == COIL_ENCODE ? 2 : 1;
#pragma unroll kCoilUnroll
          for (int c = 0; c < kCC; ++c) {
            if (c < cc) {
              const float2* Sp = a.Smaps + (long long)(c0 + c) * nvox + i3 * nplane + rowoff;
              float2 v[16];
#pragma unroll
              for (int r = 0; r < R1; ++r) v[r] = cmul(u[r], Sp[t + 16 * r]);
              reg_fft<R1, 16, SEPX>(v, t, xb, tw1);
              if (SEPX) __syncthreads();  // every wave's reads of the previous tile are done
              if (t < R1) {
#pragma unroll
                for (int k = 0; k < 16; ++k) st2(Zt + (t + R1 * k) * ZS + p, v[k]);
              }
              __syncthreads();
              if (Gv == 16) {
                // full group: no per-row guard, so the LDS reads of a sample's
                // Horner chain are issued kHornerBatch at a time
#pragma unroll
                for (int i = 0; i < kRJ; ++i) {
                  if (jk1[i] >= 0) {
                    const float2 a1 = ld2(tw2 + jk2[i]);
                    const lds_float2* zc = Zt + jk1[i] * q1 * ZS;  // the sample's column
                    float2 acc = ld2(zc + 15);
#pragma unroll
                    for (int q = 14; q >= 0; --q) {
                      const float2 z = ld2(zc + q);
                      acc = make_float2(fmaf(acc.x, a1.x, fmaf(-acc.y, a1.y, z.x)), fmaf(acc.x, a1.y, fmaf(acc.y, a1.x, z.y)));
                      if ((q & (kHornerBatch - 1)) == 0) __builtin_amdgcn_sched_barrier(0);
                    }
                    acc = cmul(acc, ld2(tw2 + ib[i]));
                    W[i][0] = cadd(W[i][0], acc);
                  }
                }
              } else {
#pragma unroll
                for (int i = 0; i < kRJ; ++i) {
                  if (jk1[i] >= 0) {
                    // sum_q W2^{k2 (r0 + q)} Z[q] = W2^{k2 r0} * (Horner in a = W2^{k2})
                    const float2 a1 = ld2(tw2 + jk2[i]);
                    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                    for (int q = 15; q >= 0; --q) {
                      if (q < Gv) {
                        const float2 z = ld2(Zt + jk1[i] * q1 * ZS + q);
                        acc = make_float2(fmaf(acc.x, a1.x, fmaf(-acc.y, a1.y, z.x)), fmaf(acc.x, a1.y, fmaf(acc.y, a1.x, z.y)));
                      }
                    }
                    acc = cmul(acc, ld2(tw2 + ib[i]));
                    W[i][0] = cadd(W[i][0], acc);
                  }
                }
              }
              if (!SEPX) __syncthreads();  // the row tile aliases the next FFT's exchange buffer
            }
#pragma unroll
            for (int i = 0; i < kRJ; ++i) {
              const float2 w0 = W[i][0];
#pragma unroll
              for (int cq = 0; cq + 1 < kCC; ++cq) W[i][cq] = W[i][cq + 1];
              W[i][kCC - 1] = w0;
            }
          }
#pragma unroll
          for (int i = 0; i < kRJ; ++i) {
            ib[i] += jstep[i];
            if (ib[i] >= m2) ib[i] -= m2;
          }
        }
#pragma unroll
        for (int i = 0; i < kRJ; ++i) {
          const int j = jg[i];
          if (jk1[i] >= 0) {
            const float2 ph = fold1_phase(jk1[i]);
#pragma unroll
            for (int c = 0; c < kCC; ++c) {
              if (c < cc) {
                const float2 Y = cscale(q1 > 1 ? cmul(W[i][c], ph) : W[i][c], inv_sqrt);
                const long long yi = j + (long long)a.Ns * (i3 + (long long)a.n3 * (c0 + c));
                if (MODE == COIL_ENCODE) a.yout[yi] = Y;
                if (MODE == COIL_ENERGY) {
                  const float2 r = csub(Y, a.yin[yi]);
                  esum[i] += r.x * r.x + r.y * r.y;
                }
                if (MODE == COIL_EHE) W[i][c] = cscale(q1 > 1 ? cmulc(Y, ph) : Y, jw[i]);
              }
            }
          }
        }
        __syncthreads();  // last sample gathers done before the adjoint reuses Zt
      }
      if (MODE == COIL_DECODE && a.Ns > 0) {
        // every load issued unconditionally (sample 0 and coil cc - 1 stand
        // in for unused slots) so that all of them are in flight together;
        // behind the per-sample and per-coil conditions each load was
        // waited for before the next one was issued
        float2 yv[kRJ][kCC];
#pragma unroll
        for (int i = 0; i < kRJ; ++i) {
          const long long jj = jk1[i] >= 0 ? jg[i] : 0;
#pragma unroll
          for (int c = 0; c < kCC; ++c)
            yv[i][c] = a.yin[jj + (long long)a.Ns * (i3 + (long long)a.n3 * (c0 + min(c, cc - 1)))];
        }
#pragma unroll
        for (int i = 0; i < kRJ; ++i) {
          if (jk1[i] >= 0) {
            const float2 ph = fold1_phase(jk1[i]);
#pragma unroll
            for (int c = 0; c < kCC; ++c)
              if (c < cc) W[i][c] = cscale(q1 > 1 ? cmulc(yv[i][c], ph) : yv[i][c], jw[i]);
          }
        }
      }

      if (ADJ && bal) {
        // Balanced adjoint from the lane table: lane L adds its (up to 3)
        // row-sorted samples into all 16 dim-2 rows of the group, writing
        // each row segment that closes in the lane straight into the tile
        // (a row begun in an earlier lane goes to the lane's carry slot
        // instead); the owner of a row that continues into later lanes adds
        // their carries in lane order once they are published.  Every lane
        // runs the same number of samples, and the sum is deterministic.
        // tile [16][NS]: columns < N the rows; Cr [16][kThreads]: column L
        // lane L's carry (both stored swapped)
        const bool deferred = (binfo & kAdjDeferred) != 0;
        const int ncont = (binfo >> kAdjNcontShift) & 1023;
        const int zrow = (binfo >> kAdjZrowShift) & 1023;
        const float fn2 = 1.0f / (float)m2;
        for (int g = 0; g < ngroups; ++g) {
          const int r0 = g * 16;
          const int Gv = min(16, n2 - r0);
          const bool row_ok = p < Gv;
          const int i2 = r0 + (row_ok ? p : 0);
          const long long rowoff = (long long)i2 * N;
          float2 o[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) o[k] = make_float2(0.f, 0.f);
          // W[i][0] holds coil c: the row of eight is rotated by one after
          // every coil (eight rotations per group, the identity)
          for (int c = 0; c < kCC; ++c) {
            if (c < cc) {
              __syncthreads();  // every wave's tile loads and FFT exchanges (sharing Cr) of the previous coil are done
              float2 acc[16];
#pragma unroll
              for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
              int lrow = 0;
#pragma unroll
              for (int i = 0; i < kRJ; ++i) {
                if (jk1[i] >= 0) {
                  const float2 w = W[i][0];  // coil c (the row is rotated after every coil)
                  const int k2 = jk2[i];
                  const int prod = k2 * (r0 + o2p);  // = k2 (r0 - o2) modulo m2
                  int idx = prod - m2 * (int)((float)prod * fn2);
                  idx += idx < 0 ? m2 : 0;
                  idx -= idx >= m2 ? m2 : 0;
                  // conj(W2^{k2 (r0+q)}) w = (w conj(W2^{k2 r0})) conj(W2^{k2})^q
                  float2 wq = cmulc(w, ld2(tw2 + idx));
                  const float2 pst = ld2(tw2 + k2);
                  // a segment closed at the previous entry restarts the sum
                  const float keep = i > 0 && ((binfo >> (kAdjCloseShift + i - 1)) & 1) ? 0.f : 1.f;
#pragma unroll
                  for (int q = 0; q < 16; ++q) {
                    acc[q] = make_float2(fmaf(acc[q].x, keep, wq.x), fmaf(acc[q].y, keep, wq.y));
                    wq = cmulc(wq, pst);
                  }
                  lrow = jk1[i];
                  if ((binfo >> (kAdjCloseShift + i)) & 1) {
                    const bool to_carry = (binfo >> (kAdjCarryShift + i)) & 1;
                    lds_float2* dst = to_carry ? Cr + tid : Zt + lrow;  // same stride NS
#pragma unroll
                    for (int q = 0; q < 16; ++q) st2(dst + q * NS, cswap(acc[q]));
                  }
                }
              }
              if (binfo & kAdjZvalid) {  // an empty row dealt to this lane
#pragma unroll
                for (int q = 0; q < 16; ++q) st2(Zt + q * NS + zrow, make_float2(0.f, 0.f));
              }
              __syncthreads();  // carries published
              if (deferred) {
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[q] = cswap(acc[q]);
                for (int l = 1; l <= ncont; ++l) {
#pragma unroll
                  for (int q = 0; q < 16; ++q) acc[q] = cadd(acc[q], ld2(Cr + q * NS + tid + l));
                }
#pragma unroll
                for (int q = 0; q < 16; ++q) st2(Zt + q * NS + lrow, acc[q]);
              }
              __syncthreads();
              float2 v[16];
#pragma unroll
              for (int r = 0; r < R1; ++r) v[r] = ld2(Zt + p * NS + t + 16 * r);
              // the exchange region holds only carries, read before the barrier above
              reg_fft<R1, 16, true>(v, t, xb, tw1);  // inverse (swapped data): lane k1 holds k1 + R1 k2
              if (t < R1) {
                const float2* Sp = a.Smaps + (long long)(c0 + c) * nvox + i3 * nplane + rowoff;
                // all 16 map loads in flight before the first use: with the
                // registers near their limit the scheduler otherwise pairs
                // every load with its wait (n1 = 224: 16 serial round trips
                // per row group and coil, the adjoint 15.5 -> 26.5 ms)
                float2 sv[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) sv[k] = Sp[t + R1 * k];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < 16; ++k) o[k] = cadd(o[k], cscale(cmulc(cswap(v[k]), sv[k]), inv_sqrt));
              }
            }
#pragma unroll
            for (int i = 0; i < kRJ; ++i) {
              const float2 w0 = W[i][0];
#pragma unroll
              for (int cq = 0; cq + 1 < kCC; ++cq) W[i][cq] = W[i][cq + 1];
              W[i][kCC - 1] = w0;
            }
          }
          if (a.tadj) {  // first translation of T^H along dim 1 (input layout of reg_fft<16,R1>)
            reg_fft<16, R1, true>(o, t, xb, tw1);
            {
              float2 ph[16];
              ramp_in<R1>(ph, ca, t);
#pragma unroll
              for (int r = 0; r < R1; ++r) o[r] = cswap(cmul(o[r], ph[r]));
            }
            reg_fft<R1, 16, true>(o, t, xb, tw1);
#pragma unroll
            for (int k = 0; k < 16; ++k) o[k] = cscale(cswap(o[k]), inv_n);
          }
          if (t < R1 && row_ok) {
            float2* Vt = Vp + rowoff;
#pragma unroll
            for (int k = 0; k < 16; ++k) Vt[t + R1 * k] = first_out ? o[k] : cadd(Vt[t + R1 * k], o[k]);
          }
        }
        first_out = false;
      } else if (ADJ) {
        // Atomic-free adjoint: the weighted samples of coil c are published
        // in LDS; lane k owns phase-encode row k1 = k and adds each of its
        // samples into all 16 dim-2 rows of the group (exclusive writes),
        // then every row is inverse-FFT'd along dim 1.
        lds_float2* wc = (lds_float2*)wcraw;
        typedef __attribute__((address_space(3))) int lds_int;
        lds_int* rpl = (lds_int*)csrraw;
        lds_int* entl = rpl + (N + 1);
        const int* rp = a.rowptr + (long long)s * (N + 1);
        const int jl0 = sc0 - o0, jl1 = sc1 - o0;
        const float fn2 = 1.0f / (float)m2;
        const int nent = o1 - o0;
        const bool csr_lds = nent <= kEntCap;
        if (csr_lds) {
          for (int k = tid; k <= N; k += kThreads) rpl[k] = rp[k];
          for (int e = tid; e < nent; e += kThreads) entl[e] = a.csr[o0 + e];
        }
        for (int g = 0; g < ngroups; ++g) {
          const int r0 = g * 16;
          const int Gv = min(16, n2 - r0);
          const bool row_ok = p < Gv;
          const int i2 = r0 + (row_ok ? p : 0);
          const long long rowoff = (long long)i2 * N;
          float2 o[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) o[k] = make_float2(0.f, 0.f);
          for (int c = 0; c < cc; ++c) {
#pragma unroll
            for (int i = 0; i < kRJ; ++i) {
              float2 w = make_float2(0.f, 0.f);
#pragma unroll
              for (int cq = 0; cq < kCC; ++cq)
                if (cq == c) w = W[i][cq];
              if (jg[i] >= 0) st2(wc + (jg[i] - sc0), w);
            }
            __syncthreads();
            for (int k1 = tid; k1 < N; k1 += kThreads) {
              float2 acc[16];
#pragma unroll
              for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
              // spectrum column k1 holds acquired row k1 / q1 when q1 divides it
              const int ka = k1 / q1;
              const bool own = k1 == ka * q1;
              const int e0 = own ? (csr_lds ? rpl[ka] : rp[ka]) : 0;
              const int e1 = own ? (csr_lds ? rpl[ka + 1] : rp[ka + 1]) : 0;
              for (int e = e0; e < e1; ++e) {
                const int ent = csr_lds ? entl[e] : a.csr[o0 + e];
                const int jl = ent & 0xffff;
                if (jl >= jl0 && jl < jl1) {
                  const int k2 = ent >> 16;
                  const float2 w = ld2(wc + (jl - jl0));
                  const int prod = k2 * (r0 + o2p);  // = k2 (r0 - o2) modulo m2
                  int idx = prod - m2 * (int)((float)prod * fn2);
                  idx += idx < 0 ? m2 : 0;
                  idx -= idx >= m2 ? m2 : 0;
                  // conj(W2^{k2 (r0+q)}) w = (w conj(W2^{k2 r0})) conj(W2^{k2})^q
                  float2 wq = cmulc(w, ld2(tw2 + idx));
                  const float2 pst = ld2(tw2 + k2);
#pragma unroll
                  for (int q = 0; q < 16; ++q) {
                    acc[q] = cadd(acc[q], wq);
                    wq = cmulc(wq, pst);
                  }
                }
              }
#pragma unroll
              for (int q = 0; q < 16; ++q) st2(Zt + q * N + k1, cswap(acc[q]));  // swapped for the inverse FFT
            }
            __syncthreads();
            float2 v[16];
#pragma unroll
            for (int r = 0; r < R1; ++r) v[r] = ld2(Zt + p * N + t + 16 * r);
            __syncthreads();  // the row tile aliases the exchange buffer
            reg_fft<R1, 16>(v, t, xb, tw1);  // inverse (swapped data): lane k1 holds k1 + R1 k2
            if (t < R1) {
              const float2* Sp = a.Smaps + (long long)(c0 + c) * nvox + i3 * nplane + rowoff;
              float2 sv[16];  // all map loads in flight before the first use
#pragma unroll
              for (int k = 0; k < 16; ++k) sv[k] = Sp[t + R1 * k];
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int k = 0; k < 16; ++k) o[k] = cadd(o[k], cscale(cmulc(cswap(v[k]), sv[k]), inv_sqrt));
            }
          }
          if (a.tadj) {  // first translation of T^H along dim 1 (input layout of reg_fft<16,R1>)
            reg_fft<16, R1>(o, t, xb, tw1);
            {
              float2 ph[16];
              ramp_in<R1>(ph, ca, t);
#pragma unroll
              for (int r = 0; r < R1; ++r) o[r] = cswap(cmul(o[r], ph[r]));
            }
            reg_fft<R1, 16>(o, t, xb, tw1);
#pragma unroll
            for (int k = 0; k < 16; ++k) o[k] = cscale(cswap(o[k]), inv_n);
          }
          if (t < R1 && row_ok) {
            float2* Vt = Vp + rowoff;
#pragma unroll
            for (int k = 0; k < 16; ++k) Vt[t + R1 * k] = first_out ? o[k] : cadd(Vt[t + R1 * k], o[k]);
          }
        }
        first_out = false;
      }
    }
    if (MODE == COIL_ENERGY) {
#pragma unroll
      for (int i = 0; i < kRJ; ++i) {
        const int j = jg[i];
        if (jk1[i] >= 0) a.epart[j + (long long)a.Ns * i3] = esum[i];
      }
    }
    if (sc1 >= o1) break;
  }
  };  // item_body
  if (PERSIST) {
    __shared__ int item_sh;
    for (;;) {
      if (tid == 0) item_sh = queue_next(a.queue, blockIdx.x & 7, per_queue);
      __syncthreads();
      const int item = item_sh;
      __syncthreads();
      if (item < 0) break;  // uniform: every thread read the same item
      const int xcd = item / per_queue, local = item - xcd * per_queue;
      const int i3 = (local / nstates) * 8 + xcd;
      if (i3 < n3e) {
        item_body(local - (local / nstates) * nstates, a.i3lo + i3);
        __syncthreads();  // the next item's LDS writes follow every read of this one
      }
    }
  } else {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int i3 = (local / nstates) * 8 + xcd;
    if (i3 < n3e) item_body(local - (local / nstates) * nstates, a.i3lo + i3);
  }
}

__global__ void zero_queue_kernel(int* q) {
  if (threadIdx.x < 8) q[threadIdx.x] = 0;
}

// the adjoint-only launch takes the lane-table instance when the plan has a
// table and dim 1 is not folded (q1 = 1)
static bool lane_table_launch(const CoilArgs& a, int mode) {
  return mode == COIL_DECODE && a.adjtab != nullptr && (a.fy1 == 0 || a.fy1 == a.n1);
}

template <int R1>
hipError_t launch_rcoil(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  const int nblk = (((a.i3n ? a.i3n : a.n3) + 7) / 8) * 8 * nstates;
  dim3 grid(nblk);
  if (lane_table_launch(a, mode)) {
    hipLaunchKernelGGL((rcoil_kernel<R1, COIL_DECODE, false, true>), grid, dim3(kThreads), 0, st, a, nstates);
    return hipGetLastError();
  }
  switch (mode) {
    case COIL_ENCODE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_ENCODE>), grid, dim3(kThreads), 0, st, a, nstates); break;
    case COIL_DECODE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_DECODE>), grid, dim3(kThreads), 0, st, a, nstates); break;
    case COIL_EHE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_EHE>), grid, dim3(kThreads), 0, st, a, nstates); break;
    case COIL_ENERGY: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_ENERGY>), grid, dim3(kThreads), 0, st, a, nstates); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int R1>
hipError_t launch_rcoil_persistent(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  static int resident[5] = {0, 0, 0, 0, 0};  // per mode, [4]: the lane-table adjoint
  const bool lt = lane_table_launch(a, mode);
  const int mi = lt ? 4 : mode - 1;
  if (mi < 0 || mi > 4) return hipErrorInvalidValue;
  if (!resident[mi]) {
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) {
      switch (lt ? 0 : mode) {
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rcoil_kernel<R1, COIL_DECODE, true, true>, kThreads, 0); break;
        case COIL_ENCODE: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rcoil_kernel<R1, COIL_ENCODE, true>, kThreads, 0); break;
        case COIL_DECODE: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rcoil_kernel<R1, COIL_DECODE, true>, kThreads, 0); break;
        case COIL_EHE: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rcoil_kernel<R1, COIL_EHE, true>, kThreads, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rcoil_kernel<R1, COIL_ENERGY, true>, kThreads, 0); break;
      }
    }
    if (e != hipSuccess) return e;
    resident[mi] = std::max(8, (ncu * std::max(1, per_cu)) / 8 * 8);
  }
  const int items = (((a.i3n ? a.i3n : a.n3) + 7) / 8) * 8 * nstates;
  const int nblk = std::max(8, std::min(resident[mi], (items + 7) / 8 * 8));
  hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, st, a.queue);
  dim3 grid(nblk);
  if (lt) {
    hipLaunchKernelGGL((rcoil_kernel<R1, COIL_DECODE, true, true>), grid, dim3(kThreads), 0, st, a, nstates);
    return hipGetLastError();
  }
  switch (mode) {
    case COIL_ENCODE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_ENCODE, true>), grid, dim3(kThreads), 0, st, a, nstates); break;
    case COIL_DECODE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_DECODE, true>), grid, dim3(kThreads), 0, st, a, nstates); break;
    case COIL_EHE: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_EHE, true>), grid, dim3(kThreads), 0, st, a, nstates); break;
    default: hipLaunchKernelGGL((rcoil_kernel<R1, COIL_ENERGY, true>), grid, dim3(kThreads), 0, st, a, nstates); break;
  }
  return hipGetLastError();
}

bool coil_fold_register(int n1, int n2, int fy1) {
  return fy1 > 0 && n1 % fy1 == 0 && n2 <= 1024 && reg_radix(n1) != 0;
}

// a folded plan stays on the register path when dim 1 folds by an integer
// factor (or not at all)
static bool fold_stage(const CoilArgs& a) {
  return a.fy1 > 0 && (a.force_generic || !coil_fold_register(a.n1, a.n2, a.fy1));
}

hipError_t launch_coil_persistent(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  if (a.i3lo < 0 || a.i3n < 0 || a.i3lo + (a.i3n ? a.i3n : a.n3) > a.n3) return hipErrorInvalidValue;
  if (fold_stage(a)) return launch_coil_fold(a, mode, nstates, st);
  if (!a.queue || a.n2 > 1024 || a.force_generic) return hipErrorInvalidValue;
  switch (reg_radix(a.n1)) {
    case 16: return launch_rcoil_persistent<16>(a, mode, nstates, st);
    case 14: return launch_rcoil_persistent<14>(a, mode, nstates, st);
    case 12: return launch_rcoil_persistent<12>(a, mode, nstates, st);
    case 10: return launch_rcoil_persistent<10>(a, mode, nstates, st);
    case 9: return launch_rcoil_persistent<9>(a, mode, nstates, st);
    case 8: return launch_rcoil_persistent<8>(a, mode, nstates, st);
    case 6: return launch_rcoil_persistent<6>(a, mode, nstates, st);
    case 4: return launch_rcoil_persistent<4>(a, mode, nstates, st);
    case 3: return launch_rcoil_persistent<3>(a, mode, nstates, st);
    case 2: return launch_rcoil_persistent<2>(a, mode, nstates, st);
    default: return hipErrorInvalidValue;
  }
}

bool coil_queue_available(int n1, int n2) { return n2 <= 1024 && reg_radix(n1) != 0; }


int coil_chunk(int n1, int n2) {
  return (n2 <= 1024 && reg_radix(n1)) ? kRJ * kThreads : kMaxJ * kThreads;
}

hipError_t launch_coil(const CoilArgs& a, int mode, int nstates, hipStream_t st) {
  if (a.i3lo < 0 || a.i3n < 0 || a.i3lo + (a.i3n ? a.i3n : a.n3) > a.n3) return hipErrorInvalidValue;
  if (fold_stage(a)) return launch_coil_fold(a, mode, nstates, st);
  if (a.n2 <= 1024 && !a.force_generic) {
    switch (reg_radix(a.n1)) {
      case 16: return launch_rcoil<16>(a, mode, nstates, st);
      case 14: return launch_rcoil<14>(a, mode, nstates, st);
      case 12: return launch_rcoil<12>(a, mode, nstates, st);
      case 10: return launch_rcoil<10>(a, mode, nstates, st);
      case 9: return launch_rcoil<9>(a, mode, nstates, st);
      case 8: return launch_rcoil<8>(a, mode, nstates, st);
      case 6: return launch_rcoil<6>(a, mode, nstates, st);
      case 4: return launch_rcoil<4>(a, mode, nstates, st);
      case 3: return launch_rcoil<3>(a, mode, nstates, st);
      case 2: return launch_rcoil<2>(a, mode, nstates, st);
      default: break;
    }
  }
  return launch_coil_generic(a, mode, nstates, st);
}

}  // namespace dis
