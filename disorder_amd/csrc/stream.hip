// Streaming-copy probes: the HBM read+write rate this GPU sustains for a
// plain device copy under several access shapes, the measured ceiling the
// HBM-bound pencil passes are compared with.  No counterpart in the
// reference (it is a measurement of the hardware, not of the method).
//
//   variant 0: grid-stride, one float2 (8 B) per lane per iteration
//   variant 1: grid-stride, one float4 (16 B) per lane per iteration
//   variant 2: grid-stride, four float4 per lane per iteration (64 B in flight)
//   variant 3: variant 2 with nontemporal loads and stores
//   variant 4: the pencil passes' shape: a workgroup copies one 32 KiB tile,
//              16 float2 loads per lane, then 16 stores
//   variant 5: variant 4 with each workgroup looping over tiles and the loads
//              of the next tile issued before the stores of the current one
//   variant 6: the readout-axis register pass's access pattern on whole
//              256 x 192 x 256 complex64 volumes (16 consecutive dim-1
//              positions at one dim-2 position, every readout plane)
//   variant 7: the dim-2 register pass's pattern on the same volumes (16
//              consecutive dim-1 positions, every dim-2 row of one plane)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/disorder.h"
#include "kernels.hpp"
#include "transform.hpp"

namespace dis {

namespace {

__global__ __launch_bounds__(256) void copy_f2(const float2* __restrict__ in, float2* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) out[i] = in[i];
}

__global__ __launch_bounds__(256) void copy_f4(const float4* __restrict__ in, float4* __restrict__ out, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) out[i] = in[i];
}

typedef float vf4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(256) void copy_f4x4(const vf4* __restrict__ in, vf4* __restrict__ out, long long n) {
  const long long stride = (long long)gridDim.x * 256;
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    vf4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = NT ? __builtin_nontemporal_load(in + i + k * stride) : in[i + k * stride];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (NT)
        __builtin_nontemporal_store(v[k], out + i + k * stride);
      else
        out[i + k * stride] = v[k];
    }
  }
  for (; i < n; i += stride) out[i] = in[i];
}

// one 4096-element (32 KiB) tile per workgroup: lanes (p = tid / 16, t =
// tid % 16) take elements p * 256 + t + 16 r, r < 16, as the contiguous
// register passes do
__global__ __launch_bounds__(256) void copy_tile(const float2* __restrict__ in, float2* __restrict__ out,
                                                 long long ntiles) {
  const long long tile = blockIdx.x;
  if (tile >= ntiles) return;
  const int p = threadIdx.x >> 4, t = threadIdx.x & 15;
  const long long base = tile * 4096 + p * 256 + t;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = in[base + 16 * r];
#pragma unroll
  for (int r = 0; r < 16; ++r) out[base + 16 * r] = v[r];
}

__global__ __launch_bounds__(256) void copy_tile_pipe(const float2* __restrict__ in, float2* __restrict__ out,
                                                      long long ntiles) {
  const int p = threadIdx.x >> 4, t = threadIdx.x & 15;
  long long tile = blockIdx.x;
  if (tile >= ntiles) return;
  float2 v[16], w[16];
  long long base = tile * 4096 + p * 256 + t;
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = in[base + 16 * r];
  for (;;) {
    const long long nxt = tile + gridDim.x;
    const long long nb = nxt * 4096 + p * 256 + t;
    if (nxt < ntiles) {
#pragma unroll
      for (int r = 0; r < 16; ++r) w[r] = in[nb + 16 * r];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) out[base + 16 * r] = v[r];
    if (nxt >= ntiles) break;
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = w[r];
    tile = nxt;
    base = nb;
  }
}

// variants 6 and 7: lane (p = tid % 16, t = tid / 16) copies element
// t + 16 r (r < R) of the pencil at dim-1 position i10 + p; the tile of
// workgroup b is dealt XCD-contiguously as in the passes (grid sizes 3072
// and 4096, multiples of 8)
constexpr int kPN1 = 256, kPN2 = 192, kPN3 = 256;
constexpr long long kPV = (long long)kPN1 * kPN2 * kPN3;

template <bool READOUT>
__global__ __launch_bounds__(256) void copy_pencils(const float2* __restrict__ in, float2* __restrict__ out) {
  constexpr int R = READOUT ? kPN3 / 16 : kPN2 / 16;
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int tile = (b & 7) * (n >> 3) + (b >> 3);
  const int line = tile / (kPN1 / 16), i10 = (tile % (kPN1 / 16)) * 16;
  const int p = threadIdx.x & 15, t = threadIdx.x >> 4;
  const long long js = READOUT ? (long long)kPN1 * kPN2 : kPN1;
  const long long base = (long long)blockIdx.y * kPV + i10 + p + (READOUT ? (long long)kPN1 * line : (long long)kPN1 * kPN2 * line);
  float2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = in[base + (long long)(t + 16 * r) * js];
#pragma unroll
  for (int r = 0; r < R; ++r) out[base + (long long)(t + 16 * r) * js] = v[r];
}

}  // namespace

}  // namespace dis

using namespace dis;

extern "C" int dis_stream_probe(dis_ctx* c, const void* in, void* out, int64_t nbytes, int32_t variant,
                                int32_t blocks, int32_t reps, double* ms) {
  if (!c) return DIS_E_INVALID;
  if (!in || !out || nbytes <= 0 || (nbytes & 32767) || variant < 0 || variant > 7 || reps <= 0 || !ms)
    return ctx_fail(c, DIS_E_INVALID, "dis_stream_probe: need nbytes a positive multiple of 32 KiB, variant 0..7");
  const long long nvol = nbytes / (long long)(kPV * sizeof(float2));
  if (variant >= 6 && (nvol < 1 || nvol > 65535))
    return ctx_fail(c, DIS_E_INVALID, "dis_stream_probe: variants 6 and 7 need at least one 256 x 192 x 256 volume");
  if (ctx_set_device(c) != hipSuccess) return ctx_fail(c, DIS_E_HIP, "hipSetDevice failed");
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_stream_probe: device query failed");
  const long long ntiles = nbytes / 32768;
  const int nb = blocks > 0 ? blocks : (variant == 4 ? (int)ntiles : ncu * 8);
  hipStream_t s = ctx_stream(c);
  auto launch = [&]() {
    switch (variant) {
      case 0: hipLaunchKernelGGL(copy_f2, dim3(nb), dim3(256), 0, s, (const float2*)in, (float2*)out, nbytes / 8); break;
      case 1: hipLaunchKernelGGL(copy_f4, dim3(nb), dim3(256), 0, s, (const float4*)in, (float4*)out, nbytes / 16); break;
      case 2: hipLaunchKernelGGL(copy_f4x4<false>, dim3(nb), dim3(256), 0, s, (const vf4*)in, (vf4*)out, nbytes / 16); break;
      case 3: hipLaunchKernelGGL(copy_f4x4<true>, dim3(nb), dim3(256), 0, s, (const vf4*)in, (vf4*)out, nbytes / 16); break;
      case 4: hipLaunchKernelGGL(copy_tile, dim3((unsigned)std::max<long long>(nb, ntiles)), dim3(256), 0, s, (const float2*)in, (float2*)out, ntiles); break;
      case 5: hipLaunchKernelGGL(copy_tile_pipe, dim3(nb), dim3(256), 0, s, (const float2*)in, (float2*)out, ntiles); break;
      case 6: hipLaunchKernelGGL(copy_pencils<true>, dim3((kPN1 / 16) * kPN2, (unsigned)nvol), dim3(256), 0, s, (const float2*)in, (float2*)out); break;
      default: hipLaunchKernelGGL(copy_pencils<false>, dim3((kPN1 / 16) * kPN3, (unsigned)nvol), dim3(256), 0, s, (const float2*)in, (float2*)out); break;
    }
  };
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return ctx_fail(c, DIS_E_HIP, "dis_stream_probe: event creation failed");
  launch();  // warm-up
  (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps; ++r) launch();
  (void)hipEventRecord(e1, s);
  float t = 0.f;
  const hipError_t er = hipEventSynchronize(e1);
  if (er == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (er != hipSuccess || hipGetLastError() != hipSuccess) return ctx_fail(c, DIS_E_HIP, "dis_stream_probe: launch failed");
  *ms = (double)t / reps;
  return DIS_OK;
}
