// Compact codestream container on the device.
//
// The reference layout (BPC_PaCo.cu:1130,1147-1157,1284) gives every codeblock a
// fixed region of int32 words: the MSB, the used codeword slots (16-bit values),
// then -1.  The compact form keeps, per codeblock, the MSB as one byte and the
// codewords as packed uint16 at offsets[cb] .. offsets[cb + 1] of one payload
// array; offsets[] is the exclusive scan of the codeword counts.
//
// All four kernels are HBM-bound byte movers: one wave per codeblock, lanes
// striding over the codewords (coalesced 256-byte reads, 128-byte writes).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bpc_internal.h"

namespace bpc {
namespace {

typedef unsigned long long u64;

constexpr int WAVES = 4;  // waves per workgroup

__device__ __forceinline__ bool msb_word_ok(int32_t w) { return (w >= 0 && w <= 27) || w == 32; }

// Per codeblock: MSB byte and codeword count (into offsets[cb + 1]).  The count is
// ncw[cb] when given, else the number of words before the first -1 after the MSB
// (the encoder's fill, BPC_PaCo.cu:1284), or stride - 1 when the region is full.
__global__ __launch_bounds__(64 * WAVES) void compact_index_kernel(const int32_t* stream, long long ncb, int stride,
                                                                   const int32_t* ncw, uint8_t* msb, u64* offsets,
                                                                   uint32_t* status) {
  const long long cb = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cb >= ncb) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  const int32_t* region = stream + cb * (long long)stride;
  const int32_t w0 = region[0];
  const bool ok = msb_word_ok(w0);
  long long n;
  bool bad = false;
  if (ncw) {
    n = ncw[cb];
    bad = n < 0 || n > stride - 1;
    if (bad) n = 0;
  } else {
    n = stride - 1;
    for (int base = 1; base < stride; base += 64) {
      const int j = base + lane;
      const u64 m = __builtin_amdgcn_ballot_w64(j < stride && region[j] == -1);
      if (m) {
        n = base + __builtin_ctzll(m) - 1;
        break;
      }
    }
  }
  if (lane == 0) {
    msb[cb] = ok ? (uint8_t)w0 : (uint8_t)255;
    offsets[cb + 1] = (u64)n;
    if (!ok) atomicOr(status, ST_MALFORMED);
    if (bad) atomicOr(status, ST_OVERFLOW);
  }
}

// In-place inclusive scan of offsets[1 .. ncb] by one workgroup (offsets[0] = 0):
// each thread sums a contiguous chunk, the chunk sums are scanned in LDS, then each
// thread rewrites its chunk as running sums.
__global__ __launch_bounds__(1024) void compact_scan_kernel(u64* offsets, long long ncb) {
  __shared__ u64 part[1024];
  const int t = threadIdx.x;
  const long long chunk = (ncb + 1023) / 1024;
  const long long b = (long long)t * chunk;
  const long long e = b + chunk < ncb ? b + chunk : ncb;
  u64 sum = 0;
  for (long long k = b; k < e; ++k) sum += offsets[1 + k];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const u64 v = t >= off ? part[t - off] : 0ull;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  u64 run = t ? part[t - 1] : 0ull;
  for (long long k = b; k < e; ++k) {
    run += offsets[1 + k];
    offsets[1 + k] = run;
  }
  if (t == 0) offsets[0] = 0;
}

// Codewords of every region into the payload (16 bits each, cap codewords).
__global__ __launch_bounds__(64 * WAVES) void compact_pack_kernel(const int32_t* stream, long long ncb, int stride,
                                                                  const u64* offsets, uint16_t* payload, u64 cap,
                                                                  uint32_t* status) {
  const long long cb = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cb >= ncb) return;
  const int lane = threadIdx.x & 63;
  const u64 o = offsets[cb];
  const long long n = (long long)(offsets[cb + 1] - o);
  if (n < 0 || n > stride - 1 || o + (u64)n > cap) {  // not an index of this stream and payload
    if (lane == 0) atomicOr(status, ST_OVERFLOW);
    return;
  }
  const int32_t* src = stream + cb * (long long)stride + 1;
  uint16_t* dst = payload + o;
  for (long long j = lane; j < n; j += 64) dst[j] = (uint16_t)src[j];
}

// Compact form back to reference regions: MSB word, codewords, -1 in the rest.
// A codeblock whose count exceeds the region keeps only its MSB (overflow status).
__global__ __launch_bounds__(64 * WAVES) void compact_unpack_kernel(const uint8_t* msb, const u64* offsets,
                                                                    const uint16_t* payload, u64 cap, long long ncb,
                                                                    int stride, int32_t* stream, uint32_t* status) {
  const long long cb = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cb >= ncb) return;
  const int lane = threadIdx.x & 63;
  const u64 o = offsets[cb];
  long long n = (long long)(offsets[cb + 1] - o);
  if (n < 0 || n > stride - 1 || o + (u64)n > cap) {
    if (lane == 0) atomicOr(status, ST_OVERFLOW);
    n = 0;
  }
  int32_t* region = stream + cb * (long long)stride;
  if (lane == 0) region[0] = msb[cb];
  const uint16_t* src = payload + o;
  for (long long j = lane; j < stride - 1; j += 64) region[1 + j] = j < n ? (int32_t)src[j] : -1;
}

}  // namespace

hipError_t launch_compact_index(const int32_t* stream, long long ncb, int stride, const int32_t* ncw, uint8_t* msb,
                                unsigned long long* offsets, uint32_t* status, hipStream_t s) {
  if (ncb <= 0) return hipSuccess;
  const long long blocks = (ncb + WAVES - 1) / WAVES;
  hipLaunchKernelGGL(compact_index_kernel, dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, stream, ncb, stride, ncw,
                     msb, offsets, status);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, s, offsets, ncb);
  return hipGetLastError();
}

hipError_t launch_compact_pack(const int32_t* stream, long long ncb, int stride, const unsigned long long* offsets,
                               uint16_t* payload, unsigned long long cap, uint32_t* status, hipStream_t s) {
  if (ncb <= 0) return hipSuccess;
  const long long blocks = (ncb + WAVES - 1) / WAVES;
  hipLaunchKernelGGL(compact_pack_kernel, dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, stream, ncb, stride,
                     offsets, payload, (u64)cap, status);
  return hipGetLastError();
}

hipError_t launch_compact_unpack(const uint8_t* msb, const unsigned long long* offsets, const uint16_t* payload,
                                 unsigned long long cap, long long ncb, int stride, int32_t* stream, uint32_t* status,
                                 hipStream_t s) {
  if (ncb <= 0) return hipSuccess;
  const long long blocks = (ncb + WAVES - 1) / WAVES;
  hipLaunchKernelGGL(compact_unpack_kernel, dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, msb, offsets, payload,
                     (u64)cap, ncb, stride, stream, status);
  return hipGetLastError();
}

}  // namespace bpc
