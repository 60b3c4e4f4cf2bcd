// Compact codestream container on the device.
//
// The reference layout (BPC_PaCo.cu:1130,1147-1157,1284) gives every codeblock a
// fixed region of int32 words: the MSB, the used codeword slots (16-bit values),
// then -1.  The compact form keeps, per codeblock, the MSB as one byte and the
// codewords as packed uint16 at offsets[cb] .. offsets[cb + 1] of one payload
// array; offsets[] is the exclusive scan of the codeword counts.
//
// All four kernels are HBM-bound byte movers: one wave per codeblock, lanes
// striding over the codewords; regions move as 16-byte word groups when they are
// 16-byte aligned.
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
  __shared__ u64 part[32];
  const int t = threadIdx.x;
  const long long chunk = (ncb + 1023) / 1024;
  const long long b = (long long)t * chunk;
  const long long e = b + chunk < ncb ? b + chunk : ncb;
  u64 sum = 0;
  for (long long k = b; k < e; ++k) sum += offsets[1 + k];
  // inclusive scan of the chunk sums: within each wave by shuffles, then over the
  // 16 wave totals by the first wave (two barriers instead of twenty)
  const int lane = t & 63, wid = t >> 6;
  u64 x = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const u64 y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) part[wid] = x;
  __syncthreads();
  if (wid == 0) {
    u64 v = lane < 16 ? part[lane] : 0ull;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const u64 y = __shfl_up(v, off, 64);
      if (lane >= off) v += y;
    }
    if (lane < 16) part[16 + lane] = v;
  }
  __syncthreads();
  u64 run = x - sum + (wid ? part[16 + wid - 1] : 0ull);  // exclusive prefix of this thread's chunk
  for (long long k = b; k < e; ++k) {
    run += offsets[1 + k];
    offsets[1 + k] = run;
  }
  if (t == 0) offsets[0] = 0;
}

// Codewords of every region into the payload (16 bits each, cap codewords).
// vec: regions are 16-byte aligned (stream aligned, stride a multiple of 4 words)
// and the payload 4-byte aligned.
__global__ __launch_bounds__(64 * WAVES) void compact_pack_kernel(const int32_t* stream, long long ncb, int stride,
                                                                  const u64* offsets, uint16_t* payload, u64 cap,
                                                                  uint32_t* status, int vec) {
  const long long cb = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cb >= ncb) return;
  const int lane = threadIdx.x & 63;
  const u64 o = offsets[cb];
  const long long n = (long long)(offsets[cb + 1] - o);
  // not an index of this stream and payload (o + n <= cap, tested without wrapping)
  if (n < 0 || n > stride - 1 || (u64)n > cap || o > cap - (u64)n) {
    if (lane == 0) atomicOr(status, ST_OVERFLOW);
    return;
  }
  const int32_t* region = stream + cb * (long long)stride;
  uint16_t* dst = payload + o;  // word w of the region (1 <= w <= n) goes to dst[w - 1]
  if (vec) {
    // 16-byte reads of word groups [4q, 4q + 4): a quarter of the memory round trips
    const long long ng = (n + 4) / 4;  // groups over words 0..n
    for (long long q = lane; q < ng; q += 64) {
      const int4 v = *reinterpret_cast<const int4*>(region + 4 * q);
      const long long w = 4 * q;
      if (w >= 1 && w + 3 <= n) {
        // a whole group: codeword pairs at even payload positions as 4-byte stores
        // (o, hence the parity, is the same for the whole wave)
        uint16_t* d = dst + (w - 1);
        if (((o + (u64)w - 1) & 1) == 0) {
          reinterpret_cast<uint32_t*>(d)[0] = ((uint32_t)v.x & 0xFFFFu) | ((uint32_t)v.y << 16);
          reinterpret_cast<uint32_t*>(d)[1] = ((uint32_t)v.z & 0xFFFFu) | ((uint32_t)v.w << 16);
        } else {
          d[0] = (uint16_t)v.x;
          *reinterpret_cast<uint32_t*>(d + 1) = ((uint32_t)v.y & 0xFFFFu) | ((uint32_t)v.z << 16);
          d[3] = (uint16_t)v.w;
        }
      } else {
        if (w >= 1 && w <= n) dst[w - 1] = (uint16_t)v.x;
        if (w + 1 <= n) dst[w] = (uint16_t)v.y;
        if (w + 2 <= n) dst[w + 1] = (uint16_t)v.z;
        if (w + 3 <= n) dst[w + 2] = (uint16_t)v.w;
      }
    }
  } else {
    for (long long j = 1 + lane; j <= n; j += 64) dst[j - 1] = (uint16_t)region[j];
  }
}

// Compact form back to reference regions: MSB word, codewords, -1 in the rest.
// A codeblock whose count exceeds the region keeps only its MSB (overflow status).
__global__ __launch_bounds__(64 * WAVES) void compact_unpack_kernel(const uint8_t* msb, const u64* offsets,
                                                                    const uint16_t* payload, u64 cap, long long ncb,
                                                                    int stride, int32_t* stream, uint32_t* status,
                                                                    int vec) {
  const long long cb = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cb >= ncb) return;
  const int lane = threadIdx.x & 63;
  const u64 o = offsets[cb];
  long long n = (long long)(offsets[cb + 1] - o);
  if (n < 0 || n > stride - 1 || (u64)n > cap || o > cap - (u64)n) {  // o + n <= cap, without wrapping
    if (lane == 0) atomicOr(status, ST_OVERFLOW);
    n = 0;
  }
  int32_t* region = stream + cb * (long long)stride;
  const uint16_t* src = payload + o;  // region word w (1 <= w <= n) is src[w - 1]
  if (vec) {
    // 16-byte stores of word groups [4q, 4q + 4): a quarter of the store instructions
    const int32_t m = msb[cb];
    for (long long q = lane; q < stride / 4; q += 64) {
      const long long w = 4 * q;
      int4 v;
      v.x = w == 0 ? m : (w <= n ? (int32_t)src[w - 1] : -1);
      v.y = w + 1 <= n ? (int32_t)src[w] : -1;
      v.z = w + 2 <= n ? (int32_t)src[w + 1] : -1;
      v.w = w + 3 <= n ? (int32_t)src[w + 2] : -1;
      *reinterpret_cast<int4*>(region + w) = v;
    }
  } else {
    if (lane == 0) region[0] = msb[cb];
    for (long long j = 1 + lane; j < stride; j += 64) region[j] = j <= n ? (int32_t)src[j - 1] : -1;
  }
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
  const int vec = ((uintptr_t)stream % 16 == 0 && stride % 4 == 0 && (uintptr_t)payload % 4 == 0) ? 1 : 0;
  hipLaunchKernelGGL(compact_pack_kernel, dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, stream, ncb, stride,
                     offsets, payload, (u64)cap, status, vec);
  return hipGetLastError();
}

hipError_t launch_compact_unpack(const uint8_t* msb, const unsigned long long* offsets, const uint16_t* payload,
                                 unsigned long long cap, long long ncb, int stride, int32_t* stream, uint32_t* status,
                                 hipStream_t s) {
  if (ncb <= 0) return hipSuccess;
  const long long blocks = (ncb + WAVES - 1) / WAVES;
  const int vec = ((uintptr_t)stream % 16 == 0 && stride % 4 == 0) ? 1 : 0;
  hipLaunchKernelGGL(compact_unpack_kernel, dim3((unsigned)blocks), dim3(64 * WAVES), 0, s, msb, offsets, payload,
                     (u64)cap, ncb, stride, stream, status, vec);
  return hipGetLastError();
}

}  // namespace bpc
