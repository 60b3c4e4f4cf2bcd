// Work queue of the batched codec (bpc_batch_run): codeblocks of all images sorted by
// MSB inside each image, paired there, and the pairs ordered by (pass mode, MSB)
// with the costliest first.  Counting sorts over 29 MSB buckets per image and 58
// pair buckets; every counter update is aggregated over the lanes of a wave that
// share a bucket.  The persistent coder that drains the queue is bpc_batch_kernel
// (bpc_kernels.hip).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bpc_internal.h"

namespace bpc {
namespace {

typedef uint64_t u64;
typedef uint32_t u32;

__device__ __forceinline__ u64 ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

__device__ __forceinline__ u32 msb_bucket(int msb) {
  return (msb >= 0 && msb <= 27) ? (u32)(27 - msb) : (u32)(NBUCKET - 1);
}

// OR of a 32-bit value over the wave (DPP row shifts and broadcasts; lane 63 ends
// with the OR of every lane).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ u32 dpp_or(u32 v) {
  return v | (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}
__device__ __forceinline__ u32 wave_or32(u32 v) {
  v = dpp_or<0x111, 0xF>(v);  // row_shr:1
  v = dpp_or<0x112, 0xF>(v);  // row_shr:2
  v = dpp_or<0x114, 0xF>(v);  // row_shr:4
  v = dpp_or<0x118, 0xF>(v);  // row_shr:8
  v = dpp_or<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_or<0x143, 0xC>(v);  // row_bcast:31
  return (u32)__builtin_amdgcn_readlane((int)v, 63);
}

// Copies the run's staging area (its size is in the run words at its head) from
// pinned host memory into device memory, 16 bytes per thread, and clears the
// queue's control words (kernels rather than a copy and a memset: the copy is a
// read of host memory by the first kernel of the run, and captured graphs hold no
// memset nodes).
__global__ __launch_bounds__(256) void batch_stage_kernel(const char* h_stage, char* d_stage, u32* ctl) {
  const BatchRun* hr = reinterpret_cast<const BatchRun*>(h_stage);
  const u32 bytes = hr->stage_bytes;
  const int words = CTL_IMG + 32 * hr->n_images;
  const u32 tid = blockIdx.x * blockDim.x + threadIdx.x, nthreads = gridDim.x * blockDim.x;
  for (u32 k = tid; k < bytes / 16u; k += nthreads)
    reinterpret_cast<uint4*>(d_stage)[k] = reinterpret_cast<const uint4*>(h_stage)[k];
  for (int k = (int)tid; k < words; k += (int)nthreads) ctl[k] = 0;
}

// Adds 1 to counter[key] for every active lane, with one atomic per distinct key in
// the wave (the keys of neighbouring threads are mostly equal here: codeblocks and
// pairs arrive in MSB order), and returns each lane's counter value: the old value
// plus the lane's rank among the lanes with the same key.
__device__ __forceinline__ u32 wave_atomic_inc(u32* counter, u32 key) {
  const int lane = threadIdx.x & 63;
  u32 result = 0;
  u64 todo = ballot(true);
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const u32 lkey = (u32)__builtin_amdgcn_readlane((int)key, leader);
    const u64 same = ballot(key == lkey) & todo;
    u32 base = 0;
    if (lane == leader) base = atomicAdd(counter + lkey, (u32)__builtin_popcountll(same));
    base = (u32)__builtin_amdgcn_readlane((int)base, leader);
    if ((same >> lane) & 1ull)
      result = base + __builtin_amdgcn_mbcnt_hi((u32)(same >> 32), __builtin_amdgcn_mbcnt_lo((u32)same, 0u));
    todo &= ~same;
  }
  return result;
}

__device__ __forceinline__ u32 mag(int v) { return v < 0 ? 0u - (u32)v : (u32)v; }

// Image owning global codeblock gcb: the last k with cb_offset[k] <= gcb.
__device__ __forceinline__ int image_of(const long long* cb_offset, int n_images, long long gcb) {
  int lo = 0, hi = n_images - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cb_offset[mid] <= gcb) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Grid-stride index of a thread and the thread count of the grid.
__device__ __forceinline__ long long thread_index() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ long long grid_threads() { return (long long)gridDim.x * blockDim.x; }

// Encoder keys: one wave per codeblock ORs its magnitudes; MSB bucket, counted per image.
__global__ __launch_bounds__(256) void queue_keys_kernel(const QueueArgs a) {
  const long long total = a.run->total_cbs;
  const int n_images = a.run->n_images, rows = a.run->rows;
  const int lane = threadIdx.x & 63;
  const long long nwaves = grid_threads() >> 6;
  for (long long gcb = thread_index() >> 6; gcb < total; gcb += nwaves) {  // wave-uniform
    const int img = image_of(a.cb_offset, n_images, gcb);
    const BatchImage& im = a.images[img];
    const int cb = (int)(gcb - a.cb_offset[img]);
    const int tx = cb % im.nx, ty = cb / im.nx;
    const int64_t base = im.layout ? (int64_t)cb * 64 * rows : (int64_t)ty * rows * im.size_x + tx * 64;
    const int rs = im.layout ? 64 : im.size_x;
    u32 m = 0;
    if (im.in16) {
      const int16_t* p0 = reinterpret_cast<const int16_t*>(im.src) + base;
      if ((((uintptr_t)p0 | ((uintptr_t)rs * 2u)) & 7u) == 0) {
        // 8 bytes per lane: four rows of 64 coefficients per wave load
        const int col = (lane & 15) * 4;
#pragma unroll 4
        for (int r = lane >> 4; r < rows; r += 4) {
          const short4 v = *reinterpret_cast<const short4*>(p0 + (int64_t)r * rs + col);
          m |= mag(v.x) | mag(v.y) | mag(v.z) | mag(v.w);
        }
      } else {
        for (int r = 0; r < rows; ++r) m |= mag(p0[(int64_t)r * rs + lane]);
      }
    } else {
      const int32_t* p0 = im.src + base;
      if ((((uintptr_t)p0 | ((uintptr_t)rs * 4u)) & 15u) == 0) {
        // 16 bytes per lane: four rows of 64 coefficients per wave load, loads in flight
        const int col = (lane & 15) * 4;
#pragma unroll 4
        for (int r = lane >> 4; r < rows; r += 4) {
          const int4 v = *reinterpret_cast<const int4*>(p0 + (int64_t)r * rs + col);
          m |= mag(v.x) | mag(v.y) | mag(v.z) | mag(v.w);
        }
      } else {
        for (int r = 0; r < rows; ++r) m |= mag(p0[(int64_t)r * rs + lane]);
      }
    }
    m = wave_or32(m);
    const int msb = m ? 31 - __clz(m) : 32;
    if (lane == 0) {
      const u32 b = msb_bucket(msb);
      a.keys[gcb] = b;
      atomicAdd(&a.ctl[CTL_IMG + img * 32 + b], 1u);
    }
  }
}

// Decoder keys: one thread per codeblock reads the MSB word of its region.
__global__ __launch_bounds__(256) void queue_keys_decode_kernel(const QueueArgs a) {
  const long long total = a.run->total_cbs;
  const int n_images = a.run->n_images;
  for (long long gcb = thread_index(); gcb < total; gcb += grid_threads()) {
    const int img = image_of(a.cb_offset, n_images, gcb);
    const BatchImage& im = a.images[img];
    const u32 b = msb_bucket(im.src[(gcb - a.cb_offset[img]) * im.stride]);
    a.keys[gcb] = b;
    (void)wave_atomic_inc(a.ctl + CTL_IMG, (u32)img * 32u + b);
  }
}

// One thread per image: bucket cursors inside the image's (even-padded) range.
__global__ void queue_image_scan_kernel(const QueueArgs a) {
  const int n_images = a.run->n_images;
  for (int img = (int)thread_index(); img < n_images; img += (int)grid_threads()) {
    u32 pos = (u32)a.sorted_base[img];
    for (int b = 0; b < NBUCKET; ++b) {
      u32* c = &a.ctl[CTL_IMG + img * 32 + b];
      const u32 n = *c;
      *c = pos;
      pos += n;
    }
    const long long ncb = a.cb_offset[img + 1] - a.cb_offset[img];
    if (ncb & 1) a.sorted[pos] = QUEUE_EMPTY;
  }
}

__global__ __launch_bounds__(256) void queue_scatter_kernel(const QueueArgs a) {
  const long long total = a.run->total_cbs;
  const int n_images = a.run->n_images;
  for (long long gcb = thread_index(); gcb < total; gcb += grid_threads()) {
    const int img = image_of(a.cb_offset, n_images, gcb);
    const u32 pos = wave_atomic_inc(a.ctl + CTL_IMG, (u32)img * 32u + a.keys[gcb]);
    a.sorted[pos] = queue_entry((u32)img, (u32)(gcb - a.cb_offset[img]));
  }
}

// One thread per pair: key (pass mode, bucket of its costlier codeblock).
__global__ __launch_bounds__(256) void queue_pair_keys_kernel(const QueueArgs a) {
  const long long total = a.run->total_pairs;
  for (long long p = thread_index(); p < total; p += grid_threads()) {
    const u32 e = a.sorted[2 * p];
    const int img = (int)queue_entry_image(e);
    const u32 b = a.keys[a.cb_offset[img] + queue_entry_cb(e)];
    const u32 k = (a.images[img].cpass3 ? NBUCKET : 0) + b;
    a.pair_keys[p] = k;
    (void)wave_atomic_inc(a.ctl + CTL_PAIRHIST, k);
  }
}

__global__ void queue_pair_scan_kernel(const QueueArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  u32 pos = 0;
  for (int k = 0; k < 2 * NBUCKET; ++k) {
    const u32 n = a.ctl[CTL_PAIRHIST + k];
    a.ctl[CTL_PAIRHIST + k] = pos;
    pos += n;
  }
}

__global__ __launch_bounds__(256) void queue_pair_scatter_kernel(const QueueArgs a) {
  const long long total = a.run->total_pairs;
  for (long long p = thread_index(); p < total; p += grid_threads()) {
    const u32 pos = wave_atomic_inc(a.ctl + CTL_PAIRHIST, a.pair_keys[p]);
    a.pair_queue[pos] = (u32)p;
  }
}

}  // namespace

// Grids of the grid-stride kernels depend on the device only (not on the run), so
// that a captured sequence replays for any batch: 32 waves per CU (a full CU) for
// the one-wave-per-codeblock encoder keys, which stream the whole input, and 4
// workgroups of 256 threads per CU for the one-thread-per-item kernels.
hipError_t launch_queue_build(const QueueArgs& a, hipStream_t stream) {
  const unsigned wave_grid = (unsigned)a.cus * 8, thread_grid = (unsigned)a.cus * 4;
  if (a.decode)
    hipLaunchKernelGGL(queue_keys_decode_kernel, dim3(thread_grid), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL(queue_keys_kernel, dim3(wave_grid), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(queue_image_scan_kernel, dim3((a.max_images + 63) / 64), dim3(64), 0, stream, a);
  hipLaunchKernelGGL(queue_scatter_kernel, dim3(thread_grid), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(queue_pair_keys_kernel, dim3(thread_grid), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(queue_pair_scan_kernel, dim3(1), dim3(64), 0, stream, a);
  hipLaunchKernelGGL(queue_pair_scatter_kernel, dim3(thread_grid), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_batch_stage(const char* h_stage, char* d_stage, uint32_t* ctl, hipStream_t stream) {
  hipLaunchKernelGGL(batch_stage_kernel, dim3(64), dim3(256), 0, stream, h_stage, d_stage, ctl);
  return hipGetLastError();
}
}  // namespace bpc
