// Reversible integer 5/3 wavelet transform (the step before encoding and after
// decoding; the reference takes transformed coefficients, main_example.cu:71-179).
//
// Lifting with whole-sample symmetric extension, for x[0..2M):
//   d[n] = x[2n+1] - floor((x[2n] + x[2n+2]) / 2)     x[2M] = x[2M-2]
//   s[n] = x[2n]   + floor((d[n-1] + d[n] + 2) / 4)   d[-1] = d[0]
// Each level transforms the columns, then the rows, of the current LL band and
// stores LL / HL / LH / HH in the Mallat layout (HL right, LH below, HH diagonal),
// the layout Find_Subband (BPC_PaCo.cu:84-111) reads.
//
// One workgroup per 64x64 tile of a level: the tile and a 2-sample halo (taken
// with the symmetric extension at the band edges) are staged in LDS, all four
// lifting steps run there, and the four 32x32 band tiles are written out, so
// every sample of a level is read from HBM once (plus the halo) and written once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bpc_internal.h"

namespace bpc {
namespace {

constexpr int TILE = 64;
constexpr int HALO = 2;
constexpr int W = TILE + 2 * HALO;  // 68 staged rows and columns
constexpr int LP = W + 1;           // LDS row pitch (odd: column walks hit distinct banks)
constexpr int THREADS = 256;
constexpr int NLOAD = (W * W + THREADS - 1) / THREADS;  // staged samples per thread

// Whole-sample symmetric extension of index g into [0, n), n >= 2.
__device__ __forceinline__ int reflect(int g, int n) {
  if ((unsigned)g < (unsigned)n) return g;
  if (g < 0 && -g < n) return -g;
  if (g >= n && g < 2 * n - 1) return 2 * (n - 1) - g;
  const int p = 2 * (n - 1);
  g %= p;
  if (g < 0) g += p;
  return g < n ? g : p - g;
}

__global__ __launch_bounds__(THREADS) void dwt53_forward_kernel(const DwtLevel a) {
  __shared__ int32_t t[W][LP];
  const int r0 = blockIdx.y * TILE, c0 = blockIdx.x * TILE;
  const int tid = threadIdx.x;
  // every load of the thread is issued before the first LDS store, so the tile
  // costs one HBM round trip rather than one per staged sample
  int32_t v[NLOAD];
#pragma unroll
  for (int q = 0; q < NLOAD; ++q) {
    const int k = tid + q * THREADS;
    if (k < W * W) {
      const int lr = k / W, lc = k - lr * W;
      const int gr = reflect(r0 - HALO + lr, a.ny), gc = reflect(c0 - HALO + lc, a.nx);
      v[q] = a.src[(long long)gr * a.src_pitch + gc];
    }
  }
#pragma unroll
  for (int q = 0; q < NLOAD; ++q) {
    const int k = tid + q * THREADS;
    if (k < W * W) t[k / W][k % W] = v[q];
  }
  __syncthreads();
  // vertical predict: odd staged rows 1..65 (every column)
  for (int k = tid; k < 33 * W; k += THREADS) {
    const int i = 1 + 2 * (k / W), c = k % W;
    t[i][c] -= (t[i - 1][c] + t[i + 1][c]) >> 1;
  }
  __syncthreads();
  // vertical update: even staged rows 2..64
  for (int k = tid; k < 32 * W; k += THREADS) {
    const int i = 2 + 2 * (k / W), c = k % W;
    t[i][c] += (t[i - 1][c] + t[i + 1][c] + 2) >> 2;
  }
  __syncthreads();
  // horizontal predict on the tile's rows 2..65: odd columns 1..65
  for (int k = tid; k < TILE * 33; k += THREADS) {
    const int r = 2 + k / 33, j = 1 + 2 * (k % 33);
    t[r][j] -= (t[r][j - 1] + t[r][j + 1]) >> 1;
  }
  __syncthreads();
  // horizontal update: even columns 2..64
  for (int k = tid; k < TILE * 32; k += THREADS) {
    const int r = 2 + k / 32, j = 2 + 2 * (k % 32);
    t[r][j] += (t[r][j - 1] + t[r][j + 1] + 2) >> 2;
  }
  __syncthreads();
  const int hy = a.ny >> 1, hx = a.nx >> 1;
  for (int k = tid; k < 4 * 32 * 32; k += THREADS) {
    const int band = k >> 10, i = (k >> 5) & 31, j = k & 31;
    const int n = (r0 >> 1) + i, m = (c0 >> 1) + j;
    if (n >= hy || m >= hx) continue;
    const int32_t v = t[2 + 2 * i + (band >> 1)][2 + 2 * j + (band & 1)];
    if (band == 0)
      a.ll[(long long)n * a.ll_pitch + m] = v;
    else
      a.det[(long long)(n + ((band >> 1) ? hy : 0)) * a.det_pitch + m + ((band & 1) ? hx : 0)] = v;
  }
}

__global__ __launch_bounds__(THREADS) void dwt53_inverse_kernel(const DwtLevel a) {
  __shared__ int32_t t[W][LP];
  const int r0 = blockIdx.y * TILE, c0 = blockIdx.x * TILE;
  const int tid = threadIdx.x;
  const int hy = a.ny >> 1, hx = a.nx >> 1;
  // interleaved samples y (y[2n] = low band n, y[2n+1] = high band n) from the
  // bands, all loads in flight before the LDS stores
  int32_t v[NLOAD];
#pragma unroll
  for (int q = 0; q < NLOAD; ++q) {
    const int k = tid + q * THREADS;
    if (k < W * W) {
      const int lr = k / W, lc = k - lr * W;
      const int gr = reflect(r0 - HALO + lr, a.ny), gc = reflect(c0 - HALO + lc, a.nx);
      const int32_t* p = (!(gr & 1) && !(gc & 1))
                             ? a.ll + (long long)(gr >> 1) * a.ll_pitch + (gc >> 1)
                             : a.det + (long long)((gr & 1) ? hy + (gr >> 1) : (gr >> 1)) * a.det_pitch +
                                   ((gc & 1) ? hx + (gc >> 1) : (gc >> 1));
      v[q] = *p;
    }
  }
#pragma unroll
  for (int q = 0; q < NLOAD; ++q) {
    const int k = tid + q * THREADS;
    if (k < W * W) t[k / W][k % W] = v[q];
  }
  __syncthreads();
  // horizontal inverse update: even columns 2..66, every staged row
  for (int k = tid; k < W * 33; k += THREADS) {
    const int r = k / 33, j = 2 + 2 * (k % 33);
    t[r][j] -= (t[r][j - 1] + t[r][j + 1] + 2) >> 2;
  }
  __syncthreads();
  // horizontal inverse predict: odd columns 3..65
  for (int k = tid; k < W * 32; k += THREADS) {
    const int r = k / 32, j = 3 + 2 * (k % 32);
    t[r][j] += (t[r][j - 1] + t[r][j + 1]) >> 1;
  }
  __syncthreads();
  // vertical inverse update: even rows 2..66 on the tile's columns 2..65
  for (int k = tid; k < 33 * TILE; k += THREADS) {
    const int i = 2 + 2 * (k / TILE), c = 2 + (k % TILE);
    t[i][c] -= (t[i - 1][c] + t[i + 1][c] + 2) >> 2;
  }
  __syncthreads();
  // vertical inverse predict: odd rows 3..65
  for (int k = tid; k < 32 * TILE; k += THREADS) {
    const int i = 3 + 2 * (k / TILE), c = 2 + (k % TILE);
    t[i][c] += (t[i - 1][c] + t[i + 1][c]) >> 1;
  }
  __syncthreads();
  for (int k = tid; k < TILE * TILE; k += THREADS) {
    const int i = k >> 6, j = k & 63;
    const int gr = r0 + i, gc = c0 + j;
    if (gr < a.ny && gc < a.nx) a.out[(long long)gr * a.out_pitch + gc] = t[2 + i][2 + j];
  }
}

}  // namespace

hipError_t launch_dwt53_level(int inverse, const DwtLevel& a, hipStream_t stream) {
  if (a.nx < 2 || a.ny < 2 || (a.nx & 1) || (a.ny & 1)) return hipErrorInvalidValue;
  const dim3 grid((a.nx + TILE - 1) / TILE, (a.ny + TILE - 1) / TILE);
  if (inverse)
    hipLaunchKernelGGL(dwt53_inverse_kernel, grid, dim3(THREADS), 0, stream, a);
  else
    hipLaunchKernelGGL(dwt53_forward_kernel, grid, dim3(THREADS), 0, stream, a);
  return hipGetLastError();
}

}  // namespace bpc
