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
constexpr int THREADS = 256;        // four waves: lane tx owns a column (or a row), wave ty a residue mod 4
constexpr int NR = W / 4;           // staged rows per thread

// Workgroups are dealt to the XCDs round robin (workgroup w runs on XCD w % 8),
// and each XCD has its own L2.  A tile's halo columns and band-edge samples share
// 128-byte lines with its neighbours, so the launch gives each XCD one contiguous
// run of tiles in raster order: horizontal (and most vertical) neighbours then
// run on the same XCD, close in time, and their shared lines come from its L2
// instead of HBM.  Returns false for the workgroups past the last tile.
constexpr int XCDS = 8;

__device__ __forceinline__ bool tile_of_workgroup(const DwtLevel& a, int& r0, int& c0) {
  const int ntx = (a.nx + TILE - 1) / TILE, nty = (a.ny + TILE - 1) / TILE;
  const int per = (ntx * nty + XCDS - 1) / XCDS;
  const int w = blockIdx.x;
  const int t = (w % XCDS) * per + w / XCDS;
  if (t >= ntx * nty) return false;
  r0 = (t / ntx) * TILE;
  c0 = (t % ntx) * TILE;
  return true;
}

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

// Thread layout shared by both directions.  Lane tx of wave ty stages column
// tx + 2 of the staged rows ty, ty + 4, ...; lanes 0..3 also stage one halo
// column each (0, 1, 66, 67).  Vertical lifting keeps that column ownership and
// splits the rows of a step by residue mod 4 over the waves; horizontal lifting
// gives lane tx a row and splits the columns the same way, so no index needs a
// division and every LDS access of a wave falls in distinct banks.
struct Lane {
  int tx, ty, hc;
  bool halo;
  __device__ Lane() : tx(threadIdx.x & 63), ty(threadIdx.x >> 6), hc((threadIdx.x & 63) < 2 ? (threadIdx.x & 63) : TILE + (threadIdx.x & 63)), halo((threadIdx.x & 63) < 4) {}
};

// d[i] -= floor((x[i-1] + x[i+1]) / 2) down a column, on odd staged rows 1 + 2k
template <int SIGN>
__device__ __forceinline__ void col_predict(int32_t (*t)[LP], int c, int k0, int kend) {
  for (int k = k0; k < kend; k += 4) {
    const int i = 1 + 2 * k + (SIGN > 0 ? 2 : 0);
    t[i][c] += SIGN * ((t[i - 1][c] + t[i + 1][c]) >> 1);
  }
}
// s[i] += floor((d[i-1] + d[i+1] + 2) / 4) down a column, on even staged rows 2 + 2k
template <int SIGN>
__device__ __forceinline__ void col_update(int32_t (*t)[LP], int c, int k0, int kend) {
  for (int k = k0; k < kend; k += 4) {
    const int i = 2 + 2 * k;
    t[i][c] += SIGN * ((t[i - 1][c] + t[i + 1][c] + 2) >> 2);
  }
}

__global__ __launch_bounds__(THREADS) void dwt53_forward_kernel(const DwtLevel a) {
  __shared__ int32_t t[W][LP];
  int r0, c0;
  if (!tile_of_workgroup(a, r0, c0)) return;
  const Lane l;
  const int gc = reflect(c0 + l.tx, a.nx), ghc = reflect(c0 - HALO + l.hc, a.nx);
  // every load of the thread is issued before the first LDS store, so the tile
  // costs one HBM round trip
  int32_t v[NR], h[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int32_t* row = a.src + (long long)reflect(r0 - HALO + l.ty + 4 * q, a.ny) * a.src_pitch;
    v[q] = row[gc];
    if (l.halo) h[q] = row[ghc];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    t[l.ty + 4 * q][l.tx + HALO] = v[q];
    if (l.halo) t[l.ty + 4 * q][l.hc] = h[q];
  }
  __syncthreads();
  // vertical predict: odd staged rows 1..65, every staged column
  col_predict<-1>(t, l.tx + HALO, l.ty, 33);
  if (l.halo) col_predict<-1>(t, l.hc, l.ty, 33);
  __syncthreads();
  // vertical update: even staged rows 2..64
  col_update<1>(t, l.tx + HALO, l.ty, 32);
  if (l.halo) col_update<1>(t, l.hc, l.ty, 32);
  __syncthreads();
  // horizontal predict on the tile's rows 2..65: odd columns 1..65
  {
    int32_t* r = t[HALO + l.tx];
    for (int k = l.ty; k < 33; k += 4) {
      const int j = 1 + 2 * k;
      r[j] -= (r[j - 1] + r[j + 1]) >> 1;
    }
  }
  __syncthreads();
  // horizontal update: even columns 2..64
  {
    int32_t* r = t[HALO + l.tx];
    for (int k = l.ty; k < 32; k += 4) {
      const int j = 2 + 2 * k;
      r[j] += (r[j - 1] + r[j + 1] + 2) >> 2;
    }
  }
  __syncthreads();
  // band rows: lanes 0..31 write the low-pass column half (LL or LH), lanes
  // 32..63 the high-pass half (HL or HH); a wave reads one staged row whole
  const int hy = a.ny >> 1, hx = a.nx >> 1;
  const int j = l.tx & 31, hi = l.tx >> 5;
  const int m = (c0 >> 1) + j;
  if (m >= hx) return;
#pragma unroll
  for (int vb = 0; vb < 2; ++vb) {
    for (int i = l.ty; i < 32; i += 4) {
      const int n = (r0 >> 1) + i;
      if (n >= hy) break;
      const int32_t val = t[HALO + 2 * i + vb][HALO + 2 * j + hi];
      if (vb == 0 && hi == 0)
        a.ll[(long long)n * a.ll_pitch + m] = val;
      else
        a.det[(long long)(n + (vb ? hy : 0)) * a.det_pitch + m + (hi ? hx : 0)] = val;
    }
  }
}

__global__ __launch_bounds__(THREADS) void dwt53_inverse_kernel(const DwtLevel a) {
  __shared__ int32_t t[W][LP];
  int r0, c0;
  if (!tile_of_workgroup(a, r0, c0)) return;
  const Lane l;
  const int hy = a.ny >> 1, hx = a.nx >> 1;
  // interleaved samples y (y[2n] = low band n, y[2n+1] = high band n) from the
  // bands, all loads in flight before the LDS stores
  const int gc = reflect(c0 + l.tx, a.nx), ghc = reflect(c0 - HALO + l.hc, a.nx);
  const int sc = (gc & 1) ? hx + (gc >> 1) : (gc >> 1), shc = (ghc & 1) ? hx + (ghc >> 1) : (ghc >> 1);
  int32_t v[NR], h[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int gr = reflect(r0 - HALO + l.ty + 4 * q, a.ny);
    const int32_t* drow = a.det + (long long)((gr & 1) ? hy + (gr >> 1) : (gr >> 1)) * a.det_pitch;
    const int32_t* lrow = a.ll + (long long)(gr >> 1) * a.ll_pitch;
    const bool even = !(gr & 1);
    v[q] = (even && !(gc & 1)) ? lrow[gc >> 1] : drow[sc];
    if (l.halo) h[q] = (even && !(ghc & 1)) ? lrow[ghc >> 1] : drow[shc];
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    t[l.ty + 4 * q][l.tx + HALO] = v[q];
    if (l.halo) t[l.ty + 4 * q][l.hc] = h[q];
  }
  __syncthreads();
  // horizontal inverse update: even columns 2..66 of every staged row (lane tx
  // takes row tx, lanes 0..3 also rows 64..67)
  for (int k = l.ty; k < 33; k += 4) {
    const int j = 2 + 2 * k;
    int32_t* r = t[l.tx];
    r[j] -= (r[j - 1] + r[j + 1] + 2) >> 2;
    if (l.halo) {
      r = t[TILE + l.tx];
      r[j] -= (r[j - 1] + r[j + 1] + 2) >> 2;
    }
  }
  __syncthreads();
  // horizontal inverse predict: odd columns 3..65
  for (int k = l.ty; k < 32; k += 4) {
    const int j = 3 + 2 * k;
    int32_t* r = t[l.tx];
    r[j] += (r[j - 1] + r[j + 1]) >> 1;
    if (l.halo) {
      r = t[TILE + l.tx];
      r[j] += (r[j - 1] + r[j + 1]) >> 1;
    }
  }
  __syncthreads();
  // vertical inverse update: even rows 2..66 on the tile's columns 2..65
  col_update<-1>(t, l.tx + HALO, l.ty, 33);
  __syncthreads();
  // vertical inverse predict: odd rows 3..65
  col_predict<1>(t, l.tx + HALO, l.ty, 32);
  __syncthreads();
  const int gcol = c0 + l.tx;
  if (gcol >= a.nx) return;
  for (int i = l.ty; i < TILE; i += 4) {
    const int gr = r0 + i;
    if (gr >= a.ny) break;
    a.out[(long long)gr * a.out_pitch + gcol] = t[HALO + i][HALO + l.tx];
  }
}

}  // namespace

hipError_t launch_dwt53_level(int inverse, const DwtLevel& a, hipStream_t stream) {
  if (a.nx < 2 || a.ny < 2 || (a.nx & 1) || (a.ny & 1)) return hipErrorInvalidValue;
  const int tiles = ((a.nx + TILE - 1) / TILE) * ((a.ny + TILE - 1) / TILE);
  const dim3 grid((tiles + XCDS - 1) / XCDS * XCDS);
  if (inverse)
    hipLaunchKernelGGL(dwt53_inverse_kernel, grid, dim3(THREADS), 0, stream, a);
  else
    hipLaunchKernelGGL(dwt53_forward_kernel, grid, dim3(THREADS), 0, stream, a);
  return hipGetLastError();
}

}  // namespace bpc
