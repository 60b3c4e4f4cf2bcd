// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime_api.h>

namespace bpc {

// Device status bits (OR-ed by the kernels into a status word).
enum : uint32_t {
  ST_RANGE = 1u,      // |coefficient| >= 2^28, LUT entry outside [0,127] or LUT row outside the array
  ST_OVERFLOW = 2u,   // codeword slots past the codeblock region
  ST_MALFORMED = 4u,  // decoder: MSB word not in [0,27] and not 32
  ST_WIDE = 8u,       // a codeblock needs more bitplanes than the instantiated kernel holds
};

// One launch over the codeblocks of one image.
struct KArgs {
  const int32_t* src;  // encoder: coefficients ; decoder: codestream regions
  int32_t* dst;        // encoder: codestream regions ; decoder: coefficients
  const int32_t* lut;  // reference-format probability LUT (int per entry)
  int32_t* ncw;        // encoder: codeword slots used per codeblock (nullable)
  uint32_t* status;    // OR of ST_* bits
  int size_x, size_y;  // image size in samples
  int nx;              // codeblocks per image row (size_x / 64)
  int ncb;             // number of codeblocks
  int rows;            // codeblock rows (CBLOCK_LENGTH)
  int stride;          // int32 words per codeblock region
  int nbp;             // LUT bitplanes (LUT_N_BITPLANES)
  int levels;          // DWT levels
  int lut_len;         // ints in lut
  int layout;          // 0: raster (reference), 1: codeblock-contiguous (tile-major)
  int fill;            // encoder: write -1 to unused region words
};

// Kernel variants hold this many bitplanes in registers.
int planes_for(int nbp);
hipError_t launch_codec(int decode, int cpass3, int maxp, const KArgs& a, hipStream_t stream);
// 16-byte aligned device-to-device copy of a multiple of 16 bytes.
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

}  // namespace bpc
