// Internal interface between the C-ABI host code and the HIP kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "../../include/bpc_paco.h"
#include "bpc_host.h"

namespace bpc {

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
  int single;          // one codeblock per wave (lanes 32-63 idle) instead of a pair
  int in16;            // coefficients are int16 (encoder input / decoder output), else int32
  void* scratch;       // per wave of the grid: 2 * MAXP * 64 uint2 (scratch_bytes_per_wave); unused
                       // by the variant with the bitplanes in registers
  // decoder only, all null unless it reads the compact container instead of src
  // (selected by cmsb): MSB byte per codeblock, offsets (ncb + 1), payload of ccap
  // codewords (null when ccap is 0)
  const uint8_t* cmsb;
  const unsigned long long* coff;
  const uint16_t* cpay;
  unsigned long long ccap;
};

// One image of a batch, as the device sees it.
struct BatchImage {
  const int32_t* src;
  int32_t* dst;
  const int32_t* lut;
  int32_t* ncw;
  int32_t size_x, size_y, nx, ncb, stride, nbp, levels, lut_len, layout, fill, cpass3;
  int32_t in16;  // coefficients as int16 (encoder input / decoder output), else int32
};

// Control words of a batch queue (device memory): CTL_IMG + 32*n_images words.
enum : int {
  CTL_PAIRHIST = 0,     // 58 pair-bucket counts, then cursors
  CTL_HEAD = 64,        // next queue slot to hand out
  CTL_STATUS = 65,      // status bits of the batch
  CTL_WIDE_COUNT = 66,  // pairs deferred to the 28-plane pass (MSB beyond the main pass's planes)
  CTL_WIDE_HEAD = 67,   // next deferred pair to hand out
  CTL_DONE = 68,        // waves of the main pass that have finished
  CTL_DONE_WIDE = 69,   // waves of the 28-plane pass that have finished
  CTL_IMG = 128,        // per image: 29 MSB-bucket counts, then cursors (32 words each)
};

constexpr int NBUCKET = 29;  // MSB buckets: 27..0, then all-zero / rejected

// Per-run words of a batch, at the head of the staging area: the kernels of a
// run read the sizes from here rather than from their arguments, so that one
// captured launch sequence serves batches of any shape.
struct BatchRun {
  uint32_t* status;       // the caller's status word (device): OR of the run's ST_* bits
  uint32_t* host_status;  // nullable, host-visible: the run's ST_* bits, stored by the fold
  int64_t total_cbs;      // codeblocks over all images
  int64_t total_pairs;    // queue entries (codeblock pairs, each image's count rounded up to even)
  int32_t n_images;
  int32_t rows;           // codeblock rows, shared by every image
  uint32_t stage_bytes;   // bytes of the staging area this run uses
  uint32_t pad0;
  uint64_t pad[2];
};
static_assert(sizeof(BatchRun) == 64, "run words fill one 64-byte line");

struct QueueArgs {
  const BatchRun* run;           // sizes of the run (device copy)
  const BatchImage* images;
  const long long* cb_offset;    // first global codeblock of each image (n_images + 1 entries)
  const long long* sorted_base;  // first entry of each image in `sorted` (even-padded ranges)
  int decode;
  int max_images;                // capacity of the batch object (grid of the per-image scan)
  int cus;                       // compute units (grids of the grid-stride kernels)
  uint32_t* keys;        // per global codeblock: MSB bucket
  uint32_t* sorted;      // 2*total_pairs entries: image << 20 | codeblock, sorted by MSB inside each image
  uint32_t* pair_keys;   // per pair: pass mode * 29 + bucket
  uint32_t* pair_queue;  // pair indices, costliest first
  uint32_t* ctl;
};

struct BatchArgs {
  void* scratch;  // per resident wave: scratch_bytes_per_wave(maxp)
  const BatchImage* images;
  const uint32_t* sorted;
  const uint32_t* pair_queue;
  const uint32_t* pair_keys;  // per pair: pass mode * 29 + MSB bucket of its costlier codeblock
  uint32_t* wide;             // pairs deferred to the 28-plane pass
  uint32_t* ctl;
  const BatchRun* run;  // total_pairs and rows of the run
  int wide_pass;  // 0: drain the pair queue (deferring pairs beyond the kernel's planes); 1: drain `wide`
  int fold;         // this launch ends the run: its last wave folds the run's status
  int report_wide;  // the 28-plane pass is left out: the fold reports deferred pairs
};

// Batched run (bpc_batch_run) with two more outputs for the server: host_status
// (nullable, pinned host memory) receives the run's ST_* bits from the last kernel
// of the run, and defer_wide != 0 leaves the 28-plane pass out of the run,
// reporting ST_WIDE_PENDING in host_status when pairs wait for it; batch_run_wide
// then runs it (on the batch's last run, before the batch object runs again).
int batch_run_ex(bpc_batch* b, int direction, const bpc_image* images, int n_images, int use_graph,
                 uint32_t* d_status, uint32_t* host_status, int defer_wide, void* stream);
int batch_run_wide(bpc_batch* b, void* stream);

// Symbol statistics of a training image (LUT trainer).
struct TrainArgs {
  const int32_t* img;  // coefficients (raster or codeblock-contiguous, as layout says)
  int size_x, size_y, nx, ncb, rows, nbp, levels, lut_len, layout;
  void* scratch;                  // per wave of the grid: scratch_bytes_per_wave(TRAIN_MAXP)
  unsigned long long* partial;    // per workgroup: lut_len packed counts (symbols << 32 | zeros)
};
constexpr int TRAIN_MAXP = 28;   // every bitplane a coefficient below 2^28 can have
constexpr int TRAIN_WAVES = 8;   // waves per trainer workgroup (one shared LDS histogram)

// Bitplane scratch of the entry point that allocates nothing (bpc_codec_device_noalloc):
// a fixed array per device; its launches use at most this many bytes of slots.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
// Device address of that array on the current device (asked once per device).
hipError_t noalloc_scratch(void** ptr);

// Kernel variants hold this many bitplanes in registers.
int planes_for(int nbp);
// Bitplane scratch of one wave: 2 row halves x MAXP planes x 64 lanes x 8 bytes.
constexpr size_t scratch_bytes_per_wave(int maxp) { return (size_t)2 * maxp * 64 * 8; }
// Waves of the per-image kernel that fit on one CU (occupancy query); reg_planes
// selects the variant that keeps the bitplanes in registers (no scratch), compact
// the decoder that reads the compact container.
int codec_resident_waves_per_cu(int decode, int cpass3, int maxp, int reg_planes, int compact);
hipError_t launch_codec(int decode, int cpass3, int maxp, int reg_planes, const KArgs& a, int grid_waves,
                        hipStream_t stream);
hipError_t launch_queue_build(const QueueArgs& a, hipStream_t stream);
// vgpr_cap_waves: 0 (compiler's register budget) or 5..8 (registers capped for that many waves per SIMD)
hipError_t launch_batch(int decode, int maxp, const BatchArgs& a, int grid_blocks, int waves_per_block,
                        int vgpr_cap_waves, hipStream_t stream);
// Copies the staging area of a run (stage_bytes, from its run words) from pinned
// host memory to the device and clears the queue's control words.
hipError_t launch_batch_stage(const char* h_stage, char* d_stage, uint32_t* ctl, hipStream_t stream);

// Workgroups of the batched kernel that fit on one CU (occupancy query).
int batch_resident_blocks_per_cu(int decode, int maxp, int waves_per_block, int vgpr_cap_waves);
// Trainer workgroups that fit on one CU with lds_bytes of LDS each.
int train_resident_blocks_per_cu(int cpass3, int lds_bytes);
// Trainer over grid_blocks workgroups, then the LUT (and, if non-null, the summed
// counts) from the partial rows.
hipError_t launch_train(int cpass3, const TrainArgs& a, int grid_blocks, int32_t* lut, uint32_t* n0, uint32_t* n,
                        hipStream_t stream);
// Compact container (bpc_container.hip): index (MSB bytes, codeword counts, scan
// into offsets[0..ncb]), pack (regions -> payload), unpack (payload -> regions).
hipError_t launch_compact_index(const int32_t* stream, long long ncb, int stride, const int32_t* ncw, uint8_t* msb,
                                unsigned long long* offsets, uint32_t* status, hipStream_t s);
hipError_t launch_compact_pack(const int32_t* stream, long long ncb, int stride, const unsigned long long* offsets,
                               uint16_t* payload, unsigned long long cap, uint32_t* status, hipStream_t s);
hipError_t launch_compact_unpack(const uint8_t* msb, const unsigned long long* offsets, const uint16_t* payload,
                                 unsigned long long cap, long long ncb, int stride, int32_t* stream, uint32_t* status,
                                 hipStream_t s);
// One level of the 5/3 transform (bpc_dwt.hip) on an nx x ny band (both even).
// Forward: src -> ll (LL band) and det (HL / LH / HH around the level's origin).
// Inverse: ll and det -> out.
struct DwtLevel {
  int nx, ny;
  const int32_t* src;
  long long src_pitch;
  int32_t* ll;
  long long ll_pitch;
  int32_t* det;
  long long det_pitch;
  int32_t* out;
  long long out_pitch;
};
hipError_t launch_dwt53_level(int inverse, const DwtLevel& a, hipStream_t stream);
// 16-byte aligned device-to-device copy of a multiple of 16 bytes.
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

}  // namespace bpc
