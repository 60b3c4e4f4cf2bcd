// Host-only part of the codec library: argument validation, the reference's LUT
// text format, the host LUT trainer, per-codeblock codestream files, the batch
// queue's key packing and the server's admission rule.  Plain C++17 with no HIP
// dependency, so that it also builds with g++ under the address and
// undefined-behaviour sanitizers.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/bpc_paco.h"

namespace bpc {

// Device status bits (OR-ed by the kernels into a status word).
enum : uint32_t {
  ST_RANGE = 1u,      // |coefficient| >= 2^28, LUT entry outside [0,127] or LUT row outside the array
  ST_OVERFLOW = 2u,   // codeword slots past the codeblock region
  ST_MALFORMED = 4u,  // decoder: MSB word not in [0,27] and not 32
  ST_WIDE = 8u,       // a codeblock needs more bitplanes than the instantiated kernel holds
  ST_NARROW = 16u,    // decoder: a coefficient outside the int16 output range (saturated)
  // batched codec, host-visible word only: pairs wait for the 28-plane pass that
  // the run left out (the server launches it when it sees this bit)
  ST_WIDE_PENDING = 0x80000000u,
};

// Sets the calling thread's message (bpc_last_error) and returns code.
int fail(int code, const char* fmt, ...);
int check_params(const bpc_params* p);
int check_image(int size_x, int size_y, const bpc_params* p);
// int32 words per codeblock region
int stride_of(const bpc_params* p);
// BPC_* code of a status word (and its message)
int status_to_code(uint32_t st);

// ---- batch work queue entries ----------------------------------------------------
// One entry per codeblock: image << 20 | codeblock.  The empty partner of an odd
// codeblock count is all ones, a value no entry takes: images are < 4096 and
// codeblocks < 2^20 - 1 (an image may hold at most 2^20 - 1 codeblocks).
constexpr int QUEUE_CB_BITS = 20;
constexpr int MAX_BATCH_IMAGES = 4096;
constexpr long long MAX_IMAGE_CODEBLOCKS = (1ll << QUEUE_CB_BITS) - 1;
constexpr uint32_t QUEUE_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t queue_entry(uint32_t image, uint32_t cb) { return (image << QUEUE_CB_BITS) | cb; }
constexpr uint32_t queue_entry_image(uint32_t e) { return e >> QUEUE_CB_BITS; }
constexpr uint32_t queue_entry_cb(uint32_t e) { return e & ((1u << QUEUE_CB_BITS) - 1u); }
// Codeblocks of an image, or a negative status when the image cannot go in a batch
// (geometry, parameters, more than MAX_IMAGE_CODEBLOCKS codeblocks).
long long batch_image_codeblocks(int size_x, int size_y, const bpc_params* p);

// ---- server admission ------------------------------------------------------------
// A request is admitted when its output fits both the byte budget of outputs held
// by admitted requests (0: no budget) and the device's free memory, of which
// `reserve` bytes stay untouched.  Output buffers come from a pool of released
// buffers when one of the same size is there; only the rest needs free memory.
enum class Admit { YES, TRIM_POOL, WAIT };
struct AdmissionRule {
  size_t budget = 0;
  size_t reserve = 0;
  // new_bytes: bytes the request must allocate (0 when the pool has its buffers);
  // free_now: device free memory; pooled_other: bytes of pooled buffers of other
  // sizes, which the server may free to make room.
  Admit decide(size_t committed, size_t out_bytes, size_t new_bytes, size_t free_now, size_t pooled_other) const;
  // Whether a request of out_bytes can ever be admitted on a device of total_bytes.
  bool possible(size_t out_bytes, size_t total_bytes) const;
};

}  // namespace bpc
