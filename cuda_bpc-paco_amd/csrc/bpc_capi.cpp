// C-ABI of the BPC-PaCo codec: drop-in run entry (CUDA_BPC, BPC_PaCo.cu:1305),
// device-resident encode/decode, LUT loading / upload / training, HIP helpers.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <atomic>
#include <mutex>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/bpc_paco.h"
#include "bpc_internal.h"

static_assert(bpc::NOALLOC_SCRATCH_BYTES == BPC_NOALLOC_SCRATCH_BYTES, "fixed scratch size in the header");

using bpc::check_image;
using bpc::check_params;
using bpc::fail;
using bpc::status_to_code;
using bpc::stride_of;

namespace {

int hip_fail(hipError_t e, const char* what) {
  return fail(BPC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
  } while (0)

// one library-owned status word per device
std::mutex g_status_mu;
std::vector<uint32_t*> g_status_words;

int default_status(uint32_t** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_status_mu);
  if ((int)g_status_words.size() <= dev) g_status_words.resize(dev + 1, nullptr);
  if (!g_status_words[dev]) {
    uint32_t* w = nullptr;
    HIP_TRY(hipMalloc(&w, sizeof(uint32_t)));
    HIP_TRY(hipMemset(w, 0, sizeof(uint32_t)));
    g_status_words[dev] = w;
  }
  *out = g_status_words[dev];
  return BPC_OK;
}

int device_cus(int dev) {
  static std::mutex mu;
  static std::vector<int> cache;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (!cache[dev]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    cache[dev] = cus;
  }
  return cache[dev];
}

// Resident waves per CU of a per-image kernel variant (an occupancy query, asked
// once per device and variant).
int resident_waves_cached(int dev, int decode, int cpass3, int maxp, int reg_planes, int compact) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int, int>, int> cache;
  const auto key = std::make_tuple(dev, decode, cpass3, maxp, reg_planes, compact);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = bpc::codec_resident_waves_per_cu(decode, cpass3, maxp, reg_planes, compact);
  if (per_cu <= 0) per_cu = 1;
  cache[key] = per_cu;
  return per_cu;
}

// Bitplane scratch of the per-image kernel, one per device.  Launches that use it
// are ordered through `last`: a launch on another stream first waits for the
// previous user.
struct DevScratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t last = nullptr;
  bool has_last = false;
  hipStream_t last_stream = nullptr;  // stream of the previous user (same stream: already ordered)
};
std::mutex g_scratch_mu;
std::vector<DevScratch> g_scratch;

// At least `need` bytes of the device's scratch for work about to be queued on
// `stream`; the caller holds g_scratch_mu and ends with scratch_release.
int scratch_acquire(int dev, size_t need, hipStream_t stream, void** ptr) {
  if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
  DevScratch& sc = g_scratch[dev];
  if (!sc.last) HIP_TRY(hipEventCreateWithFlags(&sc.last, hipEventDisableTiming));
  if (need > sc.bytes) {
    if (sc.has_last) HIP_TRY(hipEventSynchronize(sc.last));
    if (sc.ptr) HIP_TRY(hipFree(sc.ptr));
    sc.ptr = nullptr;
    sc.bytes = 0;
    HIP_TRY(hipMalloc(&sc.ptr, need));
    sc.bytes = need;
  } else if (sc.has_last && stream != sc.last_stream) {
    HIP_TRY(hipStreamWaitEvent(stream, sc.last, 0));
  }
  *ptr = sc.ptr;
  return BPC_OK;
}

int scratch_release(int dev, hipStream_t stream) {
  DevScratch& sc = g_scratch[dev];
  HIP_TRY(hipEventRecord(sc.last, stream));
  sc.has_last = true;
  sc.last_stream = stream;
  return BPC_OK;
}

// The decoder's codestream in the compact container (MSB bytes, offsets, payload).
struct CompactIn {
  const uint8_t* msb = nullptr;
  const uint64_t* offsets = nullptr;
  const uint16_t* payload = nullptr;
  uint64_t payload_len = 0;
};

// fixed_scratch (nullable): bitplane slots of fixed_bytes to use instead of the
// library's scratch; the grid is then limited to the waves whose slots fit.
// cmp (nullable, decoder): read the compact container instead of src.
int launch(int decode, const int32_t* src, int32_t* dst, int size_x, int size_y, const int32_t* d_lut, int lut_len,
           const bpc_params* p, int fill, int32_t* d_ncw, uint32_t* d_status, hipStream_t stream, int maxp,
           int in16 = 0, void* fixed_scratch = nullptr, size_t fixed_bytes = 0, const CompactIn* cmp = nullptr) {
  bpc::KArgs a{};
  a.src = src;
  a.dst = dst;
  a.lut = d_lut;
  a.ncw = d_ncw;
  a.status = d_status;
  a.size_x = size_x;
  a.size_y = size_y;
  a.nx = size_x / 64;
  a.ncb = (size_x / 64) * (size_y / p->cblock_length);
  a.rows = p->cblock_length;
  a.stride = stride_of(p);
  a.nbp = p->lut_n_bitplanes;
  a.levels = p->dwt_levels;
  a.lut_len = lut_len;
  a.layout = p->layout;
  a.fill = fill;
  a.in16 = in16;
  const int compact = decode && cmp ? 1 : 0;
  if (compact) {
    a.cmsb = cmp->msb;  // non-null: selects the decoder that reads the compact form
    a.coff = reinterpret_cast<const unsigned long long*>(cmp->offsets);
    a.cpay = cmp->payload;  // null only with payload_len 0, when no count can pass the bounds test
    a.ccap = cmp->payload_len;
  }
  if (a.ncb <= 0) return BPC_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  const long long cus = device_cus(dev);
  // An image with no more codeblocks than the device has SIMDs cannot fill it with
  // pairs: one codeblock per wave then runs each codeblock's chain on its own SIMD
  // instead of two in lock step (the launch time is the longest chain).
  a.single = (long long)a.ncb <= 4LL * cus ? 1 : 0;
  const long long units = a.single ? a.ncb : (a.ncb + 1) / 2;
  // When every wave the image needs is resident at once with the bitplanes held in
  // registers, that variant runs: no scratch slot, no plane traffic to memory.
  if (maxp <= 20 && units <= (long long)resident_waves_cached(dev, decode, p->cpass3, maxp, 1, compact) * cus) {
    a.scratch = nullptr;
    hipError_t e = bpc::launch_codec(decode, p->cpass3, maxp, 1, a, (int)units, stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return BPC_OK;
  }
  int per_cu = resident_waves_cached(dev, decode, p->cpass3, maxp, 0, compact);
  int grid = (int)std::min<long long>(units, (long long)per_cu * cus);
  if (fixed_scratch) {
    const size_t slots = fixed_bytes / bpc::scratch_bytes_per_wave(maxp);
    if (slots < 1) return fail(BPC_ERR_ARG, "fixed scratch of %zu bytes holds no %d-plane slot", fixed_bytes, maxp);
    grid = (int)std::min<size_t>((size_t)grid, slots);  // the waves grid-stride over the rest
    a.scratch = fixed_scratch;
    hipError_t e = bpc::launch_codec(decode, p->cpass3, maxp, 0, a, grid, stream);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return BPC_OK;
  }
  const size_t need = (size_t)grid * bpc::scratch_bytes_per_wave(maxp);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  int st = scratch_acquire(dev, need, stream, &a.scratch);
  if (st) return st;
  hipError_t e = bpc::launch_codec(decode, p->cpass3, maxp, 0, a, grid, stream);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return scratch_release(dev, stream);
}

int check_device_call(const void* a, const void* b, int size_x, int size_y, const int32_t* d_lut, int lut_len,
                      const bpc_params* p) {
  int st = check_params(p);
  if (st) return st;
  st = check_image(size_x, size_y, p);
  if (st) return st;
  if (!a || !b || !d_lut) return fail(BPC_ERR_ARG, "null buffer");
  if (((uintptr_t)a & 7) || ((uintptr_t)b & 7)) return fail(BPC_ERR_ARG, "buffers must be 8-byte aligned");
  if (lut_len < bpc_lut_len(p)) return fail(BPC_ERR_ARG, "lut_len %d < %d required", lut_len, bpc_lut_len(p));
  return BPC_OK;
}

}  // namespace

extern "C" {

int bpc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int bpc_encode_device(const int32_t* d_coeffs, int32_t* d_stream, int size_x, int size_y, const int32_t* d_lut,
                      int lut_len, const bpc_params* p, int fill_unused, int32_t* d_ncw, uint32_t* d_status,
                      void* stream) {
  int st = check_device_call(d_coeffs, d_stream, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_status) {
    st = default_status(&d_status);
    if (st) return st;
  }
  return launch(0, d_coeffs, d_stream, size_x, size_y, d_lut, lut_len, p, fill_unused, d_ncw, d_status,
                (hipStream_t)stream, bpc::planes_for(p->lut_n_bitplanes));
}

int bpc_decode_device(const int32_t* d_stream, int32_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                      int lut_len, const bpc_params* p, uint32_t* d_status, void* stream) {
  int st = check_device_call(d_stream, d_coeffs, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_status) {
    st = default_status(&d_status);
    if (st) return st;
  }
  return launch(1, d_stream, d_coeffs, size_x, size_y, d_lut, lut_len, p, 0, nullptr, d_status,
                (hipStream_t)stream, bpc::planes_for(p->lut_n_bitplanes));
}

int bpc_encode_device_i16(const int16_t* d_coeffs, int32_t* d_stream, int size_x, int size_y, const int32_t* d_lut,
                          int lut_len, const bpc_params* p, int fill_unused, int32_t* d_ncw, uint32_t* d_status,
                          void* stream) {
  int st = check_device_call(d_coeffs, d_stream, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_status) {
    st = default_status(&d_status);
    if (st) return st;
  }
  return launch(0, reinterpret_cast<const int32_t*>(d_coeffs), d_stream, size_x, size_y, d_lut, lut_len, p,
                fill_unused, d_ncw, d_status, (hipStream_t)stream, bpc::planes_for(p->lut_n_bitplanes), 1);
}

int bpc_decode_device_i16(const int32_t* d_stream, int16_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                          int lut_len, const bpc_params* p, uint32_t* d_status, void* stream) {
  int st = check_device_call(d_stream, d_coeffs, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_status) {
    st = default_status(&d_status);
    if (st) return st;
  }
  return launch(1, d_stream, reinterpret_cast<int32_t*>(d_coeffs), size_x, size_y, d_lut, lut_len, p, 0, nullptr,
                d_status, (hipStream_t)stream, bpc::planes_for(p->lut_n_bitplanes), 1);
}

int bpc_decode_compact_device(const uint8_t* d_msb, const uint64_t* d_offsets, const uint16_t* d_payload,
                              uint64_t payload_len, int32_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                              int lut_len, const bpc_params* p, uint32_t* d_status, void* stream) {
  int st = check_device_call(d_offsets, d_coeffs, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_msb) return fail(BPC_ERR_ARG, "null MSB array");
  if (!d_payload && payload_len) return fail(BPC_ERR_ARG, "null payload of %llu codewords",
                                             (unsigned long long)payload_len);
  if ((uintptr_t)d_payload & 1) return fail(BPC_ERR_ARG, "payload must be 2-byte aligned");
  if (!d_status) {
    st = default_status(&d_status);
    if (st) return st;
  }
  CompactIn cmp;
  cmp.msb = d_msb;
  cmp.offsets = d_offsets;
  cmp.payload = d_payload;
  cmp.payload_len = payload_len;
  return launch(1, nullptr, d_coeffs, size_x, size_y, d_lut, lut_len, p, 0, nullptr, d_status, (hipStream_t)stream,
                bpc::planes_for(p->lut_n_bitplanes), 0, nullptr, 0, &cmp);
}

int bpc_codec_device_noalloc(int direction, const int32_t* d_src, int32_t* d_dst, int size_x, int size_y,
                             const int32_t* d_lut, int lut_len, const bpc_params* p, int fill_unused,
                             int32_t* d_ncw, uint32_t* d_status, void* stream) {
  if (direction != BPC_CODE && direction != BPC_DECODE) return fail(BPC_ERR_ARG, "direction must be CODE or DECODE");
  int st = check_device_call(d_src, d_dst, size_x, size_y, d_lut, lut_len, p);
  if (st) return st;
  if (!d_status) return fail(BPC_ERR_ARG, "d_status is required (this call allocates no status word)");
  const int maxp = bpc::planes_for(p->lut_n_bitplanes);
  if (bpc::scratch_bytes_per_wave(maxp) > bpc::NOALLOC_SCRATCH_BYTES)
    return fail(BPC_ERR_ARG, "a %d-plane slot of %zu bytes exceeds the fixed scratch of %zu bytes", maxp,
                bpc::scratch_bytes_per_wave(maxp), bpc::NOALLOC_SCRATCH_BYTES);
  void* scr = nullptr;
  hipError_t e = bpc::noalloc_scratch(&scr);
  if (e != hipSuccess) return hip_fail(e, "fixed scratch address");
  return launch(direction == BPC_DECODE, d_src, d_dst, size_x, size_y, d_lut, lut_len, p,
                direction == BPC_CODE ? fill_unused : 0, direction == BPC_CODE ? d_ncw : nullptr, d_status,
                (hipStream_t)stream, maxp, 0, scr, bpc::NOALLOC_SCRATCH_BYTES);
}

int bpc_status_take(uint32_t* d_status, void* stream) {
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  uint32_t h = 0;
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  HIP_TRY(hipMemcpy(&h, d_status, sizeof h, hipMemcpyDeviceToHost));
  if (h) HIP_TRY(hipMemset(d_status, 0, sizeof h));
  return status_to_code(h);
}

int bpc_run_ex(int direction, int32_t* hdata, int size_x, int size_y, const int32_t* d_lut, int lut_len,
               const bpc_params* p) {
  if (direction != BPC_CODE && direction != BPC_DECODE) return fail(BPC_ERR_ARG, "direction must be CODE or DECODE");
  int st = check_params(p);
  if (st) return st;
  st = check_image(size_x, size_y, p);
  if (st) return st;
  if (!hdata || !d_lut) return fail(BPC_ERR_ARG, "null buffer");
  if (lut_len < bpc_lut_len(p)) return fail(BPC_ERR_ARG, "lut_len %d < %d required", lut_len, bpc_lut_len(p));
  if (bpc_device_count() <= 0) return fail(BPC_ERR_NODEVICE, "no HIP device");
  const size_t img_bytes = (size_t)size_x * size_y * sizeof(int32_t);
  const size_t cs_bytes = (size_t)bpc_stream_words(size_x, size_y, p) * sizeof(int32_t);
  const size_t in_bytes = direction == BPC_CODE ? img_bytes : cs_bytes;
  const size_t out_bytes = direction == BPC_CODE ? cs_bytes : img_bytes;
  void *d_in = nullptr, *d_out = nullptr;
  uint32_t* d_status = nullptr;
  int rc = BPC_OK;
  hipError_t e = hipMalloc(&d_in, in_bytes);
  if (e == hipSuccess) e = hipMalloc(&d_out, out_bytes);
  if (e == hipSuccess) e = hipMalloc(&d_status, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(d_status, 0, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(d_in, hdata, in_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = hip_fail(e, "device buffers");
  int maxp = bpc::planes_for(p->lut_n_bitplanes);
  for (int attempt = 0; rc == BPC_OK && attempt < 2; ++attempt) {
    rc = launch(direction == BPC_DECODE, (const int32_t*)d_in, (int32_t*)d_out, size_x, size_y, d_lut, lut_len, p, 1,
                nullptr, d_status, nullptr, maxp);
    if (rc) break;
    uint32_t h = 0;
    e = hipMemcpy(&h, d_status, sizeof h, hipMemcpyDeviceToHost);  // synchronizes the launch
    if (e != hipSuccess) {
      rc = hip_fail(e, "kernel");
      break;
    }
    if ((h & bpc::ST_WIDE) && maxp < 28) {  // rerun with every bitplane register
      maxp = 28;
      e = hipMemset(d_status, 0, sizeof(uint32_t));
      if (e != hipSuccess) rc = hip_fail(e, "hipMemset");
      continue;
    }
    rc = status_to_code(h);
    break;
  }
  if (rc == BPC_OK) {
    e = hipMemcpy(hdata, d_out, out_bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "copy back");
  }
  if (d_in) (void)hipFree(d_in);
  if (d_out) (void)hipFree(d_out);
  if (d_status) (void)hipFree(d_status);
  return rc;
}

int bpc_run(int direction, int32_t* hdata, int size_x, int size_y, const int32_t* d_lut) {
  bpc_params p;
  bpc_get_config(&p);
  return bpc_run_ex(direction, hdata, size_x, size_y, d_lut, bpc_lut_len(&p), &p);
}

int bpc_lut_upload(const int32_t* host_lut, int lut_len, int32_t** d_lut) {
  if (!host_lut || lut_len <= 0 || !d_lut) return fail(BPC_ERR_ARG, "bad arguments");
  for (int k = 0; k < lut_len; ++k)
    if (host_lut[k] < 0 || host_lut[k] > 127) return fail(BPC_ERR_RANGE, "LUT entry %d = %d outside [0,127]", k, host_lut[k]);
  int32_t* d = nullptr;
  HIP_TRY(hipMalloc(&d, (size_t)lut_len * sizeof(int32_t)));
  hipError_t e = hipMemcpy(d, host_lut, (size_t)lut_len * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "LUT upload");
  }
  *d_lut = d;
  return BPC_OK;
}

int bpc_lut_load_text(const char* folder, int cpass3, int dwt_levels, int32_t** d_lut, int* lut_len, int* nbp) {
  std::vector<int32_t> h(1 << 16);
  int len = 0, n = 0;
  int st = bpc_lut_parse_text(folder, cpass3, dwt_levels, h.data(), (int)h.size(), &len, &n);
  if (st) return st;
  st = bpc_lut_upload(h.data(), len, d_lut);
  if (st) return st;
  if (lut_len) *lut_len = len;
  if (nbp) *nbp = n;
  return BPC_OK;
}

int bpc_lut_train_device(const int32_t* d_img, int size_x, int size_y, const bpc_params* p, int32_t* d_lut,
                         uint32_t* d_n0, uint32_t* d_n, void* stream) {
  int st = check_params(p);
  if (st) return st;
  st = check_image(size_x, size_y, p);
  if (st) return st;
  if (!d_img) return fail(BPC_ERR_ARG, "null image");
  if (!d_lut && !d_n0 && !d_n) return fail(BPC_ERR_ARG, "no output buffer");
  if (bpc_device_count() <= 0) return fail(BPC_ERR_NODEVICE, "no HIP device");
  const int len = bpc_lut_len(p);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  int max_lds = 0;
  HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t lds = (size_t)len * sizeof(unsigned long long);
  if (lds > (size_t)max_lds)
    return fail(BPC_ERR_ARG, "LUT of %d entries exceeds the device trainer's LDS histogram (%d entries); use bpc_lut_train",
                len, max_lds / (int)sizeof(unsigned long long));
  bpc::TrainArgs a;
  a.img = d_img;
  a.size_x = size_x;
  a.size_y = size_y;
  a.nx = size_x / 64;
  a.ncb = (size_x / 64) * (size_y / p->cblock_length);
  a.rows = p->cblock_length;
  a.nbp = p->lut_n_bitplanes;
  a.levels = p->dwt_levels;
  a.lut_len = len;
  a.layout = p->layout;
  int per_cu = bpc::train_resident_blocks_per_cu(p->cpass3, (int)lds);
  if (per_cu <= 0) per_cu = 1;
  const long long pairs = (a.ncb + 1) / 2;
  const int grid = (int)std::max<long long>(
      1, std::min<long long>((pairs + bpc::TRAIN_WAVES - 1) / bpc::TRAIN_WAVES, (long long)per_cu * device_cus(dev)));
  const size_t plane_bytes = (size_t)grid * bpc::TRAIN_WAVES * bpc::scratch_bytes_per_wave(bpc::TRAIN_MAXP);
  const size_t need = plane_bytes + (size_t)grid * len * sizeof(unsigned long long);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  void* base = nullptr;
  st = scratch_acquire(dev, need, (hipStream_t)stream, &base);
  if (st) return st;
  a.scratch = base;
  a.partial = reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + plane_bytes);
  hipError_t e = bpc::launch_train(p->cpass3, a, grid, d_lut, d_n0, d_n, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "trainer launch");
  return scratch_release(dev, (hipStream_t)stream);
}

int bpc_lut_free(int32_t* d_lut) {
  if (d_lut) HIP_TRY(hipFree(d_lut));
  return BPC_OK;
}

int bpc_dev_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(BPC_ERR_ARG, "null pointer");
  HIP_TRY(hipMalloc(ptr, bytes));
  return BPC_OK;
}
int bpc_dev_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return BPC_OK;
}
int bpc_dev_memset(void* ptr, int value, size_t bytes, void* stream) {
  HIP_TRY(hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream));
  return BPC_OK;
}
int bpc_copy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return BPC_OK;
}
int bpc_copy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return BPC_OK;
}
int bpc_copy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  hipError_t e = bpc::launch_copy(dst, src, bytes, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "device copy (16-byte aligned, multiple of 16 bytes)");
  return BPC_OK;
}
int bpc_stream_create(void** stream) {
  hipStream_t s;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return BPC_OK;
}
int bpc_stream_destroy(void* stream) {
  if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return BPC_OK;
}
int bpc_stream_sync(void* stream) {
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return BPC_OK;
}
int bpc_device_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return BPC_OK;
}
int bpc_event_create(void** ev) {
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  *ev = e;
  return BPC_OK;
}
int bpc_event_destroy(void* ev) {
  if (ev) HIP_TRY(hipEventDestroy((hipEvent_t)ev));
  return BPC_OK;
}
int bpc_event_record(void* ev, void* stream) {
  HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return BPC_OK;
}
int bpc_event_elapsed_ms(void* start, void* stop, float* ms) {
  HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
  HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return BPC_OK;
}
int bpc_mem_info(size_t* free_bytes, size_t* total_bytes) {
  HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return BPC_OK;
}

}  // extern "C"

// ---- batched codec ---------------------------------------------------------

// Staging of one run, uploaded in one copy: the run words, the offset tables and
// the image descriptors, each at a fixed place for the batch's capacity.
struct bpc_batch {
  int device = 0;
  int max_images = 0;
  long long max_cbs = 0;
  size_t stage_bytes = 0;          // run words + offsets (2*max_images+1) + descriptors (max_images)
  char* d_stage = nullptr;         // device copy of the staging area
  char* h_stage = nullptr;         // pinned staging area (host address)
  const char* h_stage_dev = nullptr;  // its device address (read by the stage kernel)
  uint32_t* d_keys = nullptr;
  uint32_t* d_sorted = nullptr;
  uint32_t* d_pair_keys = nullptr;
  uint32_t* d_pair_queue = nullptr;
  uint32_t* d_wide = nullptr;      // pairs deferred to the 28-plane pass
  uint32_t* d_ctl = nullptr;
  hipEvent_t staged = nullptr;  // the stage kernel of the last run has read the pinned area
  bool staged_pending = false;
  hipEvent_t done = nullptr;  // the last run has finished
  bool done_pending = false;
  void* d_scratch = nullptr;  // bitplane scratch, one slot per resident wave
  size_t scratch_bytes = 0;
  int waves_per_block = 4;
  int waves_per_simd = 0;  // 0: as many as fit
  int vgpr_cap_waves = 0;  // 0: compiler's register budget; 5..8: capped for that many waves per SIMD
  int cus = 0;
  int last_direction = 0, last_maxp = 0;  // of the last run (the deferred 28-plane pass codes its pairs)
  // captured launch sequences, least recently used first; at most kMaxGraphs
  static constexpr size_t kMaxGraphs = 32;
  std::list<std::pair<std::string, hipGraphExec_t>> graphs;
  std::mutex mu;

  bpc::BatchRun* h_run() { return reinterpret_cast<bpc::BatchRun*>(h_stage); }
  long long* h_offsets() { return reinterpret_cast<long long*>(h_stage + sizeof(bpc::BatchRun)); }
  bpc::BatchImage* h_images() {
    return reinterpret_cast<bpc::BatchImage*>(h_stage + images_offset());
  }
  const bpc::BatchRun* d_run() const { return reinterpret_cast<const bpc::BatchRun*>(d_stage); }
  const long long* d_offsets() const { return reinterpret_cast<const long long*>(d_stage + sizeof(bpc::BatchRun)); }
  const bpc::BatchImage* d_images() const {
    return reinterpret_cast<const bpc::BatchImage*>(d_stage + images_offset());
  }
  size_t images_offset() const {
    const size_t o = sizeof(bpc::BatchRun) + sizeof(long long) * (2 * (size_t)max_images + 1);
    return (o + 63) & ~(size_t)63;
  }
};

namespace {

void drop_graphs(bpc_batch* b) {
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

void batch_release(bpc_batch* b) {
  drop_graphs(b);
  void* dev[] = {b->d_stage, b->d_keys, b->d_sorted, b->d_pair_keys, b->d_pair_queue, b->d_wide, b->d_ctl,
                 b->d_scratch};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  if (b->staged) (void)hipEventDestroy(b->staged);
  if (b->done) (void)hipEventDestroy(b->done);
}

struct BatchPlan {
  long long total_cbs = 0, total_pairs = 0;
  int maxp = 16;
  int rows = 0;
};

// Fills the pinned staging area: run words, offset tables, image descriptors.
int batch_stage(bpc_batch* b, int direction, const bpc_image* images, int n, BatchPlan& plan) {
  if (n <= 0 || n > b->max_images) return fail(BPC_ERR_ARG, "batch of %d images (capacity %d)", n, b->max_images);
  plan.rows = images[0].params.cblock_length;
  int max_nbp = 1;
  long long pos = 0;
  bpc::BatchImage* hi = b->h_images();
  long long* ho = b->h_offsets();
  for (int k = 0; k < n; ++k) {
    const bpc_image& im = images[k];
    const long long ncb = bpc::batch_image_codeblocks(im.size_x, im.size_y, &im.params);
    if (ncb < 0) return (int)ncb;
    if (im.params.cblock_length != plan.rows) return fail(BPC_ERR_ARG, "all images of a batch share cblock_length");
    if (!im.src || !im.dst || !im.lut) return fail(BPC_ERR_ARG, "image %d: null buffer", k);
    if (((uintptr_t)im.src & 7) || ((uintptr_t)im.dst & 7)) return fail(BPC_ERR_ARG, "image %d: buffers must be 8-byte aligned", k);
    if (im.lut_len < bpc_lut_len(&im.params)) return fail(BPC_ERR_ARG, "image %d: lut_len too small", k);
    bpc::BatchImage& d = hi[k];
    d.src = im.src;
    d.dst = im.dst;
    d.lut = im.lut;
    d.ncw = direction == BPC_CODE ? im.ncw : nullptr;
    d.size_x = im.size_x;
    d.size_y = im.size_y;
    d.nx = im.size_x / 64;
    d.ncb = (int32_t)ncb;
    d.stride = stride_of(&im.params);
    d.nbp = im.params.lut_n_bitplanes;
    d.levels = im.params.dwt_levels;
    d.lut_len = im.lut_len;
    d.layout = im.params.layout;
    d.fill = im.fill_unused;
    d.cpass3 = im.params.cpass3;
    d.in16 = im.int16_coeffs ? 1 : 0;
    ho[k] = plan.total_cbs;
    ho[b->max_images + 1 + k] = pos;  // sorted_base: a fixed place for the batch object
    plan.total_cbs += ncb;
    pos += (ncb + 1) & ~1ll;
    max_nbp = std::max(max_nbp, im.params.lut_n_bitplanes);
  }
  ho[n] = plan.total_cbs;
  plan.total_pairs = pos / 2;
  if (plan.total_cbs > b->max_cbs) return fail(BPC_ERR_ARG, "batch of %lld codeblocks (capacity %lld)", plan.total_cbs, b->max_cbs);
  plan.maxp = bpc::planes_for(max_nbp);
  bpc::BatchRun* r = b->h_run();
  r->total_cbs = plan.total_cbs;
  r->total_pairs = plan.total_pairs;
  r->n_images = n;
  r->rows = plan.rows;
  r->stage_bytes = (uint32_t)(b->images_offset() + sizeof(bpc::BatchImage) * (size_t)n);
  return BPC_OK;
}

// Workgroups of the persistent kernel that fit on the device at once.
int batch_resident_blocks(bpc_batch* b, int direction, int maxp) {
  const int wpb = b->waves_per_block;
  int per_cu = bpc::batch_resident_blocks_per_cu(direction == BPC_DECODE, maxp, wpb, b->vgpr_cap_waves);
  if (per_cu <= 0) per_cu = 1;
  if (b->waves_per_simd > 0) per_cu = std::max(1, std::min(per_cu, (4 * b->waves_per_simd) / wpb));
  return per_cu * b->cus;
}

// Workgroups of the persistent kernel: the resident capacity; an eager run takes at
// most one wave per pair (a captured sequence keeps the full grid, whose extra
// waves find the queue drained and leave).
int batch_grid_blocks(bpc_batch* b, int direction, int maxp, const BatchPlan& plan, bool captured) {
  const int full = batch_resident_blocks(b, direction, maxp);
  if (captured) return full;
  const int wpb = b->waves_per_block;
  return (int)std::min<long long>(full, std::max<long long>(1, (plan.total_pairs + wpb - 1) / wpb));
}

// Workgroups of the 28-plane pass over the deferred pairs (rare: a codeblock whose
// MSB is at least the main pass's planes): one per CU.  Its workgroups leave at
// once when nothing was deferred.
int batch_wide_blocks(bpc_batch* b, int maxp, const BatchPlan& plan, bool captured) {
  if (maxp >= 28) return 0;
  if (captured) return b->cus;
  return (int)std::max<long long>(1, std::min<long long>(b->cus, (plan.total_pairs + b->waves_per_block - 1) /
                                                                     b->waves_per_block));
}

// The run's first kernel reads the pinned staging area into device memory; once it
// has, the host may fill the area for the next run.
int batch_upload(bpc_batch* b, uint32_t* d_status, uint32_t* host_status, hipStream_t stream) {
  b->h_run()->status = d_status;
  b->h_run()->host_status = host_status;
  hipError_t e = bpc::launch_batch_stage(b->h_stage_dev, b->d_stage, b->d_ctl, stream);
  if (e != hipSuccess) return hip_fail(e, "batch stage kernel");
  HIP_TRY(hipEventRecord(b->staged, stream));
  b->staged_pending = true;
  return BPC_OK;
}

bpc::BatchArgs batch_args(bpc_batch* b) {
  bpc::BatchArgs a;
  a.scratch = b->d_scratch;
  a.images = b->d_images();
  a.sorted = b->d_sorted;
  a.pair_queue = b->d_pair_queue;
  a.pair_keys = b->d_pair_keys;
  a.wide = b->d_wide;
  a.ctl = b->d_ctl;
  a.run = b->d_run();
  a.wide_pass = 0;
  a.fold = 1;
  a.report_wide = 0;
  return a;
}

// Kernel sequence of one run after the stage kernel (captured into a graph when
// asked): queue build, main pass, 28-plane pass (unless deferred); the last wave
// of the last launch folds the run's status.
// Nothing in it depends on the run's shape but maxp: sizes are read from the run
// words, grids from the device.
int batch_kernels(bpc_batch* b, int direction, int maxp, int blocks, int wide_blocks, bool defer_wide,
                  hipStream_t stream) {
  bpc::QueueArgs q;
  q.run = b->d_run();
  q.images = b->d_images();
  q.cb_offset = b->d_offsets();
  q.sorted_base = b->d_offsets() + b->max_images + 1;
  q.decode = direction == BPC_DECODE;
  q.max_images = b->max_images;
  q.cus = b->cus;
  q.keys = b->d_keys;
  q.sorted = b->d_sorted;
  q.pair_keys = b->d_pair_keys;
  q.pair_queue = b->d_pair_queue;
  q.ctl = b->d_ctl;
  hipError_t e = bpc::launch_queue_build(q, stream);
  if (e != hipSuccess) return hip_fail(e, "queue build");
  bpc::BatchArgs a = batch_args(b);
  const bool wide = wide_blocks > 0 && !defer_wide;
  a.fold = wide ? 0 : 1;  // the last launch of the run folds its status
  a.report_wide = defer_wide && wide_blocks > 0;
  e = bpc::launch_batch(direction == BPC_DECODE, maxp, a, blocks, b->waves_per_block, b->vgpr_cap_waves, stream);
  if (e != hipSuccess) return hip_fail(e, "batched kernel launch");
  if (wide) {
    a.wide_pass = 1;
    a.fold = 1;
    a.report_wide = 0;
    e = bpc::launch_batch(direction == BPC_DECODE, 28, a, wide_blocks, b->waves_per_block, 0, stream);
    if (e != hipSuccess) return hip_fail(e, "batched kernel launch (28 planes)");
  }
  return BPC_OK;
}

// The captured sequence for (direction, maxp, tuning, deferral), captured on first use.
int batch_graph(bpc_batch* b, int direction, int maxp, int blocks, int wide_blocks, bool defer_wide,
                hipGraphExec_t* out) {
  char key[96];
  snprintf(key, sizeof key, "d%d,p%d,wpb%d,wps%d,cap%d,g%d,w%d,f%d", direction, maxp, b->waves_per_block,
           b->waves_per_simd, b->vgpr_cap_waves, blocks, wide_blocks, defer_wide ? 1 : 0);
  auto it = std::find_if(b->graphs.begin(), b->graphs.end(), [&](const auto& kv) { return kv.first == key; });
  if (it != b->graphs.end()) {
    b->graphs.splice(b->graphs.end(), b->graphs, it);  // most recently used last
    *out = b->graphs.back().second;
    return BPC_OK;
  }
  hipStream_t cap = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cap);
    return hip_fail(e, "graph capture");
  }
  int st = batch_kernels(b, direction, maxp, blocks, wide_blocks, defer_wide, cap);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(cap, &graph);
  (void)hipStreamDestroy(cap);
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "graph capture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "graph instantiate");
  if (b->graphs.size() >= bpc_batch::kMaxGraphs) {  // evict the least recently used, once idle
    if (b->done_pending) HIP_TRY(hipEventSynchronize(b->done));
    (void)hipGraphExecDestroy(b->graphs.front().second);
    b->graphs.pop_front();
  }
  b->graphs.emplace_back(key, exec);
  *out = exec;
  return BPC_OK;
}

}  // namespace

namespace bpc {

int batch_run_ex(bpc_batch* b, int direction, const bpc_image* images, int n_images, int use_graph,
                 uint32_t* d_status, uint32_t* host_status, int defer_wide, void* stream) {
  if (!b || !images) return fail(BPC_ERR_ARG, "null batch or images");
  if (direction != BPC_CODE && direction != BPC_DECODE) return fail(BPC_ERR_ARG, "bad direction");
  std::lock_guard<std::mutex> lk(b->mu);
  // the pinned staging area may still be read by the previous run's stage kernel
  if (b->staged_pending) HIP_TRY(hipEventSynchronize(b->staged));
  BatchPlan plan;
  int st = batch_stage(b, direction, images, n_images, plan);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int maxp = plan.maxp;
  const bool captured = use_graph != 0;
  const int blocks = batch_grid_blocks(b, direction, maxp, plan, captured);
  const int wide_blocks = batch_wide_blocks(b, maxp, plan, captured);
  // Bitplane slots for the full resident grid of this tuning (and the 28-plane
  // pass): grown once, when the previous run is done, so that larger batches later
  // do not grow it again (a growth drops the captured graphs, which hold the old
  // address).
  const size_t full = std::max((size_t)batch_resident_blocks(b, direction, maxp) * b->waves_per_block *
                                   bpc::scratch_bytes_per_wave(maxp),
                               (size_t)b->cus * b->waves_per_block * bpc::scratch_bytes_per_wave(28));
  if (full > b->scratch_bytes) {
    if (b->done_pending) HIP_TRY(hipEventSynchronize(b->done));
    drop_graphs(b);
    if (b->d_scratch) HIP_TRY(hipFree(b->d_scratch));
    b->d_scratch = nullptr;
    b->scratch_bytes = 0;
    HIP_TRY(hipMalloc(&b->d_scratch, full));
    b->scratch_bytes = full;
  }
  hipGraphExec_t exec = nullptr;
  if (captured) {  // before the upload: a failed capture leaves no half-run on the stream
    st = batch_graph(b, direction, maxp, blocks, wide_blocks, defer_wide != 0, &exec);
    if (st) return st;
  }
  st = batch_upload(b, d_status, host_status, s);
  if (st) return st;
  if (captured)
    HIP_TRY(hipGraphLaunch(exec, s));
  else
    st = batch_kernels(b, direction, maxp, blocks, wide_blocks, defer_wide != 0, s);
  if (st) return st;
  b->last_direction = direction;
  b->last_maxp = maxp;
  HIP_TRY(hipEventRecord(b->done, s));
  b->done_pending = true;
  return BPC_OK;
}

int batch_run_wide(bpc_batch* b, void* stream) {
  if (!b) return fail(BPC_ERR_ARG, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  if (b->last_maxp >= 28 || b->last_maxp == 0) return BPC_OK;
  hipStream_t s = (hipStream_t)stream;
  bpc::BatchArgs a = batch_args(b);
  a.wide_pass = 1;
  hipError_t e = bpc::launch_batch(b->last_direction == BPC_DECODE, 28, a, b->cus, b->waves_per_block, 0, s);
  if (e != hipSuccess) return hip_fail(e, "batched kernel launch (28 planes)");
  HIP_TRY(hipEventRecord(b->done, s));
  b->done_pending = true;
  return BPC_OK;
}

}  // namespace bpc

extern "C" {

int bpc_batch_create(bpc_batch** out, int max_images, long long max_codeblocks) {
  if (!out || max_images <= 0 || max_images > bpc::MAX_BATCH_IMAGES || max_codeblocks <= 0 ||
      max_codeblocks > ((long long)1 << 32))
    return fail(BPC_ERR_ARG, "bad batch capacity (1..%d images, 1..2^32 codeblocks)", bpc::MAX_BATCH_IMAGES);
  bpc_batch* b = new bpc_batch();
  b->max_images = max_images;
  b->max_cbs = max_codeblocks;
  hipError_t e = hipGetDevice(&b->device);
  int cus = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  b->cus = cus > 0 ? cus : 1;
  const size_t entries = (size_t)max_codeblocks + 2 * (size_t)max_images + 2;
  b->stage_bytes = b->images_offset() + sizeof(bpc::BatchImage) * (size_t)max_images;
  if (e == hipSuccess) e = hipMalloc(&b->d_stage, b->stage_bytes);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_stage, b->stage_bytes, hipHostMallocDefault);
  void* h_dev = nullptr;
  if (e == hipSuccess) e = hipHostGetDevicePointer(&h_dev, b->h_stage, 0);
  b->h_stage_dev = static_cast<const char*>(h_dev);
  if (e == hipSuccess) e = hipMalloc(&b->d_keys, sizeof(uint32_t) * max_codeblocks);
  if (e == hipSuccess) e = hipMalloc(&b->d_sorted, sizeof(uint32_t) * entries);
  if (e == hipSuccess) e = hipMalloc(&b->d_pair_keys, sizeof(uint32_t) * (entries / 2 + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_pair_queue, sizeof(uint32_t) * (entries / 2 + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_wide, sizeof(uint32_t) * (entries / 2 + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_ctl, sizeof(uint32_t) * (bpc::CTL_IMG + 32 * (size_t)max_images));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    batch_release(b);
    delete b;
    return hip_fail(e, "batch buffers");
  }
  memset(b->h_stage, 0, b->stage_bytes);
  *out = b;
  return BPC_OK;
}

int bpc_batch_destroy(bpc_batch* b) {
  if (!b) return BPC_OK;
  if (b->done_pending) (void)hipEventSynchronize(b->done);
  if (b->staged_pending) (void)hipEventSynchronize(b->staged);
  batch_release(b);
  delete b;
  return BPC_OK;
}

int bpc_batch_set_tuning(bpc_batch* b, int waves_per_block, int waves_per_simd) {
  if (!b || waves_per_block < 1 || waves_per_block > 4 || waves_per_simd < 0 || waves_per_simd > 8)
    return fail(BPC_ERR_ARG, "bad tuning");
  std::lock_guard<std::mutex> lk(b->mu);
  b->waves_per_block = waves_per_block;
  b->waves_per_simd = waves_per_simd;
  return BPC_OK;
}

int bpc_batch_set_vgpr_cap(bpc_batch* b, int waves_per_simd) {
  if (!b || (waves_per_simd != 0 && (waves_per_simd < 5 || waves_per_simd > 8)))
    return fail(BPC_ERR_ARG, "register cap must be 0 or 5..8");
  std::lock_guard<std::mutex> lk(b->mu);
  b->vgpr_cap_waves = waves_per_simd;
  return BPC_OK;
}

int bpc_batch_graph_count(bpc_batch* b) {
  if (!b) return fail(BPC_ERR_ARG, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  return (int)b->graphs.size();
}

int bpc_batch_run(bpc_batch* b, int direction, const bpc_image* images, int n_images, int use_graph,
                  uint32_t* d_status, void* stream) {
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  return bpc::batch_run_ex(b, direction, images, n_images, use_graph, d_status, nullptr, 0, stream);
}

}  // extern "C"

// ---- compact codestream container and per-codeblock files ------------------

extern "C" {

int bpc_compact_index_device(const int32_t* d_stream, int64_t ncb, int stride_words, const int32_t* d_ncw,
                             uint8_t* d_msb, uint64_t* d_offsets, uint32_t* d_status, void* stream) {
  if (ncb < 0 || stride_words < 2) return fail(BPC_ERR_ARG, "bad codeblock count or stride");
  if (ncb == 0) return BPC_OK;
  if (!d_stream || !d_msb || !d_offsets) return fail(BPC_ERR_ARG, "null buffer");
  if (ncb > ((int64_t)1 << 31)) return fail(BPC_ERR_ARG, "too many codeblocks");
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  hipError_t e = bpc::launch_compact_index(d_stream, ncb, stride_words, d_ncw, d_msb,
                                           reinterpret_cast<unsigned long long*>(d_offsets), d_status,
                                           (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "compact index launch");
  return BPC_OK;
}

int bpc_compact_pack_device(const int32_t* d_stream, int64_t ncb, int stride_words, const uint64_t* d_offsets,
                            uint16_t* d_payload, uint64_t payload_cap, uint32_t* d_status, void* stream) {
  if (ncb < 0 || stride_words < 2) return fail(BPC_ERR_ARG, "bad codeblock count or stride");
  if (ncb == 0) return BPC_OK;
  if (!d_stream || !d_offsets || (!d_payload && payload_cap)) return fail(BPC_ERR_ARG, "null buffer");
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  hipError_t e = bpc::launch_compact_pack(d_stream, ncb, stride_words,
                                          reinterpret_cast<const unsigned long long*>(d_offsets), d_payload,
                                          payload_cap, d_status, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "compact pack launch");
  return BPC_OK;
}

int bpc_compact_unpack_device(const uint8_t* d_msb, const uint64_t* d_offsets, const uint16_t* d_payload,
                              uint64_t payload_len, int64_t ncb, int stride_words, int32_t* d_stream,
                              uint32_t* d_status, void* stream) {
  if (ncb < 0 || stride_words < 2) return fail(BPC_ERR_ARG, "bad codeblock count or stride");
  if (ncb == 0) return BPC_OK;
  if (!d_msb || !d_offsets || !d_stream || (!d_payload && payload_len)) return fail(BPC_ERR_ARG, "null buffer");
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  hipError_t e = bpc::launch_compact_unpack(d_msb, reinterpret_cast<const unsigned long long*>(d_offsets), d_payload,
                                            payload_len, ncb, stride_words, d_stream, d_status, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "compact unpack launch");
  return BPC_OK;
}

}  // extern "C"

// ---- reversible 5/3 wavelet transform ------------------------------------------

namespace {

int dwt_check(const void* a, const void* b, int size_x, int size_y, int levels) {
  if (!a || !b) return fail(BPC_ERR_ARG, "null buffer");
  if (levels < 0 || levels > 16) return fail(BPC_ERR_ARG, "levels must be 0..16");
  if (size_x <= 0 || size_y <= 0 || size_x % (1 << levels) || size_y % (1 << levels) ||
      (levels > 0 && ((size_x >> levels) < 1 || (size_y >> levels) < 1)))
    return fail(BPC_ERR_ARG, "%dx%d is not divisible by 2^%d", size_x, size_y, levels);
  if ((int64_t)size_x * size_y > (int64_t)1 << 31) return fail(BPC_ERR_ARG, "image too large");
  const char *pa = (const char*)a, *pb = (const char*)b;
  const size_t bytes = (size_t)size_x * size_y * sizeof(int32_t);
  if (pa < pb + bytes && pb < pa + bytes) return fail(BPC_ERR_ARG, "input and output overlap");
  return BPC_OK;
}

// Two LL buffers of the library scratch: level l's LL goes to tmp[l % 2]
// ((X >> (l+1)) x (Y >> (l+1)) samples, so (X/2)(Y/2) and (X/4)(Y/4) suffice).
int dwt_run(int inverse, const int32_t* in, int32_t* out, int X, int Y, int L, hipStream_t stream) {
  if (L == 0) {
    HIP_TRY(hipMemcpyAsync(out, in, (size_t)X * Y * sizeof(int32_t), hipMemcpyDeviceToDevice, stream));
    return BPC_OK;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  const size_t n0 = (size_t)(X / 2) * (Y / 2), n1 = (size_t)(X / 4) * (Y / 4);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  void* base = nullptr;
  int st = scratch_acquire(dev, (n0 + n1 + 64) * sizeof(int32_t), stream, &base);
  if (st) return st;
  int32_t* tmp[2] = {static_cast<int32_t*>(base), static_cast<int32_t*>(base) + ((n0 + 63) & ~(size_t)63)};
  if (!inverse) {
    for (int l = 0; l < L; ++l) {
      bpc::DwtLevel a = {};
      a.nx = X >> l;
      a.ny = Y >> l;
      a.src = l == 0 ? in : tmp[(l - 1) % 2];
      a.src_pitch = l == 0 ? X : (X >> l);
      a.ll = l == L - 1 ? out : tmp[l % 2];
      a.ll_pitch = l == L - 1 ? X : (X >> (l + 1));
      a.det = out;
      a.det_pitch = X;
      hipError_t e = bpc::launch_dwt53_level(0, a, stream);
      if (e != hipSuccess) return hip_fail(e, "dwt forward launch");
    }
  } else {
    for (int l = L - 1; l >= 0; --l) {
      bpc::DwtLevel a = {};
      a.nx = X >> l;
      a.ny = Y >> l;
      a.ll = l == L - 1 ? const_cast<int32_t*>(in) : tmp[l % 2];
      a.ll_pitch = l == L - 1 ? X : (X >> (l + 1));
      a.det = const_cast<int32_t*>(in);
      a.det_pitch = X;
      a.out = l == 0 ? out : tmp[(l - 1) % 2];
      a.out_pitch = l == 0 ? X : (X >> l);
      hipError_t e = bpc::launch_dwt53_level(1, a, stream);
      if (e != hipSuccess) return hip_fail(e, "dwt inverse launch");
    }
  }
  return scratch_release(dev, stream);
}

}  // namespace

extern "C" {

int bpc_dwt53_forward_device(const int32_t* d_img, int32_t* d_coeffs, int size_x, int size_y, int levels,
                             void* stream) {
  int st = dwt_check(d_img, d_coeffs, size_x, size_y, levels);
  if (st) return st;
  return dwt_run(0, d_img, d_coeffs, size_x, size_y, levels, (hipStream_t)stream);
}

int bpc_dwt53_inverse_device(const int32_t* d_coeffs, int32_t* d_img, int size_x, int size_y, int levels,
                             void* stream) {
  int st = dwt_check(d_coeffs, d_img, size_x, size_y, levels);
  if (st) return st;
  return dwt_run(1, d_coeffs, d_img, size_x, size_y, levels, (hipStream_t)stream);
}

}  // extern "C"
