// C-ABI of the BPC-PaCo codec: drop-in run entry (CUDA_BPC, BPC_PaCo.cu:1305),
// device-resident encode/decode, LUT loading / upload / training, HIP helpers.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <atomic>
#include <mutex>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/bpc_paco.h"
#include "bpc_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(BPC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);   \
  } while (0)

std::mutex g_cfg_mu;
bpc_params g_cfg = {0, 5, 64, 15, 0, BPC_LAYOUT_RASTER};

int stride_of(const bpc_params* p) { return p->cb_stride_words > 0 ? p->cb_stride_words : 64 * p->cblock_length; }

int check_params(const bpc_params* p) {
  if (!p) return fail(BPC_ERR_ARG, "null params");
  if (p->cpass3 != 0 && p->cpass3 != 1) return fail(BPC_ERR_ARG, "cpass3 must be 0 or 1");
  if (p->dwt_levels < 0 || p->dwt_levels > 16) return fail(BPC_ERR_ARG, "dwt_levels out of range");
  if (p->cblock_length < 2 || p->cblock_length > 64) return fail(BPC_ERR_ARG, "cblock_length must be 2..64");
  if (p->lut_n_bitplanes < 1 || p->lut_n_bitplanes > 28) return fail(BPC_ERR_ARG, "lut_n_bitplanes must be 1..28");
  if (p->cb_stride_words < 0 || (p->cb_stride_words > 0 && p->cb_stride_words < 2) ||
      p->cb_stride_words > (1 << 26))
    return fail(BPC_ERR_ARG, "cb_stride_words must be 0 or 2..2^26");
  if (p->layout != BPC_LAYOUT_RASTER && p->layout != BPC_LAYOUT_TILED) return fail(BPC_ERR_ARG, "bad layout");
  return BPC_OK;
}

int check_image(int size_x, int size_y, const bpc_params* p) {
  if (size_x <= 0 || size_y <= 0 || size_x % 64 != 0 || size_y % p->cblock_length != 0)
    return fail(BPC_ERR_ARG, "image %dx%d is not a whole number of 64x%d codeblocks", size_x, size_y,
                p->cblock_length);
  if ((int64_t)size_x * size_y > (int64_t)1 << 31) return fail(BPC_ERR_ARG, "image too large");
  return BPC_OK;
}

int status_to_code(uint32_t st) {
  if (st & bpc::ST_MALFORMED) return fail(BPC_ERR_MALFORMED, "decoder: malformed MSB word in a codeblock region");
  if (st & bpc::ST_RANGE)
    return fail(BPC_ERR_RANGE, "coefficient magnitude >= 2^28, LUT entry outside [0,127] or LUT row outside the array");
  if (st & bpc::ST_OVERFLOW) return fail(BPC_ERR_OVERFLOW, "codestream region capacity exceeded");
  if (st & bpc::ST_WIDE) return fail(BPC_ERR_WIDE, "codeblock MSB beyond the kernel's bitplane registers");
  if (st & bpc::ST_NARROW) return fail(BPC_ERR_RANGE, "decoded coefficient outside the int16 range (saturated)");
  return BPC_OK;
}

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
int resident_waves_cached(int dev, int decode, int cpass3, int maxp) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, int> cache;
  const auto key = std::make_tuple(dev, decode, cpass3, maxp);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = bpc::codec_resident_waves_per_cu(decode, cpass3, maxp);
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

int launch(int decode, const int32_t* src, int32_t* dst, int size_x, int size_y, const int32_t* d_lut, int lut_len,
           const bpc_params* p, int fill, int32_t* d_ncw, uint32_t* d_status, hipStream_t stream, int maxp,
           int in16 = 0) {
  bpc::KArgs a;
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
  if (a.ncb <= 0) return BPC_OK;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  int per_cu = resident_waves_cached(dev, decode, p->cpass3, maxp);
  // An image with no more codeblocks than the device has SIMDs cannot fill it with
  // pairs: one codeblock per wave then runs each codeblock's chain on its own SIMD
  // instead of two in lock step (the launch time is the longest chain).
  a.single = (long long)a.ncb <= 4LL * device_cus(dev) ? 1 : 0;
  const long long units = a.single ? a.ncb : (a.ncb + 1) / 2;
  const int grid = (int)std::min<long long>(units, (long long)per_cu * device_cus(dev));
  const size_t need = (size_t)grid * bpc::scratch_bytes_per_wave(maxp);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  int st = scratch_acquire(dev, need, stream, &a.scratch);
  if (st) return st;
  hipError_t e = bpc::launch_codec(decode, p->cpass3, maxp, a, grid, stream);
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

// ---- LUT text format (main_example.cu:219-448) ----------------------------

struct LutRow {
  int bp;
  int v[9];
};

int read_lut_file(const std::string& path, int nvals, std::vector<LutRow>& rows) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return fail(BPC_ERR_ARG, "cannot open %s", path.c_str());
  int lvl, sb, bp;
  char colon[4];
  while (fscanf(f, "%d %d %d %3s", &lvl, &sb, &bp, colon) == 4) {
    LutRow r;
    r.bp = bp;
    for (int k = 0; k < nvals; ++k)
      if (fscanf(f, "%d", &r.v[k]) != 1) {
        fclose(f);
        return fail(BPC_ERR_ARG, "malformed row in %s", path.c_str());
      }
    rows.push_back(r);
  }
  fclose(f);
  return BPC_OK;
}

// Region fill with the loader's rule: a row whose bitplane is <= the previous
// one starts the next group; before moving on, (maxbp - prev)*N entries from
// i + prev*N + 1 are set to 64 (main_example.cu:292-299 and its siblings).
int place_region(const std::vector<LutRow>& rows, int N, int maxbp, int base, int32_t* lut, int cap) {
  int i = base, prev = -1;
  for (const LutRow& r : rows) {
    if (r.bp <= prev) {
      for (int z = 0; z < (maxbp - prev) * N; ++z) {
        const int idx = i + prev * N + z + 1;
        if (idx >= 0 && idx < cap) lut[idx] = 64;
      }
      i += maxbp * N;
    }
    prev = r.bp;
    for (int k = 0; k < N; ++k) {
      const int idx = i + r.bp * N + k;
      if (idx < 0 || idx >= cap) return fail(BPC_ERR_ARG, "LUT file row outside the table");
      lut[idx] = r.v[k];
    }
  }
  return BPC_OK;
}

// ---- LUT trainer --------------------------------------------------------

constexpr uint32_t S_BIT = 0x80000000u, C_BIT = 0x40000000u, R_BIT = 0x20000000u;

struct Counts {
  // [region][group][bitplane][context] -> (n0, n)
  std::vector<uint32_t> n0, n;
};

int group_of(int x, int y, int X, int Y, int levels) {
  for (int aux = 1; aux <= levels; ++aux) {
    const bool bx = x >= (X >> aux), by = y >= (Y >> aux);
    if (bx || by) return (aux - 1) * 3 + (bx ? (by ? 2 : 0) : 1);
  }
  return levels * 3;
}

int chi_of(uint32_t w) { return (w >> 31) == 0 ? 0 : ((w & 1) ? -1 : 1); }

int sign_ctx(int h, int v) {
  if (h == 0) return v == 0 ? 0 : (v > 0 ? 2 : 3);
  if (h > 0) return v == 0 ? 4 : (v > 0 ? 6 : 0);
  return v == 0 ? 5 : (v > 0 ? 1 : 7);
}

// Models one codeblock with the coder's pass and scan order (rows top to bottom,
// even columns then odd columns of each row) and counts the coded symbols.
void train_codeblock(const int32_t* img, int X, int Y, int cb, const bpc_params* p, int G, int nbp,
                     uint32_t* n0, uint32_t* n) {
  const int rows = p->cblock_length;
  const int nx = X / 64;
  const int tx = cb % nx, ty = cb / nx;
  static thread_local std::vector<uint32_t> wbuf;
  wbuf.assign((size_t)rows * 64, 0);
  uint32_t* w = wbuf.data();
  uint32_t m = 0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < 64; ++c) {
      const int32_t v = img[(int64_t)(ty * rows + r) * X + tx * 64 + c];
      const uint32_t mag = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
      if (mag >= (1u << 28)) return;
      m |= mag;
      w[r * 64 + c] = (mag << 1) | (v < 0 ? 1u : 0u) | (p->cpass3 ? C_BIT : 0u);
    }
  if (m == 0) return;
  int msb = 31;
  while (!(m >> msb)) --msb;
  int grp[64];
  for (int c = 0; c < 64; c += 2) grp[c] = grp[c + 1] = group_of(tx * 64 + c, ty * rows, X, Y, p->dwt_levels);
  // region order in the count arrays: 0 ref(1), 1 sig(9), 2 sign(4), 3 cp_sig(9), 4 cp_sign(4)
  auto count = [&](int region, int g, int bp, int ctx, uint32_t sym) {
    if (bp >= nbp) return;
    const size_t idx = (((size_t)region * G + g) * nbp + bp) * 9 + ctx;
    n[idx]++;
    if (!sym) n0[idx]++;
  };
  auto at = [&](int r, int c) -> uint32_t {
    if (r < 0 || r >= rows || c < 0 || c >= 64) return 0u;
    return w[r * 64 + c];
  };
  auto sig_pass = [&](int bp, bool clean) {
    for (int r = 0; r < rows; ++r)
      for (int par = 0; par < 2; ++par)
        for (int c = par; c < 64; c += 2) {
          uint32_t& x = w[r * 64 + c];
          const uint32_t nb[8] = {at(r - 1, c - 1), at(r - 1, c), at(r - 1, c + 1), at(r, c - 1),
                                  at(r, c + 1),     at(r + 1, c - 1), at(r + 1, c), at(r + 1, c + 1)};
          int ctx = 0;
          for (int k = 0; k < 8; ++k) ctx += (int)(nb[k] >> 31);
          if (!clean) {
            if (x & S_BIT) continue;
            if (p->cpass3 && ctx == 0) {
              x |= C_BIT;
              continue;
            }
          } else {
            if (!(x & C_BIT)) continue;
            x &= ~C_BIT;
          }
          const uint32_t sym = (x >> (bp + 1)) & 1u;
          count(clean ? 3 : 1, grp[c], bp, ctx, sym);
          if (sym) {
            x |= S_BIT;
            if (clean) x |= R_BIT;
            const int sc = sign_ctx(chi_of(nb[3]) + chi_of(nb[4]), chi_of(nb[1]) + chi_of(nb[6]));
            count(clean ? 4 : 2, grp[c], bp, sc >> 1, ((x & 1u) == (uint32_t)(sc & 1)) ? 0u : 1u);
          }
        }
  };
  auto ref_pass = [&](int bp) {
    for (int r = 0; r < rows; ++r)
      for (int par = 0; par < 2; ++par)
        for (int c = par; c < 64; c += 2) {
          uint32_t& x = w[r * 64 + c];
          if (x & R_BIT)
            count(0, grp[c], bp, 0, (x >> (bp + 1)) & 1u);
          else if (x & S_BIT)
            x |= R_BIT;
        }
  };
  int bp = msb;
  if (p->cpass3) {
    sig_pass(bp, true);
    bp--;
  }
  for (; bp >= 0; --bp) {
    sig_pass(bp, false);
    ref_pass(bp);
    if (p->cpass3) sig_pass(bp, true);
  }
}

}  // namespace

extern "C" {

void bpc_default_params(bpc_params* p) {
  if (!p) return;
  p->cpass3 = 0;
  p->dwt_levels = 5;
  p->cblock_length = 64;
  p->lut_n_bitplanes = 15;
  p->cb_stride_words = 0;
  p->layout = BPC_LAYOUT_RASTER;
}

int bpc_lut_len(const bpc_params* p) {
  if (!p) return 0;
  const int G = 3 * p->dwt_levels + 1;
  return G * p->lut_n_bitplanes * (1 + 9 + 4 + (p->cpass3 ? 9 + 4 : 0));
}

int64_t bpc_stream_words(int size_x, int size_y, const bpc_params* p) {
  if (!p || p->cblock_length <= 0 || size_x <= 0 || size_y <= 0) return 0;
  return (int64_t)(size_x / 64) * (size_y / p->cblock_length) * stride_of(p);
}

const char* bpc_last_error(void) { return g_last_error.c_str(); }

int bpc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int bpc_set_config(const bpc_params* p) {
  int st = check_params(p);
  if (st) return st;
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_cfg = *p;
  return BPC_OK;
}

int bpc_get_config(bpc_params* p) {
  if (!p) return fail(BPC_ERR_ARG, "null params");
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  *p = g_cfg;
  return BPC_OK;
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

int bpc_lut_parse_text(const char* folder, int cpass3, int dwt_levels, int32_t* host_lut, int cap, int* lut_len,
                       int* nbp) {
  if (!folder || !host_lut || cap <= 0) return fail(BPC_ERR_ARG, "bad arguments");
  if (cpass3 != 0 && cpass3 != 1) return fail(BPC_ERR_ARG, "cpass3 must be 0 or 1");
  std::string dir(folder);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const char* names[5] = {"ref.txt", "sig.txt", "sign.txt", "cp_sig.txt", "cp_sign.txt"};
  const int nv[5] = {1, 9, 4, 9, 4};
  std::vector<LutRow> rows[5];
  const int nfiles = cpass3 ? 5 : 3;
  for (int k = 0; k < nfiles; ++k) {
    int st = read_lut_file(dir + names[k], nv[k], rows[k]);
    if (st) return st;
  }
  int maxbp = -1;  // taken over ref, sig and sign (main_example.cu:231-264)
  for (int k = 0; k < 3; ++k)
    for (const LutRow& r : rows[k]) maxbp = std::max(maxbp, r.bp);
  maxbp++;
  if (maxbp < 1) return fail(BPC_ERR_ARG, "empty LUT files");
  const int G = 3 * dwt_levels + 1;
  const int refN = G * maxbp, sigN = G * maxbp * 9, signN = G * maxbp * 4;
  const int total = refN + sigN + signN + (cpass3 ? sigN + signN : 0);
  if (total > cap) return fail(BPC_ERR_ARG, "LUT needs %d ints, capacity %d", total, cap);
  for (int k = 0; k < cap; ++k) host_lut[k] = 64;
  const int bases[5] = {0, refN, refN + sigN, refN + sigN + signN, refN + 2 * sigN + signN};
  for (int k = 0; k < nfiles; ++k) {
    int st = place_region(rows[k], nv[k], maxbp, bases[k], host_lut, cap);
    if (st) return st;
  }
  if (lut_len) *lut_len = total;
  if (nbp) *nbp = maxbp;
  return BPC_OK;
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

int bpc_lut_train(const int32_t* img, int size_x, int size_y, const bpc_params* p, int32_t* host_lut, int lut_len,
                  int nthreads) {
  int st = check_params(p);
  if (st) return st;
  st = check_image(size_x, size_y, p);
  if (st) return st;
  if (!img || !host_lut) return fail(BPC_ERR_ARG, "null buffer");
  if (p->layout != BPC_LAYOUT_RASTER) return fail(BPC_ERR_ARG, "training image must be raster");
  if (lut_len < bpc_lut_len(p)) return fail(BPC_ERR_ARG, "lut_len too small");
  const int G = 3 * p->dwt_levels + 1;
  const int nbp = p->lut_n_bitplanes;
  const size_t ncnt = (size_t)5 * G * nbp * 9;
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  std::vector<uint32_t> n0(ncnt, 0), n(ncnt, 0);
  int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, std::max(1, ncb)));
  std::atomic<int> next(0);
  std::mutex mu;
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      std::vector<uint32_t> l0(ncnt, 0), ln(ncnt, 0);
      for (;;) {
        const int cb0 = next.fetch_add(8);
        if (cb0 >= ncb) break;
        for (int cb = cb0; cb < std::min(ncb, cb0 + 8); ++cb)
          train_codeblock(img, size_x, size_y, cb, p, G, nbp, l0.data(), ln.data());
      }
      std::lock_guard<std::mutex> lk(mu);
      for (size_t k = 0; k < ncnt; ++k) {
        n0[k] += l0[k];
        n[k] += ln[k];
      }
    });
  for (auto& th : pool) th.join();
  // layout: ref | sig | sign | cp_sig | cp_sign (BPC_PaCo.cu:228-250, :988-989)
  const int nctx[5] = {1, 9, 4, 9, 4};
  int base = 0;
  for (int region = 0; region < (p->cpass3 ? 5 : 3); ++region) {
    for (int g = 0; g < G; ++g)
      for (int bp = 0; bp < nbp; ++bp)
        for (int ctx = 0; ctx < nctx[region]; ++ctx) {
          const size_t idx = (((size_t)region * G + g) * nbp + bp) * 9 + ctx;
          int v = 64;
          if (n[idx]) {
            v = (int)(((uint64_t)n0[idx] * 128 + n[idx] / 2) / n[idx]);
            v = std::min(127, std::max(1, v));
          }
          host_lut[base + (g * nbp + bp) * nctx[region] + ctx] = v;
        }
    base += G * nbp * nctx[region];
  }
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

struct bpc_batch {
  int device = 0;
  int max_images = 0;
  long long max_cbs = 0;
  bpc::BatchImage* d_images = nullptr;
  bpc::BatchImage* h_images = nullptr;  // pinned staging for the descriptor upload
  long long* d_offsets = nullptr;       // cb_offset[n+1], sorted_base[n]
  long long* h_offsets = nullptr;
  uint32_t* d_keys = nullptr;
  uint32_t* d_sorted = nullptr;
  uint32_t* d_pair_keys = nullptr;
  uint32_t* d_pair_queue = nullptr;
  uint32_t* d_ctl = nullptr;
  hipEvent_t staged = nullptr;  // the last descriptor upload has been consumed
  bool staged_pending = false;
  hipEvent_t done = nullptr;  // the last run has finished
  bool done_pending = false;
  void* d_scratch = nullptr;  // bitplane scratch, one slot per resident wave
  size_t scratch_bytes = 0;
  int waves_per_block = 4;
  int waves_per_simd = 0;  // 0: as many as fit
  int vgpr_cap_waves = 0;  // 0: compiler's register budget; 5..8: capped for that many waves per SIMD
  int cus = 0;
  std::map<std::string, hipGraphExec_t> graphs;
  std::mutex mu;
};

namespace {

void batch_release(bpc_batch* b) {
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  void* dev[] = {b->d_images, b->d_offsets, b->d_keys,  b->d_sorted,
                 b->d_pair_keys, b->d_pair_queue, b->d_ctl, b->d_scratch};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (b->h_images) (void)hipHostFree(b->h_images);
  if (b->h_offsets) (void)hipHostFree(b->h_offsets);
  if (b->staged) (void)hipEventDestroy(b->staged);
  if (b->done) (void)hipEventDestroy(b->done);
}

struct BatchPlan {
  long long total_cbs = 0, total_pairs = 0;
  int maxp = 16;
  int rows = 0;
  std::string signature;
};

int batch_stage(bpc_batch* b, int direction, const bpc_image* images, int n, BatchPlan& plan) {
  if (n <= 0 || n > b->max_images) return fail(BPC_ERR_ARG, "batch of %d images (capacity %d)", n, b->max_images);
  if (n > 4096) return fail(BPC_ERR_ARG, "at most 4096 images per batch");
  plan.rows = images[0].params.cblock_length;
  int max_nbp = 1;
  char sig[128];
  plan.signature = std::to_string(direction) + ":";
  long long pos = 0;
  for (int k = 0; k < n; ++k) {
    const bpc_image& im = images[k];
    int st = check_params(&im.params);
    if (st) return st;
    if (im.params.cblock_length != plan.rows) return fail(BPC_ERR_ARG, "all images of a batch share cblock_length");
    st = check_image(im.size_x, im.size_y, &im.params);
    if (st) return st;
    if (!im.src || !im.dst || !im.lut) return fail(BPC_ERR_ARG, "image %d: null buffer", k);
    if (((uintptr_t)im.src & 7) || ((uintptr_t)im.dst & 7)) return fail(BPC_ERR_ARG, "image %d: buffers must be 8-byte aligned", k);
    if (im.lut_len < bpc_lut_len(&im.params)) return fail(BPC_ERR_ARG, "image %d: lut_len too small", k);
    const long long ncb = (long long)(im.size_x / 64) * (im.size_y / plan.rows);
    if (ncb > (1 << 20)) return fail(BPC_ERR_ARG, "image %d: more than 2^20 codeblocks", k);
    bpc::BatchImage& d = b->h_images[k];
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
    d.pad = 0;
    b->h_offsets[k] = plan.total_cbs;
    b->h_offsets[n + 1 + k] = pos;
    plan.total_cbs += ncb;
    pos += (ncb + 1) & ~1ll;
    max_nbp = std::max(max_nbp, im.params.lut_n_bitplanes);
    snprintf(sig, sizeof sig, "%dx%d/%d%d%d%d/%d;", im.size_x, im.size_y, im.params.cpass3, im.params.dwt_levels,
             im.params.lut_n_bitplanes, im.params.layout, d.stride);
    plan.signature += sig;
  }
  b->h_offsets[n] = plan.total_cbs;
  plan.total_pairs = pos / 2;
  if (plan.total_cbs > b->max_cbs) return fail(BPC_ERR_ARG, "batch of %lld codeblocks (capacity %lld)", plan.total_cbs, b->max_cbs);
  plan.maxp = bpc::planes_for(max_nbp);
  return BPC_OK;
}

int batch_grid_blocks(bpc_batch* b, int direction, int maxp, const BatchPlan& plan) {
  if (plan.total_pairs <= 0) return 0;
  const int wpb = b->waves_per_block;
  int per_cu = bpc::batch_resident_blocks_per_cu(direction == BPC_DECODE, maxp, wpb, b->vgpr_cap_waves);
  if (per_cu <= 0) per_cu = 1;
  if (b->waves_per_simd > 0) per_cu = std::max(1, std::min(per_cu, (4 * b->waves_per_simd) / wpb));
  long long blocks = (long long)per_cu * b->cus;
  blocks = std::min(blocks, (plan.total_pairs + wpb - 1) / wpb);
  return (int)blocks;
}

int batch_upload(bpc_batch* b, int n, hipStream_t stream) {
  HIP_TRY(hipMemcpyAsync(b->d_images, b->h_images, sizeof(bpc::BatchImage) * n, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(b->d_offsets, b->h_offsets, sizeof(long long) * (2 * n + 1), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipEventRecord(b->staged, stream));
  b->staged_pending = true;
  return BPC_OK;
}

// Kernel sequence of one batch run (captured into a graph when asked).
int batch_kernels(bpc_batch* b, int direction, int n, const BatchPlan& plan, int maxp, int blocks,
                  uint32_t* d_status, hipStream_t stream) {
  bpc::QueueArgs q;
  q.images = b->d_images;
  q.cb_offset = b->d_offsets;
  q.sorted_base = b->d_offsets + n + 1;
  q.n_images = n;
  q.total_cbs = plan.total_cbs;
  q.total_pairs = plan.total_pairs;
  q.rows = plan.rows;
  q.decode = direction == BPC_DECODE;
  q.keys = b->d_keys;
  q.sorted = b->d_sorted;
  q.pair_keys = b->d_pair_keys;
  q.pair_queue = b->d_pair_queue;
  q.ctl = b->d_ctl;
  hipError_t e = bpc::launch_queue_build(q, stream);
  if (e != hipSuccess) return hip_fail(e, "queue build");
  bpc::BatchArgs a;
  a.scratch = b->d_scratch;
  a.images = b->d_images;
  a.sorted = b->d_sorted;
  a.pair_queue = b->d_pair_queue;
  a.ctl = b->d_ctl;
  a.n_pairs = plan.total_pairs;
  a.rows = plan.rows;
  if (blocks > 0) {
    e = bpc::launch_batch(direction == BPC_DECODE, maxp, a, blocks, b->waves_per_block, b->vgpr_cap_waves, stream);
    if (e != hipSuccess) return hip_fail(e, "batched kernel launch");
  }
  e = bpc::launch_status_fold(b->d_ctl, d_status, stream);
  if (e != hipSuccess) return hip_fail(e, "status fold");
  return BPC_OK;
}

}  // namespace

extern "C" {

int bpc_batch_create(bpc_batch** out, int max_images, long long max_codeblocks) {
  if (!out || max_images <= 0 || max_images > 4096 || max_codeblocks <= 0)
    return fail(BPC_ERR_ARG, "bad batch capacity");
  bpc_batch* b = new bpc_batch();
  b->max_images = max_images;
  b->max_cbs = max_codeblocks;
  hipError_t e = hipGetDevice(&b->device);
  int cus = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  b->cus = cus > 0 ? cus : 1;
  const size_t entries = (size_t)max_codeblocks + 2 * (size_t)max_images + 2;
  if (e == hipSuccess) e = hipMalloc(&b->d_images, sizeof(bpc::BatchImage) * max_images);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_images, sizeof(bpc::BatchImage) * max_images, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->d_offsets, sizeof(long long) * (2 * max_images + 1));
  if (e == hipSuccess) e = hipHostMalloc(&b->h_offsets, sizeof(long long) * (2 * max_images + 1), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&b->d_keys, sizeof(uint32_t) * max_codeblocks);
  if (e == hipSuccess) e = hipMalloc(&b->d_sorted, sizeof(uint32_t) * entries);
  if (e == hipSuccess) e = hipMalloc(&b->d_pair_keys, sizeof(uint32_t) * (entries / 2 + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_pair_queue, sizeof(uint32_t) * (entries / 2 + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_ctl, sizeof(uint32_t) * (bpc::CTL_IMG + 32 * (size_t)max_images));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e != hipSuccess) {
    batch_release(b);
    delete b;
    return hip_fail(e, "batch buffers");
  }
  *out = b;
  return BPC_OK;
}

int bpc_batch_destroy(bpc_batch* b) {
  if (!b) return BPC_OK;
  if (b->done_pending) (void)hipEventSynchronize(b->done);
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

int bpc_batch_run(bpc_batch* b, int direction, const bpc_image* images, int n_images, int use_graph,
                  uint32_t* d_status, void* stream) {
  if (!b || !images) return fail(BPC_ERR_ARG, "null batch or images");
  if (direction != BPC_CODE && direction != BPC_DECODE) return fail(BPC_ERR_ARG, "bad direction");
  std::lock_guard<std::mutex> lk(b->mu);
  if (!d_status) {
    int st = default_status(&d_status);
    if (st) return st;
  }
  // the pinned staging area may still be read by the previous run's upload
  if (b->staged_pending) HIP_TRY(hipEventSynchronize(b->staged));
  BatchPlan plan;
  int st = batch_stage(b, direction, images, n_images, plan);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int maxp = plan.maxp;
  const int blocks = batch_grid_blocks(b, direction, maxp, plan);
  const size_t need = (size_t)blocks * b->waves_per_block * bpc::scratch_bytes_per_wave(maxp);
  if (need > b->scratch_bytes) {  // grow the scratch once the previous run is done
    if (b->done_pending) HIP_TRY(hipEventSynchronize(b->done));
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
    if (b->d_scratch) HIP_TRY(hipFree(b->d_scratch));
    b->d_scratch = nullptr;
    b->scratch_bytes = 0;
    HIP_TRY(hipMalloc(&b->d_scratch, need));
    b->scratch_bytes = need;
  }
  st = batch_upload(b, n_images, s);
  if (st) return st;
  if (!use_graph) {
    st = batch_kernels(b, direction, n_images, plan, maxp, blocks, d_status, s);
  } else {
    char tun[96];
    snprintf(tun, sizeof tun, "|%d,%d,%d,%d,%d,%p", b->waves_per_block, b->waves_per_simd, b->vgpr_cap_waves, maxp, blocks,
             (void*)d_status);
    const std::string key = plan.signature + tun;
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      hipStream_t cap = nullptr;
      HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        (void)hipStreamDestroy(cap);
        return hip_fail(e, "graph capture");
      }
      st = batch_kernels(b, direction, n_images, plan, maxp, blocks, d_status, cap);
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
      it = b->graphs.emplace(key, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, s));
  }
  if (st) return st;
  HIP_TRY(hipEventRecord(b->done, s));
  b->done_pending = true;
  return BPC_OK;
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

// load_Data's codeblock file order (main_example.cu:71-179): the LL band, then per
// level from the coarsest the HL, LH and HH bands, each in row-major codeblock steps.
int bpc_cb_file_order(int size_x, int size_y, int dwt_levels, int cblock_length, int32_t* tile_of_file, int cap) {
  if (size_x <= 0 || size_y <= 0 || size_x % 64 || cblock_length < 2 || cblock_length > 64 ||
      size_y % cblock_length || dwt_levels < 0 || dwt_levels > 16 || !tile_of_file)
    return fail(BPC_ERR_ARG, "bad image geometry");
  const int nx = size_x / 64;
  const int ncb = nx * (size_y / cblock_length);
  if (cap < ncb) return fail(BPC_ERR_ARG, "capacity %d < %d codeblocks", cap, ncb);
  std::vector<char> seen(ncb, 0);
  int k = 0;
  auto band = [&](int x0, int y0, int sx, int sy) -> bool {
    if (sx % 64 || sy % cblock_length) return false;
    for (int j = 0; j < sy; j += cblock_length)
      for (int i = 0; i < sx; i += 64) {
        const int t = ((y0 + j) / cblock_length) * nx + (x0 + i) / 64;
        if (k >= ncb || seen[t]) return false;
        seen[t] = 1;
        tile_of_file[k++] = t;
      }
    return true;
  };
  bool ok = band(0, 0, size_x >> dwt_levels, size_y >> dwt_levels);
  for (int d = dwt_levels; ok && d > 0; --d) {
    const int sx = size_x >> d, sy = size_y >> d;
    ok = band(sx, 0, sx, sy) && band(0, sy, sx, sy) && band(sx, sy, sx, sy);
  }
  if (!ok || k != ncb)
    return fail(BPC_ERR_ARG, "%dx%d with %d levels: subbands are not whole 64x%d codeblocks (no file order)", size_x,
                size_y, dwt_levels, cblock_length);
  return k;
}

// One file per codeblock, named by its index in the file order: the MSB and the
// codewords as big-endian 16-bit values (what load_CStream, main_example.cu:51-68,
// reads back into the reference layout, every word past the end reading as -1).
int bpc_bin_export(const char* folder, const uint8_t* msb, const uint64_t* offsets, const uint16_t* payload,
                   int size_x, int size_y, const bpc_params* p) {
  int st = check_params(p);
  if (st) return st;
  if (!folder || !msb || !offsets) return fail(BPC_ERR_ARG, "null argument");
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  std::vector<int32_t> order(ncb > 0 ? ncb : 1);
  int n = bpc_cb_file_order(size_x, size_y, p->dwt_levels, p->cblock_length, order.data(), ncb);
  if (n < 0) return n;
  std::string dir(folder);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  std::vector<uint8_t> buf;
  for (int k = 0; k < n; ++k) {
    const int t = order[k];
    const uint64_t o = offsets[t], e = offsets[t + 1];
    if (e < o || e - o > (uint64_t)stride_of(p) - 1) return fail(BPC_ERR_ARG, "codeblock %d: bad offsets", t);
    buf.resize(2 * (1 + (e - o)));
    buf[0] = 0;
    buf[1] = msb[t];
    for (uint64_t j = o; j < e; ++j) {
      buf[2 * (1 + j - o)] = (uint8_t)(payload[j] >> 8);
      buf[2 * (1 + j - o) + 1] = (uint8_t)(payload[j] & 0xFF);
    }
    const std::string path = dir + std::to_string(k) + ".bin";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return fail(BPC_ERR_ARG, "cannot write %s", path.c_str());
    const size_t w = fwrite(buf.data(), 1, buf.size(), f);
    const int c = fclose(f);
    if (w != buf.size() || c != 0) return fail(BPC_ERR_ARG, "short write to %s", path.c_str());
  }
  return BPC_OK;
}

// load_Data with CStreamDataLoad (main_example.cu:71-179, 51-68): file k fills the
// region of tile order[k] with stride big-endian 16-bit words, -1 past the end.
int bpc_bin_import(const char* folder, int size_x, int size_y, const bpc_params* p, int32_t* host_stream) {
  int st = check_params(p);
  if (st) return st;
  if (!folder || !host_stream) return fail(BPC_ERR_ARG, "null argument");
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  std::vector<int32_t> order(ncb > 0 ? ncb : 1);
  int n = bpc_cb_file_order(size_x, size_y, p->dwt_levels, p->cblock_length, order.data(), ncb);
  if (n < 0) return n;
  std::string dir(folder);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const int stride = stride_of(p);
  std::vector<uint8_t> buf((size_t)2 * stride);
  for (int k = 0; k < n; ++k) {
    const std::string path = dir + std::to_string(k) + ".bin";
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return fail(BPC_ERR_ARG, "cannot read %s", path.c_str());
    const size_t got = fread(buf.data(), 1, buf.size(), f);
    fclose(f);
    int32_t* region = host_stream + (int64_t)order[k] * stride;
    for (int i = 0; i < stride; ++i) {
      const size_t b = 2 * (size_t)i;
      const uint32_t hi = b < got ? buf[b] : 0xFFFFFFFFu;
      const uint32_t lo = b + 1 < got ? buf[b + 1] : 0xFFFFFFFFu;
      region[i] = (int32_t)((hi << 8) | lo);
    }
  }
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
