// Host-only part of the C-ABI: parameters and their validation, the reference's
// LUT text format (load_LUTs, main_example.cu:219-448), the host LUT trainer,
// per-codeblock codestream files (load_Data / load_CStream, main_example.cu:51-179),
// the batch queue's limits and the server's admission rule.  No HIP here.
#include "bpc_host.h"

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace bpc {

namespace {
thread_local std::string g_last_error;
std::mutex g_cfg_mu;
bpc_params g_cfg = {0, 5, 64, 15, 0, BPC_LAYOUT_RASTER};
}  // namespace

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

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
  if (st & ST_MALFORMED) return fail(BPC_ERR_MALFORMED, "decoder: malformed MSB word in a codeblock region");
  if (st & ST_RANGE)
    return fail(BPC_ERR_RANGE, "coefficient magnitude >= 2^28, LUT entry outside [0,127] or LUT row outside the array");
  if (st & ST_OVERFLOW) return fail(BPC_ERR_OVERFLOW, "codestream region capacity exceeded");
  if (st & ST_WIDE) return fail(BPC_ERR_WIDE, "codeblock MSB beyond the kernel's bitplane registers");
  if (st & ST_NARROW) return fail(BPC_ERR_RANGE, "decoded coefficient outside the int16 range (saturated)");
  return BPC_OK;
}

long long batch_image_codeblocks(int size_x, int size_y, const bpc_params* p) {
  int st = check_params(p);
  if (st) return st;
  st = check_image(size_x, size_y, p);
  if (st) return st;
  const long long ncb = (long long)(size_x / 64) * (size_y / p->cblock_length);
  if (ncb > MAX_IMAGE_CODEBLOCKS)
    return fail(BPC_ERR_ARG, "image of %lld codeblocks (a batched image holds at most %lld)", ncb,
                MAX_IMAGE_CODEBLOCKS);
  return ncb;
}

Admit AdmissionRule::decide(size_t committed, size_t out_bytes, size_t new_bytes, size_t free_now,
                            size_t pooled_other) const {
  if (budget && (committed > budget || out_bytes > budget - committed)) return Admit::WAIT;
  if (new_bytes == 0) return Admit::YES;  // served from the pool: no new device memory
  const size_t need = new_bytes + reserve;
  if (free_now >= need) return Admit::YES;
  // pooled buffers of other sizes count as reclaimable free memory
  if (pooled_other && free_now + pooled_other >= need) return Admit::TRIM_POOL;
  return Admit::WAIT;
}

bool AdmissionRule::possible(size_t out_bytes, size_t total_bytes) const {
  if (budget && out_bytes > budget) return false;
  return out_bytes <= total_bytes && reserve <= total_bytes - out_bytes;
}

}  // namespace bpc

using bpc::fail;

// ---- LUT text format and host trainer -----------------------------------------

namespace {

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
  int64_t i = base;
  int prev = -1;
  for (const LutRow& r : rows) {
    if (r.bp < 0) return fail(BPC_ERR_ARG, "negative bitplane in a LUT file");
    if (r.bp <= prev) {
      for (int z = 0; z < (maxbp - prev) * N; ++z) {
        const int64_t idx = i + (int64_t)prev * N + z + 1;
        if (idx >= 0 && idx < cap) lut[idx] = 64;
      }
      i += (int64_t)maxbp * N;
    }
    prev = r.bp;
    for (int k = 0; k < N; ++k) {
      const int64_t idx = i + (int64_t)r.bp * N + k;
      if (idx < 0 || idx >= cap) return fail(BPC_ERR_ARG, "LUT file row outside the table");
      lut[idx] = r.v[k];
    }
  }
  return BPC_OK;
}

constexpr uint32_t S_BIT = 0x80000000u, C_BIT = 0x40000000u, R_BIT = 0x20000000u;

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
void train_codeblock(const int32_t* img, int X, int Y, int cb, const bpc_params* p, int G, int nbp, uint32_t* n0,
                     uint32_t* n) {
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
  return (int64_t)(size_x / 64) * (size_y / p->cblock_length) * bpc::stride_of(p);
}

const char* bpc_last_error(void) { return bpc::g_last_error.c_str(); }

int bpc_set_config(const bpc_params* p) {
  int st = bpc::check_params(p);
  if (st) return st;
  std::lock_guard<std::mutex> lk(bpc::g_cfg_mu);
  bpc::g_cfg = *p;
  return BPC_OK;
}

int bpc_get_config(bpc_params* p) {
  if (!p) return fail(BPC_ERR_ARG, "null params");
  std::lock_guard<std::mutex> lk(bpc::g_cfg_mu);
  *p = bpc::g_cfg;
  return BPC_OK;
}

int bpc_lut_parse_text(const char* folder, int cpass3, int dwt_levels, int32_t* host_lut, int cap, int* lut_len,
                       int* nbp) {
  if (!folder || !host_lut || cap <= 0) return fail(BPC_ERR_ARG, "bad arguments");
  if (cpass3 != 0 && cpass3 != 1) return fail(BPC_ERR_ARG, "cpass3 must be 0 or 1");
  if (dwt_levels < 0 || dwt_levels > 16) return fail(BPC_ERR_ARG, "dwt_levels out of range");
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
  if (maxbp > 28) return fail(BPC_ERR_ARG, "LUT files name bitplane %d (at most 27)", maxbp - 1);
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

int bpc_lut_train(const int32_t* img, int size_x, int size_y, const bpc_params* p, int32_t* host_lut, int lut_len,
                  int nthreads) {
  int st = bpc::check_params(p);
  if (st) return st;
  st = bpc::check_image(size_x, size_y, p);
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

// load_Data's codeblock file order (main_example.cu:71-179): the LL band, then per
// level from the coarsest the HL, LH and HH bands, each in row-major codeblock steps.
int bpc_cb_file_order(int size_x, int size_y, int dwt_levels, int cblock_length, int32_t* tile_of_file, int cap) {
  if (size_x <= 0 || size_y <= 0 || size_x % 64 || cblock_length < 2 || cblock_length > 64 ||
      size_y % cblock_length || dwt_levels < 0 || dwt_levels > 16 || !tile_of_file)
    return fail(BPC_ERR_ARG, "bad image geometry");
  if ((int64_t)size_x * size_y > (int64_t)1 << 31) return fail(BPC_ERR_ARG, "image too large");
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
        if (k >= ncb || t < 0 || t >= ncb || seen[t]) return false;
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
  int st = bpc::check_params(p);
  if (st) return st;
  if (!folder || !msb || !offsets) return fail(BPC_ERR_ARG, "null argument");
  if (size_x <= 0 || size_y <= 0) return fail(BPC_ERR_ARG, "bad image geometry");
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  std::vector<int32_t> order(ncb > 0 ? ncb : 1);
  int n = bpc_cb_file_order(size_x, size_y, p->dwt_levels, p->cblock_length, order.data(), ncb);
  if (n < 0) return n;
  if (offsets[ncb] && !payload) return fail(BPC_ERR_ARG, "null payload");
  std::string dir(folder);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  std::vector<uint8_t> buf;
  for (int k = 0; k < n; ++k) {
    const int t = order[k];
    const uint64_t o = offsets[t], e = offsets[t + 1];
    if (e < o || e - o > (uint64_t)bpc::stride_of(p) - 1) return fail(BPC_ERR_ARG, "codeblock %d: bad offsets", t);
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
  int st = bpc::check_params(p);
  if (st) return st;
  if (!folder || !host_stream) return fail(BPC_ERR_ARG, "null argument");
  if (size_x <= 0 || size_y <= 0) return fail(BPC_ERR_ARG, "bad image geometry");
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  std::vector<int32_t> order(ncb > 0 ? ncb : 1);
  int n = bpc_cb_file_order(size_x, size_y, p->dwt_levels, p->cblock_length, order.data(), ncb);
  if (n < 0) return n;
  std::string dir(folder);
  if (!dir.empty() && dir.back() != '/') dir += '/';
  const int stride = bpc::stride_of(p);
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
