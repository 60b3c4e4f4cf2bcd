// Host-only code of the library (argument checks, batch queue entries, server
// admission rule, LUT text parser, host LUT trainer, codeblock file order and
// .bin files) and the CPU oracle, exercised in one program so that it can run
// under the address and undefined-behaviour sanitizers.  Built and run by
// tests/test_host_sanitized.py.  Usage: host_selftest <lut folder root> <tmp dir>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../cuda_bpc-paco_amd/csrc/bpc_host.h"
#include "../../oracle/bpc_oracle.h"

static int g_checks = 0, g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    ++g_checks;                                                         \
    if (!(c)) {                                                         \
      ++g_fail;                                                         \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
    }                                                                   \
  } while (0)

static uint64_t g_rng = 88172645463325252ull;
static uint64_t rnd() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return g_rng;
}

// Transform-shaped test data: magnitudes shrink with the band level, random signs.
static std::vector<int32_t> synth(int X, int Y, int levels, int bits) {
  std::vector<int32_t> v((size_t)X * Y);
  for (int y = 0; y < Y; ++y)
    for (int x = 0; x < X; ++x) {
      int lvl = levels;
      for (int a = 1; a <= levels; ++a)
        if (x >= (X >> a) || y >= (Y >> a)) {
          lvl = a - 1;
          break;
        }
      const int b = std::max(1, std::min(bits, lvl == levels ? bits : 3 + 2 * lvl));
      int32_t m = (int32_t)(rnd() % (1u << b));
      if (lvl != levels && (rnd() & 3)) m >>= (int)(rnd() % (unsigned)b);
      v[(size_t)y * X + x] = (rnd() & 1) ? -m : m;
    }
  return v;
}

static bpc_params bparams(int cp3, int levels, int rows, int nbp, int stride = 0) {
  bpc_params p;
  bpc_default_params(&p);
  p.cpass3 = cp3;
  p.dwt_levels = levels;
  p.cblock_length = rows;
  p.lut_n_bitplanes = nbp;
  p.cb_stride_words = stride;
  return p;
}

static orc_params oparams(const bpc_params& b) {
  orc_params o = {b.cpass3, b.dwt_levels, b.cblock_length, b.lut_n_bitplanes, b.cb_stride_words};
  return o;
}

static void test_checks() {
  bpc_params p = bparams(0, 5, 64, 15);
  CHECK(bpc::check_params(&p) == BPC_OK);
  CHECK(bpc::check_params(nullptr) == BPC_ERR_ARG);
  bpc_params q = p;
  q.cblock_length = 1;
  CHECK(bpc::check_params(&q) == BPC_ERR_ARG);
  q = p;
  q.lut_n_bitplanes = 29;
  CHECK(bpc::check_params(&q) == BPC_ERR_ARG);
  q = p;
  q.cb_stride_words = 1;
  CHECK(bpc::check_params(&q) == BPC_ERR_ARG);
  CHECK(bpc::check_image(2048, 2048, &p) == BPC_OK);
  CHECK(bpc::check_image(2048 + 32, 2048, &p) == BPC_ERR_ARG);
  CHECK(bpc::check_image(65536, 65536, &p) == BPC_ERR_ARG);
  CHECK(bpc::status_to_code(0) == BPC_OK);
  CHECK(bpc::status_to_code(bpc::ST_OVERFLOW) == BPC_ERR_OVERFLOW);
  CHECK(bpc::status_to_code(bpc::ST_MALFORMED | bpc::ST_RANGE) == BPC_ERR_MALFORMED);
  CHECK(strlen(bpc_last_error()) > 0);
  CHECK(bpc_lut_len(&p) == 16 * 15 * 14);
  p.cpass3 = 1;
  CHECK(bpc_lut_len(&p) == 16 * 15 * 27);
}

static void test_queue_entries() {
  // every (image, codeblock) a batch admits packs to a distinct value that is not
  // the empty marker, and unpacks to itself
  const uint32_t imgs[] = {0, 1, 2047, 4094, 4095};
  const uint32_t cbs[] = {0, 1, 0xFFFFEu, (uint32_t)bpc::MAX_IMAGE_CODEBLOCKS - 1};
  for (uint32_t i : imgs)
    for (uint32_t c : cbs) {
      const uint32_t e = bpc::queue_entry(i, c);
      CHECK(e != bpc::QUEUE_EMPTY);
      CHECK(bpc::queue_entry_image(e) == i && bpc::queue_entry_cb(e) == c);
    }
  static_assert(bpc::queue_entry(bpc::MAX_BATCH_IMAGES - 1, bpc::MAX_IMAGE_CODEBLOCKS - 1) != bpc::QUEUE_EMPTY,
                "the largest entry must differ from the empty marker");
  // an image of exactly 2^20 codeblocks would let image 4095 reach the marker: refused
  bpc_params p = bparams(0, 0, 2, 15);
  CHECK(bpc::batch_image_codeblocks(64 * 1024, 2 * 1024, &p) == BPC_ERR_ARG);
  CHECK(bpc::batch_image_codeblocks(64 * 1024, 2 * 1023, &p) == 1024LL * 1023);
  p = bparams(1, 5, 64, 16);
  CHECK(bpc::batch_image_codeblocks(16384, 16384, &p) == 65536);
  CHECK(bpc::batch_image_codeblocks(100, 64, &p) == BPC_ERR_ARG);
}

static void test_admission() {
  const size_t MiB = 1u << 20, GiB = 1u << 30;
  bpc::AdmissionRule r;
  r.budget = 0;
  r.reserve = 256 * MiB;
  // 1 GiB free, a 1 GiB output: refused while memory is short
  CHECK(r.decide(0, GiB, GiB, GiB, 0) == bpc::Admit::WAIT);
  // the same request once 2 GiB are free (another request was released)
  CHECK(r.decide(0, GiB, GiB, 2 * GiB, 0) == bpc::Admit::YES);
  // served from the pool: admitted even with no free memory
  CHECK(r.decide(5 * GiB, GiB, 0, 0, 0) == bpc::Admit::YES);
  // pooled buffers of other sizes can be freed to make room
  CHECK(r.decide(0, GiB, GiB, 512 * MiB, GiB) == bpc::Admit::TRIM_POOL);
  CHECK(r.decide(0, GiB, GiB, 512 * MiB, 512 * MiB) == bpc::Admit::WAIT);
  // reserve is kept exactly
  CHECK(r.decide(0, GiB, GiB, GiB + 256 * MiB, 0) == bpc::Admit::YES);
  CHECK(r.decide(0, GiB, GiB, GiB + 256 * MiB - 1, 0) == bpc::Admit::WAIT);
  // an output budget caps what admitted requests hold
  r.budget = 3 * GiB;
  CHECK(r.decide(2 * GiB, GiB, GiB, 100 * GiB, 0) == bpc::Admit::YES);
  CHECK(r.decide(2 * GiB + 1, GiB, GiB, 100 * GiB, 0) == bpc::Admit::WAIT);
  CHECK(r.decide(4 * GiB, GiB, 0, 100 * GiB, 0) == bpc::Admit::WAIT);
  CHECK(r.possible(3 * GiB, 288 * GiB));
  CHECK(!r.possible(3 * GiB + 1, 288 * GiB));
  r.budget = 0;
  CHECK(r.possible(288 * GiB - 256 * MiB, 288 * GiB));
  CHECK(!r.possible(288 * GiB - 256 * MiB + 1, 288 * GiB));
  CHECK(!r.possible(SIZE_MAX, 288 * GiB));
}

static void test_lut_text(const std::string& root) {
  for (int cp3 = 0; cp3 < 2; ++cp3) {
    const std::string folder = root + (cp3 ? "/3_coding_pass/" : "/2_coding_pass/");
    std::vector<int32_t> a(1 << 16), b(1 << 16);
    int len = 0, nbp = 0, len2 = 0, nbp2 = 0;
    CHECK(bpc_lut_parse_text(folder.c_str(), cp3, 5, a.data(), (int)a.size(), &len, &nbp) == BPC_OK);
    CHECK(orc_lut_load_text(folder.c_str(), cp3, 5, b.data(), (int)b.size(), &nbp2, &len2) == ORC_OK);
    CHECK(len == len2 && nbp == nbp2 && nbp == 15);
    CHECK(memcmp(a.data(), b.data(), sizeof(int32_t) * (size_t)len) == 0);
    // too small a capacity is refused, not overrun
    std::vector<int32_t> small(100);
    CHECK(bpc_lut_parse_text(folder.c_str(), cp3, 5, small.data(), (int)small.size(), &len, &nbp) == BPC_ERR_ARG);
  }
  std::vector<int32_t> a(64);
  int len = 0, nbp = 0;
  CHECK(bpc_lut_parse_text((root + "/no_such_folder/").c_str(), 0, 5, a.data(), 64, &len, &nbp) == BPC_ERR_ARG);
}

static void test_trainer_and_oracle() {
  const int X = 512, Y = 256;
  for (int cp3 = 0; cp3 < 2; ++cp3)
    for (int rows : {64, 32, 2}) {
      // 2-row codeblocks of 12-bit data need more than the default 128-word regions
      const bpc_params bp = bparams(cp3, 3, rows, 15, rows == 2 ? 512 : 0);
      const orc_params op = oparams(bp);
      const std::vector<int32_t> img = synth(X, Y, 3, 12);
      const int n = bpc_lut_len(&bp);
      std::vector<int32_t> lut(n);
      CHECK(bpc_lut_train(img.data(), X, Y, &bp, lut.data(), n, 3) == BPC_OK);
      // the same table from the oracle's symbol statistics
      std::vector<uint32_t> n0(n), nn(n);
      CHECK(orc_symbol_counts(img.data(), X, Y, &op, n0.data(), nn.data(), n, 1) == ORC_OK);
      bool same = true;
      for (int k = 0; k < n; ++k) {
        int v = 64;
        if (nn[k]) v = std::min(127, std::max(1, (int)(((uint64_t)n0[k] * 128 + nn[k] / 2) / nn[k])));
        same = same && v == lut[k];
      }
      CHECK(same);
      // round trip through the oracle, with the trace on one codeblock
      const int ncb = (X / 64) * (Y / rows);
      const int stride = bpc::stride_of(&bp);
      std::vector<int32_t> cs((size_t)ncb * stride), ncw(ncb), rec((size_t)X * Y, 7);
      CHECK(orc_encode(img.data(), X, Y, lut.data(), n, &op, cs.data(), ncw.data(), 1) == ORC_OK);
      CHECK(orc_decode(cs.data(), X, Y, lut.data(), n, &op, rec.data(), 1) == ORC_OK);
      CHECK(rec == img);
      std::vector<orc_trace_rec> tr(1 << 16);
      std::vector<int32_t> region(stride, -1);
      int32_t one = 0;
      int nt = 0;
      CHECK(orc_encode_cb_trace(img.data(), X, Y, ncb - 1, lut.data(), n, &op, region.data(), &one, tr.data(),
                                (int)tr.size(), &nt) == ORC_OK);
      CHECK(one == ncw[ncb - 1] && nt > 0);
      CHECK(memcmp(region.data(), cs.data() + (size_t)(ncb - 1) * stride, sizeof(int32_t) * stride) == 0);
      // a trace buffer that is too small is counted, not overrun
      int nt2 = 0;
      CHECK(orc_encode_cb_trace(img.data(), X, Y, ncb - 1, lut.data(), n, &op, region.data(), &one, tr.data(), 3,
                                &nt2) == ORC_OK);
      CHECK(nt2 == nt);
      // malformed MSB words are reported, nothing is read past the region
      std::vector<int32_t> bad = cs;
      bad[0] = 40;
      CHECK(orc_decode(bad.data(), X, Y, lut.data(), n, &op, rec.data(), 1) == ORC_ERR_MALFORMED);
      // regions too small for the codewords are reported, nothing is written past them
      orc_params tiny = op;
      tiny.cb_stride_words = 8;
      std::vector<int32_t> cs_tiny((size_t)ncb * 8);
      CHECK(orc_encode(img.data(), X, Y, lut.data(), n, &tiny, cs_tiny.data(), nullptr, 1) == ORC_ERR_OVERFLOW);
    }
  // magnitudes of 2^28 and more are refused by both
  std::vector<int32_t> big((size_t)64 * 64, 0);
  big[5] = 1 << 28;
  const bpc_params bp = bparams(1, 0, 64, 28, 16384);  // 28 planes of 4096 symbols at one bit each
  const orc_params op = oparams(bp);
  std::vector<int32_t> lut(bpc_lut_len(&bp), 64), cs(16384);
  CHECK(orc_encode(big.data(), 64, 64, lut.data(), (int)lut.size(), &op, cs.data(), nullptr, 1) == ORC_ERR_RANGE);
  big[5] = -(1 << 27) - 5;
  CHECK(orc_encode(big.data(), 64, 64, lut.data(), (int)lut.size(), &op, cs.data(), nullptr, 1) == ORC_OK);
}

static void test_files(const std::string& tmp) {
  const int geo[][4] = {{2048, 2048, 5, 64}, {1024, 512, 3, 32}, {3072, 3072, 5, 64}, {256, 256, 2, 16}};
  for (const auto& g : geo) {
    const int ncb = (g[0] / 64) * (g[1] / g[3]);
    std::vector<int32_t> a(ncb), b(ncb);
    const int na = bpc_cb_file_order(g[0], g[1], g[2], g[3], a.data(), ncb);
    const int nb = orc_cb_file_order(g[0], g[1], g[2], g[3], b.data(), ncb);
    CHECK(na == nb);
    if (na > 0) CHECK(a == b);
    if (na > 0) CHECK(bpc_cb_file_order(g[0], g[1], g[2], g[3], a.data(), ncb - 1) == BPC_ERR_ARG);
  }
  // .bin export of a compact codestream, import back into regions, against the
  // oracle's load_CStream of the same bytes
  const int X = 256, Y = 256;
  bpc_params p = bparams(0, 2, 64, 15);
  const int ncb = 16, stride = 64 * 64;
  std::vector<uint8_t> msb(ncb);
  std::vector<uint64_t> off(ncb + 1, 0);
  std::vector<uint16_t> payload;
  for (int c = 0; c < ncb; ++c) {
    msb[c] = (uint8_t)(c % 13);
    const int n = (int)(rnd() % 200);
    for (int k = 0; k < n; ++k) payload.push_back((uint16_t)rnd());
    off[c + 1] = payload.size();
  }
  if (payload.empty()) payload.push_back(0);
  CHECK(bpc_bin_export(tmp.c_str(), msb.data(), off.data(), payload.data(), X, Y, &p) == BPC_OK);
  std::vector<int32_t> cs((size_t)ncb * stride, 0);
  CHECK(bpc_bin_import(tmp.c_str(), X, Y, &p, cs.data()) == BPC_OK);
  std::vector<int32_t> order(ncb);
  CHECK(bpc_cb_file_order(X, Y, 2, 64, order.data(), ncb) == ncb);
  for (int k = 0; k < ncb; ++k) {
    const std::string path = tmp + "/" + std::to_string(k) + ".bin";
    FILE* f = fopen(path.c_str(), "rb");
    CHECK(f != nullptr);
    if (!f) continue;
    std::vector<uint8_t> bytes(2 * stride + 8);
    const size_t got = fread(bytes.data(), 1, bytes.size(), f);
    fclose(f);
    std::vector<int32_t> want(stride);
    orc_load_cstream(bytes.data(), got, want.data(), stride);
    CHECK(memcmp(want.data(), cs.data() + (size_t)order[k] * stride, sizeof(int32_t) * stride) == 0);
    const int t = order[k];
    CHECK(want[0] == msb[t]);
    CHECK(got == 2 * (1 + (off[t + 1] - off[t])));
  }
  // offsets that do not fit a region are refused
  off[1] = off[0] + stride;
  CHECK(bpc_bin_export(tmp.c_str(), msb.data(), off.data(), payload.data(), X, Y, &p) == BPC_ERR_ARG);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <lut folder root> <tmp dir>\n", argv[0]);
    return 2;
  }
  test_checks();
  test_queue_entries();
  test_admission();
  test_lut_text(argv[1]);
  test_trainer_and_oracle();
  test_files(argv[2]);
  printf("host selftest: %d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
