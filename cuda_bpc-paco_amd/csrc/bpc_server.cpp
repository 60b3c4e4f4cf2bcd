// Continuous-batching request server over the batched codec.
//
// Requests (encode or decode of one image, device-resident inputs) are queued on
// arrival.  A scheduler thread admits them in arrival order while the device's
// free memory (less a reserve) and the optional output budget allow their output
// buffers, groups admitted requests of one direction into a batch and launches it
// on one of several streams.  Up to `max_inflight` batches
// run on the device at once, so requests that arrive while others are on the GPU
// start as soon as a slot and memory free up, and their work overlaps on the
// device.  Each request's output is the same as coding its image alone.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bpc_paco.h"
#include "bpc_host.h"
#include "bpc_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t id = 0;
  bpc_request r;
  size_t out_bytes = 0;
  void* out = nullptr;  // server-allocated output (encode: codestream, decode: coefficients)
  int32_t* ncw = nullptr;
  Clock::time_point t_submit, t_start, t_done;
  int status = BPC_OK;
  bool admitted = false;  // its output bytes count as committed
  bool done = false;
};

// The status word a batch's last kernel has not written yet (no ST_* combination).
constexpr uint32_t kStatusUnset = 0x40000000u;

struct Slot {
  bpc_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  // the batch's status bits, stored by its last kernel into pinned host memory: read
  // once the event has completed, with no copy on the device
  volatile uint32_t* h_status = nullptr;
  uint32_t* h_status_dev = nullptr;  // its device address
  std::vector<Req*> reqs;
  bool busy = false;
};

int64_t ncb_of(const bpc_request& r) {
  return (int64_t)(r.size_x / 64) * (r.size_y / std::max(1, r.params.cblock_length));
}

}  // namespace

struct bpc_server {
  bpc_server_config cfg;
  int device = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work;   // scheduler wake-up
  std::condition_variable cv_done;   // completion notifications
  std::deque<Req*> pending;
  std::map<uint64_t, Req*> all;
  uint64_t next_id = 1;
  size_t committed = 0;  // bytes of outputs held by admitted, not yet released requests
  bpc::AdmissionRule rule;
  size_t total_mem = 0;  // device memory (mem probe at create)
  // released output buffers kept for reuse, by size (hipFree would synchronize the device)
  std::multimap<size_t, void*> pool;
  size_t pooled = 0;
  std::atomic<bool> stop{false};
  uint64_t changes = 0;  // submits and releases (what can make a pending request admissible)
  std::thread th;
  // statistics
  uint64_t batches = 0, completed = 0;
  double samples_done = 0;
  size_t max_batch_images_seen = 0;
};

namespace {

void finish_slot(bpc_server* s, Slot& sl, int st) {
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  for (Req* q : sl.reqs) {
    q->t_done = now;
    q->status = st;
    q->done = true;
    s->completed++;
    s->samples_done += (double)q->r.size_x * q->r.size_y;
  }
  sl.reqs.clear();
  sl.busy = false;
  s->cv_done.notify_all();
}

// The slot's event has completed: either the batch is done, or its main pass
// deferred pairs to the 28-plane pass, which is launched now (the slot stays busy).
void complete_slot(bpc_server* s, Slot& sl) {
  std::atomic_thread_fence(std::memory_order_acquire);
  const uint32_t bits = *sl.h_status;
  if (bits == kStatusUnset) {
    finish_slot(s, sl, bpc::fail(BPC_ERR_HIP, "server: the batch's status word was not written"));
    return;
  }
  if (bits & bpc::ST_WIDE_PENDING) {
    *sl.h_status = kStatusUnset;
    int st = bpc::batch_run_wide(sl.batch, sl.stream);
    if (!st && hipEventRecord(sl.ev, sl.stream) != hipSuccess) st = bpc::fail(BPC_ERR_HIP, "server: event record");
    if (!st) return;
    finish_slot(s, sl, st);
    return;
  }
  finish_slot(s, sl, bpc::status_to_code(bits));
}

size_t ncw_bytes(const bpc_request& r) { return sizeof(int32_t) * (size_t)std::max<int64_t>(1, ncb_of(r)); }

// Device free memory as the admission rule sees it (the injected probe, or HIP).
size_t free_memory(bpc_server* s) {
  size_t fr = 0, tot = 0;
  const int rc = s->cfg.mem_info ? s->cfg.mem_info(s->cfg.mem_info_ctx, &fr, &tot) : (int)hipMemGetInfo(&fr, &tot);
  return rc == 0 ? fr : 0;
}

// Takes a pooled buffer of exactly `bytes`, or null (caller holds s->mu).
void* pool_take(bpc_server* s, size_t bytes) {
  auto it = s->pool.find(bytes);
  if (it == s->pool.end()) return nullptr;
  void* p = it->second;
  s->pool.erase(it);
  s->pooled -= bytes;
  return p;
}

// Bytes of a request's buffers that the pool cannot serve (caller holds s->mu).
size_t new_bytes_of(bpc_server* s, const Req* q) {
  const size_t ob = q->out_bytes, cb = q->r.direction == BPC_CODE ? ncw_bytes(q->r) : 0;
  const size_t n_out = s->pool.count(ob);
  size_t need = n_out ? 0 : ob;
  if (cb && s->pool.count(cb) < 1 + (cb == ob && n_out ? 1u : 0u)) need += cb;
  return need;
}

// Frees pooled buffers, largest first, until `bytes` are released (caller holds s->mu;
// hipFree waits for the device, which only happens under memory pressure).
void trim_pool(bpc_server* s, size_t bytes) {
  size_t freed = 0;
  while (freed < bytes && !s->pool.empty()) {
    auto it = std::prev(s->pool.end());
    freed += it->first;
    s->pooled -= it->first;
    (void)hipFree(it->second);
    s->pool.erase(it);
  }
}

// Takes admissible requests of one direction off the queue (arrival order); the
// first pending request fixes the direction and the codeblock length.  Admission
// stops at the first request of that kind that does not fit (device free memory
// less the reserve, the output budget, the batch's codeblocks), so that requests
// are admitted in the order they arrived.  An admitted request claims pooled
// buffers of its sizes at once (caller holds s->mu).
std::vector<Req*> take_batch(bpc_server* s) {
  std::vector<Req*> out;
  if (s->pending.empty()) return out;
  const int dir = s->pending.front()->r.direction;
  const int rows = s->pending.front()->r.params.cblock_length;
  int64_t cbs = 0;
  size_t free_now = 0;
  bool probed = false;
  for (auto it = s->pending.begin(); it != s->pending.end() && (int)out.size() < s->cfg.max_batch_images;) {
    Req* q = *it;
    if (q->r.direction != dir || q->r.params.cblock_length != rows) {
      ++it;
      continue;
    }
    const int64_t n = ncb_of(q->r);
    if (cbs + n > s->cfg.max_batch_codeblocks) break;
    size_t nb = new_bytes_of(s, q);
    if (nb && !probed) {
      free_now = free_memory(s);
      probed = true;
    }
    bpc::Admit d = s->rule.decide(s->committed, q->out_bytes, nb, free_now, s->pooled);
    if (d == bpc::Admit::TRIM_POOL) {
      trim_pool(s, nb + s->rule.reserve - std::min(free_now, nb + s->rule.reserve));
      nb = new_bytes_of(s, q);
      free_now = free_memory(s);
      d = s->rule.decide(s->committed, q->out_bytes, nb, free_now, 0);
    }
    if (d != bpc::Admit::YES) break;  // admission in arrival order
    free_now -= std::min(free_now, nb);  // what this request will allocate
    q->out = pool_take(s, q->out_bytes);
    if (q->r.direction == BPC_CODE) q->ncw = (int32_t*)pool_take(s, ncw_bytes(q->r));
    cbs += n;
    s->committed += q->out_bytes;
    q->admitted = true;
    out.push_back(q);
    it = s->pending.erase(it);
  }
  return out;
}

// caller holds s->mu
void pool_put(bpc_server* s, void* p, size_t bytes) {
  if (!p) return;
  if (s->pooled + bytes > s->cfg.pool_limit_bytes) {
    (void)hipFree(p);
    return;
  }
  s->pool.emplace(bytes, p);
  s->pooled += bytes;
}

int launch_batch(bpc_server* s, Slot& sl, std::vector<Req*>& reqs) {
  std::vector<bpc_image> imgs(reqs.size());
  const auto now = Clock::now();
  for (size_t k = 0; k < reqs.size(); ++k) {
    Req* q = reqs[k];
    if (!q->out && hipMalloc(&q->out, q->out_bytes) != hipSuccess) {
      q->out = nullptr;
      return bpc::fail(BPC_ERR_HIP, "server: output allocation of %zu bytes", q->out_bytes);
    }
    if (q->r.direction == BPC_CODE && !q->ncw && hipMalloc(&q->ncw, ncw_bytes(q->r)) != hipSuccess) {
      q->ncw = nullptr;
      return bpc::fail(BPC_ERR_HIP, "server: count allocation");
    }
    q->t_start = now;
    bpc_image& im = imgs[k];
    im.src = q->r.src;
    im.dst = (int32_t*)q->out;
    im.lut = q->r.lut;
    im.lut_len = q->r.lut_len;
    im.ncw = q->r.direction == BPC_CODE ? q->ncw : nullptr;
    im.size_x = q->r.size_x;
    im.size_y = q->r.size_y;
    im.params = q->r.params;
    im.fill_unused = 1;
    im.int16_coeffs = q->r.int16_coeffs;
  }
  // Replayed graph (one per direction and bitplane count: the run's shape is read
  // on the device), status into the slot's pinned word, 28-plane pass only when
  // the main pass deferred pairs to it.
  *sl.h_status = kStatusUnset;
  int st = bpc::batch_run_ex(sl.batch, reqs.front()->r.direction, imgs.data(), (int)imgs.size(), 1, nullptr,
                             sl.h_status_dev, 1, sl.stream);
  if (st) return st;
  if (hipEventRecord(sl.ev, sl.stream) != hipSuccess) return BPC_ERR_HIP;
  sl.reqs = reqs;
  sl.busy = true;
  return BPC_OK;
}

void scheduler(bpc_server* s) {
  (void)hipSetDevice(s->device);
  while (!s->stop.load()) {
    bool progressed = false;
    for (Slot& sl : s->slots) {  // completions
      if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
        complete_slot(s, sl);
        progressed = true;
      }
    }
    for (Slot& sl : s->slots) {  // launches
      if (sl.busy) continue;
      std::vector<Req*> reqs;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        reqs = take_batch(s);
        if (!reqs.empty()) {
          s->batches++;
          s->max_batch_images_seen = std::max(s->max_batch_images_seen, reqs.size());
        }
      }
      if (reqs.empty()) break;
      int st = launch_batch(s, sl, reqs);
      if (st) {  // fail these requests
        std::lock_guard<std::mutex> lk(s->mu);
        for (Req* q : reqs) {
          q->status = st;
          q->done = true;
          q->t_done = Clock::now();
        }
        s->cv_done.notify_all();
      }
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> lk(s->mu);
      bool any_busy = false;
      for (Slot& sl : s->slots) any_busy |= sl.busy;
      if (any_busy) {
        s->cv_work.wait_for(lk, std::chrono::microseconds(20));  // completion polling
      } else {
        // nothing in flight: wait for a submit or a release; device memory freed by
        // others is seen at the next probe, at most 1 ms later
        const uint64_t seen = s->changes;
        s->cv_work.wait_for(lk, std::chrono::milliseconds(1), [&] { return s->stop.load() || s->changes != seen; });
      }
    }
  }
  for (Slot& sl : s->slots)
    while (sl.busy) {
      (void)hipEventSynchronize(sl.ev);
      complete_slot(s, sl);
    }
}

}  // namespace

extern "C" {

void bpc_server_default_config(bpc_server_config* c) {
  if (!c) return;
  c->max_inflight = 2;
  c->max_batch_images = 64;
  c->max_batch_codeblocks = 1 << 18;
  c->hbm_budget_bytes = 0;                 // no output budget: free device memory decides
  c->hbm_reserve_bytes = (size_t)4 << 30;  // left free for the batch workspaces and other users
  c->pool_limit_bytes = (size_t)16 << 30;
  c->waves_per_block = 4;
  c->vgpr_cap_waves = 8;  // fastest register budget of the batched kernel on MI355X
  c->mem_info = nullptr;
  c->mem_info_ctx = nullptr;
}

int bpc_server_create(bpc_server** out, const bpc_server_config* cfg) {
  if (!out) return bpc::fail(BPC_ERR_ARG, "null server");
  bpc_server_config c;
  bpc_server_default_config(&c);
  if (cfg) c = *cfg;
  if (c.max_inflight < 1 || c.max_inflight > 8 || c.max_batch_images < 1 ||
      c.max_batch_images > bpc::MAX_BATCH_IMAGES || c.max_batch_codeblocks < 1 ||
      c.max_batch_codeblocks > ((long long)1 << 32))
    return bpc::fail(BPC_ERR_ARG, "bad server configuration");
  bpc_server* s = new bpc_server();
  s->cfg = c;
  s->rule.budget = c.hbm_budget_bytes;
  s->rule.reserve = c.hbm_reserve_bytes;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return bpc::fail(BPC_ERR_NODEVICE, "no HIP device");
  }
  {
    size_t fr = 0, tot = 0;
    const int rc = c.mem_info ? c.mem_info(c.mem_info_ctx, &fr, &tot) : (int)hipMemGetInfo(&fr, &tot);
    if (rc != 0) {
      delete s;
      return bpc::fail(BPC_ERR_HIP, "server: device memory query failed");
    }
    s->total_mem = tot;
  }
  s->slots.resize(c.max_inflight);
  for (Slot& sl : s->slots) {
    int st = bpc_batch_create(&sl.batch, c.max_batch_images, c.max_batch_codeblocks);
    if (!st) st = bpc_batch_set_tuning(sl.batch, c.waves_per_block, 0);
    if (!st) st = bpc_batch_set_vgpr_cap(sl.batch, c.vgpr_cap_waves);
    if (!st && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) st = BPC_ERR_HIP;
    if (!st && hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) st = BPC_ERR_HIP;
    void* hs = nullptr;
    if (!st && hipHostMalloc(&hs, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) st = BPC_ERR_HIP;
    sl.h_status = static_cast<volatile uint32_t*>(hs);
    void* hd = nullptr;
    if (!st && hipHostGetDevicePointer(&hd, hs, 0) != hipSuccess) st = BPC_ERR_HIP;
    sl.h_status_dev = static_cast<uint32_t*>(hd);
    if (st) {
      s->stop = true;
      bpc_server_destroy(s);
      return st;
    }
  }
  s->th = std::thread(scheduler, s);
  *out = s;
  return BPC_OK;
}

// A request is checked whole here, so that a malformed one fails alone instead of
// failing the batch it would join or blocking the queue.
int bpc_server_submit(bpc_server* s, const bpc_request* r, uint64_t* id) {
  if (!s || !r || !id) return bpc::fail(BPC_ERR_ARG, "null argument");
  if (r->direction != BPC_CODE && r->direction != BPC_DECODE) return bpc::fail(BPC_ERR_ARG, "bad direction");
  const long long ncb = bpc::batch_image_codeblocks(r->size_x, r->size_y, &r->params);
  if (ncb < 0) return (int)ncb;
  if (ncb > s->cfg.max_batch_codeblocks)
    return bpc::fail(BPC_ERR_ARG, "request of %lld codeblocks exceeds the server's batch (%lld)", ncb,
                     s->cfg.max_batch_codeblocks);
  if (!r->src || !r->lut) return bpc::fail(BPC_ERR_ARG, "null buffer");
  if ((uintptr_t)r->src & 7) return bpc::fail(BPC_ERR_ARG, "input must be 8-byte aligned");
  if (r->lut_len < bpc_lut_len(&r->params))
    return bpc::fail(BPC_ERR_ARG, "lut_len %d < %d required", r->lut_len, bpc_lut_len(&r->params));
  const size_t out_bytes = r->direction == BPC_CODE ? (size_t)bpc_stream_words(r->size_x, r->size_y, &r->params) * 4
                                                    : (size_t)r->size_x * r->size_y * (r->int16_coeffs ? 2 : 4);
  if (!s->rule.possible(out_bytes, s->total_mem))
    return bpc::fail(BPC_ERR_ARG, "output of %zu bytes can never be admitted (budget %zu, reserve %zu, device %zu)",
                     out_bytes, s->rule.budget, s->rule.reserve, s->total_mem);
  Req* q = new Req();
  q->r = *r;
  q->out_bytes = out_bytes;
  q->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    s->all[q->id] = q;
    s->pending.push_back(q);
    s->changes++;
  }
  s->cv_work.notify_one();
  *id = q->id;
  return BPC_OK;
}

int bpc_server_wait(bpc_server* s, uint64_t id, int timeout_ms, bpc_result* res) {
  if (!s || !res) return bpc::fail(BPC_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return bpc::fail(BPC_ERR_ARG, "unknown request %llu", (unsigned long long)id);
  Req* q = it->second;
  if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 1 << 30 : timeout_ms),
                           [&] { return q->done; }))
    return bpc::fail(BPC_ERR_TIMEOUT, "request %llu not finished in %d ms", (unsigned long long)id, timeout_ms);
  res->status = q->status;
  res->output = q->out;
  res->ncw = q->ncw;
  res->queue_us = std::chrono::duration<double, std::micro>(q->t_start - q->t_submit).count();
  res->latency_us = std::chrono::duration<double, std::micro>(q->t_done - q->t_submit).count();
  return BPC_OK;
}

int bpc_server_release(bpc_server* s, uint64_t id) {
  if (!s) return bpc::fail(BPC_ERR_ARG, "null server");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end() || !it->second->done) return bpc::fail(BPC_ERR_ARG, "request %llu unknown or not finished", (unsigned long long)id);
    Req* q = it->second;
    s->all.erase(it);
    if (q->admitted) s->committed -= q->out_bytes;
    pool_put(s, q->out, q->out_bytes);
    if (q->ncw) pool_put(s, q->ncw, ncw_bytes(q->r));
    delete q;
    s->changes++;
  }
  s->cv_work.notify_one();
  return BPC_OK;
}

int bpc_server_stats(bpc_server* s, bpc_server_stats_t* st) {
  if (!s || !st) return BPC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  st->batches = s->batches;
  st->completed = s->completed;
  st->pending = s->pending.size();
  st->committed_bytes = s->committed;
  st->max_batch_images = s->max_batch_images_seen;
  return BPC_OK;
}

int bpc_server_destroy(bpc_server* s) {
  if (!s) return BPC_OK;
  s->stop = true;
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  for (auto& kv : s->all) {
    if (kv.second->out) (void)hipFree(kv.second->out);
    if (kv.second->ncw) (void)hipFree(kv.second->ncw);
    delete kv.second;
  }
  for (auto& kv : s->pool) (void)hipFree(kv.second);
  for (Slot& sl : s->slots) {
    if (sl.batch) bpc_batch_destroy(sl.batch);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.h_status) (void)hipHostFree(const_cast<uint32_t*>(sl.h_status));
  }
  delete s;
  return BPC_OK;
}

}  // extern "C"
