// serve.hip -- request loop with HBM admission and continuous batching.
//
// The reference has no serving layer: each ReedSolomon call runs to completion
// on the caller's thread (src/core/encode.rs:935, verify.rs:72,
// reconstruct.rs:731), and stream.rs pipelines one object at a time.  Here many
// callers' stripe operations share the device:
//
//   * admission: requests wait in FIFO order until their HBM footprint
//     ((k+m) * shard_len: inputs resident in place plus the outputs the server
//     allocates) fits under the budget and their outputs fit in the arena;
//   * continuous batching: whenever one of the batch slots (a stream with its
//     own batch-kernel batch) is idle, every admitted, not yet launched
//     request is packed into one batch and launched; batches on different slots
//     overlap on the device, so requests that arrive while others run are
//     admitted and start as soon as memory and a slot free up;
//   * completion: an event behind each batch (after a copy of its verify flags
//     to pinned host memory) is polled; finished requests report their times.
//
// Decode rows stay on the device in each codec's cache; outputs live in the
// server's arena until the caller releases them.
#include <chrono>
#include <deque>
#include <map>
#include <thread>

#include "internal.h"

namespace {

using namespace rsgpu;
using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (256-byte granules).
struct Arena {
  DeviceBuffer mem;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  bool init(size_t bytes) {
    cap = (bytes + 255) & ~(size_t)255;
    if (mem.alloc(cap) != hipSuccess) return false;
    free_.clear();
    free_[0] = cap;
    return true;
  }
  bool alloc(size_t n, size_t* off) {
    n = (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      size_t rest = it->second - n, o = it->first + n;
      free_.erase(it);
      if (rest) free_[o] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (std::max<size_t>(n, 1) + 255) & ~(size_t)255;
    auto it = free_.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Req {
  rsgpu_request r{};
  std::vector<const uint8_t*> shards;
  std::vector<uint8_t> erased;
  size_t k = 0, m = 0, footprint = 0, out_bytes = 0, arena_off = 0;
  std::vector<uint8_t*> outs;
  std::vector<rsgpu_slice> sl;
  std::vector<rsgpu_cslice> csl;
  std::vector<rsgpu_slot> slots;
  uint64_t arrival = 0, admit = 0, launch = 0, done = 0;
  int32_t status = RSGPU_OK, verify_ok = 0;
  int slot = -1, flag_idx = -1;
  bool has_arena = false, completed = false, admitted = false;
};

struct BatchSlot {
  hipStream_t stream = nullptr;
  rsgpu_batch* batch = nullptr;
  hipEvent_t ev = nullptr;
  DeviceBuffer d_flags;
  int32_t* h_flags = nullptr;
  std::vector<uint64_t> tags;
  bool busy = false;
  PresolveScratch presolve;  // job arrays of the decode solves ahead of this slot's batches
};

}  // namespace

struct rsgpu_server {
  rsgpu_server_config cfg{};
  Arena arena;
  std::vector<std::unique_ptr<BatchSlot>> slots;
  std::map<uint64_t, Req> reqs;   // by tag
  std::deque<uint64_t> queue;     // arrived, not admitted (FIFO)
  std::deque<uint64_t> pending;   // admitted, not launched
  std::deque<uint64_t> finished;  // completed, not yet polled
  size_t used = 0;                // admitted footprint
  Clock::time_point t0 = Clock::now();
  rsgpu_server_counters st{};
  uint64_t now() const {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
  }
  ~rsgpu_server() {
    for (auto& sp : slots) {
      BatchSlot& s = *sp;
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      if (s.batch) rsgpu_batch_free(s.batch);
      if (s.ev) (void)hipEventDestroy(s.ev);
      if (s.h_flags) (void)hipHostFree(s.h_flags);
      if (s.stream) (void)hipStreamDestroy(s.stream);
    }
  }
};

namespace {

int enqueue(rsgpu_server* S, const rsgpu_request& r, uint64_t arrival) {
  if (!r.codec || !r.shards || r.op < 0 || r.op > 2 || (r.op == 2 && !r.erased))
    return RSGPU_ERR_INVALID_ARGUMENT;
  if (S->reqs.count(r.tag)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = r.codec->c;
  Req q;
  q.r = r;
  q.k = c.k;
  q.m = c.m;
  q.shards.assign(r.shards, r.shards + c.k + c.m);
  if (r.op == 2) q.erased.assign(r.erased, r.erased + c.k + c.m);
  size_t n_out = r.op == 0 ? c.m : 0;
  if (r.op == 2)
    for (uint8_t e : q.erased) n_out += e ? 1 : 0;
  q.out_bytes = n_out * r.shard_len;
  q.footprint = (c.k + c.m) * r.shard_len;
  q.arrival = arrival;
  if (q.footprint > S->cfg.hbm_budget || q.out_bytes > S->arena.cap) return RSGPU_ERR_OUT_OF_MEMORY;
  S->reqs.emplace(r.tag, std::move(q));
  S->queue.push_back(r.tag);
  return RSGPU_OK;
}

void complete(rsgpu_server* S, Req& q, uint64_t t) {
  q.done = t;
  q.completed = true;
  S->finished.push_back(q.r.tag);
  S->st.completed++;
}

void release_req(rsgpu_server* S, Req& q) {
  if (q.has_arena) S->arena.release(q.arena_off, q.out_bytes);
  q.has_arena = false;
  if (q.admitted) S->used -= q.footprint;
  q.admitted = false;
}

// Admits arrived requests in FIFO order while footprint and arena allow.
void admit(rsgpu_server* S, uint64_t t) {
  while (!S->queue.empty()) {
    Req& q = S->reqs.at(S->queue.front());
    if (S->used + q.footprint > S->cfg.hbm_budget) break;
    size_t off = 0;
    if (q.out_bytes && !S->arena.alloc(q.out_bytes, &off)) break;
    q.has_arena = q.out_bytes != 0;
    q.arena_off = off;
    q.admitted = true;
    S->used += q.footprint;
    S->st.peak_footprint = std::max<uint64_t>(S->st.peak_footprint, S->used);
    q.admit = t;
    uint8_t* base = (uint8_t*)S->arena.mem.p + off;
    const size_t n = q.k + q.m, len = q.r.shard_len;
    if (q.r.op == 0) {
      for (size_t i = 0; i < q.k; ++i) q.sl.push_back({const_cast<uint8_t*>(q.shards[i]), len});
      for (size_t j = 0; j < q.m; ++j) {
        q.outs.push_back(base + j * len);
        q.sl.push_back({q.outs.back(), len});
      }
    } else if (q.r.op == 1) {
      for (size_t i = 0; i < n; ++i) q.csl.push_back({q.shards[i], len});
    } else {
      size_t o = 0;
      for (size_t i = 0; i < n; ++i) {
        if (q.erased[i]) {
          q.outs.push_back(base + (o++) * len);
          q.slots.push_back({q.outs.back(), len, 0});
        } else {
          q.slots.push_back({const_cast<uint8_t*>(q.shards[i]), len, 1});
        }
      }
    }
    S->pending.push_back(q.r.tag);
    S->queue.pop_front();
  }
}

// Packs every admitted request into free batch slots (one batch per slot).  A
// burst is shared evenly over the free slots, at least kMinShare requests each,
// so that the first batch launches once its share is built and the device works
// on it while the host builds the next; a few pending requests stay one batch.
constexpr size_t kMinShare = 64;

int launch(rsgpu_server* S, uint64_t t) {
  size_t free_slots = 0;
  for (auto& o : S->slots) free_slots += o->busy ? 0 : 1;
  for (size_t si = 0; si < S->slots.size(); ++si) {
    BatchSlot& B = *S->slots[si];
    if (S->pending.empty()) break;
    if (B.busy) continue;
    const size_t share = std::max((S->pending.size() + free_slots - 1) / free_slots, kMinShare);
    const size_t limit = std::min<size_t>(share, (size_t)S->cfg.max_batch_ops);
    --free_slots;
    int rc = rsgpu_batch_clear(B.batch);
    if (rc) return rc;
    B.tags.clear();
    int n_flags = 0;
    std::vector<uint64_t> taken;
    while (!S->pending.empty() && taken.size() < limit) {
      taken.push_back(S->pending.front());
      S->pending.pop_front();
    }
    // A lone request runs as its own per-stripe call on the slot's stream: no
    // descriptor upload and no batch kernel, so a lightly loaded server answers
    // at the single-call latency.
    if (taken.size() == 1) {
      Req& q = S->reqs.at(taken[0]);
      const rsgpu_codec* c = q.r.codec;
      if (q.r.op == 0) {
        rc = rsgpu_encode(c, q.sl.data(), q.sl.size(), B.stream);
      } else if (q.r.op == 1) {
        q.flag_idx = n_flags++;
        rc = rsgpu_verify_async(c, q.csl.data(), q.csl.size(), (int32_t*)B.d_flags.p + q.flag_idx, B.stream);
      } else {
        rc = rsgpu_reconstruct(c, q.slots.data(), q.slots.size(), B.stream);
      }
      q.launch = t;
      if (rc) {  // validation error: reported per request, never launched
        q.status = rc;
        complete(S, q, t);
        continue;
      }
      q.slot = (int)si;
      B.tags.push_back(q.r.tag);
      if (n_flags && hipMemcpyAsync(B.h_flags, B.d_flags.p, 4, hipMemcpyDeviceToHost, B.stream) != hipSuccess)
        return RSGPU_ERR_HIP;
      if (hipEventRecord(B.ev, B.stream) != hipSuccess) return RSGPU_ERR_HIP;
      B.busy = true;
      S->st.batches++;
      S->st.launched_ops += 1;
      uint64_t inflight = 0;
      for (auto& o : S->slots) inflight += o->busy ? 1 : 0;
      S->st.max_inflight_batches = std::max<uint64_t>(S->st.max_inflight_batches, inflight);
      continue;
    }
    // the new erasure patterns among them, solved together ahead of the adds
    std::vector<PresolveCall> calls;
    for (uint64_t tag : taken) {
      const Req& q = S->reqs.at(tag);
      if (q.r.op == 2) calls.push_back({&q.r.codec->c, q.slots.data(), q.slots.size()});
    }
    // best effort: on failure each add solves its own pattern (and reports its error)
    if (!calls.empty()) (void)presolve_decode(calls, B.stream, B.presolve);
    for (uint64_t tag : taken) {
      Req& q = S->reqs.at(tag);
      const rsgpu_codec* c = q.r.codec;
      if (q.r.op == 0) {
        rc = rsgpu_batch_add_encode(B.batch, c, q.sl.data(), q.sl.size());
      } else if (q.r.op == 1) {
        q.flag_idx = n_flags++;
        rc = rsgpu_batch_add_verify(B.batch, c, q.csl.data(), q.csl.size(),
                                    (int32_t*)B.d_flags.p + q.flag_idx);
      } else {
        rc = rsgpu_batch_add_reconstruct(B.batch, c, q.slots.data(), q.slots.size(), 0, B.stream);
      }
      q.launch = t;
      if (rc) {  // validation error: reported per request, never launched
        q.status = rc;
        complete(S, q, t);
        continue;
      }
      q.slot = (int)si;
      B.tags.push_back(q.r.tag);
    }
    if (B.tags.empty()) continue;
    if (n_flags && launch_fill_i32((int32_t*)B.d_flags.p, 0, n_flags, B.stream) != hipSuccess)
      return RSGPU_ERR_HIP;
    if ((rc = rsgpu_batch_run(B.batch, B.stream))) return rc;
    if (n_flags && hipMemcpyAsync(B.h_flags, B.d_flags.p, (size_t)n_flags * 4, hipMemcpyDeviceToHost,
                                  B.stream) != hipSuccess)
      return RSGPU_ERR_HIP;
    if (hipEventRecord(B.ev, B.stream) != hipSuccess) return RSGPU_ERR_HIP;
    B.busy = true;
    S->st.batches++;
    S->st.launched_ops += B.tags.size();
    uint64_t inflight = 0;
    for (auto& o : S->slots) inflight += o->busy ? 1 : 0;
    S->st.max_inflight_batches = std::max<uint64_t>(S->st.max_inflight_batches, inflight);
  }
  return RSGPU_OK;
}

int retire(rsgpu_server* S, uint64_t t) {
  for (auto& Bp : S->slots) {
    BatchSlot& B = *Bp;
    if (!B.busy) continue;
    hipError_t e = hipEventQuery(B.ev);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) {
      set_hip_error(e, "server batch");
      return RSGPU_ERR_HIP;
    }
    for (uint64_t tag : B.tags) {
      Req& q = S->reqs.at(tag);
      if (q.r.op == 1) q.verify_ok = B.h_flags[q.flag_idx] == 0 ? 1 : 0;
      complete(S, q, t);
    }
    B.tags.clear();
    B.busy = false;
  }
  return RSGPU_OK;
}

int pump(rsgpu_server* S, uint64_t t) {
  int rc = retire(S, t);
  if (rc) return rc;
  admit(S, t);
  return launch(S, t);
}

// With no batch in flight and nothing admitted, a request still at the head of
// the queue waits on memory that only a caller's release can free: outputs kept
// past completion hold the budget and the arena.  A replay never releases kept
// outputs, so it completes that request with RSGPU_ERR_OUT_OF_MEMORY instead of
// waiting forever.  Returns true when it did.
bool fail_blocked_head(rsgpu_server* S, uint64_t t) {
  // completions not yet polled may still release their outputs (keep_outputs = 0)
  if (S->queue.empty() || !S->pending.empty() || !S->finished.empty()) return false;
  for (auto& B : S->slots)
    if (B->busy) return false;
  Req& q = S->reqs.at(S->queue.front());
  S->queue.pop_front();
  q.status = RSGPU_ERR_OUT_OF_MEMORY;
  q.admit = q.launch = t;
  complete(S, q, t);
  return true;
}

}  // namespace

extern "C" {

int rsgpu_server_new(const rsgpu_server_config* cfg, rsgpu_server** out) {
  if (!cfg || !out || cfg->hbm_budget == 0 || cfg->arena_bytes == 0 || cfg->max_inflight_batches < 1 ||
      cfg->max_inflight_batches > 16 || cfg->max_batch_ops < 1)
    return RSGPU_ERR_INVALID_ARGUMENT;
  auto* S = new rsgpu_server();
  S->cfg = *cfg;
  if (!S->arena.init(cfg->arena_bytes)) {
    delete S;
    return RSGPU_ERR_OUT_OF_MEMORY;
  }
  for (int i = 0; i < cfg->max_inflight_batches; ++i) {
    S->slots.push_back(std::make_unique<BatchSlot>());
    BatchSlot& B = *S->slots.back();
    if (hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&B.ev, hipEventDisableTiming) != hipSuccess ||
        B.d_flags.alloc((size_t)cfg->max_batch_ops * 4) != hipSuccess ||
        hipHostMalloc((void**)&B.h_flags, (size_t)cfg->max_batch_ops * 4, hipHostMallocDefault) != hipSuccess ||
        rsgpu_batch_new(&B.batch) != RSGPU_OK ||
        // staging for a full batch up front: serving then allocates nothing
        batch_reserve(B.batch, (size_t)cfg->max_batch_ops * 2, (size_t)cfg->max_batch_ops * 24) ||
        B.presolve.ensure((size_t)cfg->max_batch_ops * sizeof(DecodeMultiJob) + 4096)) {
      delete S;
      return RSGPU_ERR_OUT_OF_MEMORY;
    }
  }
  *out = S;
  return RSGPU_OK;
}

void rsgpu_server_free(rsgpu_server* S) { delete S; }

int rsgpu_server_submit(rsgpu_server* S, const rsgpu_request* r) {
  if (!S || !r) return RSGPU_ERR_INVALID_ARGUMENT;
  return enqueue(S, *r, S->now());
}

int rsgpu_server_pump(rsgpu_server* S, uint32_t* in_flight) {
  if (!S) return RSGPU_ERR_INVALID_ARGUMENT;
  int rc = pump(S, S->now());
  if (in_flight) {
    uint32_t n = 0;
    for (auto& B : S->slots) n += B->busy ? 1 : 0;
    *in_flight = n;
  }
  return rc;
}

int rsgpu_server_poll(rsgpu_server* S, rsgpu_completion* out, size_t max, size_t* n) {
  if (!S || (!out && max) || !n) return RSGPU_ERR_INVALID_ARGUMENT;
  size_t i = 0;
  while (i < max && !S->finished.empty()) {
    Req& q = S->reqs.at(S->finished.front());
    S->finished.pop_front();
    out[i++] = {q.r.tag, q.status, q.verify_ok, q.arrival, q.admit, q.launch, q.done, q.footprint};
  }
  *n = i;
  return RSGPU_OK;
}

int rsgpu_server_outputs(const rsgpu_server* S, uint64_t tag, uint8_t** outs, size_t n) {
  if (!S || (!outs && n)) return RSGPU_ERR_INVALID_ARGUMENT;
  auto it = S->reqs.find(tag);
  if (it == S->reqs.end() || !it->second.completed || n != it->second.outs.size())
    return RSGPU_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < n; ++i) outs[i] = it->second.outs[i];
  return RSGPU_OK;
}

int rsgpu_server_release(rsgpu_server* S, uint64_t tag) {
  if (!S) return RSGPU_ERR_INVALID_ARGUMENT;
  auto it = S->reqs.find(tag);
  if (it == S->reqs.end() || !it->second.completed) return RSGPU_ERR_INVALID_ARGUMENT;
  release_req(S, it->second);
  S->reqs.erase(it);
  return RSGPU_OK;
}

int rsgpu_server_replay(rsgpu_server* S, const rsgpu_request* reqs, size_t n, int32_t keep_outputs,
                        rsgpu_completion* out) {
  if (!S || (!reqs && n) || (!out && n)) return RSGPU_ERR_INVALID_ARGUMENT;
  std::map<uint64_t, size_t> index;
  for (size_t i = 0; i < n; ++i) {
    if (i && reqs[i].arrival_ns < reqs[i - 1].arrival_ns) return RSGPU_ERR_INVALID_ARGUMENT;
    if (!index.emplace(reqs[i].tag, i).second) return RSGPU_ERR_INVALID_ARGUMENT;
  }
  const uint64_t base = S->now();
  size_t next = 0, done = 0;
  std::vector<rsgpu_completion> buf(256);
  while (done < n) {
    const uint64_t t = S->now();
    bool progress = false;
    while (next < n && base + reqs[next].arrival_ns <= t) {
      int rc = enqueue(S, reqs[next], base + reqs[next].arrival_ns);
      if (rc) return rc;
      ++next;
      progress = true;
    }
    int rc = pump(S, t);
    if (rc) return rc;
    progress |= fail_blocked_head(S, t);
    size_t got = 0;
    rsgpu_server_poll(S, buf.data(), buf.size(), &got);
    for (size_t i = 0; i < got; ++i) {
      rsgpu_completion c = buf[i];
      c.arrival_ns -= base;
      c.admit_ns -= base;
      c.launch_ns -= base;
      c.done_ns -= base;
      out[index.at(c.tag)] = c;
      if (!keep_outputs) rsgpu_server_release(S, c.tag);
      ++done;
      progress = true;
    }
    if (!progress) std::this_thread::yield();
  }
  return RSGPU_OK;
}

int rsgpu_server_stats(const rsgpu_server* S, rsgpu_server_counters* out) {
  if (!S || !out) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = S->st;
  out->footprint = S->used;
  return RSGPU_OK;
}

}  // extern "C"
