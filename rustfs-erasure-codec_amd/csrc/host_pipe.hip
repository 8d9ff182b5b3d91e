// host_pipe.hip -- the host-pointer entry points (shards in caller memory, the
// reference's normal case): a chunked pipeline over the shard columns.
//
// Encode, verify and reconstruct act on every byte column independently (64-byte
// blocks for Leopard GF(2^16)), so a call is cut into column chunks (an eighth
// of the shard, 256 KiB to 4 MiB).  Each chunk goes through one of two staging
// slots:
// worker threads copy the caller's bytes into pinned memory, the slot's stream
// copies them to the device, runs the operation on device pointers and copies
// the outputs back, and the workers copy those to the caller.  While one slot's
// chunk is on the device, the host fills the other, so host copies, both PCIe
// directions and the kernels overlap.  Staging (pinned and device) belongs to
// the calling thread and is reused across calls.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"

namespace rsgpu {
namespace {

// Bytes per shard per chunk: about 8 chunks per call for pipelining, at least
// 256 KiB (per-chunk call overheads) and at most 4 MiB (staging size); a
// multiple of 64 KiB, so chunk boundaries keep Leopard's 64-byte blocks whole.
size_t chunk_for(size_t len) {
  size_t c = (len / 8 + (64 << 10) - 1) / (64 << 10) * (64 << 10);
  return std::min<size_t>(std::max<size_t>(c, 256 << 10), 4 << 20);
}

// Persistent workers for the copies between caller memory and pinned staging.
// run(n, fn) calls fn(0..n-1) on the workers and the calling thread and returns
// when all are done; calls from several threads queue their jobs together.  A
// task is claimed in the same critical section that finds its job, and the job
// (on its caller's stack) outlives every claimed task: run() returns only when
// all n have finished.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool pool;
    return pool;
  }
  void run(size_t n, const std::function<void(size_t)>& fn) {
    if (n == 0) return;
    Job job{&fn, n};
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_.push_back(&job);
    }
    cv_.notify_all();
    std::unique_lock<std::mutex> g(mu_);
    while (true) {  // the caller takes tasks of its own job too
      Job* j = nullptr;
      size_t i = 0;
      if (!claim(&job, j, i)) break;
      g.unlock();
      (*j->fn)(i);
      g.lock();
      finish(j);
    }
    done_cv_.wait(g, [&] { return job.finished == job.n; });
  }

 private:
  struct Job {
    const std::function<void(size_t)>* fn;
    size_t n, next = 0, finished = 0;
  };
  CopyPool() {
    unsigned hw = std::thread::hardware_concurrency();
    const unsigned workers = std::max(1u, std::min(8u, hw ? hw - 1 : 1u));
    for (unsigned w = 0; w < workers; ++w)
      threads_.emplace_back([this] {
        std::unique_lock<std::mutex> g(mu_);
        while (true) {
          cv_.wait(g, [&] { return stop_ || !jobs_.empty(); });
          if (stop_) return;
          Job* j = nullptr;
          size_t i = 0;
          if (!claim(nullptr, j, i)) continue;
          g.unlock();
          (*j->fn)(i);
          g.lock();
          finish(j);
        }
      });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // With mu_ held: the next task of `only` (or of the oldest queued job); a
  // job leaves the queue when its last task is handed out.
  bool claim(Job* only, Job*& j, size_t& i) {
    j = only ? only : (jobs_.empty() ? nullptr : jobs_.front());
    if (!j || j->next >= j->n) return false;
    i = j->next++;
    if (j->next == j->n)
      for (auto it = jobs_.begin(); it != jobs_.end(); ++it)
        if (*it == j) {
          jobs_.erase(it);
          break;
        }
    return true;
  }
  void finish(Job* j) {  // with mu_ held
    if (++j->finished == j->n) done_cv_.notify_all();
  }
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<Job*> jobs_;
  std::vector<std::thread> threads_;
  bool stop_ = false;
};

// Copies w bytes from src(r) to dst(r) for every row r on the pool, in pieces of
// at most 256 KiB.
void par_copy(size_t rows, size_t w, const std::function<const uint8_t*(size_t)>& src,
              const std::function<uint8_t*(size_t)>& dst) {
  const size_t piece = 256 << 10, per_row = (w + piece - 1) / piece;
  CopyPool::get().run(rows * per_row, [&](size_t t) {
    const size_t r = t / per_row, off = (t % per_row) * piece;
    std::memcpy(dst(r) + off, src(r) + off, std::min(piece, w - off));
  });
}

struct HostPipe {
  int device = -1;
  size_t cap = 0;           // bytes per slot
  size_t pitch = 0;         // this call's bytes per shard row (its chunk size)
  uint8_t* pinned[2] = {};  // [shards][pitch]
  DeviceBuffer dev[2];      // [shards][pitch]
  DeviceBuffer flags;       // verify: one flag per chunk
  int32_t* h_flags = nullptr;
  size_t n_flags = 0;
  hipStream_t stream[2] = {};
  hipEvent_t ev[2] = {};
  ~HostPipe() { release(); }
  void release() {
    for (int i = 0; i < 2; ++i) {
      if (stream[i]) (void)hipStreamSynchronize(stream[i]);
      if (pinned[i]) (void)hipHostFree(pinned[i]);
      if (dev[i].p) (void)hipFree(dev[i].p);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (stream[i]) (void)hipStreamDestroy(stream[i]);
      pinned[i] = nullptr, dev[i].p = nullptr, dev[i].bytes = 0, ev[i] = nullptr, stream[i] = nullptr;
    }
    if (h_flags) (void)hipHostFree(h_flags);
    if (flags.p) (void)hipFree(flags.p);
    h_flags = nullptr, flags.p = nullptr, flags.bytes = 0;
    cap = pitch = n_flags = 0;
    device = -1;
  }
  int ensure(size_t n, size_t chunk, size_t chunks) {
    int d = 0;
    HIP_TRY(hipGetDevice(&d));
    if (d != device) {  // staging is per device: start over on another one
      release();
      for (int i = 0; i < 2; ++i) {
        HIP_TRY(hipStreamCreateWithFlags(&stream[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
      }
      device = d;
    }
    if (n * chunk > cap) {
      for (int i = 0; i < 2; ++i) {
        HIP_TRY(hipStreamSynchronize(stream[i]));
        if (pinned[i]) (void)hipHostFree(pinned[i]);
        pinned[i] = nullptr;
        HIP_TRY(hipHostMalloc((void**)&pinned[i], n * chunk, hipHostMallocDefault));
        HIP_TRY(dev[i].alloc(n * chunk));
      }
      cap = n * chunk;
    }
    pitch = chunk;
    if (chunks > n_flags) {
      for (int i = 0; i < 2; ++i) HIP_TRY(hipStreamSynchronize(stream[i]));
      if (h_flags) (void)hipHostFree(h_flags);
      h_flags = nullptr;
      HIP_TRY(hipHostMalloc((void**)&h_flags, chunks * 4, hipHostMallocDefault));
      HIP_TRY(flags.alloc(chunks * 4));
      n_flags = chunks;
    }
    return RSGPU_OK;
  }
};

thread_local HostPipe t_pipe;

// Runs `op` over the column chunks of a call.  For chunk c in slot s:
// in_rows are copied in (caller -> pinned -> device), op(c, s, width) enqueues
// the device work on the slot's stream, out_rows are copied back after it.
enum class HipDir { kIn, kOut };

struct ChunkPlan {
  size_t len, chunk;
  std::vector<size_t> in_rows, out_rows;  // shard indices (slot row = shard index)
  std::function<const uint8_t*(size_t)> host_in;
  std::function<uint8_t*(size_t)> host_out;
  std::function<int(size_t chunk, int slot, size_t off, size_t w)> op;
  std::function<void()> after_first;  // once the first chunk's op is enqueued
};

// One transfer per run of consecutive rows (rows a..b: (b - a) * pitch + w bytes).
int copy_rows(HipDir dir, HostPipe& P, int s, const std::vector<size_t>& rows, size_t w) {
  for (size_t i = 0; i < rows.size();) {
    size_t j = i;
    while (j + 1 < rows.size() && rows[j + 1] == rows[j] + 1) ++j;
    const size_t off = rows[i] * P.pitch, bytes = (rows[j] - rows[i]) * P.pitch + w;
    uint8_t* d = (uint8_t*)P.dev[s].p + off;
    uint8_t* h = P.pinned[s] + off;
    if (dir == HipDir::kIn)
      HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, P.stream[s]));
    else
      HIP_TRY(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, P.stream[s]));
    i = j + 1;
  }
  return RSGPU_OK;
}

int run_chunks(HostPipe& P, ChunkPlan& cp) {
  const size_t chunks = (cp.len + cp.chunk - 1) / cp.chunk;
  bool pending[2] = {false, false};
  size_t pend_off[2] = {0, 0}, pend_w[2] = {0, 0};
  auto drain = [&](int s) -> int {  // outputs of the slot's last chunk -> caller
    if (!pending[s]) return RSGPU_OK;
    HIP_TRY(hipEventSynchronize(P.ev[s]));
    const size_t off = pend_off[s], w = pend_w[s];
    par_copy(cp.out_rows.size(), w, [&](size_t r) { return P.pinned[s] + cp.out_rows[r] * P.pitch; },
             [&](size_t r) { return cp.host_out(cp.out_rows[r]) + off; });
    pending[s] = false;
    return RSGPU_OK;
  };
  // one chunk; an error returns from here only, so both slots are still finished below
  auto step = [&](size_t c) -> int {
    const int s = (int)(c & 1);
    const size_t off = c * cp.chunk, w = std::min(cp.chunk, cp.len - off);
    int r = drain(s);
    if (r) return r;
    HIP_TRY(hipEventSynchronize(P.ev[s]));  // the slot's previous transfers are done
    par_copy(cp.in_rows.size(), w, [&](size_t i) { return cp.host_in(cp.in_rows[i]) + off; },
             [&](size_t i) { return P.pinned[s] + cp.in_rows[i] * P.pitch; });
    if ((r = copy_rows(HipDir::kIn, P, s, cp.in_rows, w))) return r;
    if ((r = cp.op(c, s, off, w))) return r;
    if (c == 0 && cp.after_first) cp.after_first();
    if ((r = copy_rows(HipDir::kOut, P, s, cp.out_rows, w))) return r;
    HIP_TRY(hipEventRecord(P.ev[s], P.stream[s]));
    pending[s] = !cp.out_rows.empty();
    pend_off[s] = off;
    pend_w[s] = w;
    return RSGPU_OK;
  };
  int rc = RSGPU_OK;
  for (size_t c = 0; c < chunks && !rc; ++c) rc = step(c);
  // finish both slots even after an error, so no transfer into staging is left
  for (int s = 0; s < 2; ++s) {
    (void)hipStreamSynchronize(P.stream[s]);
    if (!rc) rc = drain(s);
  }
  return rc;
}

uint8_t* slot_dev(HostPipe& P, int s, size_t r) { return (uint8_t*)P.dev[s].p + r * P.pitch; }

// The length rules of the device calls, checked before any chunk runs (only the
// last chunk can break them: chunk sizes are multiples of 64).
int check_len(const Codec& c, size_t len) {
  if (len == 0) return RSGPU_EMPTY_SHARD;
  if (c.family != RSGPU_FAMILY_CLASSIC && (len % 64)) return RSGPU_INCORRECT_SHARD_SIZE;
  if (c.family == RSGPU_FAMILY_CLASSIC && c.field == RSGPU_FIELD_GF16 && (len & 1))
    return RSGPU_INCORRECT_SHARD_SIZE;
  return RSGPU_OK;
}

}  // namespace

void host_parallel_for(size_t n, const std::function<void(size_t)>& fn) { CopyPool::get().run(n, fn); }

}  // namespace rsgpu

using namespace rsgpu;

extern "C" {

int rsgpu_encode_host(const rsgpu_codec* h, uint8_t* const* shards, size_t n, size_t len) {
  if (!h || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (n < c.k + c.m) return RSGPU_TOO_FEW_SHARDS;
  if (n > c.k + c.m) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_len(c, len);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i)
    if (!shards[i]) return RSGPU_ERR_INVALID_ARGUMENT;
  HostPipe& P = t_pipe;
  ChunkPlan cp;
  cp.len = len;
  cp.chunk = chunk_for(len);
  if ((rc = P.ensure(n, cp.chunk, 1))) return rc;
  for (size_t i = 0; i < c.k; ++i) cp.in_rows.push_back(i);
  for (size_t i = c.k; i < n; ++i) cp.out_rows.push_back(i);
  cp.host_in = [&](size_t i) { return (const uint8_t*)shards[i]; };
  cp.host_out = [&](size_t i) { return shards[i]; };
  std::vector<rsgpu_slice> sl(n);
  cp.op = [&](size_t, int s, size_t, size_t w) {
    for (size_t i = 0; i < n; ++i) sl[i] = {slot_dev(P, s, i), w};
    return rsgpu_encode(h, sl.data(), n, P.stream[s]);
  };
  return run_chunks(P, cp);
}

int rsgpu_verify_host(const rsgpu_codec* h, const uint8_t* const* shards, size_t n, size_t len,
                      int32_t* ok) {
  if (!h || !shards || !ok) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (n < c.k + c.m) return RSGPU_TOO_FEW_SHARDS;
  if (n > c.k + c.m) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_len(c, len);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i)
    if (!shards[i]) return RSGPU_ERR_INVALID_ARGUMENT;
  HostPipe& P = t_pipe;
  ChunkPlan cp;
  cp.len = len;
  cp.chunk = chunk_for(len);
  const size_t chunks = (len + cp.chunk - 1) / cp.chunk;
  if ((rc = P.ensure(n, cp.chunk, chunks))) return rc;
  for (size_t i = 0; i < n; ++i) cp.in_rows.push_back(i);
  cp.host_in = [&](size_t i) { return shards[i]; };
  std::vector<rsgpu_cslice> sl(n);
  cp.op = [&](size_t chunk, int s, size_t, size_t w) {
    for (size_t i = 0; i < n; ++i) sl[i] = {slot_dev(P, s, i), w};
    return rsgpu_verify_async(h, sl.data(), n, (int32_t*)P.flags.p + chunk, P.stream[s]);
  };
  if ((rc = run_chunks(P, cp))) return rc;
  HIP_TRY(hipMemcpy(P.h_flags, P.flags.p, chunks * 4, hipMemcpyDeviceToHost));
  int32_t bad = 0;
  for (size_t i = 0; i < chunks; ++i) bad |= P.h_flags[i];
  *ok = bad == 0 ? 1 : 0;
  return RSGPU_OK;
}

int rsgpu_reconstruct_host(const rsgpu_codec* h, uint8_t* const* shards, uint8_t* present,
                           size_t n, size_t len, int32_t data_only) {
  if (!h || !shards || !present) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (n < c.k + c.m) return RSGPU_TOO_FEW_SHARDS;
  if (n > c.k + c.m) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_len(c, len);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i)
    if (!shards[i]) return RSGPU_ERR_INVALID_ARGUMENT;
  HostPipe& P = t_pipe;
  ChunkPlan cp;
  cp.len = len;
  cp.chunk = chunk_for(len);
  if ((rc = P.ensure(n, cp.chunk, 1))) return rc;
  for (size_t i = 0; i < n; ++i)
    if (present[i]) cp.in_rows.push_back(i);
  cp.host_in = [&](size_t i) { return (const uint8_t*)shards[i]; };
  cp.host_out = [&](size_t i) { return shards[i]; };
  std::vector<rsgpu_slot> sl(n);
  cp.op = [&](size_t, int s, size_t, size_t w) {
    for (size_t i = 0; i < n; ++i) sl[i] = {slot_dev(P, s, i), w, present[i] ? 1 : 0};
    return data_only ? rsgpu_reconstruct_data(h, sl.data(), n, P.stream[s])
                     : rsgpu_reconstruct(h, sl.data(), n, P.stream[s]);
  };
  // the shards the first chunk rebuilt are the outputs of every chunk
  cp.after_first = [&] {
    for (size_t i = 0; i < n; ++i)
      if (!present[i] && sl[i].present) cp.out_rows.push_back(i);
  };
  if ((rc = run_chunks(P, cp))) return rc;
  for (size_t r : cp.out_rows) present[r] = 1;
  return RSGPU_OK;
}

}  // extern "C"
