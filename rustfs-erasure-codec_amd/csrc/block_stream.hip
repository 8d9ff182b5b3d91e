// block_stream.hip -- block-wise encode / verify / reconstruct of shard streams
// (ReedSolomon::encode_stream, verify_stream, reconstruct_stream_from,
// src/core/stream.rs:677-1031).
//
// Shard streams are read and written through caller callbacks one block at a
// time.  Each block goes host -> pinned staging -> device -> kernel -> pinned ->
// host.  Two staging sets alternate: while the device copies and computes block
// i, the host reads block i+1 into the other set; once block i is done, block
// i+1 is sent to the device before the host writes block i's outputs, so the
// writes overlap the device work too.  Semantics follow the reference:
//   * the block size is clamped to [1 KiB, 16 MiB] (stream.rs:41-44, 686-688);
//   * Leopard codecs return UnsupportedCodecFamily (stream.rs:695-697);
//   * encode: each block is as long as the longest data read; shorter data
//     blocks are zero-padded (stream.rs:259-283); parity writes carry the shard
//     index k + j on error (write_block, stream.rs:473-493);
//   * verify: every shard must read the same length per block, else
//     IncorrectShardSize at the first short shard; the first mismatching block
//     returns false (stream.rs:777-810);
//   * reconstruct: reader/writer layout and present count are checked before
//     any I/O (stream.rs:956-974); present streams must read equal lengths per
//     block (stream.rs:379-397); missing shards are rebuilt block by block;
//   * io_mode SERIAL / PARALLEL / AUTO (stream.rs:214-226): PARALLEL runs one
//     block's callbacks on the library's host worker threads.
// The outputs of block i are written before an error of block i+1 is returned.
// When a write of block i fails, or verify rejects block i, the readers have
// also consumed block i+1, one block more than the reference reads.
#include <algorithm>
#include <cstring>

#include "internal.h"

namespace {

using namespace rsgpu;

constexpr size_t kMinBlock = 1024, kMaxBlock = 16u << 20, kDefaultBlock = 4u << 20;
constexpr size_t kAutoParallelBlock = 4u << 20, kAutoParallelStreams = 10;

// Pinned and device staging for two blocks of n shards.  Each thread keeps one
// (reused across calls, grown on demand, rebuilt on another device); a call made
// while the thread's own is busy (a callback that streams again) uses its own.
struct StreamCtx {
  int device = -1;
  size_t cap = 0;                         // bytes per set
  size_t block = 0;                       // this call's row pitch
  bool busy = false;
  hipStream_t s = nullptr;
  uint8_t* host[2] = {nullptr, nullptr};  // pinned, [shards][block]
  DeviceBuffer dev[2];
  DeviceBuffer flag;                      // verify flags, one per set
  int32_t* host_flag = nullptr;           // pinned, one per set
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~StreamCtx() { release(); }
  void release() {
    if (s) (void)hipStreamSynchronize(s);
    for (int i = 0; i < 2; ++i) {
      if (host[i]) (void)hipHostFree(host[i]);
      if (dev[i].p) (void)hipFree(dev[i].p);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      host[i] = nullptr, dev[i].p = nullptr, dev[i].bytes = 0, ev[i] = nullptr;
    }
    if (host_flag) (void)hipHostFree(host_flag);
    if (flag.p) (void)hipFree(flag.p);
    if (s) (void)hipStreamDestroy(s);
    host_flag = nullptr, flag.p = nullptr, flag.bytes = 0, s = nullptr;
    cap = block = 0;
    device = -1;
  }
  int ensure(size_t n, size_t block_) {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return RSGPU_ERR_HIP;
    if (d != device) {
      release();
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return RSGPU_ERR_HIP;
      for (int i = 0; i < 2; ++i)
        if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return RSGPU_ERR_HIP;
      if (flag.alloc(2 * sizeof(int32_t)) != hipSuccess ||
          hipHostMalloc((void**)&host_flag, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return RSGPU_ERR_OUT_OF_MEMORY;
      device = d;
    }
    if (n * block_ > cap) {
      if (hipStreamSynchronize(s) != hipSuccess) return RSGPU_ERR_HIP;
      cap = 0;
      for (int i = 0; i < 2; ++i) {
        if (host[i]) (void)hipHostFree(host[i]);
        host[i] = nullptr;
        if (hipHostMalloc((void**)&host[i], n * block_, hipHostMallocDefault) != hipSuccess ||
            dev[i].alloc(n * block_) != hipSuccess)
          return RSGPU_ERR_OUT_OF_MEMORY;
      }
      cap = n * block_;
    }
    block = block_;
    return RSGPU_OK;
  }
  uint8_t* h(int set, size_t shard) { return host[set] + shard * block; }
  uint8_t* d(int set, size_t shard) { return (uint8_t*)dev[set].p + shard * block; }
};

thread_local StreamCtx t_ctx;

// The thread's staging, or a call-local one while the thread's is in use.
class Staging {
 public:
  Staging() : ctx_(t_ctx.busy ? &own_ : &t_ctx) { ctx_->busy = true; }
  ~Staging() {
    if (ctx_->s) (void)hipStreamSynchronize(ctx_->s);  // no copy may outlive the call
    ctx_->busy = false;
  }
  StreamCtx& operator*() { return *ctx_; }

 private:
  StreamCtx own_;
  StreamCtx* ctx_;
};

void set_err(rsgpu_stream_error* e, int kind, size_t index, int64_t code) {
  if (e) {
    e->kind = kind;
    e->shard_index = (int32_t)index;
    e->io_error = code;
  }
}

int codec_err(rsgpu_stream_error* e, size_t index, int rc) {
  set_err(e, RSGPU_STREAM_CODEC, index, rc);
  return rc;
}

// read_one_stream (stream.rs:303-321): fill up to max bytes, stop at EOF.
int read_full(rsgpu_read_cb rd, void* user, size_t shard, uint8_t* buf, size_t max, size_t* got,
              rsgpu_stream_error* e) {
  size_t total = 0;
  while (total < max) {
    int64_t r = rd(user, shard, buf + total, max - total);
    if (r < 0) {
      set_err(e, RSGPU_STREAM_READ, shard, r);
      return RSGPU_ERR_IO;
    }
    if (r == 0) break;
    total += (size_t)r;
  }
  *got = total;
  return RSGPU_OK;
}

// write_block (stream.rs:473-493): all of it, a zero-length write is an error.
int write_full(rsgpu_write_cb wr, void* user, size_t shard, size_t index, const uint8_t* buf,
               size_t len, rsgpu_stream_error* e) {
  size_t done = 0;
  while (done < len) {
    int64_t r = wr(user, shard, buf + done, len - done);
    if (r <= 0) {
      set_err(e, RSGPU_STREAM_WRITE, index, r == 0 ? 0 : r);
      return RSGPU_ERR_IO;
    }
    done += (size_t)r;
  }
  return RSGPU_OK;
}

struct Options {
  size_t requested = kDefaultBlock;  // the caller's value, for the AUTO rule
  size_t block = kDefaultBlock;      // clamped
  int32_t mode = RSGPU_STREAM_IO_AUTO;
};

int resolve(const rsgpu_stream_options* o, Options* out) {
  if (o) {
    if (o->io_mode < RSGPU_STREAM_IO_AUTO || o->io_mode > RSGPU_STREAM_IO_PARALLEL)
      return RSGPU_ERR_INVALID_ARGUMENT;
    out->requested = o->block_size;
    out->mode = o->io_mode;
  }
  out->block = std::min(std::max(out->requested, kMinBlock), kMaxBlock);
  return RSGPU_OK;
}

// use_parallel_stream_io (stream.rs:214-226): AUTO compares the unclamped size.
bool parallel_io(const Options& o, size_t streams) {
  if (o.mode == RSGPU_STREAM_IO_SERIAL) return false;
  if (o.mode == RSGPU_STREAM_IO_PARALLEL) return true;
  return o.requested >= kAutoParallelBlock && streams >= kAutoParallelStreams;
}

// Reads one block of each listed shard into set `set`; got[t] = bytes of
// shards[t].  A parallel block reports the failing shard listed first.
int read_shards(rsgpu_read_cb rd, void* user, const std::vector<size_t>& shards, bool par,
                StreamCtx& X, int set, std::vector<size_t>& got, rsgpu_stream_error* err) {
  got.assign(shards.size(), 0);
  if (!par) {
    for (size_t t = 0; t < shards.size(); ++t) {
      int r = read_full(rd, user, shards[t], X.h(set, shards[t]), X.block, &got[t], err);
      if (r) return r;
    }
    return RSGPU_OK;
  }
  std::vector<int> rc(shards.size(), RSGPU_OK);
  std::vector<rsgpu_stream_error> e(shards.size());
  host_parallel_for(shards.size(), [&](size_t t) {
    rc[t] = read_full(rd, user, shards[t], X.h(set, shards[t]), X.block, &got[t], &e[t]);
  });
  for (size_t t = 0; t < shards.size(); ++t)
    if (rc[t]) {
      if (err) *err = e[t];
      return rc[t];
    }
  return RSGPU_OK;
}

// Writes len bytes of each listed row of set `set`: row rows[t] goes to writer
// writer[t], reported as shard index[t] on error.
int write_shards(rsgpu_write_cb wr, void* user, const std::vector<size_t>& rows,
                 const std::vector<size_t>& writer, const std::vector<size_t>& index, bool par,
                 StreamCtx& X, int set, size_t len, rsgpu_stream_error* err) {
  if (!par) {
    for (size_t t = 0; t < rows.size(); ++t) {
      int r = write_full(wr, user, writer[t], index[t], X.h(set, rows[t]), len, err);
      if (r) return r;
    }
    return RSGPU_OK;
  }
  std::vector<int> rc(rows.size(), RSGPU_OK);
  std::vector<rsgpu_stream_error> e(rows.size());
  host_parallel_for(rows.size(), [&](size_t t) {
    rc[t] = write_full(wr, user, writer[t], index[t], X.h(set, rows[t]), len, &e[t]);
  });
  for (size_t t = 0; t < rows.size(); ++t)
    if (rc[t]) {
      if (err) *err = e[t];
      return rc[t];
    }
  return RSGPU_OK;
}

int count_check(size_t actual, size_t expected, rsgpu_stream_error* e) {
  if (actual < expected) return codec_err(e, 0, RSGPU_TOO_FEW_SHARDS);
  if (actual > expected) return codec_err(e, 0, RSGPU_TOO_MANY_SHARDS);
  return RSGPU_OK;
}

bool is_leopard(const Codec& c) { return c.family != RSGPU_FAMILY_CLASSIC; }

int hip_fail(hipError_t e, const char* what) {
  set_hip_error(e, what);
  return RSGPU_ERR_HIP;
}

#define HIPS(expr)                                          \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return hip_fail(_e, #expr);       \
  } while (0)

// Callbacks of rsgpu_reconstruct_stream over the caller's byte vectors.  A read
// copies from data[pos..len) and advances pos.  A write appends at data[len..),
// first asking `grow` for room: the first write of a shard asks for at least
// `reserve` bytes (the reference's reserve_exact of the common length,
// stream.rs:423-432), later ones for exactly what the block needs.  Each call
// touches one shard's vector (and its `reserved` byte), so PARALLEL I/O is safe.
struct VecStreams {
  rsgpu_vec* v;
  rsgpu_grow_cb grow;
  void* user;
  size_t reserve;
  std::vector<uint8_t> reserved;
};

int64_t vec_read(void* u, size_t shard, uint8_t* buf, size_t len) {
  rsgpu_vec& v = static_cast<VecStreams*>(u)->v[shard];
  const size_t n = v.pos < v.len ? std::min(len, v.len - v.pos) : 0;
  if (n) std::memcpy(buf, v.data + v.pos, n);
  v.pos += n;
  return (int64_t)n;
}

int64_t vec_write(void* u, size_t shard, const uint8_t* buf, size_t len) {
  VecStreams& S = *static_cast<VecStreams*>(u);
  rsgpu_vec& v = S.v[shard];
  size_t need = v.len + len;
  if (!S.reserved[shard]) {
    S.reserved[shard] = 1;
    need = std::max(need, v.len + S.reserve);
  }
  if (need > v.cap) {
    if (!S.grow || S.grow(S.user, shard, &v, need) != 0 || v.cap < need || !v.data) return -1;
  }
  std::memcpy(v.data + v.len, buf, len);
  v.len += len;
  return (int64_t)len;
}

std::vector<size_t> iota(size_t n, size_t from = 0) {
  std::vector<size_t> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = from + i;
  return v;
}

// The two-set pipeline shared by encode and reconstruct: read(set, &len) fills
// a set, submit(set, len) enqueues its device work and records ev[set],
// write(set, len) hands its outputs to the writers.  Block i+1 is read while
// block i runs and is submitted before block i's outputs are written.
template <class Read, class Submit, class Write>
int pipeline(StreamCtx& X, Read read, Submit submit, Write write) {
  size_t lens[2] = {0, 0};
  int rc = read(0, &lens[0]);
  if (rc || !lens[0]) return rc;
  if ((rc = submit(0, lens[0]))) return rc;
  for (int cur = 0;; cur ^= 1) {
    const int nxt = cur ^ 1;
    const int read_rc = read(nxt, &lens[nxt]);  // overlaps the device work of `cur`
    HIPS(hipEventSynchronize(X.ev[cur]));
    const int sub_rc = (!read_rc && lens[nxt]) ? submit(nxt, lens[nxt]) : RSGPU_OK;
    if ((rc = write(cur, lens[cur]))) return rc;  // overlaps the device work of `nxt`
    if (read_rc) return read_rc;
    if (sub_rc) return sub_rc;
    if (!lens[nxt]) return RSGPU_OK;
  }
}

}  // namespace

extern "C" {

void rsgpu_stream_options_default(rsgpu_stream_options* o) {
  if (!o) return;
  o->block_size = kDefaultBlock;
  o->io_mode = RSGPU_STREAM_IO_AUTO;
  o->reserved = 0;
}

int64_t rsgpu_mem_stream_read(void* user, size_t shard, uint8_t* buf, size_t len) {
  rsgpu_mem_stream& r = static_cast<rsgpu_mem_streams*>(user)->readers[shard];
  const size_t n = r.pos < r.len ? std::min(len, r.len - r.pos) : 0;
  if (n) std::memcpy(buf, r.data + r.pos, n);
  r.pos += n;
  return (int64_t)n;
}

int64_t rsgpu_mem_stream_write(void* user, size_t shard, const uint8_t* buf, size_t len) {
  rsgpu_mem_stream& w = static_cast<rsgpu_mem_streams*>(user)->writers[shard];
  const size_t n = w.pos < w.len ? std::min(len, w.len - w.pos) : 0;
  if (n) std::memcpy(w.data + w.pos, buf, n);
  w.pos += n;
  return (int64_t)n;
}

int rsgpu_encode_stream(const rsgpu_codec* h, size_t n_data, size_t n_parity, rsgpu_read_cb rd,
                        rsgpu_write_cb wr, void* user, const rsgpu_stream_options* options,
                        rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  Options o;
  if (!h || !rd || !wr || resolve(options, &o)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  int rc = count_check(n_data, c.k, err);
  if (rc || (rc = count_check(n_parity, c.m, err))) return rc;
  const size_t k = c.k, m = c.m;
  Staging st;
  StreamCtx& X = *st;
  if ((rc = X.ensure(k + m, o.block))) return rc;
  const std::vector<size_t> data = iota(k), rows = iota(m, k), writers = iota(m), index = iota(m, k);
  const bool par_rd = parallel_io(o, k), par_wr = parallel_io(o, m);
  std::vector<size_t> got;
  // a block is as long as the longest data read; shorter reads are zero-padded
  auto read = [&](int set, size_t* len) -> int {
    *len = 0;
    int r = read_shards(rd, user, data, par_rd, X, set, got, err);
    if (r) return r;
    const size_t longest = got.empty() ? 0 : *std::max_element(got.begin(), got.end());
    for (size_t i = 0; i < k; ++i)
      if (got[i] < longest) std::memset(X.h(set, i) + got[i], 0, longest - got[i]);
    *len = longest;
    return RSGPU_OK;
  };
  auto submit = [&](int set, size_t len) -> int {
    HIPS(hipMemcpy2DAsync(X.d(set, 0), X.block, X.h(set, 0), X.block, len, k, hipMemcpyHostToDevice, X.s));
    std::vector<rsgpu_cslice> din(k);
    std::vector<rsgpu_slice> dout(m);
    for (size_t i = 0; i < k; ++i) din[i] = {X.d(set, i), len};
    for (size_t j = 0; j < m; ++j) dout[j] = {X.d(set, k + j), len};
    int r = rsgpu_encode_sep(h, din.data(), k, dout.data(), m, X.s);
    if (r) return codec_err(err, 0, r);
    HIPS(hipMemcpy2DAsync(X.h(set, k), X.block, X.d(set, k), X.block, len, m, hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[set], X.s));
    return RSGPU_OK;
  };
  auto write = [&](int set, size_t len) -> int {
    return write_shards(wr, user, rows, writers, index, par_wr, X, set, len, err);
  };
  return pipeline(X, read, submit, write);
}

int rsgpu_verify_stream(const rsgpu_codec* h, size_t n_shards, rsgpu_read_cb rd, void* user,
                        const rsgpu_stream_options* options, int32_t* ok, rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  Options o;
  if (!h || !rd || !ok || resolve(options, &o)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  const size_t n = c.k + c.m;
  int rc = count_check(n_shards, n, err);
  if (rc) return rc;
  Staging st;
  StreamCtx& X = *st;
  if ((rc = X.ensure(n, o.block))) return rc;
  *ok = 1;
  const std::vector<size_t> all = iota(n);
  const bool par_rd = parallel_io(o, n);
  std::vector<size_t> got;
  size_t lens[2] = {0, 0};
  // one block of every shard; returns the shard-size error of stream.rs:788-790
  auto read = [&](int set, size_t* len) -> int {
    *len = 0;
    int r = read_shards(rd, user, all, par_rd, X, set, got, err);
    if (r) return r;
    const size_t longest = *std::max_element(got.begin(), got.end());
    *len = longest;
    if (longest)
      for (size_t i = 0; i < n; ++i)
        if (got[i] != longest) return codec_err(err, i, RSGPU_INCORRECT_SHARD_SIZE);
    return RSGPU_OK;
  };
  auto submit = [&](int set, size_t len) -> int {
    HIPS(hipMemcpy2DAsync(X.d(set, 0), X.block, X.h(set, 0), X.block, len, n, hipMemcpyHostToDevice, X.s));
    std::vector<rsgpu_cslice> sh(n);
    for (size_t i = 0; i < n; ++i) sh[i] = {X.d(set, i), len};
    int32_t* dflag = (int32_t*)X.flag.p + set;
    int r = rsgpu_verify_async(h, sh.data(), n, dflag, X.s);
    if (r) return codec_err(err, 0, r);
    HIPS(hipMemcpyAsync(X.host_flag + set, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[set], X.s));
    return RSGPU_OK;
  };
  if ((rc = read(0, &lens[0])) || !lens[0]) return rc;
  if ((rc = submit(0, lens[0]))) return rc;
  for (int cur = 0;; cur ^= 1) {
    const int nxt = cur ^ 1;
    const int read_rc = read(nxt, &lens[nxt]);
    HIPS(hipEventSynchronize(X.ev[cur]));
    if (X.host_flag[cur] != 0) {  // this block fails: later blocks are not examined
      *ok = 0;
      set_err(err, RSGPU_STREAM_OK, 0, 0);
      return RSGPU_OK;
    }
    if (read_rc) return read_rc;
    if (!lens[nxt]) return RSGPU_OK;
    if ((rc = submit(nxt, lens[nxt]))) return rc;
  }
}

int rsgpu_reconstruct_stream_from(const rsgpu_codec* h, size_t n_readers, size_t n_writers,
                                  const uint8_t* has_reader, const uint8_t* has_writer, rsgpu_read_cb rd,
                                  rsgpu_write_cb wr, void* user, const rsgpu_stream_options* options,
                                  rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  Options o;
  if (!h || !rd || !wr || (!has_reader && n_readers) || (!has_writer && n_writers) || resolve(options, &o))
    return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t n = c.k + c.m;
  int rc = count_check(n_readers, n, err);
  if (rc || (rc = count_check(n_writers, n, err))) return rc;
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  std::vector<size_t> present, missing;
  for (size_t i = 0; i < n; ++i) {
    if ((has_reader[i] != 0) == (has_writer[i] != 0)) return codec_err(err, i, RSGPU_INVALID_SHARD_FLAGS);
    (has_reader[i] ? present : missing).push_back(i);
  }
  if (present.size() < c.k) return codec_err(err, 0, RSGPU_TOO_FEW_SHARDS_PRESENT);
  Staging st;
  StreamCtx& X = *st;
  if ((rc = X.ensure(n, o.block))) return rc;
  const bool par_rd = parallel_io(o, present.size()), par_wr = parallel_io(o, missing.size());
  std::vector<size_t> got;
  // present streams only; every present stream must read the block's length
  auto read = [&](int set, size_t* len) -> int {
    *len = 0;
    int r = read_shards(rd, user, present, par_rd, X, set, got, err);
    if (r) return r;
    const size_t longest = *std::max_element(got.begin(), got.end());
    *len = longest;
    if (longest)
      for (size_t t = 0; t < present.size(); ++t)
        if (got[t] != longest) return codec_err(err, present[t], RSGPU_INCORRECT_SHARD_SIZE);
    return RSGPU_OK;
  };
  auto submit = [&](int set, size_t len) -> int {
    std::vector<rsgpu_slot> slots(n);
    for (size_t i = 0; i < n; ++i) slots[i] = {X.d(set, i), len, has_reader[i] ? 1 : 0};
    // one copy per run of consecutive present rows
    for (size_t t = 0; t < present.size();) {
      size_t u = t + 1;
      while (u < present.size() && present[u] == present[u - 1] + 1) ++u;
      HIPS(hipMemcpy2DAsync(X.d(set, present[t]), X.block, X.h(set, present[t]), X.block, len, u - t,
                            hipMemcpyHostToDevice, X.s));
      t = u;
    }
    int r = rsgpu_reconstruct(h, slots.data(), n, X.s);
    if (r) return codec_err(err, 0, r);
    for (size_t i : missing)
      HIPS(hipMemcpyAsync(X.h(set, i), X.d(set, i), len, hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[set], X.s));
    return RSGPU_OK;
  };
  auto write = [&](int set, size_t len) -> int {
    return write_shards(wr, user, missing, missing, missing, par_wr, X, set, len, err);
  };
  return pipeline(X, read, submit, write);
}

// ReedSolomon::reconstruct_stream (stream.rs:839-900): the count check, the
// Leopard refusal, then presence by length (an empty vector is missing); no
// present shard is Ok with nothing written, fewer than k is TooFewShardsPresent.
// Present vectors are read from position 0; missing ones are appended to, with
// the common length as their first reservation when every present shard has
// the same length.  The rest is reconstruct_stream_from.
int rsgpu_reconstruct_stream(const rsgpu_codec* h, rsgpu_vec* shards, size_t n_shards, rsgpu_grow_cb grow,
                             void* user, const rsgpu_stream_options* options, rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  Options o;
  if (!h || (!shards && n_shards) || resolve(options, &o)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t n = c.k + c.m;
  int rc = count_check(n_shards, n, err);
  if (rc) return rc;
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  size_t present = 0, known = 0;
  bool equal = true;
  for (size_t i = 0; i < n; ++i) {
    const size_t len = shards[i].len;
    if (!len) continue;
    ++present;
    if (!known) known = len;
    equal = equal && len == known;
  }
  if (!present) return RSGPU_OK;
  if (present < c.k) return codec_err(err, 0, RSGPU_TOO_FEW_SHARDS_PRESENT);
  std::vector<uint8_t> has_reader(n), has_writer(n);
  for (size_t i = 0; i < n; ++i) {
    has_reader[i] = shards[i].len != 0;
    has_writer[i] = !has_reader[i];
    if (has_reader[i]) shards[i].pos = 0;
  }
  VecStreams S{shards, grow, user, equal ? known : 0, std::vector<uint8_t>(n, 0)};
  return rsgpu_reconstruct_stream_from(h, n, n, has_reader.data(), has_writer.data(), vec_read, vec_write, &S,
                                       options, err);
}

}  // extern "C"
