// block_stream.hip -- block-wise encode / verify / reconstruct of shard streams
// (ReedSolomon::encode_stream, verify_stream, reconstruct_stream_from,
// src/core/stream.rs:677-1031).
//
// Shard streams are read and written through caller callbacks one block at a
// time.  Each block goes host -> pinned staging -> device -> kernel -> pinned ->
// host.  Two staging sets alternate, so while the device copies and computes
// block i the host already reads block i+1 into the other set.  Semantics follow
// the reference:
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
//     block (stream.rs:379-397); missing shards are rebuilt block by block.
// On an I/O error the outputs of every earlier block have been written; the
// readers may have advanced one block further than the reference would read.
#include <algorithm>
#include <cstring>

#include "internal.h"

namespace {

using namespace rsgpu;

constexpr size_t kMinBlock = 1024, kMaxBlock = 16u << 20;

struct StreamCtx {
  size_t n = 0, block = 0;
  hipStream_t s = nullptr;
  uint8_t* host[2] = {nullptr, nullptr};  // pinned, n * block each
  DeviceBuffer dev[2];
  DeviceBuffer flag;                      // verify flags, one per set
  int32_t* host_flag = nullptr;           // pinned, one per set
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~StreamCtx() {
    if (s) (void)hipStreamSynchronize(s);
    for (int i = 0; i < 2; ++i) {
      if (host[i]) (void)hipHostFree(host[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    if (host_flag) (void)hipHostFree(host_flag);
    if (s) (void)hipStreamDestroy(s);
  }
  int init(size_t n_, size_t block_) {
    n = n_;
    block = block_;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return RSGPU_ERR_HIP;
    for (int i = 0; i < 2; ++i) {
      if (hipHostMalloc((void**)&host[i], n * block, hipHostMallocDefault) != hipSuccess ||
          dev[i].alloc(n * block) != hipSuccess ||
          hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
        return RSGPU_ERR_OUT_OF_MEMORY;
    }
    if (flag.alloc(2 * sizeof(int32_t)) != hipSuccess ||
        hipHostMalloc((void**)&host_flag, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
      return RSGPU_ERR_OUT_OF_MEMORY;
    return RSGPU_OK;
  }
  uint8_t* h(int set, size_t shard) { return host[set] + shard * block; }
  uint8_t* d(int set, size_t shard) { return (uint8_t*)dev[set].p + shard * block; }
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

size_t clamp_block(size_t b) { return std::min(std::max(b, kMinBlock), kMaxBlock); }

int count_check(size_t actual, size_t expected, rsgpu_stream_error* e) {
  if (actual < expected) return codec_err(e, 0, RSGPU_TOO_FEW_SHARDS);
  if (actual > expected) return codec_err(e, 0, RSGPU_TOO_MANY_SHARDS);
  return RSGPU_OK;
}

bool is_leopard(const Codec& c) { return c.family != RSGPU_FAMILY_CLASSIC; }

#define HIPS(expr)                          \
  do {                                      \
    hipError_t _e = (expr);                 \
    if (_e != hipSuccess) {                 \
      set_hip_error(_e, #expr);             \
      return RSGPU_ERR_HIP;                 \
    }                                       \
  } while (0)

}  // namespace

extern "C" {

int rsgpu_encode_stream(const rsgpu_codec* h, size_t n_data, size_t n_parity, rsgpu_read_cb rd,
                        rsgpu_write_cb wr, void* user, size_t block_size, rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  if (!h || !rd || !wr) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t block = clamp_block(block_size);
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  int rc = count_check(n_data, c.k, err);
  if (rc || (rc = count_check(n_parity, c.m, err))) return rc;
  const size_t k = c.k, m = c.m;
  StreamCtx X;
  if ((rc = X.init(k + m, block))) return rc;
  size_t lens[2] = {0, 0};
  // reads one block of every data stream into set `set`; the block length is
  // the longest read and shorter reads are zero-padded to it
  auto read_block = [&](int set, size_t* len) -> int {
    size_t longest = 0;
    std::vector<size_t> got(k);
    for (size_t i = 0; i < k; ++i) {
      int r = read_full(rd, user, i, X.h(set, i), block, &got[i], err);
      if (r) return r;
      longest = std::max(longest, got[i]);
    }
    for (size_t i = 0; i < k; ++i)
      if (got[i] < longest) std::memset(X.h(set, i) + got[i], 0, longest - got[i]);
    *len = longest;
    return RSGPU_OK;
  };
  int cur = 0;
  if ((rc = read_block(cur, &lens[cur]))) return rc;
  while (lens[cur]) {
    const size_t len = lens[cur];
    HIPS(hipMemcpy2DAsync(X.d(cur, 0), block, X.h(cur, 0), block, len, k, hipMemcpyHostToDevice, X.s));
    std::vector<rsgpu_cslice> din(k);
    std::vector<rsgpu_slice> dout(m);
    for (size_t i = 0; i < k; ++i) din[i] = {X.d(cur, i), len};
    for (size_t j = 0; j < m; ++j) dout[j] = {X.d(cur, k + j), len};
    if ((rc = rsgpu_encode_sep(h, din.data(), k, dout.data(), m, X.s))) return codec_err(err, 0, rc);
    HIPS(hipMemcpy2DAsync(X.h(cur, k), block, X.d(cur, k), block, len, m, hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[cur], X.s));
    const int nxt = cur ^ 1;
    const int read_rc = read_block(nxt, &lens[nxt]);  // overlaps the device work
    HIPS(hipEventSynchronize(X.ev[cur]));
    for (size_t j = 0; j < m; ++j)
      if ((rc = write_full(wr, user, j, k + j, X.h(cur, k + j), len, err))) return rc;
    if (read_rc) return read_rc;
    cur = nxt;
  }
  return RSGPU_OK;
}

int rsgpu_verify_stream(const rsgpu_codec* h, size_t n_shards, rsgpu_read_cb rd, void* user,
                        size_t block_size, int32_t* ok, rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  if (!h || !rd || !ok) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t block = clamp_block(block_size);
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  const size_t n = c.k + c.m;
  int rc = count_check(n_shards, n, err);
  if (rc) return rc;
  StreamCtx X;
  if ((rc = X.init(n, block))) return rc;
  *ok = 1;
  size_t lens[2] = {0, 0};
  // one block of every shard; returns the shard-size error of stream.rs:788-790
  auto read_block = [&](int set, size_t* len) -> int {
    std::vector<size_t> got(n);
    size_t longest = 0;
    for (size_t i = 0; i < n; ++i) {
      int r = read_full(rd, user, i, X.h(set, i), block, &got[i], err);
      if (r) return r;
      longest = std::max(longest, got[i]);
    }
    *len = longest;
    if (longest)
      for (size_t i = 0; i < n; ++i)
        if (got[i] != longest) return codec_err(err, i, RSGPU_INCORRECT_SHARD_SIZE);
    return RSGPU_OK;
  };
  int cur = 0;
  if ((rc = read_block(cur, &lens[cur]))) return rc;
  while (lens[cur]) {
    const size_t len = lens[cur];
    HIPS(hipMemcpy2DAsync(X.d(cur, 0), block, X.h(cur, 0), block, len, n, hipMemcpyHostToDevice, X.s));
    std::vector<rsgpu_cslice> sh(n);
    for (size_t i = 0; i < n; ++i) sh[i] = {X.d(cur, i), len};
    int32_t* dflag = (int32_t*)X.flag.p + cur;
    if ((rc = rsgpu_verify_async(h, sh.data(), n, dflag, X.s))) return codec_err(err, 0, rc);
    HIPS(hipMemcpyAsync(X.host_flag + cur, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[cur], X.s));
    const int nxt = cur ^ 1;
    const int read_rc = read_block(nxt, &lens[nxt]);
    HIPS(hipEventSynchronize(X.ev[cur]));
    if (X.host_flag[cur] != 0) {  // this block fails: later blocks are not examined
      *ok = 0;
      set_err(err, RSGPU_STREAM_OK, 0, 0);
      return RSGPU_OK;
    }
    if (read_rc) return read_rc;
    cur = nxt;
  }
  return RSGPU_OK;
}

int rsgpu_reconstruct_stream_from(const rsgpu_codec* h, size_t n_readers, size_t n_writers,
                                  const uint8_t* has_reader, const uint8_t* has_writer, rsgpu_read_cb rd,
                                  rsgpu_write_cb wr, void* user, size_t block_size,
                                  rsgpu_stream_error* err) {
  set_err(err, RSGPU_STREAM_OK, 0, 0);
  if (!h || !rd || !wr || (!has_reader && n_readers) || (!has_writer && n_writers))
    return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t n = c.k + c.m;
  int rc = count_check(n_readers, n, err);
  if (rc || (rc = count_check(n_writers, n, err))) return rc;
  if (is_leopard(c)) return codec_err(err, 0, RSGPU_UNSUPPORTED_CODEC_FAMILY);
  std::vector<size_t> missing;
  for (size_t i = 0; i < n; ++i) {
    if ((has_reader[i] != 0) == (has_writer[i] != 0)) return codec_err(err, i, RSGPU_INVALID_SHARD_FLAGS);
    if (!has_reader[i]) missing.push_back(i);
  }
  if (n - missing.size() < c.k) return codec_err(err, 0, RSGPU_TOO_FEW_SHARDS_PRESENT);
  const size_t block = clamp_block(block_size);
  StreamCtx X;
  if ((rc = X.init(n, block))) return rc;
  size_t lens[2] = {0, 0};
  // present streams only; every present stream must read the block's length
  auto read_block = [&](int set, size_t* len) -> int {
    std::vector<size_t> got(n, 0);
    size_t longest = 0;
    for (size_t i = 0; i < n; ++i) {
      if (!has_reader[i]) continue;
      int r = read_full(rd, user, i, X.h(set, i), block, &got[i], err);
      if (r) return r;
      longest = std::max(longest, got[i]);
    }
    *len = longest;
    if (longest)
      for (size_t i = 0; i < n; ++i)
        if (has_reader[i] && got[i] != longest) return codec_err(err, i, RSGPU_INCORRECT_SHARD_SIZE);
    return RSGPU_OK;
  };
  int cur = 0;
  if ((rc = read_block(cur, &lens[cur]))) return rc;
  while (lens[cur]) {
    const size_t len = lens[cur];
    std::vector<rsgpu_slot> slots(n);
    for (size_t i = 0; i < n; ++i) {
      slots[i] = {X.d(cur, i), len, has_reader[i] ? 1 : 0};
      if (has_reader[i])
        HIPS(hipMemcpyAsync(X.d(cur, i), X.h(cur, i), len, hipMemcpyHostToDevice, X.s));
    }
    if ((rc = rsgpu_reconstruct(h, slots.data(), n, X.s))) return codec_err(err, 0, rc);
    for (size_t i : missing)
      HIPS(hipMemcpyAsync(X.h(cur, i), X.d(cur, i), len, hipMemcpyDeviceToHost, X.s));
    HIPS(hipEventRecord(X.ev[cur], X.s));
    const int nxt = cur ^ 1;
    const int read_rc = read_block(nxt, &lens[nxt]);
    HIPS(hipEventSynchronize(X.ev[cur]));
    for (size_t i : missing)
      if ((rc = write_full(wr, user, i, i, X.h(cur, i), len, err))) return rc;
    if (read_rc) return read_rc;
    cur = nxt;
  }
  return RSGPU_OK;
}

}  // extern "C"
