// internal.h -- types shared by the rsgpu host code and its HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cstddef>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rsgpu.h"
#include "gf.h"

namespace rsgpu {

// Pointers handed to one matrix-apply launch travel in the kernel arguments
// (scalar-loaded from the kernarg segment), so a launch needs no host-to-device
// copy.  Wider codecs are applied in column chunks of kMaxIn inputs and row
// chunks of kMaxOut outputs.
constexpr int kMaxIn = 64;
constexpr int kMaxOut = 8;
constexpr int kBlock = 256;

enum ApplyMode : int { kWrite = 0, kAccumulate = 1, kVerify = 2 };

// Coefficient-table words that can travel inside ApplyArgs itself (2 KiB: 64
// GF(2^8) coefficients, e.g. 10 inputs x 4 outputs), so a launch needs no table
// in device memory.
constexpr int kInlineTabWords = 512;

struct ApplyArgs {
  const uint8_t* in[kMaxIn];
  uint8_t* out[kMaxOut];   // destinations, or the stored parity when verifying
  const uint32_t* tables;  // row-major [row][input][kTabWords]; nullptr: tab_inline
  int32_t* flag;           // verify: set nonzero on any mismatch
  uint64_t n16;            // 16-byte column vectors per shard
  uint64_t v_begin;        // first vector handled by this launch
  int32_t k;               // inputs in this launch
  int32_t tab_row_stride;  // dwords between consecutive table rows
  int32_t tab_col0;        // table column of in[0]
  int32_t r;               // outputs in this launch (<= kMaxOut)
  // verify only: the flag protocol of one call (verify_claim / verify_report);
  // null: the caller cleared the flag
  uint64_t* clear_ws;
  uint64_t call_id;
  // Split over inputs (in_dev set, write mode): blockIdx.z takes inputs
  // [z * k, z * k + k) of the k_total device pointers in_dev (the last chunk
  // fewer) and writes its partial sums to rows z * r + row of `part`, part_pitch
  // bytes apart; xor_partials then folds the chunks into the outputs.
  const uint8_t* const* in_dev;
  uint8_t* part;
  uint64_t part_pitch;
  int32_t k_total;
  int32_t pad_;
  // the tables when `tables` is null (same layout, rows tab_row_stride apart);
  // the kernel reads them straight from its kernel-argument segment
  uint32_t tab_inline[kInlineTabWords];
};
static_assert(offsetof(ApplyArgs, tab_inline) % 16 == 0, "16-byte direct loads of the inline tables");

struct ByteArgs {
  const uint8_t* in[kMaxIn];
  uint8_t* out[kMaxOut];
  const uint32_t* tables;
  int32_t* flag;
  uint64_t begin, end;  // byte range [begin, end)
  int32_t k, tab_row_stride, tab_col0, r;
  int32_t mode;
  int32_t tower;  // 1: GF(2^16) tower elements (byte pairs)
};

// Kernel variants a call can be dispatched to.  Every variant computes the same
// bytes; they differ only in how the work is tiled.
//   0 .. kNumApplyVariants-1: matrix apply with tile shape kTileShapes[v % kNumTiles]
//       and kRowSplit[v / kNumTiles] outputs per workgroup (0: all of them).  On a
//       Leopard GF(2^8) codec these apply its generator matrix instead of the FFT.
//   kNumApplyVariants ..: Leopard FFT kernels with a column tile of
//       kLeoTileWords[v - kNumApplyVariants] words per workgroup (clamped to LDS).
struct TileShape {
  int lanes;  // lanes per workgroup
  int nv;     // 16-byte vectors per lane
};
constexpr int kNumTiles = 5;
constexpr TileShape kTileShapes[kNumTiles] = {{256, 1}, {256, 2}, {64, 1}, {64, 2}, {64, 4}};
// The default tiling of GF(2^8) calls (no variant chosen): 256-lane workgroups,
// one dword of every shard per lane (apply_dword_kernel).
constexpr int kTileDword = kNumTiles;
constexpr int kNumRowSplits = 3;
constexpr int kRowSplit[kNumRowSplits] = {0, 2, 1};
constexpr int kNumApplyVariants = kNumTiles * kNumRowSplits;
constexpr int kNumLeoTiles = 5;
constexpr int kLeoTileWords[kNumLeoTiles] = {4, 8, 16, 32, 64};
constexpr int kNumVariants = kNumApplyVariants + kNumLeoTiles;
// Dispatch-table shard-size buckets: round(2 * log2(len)) - kBucket0, clamped.
constexpr int kBucket0 = 12;
constexpr int kNumBuckets = 50;
enum DispatchOp : int { kOpEncode = 0, kOpVerify = 1, kOpReconstruct = 2, kNumOps = 3 };

// Launchers (classic_kernels.hip; the apply kernels are compiled from apply_inst.hip).  `tile` indexes
// kTileShapes, `rows` is the number of outputs per workgroup.  `tower` 1 selects
// GF((2^8)^2) element pairs, 2 the split-layout GF(2^16) elements of Leopard
// shards; tables are then four byte maps per coefficient [row][input][4][kTabWords].
hipError_t launch_apply(const ApplyArgs& a, int mode, int tile, int rows, int tower, hipStream_t s);
hipError_t launch_apply_bytes(const ByteArgs& a, hipStream_t s);
// out[row] (^)= XOR over z < chunks of part[(z * rows + row) * pitch ..], len bytes
// (a multiple of 16) per row; accumulate XORs into the outputs instead of writing.
hipError_t launch_xor_partials(const uint8_t* part, uint64_t pitch, int chunks, int rows,
                               uint8_t* const* out, uint64_t len, bool accumulate, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int32_t v, int n, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t s);

// Decode rows on the device (classic_kernels.hip).  One workgroup solves the
// missing-data equations of the selected survivors and composes the requested
// rows, then writes their byte-permute tables.
struct DecodeJob {
  uint16_t valid[256];    // survivor shard indices (first k present), k entries
  uint16_t targets[256];  // requested outputs: missing data first, then parity; a target
                          // that is itself a survivor gets its unit row
  uint16_t cols[256];     // survivor columns kept in the rows (n_cols of them, ascending)
  int32_t k, m, n_targets, tower;
  const uint8_t* field;   // device FieldTables log[256] + exp[512] (nullptr: 0x11d field)
  int32_t one;            // the field's multiplicative identity
  int32_t n_cols;         // 0: every column 0..k-1
};
// Sets *flag when row r of a differs from row r of b anywhere in [0, len).
struct CompareArgs {
  const uint8_t* a[kMaxOut];
  const uint8_t* b[kMaxOut];
  uint64_t len;
  int32_t* flag;
  int32_t rows;
};
hipError_t launch_compare_rows(const CompareArgs& c, hipStream_t s);

// dst[i] ^= src[i] over len[i] bytes for up to kXorPairs pairs per launch.
constexpr int kXorPairs = 16;
struct XorPairs {
  const uint8_t* src[kXorPairs];
  uint8_t* dst[kXorPairs];
  uint64_t len[kXorPairs];
  int32_t n;
};
hipError_t launch_xor_pairs(const XorPairs& p, hipStream_t s);

hipError_t launch_decode_rows(const uint16_t* d_matrix, const DecodeJob& job, uint8_t* d_rows,
                              uint32_t* d_tables, int32_t* d_status, hipStream_t s);
// Dynamic LDS of one solve (0: more than kMaxSolve missing data rows).
size_t decode_rows_lds(const DecodeJob& job);
// Independent solves of one codec in one launch (a job array in device memory).
struct DecodeMultiJob {
  DecodeJob job;
  uint8_t* rows;
  uint32_t* tables;
};
hipError_t launch_decode_rows_multi(const uint16_t* d_matrix, const DecodeMultiJob* d_jobs, int n,
                                    size_t lds, int32_t* d_status, hipStream_t s);

// Decode rows of codecs beyond DecodeJob's limits (classic GF(2^16) with k > 256,
// more than 256 targets, or more than kMaxSolve missing data shards): C^-1 of
// the missing-data system by one workgroup -- in LDS up to kMaxSolve rows, in a
// device work matrix beyond -- then every (target, column) coefficient in a
// grid.  Survivors are the present data (every index below k not in `missing`)
// followed by `par_surv`, as make_plan's first-k-present rule orders them.
constexpr int kMaxSolve = 128;        // missing data rows solvable in LDS
constexpr int kMaxBigSolve = 8192;    // missing data rows of the device-memory solve
constexpr int kBigTargetChunk = 1024;  // targets per rows launch
struct DecodeBigJob {
  int32_t k, e, tower, one;
  const uint8_t* field;      // as in DecodeJob
  const uint16_t* missing;   // device: e missing data indices, ascending
  const uint16_t* par_surv;  // device: e parity survivors solving for them, ascending
  const uint16_t* present;   // device: the k - e present data indices, ascending
};
// d_work: 2 * e * e elements of device scratch, needed when e > kMaxSolve.
hipError_t launch_decode_big(const uint16_t* d_matrix, const DecodeBigJob& job,
                             const std::vector<int>& targets, uint16_t* d_cinv, uint16_t* d_work,
                             uint8_t* d_rows, uint32_t* d_tables, int32_t* d_status, hipStream_t s);
// Writes n 16-bit values from the host to device memory on stream s through
// kernel arguments (no staging buffer, nothing to keep alive after the call).
hipError_t launch_store_u16(uint16_t* dst, const uint16_t* vals, size_t n, hipStream_t s);
// dst[0..n) = vals, the values carried in kernel arguments (no staging buffer)
hipError_t launch_store_u32(uint32_t* dst, const uint32_t* vals, size_t n, hipStream_t s);

// Calls fn(0..n-1) on the library's persistent host worker threads and the
// calling thread; returns when every call has finished.
void host_parallel_for(size_t n, const std::function<void(size_t)>& fn);

// Tables for a host-known row set (encode rows, apply_rows primitive).
void build_tables_host(const uint16_t* rows, int r, int k, bool tower, std::vector<uint32_t>& out,
                       const FieldTables* field = nullptr);

// Least-recently-used map of shared values with an entry capacity.  An insert
// past the capacity evicts from the least recent end into `evicted`, which the
// caller drops after releasing its lock: a value's destructor may free device
// memory, and a hipFree waits for the whole device.
template <class V>
struct LruMap {
  using Item = std::pair<std::string, std::shared_ptr<V>>;
  size_t cap = 64;
  std::list<Item> order;  // front: most recent
  std::unordered_map<std::string, typename std::list<Item>::iterator> index;
  std::shared_ptr<V> get(const std::string& key) {
    auto it = index.find(key);
    if (it == index.end()) return nullptr;
    order.splice(order.begin(), order, it->second);
    return it->second->second;
  }
  void put(const std::string& key, std::shared_ptr<V> v, std::vector<std::shared_ptr<V>>& evicted) {
    auto it = index.find(key);
    if (it != index.end()) {
      evicted.push_back(std::move(it->second->second));
      it->second->second = std::move(v);
      order.splice(order.begin(), order, it->second);
      return;
    }
    while (!order.empty() && index.size() >= cap) {
      evicted.push_back(std::move(order.back().second));
      index.erase(order.back().first);
      order.pop_back();
    }
    order.emplace_front(key, std::move(v));
    index[key] = order.begin();
  }
  size_t size() const { return index.size(); }
};

struct DeviceBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer();
  hipError_t alloc(size_t n);
};

// Device slots for decode entries of one codec, each with its own event.  A
// released slot retires: kernels already queued may still read it.  Retired
// slots become free again after one device-wide synchronisation, taken only
// when no free slot is left and at least kRecycleBatch have retired.  New
// slots come in slabs of one hipMalloc each (doubling up to kMaxSlab slots and
// 64 MiB; a slot's event is made on its first use), so
// a stream of cache misses costs neither a hipMalloc nor a synchronising
// hipFree per entry.
struct DecodePool {
  static constexpr size_t kRecycleBatch = 64;
  static constexpr size_t kMaxSlab = 32;
  struct Slot {
    void* mem;
    hipEvent_t ready;
  };
  std::mutex mu;
  size_t slot_bytes = 0;
  size_t next_slab = 8;  // slots in the next slab
  std::vector<void*> slabs;
  std::vector<Slot> all, free_slots, retired;
  ~DecodePool();
  hipError_t acquire(Slot& out);
  void release(const Slot& s);
  hipError_t ready_event(Slot& sl);  // with mu held
};

// One cached decode plan: the composed rows for a (present, targets) pattern.
// `slot` holds the tables [n_targets][k][words] followed by the rows
// [n_targets][k]; `slot.ready` is recorded on `solve_stream` behind the solve,
// and a call on another stream waits for it before reading the entry.
// An entry wider than the pool's slots (decode_idx rows for more targets than
// the codec has parity shards) owns its buffer and event instead.
struct DecodeEntry {
  std::shared_ptr<DecodePool> pool;
  DecodePool::Slot slot{nullptr, nullptr};
  std::unique_ptr<DeviceBuffer> own;  // set when the entry does not use a pool slot
  uint32_t* tables = nullptr;  // device tables; null while only h_tables exist
  uint8_t* rows = nullptr;
  // [targets][k][words] solved on the host, for patterns whose tables travel in
  // the kernel arguments (empty otherwise)
  std::vector<uint32_t> h_tables;
  std::mutex dev_mu;  // guards the device copy made for such an entry on demand
  hipStream_t solve_stream = nullptr;
  std::vector<int> valid;  // survivor indices
  std::vector<int> targets;
  size_t bytes = 0;
  ~DecodeEntry();
};

// A Leopard GF(2^8) decode as a matrix over the present shards (codec.hip,
// leopard_decoder_matrix_reconstruct): coefficient tables for every target.
struct DecoderMatrix {
  DeviceBuffer tables;  // [targets][inputs][kTabWords]
  std::vector<int> inputs, targets;
};

struct VariantTable {
  int n_erasures = 0;      // erasure counts 0 .. n_erasures-1 (larger ones clamp)
  std::vector<int16_t> v;  // [kNumOps][n_erasures][kNumBuckets]
};

struct Codec {
  size_t k = 0, m = 0;
  int field = RSGPU_FIELD_GF8;
  int family = RSGPU_FAMILY_CLASSIC;
  int matrix_mode = RSGPU_MATRIX_VANDERMONDE;
  bool fast_one_parity = false;
  bool use_cache = true;
  size_t cache_capacity = 0, cache_max_bytes = 0;
  int device = 0;
  std::vector<uint16_t> matrix;  // (k+m) x k elements
  DeviceBuffer d_matrix;         // same, on the device
  DeviceBuffer d_parity_tables;  // [m][k][words]
  std::vector<uint32_t> h_parity_tables;
  // decode cache (LRU), keyed by present mask + target mask
  mutable std::mutex mu;
  mutable std::list<std::pair<std::string, std::shared_ptr<DecodeEntry>>> lru;
  mutable std::unordered_map<std::string,
                             std::list<std::pair<std::string, std::shared_ptr<DecodeEntry>>>::iterator>
      index;
  mutable size_t cache_bytes = 0;
  mutable uint64_t requests = 0, hits = 0, misses = 0;
  mutable std::shared_ptr<DecodePool> decode_pool;  // created on the first miss
  // Leopard device state (leopard.hip)
  std::shared_ptr<void> leopard;
  // Kernel-variant dispatch (codec.hip): a forced variant (>= 0), else the
  // installed table [op][erasures][size bucket] (entries < 0: default tiling).
  std::atomic<int> forced_variant{-1};
  std::shared_ptr<const VariantTable> variant_table;  // std::atomic_load/store
  // Leopard GF(2^8): classic codec over the generator matrix [I; G], built on the
  // first call that dispatches to a matrix-apply variant (guarded by mu).
  mutable std::shared_ptr<Codec> leo_matrix;
  // Leopard GF(2^8): decoder matrices by (present, target) pattern, least
  // recently used evicted past 256 (guarded by mu).
  mutable LruMap<const DecoderMatrix> dec_matrix{256};
  // Byte field of the matrix when it is not the 0x11d field (the generator-matrix
  // codec of a Leopard GF(2^8) codec), with a device copy of log + exp.
  std::shared_ptr<const FieldTables> ftab;
  DeviceBuffer d_ftab;
  // A operands of the GF(2) int8-MFMA encode (gf2_mfma.hip), built on first use
  // (guarded by mu).
  mutable DeviceBuffer d_gf2_bits;
};

int words_per_coef(int tower);  // kTabWords or 4 * kTabWords



// Staging for presolve_decode's job arrays, owned by the caller of one stream
// and reused across calls; `done` is recorded behind the launches that read it.
// It grows geometrically, and outgrown buffers are freed only with the scratch
// (a hipFree waits for the whole device, which would stall a serving loop).
struct PresolveScratch {
  uint8_t* pinned = nullptr;
  size_t cap = 0;
  DeviceBuffer dev;
  std::vector<std::unique_ptr<DeviceBuffer>> old_dev;
  std::vector<uint8_t*> old_pinned;
  hipEvent_t done = nullptr;
  ~PresolveScratch();
  int ensure(size_t bytes);
};

// A reconstruct call to be solved ahead: its codec and its k + m slots.
struct PresolveCall {
  const Codec* codec;
  const rsgpu_slot* shards;
  size_t n;
};

// Solves, on stream s, the decode rows of every pattern among `calls` (goal: all
// missing shards) that the codec's cache does not hold yet: one launch per codec
// with one workgroup per pattern.  The entries enter the cache as misses would;
// calls then find them there.  Codecs outside the one-workgroup solve (Leopard,
// more than 256 data shards) are left to the calls themselves.
int presolve_decode(const std::vector<PresolveCall>& calls, hipStream_t s, PresolveScratch& scr);
constexpr size_t kPresolveMin = 8;  // fewer new patterns of a codec: solved by their calls

// Sizes a batch's descriptor staging for n_stripes descriptors and n_ptrs shard
// pointers up front, so later uploads of that size allocate nothing.
int batch_reserve(rsgpu_batch* b, size_t n_stripes, size_t n_ptrs);

// thread-local last HIP error
void set_hip_error(hipError_t e, const char* where);

// Returns the status of a failed HIP call from the enclosing C-ABI function.
#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::rsgpu::set_hip_error(_e, #expr);                                          \
      return _e == hipErrorOutOfMemory ? RSGPU_ERR_OUT_OF_MEMORY : RSGPU_ERR_HIP; \
    }                                                                             \
  } while (0)

}  // namespace rsgpu

struct rsgpu_codec {
  rsgpu::Codec c;
};
