// codec.hip -- host side of the classic data path behind the C ABI: codec
// construction (encoding matrices), argument validation in the reference's
// order, reconstruction planning, the device-resident decode-row cache, and the
// launch plumbing for the matrix-apply kernels.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <tuple>

#include "batch.h"
#include "internal.h"
#include "leopard.h"

namespace rsgpu {

// ------------------------------------------------------------------ errors
static thread_local std::string t_hip_error;

void set_hip_error(hipError_t e, const char* where) {
  t_hip_error = std::string(where) + ": " + hipGetErrorString(e);
}

DeviceBuffer::~DeviceBuffer() {
  if (p) (void)hipFree(p);
}
DecodePool::~DecodePool() {
  for (auto& sl : all)
    if (sl.ready) (void)hipEventDestroy(sl.ready);
  for (void* p : slabs) (void)hipFree(p);
}

// A slot's event is made on its first use (a slab's worth at once would stall
// the call that grows the pool).  Free and retired copies of a slot carry it.
hipError_t DecodePool::ready_event(Slot& sl) {
  if (sl.ready) return hipSuccess;
  hipError_t e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
  if (e != hipSuccess) return e;
  all.push_back(sl);
  return hipSuccess;
}

hipError_t DecodePool::acquire(Slot& out) {
  std::lock_guard<std::mutex> g(mu);
  if (free_slots.empty() && retired.size() >= kRecycleBatch) {
    hipError_t e = hipDeviceSynchronize();  // every queued reader of a retired slot is done
    if (e != hipSuccess) return e;
    free_slots.insert(free_slots.end(), retired.begin(), retired.end());
    retired.clear();
  }
  if (!free_slots.empty()) {
    out = free_slots.back();
    free_slots.pop_back();
    return ready_event(out);
  }
  // at most 64 MiB per slab (large GF(2^16) codecs have multi-MB slots)
  const size_t n_slots = std::max<size_t>(1, std::min(next_slab, (size_t(64) << 20) / slot_bytes));
  void* slab = nullptr;
  hipError_t e = hipMalloc(&slab, n_slots * slot_bytes);
  if (e != hipSuccess) return e;
  slabs.push_back(slab);
  for (size_t i = 0; i < n_slots; ++i) free_slots.push_back(Slot{(uint8_t*)slab + i * slot_bytes, nullptr});
  next_slab = std::min(next_slab * 2, kMaxSlab);
  out = free_slots.back();
  free_slots.pop_back();
  return ready_event(out);
}

void DecodePool::release(const Slot& s) {
  std::lock_guard<std::mutex> g(mu);
  retired.push_back(s);
}

DecodeEntry::~DecodeEntry() {
  if (pool && slot.mem) {
    pool->release(slot);
  } else if (own && slot.ready) {
    (void)hipEventSynchronize(slot.ready);  // the buffer is freed below
    (void)hipEventDestroy(slot.ready);
  }
}

hipError_t DeviceBuffer::alloc(size_t n) {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(&p, n);
  if (e == hipSuccess) bytes = n;
  return e;
}

// ------------------------------------------------------- field abstraction
struct Field {
  bool tower;
  uint16_t one() const { return 1; }
  uint16_t mul(uint16_t a, uint16_t b) const { return tower ? gf16_mul(a, b) : gf8_mul((uint8_t)a, (uint8_t)b); }
  uint16_t div(uint16_t a, uint16_t b) const { return tower ? gf16_div(a, b) : gf8_div((uint8_t)a, (uint8_t)b); }
  uint16_t exp(uint16_t a, size_t n) const { return tower ? gf16_exp(a, n) : gf8_exp((uint8_t)a, n); }
  uint16_t nth(size_t n) const { return tower ? (uint16_t)n : (uint16_t)(uint8_t)n; }
  size_t order() const { return tower ? 65536 : 256; }
};

// Gauss-Jordan elimination with the first-nonzero-below pivot rule
// (src/matrix.rs:373-422).  Returns false for a singular matrix.
static bool gauss_jordan(const Field& F, uint16_t* w, int rows, int cols) {
  for (int r = 0; r < rows; ++r) {
    if (w[r * cols + r] == 0)
      for (int rb = r + 1; rb < rows; ++rb)
        if (w[rb * cols + r] != 0) {
          for (int c = 0; c < cols; ++c) std::swap(w[r * cols + c], w[rb * cols + c]);
          break;
        }
    if (w[r * cols + r] == 0) return false;
    if (w[r * cols + r] != 1) {
      uint16_t s = F.div(1, w[r * cols + r]);
      for (int c = 0; c < cols; ++c) w[r * cols + c] = F.mul(s, w[r * cols + c]);
    }
    for (int rb = 0; rb < rows; ++rb) {
      if (rb == r) continue;
      uint16_t f = w[rb * cols + r];
      if (f)
        for (int c = 0; c < cols; ++c) w[rb * cols + c] ^= F.mul(f, w[r * cols + c]);
    }
  }
  return true;
}

static bool invert(const Field& F, const std::vector<uint16_t>& a, int n, std::vector<uint16_t>& out) {
  std::vector<uint16_t> w((size_t)n * 2 * n, 0);
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < n; ++c) w[(size_t)r * 2 * n + c] = a[(size_t)r * n + c];
    w[(size_t)r * 2 * n + n + r] = 1;
  }
  if (!gauss_jordan(F, w.data(), n, 2 * n)) return false;
  out.assign((size_t)n * n, 0);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) out[(size_t)r * n + c] = w[(size_t)r * 2 * n + n + c];
  return true;
}

static void vandermonde(const Field& F, int rows, int cols, std::vector<uint16_t>& v) {
  v.assign((size_t)rows * cols, 0);
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) v[(size_t)r * cols + c] = F.exp(F.nth(r), c);
}

// Encoding matrices (src/core/codec.rs:67-258).
static int build_matrix(const Field& F, int mode, bool fast_one, int k, int m,
                        std::vector<uint16_t>& out) {
  const int total = k + m;
  out.assign((size_t)total * k, 0);
  if (fast_one && m == 1) {  // codec.rs:240-250
    for (int c = 0; c < k; ++c) out[(size_t)c * k + c] = out[(size_t)k * k + c] = 1;
    return RSGPU_OK;
  }
  if (mode == RSGPU_MATRIX_VANDERMONDE) {  // V * inv(V_top), codec.rs:67-89
    std::vector<uint16_t> v, top((size_t)k * k), inv;
    vandermonde(F, total, k, v);
    std::copy(v.begin(), v.begin() + (size_t)k * k, top.begin());
    if (!invert(F, top, k, inv)) return RSGPU_INVALID_CUSTOM_MATRIX;
    for (int r = 0; r < total; ++r)
      for (int c = 0; c < k; ++c) {
        uint16_t s = 0;
        for (int i = 0; i < k; ++i) s ^= F.mul(v[(size_t)r * k + i], inv[(size_t)i * k + c]);
        out[(size_t)r * k + c] = s;
      }
    return RSGPU_OK;
  }
  if (mode == RSGPU_MATRIX_CAUCHY) {  // codec.rs:91-112
    for (int r = 0; r < total; ++r) {
      if (r < k) {
        out[(size_t)r * k + r] = 1;
        continue;
      }
      for (int c = 0; c < k; ++c) {
        uint16_t den = F.nth(r) ^ F.nth(c);
        if (den == 0) return RSGPU_INVALID_CUSTOM_MATRIX;
        out[(size_t)r * k + c] = F.div(1, den);
      }
    }
    return RSGPU_OK;
  }
  if (mode == RSGPU_MATRIX_JERASURE_LIKE) {  // codec.rs:114-197
    std::vector<uint16_t>& vm = out;
    vandermonde(F, total, k, vm);
    auto at = [&](int r, int c) -> uint16_t& { return vm[(size_t)r * k + c]; };
    at(0, 0) = 1;
    for (int i = 1; i < k; ++i) at(0, i) = 0;
    for (int i = 0; i + 1 < k; ++i) at(total - 1, i) = 0;
    at(total - 1, k - 1) = 1;
    for (int i = 0; i < k; ++i) {
      int r = i;
      while (r < total && at(r, i) == 0) ++r;
      if (r != i) {
        if (r >= total) return RSGPU_INVALID_CUSTOM_MATRIX;
        for (int c = 0; c < k; ++c) std::swap(at(r, c), at(i, c));
      } else if (at(i, i) == 0) {
        return RSGPU_INVALID_CUSTOM_MATRIX;
      }
      uint16_t diag = at(i, i);
      if (diag == 0) return RSGPU_INVALID_CUSTOM_MATRIX;
      uint16_t scale = diag != 1 ? F.div(1, diag) : 1;
      if (scale != 1)
        for (int row = 0; row < total; ++row) at(row, i) = F.mul(at(row, i), scale);
      for (int j = 0; j < k; ++j) {
        uint16_t value = at(i, j);
        if (j != i && value != 0)
          for (int row = 0; row < total; ++row) at(row, j) ^= F.mul(value, at(row, i));
      }
    }
    for (int j = 0; j < k; ++j) {
      uint16_t value = at(k, j);
      if (value == 0) return RSGPU_INVALID_CUSTOM_MATRIX;
      if (value != 1) {
        uint16_t s = F.div(1, value);
        for (int row = k; row < total; ++row) at(row, j) = F.mul(at(row, j), s);
      }
    }
    for (int row = k + 1; row < total; ++row) {
      uint16_t value = at(row, 0);
      if (value == 0) return RSGPU_INVALID_CUSTOM_MATRIX;
      if (value != 1) {
        uint16_t s = F.div(1, value);
        for (int col = 0; col < k; ++col) at(row, col) = F.mul(at(row, col), s);
      }
    }
    return RSGPU_OK;
  }
  return RSGPU_INVALID_CUSTOM_MATRIX;
}

static size_t derive_cache_capacity(size_t k, size_t m) {  // codec.rs:44-59
  size_t h = (k + m) * std::max<size_t>(m, 1) * 2, p = 1;
  while (p < h) p <<= 1;
  return std::min<size_t>(std::max<size_t>(p, 128), 4096);
}

static int upload_codec(Codec& c) {
  HIP_TRY(hipGetDevice(&c.device));
  if (c.family != RSGPU_FAMILY_CLASSIC) return init_leopard(c);
  HIP_TRY(c.d_matrix.alloc(c.matrix.size() * sizeof(uint16_t)));
  HIP_TRY(hipMemcpy(c.d_matrix.p, c.matrix.data(), c.matrix.size() * sizeof(uint16_t),
                    hipMemcpyHostToDevice));
  build_tables_host(c.matrix.data() + c.k * c.k, (int)c.m, (int)c.k, c.field == RSGPU_FIELD_GF16,
                    c.h_parity_tables, c.ftab.get());
  if (c.ftab) {
    HIP_TRY(c.d_ftab.alloc(768));
    HIP_TRY(hipMemcpy(c.d_ftab.p, c.ftab->log, 256, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy((uint8_t*)c.d_ftab.p + 256, c.ftab->exp, 512, hipMemcpyHostToDevice));
  }
  HIP_TRY(c.d_parity_tables.alloc(c.h_parity_tables.size() * 4));
  HIP_TRY(hipMemcpy(c.d_parity_tables.p, c.h_parity_tables.data(), c.h_parity_tables.size() * 4,
                    hipMemcpyHostToDevice));
  return RSGPU_OK;
}

// ------------------------------------------------------------- apply plumbing
struct Views {
  std::vector<const uint8_t*> in;
  std::vector<uint8_t*> out;
};

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// Apply rows (tables [r][k_total][W], row stride k_total*W) to inputs -> outputs.
// len is in bytes.  Splits k into kMaxIn chunks (later chunks accumulate) and r
// into kMaxOut chunks.  mode kVerify compares with `out` and sets *flag.
// `table_cols` (default k_total) is the column count of the table rows and
// `col0` the table column of in[0], so a subset of a wider table can be applied.
// mode kAccumulate XORs into the outputs from the first input on.
// Takes plain arrays of k_total inputs and r_total outputs and allocates nothing.
// True when one launch of the vector kernel covers the whole call with its
// tables inline in the kernel arguments: 16-byte aligned buffers and length, at
// most kMaxIn inputs and kMaxOut outputs, and r x k tables within kInlineTabWords.
static bool apply_inline_fits(int k_total, int r_total, bool tower, const uint8_t* const* in,
                              uint8_t* const* out, size_t len) {
  if (len % 16 || k_total > kMaxIn || r_total > kMaxOut ||
      (size_t)k_total * r_total * words_per_coef(tower) > (size_t)kInlineTabWords)
    return false;
  for (int i = 0; i < k_total; ++i)
    if (!aligned16(in[i])) return false;
  for (int j = 0; j < r_total; ++j)
    if (!aligned16(out[j])) return false;
  return true;
}

// The in-kernel flag clear of one verify call (verify_claim in kernel_common.h).
struct VerifyClear {
  uint64_t* ws;
  uint64_t id;
};

static int run_apply(const uint32_t* d_tables, int k_total, int r_total, bool tower,
                     const uint8_t* const* in, uint8_t* const* out,
                     size_t len, int mode, int32_t* flag, hipStream_t s, int variant = -1,
                     int table_cols = -1, int col0 = 0, const uint32_t* h_tables = nullptr,
                     const VerifyClear* vc = nullptr, bool leo16 = false) {
  if (table_cols < 0) table_cols = k_total;
  const int W = words_per_coef(tower);
  bool vec_ok = true;
  for (int i = 0; i < k_total; ++i) vec_ok &= aligned16(in[i]);
  for (int j = 0; j < r_total; ++j) vec_ok &= aligned16(out[j]);
  // Leopard GF(2^16) split layout (tower tables): vector tiles only, no byte tail
  if (leo16 && (!tower || !vec_ok || len % 16)) return RSGPU_ERR_UNSUPPORTED;
  if (mode == kVerify && k_total > kMaxIn) return RSGPU_ERR_UNSUPPORTED;  // handled by caller
  const uint64_t n16 = vec_ok ? len / 16 : 0;
  const size_t tail_begin = (size_t)n16 * 16;
  // host tables (same layout as d_tables) travel inline when one launch covers the call
  const bool inline_tabs = h_tables && apply_inline_fits(k_total, r_total, tower, in, out, len);
  if (!inline_tabs && !d_tables) return RSGPU_ERR_UNSUPPORTED;
  // the byte kernel alone does not clear a verify flag
  if (mode == kVerify && vc && !n16) HIP_TRY(launch_fill_i32(flag, 0, 1, s));
  // Default tiling: GF(2^8) one dword per lane (apply_dword_kernel).  GF(2^16)
  // elements take four byte maps per coefficient, so a workgroup holding every
  // output runs short of registers (two waves per SIMD at 4 outputs, one at 8) and
  // of workgroups on short shards: one output per workgroup up to 256 KiB
  // (64-lane tiles up to 64 KiB), two above; two vectors per lane from 16 MiB.
  // Measured per call (back to back) on MI355X against every output per
  // workgroup: 20+8 encode at 16 KiB 6.5 us against 23.6, 10+4 at 256 KiB 4.8
  // against 8.5, 10+4 at 1 MiB 8.1 against 9.2; at 4 MiB within the run-to-run
  // spread (about 10%) of the fastest tiling.
  int tile, split;
  if (variant < 0 || variant >= kNumApplyVariants) {
    if (tower) {
      tile = len <= ((size_t)64 << 10) ? 2 : (len >= ((size_t)16 << 20) ? 1 : 0);
      split = len <= ((size_t)256 << 10) ? 1 : 2;
    } else {
      tile = kTileDword;
      split = 0;
    }
  } else {
    tile = variant % kNumTiles;
    split = kRowSplit[variant / kNumTiles];
  }
  for (int r0 = 0; r0 < r_total; r0 += kMaxOut) {
    const int rn = std::min(kMaxOut, r_total - r0);
    for (int k0 = 0; k0 < k_total; k0 += kMaxIn) {
      const int kn = std::min(kMaxIn, k_total - k0);
      const int m = (mode == kVerify) ? kVerify : (k0 == 0 && mode == kWrite ? kWrite : kAccumulate);
      if (n16) {
        ApplyArgs a;
        std::memset(&a, 0, offsetof(ApplyArgs, tab_inline));
        for (int i = 0; i < kn; ++i) a.in[i] = in[k0 + i];
        for (int j = 0; j < rn; ++j) a.out[j] = out[r0 + j];
        a.flag = flag;
        a.n16 = n16;
        a.v_begin = 0;
        a.k = kn;
        a.r = rn;
        if (mode == kVerify && vc) {
          a.clear_ws = vc->ws;
          a.call_id = vc->id;
        }
        if (inline_tabs) {  // this launch's rn x kn block, rows kn * W apart
          for (int j = 0; j < rn; ++j)
            std::memcpy(a.tab_inline + (size_t)j * kn * W,
                        h_tables + ((size_t)(r0 + j) * table_cols + col0 + k0) * W, (size_t)kn * W * 4);
          a.tables = nullptr;
          a.tab_row_stride = kn * W;
          a.tab_col0 = 0;
        } else {
          a.tables = d_tables + (size_t)r0 * table_cols * W;
          a.tab_row_stride = table_cols * W;
          a.tab_col0 = col0 + k0;
        }
        HIP_TRY(launch_apply(a, m, tile, split ? std::min(split, rn) : rn, leo16 ? 2 : (tower ? 1 : 0), s));
      }
      if (tail_begin < len) {
        ByteArgs b{};
        for (int i = 0; i < kn; ++i) b.in[i] = in[k0 + i];
        for (int j = 0; j < rn; ++j) b.out[j] = out[r0 + j];
        b.tables = d_tables + (size_t)r0 * table_cols * W;
        b.flag = flag;
        b.begin = tail_begin;
        b.end = len;
        b.k = kn;
        b.tab_row_stride = table_cols * W;
        b.tab_col0 = col0 + k0;
        b.r = rn;
        b.mode = m;
        b.tower = tower ? 1 : 0;
        HIP_TRY(launch_apply_bytes(b, s));
      }
    }
  }
  return RSGPU_OK;
}

// Per (device, stream): the partial-sum rows of split-over-inputs applies.  Calls
// on one stream reuse it in stream order (the slot's mutex is held until a call's
// launches are enqueued); it only grows, after waiting for its stream, the only
// one that uses it.  Kept for the life of the process.
struct PartSlot {
  std::mutex mu;
  DeviceBuffer buf;
};
static std::mutex g_part_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<PartSlot>>* g_part_slots =
    new std::map<std::pair<int, hipStream_t>, std::unique_ptr<PartSlot>>();

static PartSlot* part_slot(hipStream_t s) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(g_part_mu);
  auto& slot = (*g_part_slots)[std::make_pair(device, s)];
  if (!slot) slot = std::make_unique<PartSlot>();
  return slot.get();
}

// out (r_total rows) = M (r_total x k_total, tables as run_apply) applied to
// k_total > kMaxIn inputs: one launch per row group in which workgroup column z
// takes inputs [z * kMaxIn, z * kMaxIn + kMaxIn) from a staged pointer array and
// writes partial rows, then xor_partials folds the chunks into the outputs.  A
// long, narrow apply (thousands of inputs, a few outputs) then fills the chip
// instead of running kMaxIn-input launches one after another.  Vector tiles only:
// 16-byte aligned buffers, len a multiple of 16.  elem: 0 GF(2^8), 1 tower, 2
// split-layout GF(2^16).
static int run_apply_splitk(const uint32_t* d_tables, int k_total, int r_total, int elem,
                            const uint8_t* const* in, uint8_t* const* out, size_t len, hipStream_t s,
                            int variant) {
  const int W = words_per_coef(elem != 0);
  const int chunks = (k_total + kMaxIn - 1) / kMaxIn;
  int tile = 0, split = 0;
  if (variant >= 0 && variant < kNumApplyVariants) {
    tile = variant % kNumTiles;
    split = kRowSplit[variant / kNumTiles];
  }
  PartSlot* ps = part_slot(s);
  if (!ps) return RSGPU_ERR_HIP;
  std::lock_guard<std::mutex> g(ps->mu);
  const size_t pitch = (len + 255) / 256 * 256;
  const size_t need = (size_t)chunks * std::min(kMaxOut, r_total) * pitch;
  if (ps->buf.bytes < need) {
    HIP_TRY(hipStreamSynchronize(s));  // the old rows may still be read on s
    if (ps->buf.p) {
      (void)hipFree(ps->buf.p);
      ps->buf.p = nullptr;
      ps->buf.bytes = 0;
    }
    HIP_TRY(ps->buf.alloc(need));
  }
  std::vector<const uint8_t*> ptrs(in, in + k_total);
  const uint8_t* const* d_in = nullptr;
  size_t off = 0, bytes = 0;
  int rc = stage_pointer_table(ptrs, s, &d_in, &off, &bytes);
  if (rc) return rc;
  for (int r0 = 0; r0 < r_total; r0 += kMaxOut) {
    const int rn = std::min(kMaxOut, r_total - r0);
    ApplyArgs a;
    std::memset(&a, 0, offsetof(ApplyArgs, tab_inline));
    for (int j = 0; j < rn; ++j) a.out[j] = out[r0 + j];
    a.in[0] = in[0];
    a.n16 = len / 16;
    a.k = kMaxIn;
    a.r = rn;
    a.tables = d_tables + (size_t)r0 * k_total * W;
    a.tab_row_stride = k_total * W;
    a.in_dev = d_in;
    a.part = (uint8_t*)ps->buf.p;
    a.part_pitch = pitch;
    a.k_total = k_total;
    HIP_TRY(launch_apply(a, kWrite, tile, split ? std::min(split, rn) : rn, elem, s));
    HIP_TRY(launch_xor_partials((const uint8_t*)ps->buf.p, pitch, chunks, rn, out + r0, len, false, s));
  }
  return release_pointer_table(off, bytes, s);
}

static int run_apply(const uint32_t* d_tables, int k_total, int r_total, bool tower,
                     const std::vector<const uint8_t*>& in, const std::vector<uint8_t*>& out,
                     size_t len, int mode, int32_t* flag, hipStream_t s, int variant = -1,
                     int table_cols = -1, int col0 = 0) {
  return run_apply(d_tables, k_total, r_total, tower, in.data(), out.data(), len, mode, flag, s, variant,
                   table_cols, col0);
}

// ---------------------------------------------------------------- validation
static int elem_bytes(const Codec& c) { return c.field == RSGPU_FIELD_GF16 ? 2 : 1; }

// check_slices!(multi) (macros.rs:142-155) for a run of lengths.
template <class S>
static int check_slices(const S* v, size_t n) {
  size_t size = v[0].len;
  if (size == 0) return RSGPU_EMPTY_SHARD;
  for (size_t i = 0; i < n; ++i)
    if (v[i].len != size) return RSGPU_INCORRECT_SHARD_SIZE;
  return RSGPU_OK;
}

static int check_elem_len(const Codec& c, size_t len) {
  if (c.field == RSGPU_FIELD_GF16 && (len & 1)) return RSGPU_INCORRECT_SHARD_SIZE;
  return RSGPU_OK;
}

static thread_local DeviceBuffer t_flag;
static thread_local int t_flag_device = -1;

static int32_t* thread_flag(int device) {
  if (t_flag_device != device || !t_flag.p) {
    if (t_flag.alloc(64) != hipSuccess) return nullptr;
    t_flag_device = device;
  }
  return (int32_t*)t_flag.p;
}

// ------------------------------------------------------------ variant dispatch
constexpr int kPick = -2;  // impl argument: take the variant from the codec's dispatch state
constexpr size_t kLeo16MatrixErasures = 32;

static bool variant_valid(const Codec& c, int op, size_t erasures, int v) {
  if (v < 0 || v >= kNumVariants) return false;
  if (c.family == RSGPU_FAMILY_CLASSIC) return v < kNumApplyVariants;
  if (v >= kNumApplyVariants) return true;  // Leopard FFT column tiles
  // Leopard GF(2^8): matrix variants for every operation; reconstruct applies
  // the generator-matrix solve (exactly k shards present) or the FFT decoder's
  // own matrix (more), both bit-identical to the FFT decoder.  Leopard GF(2^16):
  // reconstruct of at most kLeo16MatrixErasures shards, through the FFT decoder's
  // own matrix (its apply grows with erasures x present shards, the FFT does not:
  // past a few dozen erasures the FFT is faster at every shape).
  return c.family == RSGPU_FAMILY_LEOPARD_GF8 ||
         (c.family == RSGPU_FAMILY_LEOPARD_GF16 && op == kOpReconstruct && erasures <= kLeo16MatrixErasures);
}

static int size_bucket(size_t len) {
  int b = (int)std::lround(2.0 * std::log2((double)std::max<size_t>(len, 1))) - kBucket0;
  return std::min(std::max(b, 0), kNumBuckets - 1);
}

// The default choice when neither a forced variant nor a dispatch-table entry
// applies.  A Leopard reconstruct of 1..4 erasures runs its matrix form (the
// decoder's own matrix, or the [I; G] solve with exactly k present) with 64-lane
// tiles and one output per workgroup, since the FFT decoder costs the same for
// one erasure as for m.  In the MI355X timing set that variant beat the FFT on
// all 31 such GF(2^16) calls (10+4 to 4096+1024, 5-260 KiB shards; 1.24x at
// least, 2.69x geometric mean) and on 44 of 46 GF(2^8) calls (2.06x geometric
// mean, 0.87x at worst).  Everything else takes the default tiling (-1): the
// matrix encode and verify lose to the FFT on most Leopard shapes.
constexpr int kLeoFewErasureVariant = 12;
static int default_variant(const Codec& c, int op, size_t erasures) {
  if (c.family != RSGPU_FAMILY_CLASSIC && op == kOpReconstruct && erasures >= 1 && erasures <= 4)
    return kLeoFewErasureVariant;
  return -1;
}

// The variant for one call: the forced one, else the dispatch table's entry,
// else (no table, an entry of -1, or one not exact here) the default choice.
static int pick_variant(const Codec& c, int op, size_t erasures, size_t len) {
  int v = c.forced_variant.load(std::memory_order_relaxed);
  if (v < 0) {
    std::shared_ptr<const VariantTable> t = std::atomic_load(&c.variant_table);
    if (t && t->n_erasures > 0) {
      const size_t e = std::min<size_t>(erasures, (size_t)t->n_erasures - 1);
      v = t->v[((size_t)op * t->n_erasures + e) * kNumBuckets + size_bucket(len)];
    }
  }
  return variant_valid(c, op, erasures, v) ? v : default_variant(c, op, erasures);
}

static bool is_matrix_variant(const Codec& c, int v) {
  return c.family == RSGPU_FAMILY_LEOPARD_GF8 && v >= 0 && v < kNumApplyVariants;
}

// Classic codec over a Leopard GF(2^8) codec's generator matrix [I; G], in
// Leopard's byte field.  G is read off one FFT encode of unit columns: data
// shard i holds the field's identity at byte i (zeros elsewhere), so parity
// shard j holds G[j][i] at byte i.
static int leo_matrix_codec(const Codec& c, hipStream_t s, std::shared_ptr<Codec>& out) {
  std::lock_guard<std::mutex> g(c.mu);
  if (c.leo_matrix) {
    out = c.leo_matrix;
    return RSGPU_OK;
  }
  const size_t k = c.k, m = c.m, len = (k + 63) / 64 * 64;
  DeviceBuffer buf;
  HIP_TRY(buf.alloc((k + m) * len));
  std::shared_ptr<const FieldTables> ft = leopard8_field_tables();
  std::vector<uint8_t> host((k + m) * len, 0);
  for (size_t i = 0; i < k; ++i) host[i * len + i] = ft->one;
  HIP_TRY(hipMemcpyAsync(buf.p, host.data(), k * len, hipMemcpyHostToDevice, s));
  std::vector<rsgpu_cslice> data(k);
  std::vector<rsgpu_slice> par(m);
  for (size_t i = 0; i < k; ++i) data[i] = {(const uint8_t*)buf.p + i * len, len};
  for (size_t j = 0; j < m; ++j) par[j] = {(uint8_t*)buf.p + (k + j) * len, len};
  int rc = leopard_encode(c, data.data(), par.data(), len, s, -1);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(host.data() + k * len, (uint8_t*)buf.p + k * len, m * len,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  auto mc = std::make_shared<Codec>();
  mc->k = k;
  mc->m = m;
  mc->field = RSGPU_FIELD_GF8;
  mc->family = RSGPU_FAMILY_CLASSIC;
  mc->matrix_mode = RSGPU_MATRIX_VANDERMONDE;  // MDS: every k survivors solve
  mc->use_cache = c.use_cache;
  mc->cache_capacity = c.cache_capacity;
  mc->cache_max_bytes = c.cache_max_bytes;
  mc->ftab = ft;
  mc->matrix.assign((k + m) * k, 0);
  for (size_t i = 0; i < k; ++i) mc->matrix[i * k + i] = ft->one;
  for (size_t j = 0; j < m; ++j)
    for (size_t i = 0; i < k; ++i) mc->matrix[(k + j) * k + i] = host[(k + j) * len + i];
  if ((rc = upload_codec(*mc))) return rc;
  c.leo_matrix = mc;
  out = mc;
  return RSGPU_OK;
}

// ------------------------------------------------------------ encode / verify
static int encode_sep_impl(const Codec& c, const rsgpu_cslice* data, size_t nd,
                           const rsgpu_slice* parity, size_t np, hipStream_t s, int variant = kPick) {
  if (nd < c.k) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (nd > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  if (np < c.m) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (np > c.m) return RSGPU_TOO_MANY_PARITY_SHARDS;
  int rc = check_slices(data, nd);
  if (rc) return rc;
  if ((rc = check_slices(parity, np))) return rc;
  if (data[0].len != parity[0].len) return RSGPU_INCORRECT_SHARD_SIZE;
  if ((rc = check_elem_len(c, data[0].len))) return rc;
  for (size_t i = 0; i < nd; ++i)
    if (!data[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < np; ++i)
    if (!parity[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  const size_t len = data[0].len;
  if (variant == kPick) variant = pick_variant(c, kOpEncode, 0, len);
  if (is_matrix_variant(c, variant)) {
    if (len % 64) return RSGPU_INCORRECT_SHARD_SIZE;  // leopard.rs:50-56
    std::shared_ptr<Codec> mc;
    if ((rc = leo_matrix_codec(c, s, mc))) return rc;
    return encode_sep_impl(*mc, data, nd, parity, np, s, variant);
  }
  if (c.family != RSGPU_FAMILY_CLASSIC) return leopard_encode(c, data, parity, len, s, variant);
  std::vector<const uint8_t*> in(nd);
  std::vector<uint8_t*> out(np);
  for (size_t i = 0; i < nd; ++i) in[i] = data[i].ptr;
  for (size_t i = 0; i < np; ++i) out[i] = parity[i].ptr;
  return run_apply((const uint32_t*)c.d_parity_tables.p, (int)c.k, (int)c.m,
                   c.field == RSGPU_FIELD_GF16, in, out, len, kWrite, nullptr, s, variant);
}

// ------------------------------------------------------------ verify tickets
// Per (device, stream): the device workspace of the in-kernel flag protocol
// (words: last id claimed, cleared, reported; one padding word) and the
// stream's next call id.  Work on one stream runs in order, and the slot's mutex
// is held from taking an id until the call's launches are enqueued, so ids reach
// the device in increasing order.  Workspaces come from 32 KiB device chunks
// zeroed once, kept for the life of the process.
struct VerifySlot {
  std::mutex mu;
  uint64_t* ws = nullptr;
  uint64_t next = 0;
};
constexpr size_t kVerifyChunkSlots = 1024;
static std::mutex g_verify_mu;
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<VerifySlot>> g_verify_slots;
static std::vector<std::tuple<int, uint64_t*, size_t>> g_verify_chunks;  // device, base, slots used

static VerifySlot* verify_slot(hipStream_t s) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return nullptr;
  // the calling thread's last (device, stream): no global lock on a repeat
  thread_local int t_device = -1;
  thread_local hipStream_t t_stream = nullptr;
  thread_local VerifySlot* t_slot = nullptr;
  if (t_slot && t_device == device && t_stream == s) return t_slot;
  std::lock_guard<std::mutex> g(g_verify_mu);
  const auto key = std::make_pair(device, s);
  auto it = g_verify_slots.find(key);
  if (it != g_verify_slots.end()) {
    t_device = device, t_stream = s, t_slot = it->second.get();
    return t_slot;
  }
  std::tuple<int, uint64_t*, size_t>* chunk = nullptr;
  for (auto& ch : g_verify_chunks)
    if (std::get<0>(ch) == device && std::get<2>(ch) < kVerifyChunkSlots) chunk = &ch;
  if (!chunk) {
    void* p = nullptr;
    if (hipMalloc(&p, kVerifyChunkSlots * 32) != hipSuccess) return nullptr;
    if (hipMemset(p, 0, kVerifyChunkSlots * 32) != hipSuccess) return nullptr;  // once per chunk
    g_verify_chunks.emplace_back(device, (uint64_t*)p, 0);
    chunk = &g_verify_chunks.back();
  }
  auto slot = std::make_unique<VerifySlot>();
  slot->ws = std::get<1>(*chunk) + 4 * std::get<2>(*chunk)++;  // words: claimed, cleared, reported
  VerifySlot* out = slot.get();
  g_verify_slots.emplace(key, std::move(slot));
  t_device = device, t_stream = s, t_slot = out;
  return out;
}

static int verify_impl(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* d_flag,
                       hipStream_t s, int variant = kPick) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  if ((rc = check_elem_len(c, shards[0].len))) return rc;
  for (size_t i = 0; i < n; ++i)
    if (!shards[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  if (variant == kPick) variant = pick_variant(c, kOpVerify, 0, shards[0].len);
  if (is_matrix_variant(c, variant)) {
    if (shards[0].len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
    std::shared_ptr<Codec> mc;
    if ((rc = leo_matrix_codec(c, s, mc))) return rc;
    return verify_impl(*mc, shards, n, d_flag, s, variant);
  }
  std::vector<const uint8_t*> in(c.k);
  std::vector<uint8_t*> out(c.m);
  for (size_t i = 0; i < c.k; ++i) in[i] = shards[i].ptr;
  for (size_t i = 0; i < c.m; ++i) out[i] = const_cast<uint8_t*>(shards[c.k + i].ptr);
  const bool tower = c.field == RSGPU_FIELD_GF16;
  if (c.family == RSGPU_FAMILY_CLASSIC && c.k <= (size_t)kMaxIn) {
    // one launch clears the flag and compares (verify_claim / verify_report)
    if (VerifySlot* vs = verify_slot(s)) {
      std::lock_guard<std::mutex> g(vs->mu);
      const VerifyClear vc{vs->ws, ++vs->next};
      return run_apply((const uint32_t*)c.d_parity_tables.p, (int)c.k, (int)c.m, tower, in.data(), out.data(),
                       shards[0].len, kVerify, d_flag, s, variant, -1, 0, nullptr, &vc);
    }
  }
  HIP_TRY(launch_fill_i32(d_flag, 0, 1, s));
  if (c.family != RSGPU_FAMILY_CLASSIC)
    return leopard_verify(c, shards, shards[0].len, d_flag, s, variant);
  if (c.k <= (size_t)kMaxIn)
    return run_apply((const uint32_t*)c.d_parity_tables.p, (int)c.k, (int)c.m, tower, in, out,
                     shards[0].len, kVerify, d_flag, s, variant);
  // Wide codecs: recompute into scratch, then compare (two passes).
  const size_t len = shards[0].len;
  DeviceBuffer scratch;
  HIP_TRY(scratch.alloc(len * c.m));
  std::vector<uint8_t*> tmp(c.m);
  for (size_t j = 0; j < c.m; ++j) tmp[j] = (uint8_t*)scratch.p + j * len;
  rc = run_apply((const uint32_t*)c.d_parity_tables.p, (int)c.k, (int)c.m, tower, in, tmp, len,
                 kWrite, nullptr, s, variant);
  if (rc) return rc;
  // compare the recomputed parity with the stored parity, kMaxOut rows a launch
  for (size_t r0 = 0; r0 < c.m; r0 += kMaxOut) {
    CompareArgs ca{};
    ca.rows = (int32_t)std::min<size_t>(kMaxOut, c.m - r0);
    for (int j = 0; j < ca.rows; ++j) ca.a[j] = tmp[r0 + j], ca.b[j] = out[r0 + j];
    ca.len = len;
    ca.flag = d_flag;
    HIP_TRY(launch_compare_rows(ca, s));
  }
  HIP_TRY(hipStreamSynchronize(s));  // the scratch buffer is freed on return
  return RSGPU_OK;
}

// ------------------------------------------------------------ reconstruction
struct Plan {
  std::vector<int> valid;    // first k present
  std::vector<int> targets;  // missing requested: data ascending, then parity ascending
  size_t shard_len = 0;
  bool all_present = false;
};

// with_classic_reconstruct_plan / fill_classic_reconstruct_shape_from_suffix
// (recovery_plan.rs:262-387) and the target selection of
// fill_classic_reconstruct_plan (recovery_plan.rs:425-517).
// Index list of fixed capacity for the planner of the allocation-free entry point.
struct FixedIdx {
  static constexpr size_t kCap = 256;
  int v[kCap];
  size_t n = 0;
  void push_back(int x) { v[n++] = x; }
  size_t size() const { return n; }
  int operator[](size_t i) const { return v[i]; }
  const int* begin() const { return v; }
  const int* end() const { return v + n; }
};

// The planner over any index list (std::vector<int>, or FixedIdx for codecs of at
// most FixedIdx::kCap shards).
template <class V>
static int plan_shards(const Codec& c, const rsgpu_slot* shards, size_t n, int goal,
                       const uint8_t* required, size_t n_required, V& valid, V& targets,
                       size_t& shard_len, bool& all_present, bool check_buffers = true) {
  const size_t total = c.k + c.m;
  if (goal == 2 && n_required != c.k && n_required != total) return RSGPU_INVALID_SHARD_FLAGS;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  size_t present = 0;
  shard_len = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!shards[i].present) continue;
    if (shards[i].len == 0) return RSGPU_EMPTY_SHARD;
    if (shard_len != 0 && shards[i].len != shard_len) return RSGPU_INCORRECT_SHARD_SIZE;
    shard_len = shards[i].len;
    ++present;
  }
  if (present == total) {
    all_present = true;
    return RSGPU_OK;
  }
  if (present < c.k) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  int rc = check_elem_len(c, shard_len);
  if (rc) return rc;
  for (size_t i = 0; i < n && valid.size() < c.k; ++i)
    if (shards[i].present) valid.push_back((int)i);
  for (size_t i = 0; i < n; ++i) {
    if (shards[i].present) continue;
    bool want;
    if (goal == 2) want = i < n_required && required[i];
    else want = (i < c.k) || goal == 0;
    if (want) targets.push_back((int)i);
  }
  if (!check_buffers) return RSGPU_OK;
  for (int t : targets) {
    if (!shards[t].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
    if (shards[t].len != shard_len) return RSGPU_INCORRECT_SHARD_SIZE;
  }
  for (int v : valid)
    if (!shards[v].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  return RSGPU_OK;
}

static int make_plan(const Codec& c, const rsgpu_slot* shards, size_t n, int goal,
                     const uint8_t* required, size_t n_required, Plan& plan) {
  return plan_shards(c, shards, n, goal, required, n_required, plan.valid, plan.targets, plan.shard_len,
                     plan.all_present);
}

static std::string plan_key(const Codec& c, const rsgpu_slot* shards, const Plan& plan) {
  std::string key(c.k + c.m, '0');
  for (size_t i = 0; i < c.k + c.m; ++i) key[i] = shards[i].present ? '1' : '0';
  for (int t : plan.targets) key[t] = 't';
  return key;
}

// Host-side invertibility check of the missing-data system, used only for
// custom matrices: the library's own constructions are MDS, so every survivor
// set is solvable, but a user matrix may not be, and the reference reports that
// as InvalidCustomMatrix before any output is written (reconstruct.rs:1130-1160).
template <class V>
static bool custom_solvable_valid(const Codec& c, const V& valid) {
  const Field F{c.field == RSGPU_FIELD_GF16};
  int p = 0;
  while (p < (int)c.k && valid[p] < (int)c.k) ++p;
  int e = (int)c.k - p;
  if (e == 0) return true;
  std::vector<int> missing;
  for (int d = 0, kn = 0; d < (int)c.k; ++d) {
    if (kn < p && valid[kn] == d) ++kn;
    else missing.push_back(d);
  }
  std::vector<uint16_t> w((size_t)e * e);
  for (int r = 0; r < e; ++r)
    for (int col = 0; col < e; ++col) w[(size_t)r * e + col] = c.matrix[(size_t)valid[p + r] * c.k + missing[col]];
  return gauss_jordan(F, w.data(), e, e);
}

static bool custom_solvable(const Codec& c, const Plan& plan) { return custom_solvable_valid(c, plan.valid); }

// The same check on the stack (at most kMaxSolve missing data shards), for the
// allocation-free entry point.
static bool custom_solvable_fixed(const Codec& c, const FixedIdx& valid) {
  const Field F{c.field == RSGPU_FIELD_GF16};
  int p = 0;
  while (p < (int)c.k && valid[p] < (int)c.k) ++p;
  const int e = (int)c.k - p;
  if (e == 0) return true;
  if (e > kMaxSolve) return false;
  int missing[kMaxSolve];
  for (int d = 0, kn = 0, j = 0; d < (int)c.k; ++d) {
    if (kn < p && valid[kn] == d) ++kn;
    else missing[j++] = d;
  }
  uint16_t w[kMaxSolve * kMaxSolve];
  for (int r = 0; r < e; ++r)
    for (int col = 0; col < e; ++col) w[r * e + col] = c.matrix[(size_t)valid[p + r] * c.k + missing[col]];
  return gauss_jordan(F, w, e, e);
}

// Inserts a new entry at the front of the LRU, evicting from the back to the
// capacity and byte budget (with c.mu held).  A victim still read by queued
// kernels stays alive through the shared_ptr of the caller that launched them;
// its slot retires when the last owner lets go.
static void cache_insert(const Codec& c, const std::string& key, const std::shared_ptr<DecodeEntry>& e) {
  if (!c.use_cache) return;
  const size_t cap = c.cache_capacity;
  while (!c.lru.empty() && (c.index.size() >= cap || c.cache_bytes + e->bytes > c.cache_max_bytes)) {
    auto& victim = c.lru.back();
    c.cache_bytes -= victim.second->bytes;
    c.index.erase(victim.first);
    c.lru.pop_back();
  }
  if (e->bytes <= c.cache_max_bytes) {
    c.lru.emplace_front(key, e);
    c.index[key] = c.lru.begin();
    c.cache_bytes += e->bytes;
  }
}

// The one-workgroup solve's job for `plan` (k <= 256, at most 256 targets).
template <class V>
static void fill_decode_job_lists(const Codec& c, const V& valid, const V& targets, DecodeJob& job) {
  job.k = (int)c.k;
  job.m = (int)c.m;
  job.n_targets = (int)targets.size();
  job.tower = c.field == RSGPU_FIELD_GF16 ? 1 : 0;
  job.field = c.ftab ? (const uint8_t*)c.d_ftab.p : nullptr;
  job.one = c.ftab ? c.ftab->one : 1;
  for (size_t i = 0; i < c.k; ++i) job.valid[i] = (uint16_t)valid[i];
  for (size_t i = 0; i < targets.size(); ++i) job.targets[i] = (uint16_t)targets[i];
}

static void fill_decode_job(const Codec& c, const Plan& plan, DecodeJob& job) {
  fill_decode_job_lists(c, plan.valid, plan.targets, job);
}

// A decode entry for `plan` on a pool slot (with c.mu held); its solve is
// enqueued on s by the caller, who records slot.ready behind it.
// Byte offsets inside a pool slot: tables and rows of m targets, then for the
// large-codec solve C^-1, the index lists (missing, parity survivors, present
// data) and, beyond kMaxSolve missing rows, the [C | I] work matrix.
struct SlotLayout {
  size_t rows, cinv, lists, work, total;
  explicit SlotLayout(const Codec& c) {
    const bool tower = c.field == RSGPU_FIELD_GF16;
    const size_t max_e = std::min<size_t>(std::min(c.k, c.m), kMaxBigSolve);
    auto up = [](size_t n) { return (n + 255) / 256 * 256; };
    rows = up(c.m * c.k * (size_t)words_per_coef(tower) * 4);
    cinv = rows + up(c.m * c.k * 2);
    lists = cinv + up(max_e * max_e * 2);
    work = lists + up((c.k + max_e) * 2);
    total = work + (max_e > (size_t)kMaxSolve ? up(2 * max_e * max_e * 2) : 0);
  }
};

// `n_coefs` (default: targets x k) is the number of row coefficients the solve
// writes; an entry with more than the pool's m x k gets a buffer of its own.
static int new_decode_entry(const Codec& c, const Plan& plan, const std::string& key, hipStream_t s,
                            std::shared_ptr<DecodeEntry>& out, size_t n_coefs = 0) {
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const size_t W4 = (size_t)words_per_coef(tower) * 4;
  const SlotLayout lay(c);
  const size_t max_tab = lay.rows;
  if (!c.decode_pool) {  // one slot size per codec
    c.decode_pool = std::make_shared<DecodePool>();
    c.decode_pool->slot_bytes = lay.total;
  }
  if (!n_coefs) n_coefs = plan.targets.size() * c.k;
  auto e = std::make_shared<DecodeEntry>();
  e->valid = plan.valid;
  e->targets = plan.targets;
  if (n_coefs <= c.m * c.k) {
    HIP_TRY(c.decode_pool->acquire(e->slot));
    e->pool = c.decode_pool;
    e->tables = (uint32_t*)e->slot.mem;
    e->rows = (uint8_t*)e->slot.mem + max_tab;
    e->bytes = c.decode_pool->slot_bytes + key.size();
  } else {
    const size_t tab = (n_coefs * W4 + 255) / 256 * 256;
    e->own = std::make_unique<DeviceBuffer>();
    HIP_TRY(e->own->alloc(tab + n_coefs * 2));
    HIP_TRY(hipEventCreateWithFlags(&e->slot.ready, hipEventDisableTiming));
    e->tables = (uint32_t*)e->own->p;
    e->rows = (uint8_t*)e->own->p + tab;
    e->bytes = e->own->bytes + key.size();
  }
  e->solve_stream = s;
  out = std::move(e);
  return RSGPU_OK;
}

// Orders a read of `e` on stream s after its solve when the solve was enqueued
// on another stream and has not completed yet.
static int wait_ready(const DecodeEntry& e, hipStream_t s) {
  if (!e.slot.ready || e.solve_stream == s || hipEventQuery(e.slot.ready) == hipSuccess) return RSGPU_OK;
  HIP_TRY(hipStreamWaitEvent(s, e.slot.ready, 0));
  return RSGPU_OK;
}

// Patterns whose decode rows are solved on the host and whose tables travel in
// the kernel arguments: the library's own byte fields (no Leopard generator-
// matrix field), and targets x k tables within kInlineTabWords.
static bool host_solve_eligible(const Codec& c, size_t n_targets) {
  return c.family == RSGPU_FAMILY_CLASSIC && !c.ftab && c.k <= (size_t)kMaxIn &&
         n_targets <= (size_t)kMaxOut &&
         n_targets * c.k * (size_t)words_per_coef(c.field == RSGPU_FIELD_GF16) <= (size_t)kInlineTabWords;
}

// The decode rows of decode_rows_body on the host (the same solve: [C | B | I]
// Gauss-Jordan over the first k survivors, recovery_plan.rs:329-387, composed
// parity rows per recovery_plan.rs:483-515), then their v_perm tables.  The rows
// are unique for an invertible C, so they equal the device solve's bytes.
static bool host_decode_tables(const Codec& c, const Plan& plan, std::vector<uint32_t>& tables) {
  const size_t k = c.k, nt = plan.targets.size();
  const Field F{c.field == RSGPU_FIELD_GF16};
  size_t p = 0;  // present data among the survivors (a sorted prefix of valid)
  while (p < k && plan.valid[p] < (int)k) ++p;
  const size_t e = k - p, cols = e + p + e;
  std::vector<int> missing;
  for (size_t d = 0, kn = 0; d < k && missing.size() < e; ++d) {
    if (kn < p && plan.valid[kn] == (int)d) ++kn;
    else missing.push_back((int)d);
  }
  std::vector<uint16_t> wm(e * cols, 0);
  for (size_t r = 0; r < e; ++r) {
    const uint16_t* prow = c.matrix.data() + (size_t)plan.valid[p + r] * k;
    for (size_t j = 0; j < e; ++j) wm[r * cols + j] = prow[missing[j]];
    for (size_t j = 0; j < p; ++j) wm[r * cols + e + j] = prow[plan.valid[j]];
    wm[r * cols + e + p + r] = 1;
  }
  if (e && !gauss_jordan(F, wm.data(), (int)e, (int)cols)) return false;
  std::vector<uint16_t> rows(nt * k, 0);
  for (size_t t = 0; t < nt; ++t) {
    const int target = plan.targets[t];
    uint16_t* row = rows.data() + t * k;
    if (target < (int)k) {
      const size_t j = std::find(missing.begin(), missing.end(), target) - missing.begin();
      for (size_t col = 0; col < k; ++col)
        row[col] = j < e ? wm[j * cols + e + col] : (plan.valid[col] == target ? 1 : 0);
    } else {
      const uint16_t* prow = c.matrix.data() + (size_t)target * k;
      for (size_t col = 0; col < k; ++col) {
        uint16_t v = col < p ? prow[plan.valid[col]] : 0;
        for (size_t j = 0; j < e; ++j) v ^= F.mul(prow[missing[j]], wm[j * cols + e + col]);
        row[col] = v;
      }
    }
  }
  build_tables_host(rows.data(), (int)nt, (int)k, c.field == RSGPU_FIELD_GF16, tables, nullptr);
  return true;
}

// A device copy of a host-solved entry's tables, made on first need (a batch,
// an unaligned or wide call): a pool slot filled by a kernel that carries the
// words in its arguments, ordered on s.
static int ensure_device_tables(const Codec& c, DecodeEntry& e, hipStream_t s) {
  std::lock_guard<std::mutex> g(e.dev_mu);
  if (e.tables) return wait_ready(e, s);
  {
    std::lock_guard<std::mutex> cg(c.mu);
    if (!c.decode_pool) {
      c.decode_pool = std::make_shared<DecodePool>();
      c.decode_pool->slot_bytes = SlotLayout(c).total;
    }
    e.pool = c.decode_pool;
  }
  HIP_TRY(e.pool->acquire(e.slot));
  HIP_TRY(launch_store_u32((uint32_t*)e.slot.mem, e.h_tables.data(), e.h_tables.size(), s));
  HIP_TRY(hipEventRecord(e.slot.ready, s));
  e.solve_stream = s;
  e.rows = nullptr;  // only the tables are copied
  e.tables = (uint32_t*)e.slot.mem;
  return RSGPU_OK;
}

// The cache entry of `plan`.  With inline_ok the caller can take host tables
// (then a miss is solved on the host when eligible); otherwise the entry has
// device tables on return, ordered before work on s.
static int get_decode_entry(const Codec& c, const rsgpu_slot* shards, const Plan& plan,
                            hipStream_t s, std::shared_ptr<DecodeEntry>& out, bool inline_ok = false) {
  const std::string key = plan_key(c, shards, plan);
  std::unique_lock<std::mutex> g(c.mu);
  ++c.requests;
  auto it = c.index.find(key);
  if (it != c.index.end()) {
    ++c.hits;
    c.lru.splice(c.lru.begin(), c.lru, it->second);
    out = it->second->second;
    g.unlock();
    if (out->h_tables.empty()) return wait_ready(*out, s);  // a device solve's entry
    return inline_ok ? RSGPU_OK : ensure_device_tables(c, *out, s);
  }
  ++c.misses;
  if (c.matrix_mode == RSGPU_MATRIX_CUSTOM && !custom_solvable(c, plan)) return RSGPU_INVALID_CUSTOM_MATRIX;
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const size_t nt = plan.targets.size();
  if (inline_ok && host_solve_eligible(c, nt)) {
    auto h = std::make_shared<DecodeEntry>();
    h->valid = plan.valid;
    h->targets = plan.targets;
    if (!host_decode_tables(c, plan, h->h_tables)) return RSGPU_INVALID_CUSTOM_MATRIX;
    h->bytes = h->h_tables.size() * 4 + nt * c.k * (tower ? 2 : 1) + key.size();
    out = h;
    cache_insert(c, key, h);
    return RSGPU_OK;
  }
  std::shared_ptr<DecodeEntry> e;
  int rc = new_decode_entry(c, plan, key, s, e);
  if (rc) return rc;
  int32_t* st = thread_flag(c.device);
  if (!st) return RSGPU_ERR_OUT_OF_MEMORY;
  hipError_t he;
  size_t p = 0;  // present data among the survivors (a sorted prefix)
  while (p < c.k && plan.valid[p] < (int)c.k) ++p;
  const size_t n_missing = c.k - p;
  if (c.k > 256 || nt > 256 || n_missing > (size_t)kMaxSolve) {  // beyond DecodeJob's limits
    if (n_missing > (size_t)kMaxBigSolve || e->own) return RSGPU_ERR_UNSUPPORTED;
    const SlotLayout lay(c);
    uint8_t* base = (uint8_t*)e->slot.mem;
    DecodeBigJob bj{};
    bj.k = (int)c.k;
    bj.e = (int)n_missing;
    bj.tower = tower ? 1 : 0;
    bj.field = c.ftab ? (const uint8_t*)c.d_ftab.p : nullptr;
    bj.one = c.ftab ? c.ftab->one : 1;
    // lists [missing e][parity survivors e][present data k - e] into the slot
    std::vector<uint16_t> lists(c.k + n_missing);
    for (size_t d = 0, kn = 0, j = 0; d < c.k; ++d) {
      if (kn < p && plan.valid[kn] == (int)d) lists[2 * n_missing + kn++] = (uint16_t)d;
      else lists[j++] = (uint16_t)d;
    }
    for (size_t i = 0; i < n_missing; ++i) lists[n_missing + i] = (uint16_t)plan.valid[p + i];
    uint16_t* dl = (uint16_t*)(base + lay.lists);
    HIP_TRY(launch_store_u16(dl, lists.data(), lists.size(), s));
    bj.missing = dl;
    bj.par_surv = dl + n_missing;
    bj.present = dl + 2 * n_missing;
    he = launch_decode_big((const uint16_t*)c.d_matrix.p, bj, plan.targets, (uint16_t*)(base + lay.cinv),
                           (uint16_t*)(base + lay.work), e->rows, e->tables, st + 1, s);
  } else {
    DecodeJob job{};
    fill_decode_job(c, plan, job);
    he = launch_decode_rows((const uint16_t*)c.d_matrix.p, job, e->rows, e->tables, st + 1, s);
  }
  if (he != hipSuccess) {
    set_hip_error(he, "decode_rows");
    return he == hipErrorInvalidValue ? RSGPU_ERR_UNSUPPORTED : RSGPU_ERR_HIP;
  }
  HIP_TRY(hipEventRecord(e->slot.ready, s));
  out = e;
  cache_insert(c, key, e);
  return RSGPU_OK;
}

// ---------------------------------------------- decode solves ahead of a batch
PresolveScratch::~PresolveScratch() {
  if (done) (void)hipEventSynchronize(done);
  if (pinned) (void)hipHostFree(pinned);
  for (uint8_t* p : old_pinned) (void)hipHostFree(p);
  if (done) (void)hipEventDestroy(done);
}

int PresolveScratch::ensure(size_t bytes) {
  if (bytes <= cap) return RSGPU_OK;
  const size_t want = std::max<size_t>(std::max(bytes, 2 * cap), 256 << 10);
  if (pinned) old_pinned.push_back(pinned);
  pinned = nullptr;
  if (dev.p) {
    auto o = std::make_unique<DeviceBuffer>();
    std::swap(o->p, dev.p);
    std::swap(o->bytes, dev.bytes);
    old_dev.push_back(std::move(o));
  }
  cap = 0;
  HIP_TRY(hipHostMalloc((void**)&pinned, want, hipHostMallocDefault));
  HIP_TRY(dev.alloc(want));
  cap = want;
  return RSGPU_OK;
}

int presolve_decode(const std::vector<PresolveCall>& calls, hipStream_t s, PresolveScratch& scr) {
  std::vector<const Codec*> order;  // codecs in first-seen order
  for (auto& cl : calls)
    if (std::find(order.begin(), order.end(), cl.codec) == order.end()) order.push_back(cl.codec);
  if (!scr.done) HIP_TRY(hipEventCreateWithFlags(&scr.done, hipEventDisableTiming));
  HIP_TRY(hipEventSynchronize(scr.done));  // the previous call's jobs were read
  // room for every call's job up front: no reallocation while copies are queued
  int rc0 = scr.ensure(calls.size() * sizeof(DecodeMultiJob) + order.size() * 256);
  if (rc0) return rc0;
  size_t off = 0;
  for (const Codec* cp : order) {
    const Codec& c = *cp;
    if (c.family != RSGPU_FAMILY_CLASSIC || c.k > 256) continue;
    std::lock_guard<std::mutex> g(c.mu);  // held until the new entries' events are recorded
    std::vector<DecodeMultiJob> jobs;
    std::vector<std::pair<std::string, std::shared_ptr<DecodeEntry>>> made;
    // the codec's new patterns first; a few are left to their calls, whose own
    // solves need no job-array transfer
    std::vector<std::pair<std::string, Plan>> fresh;
    for (auto& cl : calls) {
      if (cl.codec != cp) continue;
      Plan plan;
      if (make_plan(c, cl.shards, cl.n, 0, nullptr, 0, plan) || plan.all_present ||
          plan.targets.empty() || plan.targets.size() > 256)
        continue;  // errors are reported by the call itself
      std::string key = plan_key(c, cl.shards, plan);
      if (c.index.count(key)) continue;
      bool dup = false;
      for (auto& f : fresh) dup |= f.first == key;
      if (dup) continue;
      if (c.matrix_mode == RSGPU_MATRIX_CUSTOM && !custom_solvable(c, plan)) continue;
      fresh.emplace_back(std::move(key), std::move(plan));
    }
    if (fresh.size() < kPresolveMin) continue;
    size_t lds = 0;
    for (auto& f : fresh) {
      DecodeMultiJob J{};
      fill_decode_job(c, f.second, J.job);
      const size_t l = decode_rows_lds(J.job);
      if (!l) continue;
      lds = std::max(lds, l);
      std::shared_ptr<DecodeEntry> e;
      int rc = new_decode_entry(c, f.second, f.first, s, e);
      if (rc) return rc;
      J.rows = e->rows;
      J.tables = e->tables;
      jobs.push_back(J);
      made.emplace_back(f.first, std::move(e));
    }
    if (jobs.empty()) continue;
    const size_t bytes = jobs.size() * sizeof(DecodeMultiJob);
    std::memcpy(scr.pinned + off, jobs.data(), bytes);
    HIP_TRY(hipMemcpyAsync((uint8_t*)scr.dev.p + off, scr.pinned + off, bytes, hipMemcpyHostToDevice, s));
    int32_t* st = thread_flag(c.device);
    if (!st) return RSGPU_ERR_OUT_OF_MEMORY;
    HIP_TRY(launch_decode_rows_multi((const uint16_t*)c.d_matrix.p,
                                     (const DecodeMultiJob*)((uint8_t*)scr.dev.p + off),
                                     (int)jobs.size(), lds, st + 1, s));
    off += (bytes + 255) / 256 * 256;
    for (auto& mk : made) {
      HIP_TRY(hipEventRecord(mk.second->slot.ready, s));
      c.requests++;
      c.misses++;
      cache_insert(c, mk.first, mk.second);
    }
  }
  HIP_TRY(hipEventRecord(scr.done, s));
  return RSGPU_OK;
}

static int reconstruct_impl(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                            const uint8_t* required, size_t n_required, hipStream_t s,
                            int variant = kPick) {
  if (variant == kPick) {
    size_t erased = 0, len = 0;
    for (size_t i = 0; i < std::min(n, c.k + c.m); ++i) {
      if (!shards[i].present) ++erased;
      else if (!len) len = shards[i].len;
    }
    variant = pick_variant(c, kOpReconstruct, erased, len);
  }
  if (c.family != RSGPU_FAMILY_CLASSIC)
    return leopard_reconstruct(c, shards, n, goal, required, n_required, s, variant);
  Plan plan;
  int rc = make_plan(c, shards, n, goal, required, n_required, plan);
  if (rc || plan.all_present || plan.targets.empty()) return rc;
  const size_t nt = plan.targets.size();
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const uint8_t* in[kMaxIn];
  uint8_t* out[kMaxOut];
  std::vector<const uint8_t*> vin;
  std::vector<uint8_t*> vout;
  const uint8_t** pin = in;
  uint8_t** pout = out;
  if (c.k > (size_t)kMaxIn || nt > (size_t)kMaxOut) {  // wide codecs: heap lists
    vin.resize(c.k);
    vout.resize(nt);
    pin = vin.data();
    pout = vout.data();
  }
  for (size_t i = 0; i < c.k; ++i) pin[i] = shards[plan.valid[i]].ptr;
  for (size_t t = 0; t < nt; ++t) pout[t] = shards[plan.targets[t]].ptr;
  const bool inline_ok = apply_inline_fits((int)c.k, (int)nt, tower, pin, pout, plan.shard_len);
  std::shared_ptr<DecodeEntry> e;
  if ((rc = get_decode_entry(c, shards, plan, s, e, inline_ok))) return rc;
  const bool inline_tabs = inline_ok && !e->h_tables.empty();
  rc = run_apply(inline_tabs ? nullptr : e->tables, (int)c.k, (int)nt, tower, pin, pout, plan.shard_len, kWrite,
                 nullptr, s, variant, -1, 0, inline_tabs ? e->h_tables.data() : nullptr);
  if (rc) return rc;
  // An entry evicted meanwhile retires its slot when `e` goes out of scope; the
  // slot is reused only after a device-wide synchronisation.
  for (int t : plan.targets) shards[t].present = 1;
  return RSGPU_OK;
}

}  // namespace rsgpu

// ------------------------------------------------- prepared reconstruction
// ReconstructWorkspace (src/core/workspace.rs:85-140): the effective format it
// was prepared for, the validated erasure layout, the targets and the device
// decode entry, owned by the workspace so a cache eviction does not drop it.
struct rsgpu_reconstruct_workspace {
  size_t k = 0, m = 0;
  int field = 0, family = 0;
  std::vector<uint16_t> matrix;
  std::vector<uint8_t> present;  // prepared layout: 1 = present
  size_t shard_len = 0;
  bool all_present = false;
  std::vector<int> valid, targets;  // targets: data ascending, then parity ascending
  std::shared_ptr<rsgpu::DecodeEntry> entry;
};

namespace rsgpu {

// prepare_reconstruct_workspace_with_goal (reconstruct.rs:598-611): counts,
// family, the shape of the present lengths, then the plan and its decode solve.
static int prepare_workspace(const Codec& c, const rsgpu_slot* shards, size_t n, int goal,
                             const uint8_t* required, size_t n_required, hipStream_t s,
                             rsgpu_reconstruct_workspace** out) {
  const size_t total = c.k + c.m;
  if (goal == 2 && n_required != c.k && n_required != total) return RSGPU_INVALID_SHARD_FLAGS;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  Plan plan;
  int rc = plan_shards(c, shards, n, goal, required, n_required, plan.valid, plan.targets, plan.shard_len,
                       plan.all_present, false);
  if (rc) return rc;
  auto w = std::make_unique<rsgpu_reconstruct_workspace>();
  w->k = c.k;
  w->m = c.m;
  w->field = c.field;
  w->family = c.family;
  w->matrix = c.matrix;
  w->present.resize(total);
  for (size_t i = 0; i < total; ++i) w->present[i] = shards[i].present ? 1 : 0;
  w->shard_len = plan.shard_len;
  w->all_present = plan.all_present;
  w->valid = plan.valid;
  w->targets = plan.targets;
  // an empty request is a validated no-op without a solve (recovery_plan.rs:488-495)
  if (!plan.all_present && !plan.targets.empty())
    if ((rc = get_decode_entry(c, shards, plan, s, w->entry, host_solve_eligible(c, plan.targets.size()))))
      return rc;
  *out = w.release();
  return RSGPU_OK;
}

// validate_reconstruct_workspace (reconstruct.rs:669-700) then the prepared
// launches: no planning, cache lookup or solve here.
static int reconstruct_with_workspace_impl(const Codec& c, rsgpu_slot* shards, size_t n,
                                           const rsgpu_reconstruct_workspace& w, hipStream_t s) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (w.k < c.k) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (w.k > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  if (w.m < c.m) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (w.m > c.m) return RSGPU_TOO_MANY_PARITY_SHARDS;
  if (w.field != c.field || w.family != c.family || w.matrix != c.matrix) return RSGPU_INVALID_CUSTOM_MATRIX;
  // validate_classic_plan_layout (recovery_plan.rs:390-410)
  size_t len = 0;
  bool same = true;
  for (size_t i = 0; i < total; ++i) {
    const bool present = shards[i].present != 0;
    if (present) {
      if (shards[i].len == 0) return RSGPU_EMPTY_SHARD;
      if (len != 0 && shards[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
      len = shards[i].len;
    }
    same &= present == (w.present[i] != 0);
  }
  if (!same) return RSGPU_INVALID_SHARD_FLAGS;
  if (len != w.shard_len) return RSGPU_INCORRECT_SHARD_SIZE;
  if (w.all_present || w.targets.empty()) return RSGPU_OK;
  for (int t : w.targets) {
    if (!shards[t].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
    if (shards[t].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  }
  const size_t nt = w.targets.size();
  const uint8_t* in_buf[kMaxIn];
  uint8_t* out_buf[kMaxOut];
  std::vector<const uint8_t*> vin;
  std::vector<uint8_t*> vout;
  const uint8_t** pin = in_buf;
  uint8_t** pout = out_buf;
  if (c.k > (size_t)kMaxIn || nt > (size_t)kMaxOut) {  // wide codecs: heap lists
    vin.resize(c.k);
    vout.resize(nt);
    pin = vin.data();
    pout = vout.data();
  }
  for (size_t i = 0; i < c.k; ++i) {
    pin[i] = shards[w.valid[i]].ptr;
    if (!pin[i]) return RSGPU_ERR_INVALID_ARGUMENT;
  }
  for (size_t t = 0; t < nt; ++t) pout[t] = shards[w.targets[t]].ptr;
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const DecodeEntry& e = *w.entry;
  const bool inline_tabs = !e.h_tables.empty() && apply_inline_fits((int)c.k, (int)nt, tower, pin, pout, len);
  int rc = inline_tabs ? RSGPU_OK
                       : (e.h_tables.empty() ? wait_ready(e, s) : ensure_device_tables(c, *w.entry, s));
  if (rc) return rc;
  rc = run_apply(inline_tabs ? nullptr : e.tables, (int)c.k, (int)nt, tower, pin, pout, len, kWrite, nullptr, s,
                 pick_variant(c, kOpReconstruct, total - c.k, len), -1, 0,
                 inline_tabs ? e.h_tables.data() : nullptr);
  if (rc) return rc;
  for (int t : w.targets) shards[t].present = 1;
  return RSGPU_OK;
}

// ------------------------------------------------------- progressive recovery
// decode_idx (src/galois_8/policy.rs:543-704).  With expect != NULL the k
// expected survivors define the decode; this call's inputs are a subset of them
// and each requested destination accumulates sum_col row[col] * input[col] over
// that subset, so calls over a partition of the survivors add up to the
// reconstruction.  With expect == NULL partial results merge by XOR.  NULL ptr =
// None.  Every pair or length is validated before any device work.
static int decode_idx_impl(const Codec& c, const rsgpu_slice* dst, size_t n_dst, const uint8_t* expect,
                           size_t n_expect, const rsgpu_cslice* input, size_t n_input, hipStream_t s) {
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  const size_t total = c.k + c.m;
  for (size_t count : {n_dst, n_input}) {  // validate_opt_shard_count (policy.rs:40-48)
    if (count < total) return RSGPU_TOO_FEW_SHARDS;
    if (count > total) return RSGPU_TOO_MANY_SHARDS;
  }
  if (!expect) {  // merge mode (policy.rs:675-700): validate every pair, then XOR
    for (size_t i = 0; i < total; ++i) {
      if (!input[i].ptr) continue;
      if (!dst[i].ptr) return RSGPU_TOO_FEW_SHARDS;
      if (dst[i].len != input[i].len) return RSGPU_INCORRECT_SHARD_SIZE;
    }
    XorPairs xp{};
    for (size_t i = 0; i <= total; ++i) {
      if (i < total && input[i].ptr && input[i].len) {
        xp.src[xp.n] = input[i].ptr;
        xp.dst[xp.n] = dst[i].ptr;
        xp.len[xp.n] = input[i].len;
        ++xp.n;
      }
      if (xp.n == kXorPairs || (i == total && xp.n)) {
        HIP_TRY(launch_xor_pairs(xp, s));
        xp.n = 0;
      }
    }
    return RSGPU_OK;
  }
  if (n_expect != total) return RSGPU_INVALID_SHARD_FLAGS;
  std::vector<int> valid;
  for (size_t i = 0; i < total; ++i)
    if (expect[i]) valid.push_back((int)i);
  if (valid.size() < c.k) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  // the shard length: the first input, else the first destination (policy.rs:579-583)
  size_t len = 0;
  bool found = false;
  for (size_t i = 0; i < total && !found; ++i)
    if (input[i].ptr) len = input[i].len, found = true;
  for (size_t i = 0; i < total && !found; ++i)
    if (dst[i].ptr) len = dst[i].len, found = true;
  if (!found) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  for (size_t i = 0; i < total; ++i)
    if (input[i].ptr && input[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  for (size_t i = 0; i < total; ++i)
    if (dst[i].ptr && dst[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  std::vector<int> outputs, cols;
  for (size_t i = 0; i < total; ++i)
    if (dst[i].ptr) outputs.push_back((int)i);
  if (outputs.empty()) return RSGPU_OK;
  for (size_t col = 0; col < c.k; ++col)
    if (input[valid[col]].ptr) cols.push_back((int)col);
  if (cols.empty()) return RSGPU_OK;
  // get_data_decode_matrix takes exactly k survivors (reconstruct.rs:1181-1183)
  if (valid.size() != c.k) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  if (c.matrix_mode == RSGPU_MATRIX_CUSTOM && !custom_solvable_valid(c, valid)) return RSGPU_INVALID_CUSTOM_MATRIX;
  if (len == 0) return RSGPU_OK;  // decode_idx_accumulate_reduced_outputs: nothing to add
  int rc = check_elem_len(c, len);
  if (rc) return rc;
  if (c.k > 256 || outputs.size() > 256) return RSGPU_ERR_UNSUPPORTED;  // one-workgroup solve
  // cached rows: key = expected set, destinations and the columns of this call
  std::string key = "x";
  for (size_t i = 0; i < total; ++i) key.push_back((char)('a' + (expect[i] ? 1 : 0) + (dst[i].ptr ? 2 : 0)));
  key.push_back('|');
  for (size_t col = 0, q = 0; col < c.k; ++col) {
    const bool on = q < cols.size() && cols[q] == (int)col;
    q += on ? 1 : 0;
    key.push_back(on ? '1' : '0');
  }
  std::shared_ptr<DecodeEntry> e;
  {
    std::lock_guard<std::mutex> g(c.mu);
    ++c.requests;
    auto it = c.index.find(key);
    if (it != c.index.end()) {
      ++c.hits;
      c.lru.splice(c.lru.begin(), c.lru, it->second);
      e = it->second->second;
      if ((rc = wait_ready(*e, s))) return rc;
    } else {
      ++c.misses;
      Plan plan;
      plan.valid = valid;
      plan.targets = outputs;
      if ((rc = new_decode_entry(c, plan, key, s, e, outputs.size() * cols.size()))) return rc;
      int32_t* st = thread_flag(c.device);
      if (!st) return RSGPU_ERR_OUT_OF_MEMORY;
      DecodeJob job{};
      fill_decode_job(c, plan, job);
      job.n_cols = (int32_t)cols.size();
      for (size_t q = 0; q < cols.size(); ++q) job.cols[q] = (uint16_t)cols[q];
      hipError_t he = launch_decode_rows((const uint16_t*)c.d_matrix.p, job, e->rows, e->tables, st + 1, s);
      if (he != hipSuccess) {
        set_hip_error(he, "decode_idx rows");
        return he == hipErrorInvalidValue ? RSGPU_ERR_UNSUPPORTED : RSGPU_ERR_HIP;
      }
      HIP_TRY(hipEventRecord(e->slot.ready, s));
      cache_insert(c, key, e);
    }
  }
  std::vector<const uint8_t*> in(cols.size());
  std::vector<uint8_t*> out(outputs.size());
  for (size_t q = 0; q < cols.size(); ++q) in[q] = input[valid[cols[q]]].ptr;
  for (size_t t = 0; t < outputs.size(); ++t) out[t] = dst[outputs[t]].ptr;
  return run_apply(e->tables, (int)cols.size(), (int)outputs.size(), c.field == RSGPU_FIELD_GF16, in, out, len,
                   kAccumulate, nullptr, s, pick_variant(c, kOpReconstruct, c.m, len));
}

int leopard_matrix_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                               const uint8_t* required, size_t n_required, hipStream_t s, int variant) {
  std::shared_ptr<Codec> mc;
  int rc = leo_matrix_codec(c, s, mc);
  if (rc) return rc;
  return reconstruct_impl(*mc, shards, n, goal, required, n_required, s, variant);
}

// The FFT decoder is linear over the field (constant multiplies, XORs and the
// transforms), so for one (present, target) pattern its output is M * (present
// shards).  M is read off one decode of unit columns: present shard p holds the
// field's identity at element p, so target t comes back with M[t][p] at element p.
// Applying M then gives the decoder's bytes for any input, codewords or not.
// GF(2^16) elements sit in the split layout (element p of a shard: low byte at
// (p / 32) * 64 + p % 32, high byte 32 bytes on).
static int build_decoder_matrix(const Codec& c, const uint8_t* flags, hipStream_t s,
                                std::shared_ptr<const DecoderMatrix>& out) {
  const size_t total = c.k + c.m;
  const bool g16 = c.family == RSGPU_FAMILY_LEOPARD_GF16;
  auto dm = std::make_shared<DecoderMatrix>();
  for (size_t i = 0; i < total; ++i) {
    if (flags[i] & 1) dm->inputs.push_back((int)i);
    else if (flags[i] & 2) dm->targets.push_back((int)i);
  }
  const size_t np = dm->inputs.size(), nt = dm->targets.size();
  const size_t L = g16 ? std::max<size_t>(64, (np + 31) / 32 * 64) : std::max<size_t>(64, (np + 63) / 64 * 64);
  auto at = [&](size_t p) { return g16 ? (p / 32) * 64 + p % 32 : p; };  // low (or only) byte of element p
  std::shared_ptr<const FieldTables> ft = g16 ? nullptr : leopard8_field_tables();
  const uint16_t one = g16 ? leopard16_one() : ft->one;
  DeviceBuffer buf;
  HIP_TRY(buf.alloc(total * L));
  std::vector<uint8_t> host(total * L, 0);
  for (size_t p = 0; p < np; ++p) {
    uint8_t* e = host.data() + (size_t)dm->inputs[p] * L + at(p);
    e[0] = (uint8_t)one;
    if (g16) e[32] = (uint8_t)(one >> 8);
  }
  HIP_TRY(hipMemcpyAsync(buf.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
  std::vector<rsgpu_slot> sl(total);
  std::vector<uint8_t> req(total);
  for (size_t i = 0; i < total; ++i) {
    sl[i] = {(uint8_t*)buf.p + i * L, L, (int32_t)(flags[i] & 1)};
    req[i] = (flags[i] & 2) ? 1 : 0;
  }
  int rc = leopard_reconstruct(c, sl.data(), total, 2, req.data(), total, s, -1);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(host.data(), buf.p, host.size(), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<uint16_t> rows(nt * np);
  for (size_t t = 0; t < nt; ++t)
    for (size_t p = 0; p < np; ++p) {
      const uint8_t* e = host.data() + (size_t)dm->targets[t] * L + at(p);
      rows[t * np + p] = g16 ? (uint16_t)(e[0] | e[32] << 8) : e[0];
    }
  std::vector<uint32_t> ht;
  if (g16) {
    ht.resize(rows.size() * 4 * kTabWords);
    leopard16_apply_tables(rows.data(), rows.size(), ht.data());
  } else {
    build_tables_host(rows.data(), (int)nt, (int)np, false, ht, ft.get());
  }
  HIP_TRY(dm->tables.alloc(ht.size() * 4));
  HIP_TRY(hipMemcpy(dm->tables.p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice));
  out = dm;
  return RSGPU_OK;
}

int leopard_decoder_matrix_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, const uint8_t* flags,
                                       size_t len, hipStream_t s, int variant) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  const std::string key((const char*)flags, total);
  std::shared_ptr<const DecoderMatrix> dm;
  {
    std::lock_guard<std::mutex> g(c.mu);
    dm = c.dec_matrix.get(key);
  }
  if (!dm) {
    int rc = build_decoder_matrix(c, flags, s, dm);
    if (rc) return rc;
    std::vector<std::shared_ptr<const DecoderMatrix>> evicted;  // freed after the lock
    std::lock_guard<std::mutex> g(c.mu);
    c.dec_matrix.put(key, dm, evicted);
  }
  std::vector<const uint8_t*> in;
  std::vector<uint8_t*> out;
  for (int i : dm->inputs) in.push_back(shards[i].ptr);
  for (int t : dm->targets) out.push_back(shards[t].ptr);
  const bool g16 = c.family == RSGPU_FAMILY_LEOPARD_GF16;
  bool vec = len % 16 == 0;
  for (auto p : in) vec = vec && aligned16(p);
  for (auto p : out) vec = vec && aligned16(p);
  // split over inputs from three input chunks on (two chunks run as fast in turn)
  int rc = (in.size() > (size_t)2 * kMaxIn && vec)
               ? run_apply_splitk((const uint32_t*)dm->tables.p, (int)in.size(), (int)out.size(), g16 ? 2 : 0,
                                  in.data(), out.data(), len, s, variant)
               : run_apply((const uint32_t*)dm->tables.p, (int)in.size(), (int)out.size(), g16, in.data(),
                           out.data(), len, kWrite, nullptr, s, variant, -1, 0, nullptr, nullptr, g16);
  if (rc) return rc;
  for (int t : dm->targets) shards[t].present = 1;
  return RSGPU_OK;
}

// ------------------------------------------------------------ batch stripes
// Adds one stripe operation to a batch: descriptors of at most kMaxOut
// (GF(2^16): kMaxOut16) outputs each, plus byte-kernel jobs for unaligned views
// and for the length % 16 tail.  The split layout takes whole aligned 64-byte
// blocks only (Leopard shard lengths are multiples of 64).
static int add_stripe(BatchPlan& plan, const std::vector<const uint8_t*>& in,
                      const std::vector<uint8_t*>& out, const uint32_t* tables, size_t len, int mode,
                      int32_t* flag, int elem = kElemByte) {
  const int k = (int)in.size(), r = (int)out.size();
  const int W = words_per_coef(elem != kElemByte);
  const int rstep = elem == kElemByte ? kMaxOut : kMaxOut16;
  bool vec_ok = true;
  for (auto p : in) vec_ok &= aligned16(p);
  for (auto p : out) vec_ok &= aligned16(p);
  if (elem == kElemSplit && (!vec_ok || len % 64)) return RSGPU_ERR_UNSUPPORTED;
  const uint64_t n16 = vec_ok ? len / 16 : 0;
  const size_t tail = (size_t)n16 * 16;
  // a verify tail compares the complete sum, so its inputs cannot be split
  if (tail < len && k > kMaxIn && mode == kVerify) return RSGPU_ERR_UNSUPPORTED;
  for (int r0 = 0; r0 < r; r0 += rstep) {
    const int rn = std::min(rstep, r - r0);
    if (n16) {
      BatchStripe s{};
      s.n16 = n16;
      s.tables = tables + (size_t)r0 * k * W;
      s.flag = flag;
      s.ptr0 = (uint32_t)plan.ptrs.size();
      s.tab_row_stride = (uint32_t)(k * W);
      s.k = (uint16_t)k;
      s.r = (uint16_t)rn;
      s.mode = (uint16_t)mode;
      s.elem = (uint16_t)elem;
      for (auto p : in) plan.ptrs.push_back(p);
      for (int j = 0; j < rn; ++j) plan.ptrs.push_back(out[r0 + j]);
      plan.stripes[elem].push_back(s);
    }
    // the len % 16 tail (or all of an unaligned stripe) in byte jobs of at most
    // kMaxIn inputs; after the first, each accumulates into the outputs
    for (int k0 = 0; tail < len && k0 < k; k0 += kMaxIn) {
      ByteArgs b{};
      const int kn = std::min(kMaxIn, k - k0);
      for (int i = 0; i < kn; ++i) b.in[i] = in[k0 + i];
      for (int j = 0; j < rn; ++j) b.out[j] = out[r0 + j];
      b.tables = tables + (size_t)r0 * k * W;
      b.flag = flag;
      b.begin = tail;
      b.end = len;
      b.k = kn;
      b.tab_row_stride = k * W;
      b.tab_col0 = k0;
      b.r = rn;
      b.mode = (k0 == 0 || mode == kVerify) ? mode : kAccumulate;
      b.tower = elem == kElemTower ? 1 : 0;
      plan.byte_jobs.push_back(b);
    }
  }
  plan.bytes_moved += (uint64_t)(k + r) * len;
  plan.data_bytes += (uint64_t)k * len;
  plan.n_ops++;
  return RSGPU_OK;
}

// A classic codec's element type in the batch kernels.
static int batch_elem(const Codec& c) { return c.field == RSGPU_FIELD_GF16 ? kElemTower : kElemByte; }

int batch_encode_stripe(const Codec& c, const rsgpu_slice* shards, size_t n, BatchPlan& plan) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  std::vector<const uint8_t*> in(c.k);
  std::vector<uint8_t*> out(c.m);
  for (size_t i = 0; i < c.k; ++i) in[i] = shards[i].ptr;
  for (size_t j = 0; j < c.m; ++j) out[j] = shards[c.k + j].ptr;
  for (size_t i = 0; i < n; ++i)
    if (!shards[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  if ((rc = check_elem_len(c, shards[0].len))) return rc;
  return add_stripe(plan, in, out, (const uint32_t*)c.d_parity_tables.p, shards[0].len, kWrite, nullptr,
                    batch_elem(c));
}

int batch_verify_stripe(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* flag,
                        BatchPlan& plan) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  std::vector<const uint8_t*> in(c.k);
  std::vector<uint8_t*> out(c.m);
  for (size_t i = 0; i < c.k; ++i) in[i] = shards[i].ptr;
  for (size_t j = 0; j < c.m; ++j) out[j] = const_cast<uint8_t*>(shards[c.k + j].ptr);
  for (size_t i = 0; i < n; ++i)
    if (!shards[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  if ((rc = check_elem_len(c, shards[0].len))) return rc;
  return add_stripe(plan, in, out, (const uint32_t*)c.d_parity_tables.p, shards[0].len, kVerify, flag,
                    batch_elem(c));
}

int batch_reconstruct_stripe(const Codec& c, rsgpu_slot* shards, size_t n, int goal, hipStream_t s,
                             BatchPlan& plan) {
  Plan p;
  int rc = make_plan(c, shards, n, goal, nullptr, 0, p);
  if (rc || p.all_present || p.targets.empty()) return rc;
  std::shared_ptr<DecodeEntry> e;
  if ((rc = get_decode_entry(c, shards, p, s, e))) return rc;
  std::vector<const uint8_t*> in(c.k);
  std::vector<uint8_t*> out(p.targets.size());
  for (size_t i = 0; i < c.k; ++i) in[i] = shards[p.valid[i]].ptr;
  for (size_t t = 0; t < p.targets.size(); ++t) out[t] = shards[p.targets[t]].ptr;
  if ((rc = add_stripe(plan, in, out, e->tables, p.shard_len, kWrite, nullptr, batch_elem(c))))
    return rc;
  plan.keep.push_back(e);
  for (int t : p.targets) shards[t].present = 1;
  return RSGPU_OK;
}

// ------------------------------------------- Leopard codecs in the batch kernels
// The matrix forms of a Leopard GF(2^8) codec are GF(2^8) byte applies with
// tables in Leopard's field, bit-identical to the FFT: encode and verify over
// [I; G], reconstruct with exactly k shards present as the [I; G] solve, with
// more as the decoder's own matrix over every present shard.  A batch takes them
// into its kernel when the codec's dispatch picks a matrix variant for the call,
// or, with no dispatch state, for k <= 32: there the FFT costs a launch per op
// and at least as much per call (10+4 x 1 MiB on MI355X: encode 7.5 us FFT, 5.9
// matrix; reconstruct of 1 or 4 erasures 19 us FFT, 4.8 and 6.4 matrix), while the
// matrix forms share the batch's one launch.  A Leopard GF(2^16) reconstruct goes into the
// split-layout kernel as the decoder's own matrix when the dispatch picks a
// matrix variant for it (at most 32 erasures), or, with no dispatch state, for
// at most 4 erasures: the FFT costs the same for one erasure as for m, and per
// call the decoder matrix measured faster at every shape up to 4 (1024+256 x
// 256 KiB: 1 / 4 erasures 92 / 161 us against 930-1000 FFT; 4096+1024 x 64 KiB:
// 133 / 201 against 1750; 300+45 x 64 KiB: 31 / 40 against 62-67).  Otherwise
// RSGPU_ERR_UNSUPPORTED tells the batch to run the FFT through its own launch.
static bool leo_batch_matrix(const Codec& c, int op, size_t erasures, size_t len) {
  if (c.family == RSGPU_FAMILY_LEOPARD_GF16) {
    const int v = pick_variant(c, op, erasures, len);  // with no state: 1..4 erasures
    return v >= 0 && v < kNumApplyVariants;
  }
  if (c.family != RSGPU_FAMILY_LEOPARD_GF8) return false;
  const bool no_state = c.forced_variant.load(std::memory_order_relaxed) < 0 &&
                        !std::atomic_load(&c.variant_table);
  if (no_state) return c.k <= 32;
  return is_matrix_variant(c, pick_variant(c, op, erasures, len));
}

int batch_leopard8_encode(const Codec& c, const rsgpu_slice* shards, size_t n, BatchPlan& plan) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  if (shards[0].len % 64) return RSGPU_INCORRECT_SHARD_SIZE;  // leopard.rs:50-56
  if (!leo_batch_matrix(c, kOpEncode, 0, shards[0].len)) return RSGPU_ERR_UNSUPPORTED;
  std::shared_ptr<Codec> mc;
  if ((rc = leo_matrix_codec(c, nullptr, mc))) return rc;
  return batch_encode_stripe(*mc, shards, n, plan);
}

int batch_leopard8_verify(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* flag,
                          BatchPlan& plan) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  if (shards[0].len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
  if (!leo_batch_matrix(c, kOpVerify, 0, shards[0].len)) return RSGPU_ERR_UNSUPPORTED;
  std::shared_ptr<Codec> mc;
  if ((rc = leo_matrix_codec(c, nullptr, mc))) return rc;
  return batch_verify_stripe(*mc, shards, n, flag, plan);
}

// Validation as leopard_reconstruct (reconstruct.rs:795-1040) for goals 0 (all)
// and 1 (data only); both Leopard fields.
int batch_leopard_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal, hipStream_t s,
                              BatchPlan& plan) {
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  size_t len = 0, present = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!shards[i].present) continue;
    if (shards[i].len == 0) return RSGPU_EMPTY_SHARD;
    if ((len && shards[i].len != len) || shards[i].len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
    len = shards[i].len;
    ++present;
  }
  if (present < c.k) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  std::vector<uint8_t> flags(total, 0);
  bool need = false;
  for (size_t i = 0; i < total; ++i) {
    const bool req = goal == 0 || i < c.k;
    flags[i] = (uint8_t)((shards[i].present ? 1 : 0) | (!shards[i].present && req ? 2 : 0));
    if (!shards[i].present && req) {
      need = true;
      if (!shards[i].ptr || shards[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
    }
  }
  if (!need) return RSGPU_OK;
  if (total - present > c.m) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  if (!leo_batch_matrix(c, kOpReconstruct, total - present, len)) return RSGPU_ERR_UNSUPPORTED;
  const bool g16 = c.family == RSGPU_FAMILY_LEOPARD_GF16;
  if (g16) {  // the split-layout kernel takes aligned shards only
    for (size_t i = 0; i < total; ++i)
      if ((flags[i] & 3) && !aligned16(shards[i].ptr)) return RSGPU_ERR_UNSUPPORTED;
  }
  int rc;
  if (present == c.k && !g16) {
    std::shared_ptr<Codec> mc;
    if ((rc = leo_matrix_codec(c, s, mc))) return rc;
    return batch_reconstruct_stripe(*mc, shards, n, goal, s, plan);
  }
  const std::string key((const char*)flags.data(), total);
  std::shared_ptr<const DecoderMatrix> dm;
  {
    std::lock_guard<std::mutex> g(c.mu);
    dm = c.dec_matrix.get(key);
  }
  if (!dm) {
    if ((rc = build_decoder_matrix(c, flags.data(), s, dm))) return rc;
    std::vector<std::shared_ptr<const DecoderMatrix>> evicted;  // freed after the lock
    std::lock_guard<std::mutex> g(c.mu);
    c.dec_matrix.put(key, dm, evicted);
  }
  std::vector<const uint8_t*> in;
  std::vector<uint8_t*> out;
  for (int i : dm->inputs) in.push_back(shards[i].ptr);
  for (int t : dm->targets) out.push_back(shards[t].ptr);
  if ((rc = add_stripe(plan, in, out, (const uint32_t*)dm->tables.p, len, kWrite, nullptr,
                       g16 ? kElemSplit : kElemByte)))
    return rc;
  plan.keep.push_back(std::const_pointer_cast<DecoderMatrix>(dm));
  for (int t : dm->targets) shards[t].present = 1;
  return RSGPU_OK;
}

// ---------------------------------------------- allocation-free reconstruct
// Device scratch of rsgpu_reconstruct_noalloc: one decode solve's status word,
// coefficient tables and rows.  Each device holds its own copy (for the life of
// the process); host code reaches it only through the address that
// hipGetSymbolAddress returns for that device, kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = (size_t)2 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
static std::atomic<uint8_t*> g_scratch_ptr[kMaxScratchDevices];
static std::mutex g_scratch_mu;

// The device address of `device`'s scratch, looked up once per device with that
// device current.
static int scratch_for_device(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxScratchDevices) return RSGPU_ERR_UNSUPPORTED;
  uint8_t* p = g_scratch_ptr[device].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    p = g_scratch_ptr[device].load(std::memory_order_relaxed);
    if (!p) {
      int cur = 0;
      HIP_TRY(hipGetDevice(&cur));
      if (cur != device) HIP_TRY(hipSetDevice(device));
      void* addr = nullptr;
      hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (cur != device) (void)hipSetDevice(cur);
      HIP_TRY(e);
      p = (uint8_t*)addr;
      g_scratch_ptr[device].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return RSGPU_OK;
}

static size_t round256(size_t n) { return (n + 255) / 256 * 256; }

// reconstruct / reconstruct_data with the decode solve written into the device
// scratch instead of the decode cache: validation, planning and launches use
// stack storage only.
static int reconstruct_noalloc_impl(const Codec& c, rsgpu_slot* shards, size_t n, int goal, hipStream_t s) {
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (c.k + c.m > FixedIdx::kCap) {
    set_hip_error(hipErrorInvalidValue, "reconstruct_noalloc: more than 256 shards");
    return RSGPU_ERR_UNSUPPORTED;
  }
  FixedIdx valid, targets;
  size_t len = 0;
  bool all_present = false;
  int rc = plan_shards(c, shards, n, goal, nullptr, 0, valid, targets, len, all_present);
  if (rc || all_present || targets.size() == 0) return rc;
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const size_t nt = targets.size();
  const size_t tab_bytes = round256(nt * c.k * (size_t)words_per_coef(tower) * 4);
  const size_t row_bytes = round256(nt * c.k * (tower ? 2 : 1));
  if (256 + tab_bytes + row_bytes > kNoAllocScratchBytes) {
    set_hip_error(hipErrorInvalidValue, "reconstruct_noalloc: decode tables exceed the 2 MiB scratch");
    return RSGPU_ERR_UNSUPPORTED;
  }
  if (c.matrix_mode == RSGPU_MATRIX_CUSTOM && !custom_solvable_fixed(c, valid)) return RSGPU_INVALID_CUSTOM_MATRIX;
  uint8_t* base = nullptr;
  if ((rc = scratch_for_device(c.device, &base))) return rc;
  int32_t* status = (int32_t*)base;
  uint32_t* tables = (uint32_t*)(base + 256);
  uint8_t* rows = base + 256 + tab_bytes;
  DecodeJob job{};
  fill_decode_job_lists(c, valid, targets, job);
  hipError_t he = launch_decode_rows((const uint16_t*)c.d_matrix.p, job, rows, tables, status, s);
  if (he != hipSuccess) {
    set_hip_error(he, "reconstruct_noalloc decode_rows");
    return he == hipErrorInvalidValue ? RSGPU_ERR_UNSUPPORTED : RSGPU_ERR_HIP;
  }
  const uint8_t* in[FixedIdx::kCap];
  uint8_t* out[FixedIdx::kCap];
  for (size_t i = 0; i < c.k; ++i) in[i] = shards[valid[i]].ptr;
  for (size_t t = 0; t < nt; ++t) out[t] = shards[targets[t]].ptr;
  size_t erased = 0;
  for (size_t i = 0; i < n; ++i) erased += shards[i].present ? 0 : 1;
  rc = run_apply(tables, (int)c.k, (int)nt, tower, in, out, len, kWrite, nullptr, s,
                 pick_variant(c, kOpReconstruct, erased, len));
  if (rc) return rc;
  for (int t : targets) shards[t].present = 1;
  return RSGPU_OK;
}

}  // namespace rsgpu

using namespace rsgpu;

// ================================================================== C ABI
extern "C" {

void rsgpu_options_default(rsgpu_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->field = RSGPU_FIELD_GF8;
  o->codec_family = RSGPU_FAMILY_CLASSIC;
  o->matrix_mode = RSGPU_MATRIX_VANDERMONDE;
  o->leopard_mode = RSGPU_LEOPARD_DISABLED;
  o->inversion_cache = 1;
}

// resolve_codec_family (src/core/leopard.rs:191-223)
static int resolve_family(int requested, int mode, bool byte_field, size_t total, size_t parity) {
  if (requested != RSGPU_FAMILY_CLASSIC || !byte_field) return requested;
  switch (mode) {
    case RSGPU_LEOPARD_AS_NEEDED: return total <= 256 ? RSGPU_FAMILY_CLASSIC : RSGPU_FAMILY_LEOPARD_GF16;
    case RSGPU_LEOPARD_PREFER_GF16: return RSGPU_FAMILY_LEOPARD_GF16;
    case RSGPU_LEOPARD_PREFER_LEOPARD:
      return (total <= 256 && parity <= 128) ? RSGPU_FAMILY_LEOPARD_GF8 : RSGPU_FAMILY_LEOPARD_GF16;
    default: return RSGPU_FAMILY_CLASSIC;
  }
}

// leopard_shard_combo_fits (src/core/leopard.rs:225-236)
static bool leopard_fits(size_t k, size_t m, size_t order) {
  size_t p = 1;
  while (p < m) p <<= 1;
  if (k == 0 || m == 0 || p > order) return false;
  size_t rounded = (k + p - 1) / p * p;
  return rounded <= order - p;
}

static int codec_common(size_t k, size_t m, const rsgpu_options* opts, Codec& c, bool custom) {
  rsgpu_options o;
  rsgpu_options_default(&o);
  if (opts) o = *opts;
  if (k == 0) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (m == 0) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (o.field != RSGPU_FIELD_GF8 && o.field != RSGPU_FIELD_GF16) return RSGPU_ERR_INVALID_ARGUMENT;
  const size_t total = k + m;
  const bool byte_field = o.field == RSGPU_FIELD_GF8;
  if (custom) {
    if (total > (byte_field ? 256u : 65536u)) return RSGPU_TOO_MANY_SHARDS;
    if (o.codec_family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
    o.matrix_mode = RSGPU_MATRIX_CUSTOM;
  } else {
    o.codec_family = resolve_family(o.codec_family, o.leopard_mode, byte_field, total, m);
    size_t cap = o.codec_family == RSGPU_FAMILY_CLASSIC ? (byte_field ? 256 : 65536)
                 : o.codec_family == RSGPU_FAMILY_LEOPARD_GF8 ? 256 : 65536;
    if (total > cap) return RSGPU_TOO_MANY_SHARDS;
    if (o.codec_family != RSGPU_FAMILY_CLASSIC) {  // validate_leopard_family (leopard.rs:144-270)
      if (!byte_field) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
      if (!leopard_fits(k, m, o.codec_family == RSGPU_FAMILY_LEOPARD_GF8 ? 256 : 65536))
        return RSGPU_TOO_MANY_SHARDS;
    }
  }
  c.k = k;
  c.m = m;
  c.field = o.field;
  c.family = o.codec_family;
  c.matrix_mode = o.matrix_mode;
  c.fast_one_parity = o.fast_one_parity != 0;
  c.use_cache = o.inversion_cache != 0;
  c.cache_capacity = o.inversion_cache_capacity ? o.inversion_cache_capacity : derive_cache_capacity(k, m);
  c.cache_max_bytes = o.inversion_cache_max_bytes ? o.inversion_cache_max_bytes : (8u << 20);
  return RSGPU_OK;
}

int rsgpu_codec_new(size_t k, size_t m, const rsgpu_options* opts, rsgpu_codec** out) {
  if (!out) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto h = std::make_unique<rsgpu_codec>();
  Codec& c = h->c;
  int rc = codec_common(k, m, opts, c, false);
  if (rc) return rc;
  if (c.family == RSGPU_FAMILY_CLASSIC) {
    if (c.matrix_mode == RSGPU_MATRIX_CUSTOM) return RSGPU_INVALID_CUSTOM_MATRIX;
    rc = build_matrix(Field{c.field == RSGPU_FIELD_GF16}, c.matrix_mode, c.fast_one_parity, (int)k,
                      (int)m, c.matrix);
    if (rc) return rc;
  }
  if ((rc = upload_codec(c))) return rc;
  *out = h.release();
  return RSGPU_OK;
}

int rsgpu_codec_with_custom_matrix(size_t k, size_t m, const uint8_t* rows, size_t n_rows,
                                   size_t row_len, const rsgpu_options* opts, rsgpu_codec** out) {
  if (!out) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto h = std::make_unique<rsgpu_codec>();
  Codec& c = h->c;
  int rc = codec_common(k, m, opts, c, true);
  if (rc) return rc;
  if (n_rows < m || row_len < k || !rows) return RSGPU_INVALID_CUSTOM_MATRIX;  // codec.rs:204-215
  const int eb = c.field == RSGPU_FIELD_GF16 ? 2 : 1;
  c.matrix.assign((k + m) * k, 0);
  for (size_t r = 0; r < k; ++r) c.matrix[r * k + r] = 1;
  for (size_t r = 0; r < m; ++r)
    for (size_t col = 0; col < k; ++col) {
      const uint8_t* e = rows + (r * row_len + col) * eb;
      c.matrix[(k + r) * k + col] = eb == 2 ? (uint16_t)((e[0] << 8) | e[1]) : e[0];
    }
  if (c.fast_one_parity && m == 1)  // codec.rs:593-600
    for (size_t col = 0; col < k; ++col)
      if (c.matrix[k * k + col] != 1) return RSGPU_INVALID_CUSTOM_MATRIX;
  if ((rc = upload_codec(c))) return rc;
  *out = h.release();
  return RSGPU_OK;
}

void rsgpu_codec_free(rsgpu_codec* codec) { delete codec; }

int rsgpu_codec_info(const rsgpu_codec* h, size_t* k, size_t* m, int32_t* field, int32_t* fam) {
  if (!h) return RSGPU_ERR_INVALID_ARGUMENT;
  if (k) *k = h->c.k;
  if (m) *m = h->c.m;
  if (field) *field = h->c.field;
  if (fam) *fam = h->c.family;
  return RSGPU_OK;
}

int rsgpu_codec_matrix(const rsgpu_codec* h, uint8_t* out, size_t out_len) {
  if (!h || !out) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t eb = c.field == RSGPU_FIELD_GF16 ? 2 : 1;
  if (c.matrix.empty()) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (out_len < c.matrix.size() * eb) return RSGPU_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < c.matrix.size(); ++i) {
    if (eb == 2) {
      out[2 * i] = (uint8_t)(c.matrix[i] >> 8);
      out[2 * i + 1] = (uint8_t)c.matrix[i];
    } else {
      out[i] = (uint8_t)c.matrix[i];
    }
  }
  return RSGPU_OK;
}

int rsgpu_encode(const rsgpu_codec* h, const rsgpu_slice* shards, size_t n, void* stream) {
  if (!h || (!shards && n)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  std::vector<rsgpu_cslice> data(c.k);
  for (size_t i = 0; i < c.k; ++i) data[i] = {shards[i].ptr, shards[i].len};
  return encode_sep_impl(c, data.data(), c.k, shards + c.k, c.m, (hipStream_t)stream);
}

int rsgpu_encode_sep(const rsgpu_codec* h, const rsgpu_cslice* data, size_t nd,
                     const rsgpu_slice* parity, size_t np, void* stream) {
  if (!h || (!data && nd) || (!parity && np)) return RSGPU_ERR_INVALID_ARGUMENT;
  if (nd == 0) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (np == 0) return h->c.m ? RSGPU_TOO_FEW_PARITY_SHARDS : RSGPU_ERR_INVALID_ARGUMENT;
  return encode_sep_impl(h->c, data, nd, parity, np, (hipStream_t)stream);
}

int rsgpu_verify_async(const rsgpu_codec* h, const rsgpu_cslice* shards, size_t n,
                       int32_t* device_flag, void* stream) {
  if (!h || !shards || !device_flag) return RSGPU_ERR_INVALID_ARGUMENT;
  return verify_impl(h->c, shards, n, device_flag, (hipStream_t)stream);
}

int rsgpu_verify(const rsgpu_codec* h, const rsgpu_cslice* shards, size_t n, int32_t* ok,
                 void* stream) {
  if (!h || !shards || !ok) return RSGPU_ERR_INVALID_ARGUMENT;
  int32_t* flag = thread_flag(h->c.device);
  if (!flag) return RSGPU_ERR_OUT_OF_MEMORY;
  hipStream_t s = (hipStream_t)stream;
  int rc = verify_impl(h->c, shards, n, flag, s);
  if (rc) return rc;
  int32_t host = 0;
  HIP_TRY(hipMemcpyAsync(&host, flag, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *ok = host == 0 ? 1 : 0;
  return RSGPU_OK;
}

int rsgpu_verify_with_buffer(const rsgpu_codec* h, const rsgpu_cslice* shards, size_t n,
                             const rsgpu_slice* buffer, size_t nb, int32_t* ok, void* stream) {
  if (!h || !shards || !ok || (!buffer && nb)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t total = c.k + c.m;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  if (nb < c.m) return RSGPU_TOO_FEW_BUFFER_SHARDS;
  if (nb > c.m) return RSGPU_TOO_MANY_BUFFER_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  if ((rc = check_slices(buffer, nb))) return rc;
  if (shards[0].len != buffer[0].len) return RSGPU_INCORRECT_SHARD_SIZE;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = encode_sep_impl(c, shards, c.k, buffer, c.m, s))) return rc;
  // compare recomputed parity with the stored parity through the identity rows
  int32_t* flag = thread_flag(c.device);
  if (!flag) return RSGPU_ERR_OUT_OF_MEMORY;
  HIP_TRY(launch_fill_i32(flag, 0, 1, s));
  const bool tower = c.field == RSGPU_FIELD_GF16;
  std::vector<uint16_t> eye(c.m * c.m, 0);
  for (size_t j = 0; j < c.m; ++j) eye[j * c.m + j] = 1;
  std::vector<uint32_t> ht;
  build_tables_host(eye.data(), (int)c.m, (int)c.m, tower, ht);
  DeviceBuffer dt;
  HIP_TRY(dt.alloc(ht.size() * 4));
  HIP_TRY(hipMemcpyAsync(dt.p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, s));
  std::vector<const uint8_t*> in(c.m);
  std::vector<uint8_t*> out(c.m);
  for (size_t j = 0; j < c.m; ++j) {
    in[j] = buffer[j].ptr;
    out[j] = const_cast<uint8_t*>(shards[c.k + j].ptr);
  }
  if (c.m > (size_t)kMaxIn) return RSGPU_ERR_UNSUPPORTED;
  if ((rc = run_apply((const uint32_t*)dt.p, (int)c.m, (int)c.m, tower, in, out, shards[0].len,
                      kVerify, flag, s)))
    return rc;
  int32_t host = 0;
  HIP_TRY(hipMemcpyAsync(&host, flag, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *ok = host == 0 ? 1 : 0;
  return RSGPU_OK;
}

// encode_single_sep (src/core/encode.rs:913-931): parity = c * data for i_data
// == 0, parity ^= c * data otherwise (code_single_slice, encode.rs:676-708).
int rsgpu_encode_single_sep(const rsgpu_codec* h, size_t i_data, rsgpu_cslice single,
                            const rsgpu_slice* parity, size_t np, void* stream) {
  if (!h || (!parity && np)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (i_data >= c.k) return RSGPU_INVALID_INDEX;
  if (np < c.m) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (np > c.m) return RSGPU_TOO_MANY_PARITY_SHARDS;
  int rc = check_slices(parity, np);
  if (rc) return rc;
  if (single.len != parity[0].len) return RSGPU_INCORRECT_SHARD_SIZE;
  if ((rc = check_elem_len(c, single.len))) return rc;
  if (!single.ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t*> out(np);
  for (size_t j = 0; j < np; ++j) {
    if (!parity[j].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
    out[j] = parity[j].ptr;
  }
  const size_t len = single.len;
  return run_apply((const uint32_t*)c.d_parity_tables.p, 1, (int)c.m, c.field == RSGPU_FIELD_GF16,
                   {single.ptr}, out, len, i_data == 0 ? kWrite : kAccumulate, nullptr,
                   (hipStream_t)stream, pick_variant(c, kOpEncode, 0, len), (int)c.k, (int)i_data);
}

// encode_single (src/core/encode.rs:892-911): shards[i_data] into shards[k..].
int rsgpu_encode_single(const rsgpu_codec* h, size_t i_data, const rsgpu_slice* shards, size_t n,
                        void* stream) {
  if (!h || (!shards && n)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (i_data >= c.k) return RSGPU_INVALID_INDEX;
  if (n < c.k + c.m) return RSGPU_TOO_FEW_SHARDS;
  if (n > c.k + c.m) return RSGPU_TOO_MANY_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  return rsgpu_encode_single_sep(h, i_data, {shards[i_data].ptr, shards[i_data].len}, shards + c.k, c.m,
                                 stream);
}

// update (src/core/encode.rs:1048-1132): parity ^= c * (old ^ new) for every
// changed data shard, as one accumulate pass over the inputs [old_i, new_i].
int rsgpu_update(const rsgpu_codec* h, const rsgpu_cslice* old_data, size_t n_old,
                 const rsgpu_cslice* new_data, size_t n_new, const rsgpu_slice* parity, size_t np,
                 void* stream) {
  if (!h || (!old_data && n_old) || (!new_data && n_new) || (!parity && np))
    return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (c.family != RSGPU_FAMILY_CLASSIC) return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  if (n_old < c.k) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (n_old > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  if (np < c.m) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (np > c.m) return RSGPU_TOO_MANY_PARITY_SHARDS;
  if (n_new < c.k) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (n_new > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  int rc = check_slices(old_data, n_old);
  if (rc || (rc = check_slices(parity, np))) return rc;
  const size_t len = old_data[0].len;
  if (parity[0].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  std::vector<size_t> changed;
  for (size_t i = 0; i < n_new; ++i) {
    if (!new_data[i].ptr) continue;  // None: unchanged
    if (new_data[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
    changed.push_back(i);
  }
  if (changed.empty()) return RSGPU_OK;
  if ((rc = check_elem_len(c, len))) return rc;
  for (size_t i = 0; i < n_old; ++i)
    if (!old_data[i].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t*> out(np);
  for (size_t j = 0; j < np; ++j) {
    if (!parity[j].ptr) return RSGPU_ERR_INVALID_ARGUMENT;
    out[j] = parity[j].ptr;
  }
  // parity_j ^= c_ji * old_i ^ c_ji * new_i for every changed i
  const bool tower = c.field == RSGPU_FIELD_GF16;
  const size_t W = (size_t)words_per_coef(tower), nc = 2 * changed.size();
  std::vector<const uint8_t*> in;
  for (size_t i : changed) {
    in.push_back(old_data[i].ptr);
    in.push_back(new_data[i].ptr);
  }
  hipStream_t s = (hipStream_t)stream;
  const int variant = pick_variant(c, kOpEncode, 0, len);
  if (apply_inline_fits((int)nc, (int)c.m, tower, in.data(), out.data(), len)) {
    // one launch, the gathered columns [i, i] inline in its arguments
    std::vector<uint32_t> tab(c.m * nc * W);
    for (size_t q = 0; q < changed.size(); ++q)
      for (size_t j = 0; j < c.m; ++j)
        for (int t = 0; t < 2; ++t)
          std::memcpy(&tab[(j * nc + 2 * q + t) * W], &c.h_parity_tables[(j * c.k + changed[q]) * W], W * 4);
    return run_apply(nullptr, (int)nc, (int)c.m, tower, in.data(), out.data(), len, kAccumulate, nullptr, s,
                     variant, -1, 0, tab.data());
  }
  // otherwise column i of the codec's own device tables, old then new: two
  // accumulating launches per changed shard, nothing allocated
  for (size_t q = 0; q < changed.size(); ++q)
    for (int t = 0; t < 2; ++t)
      if ((rc = run_apply((const uint32_t*)c.d_parity_tables.p, 1, (int)c.m, tower, &in[2 * q + t], out.data(),
                          len, kAccumulate, nullptr, s, variant, (int)c.k, (int)changed[q])))
        return rc;
  return RSGPU_OK;
}

// ------------------------------------------------- format identity, split, join
// EncodingFormat::stable_id_bytes (src/core/format.rs:133-219), version 1:
// magic "RSEFMT\0\0", u16 version, u8 family, u32 + field id, u32 element width,
// u64 field order, u64 data, u64 parity, u64 rows, u64 cols, then the matrix
// coefficients row-major (Leopard formats carry no matrix: 0 x 0).  Integers are
// big-endian.
int rsgpu_codec_format_id(const rsgpu_codec* h, uint8_t* out, size_t cap, size_t* len) {
  if (!h || !len || (!out && cap)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const bool tower = c.field == RSGPU_FIELD_GF16 && c.family == RSGPU_FAMILY_CLASSIC;
  // galois_8::Field / galois_16::Field stable_format (galois_8/mod.rs:73-83, galois_16.rs:27-37);
  // Leopard codecs are ReedSolomon<galois_8::Field> over byte input.
  const char* field_id = tower ? "rustfs.gf16.tower-11d-x2-02x-80.ax-ac.v1" : "rustfs.gf8.poly-11d.u8.v1";
  const uint32_t width = tower ? 2 : 1;
  const uint64_t order = tower ? 65536 : 256;
  const bool leo = c.family != RSGPU_FAMILY_CLASSIC;
  const uint64_t rows = leo ? 0 : c.k + c.m, cols = leo ? 0 : c.k;
  std::vector<uint8_t> b;
  auto put = [&](uint64_t v, int n) {
    for (int i = n - 1; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
  };
  const char magic[8] = {'R', 'S', 'E', 'F', 'M', 'T', 0, 0};
  b.insert(b.end(), magic, magic + 8);
  put(1, 2);
  b.push_back((uint8_t)(c.family == RSGPU_FAMILY_CLASSIC ? 0 : c.family == RSGPU_FAMILY_LEOPARD_GF8 ? 1 : 2));
  const size_t fl = std::strlen(field_id);
  put(fl, 4);
  b.insert(b.end(), field_id, field_id + fl);
  put(width, 4);
  for (uint64_t v : {order, (uint64_t)c.k, (uint64_t)c.m, rows, cols}) put(v, 8);
  for (uint64_t i = 0; i < rows * cols; ++i) {
    const uint16_t e = c.matrix[i];
    if (tower) {  // [x coefficient, constant] (galois_16.rs:31-37)
      b.push_back((uint8_t)(e >> 8));
      b.push_back((uint8_t)e);
    } else {
      b.push_back((uint8_t)e);
    }
  }
  *len = b.size();
  if (cap < b.size()) return out ? RSGPU_ERR_INVALID_ARGUMENT : RSGPU_OK;
  std::memcpy(out, b.data(), b.size());
  return RSGPU_OK;
}

// leopard_aligned_shard_len (src/core/leopard.rs:75-98)
static size_t leopard_aligned_len(size_t data_len, size_t shards) {
  if (shards == 0) shards = 1;
  const size_t per = data_len / shards + (data_len % shards ? 1 : 0);
  size_t r = per % 64 ? per + (64 - per % 64) : per;
  if (r < per) r = SIZE_MAX;  // saturate
  r = std::max<size_t>(r, 64);
  return r - r % 64;
}

// Shard length in bytes that split() gives `data_len` bytes (codec.rs:476-505):
// classic ceil(elements / k) elements, Leopard a nonzero multiple of 64 bytes.
int rsgpu_split_len(const rsgpu_codec* h, size_t data_len, size_t* shard_len) {
  if (!h || !shard_len) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  const size_t eb = c.field == RSGPU_FIELD_GF16 && c.family == RSGPU_FAMILY_CLASSIC ? 2 : 1;
  if (data_len % eb) return RSGPU_ERR_INVALID_ARGUMENT;
  if (c.family != RSGPU_FAMILY_CLASSIC) {
    *shard_len = leopard_aligned_len(data_len, c.k);
  } else {
    const size_t el = data_len / eb;
    *shard_len = (el / c.k + (el % c.k ? 1 : 0)) * eb;
  }
  return RSGPU_OK;
}

// ReedSolomon::split (codec.rs:476-505) into caller shards of split_len bytes:
// shard i takes bytes [i * len, (i + 1) * len) of data, zero-padded.
int rsgpu_split(const rsgpu_codec* h, const uint8_t* data, size_t data_len, const rsgpu_slice* shards,
                size_t n, void* stream) {
  if (!h || (!data && data_len) || (!shards && n)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  size_t len = 0;
  int rc = rsgpu_split_len(h, data_len, &len);
  if (rc) return rc;
  if (n != c.k) return RSGPU_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < n; ++i)
    if (shards[i].len != len || (!shards[i].ptr && len)) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  for (size_t i = 0; i < n; ++i) {
    const size_t start = i * len, end = std::min(start + len, data_len);
    const size_t have = start < data_len ? end - start : 0;
    if (have) HIP_TRY(hipMemcpyAsync(shards[i].ptr, data + start, have, hipMemcpyDeviceToDevice, s));
    if (have < len) HIP_TRY(launch_fill_u8(shards[i].ptr + have, 0, len - have, s));
  }
  return RSGPU_OK;
}

// ReedSolomon::join (codec.rs:507-538): the first min(out_len, k * len) bytes of
// the data shards, in order; *written receives that count.
int rsgpu_join(const rsgpu_codec* h, const rsgpu_cslice* shards, size_t n, size_t out_len, uint8_t* out,
               size_t* written, void* stream) {
  if (!h || (!shards && n) || !written) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (n < c.k) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (n > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  int rc = check_slices(shards, n);
  if (rc) return rc;
  const size_t len = shards[0].len;
  const size_t target = std::min(out_len, len * n);
  if (!out && target) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  size_t done = 0;
  for (size_t i = 0; i < n && done < target; ++i) {
    const size_t take = std::min(len, target - done);
    HIP_TRY(hipMemcpyAsync(out + done, shards[i].ptr, take, hipMemcpyDeviceToDevice, s));
    done += take;
  }
  *written = target;
  return RSGPU_OK;
}

int rsgpu_variant_count(void) { return kNumVariants; }

int rsgpu_dispatch_dims(int32_t* n_ops, int32_t* n_buckets, int32_t* bucket0) {
  if (!n_ops || !n_buckets || !bucket0) return RSGPU_ERR_INVALID_ARGUMENT;
  *n_ops = kNumOps;
  *n_buckets = kNumBuckets;
  *bucket0 = kBucket0;
  return RSGPU_OK;
}

int rsgpu_variant_info(int32_t variant, int32_t* kind, int32_t* lanes, int32_t* vectors_per_lane,
                       int32_t* rows_per_group, int32_t* leopard_tile_words) {
  if (variant < 0 || variant >= kNumVariants || !kind || !lanes || !vectors_per_lane ||
      !rows_per_group || !leopard_tile_words)
    return RSGPU_ERR_INVALID_ARGUMENT;
  *kind = variant < kNumApplyVariants ? 0 : 1;
  *lanes = variant < kNumApplyVariants ? kTileShapes[variant % kNumTiles].lanes : 256;
  *vectors_per_lane = variant < kNumApplyVariants ? kTileShapes[variant % kNumTiles].nv : 0;
  *rows_per_group = variant < kNumApplyVariants ? kRowSplit[variant / kNumTiles] : 0;
  *leopard_tile_words = variant < kNumApplyVariants ? 0 : kLeoTileWords[variant - kNumApplyVariants];
  return RSGPU_OK;
}

int rsgpu_variant_valid(const rsgpu_codec* h, int32_t op, size_t erasures, int32_t variant) {
  if (!h || op < 0 || op >= kNumOps) return 0;
  return variant_valid(h->c, op, erasures, variant) ? 1 : 0;
}

int rsgpu_codec_set_variant(rsgpu_codec* h, int32_t variant) {
  if (!h || variant < -1 || variant >= kNumVariants) return RSGPU_ERR_INVALID_ARGUMENT;
  h->c.forced_variant.store(variant, std::memory_order_relaxed);
  return RSGPU_OK;
}

int rsgpu_codec_set_variant_table(rsgpu_codec* h, const int16_t* table, size_t n_erasures) {
  if (!h || (table && n_erasures == 0)) return RSGPU_ERR_INVALID_ARGUMENT;
  std::shared_ptr<const VariantTable> t;
  if (table) {
    auto nt = std::make_shared<VariantTable>();
    nt->n_erasures = (int)n_erasures;
    nt->v.assign(table, table + (size_t)kNumOps * n_erasures * kNumBuckets);
    for (int16_t v : nt->v)
      if (v < -1 || v >= kNumVariants) return RSGPU_ERR_INVALID_ARGUMENT;
    t = nt;
  }
  std::atomic_store(&h->c.variant_table, t);
  return RSGPU_OK;
}

// Network inputs of one call; the same 14 features as features() in
// rsgpu/dispatch.py (float64 arithmetic, rounded to float at the end).
static void dispatch_features(int kind, size_t k, size_t m, size_t len, int op, size_t e, float* f) {
  const double l2 = (std::log2((double)std::max<size_t>(len, 1)) - 16.0) / 6.0;
  const double v[14] = {kind == 1 ? 1.0 : 0.0, kind == 2 ? 1.0 : 0.0, kind == 3 ? 1.0 : 0.0,
                        std::log2((double)k) / 12.0, m > 0 ? std::log2((double)m) / 10.0 : 0.0,
                        std::log2((double)(k + m)) / 12.0, l2, l2 * l2,
                        op == 0 ? 1.0 : 0.0, op == 1 ? 1.0 : 0.0, op == 2 ? 1.0 : 0.0,
                        (double)e / (double)std::max<size_t>(m, 1), std::log2(1.0 + (double)e) / 10.0,
                        (std::log2((double)std::max<size_t>(len, 1)) + std::log2((double)(k + m)) - 24.0) / 6.0};
  for (int i = 0; i < 14; ++i) f[i] = (float)v[i];
}

// Fills the codec's variant table from the dispatcher network: one forward pass
// over every (op, erasures, size bucket) and the argmin over the variants valid
// for that call.  A network with one output more than there are variants
// predicts the default tiling there too; where that is the argmin the entry
// stays -1 (default tiling).  `envelope` ([4][6] int64: k_min, k_max, m_min, m_max, len_min,
// len_max per kind classic-GF8, classic-GF16, Leopard GF8, Leopard GF16; a row
// with k_max == 0 is unknown) limits it to trained shapes; other entries keep the
// default tiling.
int rsgpu_codec_install_dispatch(rsgpu_codec* h, rsgpu_mlp* net, const int64_t* envelope, void* stream) {
  if (!h || !net) return RSGPU_ERR_INVALID_ARGUMENT;
  Codec& c = h->c;
  int32_t n_in = 0, n_hidden = 0, n_out = 0;
  size_t n_params = 0;
  int rc = rsgpu_mlp_dims(net, &n_in, &n_hidden, &n_out, &n_params);
  if (rc) return rc;
  if (n_in != 14 || (n_out != kNumVariants && n_out != kNumVariants + 1)) return RSGPU_ERR_INVALID_ARGUMENT;
  const int kind = c.family == RSGPU_FAMILY_LEOPARD_GF8    ? 2
                   : c.family == RSGPU_FAMILY_LEOPARD_GF16 ? 3
                   : c.field == RSGPU_FIELD_GF16           ? 1
                                                           : 0;
  struct Row {
    int op;
    size_t e;
    int b;
    size_t len;
  };
  std::vector<Row> rows;
  for (int op = 0; op < kNumOps; ++op)
    for (size_t e = 0; e <= c.m; ++e) {
      if ((op < 2) != (e == 0)) continue;
      for (int b = 0; b < kNumBuckets; ++b)
        rows.push_back({op, e, b, (size_t)std::llround(std::pow(2.0, (b + kBucket0) / 2.0))});
    }
  std::vector<float> x(rows.size() * 14), y(rows.size() * (size_t)n_out);
  for (size_t i = 0; i < rows.size(); ++i)
    dispatch_features(kind, c.k, c.m, rows[i].len, rows[i].op, rows[i].e, &x[i * 14]);
  hipStream_t s = (hipStream_t)stream;
  DeviceBuffer dx, dy;
  HIP_TRY(dx.alloc(x.size() * 4));
  HIP_TRY(dy.alloc(y.size() * 4));
  HIP_TRY(hipMemcpyAsync(dx.p, x.data(), x.size() * 4, hipMemcpyHostToDevice, s));
  if ((rc = rsgpu_mlp_forward(net, (const float*)dx.p, rows.size(), (float*)dy.p, s))) return rc;
  HIP_TRY(hipMemcpyAsync(y.data(), dy.p, y.size() * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const int64_t* env = envelope ? envelope + 6 * kind : nullptr;
  auto inside = [&](size_t len) {
    if (!env) return true;
    if (env[1] == 0) return false;
    return (int64_t)c.k >= env[0] && (int64_t)c.k <= env[1] && (int64_t)c.m >= env[2] &&
           (int64_t)c.m <= env[3] && (int64_t)len >= env[4] / 2 && (int64_t)len <= env[5] * 2;
  };
  const size_t ne = c.m + 1;
  std::vector<int16_t> tab((size_t)kNumOps * ne * kNumBuckets, -1);
  for (size_t i = 0; i < rows.size(); ++i) {
    const Row& r = rows[i];
    if (!inside(r.len)) continue;
    int best = -1;
    float bv = 0;
    for (int v = 0; v < kNumVariants; ++v) {
      if (!variant_valid(c, r.op, r.e, v)) continue;
      const float t = y[i * n_out + v];
      if (best < 0 || t < bv) best = v, bv = t;
    }
    if (n_out == kNumVariants + 1 && (best < 0 || y[i * n_out + kNumVariants] < bv)) best = -1;
    tab[((size_t)r.op * ne + r.e) * kNumBuckets + r.b] = (int16_t)best;
  }
  // encode / verify rows are keyed by e = 0; copy them so every erasure index resolves
  for (int op = 0; op < 2; ++op)
    for (size_t e = 1; e < ne; ++e)
      std::memcpy(&tab[((size_t)op * ne + e) * kNumBuckets], &tab[(size_t)op * ne * kNumBuckets],
                  kNumBuckets * sizeof(int16_t));
  return rsgpu_codec_set_variant_table(h, tab.data(), ne);
}

int rsgpu_codec_variant_for(const rsgpu_codec* h, int32_t op, size_t erasures, size_t shard_len,
                            int32_t* variant) {
  if (!h || !variant || op < 0 || op >= kNumOps) return RSGPU_ERR_INVALID_ARGUMENT;
  *variant = pick_variant(h->c, op, erasures, shard_len);
  return RSGPU_OK;
}

int rsgpu_reconstruct(const rsgpu_codec* h, rsgpu_slot* shards, size_t n, void* stream) {
  if (!h || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  return reconstruct_impl(h->c, shards, n, 0, nullptr, 0, (hipStream_t)stream);
}

int rsgpu_reconstruct_data(const rsgpu_codec* h, rsgpu_slot* shards, size_t n, void* stream) {
  if (!h || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  return reconstruct_impl(h->c, shards, n, 1, nullptr, 0, (hipStream_t)stream);
}

int rsgpu_reconstruct_some(const rsgpu_codec* h, rsgpu_slot* shards, size_t n,
                           const uint8_t* required, size_t n_required, void* stream) {
  if (!h || !shards || (!required && n_required)) return RSGPU_ERR_INVALID_ARGUMENT;
  return reconstruct_impl(h->c, shards, n, 2, required, n_required, (hipStream_t)stream);
}

int rsgpu_prepare_reconstruct_workspace(const rsgpu_codec* h, const rsgpu_slot* shards, size_t n, void* stream,
                                        rsgpu_reconstruct_workspace** out) {
  if (!h || !shards || !out) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  return prepare_workspace(h->c, shards, n, 0, nullptr, 0, (hipStream_t)stream, out);
}

int rsgpu_prepare_reconstruct_data_workspace(const rsgpu_codec* h, const rsgpu_slot* shards, size_t n,
                                             void* stream, rsgpu_reconstruct_workspace** out) {
  if (!h || !shards || !out) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  return prepare_workspace(h->c, shards, n, 1, nullptr, 0, (hipStream_t)stream, out);
}

int rsgpu_prepare_reconstruct_some_workspace(const rsgpu_codec* h, const rsgpu_slot* shards, size_t n,
                                             const uint8_t* required, size_t n_required, void* stream,
                                             rsgpu_reconstruct_workspace** out) {
  if (!h || !shards || !out || (!required && n_required)) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  return prepare_workspace(h->c, shards, n, 2, required, n_required, (hipStream_t)stream, out);
}

int rsgpu_reconstruct_with_workspace(const rsgpu_codec* h, rsgpu_slot* shards, size_t n,
                                     const rsgpu_reconstruct_workspace* w, void* stream) {
  if (!h || !shards || !w) return RSGPU_ERR_INVALID_ARGUMENT;
  return reconstruct_with_workspace_impl(h->c, shards, n, *w, (hipStream_t)stream);
}

int rsgpu_reconstruct_workspace_targets(const rsgpu_reconstruct_workspace* w, size_t* targets, size_t cap,
                                        size_t* n_data, size_t* n_parity) {
  if (!w || (!targets && cap) || !n_data || !n_parity) return RSGPU_ERR_INVALID_ARGUMENT;
  size_t nd = 0;
  for (int t : w->targets) nd += (size_t)t < w->k ? 1 : 0;
  *n_data = nd;
  *n_parity = w->targets.size() - nd;
  for (size_t i = 0; i < w->targets.size() && i < cap; ++i) targets[i] = (size_t)w->targets[i];
  return RSGPU_OK;
}

void rsgpu_reconstruct_workspace_free(rsgpu_reconstruct_workspace* w) { delete w; }

int rsgpu_decode_idx(const rsgpu_codec* h, const rsgpu_slice* dst, size_t n_dst, const uint8_t* expect_input,
                     size_t n_expect, const rsgpu_cslice* input, size_t n_input, void* stream) {
  if (!h || (!dst && n_dst) || (!input && n_input)) return RSGPU_ERR_INVALID_ARGUMENT;
  return decode_idx_impl(h->c, dst, n_dst, expect_input, n_expect, input, n_input, (hipStream_t)stream);
}

int rsgpu_reconstruct_noalloc(const rsgpu_codec* h, rsgpu_slot* shards, size_t n, int32_t data_only,
                              void* stream) {
  if (!h || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  return reconstruct_noalloc_impl(h->c, shards, n, data_only ? 1 : 0, (hipStream_t)stream);
}

// Coefficient tables of rsgpu_apply_rows, kept on the device by (device, field,
// shape, row bytes): callers pass the same parity rows call after call, which
// then only enqueue the kernels (no allocation, upload or synchronisation).  The
// 64 most recently used row sets are kept; an evicted set is freed (after the
// cache lock is released) by hipFree, which waits for the device, so no launch
// still reads it.
static int apply_rows_tables(int32_t field, const uint8_t* rows, size_t r, size_t k,
                             std::shared_ptr<const DeviceBuffer>& out) {
  const bool tower = field == RSGPU_FIELD_GF16;
  const size_t nbytes = r * k * (tower ? 2 : 1);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::string key(reinterpret_cast<const char*>(&dev), sizeof dev);
  const uint64_t shape[3] = {(uint64_t)field, (uint64_t)r, (uint64_t)k};
  key.append(reinterpret_cast<const char*>(shape), sizeof shape);
  key.append(reinterpret_cast<const char*>(rows), nbytes);
  // never destroyed: no hipFree may run in static destruction at process exit
  static std::mutex& mu = *new std::mutex;
  static auto& cache = *new LruMap<const DeviceBuffer>{64};
  {
    std::lock_guard<std::mutex> g(mu);
    if ((out = cache.get(key))) return RSGPU_OK;
  }
  std::vector<uint16_t> rr(r * k);
  for (size_t i = 0; i < r * k; ++i) rr[i] = tower ? (uint16_t)((rows[2 * i] << 8) | rows[2 * i + 1]) : rows[i];
  std::vector<uint32_t> ht;
  build_tables_host(rr.data(), (int)r, (int)k, tower, ht);
  auto dt = std::make_shared<DeviceBuffer>();
  HIP_TRY(dt->alloc(ht.size() * 4));
  HIP_TRY(hipMemcpy(dt->p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice));
  std::vector<std::shared_ptr<const DeviceBuffer>> evicted;  // freed after the lock
  std::lock_guard<std::mutex> g(mu);
  cache.put(key, dt, evicted);
  out = dt;
  return RSGPU_OK;
}

int rsgpu_apply_rows(int32_t field, const uint8_t* rows, size_t r, size_t k, const rsgpu_cslice* in,
                     const rsgpu_slice* out, void* stream) {
  if (!rows || !in || !out || r == 0 || k == 0) return RSGPU_ERR_INVALID_ARGUMENT;
  if (field != RSGPU_FIELD_GF8 && field != RSGPU_FIELD_GF16) return RSGPU_ERR_INVALID_ARGUMENT;
  // SelectedEncodeKernel::execute validation (codegen.rs:46-79): lengths first
  const size_t len = in[0].len;
  if (len == 0) return RSGPU_EMPTY_SHARD;
  for (size_t i = 0; i < k; ++i)
    if (in[i].len != len || !in[i].ptr) return RSGPU_INCORRECT_SHARD_SIZE;
  for (size_t j = 0; j < r; ++j)
    if (out[j].len != len || !out[j].ptr) return RSGPU_INCORRECT_SHARD_SIZE;
  const bool tower = field == RSGPU_FIELD_GF16;
  if (tower && (len & 1)) return RSGPU_INCORRECT_SHARD_SIZE;
  hipStream_t s = (hipStream_t)stream;
  std::shared_ptr<const DeviceBuffer> dt;
  int rc = apply_rows_tables(field, rows, r, k, dt);
  if (rc) return rc;
  std::vector<const uint8_t*> vin(k);
  std::vector<uint8_t*> vout(r);
  for (size_t i = 0; i < k; ++i) vin[i] = in[i].ptr;
  for (size_t j = 0; j < r; ++j) vout[j] = out[j].ptr;
  return run_apply((const uint32_t*)dt->p, (int)k, (int)r, tower, vin, vout, len, kWrite, nullptr, s);
}

// ---- Leopard plug points (leopard_gf8_encode / leopard_gf16_encode /
// reconstruct_with_tables, src/core/leopard.rs:277-317): the FFT engines by
// (field, k, m), without a codec handle.  One Leopard codec per (device, field,
// k, m) is built on first use and kept for the process, as the reference keeps
// its tables and per-shape plans.
static int leopard_plug_codec(int32_t bits, size_t k, size_t m, const rsgpu_codec** out) {
  if (bits != 8 && bits != 16) return RSGPU_ERR_INVALID_ARGUMENT;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  // never destroyed: a host thread may still call in while the process exits
  static std::mutex& mu = *new std::mutex;
  static auto& codecs = *new std::map<std::tuple<int, int, size_t, size_t>, rsgpu_codec*>;
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_tuple(dev, (int)bits, k, m);
  auto it = codecs.find(key);
  if (it == codecs.end()) {
    rsgpu_options o;
    rsgpu_options_default(&o);
    o.codec_family = bits == 8 ? RSGPU_FAMILY_LEOPARD_GF8 : RSGPU_FAMILY_LEOPARD_GF16;
    rsgpu_codec* c = nullptr;
    int rc = rsgpu_codec_new(k, m, &o, &c);
    if (rc) return rc;
    it = codecs.emplace(key, c).first;
  }
  *out = it->second;
  return RSGPU_OK;
}

int rsgpu_leopard_encode(int32_t bits, size_t data_shards, size_t parity_shards, const rsgpu_cslice* data,
                         size_t n_data, const rsgpu_slice* parity, size_t n_parity, void* stream) {
  if ((!data && n_data) || (!parity && n_parity)) return RSGPU_ERR_INVALID_ARGUMENT;
  // encode_with_tables (leopard_gf8/encode.rs:201-210), then the driver's length rule
  if (n_data != data_shards || n_parity != parity_shards || n_data == 0) return RSGPU_TOO_FEW_SHARDS;
  const size_t len = data[0].len;
  if (len == 0 || len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
  for (size_t i = 0; i < n_data; ++i)
    if (data[i].len != len || !data[i].ptr) return RSGPU_INCORRECT_SHARD_SIZE;
  for (size_t j = 0; j < n_parity; ++j)
    if (parity[j].len != len || !parity[j].ptr) return RSGPU_INCORRECT_SHARD_SIZE;
  const rsgpu_codec* h = nullptr;
  int rc = leopard_plug_codec(bits, data_shards, parity_shards, &h);
  if (rc) return rc;
  return rsgpu_encode_sep(h, data, n_data, parity, n_parity, stream);
}

int rsgpu_leopard_reconstruct(int32_t bits, size_t data_shards, size_t parity_shards, const uint8_t* present,
                              const rsgpu_cslice* inputs, const rsgpu_slice* outputs, void* stream) {
  if (!present || !inputs || !outputs) return RSGPU_ERR_INVALID_ARGUMENT;
  // validate_leopard_decode_buffers (src/core/reconstruct.rs:140-185)
  if (data_shards == 0) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (parity_shards == 0) return RSGPU_TOO_FEW_PARITY_SHARDS;
  const size_t total = data_shards + parity_shards;
  size_t len = 0, n_present = 0;
  for (size_t i = 0; i < total; ++i) {
    if (!present[i]) continue;
    if (!inputs[i].ptr) return RSGPU_TOO_FEW_SHARDS_PRESENT;
    if (inputs[i].len == 0) return RSGPU_EMPTY_SHARD;
    if (inputs[i].len % 64 || (len && inputs[i].len != len)) return RSGPU_INCORRECT_SHARD_SIZE;
    len = inputs[i].len;
    ++n_present;
  }
  if (n_present < data_shards || !len) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  bool any_output = false;
  for (size_t i = 0; i < total; ++i) {
    if (outputs[i].len != 0 && outputs[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
    any_output |= outputs[i].len != 0;
  }
  // decode.rs:305-309: nothing missing or nothing requested
  if (n_present == total || !any_output) return RSGPU_OK;
  const rsgpu_codec* h = nullptr;
  int rc = leopard_plug_codec(bits, data_shards, parity_shards, &h);
  if (rc) return rc;
  std::vector<rsgpu_slot> slots(total);
  std::vector<uint8_t> required(total, 0);
  for (size_t i = 0; i < total; ++i) {
    if (present[i]) {
      slots[i] = {const_cast<uint8_t*>(inputs[i].ptr), len, 1};
    } else {
      slots[i] = {outputs[i].ptr, len, 0};
      required[i] = outputs[i].len != 0 ? 1 : 0;
    }
  }
  return rsgpu_reconstruct_some(h, slots.data(), total, required.data(), total, stream);
}

// ---- host-pointer variants
int rsgpu_cache_stats(const rsgpu_codec* h, uint64_t* requests, uint64_t* hits, uint64_t* misses,
                      uint64_t* entries) {
  if (!h) return RSGPU_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> g(h->c.mu);
  if (requests) *requests = h->c.requests;
  if (hits) *hits = h->c.hits;
  if (misses) *misses = h->c.misses;
  if (entries) *entries = h->c.index.size();
  return RSGPU_OK;
}

const char* rsgpu_status_name(int s) {
  static const char* names[] = {"Ok",
                                "TooFewShards",
                                "TooManyShards",
                                "TooFewDataShards",
                                "TooManyDataShards",
                                "TooFewParityShards",
                                "TooManyParityShards",
                                "TooFewBufferShards",
                                "TooManyBufferShards",
                                "IncorrectShardSize",
                                "TooFewShardsPresent",
                                "EmptyShard",
                                "InvalidShardFlags",
                                "InvalidIndex",
                                "InvalidCustomMatrix",
                                "UnsupportedCodecFamily",
                                "UnsupportedLeopardPrototype"};
  if (s >= 0 && s <= 16) return names[s];
  switch (s) {
    case RSGPU_ERR_HIP: return "HipError";
    case RSGPU_ERR_INVALID_ARGUMENT: return "InvalidArgument";
    case RSGPU_ERR_NO_DEVICE: return "NoDevice";
    case RSGPU_ERR_UNSUPPORTED: return "Unsupported";
    case RSGPU_ERR_OUT_OF_MEMORY: return "OutOfMemory";
    case RSGPU_ERR_IO: return "IoError";
  }
  return "Unknown";
}

int rsgpu_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RSGPU_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return RSGPU_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(device));
  uint8_t* scratch_ptr = nullptr;  // the allocation-free entry point's scratch address
  return device < kMaxScratchDevices ? scratch_for_device(device, &scratch_ptr) : RSGPU_OK;
}

const char* rsgpu_last_hip_error(void) { return t_hip_error.c_str(); }

}  // extern "C"
