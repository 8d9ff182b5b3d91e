// leopard.hip -- Leopard GF(2^8) and GF(2^16) FFT codecs on gfx950.
//
// Field tables, FFT skews and the transform schedules follow
// src/core/leopard_gf8/{tables,mod,encode,decode}.rs and
// src/core/leopard_gf16/{tables,mod,encode,decode,ops}.rs.
//
// Device mapping.  Every Leopard operation is independent per column: a byte
// (GF8) or a split-layout element (GF16, ops.rs:520-568: in each 64-byte block
// element i = byte[i] | byte[i+32] << 8).  A workgroup takes a column tile of C
// "words" -- a dword of 4 bytes (GF8), or the pair (low-byte dword, high-byte
// dword) of 4 elements (GF16) -- from every work lane, keeps the whole lane x
// tile matrix in LDS, and runs all transform layers of the plan on it with one
// barrier per layer.  Items of a layer are (butterfly, word) pairs with the word
// fastest, so the lanes of a wave share a butterfly's twiddle tables.
//
// Multiplication by a fixed element is GF(2)-linear in the element bits (the
// Cantor-basis representation is a linear change of basis), so it uses the same
// three-v_perm byte tables as the classic kernels: one table per byte map
// (GF16: low->low, low->high, high->low, high->high).  The reference's
// butterfly shortcut for log_m == 2^n - 1 (ops.rs:466-498 GF8, :382-452 GF16)
// is exactly a multiplication by zero, so such twiddles get an all-zero table.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <map>

#include "kernel_common.h"
#include "leopard.h"

namespace rsgpu {

// ------------------------------------------------------------- host tables
struct LeoField {
  bool gf16;
  unsigned order, modulus;
  std::vector<uint32_t> log, exp;  // exp has 2*order entries for GF16 (ops.rs index range)
  std::vector<uint32_t> skew, walsh;
  uint32_t add_mod(uint32_t a, uint32_t b) const {
    uint32_t sum = a + b;
    return (sum + (sum >> (gf16 ? 16 : 8))) & (order - 1);
  }
  uint32_t mul_log(uint32_t a, uint32_t log_b) const {
    if (a == 0) return 0;
    return exp[add_mod(log[a], log_b)];
  }
};

// init_luts8/16 + init_fft_skew8/16 (leopard_gf8/tables.rs:25-104,
// leopard_gf16/tables.rs:22-116).
static void build_field(LeoField& F, bool gf16) {
  F.gf16 = gf16;
  const unsigned bits = gf16 ? 16 : 8;
  F.order = 1u << bits;
  F.modulus = F.order - 1;
  const uint32_t poly = gf16 ? 0x1002Du : 0x11Du;
  static const uint32_t cantor8[8] = {1, 214, 152, 146, 86, 200, 88, 230};
  static const uint32_t cantor16[16] = {0x0001, 0xACCA, 0x3C0E, 0x163E, 0xC582, 0xED2E, 0x914C, 0x4012,
                                        0x6C98, 0x10D8, 0x6A72, 0xB900, 0xFDB8, 0xFB34, 0xFF38, 0x991E};
  const uint32_t* cantor = gf16 ? cantor16 : cantor8;
  F.exp.assign(gf16 ? 2 * F.order : F.order, 0);
  F.log.assign(F.order, 0);
  uint32_t state = 1;
  for (uint32_t i = 0; i < F.modulus; ++i) {
    F.exp[state] = i;
    state <<= 1;
    if (state >= F.order) state ^= poly;
  }
  F.exp[0] = F.modulus;
  F.log[0] = 0;
  for (unsigned i = 0; i < bits; ++i) {
    uint32_t width = 1u << i;
    for (uint32_t j = 0; j < width; ++j) F.log[j + width] = F.log[j] ^ cantor[i];
  }
  for (uint32_t i = 0; i < F.order; ++i) F.log[i] = F.exp[F.log[i]];
  for (uint32_t i = 0; i < F.order; ++i) F.exp[F.log[i]] = i;
  F.exp[F.modulus] = F.exp[0];

  std::vector<uint32_t> temp(bits - 1);
  for (unsigned i = 1; i < bits; ++i) temp[i - 1] = 1u << i;
  F.skew.assign(F.modulus, 0);
  for (unsigned m = 0; m + 1 < bits; ++m) {
    uint32_t step = 1u << (m + 1);
    F.skew[(1u << m) - 1] = 0;
    for (unsigned i = m; i + 1 < bits; ++i) {
      uint32_t s = 1u << (i + 1);
      for (uint32_t j = (1u << m) - 1; j < s; j += step) F.skew[j + s] = F.skew[j] ^ temp[i];
    }
    uint32_t prod = F.mul_log(temp[m], F.log[temp[m] ^ 1]);
    temp[m] = F.modulus - F.log[prod];
    for (unsigned i = m + 1; i + 1 < bits; ++i) {
      uint32_t sum = F.add_mod(F.log[temp[i] ^ 1], temp[m]);
      temp[i] = F.mul_log(temp[i], sum);
    }
  }
  for (uint32_t i = 0; i < F.modulus; ++i) F.skew[i] = F.log[F.skew[i]];
}

static const LeoField& field8() {
  static LeoField f;
  static std::once_flag once;
  std::call_once(once, [] { build_field(f, false); });
  return f;
}
static const LeoField& field16() {
  static LeoField f;
  static std::once_flag once;
  std::call_once(once, [] { build_field(f, true); });
  return f;
}

// Byte-permute tables of "multiply by g^log_m": GF8 one 8-dword table, GF16 four
// (low->low, low->high, high->low, high->high).  `twiddle` applies the
// butterfly shortcut (log_m == modulus -> zero).
constexpr int kLeoTabWords8 = kTabWords;
constexpr int kLeoTabWords16 = 4 * kTabWords;

static void pack_table(const uint8_t f[256], uint32_t* w) {
  uint32_t v[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    v[0] |= (uint32_t)f[i] << (8 * i);
    v[1] |= (uint32_t)f[i + 4] << (8 * i);
    v[2] |= (uint32_t)f[8 * i] << (8 * i);
    v[3] |= (uint32_t)f[8 * (i + 4)] << (8 * i);
    v[4] |= (uint32_t)f[64 * i] << (8 * i);
  }
  for (int i = 0; i < 5; ++i) w[i] = v[i];
  w[5] = w[6] = w[7] = 0;
}

static void leo_table(const LeoField& F, uint32_t log_m, bool twiddle, uint32_t* out) {
  const bool zero = twiddle && log_m == F.modulus;
  if (!F.gf16) {
    uint8_t f[256];
    for (int v = 0; v < 256; ++v) f[v] = zero ? 0 : (uint8_t)F.mul_log((uint32_t)v, log_m);
    pack_table(f, out);
    return;
  }
  uint8_t ll[256], lh[256], hl[256], hh[256];
  for (int v = 0; v < 256; ++v) {
    uint32_t lo = zero ? 0 : F.mul_log((uint32_t)v, log_m);
    uint32_t hi = zero ? 0 : F.mul_log((uint32_t)v << 8, log_m);
    ll[v] = lo & 0xff;
    lh[v] = lo >> 8;
    hl[v] = hi & 0xff;
    hh[v] = hi >> 8;
  }
  pack_table(ll, out);
  pack_table(lh, out + kTabWords);
  pack_table(hl, out + 2 * kTabWords);
  pack_table(hh, out + 3 * kTabWords);
}

uint16_t leopard16_one() { return (uint16_t)field16().exp[0]; }

// pack_table needs f at 20 arguments only: 0..7, 8*i, 64*i.
void leopard16_apply_tables(const uint16_t* coef, size_t n, uint32_t* out) {
  const LeoField& F = field16();
  static const int kArgs[] = {1, 2, 3, 4, 5, 6, 7, 8, 16, 24, 32, 40, 48, 56, 64, 128, 192};
  for (size_t t = 0; t < n; ++t) {
    uint32_t* dst = out + t * kLeoTabWords16;
    if (coef[t] == 0) {
      std::fill(dst, dst + kLeoTabWords16, 0u);
      continue;
    }
    const uint32_t lg = F.log[coef[t]];
    uint8_t ll[256] = {}, lh[256] = {}, hl[256] = {}, hh[256] = {};
    for (int v : kArgs) {
      const uint32_t lo = F.mul_log((uint32_t)v, lg), hi = F.mul_log((uint32_t)v << 8, lg);
      ll[v] = lo & 0xff, lh[v] = lo >> 8, hl[v] = hi & 0xff, hh[v] = hi >> 8;
    }
    // the kernel's planes: low bytes first, so x' = A x ^ B c is low = ll(low) ^ hl(high)
    pack_table(ll, dst);
    pack_table(hl, dst + kTabWords);
    pack_table(lh, dst + 2 * kTabWords);
    pack_table(hh, dst + 3 * kTabWords);
  }
}

// ------------------------------------------------------------ transform plans
// A butterfly block: lanes r .. r + radix * cnt, with the twiddle logs of its
// three multiplies (radix 2: one twiddle, stored three times).  Host only: the
// device reads the tables of each block (BlockTables) and computes r.
struct LeoBlock {
  uint32_t r, cnt, l01, l23, l02;
};
// One layer.  Its blocks are block0 .. block0 + n_blocks of the plan, block b
// at lane b << span_shift; cnt = 1 << lg_cnt butterflies per block.
struct LeoLayer {
  uint32_t dist, n_blocks, block0, cnt, radix, inverse, lg_cnt, span_shift;
};
// Tables of one block, in the order the butterflies use them (t01, t23, t02),
// each the five dwords q = 0..4 of the byte maps: GF(2^16) as five 16-byte
// vectors (map0[q], map1[q], map2[q], map3[q]) -- maps low->low, low->high,
// high->low, high->high -- and GF(2^8) as the five dwords of its one map.
constexpr uint32_t kBlockTabWords16 = 60;
constexpr uint32_t kBlockTabWords8 = 16;  // 15 used
struct LeoGroup {
  uint32_t first, mtrunc, layer0, n_layers;
};

static uint32_t log2_exact(uint32_t v) {
  uint32_t l = 0;
  while ((1u << l) < v) ++l;
  return l;
}

struct PlanBuilder {
  const LeoField& F;
  std::vector<LeoLayer> layers;
  std::vector<LeoBlock> blocks;
  explicit PlanBuilder(const LeoField& f) : F(f) {}
  void begin_layer(uint32_t dist, uint32_t cnt, uint32_t radix, bool inverse) {
    LeoLayer L{};
    L.dist = dist;
    L.cnt = cnt;
    L.radix = radix;
    L.inverse = inverse ? 1 : 0;
    L.lg_cnt = log2_exact(cnt);
    L.span_shift = log2_exact(radix * dist);
    L.block0 = (uint32_t)blocks.size();
    layers.push_back(L);
  }
  void block(uint32_t r, uint32_t cnt, uint32_t l01, uint32_t l23, uint32_t l02) {
    blocks.push_back(LeoBlock{r, cnt, l01, l23, l02});
    layers.back().n_blocks++;
  }
  // IFFT, encoder form (skew offset base; build_ifft_dit8_plan mod.rs:376-447) or
  // decoder form (full skew with i_end - 1; build_ifft_decode_dit8_plan :455-526).
  uint32_t ifft(uint32_t mtrunc, uint32_t m, const uint32_t* skew, bool decoder) {
    const uint32_t first = (uint32_t)layers.size();
    const uint32_t o = decoder ? 1 : 0;
    uint32_t dist = 1, dist4 = 4;
    while (dist4 <= m) {
      begin_layer(dist, dist, 4, true);
      for (uint32_t r = 0; r < mtrunc; r += dist4) {
        uint32_t iend = r + dist;
        block(r, dist, skew[iend - o], skew[iend + 2 * dist - o], skew[iend + dist - o]);
      }
      dist = dist4;
      dist4 <<= 2;
    }
    if (dist < m) {
      begin_layer(dist, dist, 2, true);
      uint32_t l = skew[dist - o];
      block(0, dist, l, l, l);
    }
    return (uint32_t)layers.size() - first;
  }
  // FFT (build_fft_dit8_plan mod.rs:334-374): full skew, i_end - 1 indices.
  uint32_t fft(uint32_t mtrunc, uint32_t m, const uint32_t* skew) {
    const uint32_t first = (uint32_t)layers.size();
    uint32_t dist4 = m, dist = m >> 2;
    while (dist != 0) {
      begin_layer(dist, dist, 4, false);
      for (uint32_t r = 0; r < mtrunc; r += dist4) {
        uint32_t iend = r + dist;
        block(r, dist, skew[iend - 1], skew[iend + 2 * dist - 1], skew[iend + dist - 1]);
      }
      dist4 = dist;
      dist >>= 2;
    }
    if (dist4 == 2) {
      begin_layer(1, 1, 2, false);
      for (uint32_t r = 0; r < mtrunc; r += 2) block(r, 1, skew[r], skew[r], skew[r]);
    }
    return (uint32_t)layers.size() - first;
  }
};

static uint32_t ceil_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

struct LeoDevPlan {
  DeviceBuffer layers, btab;  // LeoLayer[n_layers], block tables [n_blocks][kBlockTabWords*]
  uint32_t n_layers = 0, n_btab_words = 0;
};

// Uploads the layers and, per block, its three twiddle tables in the layout of
// BlockTables.  Every block must sit at lane b << span_shift of its layer, which
// is how the schedules lay them out (r steps by radix * dist from 0).
static int upload_plan(const PlanBuilder& b, LeoDevPlan& p) {
  const LeoField& F = b.F;
  const uint32_t BW = F.gf16 ? kBlockTabWords16 : kBlockTabWords8;
  for (const LeoLayer& L : b.layers)
    for (uint32_t j = 0; j < L.n_blocks; ++j)
      if (b.blocks[L.block0 + j].r != (j << L.span_shift) || (1u << L.lg_cnt) != L.cnt)
        return RSGPU_ERR_UNSUPPORTED;
  std::vector<uint32_t> btab((size_t)b.blocks.size() * BW, 0);
  std::map<uint32_t, std::vector<uint32_t>> made;  // log -> leo_table words
  auto table = [&](uint32_t log_m) -> const std::vector<uint32_t>& {
    auto it = made.find(log_m);
    if (it != made.end()) return it->second;
    std::vector<uint32_t> w(F.gf16 ? kLeoTabWords16 : kLeoTabWords8);
    leo_table(F, log_m, true, w.data());
    return made.emplace(log_m, std::move(w)).first->second;
  };
  for (size_t j = 0; j < b.blocks.size(); ++j) {
    const LeoBlock& B = b.blocks[j];
    const uint32_t logs[3] = {B.l01, B.l23, B.l02};
    uint32_t* dst = btab.data() + j * BW;
    for (int t = 0; t < 3; ++t) {
      const std::vector<uint32_t>& w = table(logs[t]);
      if (F.gf16) {
        for (int q = 0; q < 5; ++q)
          for (int m = 0; m < 4; ++m) dst[t * 20 + q * 4 + m] = w[m * kTabWords + q];
      } else {
        for (int q = 0; q < 5; ++q) dst[t * 5 + q] = w[q];
      }
    }
  }
  if (p.layers.alloc(std::max<size_t>(1, b.layers.size()) * sizeof(LeoLayer)) != hipSuccess ||
      p.btab.alloc(std::max<size_t>(1, btab.size()) * 4) != hipSuccess)
    return RSGPU_ERR_OUT_OF_MEMORY;
  if (!b.layers.empty() && hipMemcpy(p.layers.p, b.layers.data(), b.layers.size() * sizeof(LeoLayer),
                                     hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (!btab.empty() && hipMemcpy(p.btab.p, btab.data(), btab.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  p.n_layers = (uint32_t)b.layers.size();
  p.n_btab_words = (uint32_t)btab.size();
  return RSGPU_OK;
}

// Per-pattern decode state: transform plan, per-lane multiply tables (error
// locator and its inverse) built on the device.
// The locator tables are written by a kernel on `stream`; `ready` is recorded
// behind it, and a call on another stream waits for it while it is pending.
struct LeoDecodeEntry {
  LeoDevPlan plan;
  uint32_t n_ifft = 0, n_fft = 0, input_count = 0;
  DeviceBuffer err_tabs;  // [n][words] multiply-by-err_loc tables (lanes < work_size)
  DeviceBuffer inv_tabs;  // [n][words] multiply-by-(modulus - err_loc)
  DeviceBuffer err_locs;  // [work_size] u32, kept for inspection
  DeviceBuffer pos;       // erased positions, read by the locator kernel
  hipEvent_t ready = nullptr;
  hipStream_t stream = nullptr;
  ~LeoDecodeEntry() {
    if (ready) {
      (void)hipEventSynchronize(ready);
      (void)hipEventDestroy(ready);
    }
  }
};

// A decode plan being built by one thread; others asking for the same pattern
// wait for it instead of building it again.
struct LeoPending {
  std::mutex m;
  std::condition_variable cv;
  bool done = false;
  int rc = RSGPU_OK;
  std::shared_ptr<LeoDecodeEntry> e;
};

struct LeoState {
  bool gf16 = false;
  uint32_t k = 0, parity = 0, m = 0, n = 0, work_size = 0;
  // encode
  LeoDevPlan enc;
  std::vector<LeoGroup> groups;
  uint32_t fft_layer0 = 0, fft_layers = 0;
  uint32_t fused_layer0 = 0, fused_layers = 0;  // every group's IFFT in one layer list
  DeviceBuffer d_groups;
  // device copies of the field log/exp tables (for the error locator and tables)
  DeviceBuffer d_log, d_exp;
  size_t lds_limit = 65536;
  // decode plans by presence pattern: the 64 most recently used, plus the ones
  // in flight (guarded by mu; built outside it)
  std::mutex mu;
  LruMap<LeoDecodeEntry> decode_cache{64};
  std::map<std::string, std::shared_ptr<LeoPending>> in_flight;
};

// ---------------------------------------------------------------- kernels
constexpr uint32_t kLeoInline = 256;

struct LeoArgs {
  const LeoLayer* layers;
  const uint32_t* btab;        // block tables (BlockTables layout)
  const uint8_t* const* ptrs;  // k data then parity shard pointers
  const LeoGroup* groups;
  const uint32_t* err_tabs;
  const uint32_t* inv_tabs;
  const uint8_t* present;      // decode: per shard (device copy), required in bit 1
  uint64_t words;              // words per shard (GF8: len/4, GF16: len/8)
  uint32_t C;                  // words per tile
  uint32_t rows;               // LDS lanes
  uint32_t k, parity, m, n, work_size;
  uint32_t n_groups, fft_layer0, fft_layers, ifft_layer0, ifft_layers;
  int32_t* flag;               // verify
  uint32_t mode;               // 0 encode, 1 verify
  uint32_t vec;                // every shard pointer 16-byte aligned
  // A small plan is copied into LDS at dword plan_off when the kernel starts
  // (plan_lds = 1): its layers then read blocks and tables from LDS.
  uint32_t plan_lds, plan_off, plan_layers, plan_btab_words;
  // Encode with every data group resident in LDS (group g on lanes g*m ..):
  // one load phase and the fused IFFT layers instead of one pass per group.
  uint32_t resident, fused_layer0, fused_layers;
  // Up to kLeoInline shards: pointers and flags travel in the kernel arguments
  // (inline_args = 1) instead of through the staged copy above, which saves a
  // host-to-device copy per call.
  uint32_t inline_args;
  const uint8_t* iptr[kLeoInline];
  uint8_t ipresent[kLeoInline];
};

__device__ __forceinline__ const uint8_t* shard_ptr(const LeoArgs& a, uint32_t i) {
  return a.inline_args ? a.iptr[i] : a.ptrs[i];
}
__device__ __forceinline__ uint8_t shard_flags(const LeoArgs& a, uint32_t i) {
  return a.inline_args ? a.ipresent[i] : a.present[i];
}

template <bool G16>
struct Word;
template <>
struct Word<false> {
  uint32_t v;
};
template <>
struct Word<true> {
  uint32_t lo, hi;
};

// G consecutive words of one LDS row (G = 1, 2 or 4; the row stride C is a
// multiple of G), at word index idx of a buffer.  G = 4 moves them as 16-byte
// LDS accesses.  A GF16 group of four words is 32 bytes, so its two halves are
// stored apart: groups are laid out in blocks of 16 (512 bytes), the words 0-1
// of the 16 groups first, then their words 2-3.  Consecutive lanes (consecutive
// groups) then read consecutive 16-byte slots in both accesses: conflict free.
// Buffers in this layout start on a block and span whole blocks
// (gf16_lds_dwords).
__host__ __device__ constexpr uint32_t gf16_lds_dwords(uint32_t words) {  // whole 16-group blocks
  return (words / 4 + 15) / 16 * 128;
}
__device__ __forceinline__ uint32_t gf16_half0(uint32_t idx) {
  const uint32_t gi = idx >> 2;
  return ((gi & ~15u) << 3) + ((gi & 15u) << 2);
}
template <bool G16, int G>
__device__ __forceinline__ void lds_getv(const uint32_t* lds, uint32_t idx, Word<G16> (&w)[G]) {
  if constexpr (G == 4 && !G16) {
    uint4 t = *reinterpret_cast<const uint4*>(lds + idx);
    w[0].v = t.x, w[1].v = t.y, w[2].v = t.z, w[3].v = t.w;
  } else if constexpr (G == 4 && G16) {
    const uint32_t d0 = gf16_half0(idx);
    uint4 t0 = *reinterpret_cast<const uint4*>(lds + d0);
    uint4 t1 = *reinterpret_cast<const uint4*>(lds + d0 + 64);
    w[0].lo = t0.x, w[0].hi = t0.y, w[1].lo = t0.z, w[1].hi = t0.w;
    w[2].lo = t1.x, w[2].hi = t1.y, w[3].lo = t1.z, w[3].hi = t1.w;
  } else if constexpr (G == 2 && !G16) {
    uint2 t = *reinterpret_cast<const uint2*>(lds + idx);
    w[0].v = t.x, w[1].v = t.y;
  } else if constexpr (G == 2 && G16) {
    uint4 t = *reinterpret_cast<const uint4*>(lds + 2 * idx);
    w[0].lo = t.x, w[0].hi = t.y, w[1].lo = t.z, w[1].hi = t.w;
  } else {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if constexpr (G16) {
        uint2 t = reinterpret_cast<const uint2*>(lds)[idx + g];
        w[g].lo = t.x, w[g].hi = t.y;
      } else {
        w[g].v = lds[idx + g];
      }
    }
  }
}
template <bool G16, int G>
__device__ __forceinline__ void lds_putv(uint32_t* lds, uint32_t idx, const Word<G16> (&w)[G]) {
  if constexpr (G == 4 && !G16) {
    *reinterpret_cast<uint4*>(lds + idx) = make_uint4(w[0].v, w[1].v, w[2].v, w[3].v);
  } else if constexpr (G == 4 && G16) {
    const uint32_t d0 = gf16_half0(idx);
    *reinterpret_cast<uint4*>(lds + d0) = make_uint4(w[0].lo, w[0].hi, w[1].lo, w[1].hi);
    *reinterpret_cast<uint4*>(lds + d0 + 64) = make_uint4(w[2].lo, w[2].hi, w[3].lo, w[3].hi);
  } else if constexpr (G == 2 && !G16) {
    *reinterpret_cast<uint2*>(lds + idx) = make_uint2(w[0].v, w[1].v);
  } else if constexpr (G == 2 && G16) {
    *reinterpret_cast<uint4*>(lds + 2 * idx) = make_uint4(w[0].lo, w[0].hi, w[1].lo, w[1].hi);
  } else {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if constexpr (G16) reinterpret_cast<uint2*>(lds)[idx + g] = make_uint2(w[g].lo, w[g].hi);
      else lds[idx + g] = w[g].v;
    }
  }
}

// A twiddle's multiply tables held in registers for the words of one item:
// GF8 one 5-dword table, GF16 the four byte maps (low->low, low->high,
// high->low, high->high) at dword offsets 0, 8, 16, 24.
template <bool G16>
struct Tab;
template <>
struct Tab<false> {
  uint32_t a[5];
};
template <>
struct Tab<true> {
  uint32_t a[4][5];
};

__device__ __forceinline__ void load5(const uint32_t* p, uint32_t (&a)[5]) {
  uint4 q = *reinterpret_cast<const uint4*>(p);  // tables are 32-byte aligned
  a[0] = q.x, a[1] = q.y, a[2] = q.z, a[3] = q.w, a[4] = p[4];
}
template <bool G16>
__device__ __forceinline__ Tab<G16> load_tab(const uint32_t* p) {
  Tab<G16> t;
  if constexpr (G16) {
#pragma unroll
    for (int j = 0; j < 4; ++j) load5(p + j * kTabWords, t.a[j]);
  } else {
    load5(p, t.a);
  }
  return t;
}

__device__ __forceinline__ uint32_t tmul5(const uint32_t (&t)[5], const Sel& s) {
  uint32_t a = __builtin_amdgcn_perm(t[1], t[0], s.s0);
  uint32_t b = __builtin_amdgcn_perm(t[3], t[2], s.s1);
  uint32_t c = __builtin_amdgcn_perm(t[4], t[4], s.s2);
  return xor3(a, b, c);
}

// y * g^log for the twiddle's tables
template <bool G16>
__device__ __forceinline__ Word<G16> wmul(const Tab<G16>& t, Word<G16> y) {
  Word<G16> r;
  if constexpr (G16) {
    Sel sl = selectors(y.lo), sh = selectors(y.hi);
    r.lo = tmul5(t.a[0], sl) ^ tmul5(t.a[2], sh);
    r.hi = tmul5(t.a[1], sl) ^ tmul5(t.a[3], sh);
  } else {
    r.v = tmul5(t.a, selectors(y.v));
  }
  return r;
}
template <bool G16>
__device__ __forceinline__ void wxor(Word<G16>& a, Word<G16> b) {
  if constexpr (G16) {
    a.lo ^= b.lo;
    a.hi ^= b.hi;
  } else {
    a.v ^= b.v;
  }
}
// One half of a GF(2^16) product folded into acc: acc ^ (the six byte lookups
// of the low map tl on sl and the high map th on sh), three xor3s in all.
__device__ __forceinline__ uint32_t fold6(uint32_t acc, const uint32_t (&tl)[5], const Sel& sl,
                                          const uint32_t (&th)[5], const Sel& sh) {
  acc = xor3(acc, __builtin_amdgcn_perm(tl[1], tl[0], sl.s0), __builtin_amdgcn_perm(tl[3], tl[2], sl.s1));
  acc = xor3(acc, __builtin_amdgcn_perm(tl[4], tl[4], sl.s2), __builtin_amdgcn_perm(th[1], th[0], sh.s0));
  return xor3(acc, __builtin_amdgcn_perm(th[3], th[2], sh.s1), __builtin_amdgcn_perm(th[4], th[4], sh.s2));
}
// x ^= y * g^log for the twiddle's tables
template <bool G16>
__device__ __forceinline__ void wmul_acc(const Tab<G16>& t, Word<G16> y, Word<G16>& x) {
  if constexpr (G16) {
    const Sel sl = selectors(y.lo), sh = selectors(y.hi);
    x.lo = fold6(x.lo, t.a[0], sl, t.a[2], sh);
    x.hi = fold6(x.hi, t.a[1], sl, t.a[3], sh);
  } else {
    wxor(x, wmul<G16>(t, y));
  }
}
// fft_dit2: x ^= y * t; y ^= x      ifft_dit2: y ^= x; x ^= y * t
template <bool G16>
__device__ __forceinline__ void fwd2(Word<G16>& x, Word<G16>& y, const Tab<G16>& t) {
  wmul_acc<G16>(t, y, x);
  wxor(y, x);
}
template <bool G16>
__device__ __forceinline__ void inv2(Word<G16>& x, Word<G16>& y, const Tab<G16>& t) {
  wxor(y, x);
  wmul_acc<G16>(t, y, x);
}

typedef const __attribute__((address_space(4))) uint32_t* cword_t;

// The transform plan as the layers read it: the arrays in global memory, or
// their copy in LDS.
struct PlanPtrs {
  const LeoLayer* layers;
  const uint32_t* btab;
};

// Copies the plan (layers, block tables: whole 16-byte vectors each) into LDS at
// dst.  The first barrier of the kernel body orders the copy before any use.
__device__ __forceinline__ PlanPtrs stage_plan(const LeoArgs& a, uint32_t* dst) {
  const uint32_t nl = a.plan_layers * 8, nt = a.plan_btab_words;
  auto copy = [&](uint32_t* to, const uint32_t* from, uint32_t dwords) {
    for (uint32_t i = threadIdx.x; i < dwords / 4; i += blockDim.x)
      reinterpret_cast<uint4*>(to)[i] = reinterpret_cast<const uint4*>(from)[i];
  };
  copy(dst, reinterpret_cast<const uint32_t*>(a.layers), nl);
  copy(dst + nl, a.btab, nt);
  return PlanPtrs{reinterpret_cast<const LeoLayer*>(dst), dst + nl};
}

// One twiddle's tables from a block's tables (BlockTables layout): scalar loads
// when the whole wave shares the block (S), else vector loads (GF(2^16): five
// 16-byte loads).
template <bool G16, bool S>
__device__ __forceinline__ Tab<G16> block_tab(const uint32_t* p) {
  Tab<G16> t;
  if constexpr (S) {
    const cword_t c = (cword_t)p;
    if constexpr (G16) {
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) t.a[j][q] = c[q * 4 + j];
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) t.a[q] = c[q];
    }
  } else {
    if constexpr (G16) {
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const uint4 v = reinterpret_cast<const uint4*>(p)[q];
        t.a[0][q] = v.x, t.a[1][q] = v.y, t.a[2][q] = v.z, t.a[3][q] = v.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 5; ++q) t.a[q] = p[q];
    }
  }
  return t;
}

// The items of one layer: item = (butterfly, group of G words), the word group
// fastest, so consecutive lanes share a butterfly.  Counts are powers of two
// (cnt = 1 << lg_cnt butterflies per block, ng = 1 << lg_ng word groups), so the
// item decomposes with shifts, and block b sits at lane b << span_shift.  UNI:
// every wave's 64 items lie in one block (cnt * ng a multiple of 64), so its
// tables are read once per wave -- scalar loads from global memory, or
// broadcast LDS reads when the plan is in LDS (PL).  Each stage of the
// butterfly loads its one table and applies it to all G words, so only one
// table is live at a time.
template <bool G16, int G, bool UNI, bool PL>
__device__ __forceinline__ void layer_items(uint32_t* lds, const LeoLayer& L, uint32_t lg_ng,
                                            const LeoArgs& a, const PlanPtrs& P) {
  constexpr uint32_t BW = G16 ? kBlockTabWords16 : kBlockTabWords8;
  constexpr uint32_t TS = G16 ? 20 : 5;  // dwords per table
  constexpr bool S = UNI && !PL;
  const uint32_t C = a.C;
  const uint32_t items = L.n_blocks << (L.lg_cnt + lg_ng);
  const uint32_t ng_mask = (1u << lg_ng) - 1, cnt_mask = L.cnt - 1;
  for (uint32_t it = threadIdx.x; it < items; it += blockDim.x) {
    const uint32_t bf = it >> lg_ng, wq = (it & ng_mask) * G;
    uint32_t blk = bf >> L.lg_cnt;
    if constexpr (UNI) blk = __builtin_amdgcn_readfirstlane(blk);
    const uint32_t i = (blk << L.span_shift) + (bf & cnt_mask);
    const uint32_t d = L.dist;
    const uint32_t* tb = P.btab + (size_t)(L.block0 + blk) * BW;
    if (L.radix == 2) {
      Word<G16> x[G], y[G];
      lds_getv<G16, G>(lds, i * C + wq, x);
      lds_getv<G16, G>(lds, (i + d) * C + wq, y);
      const Tab<G16> t = block_tab<G16, S>(tb);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (L.inverse) inv2<G16>(x[g], y[g], t);
        else fwd2<G16>(x[g], y[g], t);
      }
      lds_putv<G16, G>(lds, i * C + wq, x);
      lds_putv<G16, G>(lds, (i + d) * C + wq, y);
    } else {
      Word<G16> x0[G], x1[G], x2[G], x3[G];
      lds_getv<G16, G>(lds, i * C + wq, x0);
      lds_getv<G16, G>(lds, (i + d) * C + wq, x1);
      lds_getv<G16, G>(lds, (i + 2 * d) * C + wq, x2);
      lds_getv<G16, G>(lds, (i + 3 * d) * C + wq, x3);
      if (L.inverse) {  // ifft_dit4 (ops.rs:556-586)
        {
          const Tab<G16> t01 = block_tab<G16, S>(tb);
#pragma unroll
          for (int g = 0; g < G; ++g) inv2<G16>(x0[g], x1[g], t01);
        }
        {
          const Tab<G16> t23 = block_tab<G16, S>(tb + TS);
#pragma unroll
          for (int g = 0; g < G; ++g) inv2<G16>(x2[g], x3[g], t23);
        }
        const Tab<G16> t02 = block_tab<G16, S>(tb + 2 * TS);
#pragma unroll
        for (int g = 0; g < G; ++g) {
          inv2<G16>(x1[g], x3[g], t02);
          inv2<G16>(x0[g], x2[g], t02);
        }
      } else {  // fft_dit4 (ops.rs:516-548)
        {
          const Tab<G16> t02 = block_tab<G16, S>(tb + 2 * TS);
#pragma unroll
          for (int g = 0; g < G; ++g) {
            fwd2<G16>(x0[g], x2[g], t02);
            fwd2<G16>(x1[g], x3[g], t02);
          }
        }
        {
          const Tab<G16> t01 = block_tab<G16, S>(tb);
#pragma unroll
          for (int g = 0; g < G; ++g) fwd2<G16>(x0[g], x1[g], t01);
        }
        const Tab<G16> t23 = block_tab<G16, S>(tb + TS);
#pragma unroll
        for (int g = 0; g < G; ++g) fwd2<G16>(x2[g], x3[g], t23);
      }
      lds_putv<G16, G>(lds, i * C + wq, x0);
      lds_putv<G16, G>(lds, (i + d) * C + wq, x1);
      lds_putv<G16, G>(lds, (i + 2 * d) * C + wq, x2);
      lds_putv<G16, G>(lds, (i + 3 * d) * C + wq, x3);
    }
  }
}

// log2 of the word groups per lane row: C and G are powers of two.
template <int G>
__device__ __forceinline__ uint32_t lg_groups(const LeoArgs& a) {
  return (uint32_t)__builtin_ctz(a.C) - (G == 4 ? 2u : G == 2 ? 1u : 0u);
}

// Runs layers [l0, l0 + nl) over the LDS lanes.  An item is one butterfly on G
// consecutive words, so its twiddle tables are loaded once for G words.
template <bool G16, int G, bool PL>
__device__ __forceinline__ void run_layers(uint32_t* lds, uint32_t l0, uint32_t nl, const LeoArgs& a,
                                           const PlanPtrs& P) {
  const uint32_t lg_ng = lg_groups<G>(a);
  for (uint32_t l = l0; l < l0 + nl; ++l) {
    const LeoLayer L = P.layers[l];
    if (L.lg_cnt + lg_ng >= 6) layer_items<G16, G, true, PL>(lds, L, lg_ng, a, P);
    else layer_items<G16, G, false, PL>(lds, L, lg_ng, a, P);
    __syncthreads();
  }
}

// Global access to words [word, word + G) of a shard; words past `words` read as
// zero and are not written.  G = 4 with `vec` (16-byte aligned shards, word a
// multiple of 4) uses 16-byte loads/stores: GF8 words are 4 bytes, GF16 words are
// the (low-byte, high-byte) dword pairs of the split layout (ops.rs:520-568:
// element i of a 64-byte block = byte[i] | byte[i + 32] << 8), and 4 aligned
// words sit in one block's low and high halves.
template <bool G16>
__device__ __forceinline__ Word<G16> gload1(const uint8_t* shard, uint64_t word) {
  Word<G16> r;
  if constexpr (G16) {
    const uint64_t b = word >> 3, j = word & 7;
    r.lo = *reinterpret_cast<const uint32_t*>(shard + b * 64 + j * 4);
    r.hi = *reinterpret_cast<const uint32_t*>(shard + b * 64 + 32 + j * 4);
  } else {
    r.v = *reinterpret_cast<const uint32_t*>(shard + word * 4);
  }
  return r;
}
template <bool G16>
__device__ __forceinline__ void gstore1(uint8_t* shard, uint64_t word, Word<G16> v) {
  if constexpr (G16) {
    const uint64_t b = word >> 3, j = word & 7;
    *reinterpret_cast<uint32_t*>(shard + b * 64 + j * 4) = v.lo;
    *reinterpret_cast<uint32_t*>(shard + b * 64 + 32 + j * 4) = v.hi;
  } else {
    *reinterpret_cast<uint32_t*>(shard + word * 4) = v.v;
  }
}
template <bool G16, int G>
__device__ __forceinline__ void gloadv(const uint8_t* shard, uint64_t word, uint64_t words, bool vec,
                                       Word<G16> (&w)[G]) {
  if constexpr (G == 4) {
    if (vec && word + 4 <= words) {
      if constexpr (G16) {
        const uint8_t* p = shard + (word >> 3) * 64 + (word & 7) * 4;
        uint4 lo = load16(p), hi = load16(p + 32);
        w[0] = {lo.x, hi.x}, w[1] = {lo.y, hi.y}, w[2] = {lo.z, hi.z}, w[3] = {lo.w, hi.w};
      } else {
        uint4 q = load16(shard + word * 4);
        w[0].v = q.x, w[1].v = q.y, w[2].v = q.z, w[3].v = q.w;
      }
      return;
    }
  }
  w[0] = word < words ? gload1<G16>(shard, word) : Word<G16>{};
  if constexpr (G > 1) w[1] = word + 1 < words ? gload1<G16>(shard, word + 1) : Word<G16>{};
  if constexpr (G > 2) {
    w[2] = word + 2 < words ? gload1<G16>(shard, word + 2) : Word<G16>{};
    w[3] = word + 3 < words ? gload1<G16>(shard, word + 3) : Word<G16>{};
  }
}
template <bool G16, int G>
__device__ __forceinline__ void gstorev(uint8_t* shard, uint64_t word, uint64_t words, bool vec,
                                        const Word<G16> (&w)[G]) {
  if constexpr (G == 4) {
    if (vec && word + 4 <= words) {
      if constexpr (G16) {
        uint8_t* p = shard + (word >> 3) * 64 + (word & 7) * 4;
        store16(p, make_uint4(w[0].lo, w[1].lo, w[2].lo, w[3].lo));
        store16(p + 32, make_uint4(w[0].hi, w[1].hi, w[2].hi, w[3].hi));
      } else {
        store16(shard + word * 4, make_uint4(w[0].v, w[1].v, w[2].v, w[3].v));
      }
      return;
    }
  }
  // partial last vector or unaligned shards: word by word (no indexed loop, so
  // the words stay in registers)
  if (word < words) gstore1<G16>(shard, word, w[0]);
  if constexpr (G > 1) {
    if (word + 1 < words) gstore1<G16>(shard, word + 1, w[1]);
  }
  if constexpr (G > 2) {
    if (word + 2 < words) gstore1<G16>(shard, word + 2, w[2]);
    if (word + 3 < words) gstore1<G16>(shard, word + 3, w[3]);
  }
}
template <bool G16>
__device__ __forceinline__ bool wequal(Word<G16> a, Word<G16> b) {
  if constexpr (G16) return a.lo == b.lo && a.hi == b.hi;
  else return a.v == b.v;
}

// encode_with_tables (leopard_gf8/encode.rs:194-372; leopard_gf16/encode.rs:109-253):
// groups of m data lanes are IFFT'd (skew offset m-1+group*m) and XORed into the
// first m lanes, which are then FFT'd; lanes 0..parity are the parity.
template <bool G16, int G, bool PL>
__device__ __forceinline__ void leo_encode_resident(const LeoArgs& a, uint32_t* lds, const PlanPtrs& P) {
  const uint32_t C = a.C, m = a.m, ng = C / G, lg_ng = lg_groups<G>(a);
  const uint64_t w0 = (uint64_t)blockIdx.x * C;
  // group g holds data shards g*m .. g*m + min(m, k - g*m) - 1; the rest are zero
  const uint32_t lanes = a.n_groups * m;
  for (uint32_t it = threadIdx.x; it < lanes * ng; it += blockDim.x) {
    const uint32_t lane = it >> lg_ng, wq = (it & (ng - 1)) * G;
    Word<G16> v[G];
    if (lane < a.k) {
      gloadv<G16, G>(shard_ptr(a, lane), w0 + wq, a.words, a.vec != 0, v);
    } else {
#pragma unroll
      for (int q = 0; q < G; ++q) v[q] = Word<G16>{};
    }
    lds_putv<G16, G>(lds, lane * C + wq, v);
  }
  __syncthreads();
  run_layers<G16, G, PL>(lds, a.fused_layer0, a.fused_layers, a, P);
  for (uint32_t it = threadIdx.x; it < m * ng; it += blockDim.x) {
    Word<G16> x[G], y[G];
    lds_getv<G16, G>(lds, it * G, x);
    for (uint32_t g = 1; g < a.n_groups; ++g) {
      lds_getv<G16, G>(lds, g * m * C + it * G, y);
#pragma unroll
      for (int q = 0; q < G; ++q) wxor(x[q], y[q]);
    }
    lds_putv<G16, G>(lds, it * G, x);
  }
  __syncthreads();
}

template <bool G16, int G, bool PL>
__device__ __forceinline__ void leo_encode_body(const LeoArgs& a, uint32_t* lds, const PlanPtrs& P) {
  const uint32_t C = a.C, m = a.m, ng = C / G, lg_ng = lg_groups<G>(a);
  const uint64_t w0 = (uint64_t)blockIdx.x * C;
  uint32_t* A = lds;
  uint32_t* Bv = lds + (G16 && G == 4 ? gf16_lds_dwords(m * C) : m * C * (G16 ? 2u : 1u));
  if (a.resident) leo_encode_resident<G16, G, PL>(a, lds, P);
  for (uint32_t g = 0; g < (a.resident ? 0u : a.n_groups); ++g) {
    const LeoGroup Gp = a.groups[g];
    uint32_t* dst = g == 0 ? A : Bv;
    for (uint32_t it = threadIdx.x; it < m * ng; it += blockDim.x) {
      const uint32_t lane = it >> lg_ng, wq = (it & (ng - 1)) * G;
      Word<G16> v[G];
      if (lane < Gp.mtrunc) {
        gloadv<G16, G>(shard_ptr(a, Gp.first + lane), w0 + wq, a.words, a.vec != 0, v);
      } else {
#pragma unroll
        for (int q = 0; q < G; ++q) v[q] = Word<G16>{};
      }
      lds_putv<G16, G>(dst, lane * C + wq, v);
    }
    __syncthreads();
    run_layers<G16, G, PL>(dst, Gp.layer0, Gp.n_layers, a, P);
    if (g > 0) {
      for (uint32_t it = threadIdx.x; it < m * ng; it += blockDim.x) {
        Word<G16> x[G], y[G];
        lds_getv<G16, G>(A, it * G, x);
        lds_getv<G16, G>(Bv, it * G, y);
#pragma unroll
        for (int q = 0; q < G; ++q) wxor(x[q], y[q]);
        lds_putv<G16, G>(A, it * G, x);
      }
      __syncthreads();
    }
  }
  run_layers<G16, G, PL>(A, a.fft_layer0, a.fft_layers, a, P);
  bool mismatch = false;
  for (uint32_t it = threadIdx.x; it < a.parity * ng; it += blockDim.x) {
    const uint32_t lane = it >> lg_ng, wq = (it & (ng - 1)) * G;
    if (w0 + wq >= a.words) continue;
    uint8_t* out = const_cast<uint8_t*>(shard_ptr(a, a.k + lane));
    Word<G16> v[G];
    lds_getv<G16, G>(A, lane * C + wq, v);
    if (a.mode == 0) {
      gstorev<G16, G>(out, w0 + wq, a.words, a.vec != 0, v);
    } else {
      Word<G16> s[G];
      gloadv<G16, G>(out, w0 + wq, a.words, a.vec != 0, s);
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (w0 + wq + q < a.words) mismatch |= !wequal<G16>(v[q], s[q]);
    }
  }
  if (wave_leader(mismatch)) raise_flag(a.flag);
}

template <bool G16, int NT, int MINW = NT / 256>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 2 * NT / 256))) void leo_encode_kernel(const LeoArgs a) {
  extern __shared__ uint32_t lds[];
  if (a.plan_lds) {
    const PlanPtrs P = stage_plan(a, lds + a.plan_off);
    if (a.C % 4 == 0) leo_encode_body<G16, 4, true>(a, lds, P);
    else if (a.C % 2 == 0) leo_encode_body<G16, 2, true>(a, lds, P);
    else leo_encode_body<G16, 1, true>(a, lds, P);
  } else {
    const PlanPtrs P{a.layers, a.btab};
    if (a.C % 4 == 0) leo_encode_body<G16, 4, false>(a, lds, P);
    else if (a.C % 2 == 0) leo_encode_body<G16, 2, false>(a, lds, P);
    else leo_encode_body<G16, 1, false>(a, lds, P);
  }
}

// The decode's input rows (compute_reconstruct_chunk step 1, leopard_gf16/decode.rs:858-893):
// lanes 0..parity are the parity shards, lanes m..m+k the data shards; a
// present shard's words, scaled by its lane's error-locator table, else zero.
template <bool G16, int G>
__device__ __forceinline__ void decode_row(const LeoArgs& a, uint32_t lane, uint64_t word, Word<G16> (&v)[G]) {
  constexpr int TW = G16 ? kLeoTabWords16 : kLeoTabWords8;
  int shard = -1;
  if (lane < a.parity) shard = (int)(a.k + lane);
  else if (lane >= a.m && lane < a.m + a.k) shard = (int)(lane - a.m);
  if (shard >= 0 && (shard_flags(a, shard) & 1)) {
    gloadv<G16, G>(shard_ptr(a, shard), word, a.words, a.vec != 0, v);
    const Tab<G16> te = load_tab<G16>(a.err_tabs + (size_t)lane * TW);
#pragma unroll
    for (int q = 0; q < G; ++q) v[q] = wmul<G16>(te, v[q]);
  } else {
#pragma unroll
    for (int q = 0; q < G; ++q) v[q] = Word<G16>{};
  }
}

// Formal derivative (decode.rs:537-551, compute_formal_derivative16
// leopard_gf16/decode.rs:1100-1116) in closed form: row j takes the XOR of the
// rows j + 2^b for the clear bits b of j (j + 2^b < n), all read before any
// update.  With Q = n / 4, an item is the four rows g, g + Q, g + 2Q, g + 3Q
// (g < Q) times G words: the two top bits stay inside the item, and a clear bit
// b of g adds item g + 2^b row for row.  Consecutive lanes take consecutive g,
// so every LDS access is a run of consecutive rows (conflict free).  A row
// depends only on itself and higher rows, so chunks of items run in increasing
// g, each reading before any of its writes.
template <bool G16, int G>
__device__ __forceinline__ void formal_derivative(const LeoArgs& a, uint32_t* lds) {
  const uint32_t C = a.C, ng = C / G, lg_ng = lg_groups<G>(a), Q = a.n >> 2;
  for (uint32_t it0 = 0; it0 < Q * ng; it0 += blockDim.x) {
    const uint32_t it = it0 + threadIdx.x;
    const bool live = it < Q * ng;
    const uint32_t g = it >> lg_ng, wq = (it & (ng - 1)) * G;
    Word<G16> r0[G], r1[G], r2[G], r3[G];
    if (live) {
      lds_getv<G16, G>(lds, g * C + wq, r0);
      lds_getv<G16, G>(lds, (g + Q) * C + wq, r1);
      lds_getv<G16, G>(lds, (g + 2 * Q) * C + wq, r2);
      lds_getv<G16, G>(lds, (g + 3 * Q) * C + wq, r3);
#pragma unroll
      for (int q = 0; q < G; ++q) {
        wxor(r0[q], r1[q]);
        wxor(r0[q], r2[q]);
        wxor(r1[q], r3[q]);
        wxor(r2[q], r3[q]);
      }
      for (uint32_t step = 1; step < Q; step <<= 1) {
        if ((g & step) || g + step >= Q) continue;
        const uint32_t h = g + step;
        Word<G16> y[G];
        lds_getv<G16, G>(lds, h * C + wq, y);
#pragma unroll
        for (int q = 0; q < G; ++q) wxor(r0[q], y[q]);
        lds_getv<G16, G>(lds, (h + Q) * C + wq, y);
#pragma unroll
        for (int q = 0; q < G; ++q) wxor(r1[q], y[q]);
        lds_getv<G16, G>(lds, (h + 2 * Q) * C + wq, y);
#pragma unroll
        for (int q = 0; q < G; ++q) wxor(r2[q], y[q]);
        lds_getv<G16, G>(lds, (h + 3 * Q) * C + wq, y);
#pragma unroll
        for (int q = 0; q < G; ++q) wxor(r3[q], y[q]);
      }
    }
    __syncthreads();
    if (live) {
      lds_putv<G16, G>(lds, g * C + wq, r0);
      lds_putv<G16, G>(lds, (g + Q) * C + wq, r1);
      lds_putv<G16, G>(lds, (g + 2 * Q) * C + wq, r2);
      lds_putv<G16, G>(lds, (g + 3 * Q) * C + wq, r3);
    }
    __syncthreads();
  }
}

// reconstruct_with_resources (leopard_gf8/decode.rs:319-429; leopard_gf16/decode.rs:816-905).
template <bool G16, int G, bool PL>
__device__ __forceinline__ void leo_decode_body(const LeoArgs& a, uint32_t* lds, const PlanPtrs& P) {
  constexpr int TW = G16 ? kLeoTabWords16 : kLeoTabWords8;  // inverse tables below
  const uint32_t C = a.C, n = a.n, m = a.m, ng = C / G, lg_ng = lg_groups<G>(a);
  const uint64_t w0 = (uint64_t)blockIdx.x * C;
  for (uint32_t it = threadIdx.x; it < n * ng; it += blockDim.x) {
    const uint32_t lane = it >> lg_ng, wq = (it & (ng - 1)) * G;
    Word<G16> v[G];
    decode_row<G16, G>(a, lane, w0 + wq, v);
    lds_putv<G16, G>(lds, lane * C + wq, v);
  }
  __syncthreads();
  run_layers<G16, G, PL>(lds, a.ifft_layer0, a.ifft_layers, a, P);
  if (n >= 4) {
    formal_derivative<G16, G>(a, lds);
  } else {
    // formal derivative (decode.rs:537-551): in closed form every lane j takes
    // lane j ^ XOR_{b : bit b of j clear, j + 2^b < n} lane (j + 2^b), all of
    // them before the update.
    for (uint32_t it0 = 0; it0 < n * ng; it0 += blockDim.x) {
      const uint32_t it = it0 + threadIdx.x;
      Word<G16> acc[G];
      const bool live = it < n * ng;
      uint32_t j = 0, wq = 0;
      if (live) {
        j = it >> lg_ng, wq = (it & (ng - 1)) * G;
        lds_getv<G16, G>(lds, j * C + wq, acc);
        for (uint32_t bit = 1; bit < n; bit <<= 1) {
          if ((j & bit) || j + bit >= n) continue;
          Word<G16> y[G];
          lds_getv<G16, G>(lds, (j + bit) * C + wq, y);
#pragma unroll
          for (int q = 0; q < G; ++q) wxor(acc[q], y[q]);
        }
      }
      __syncthreads();
      if (live) lds_putv<G16, G>(lds, j * C + wq, acc);
      __syncthreads();
    }
  }
  run_layers<G16, G, PL>(lds, a.fft_layer0, a.fft_layers, a, P);
  for (uint32_t it = threadIdx.x; it < (a.k + a.parity) * ng; it += blockDim.x) {
    const uint32_t s = it >> lg_ng, wq = (it & (ng - 1)) * G;
    if ((shard_flags(a, s) & 3) != 2 || w0 + wq >= a.words) continue;  // missing and required
    const uint32_t lane = s >= a.k ? s - a.k : m + s;
    const Tab<G16> ti = load_tab<G16>(a.inv_tabs + (size_t)lane * TW);
    Word<G16> v[G];
    lds_getv<G16, G>(lds, lane * C + wq, v);
#pragma unroll
    for (int q = 0; q < G; ++q) v[q] = wmul<G16>(ti, v[q]);
    gstorev<G16, G>(const_cast<uint8_t*>(shard_ptr(a, s)), w0 + wq, a.words, a.vec != 0, v);
  }
}

template <bool G16, int NT, int MINW = NT / 256>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 2 * NT / 256))) void leo_decode_kernel(const LeoArgs a) {
  extern __shared__ uint32_t lds[];
  if (a.plan_lds) {
    const PlanPtrs P = stage_plan(a, lds + a.plan_off);
    if (a.C % 4 == 0) leo_decode_body<G16, 4, true>(a, lds, P);
    else if (a.C % 2 == 0) leo_decode_body<G16, 2, true>(a, lds, P);
    else leo_decode_body<G16, 1, true>(a, lds, P);
  } else {
    const PlanPtrs P{a.layers, a.btab};
    if (a.C % 4 == 0) leo_decode_body<G16, 4, false>(a, lds, P);
    else if (a.C % 2 == 0) leo_decode_body<G16, 2, false>(a, lds, P);
    else leo_decode_body<G16, 1, false>(a, lds, P);
  }
}

// Error locator in closed form: err[i] = sum_{j erased} log[i ^ j] mod (2^n - 1)
// over the erased positions of compute_error_locs (leopard_gf8/decode.rs:68-114:
// missing parity p -> p, padding parity..m, missing data d -> m + d).  This is
// the FWHT convolution of the reference evaluated directly (the Walsh transform
// is its own inverse up to 2^n == 1 mod 2^n - 1, the identity
// leopard_gf16/decode.rs:412-436 relies on).  Both representations of zero
// (0 and 2^n - 1) multiply as the identity, so outputs are unchanged.  Then
// writes the per-lane multiply tables for err and (modulus - err).
struct ErrJob {
  uint32_t work_size, n_pos, gf16, modulus;
};

__device__ __forceinline__ uint32_t d_mul_log(const uint32_t* lg, const uint32_t* ex, uint32_t a,
                                              uint32_t log_b, uint32_t bits, uint32_t order) {
  if (a == 0) return 0;
  uint32_t sum = lg[a] + log_b;
  return ex[(sum + (sum >> bits)) & (order - 1)];
}

__device__ void d_pack(const uint32_t* lg, const uint32_t* ex, uint32_t log_m, uint32_t shift,
                       uint32_t byte_sel, uint32_t bits, uint32_t order, uint32_t* w) {
  uint32_t v[5] = {0, 0, 0, 0, 0};
  auto f = [&](uint32_t x) -> uint32_t {
    return (d_mul_log(lg, ex, x << shift, log_m, bits, order) >> (8 * byte_sel)) & 0xff;
  };
  for (int i = 0; i < 4; ++i) {
    v[0] |= f(i) << (8 * i);
    v[1] |= f(i + 4) << (8 * i);
    v[2] |= f(8 * i) << (8 * i);
    v[3] |= f(8 * (i + 4)) << (8 * i);
    v[4] |= f(64 * i) << (8 * i);
  }
  for (int i = 0; i < 5; ++i) w[i] = v[i];
  w[5] = w[6] = w[7] = 0;
}

__global__ void leo_errloc_kernel(const uint32_t* __restrict__ lg, const uint32_t* __restrict__ ex,
                                  const uint32_t* __restrict__ pos, ErrJob job, uint32_t* err,
                                  uint32_t* err_tabs, uint32_t* inv_tabs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= job.work_size) return;
  uint64_t sum = 0;
  for (uint32_t t = 0; t < job.n_pos; ++t) sum += lg[i ^ pos[t]];
  const uint32_t e = (uint32_t)(sum % job.modulus);
  err[i] = e;
  const uint32_t bits = job.gf16 ? 16 : 8, order = job.modulus + 1;
  const uint32_t inv = job.modulus - e;
  if (job.gf16) {
    uint32_t* te = err_tabs + (size_t)i * kLeoTabWords16;
    uint32_t* ti = inv_tabs + (size_t)i * kLeoTabWords16;
    d_pack(lg, ex, e, 0, 0, bits, order, te);
    d_pack(lg, ex, e, 0, 1, bits, order, te + kTabWords);
    d_pack(lg, ex, e, 8, 0, bits, order, te + 2 * kTabWords);
    d_pack(lg, ex, e, 8, 1, bits, order, te + 3 * kTabWords);
    d_pack(lg, ex, inv, 0, 0, bits, order, ti);
    d_pack(lg, ex, inv, 0, 1, bits, order, ti + kTabWords);
    d_pack(lg, ex, inv, 8, 0, bits, order, ti + 2 * kTabWords);
    d_pack(lg, ex, inv, 8, 1, bits, order, ti + 3 * kTabWords);
  } else {
    d_pack(lg, ex, e, 0, 0, bits, order, err_tabs + (size_t)i * kLeoTabWords8);
    d_pack(lg, ex, inv, 0, 0, bits, order, inv_tabs + (size_t)i * kLeoTabWords8);
  }
}

// ------------------------------------------------------------------- host
std::shared_ptr<const FieldTables> leopard8_field_tables() {
  static std::shared_ptr<const FieldTables> t = [] {
    const LeoField& F = field8();
    auto ft = std::make_shared<FieldTables>();
    std::memset(ft.get(), 0, sizeof(FieldTables));
    for (uint32_t a = 1; a < F.order; ++a) ft->log[a] = (uint8_t)(F.log[a] % F.modulus);
    for (uint32_t i = 0; i < 512; ++i) ft->exp[i] = (uint8_t)F.exp[i % F.modulus];
    ft->one = (uint8_t)F.exp[0];
    return ft;
  }();
  return t;
}

static LeoState& state(const Codec& c) { return *static_cast<LeoState*>(c.leopard.get()); }

// Column tile of a launch: as many words as fit LDS for `rows` lanes, at most
// max(64, 4096 / rows) (rows = cap_rows when given) -- so that a layer keeps
// the threads busy even when the codec has few work lanes -- or the smaller
// `want` a dispatched variant asks for.
static uint32_t tile_words(const LeoState& s, uint32_t rows, uint64_t words, uint32_t want = 0,
                           size_t reserve = 0, uint32_t cap_rows = 0, uint32_t min_half = 8) {
  const uint32_t wb = s.gf16 ? 8 : 4;
  // room for the GF16 layout's block padding and for a plan copied into LDS
  const uint64_t usable = s.lds_limit - 1024 - reserve;
  uint64_t C = usable / ((uint64_t)rows * wb);
  const uint64_t cap = std::max<uint64_t>(64, 4096 / (cap_rows ? cap_rows : rows));
  if (C > cap) C = cap;
  // Default: two workgroups per CU when that still leaves min_half words a lane
  // -- a lone workgroup leaves its table loads' and barriers' latency exposed.
  const uint64_t half_room = (s.lds_limit - 1024) / 2;
  const uint64_t half = half_room > reserve ? (half_room - reserve) / ((uint64_t)rows * wb) : 0;
  if (!want && C > half && half >= min_half) C = half;
  if (want && C > want) C = want;
  if (C < 1) C = 1;
  if (C > words) C = (uint32_t)words;
  // a power of two (items decompose with shifts); from 4 words on, whole 4-word
  // vectors (16-byte LDS / global accesses)
  uint64_t p2 = 1;
  while (p2 * 2 <= C) p2 *= 2;
  return (uint32_t)p2;
}

// Threads per workgroup.  The GF8 kernels take 72 VGPRs, so two 512-thread
// workgroups share a CU at 4 waves per SIMD.  The GF16 kernels take about 137
// (3 waves per SIMD at most) unless built to 128 (a small spill) for 4 waves per
// SIMD.  GF16 encode runs 512 threads: built to 128 VGPRs while two workgroups
// fit the LDS, at its natural count when one fills it.  GF16 decode is built to
// 128 VGPRs: 512 threads while two workgroups fit, 1024 when one fills the LDS.
// Measured on MI355X against 256 threads everywhere: GF8 128+64 encode /
// reconstruct 672 -> 787 / 261 -> 357 GB/s, GF16 300+45 reconstruct 268 -> 315,
// GF16 1024+256 reconstruct 140 -> 207, GF16 4096+1024 encode 418 -> 505; GF16
// encode at 512 threads and 128 VGPRs against 256 threads at 136 (two
// workgroups per CU): 1024+256 x 256 KiB 302 -> 289 us, 300+45 x 256 KiB 81.5 ->
// 76.4, 64+16 x 256 KiB 18.1 -> 16.6, 128+64 x 256 KiB 37.7 -> 36.4.  (Half the
// threads of a 1024+256 tile idle in its layers; two-word items to occupy them
// ran 16% slower: each item loads its twiddle tables for half the words.)
static int leo_threads(const LeoState& s, size_t lds, bool decode) {
  if (!s.gf16) return 512;
  const bool one_fills = 2 * lds > s.lds_limit;
  if (decode) return one_fills ? 1024 : 512;
  return 512;
}

// Bytes of LDS a plan takes when copied there, or 0 for a plan kept in global
// memory.  Up to 16 KiB (small codecs, whose tiles are small) the copy replaces
// each layer's chain of dependent global loads (layer, block, tables) with LDS
// reads.
constexpr size_t kPlanLdsMax = 16 << 10;
static size_t plan_lds_bytes(const LeoState& s, const LeoDevPlan& p, uint32_t rows) {
  const size_t b = (size_t)p.n_layers * 32 + (size_t)p.n_btab_words * 4;
  if (b > kPlanLdsMax) return 0;
  // the tile keeps at least 8 words a lane beside it
  const size_t wb = s.gf16 ? 8 : 4;
  return s.lds_limit > 1024 + b && (s.lds_limit - 1024 - b) / (rows * wb) >= 8 ? b : 0;
}

// Puts the plan behind `data_bytes` of LDS when it is small; returns the total.
static size_t place_plan(const LeoDevPlan& p, size_t plan_bytes, size_t data_bytes, LeoArgs& a) {
  if (!plan_bytes) return data_bytes;
  a.plan_lds = 1;
  a.plan_off = (uint32_t)((data_bytes + 15) / 16 * 4);
  a.plan_layers = p.n_layers;
  a.plan_btab_words = p.n_btab_words;
  return (size_t)a.plan_off * 4 + plan_bytes;
}

static bool all_aligned16(const std::vector<const uint8_t*>& ptrs) {
  for (auto p : ptrs)
    if ((uintptr_t)p & 15) return false;
  return true;
}

static uint32_t variant_tile(int variant) {
  return (variant >= kNumApplyVariants && variant < kNumVariants)
             ? (uint32_t)kLeoTileWords[variant - kNumApplyVariants]
             : 0;
}

// Per-thread staging for the shard-pointer and flag tables of Leopard launches
// (too large for kernel arguments at up to 65536 shards): pinned host memory
// copied on the launch stream into a device ring.  A span is reused only after
// the event recorded behind the kernel that reads it has completed, so calls
// neither allocate nor synchronise in the steady state.
struct ArgRing {
  struct Span {
    size_t b, e;
    hipEvent_t ev;
  };
  int device = -1;
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
  size_t cap = 0, head = 0;
  std::vector<Span> live;
  std::vector<hipEvent_t> spare;

  hipError_t drain() {
    for (auto& sp : live) {
      hipError_t e = hipEventSynchronize(sp.ev);
      spare.push_back(sp.ev);
      if (e != hipSuccess) return e;
    }
    live.clear();
    return hipSuccess;
  }
  // Reserves n bytes; *off receives the offset (host + off, dev + off).
  hipError_t reserve(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t)255;
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return e;
    if (d != device || n > cap) {
      if ((e = drain()) != hipSuccess) return e;
      if (host) (void)hipHostFree(host);
      if (dev) (void)hipFree(dev);
      host = dev = nullptr;
      for (auto ev : spare) (void)hipEventDestroy(ev);
      spare.clear();
      cap = std::max<size_t>(n, (size_t)1 << 20);
      head = 0;
      device = d;
      if ((e = hipHostMalloc((void**)&host, cap, hipHostMallocDefault)) != hipSuccess) return e;
      if ((e = hipMalloc((void**)&dev, cap)) != hipSuccess) return e;
    }
    if (head + n > cap) head = 0;
    const size_t b = head, en = head + n;
    size_t keep = 0;
    for (size_t i = 0; i < live.size(); ++i) {
      Span sp = live[i];
      bool overlap = sp.b < en && b < sp.e;
      if (overlap || hipEventQuery(sp.ev) == hipSuccess) {
        if (overlap && (e = hipEventSynchronize(sp.ev)) != hipSuccess) return e;
        spare.push_back(sp.ev);
      } else {
        live[keep++] = sp;
      }
    }
    live.resize(keep);
    head = en;
    *off = b;
    return hipSuccess;
  }
  // Marks [off, off + n) busy until the work queued so far on s completes.
  hipError_t commit(size_t off, size_t n, hipStream_t s) {
    hipEvent_t ev;
    if (!spare.empty()) {
      ev = spare.back();
      spare.pop_back();
    } else {
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) {
      spare.push_back(ev);
      return e;
    }
    live.push_back({off, off + ((n + 255) & ~(size_t)255), ev});
    return hipSuccess;
  }
};

// Owned by its thread for the life of the process (not destroyed at thread exit,
// when the HIP runtime may already be gone).
static ArgRing& arg_ring() {
  static thread_local ArgRing* r = new ArgRing();
  return *r;
}

// Copies the pointer table (and optional flag bytes behind it) to the device on
// stream s; *dptrs / *dflags receive the device addresses.
static int stage_args(const std::vector<const uint8_t*>& ptrs, const std::vector<uint8_t>* flags,
                      hipStream_t s, const uint8_t* const** dptrs, const uint8_t** dflags, size_t* off,
                      size_t* bytes) {
  ArgRing& ring = arg_ring();
  const size_t pb = ptrs.size() * sizeof(void*);
  const size_t n = pb + (flags ? flags->size() : 0);
  hipError_t e = ring.reserve(n, off);
  if (e != hipSuccess) {
    set_hip_error(e, "leopard arg ring");
    return RSGPU_ERR_OUT_OF_MEMORY;
  }
  std::memcpy(ring.host + *off, ptrs.data(), pb);
  if (flags) std::memcpy(ring.host + *off + pb, flags->data(), flags->size());
  e = hipMemcpyAsync(ring.dev + *off, ring.host + *off, n, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    set_hip_error(e, "leopard arg copy");
    return RSGPU_ERR_HIP;
  }
  *dptrs = (const uint8_t* const*)(ring.dev + *off);
  if (dflags) *dflags = ring.dev + *off + pb;
  *bytes = n;
  return RSGPU_OK;
}

// Puts the pointer table (and flags) of a launch into `a`: inline when it fits,
// else staged on stream s (bytes > 0 then, to be released after the launch).
static int pass_args(LeoArgs& a, const std::vector<const uint8_t*>& ptrs,
                     const std::vector<uint8_t>* flags, hipStream_t s, size_t* off, size_t* bytes) {
  *off = *bytes = 0;
  if (ptrs.size() <= kLeoInline) {
    a.inline_args = 1;
    for (size_t i = 0; i < ptrs.size(); ++i) a.iptr[i] = ptrs[i];
    if (flags)
      for (size_t i = 0; i < flags->size() && i < kLeoInline; ++i) a.ipresent[i] = (*flags)[i];
    return RSGPU_OK;
  }
  return stage_args(ptrs, flags, s, &a.ptrs, flags ? &a.present : nullptr, off, bytes);
}

static int release_args(size_t off, size_t bytes, hipStream_t s) {
  if (bytes == 0) return RSGPU_OK;  // passed inline
  hipError_t e = arg_ring().commit(off, bytes, s);
  if (e != hipSuccess) {
    set_hip_error(e, "leopard arg ring");
    return RSGPU_ERR_HIP;
  }
  return RSGPU_OK;
}

int stage_pointer_table(const std::vector<const uint8_t*>& ptrs, hipStream_t s, const uint8_t* const** d,
                        size_t* off, size_t* bytes) {
  return stage_args(ptrs, nullptr, s, d, nullptr, off, bytes);
}
int release_pointer_table(size_t off, size_t bytes, hipStream_t s) { return release_args(off, bytes, s); }

int init_leopard(Codec& c) {
  auto st = std::make_shared<LeoState>();
  st->gf16 = c.family == RSGPU_FAMILY_LEOPARD_GF16;
  const LeoField& F = st->gf16 ? field16() : field8();
  st->k = (uint32_t)c.k;
  st->parity = (uint32_t)c.m;
  st->m = ceil_pow2(st->parity);
  if (st->m > F.modulus) return RSGPU_TOO_MANY_SHARDS;
  st->work_size = st->m + st->k;
  st->n = ceil_pow2(st->work_size);
  int dev = 0, lds = 65536;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
    st->lds_limit = (size_t)lds;
  if ((size_t)st->n * (st->gf16 ? 8 : 4) > st->lds_limit ||
      (size_t)2 * st->m * (st->gf16 ? 8 : 4) > st->lds_limit)
    return RSGPU_ERR_UNSUPPORTED;
  // encode plan (EncodePlans::new, leopard_gf8/encode.rs:109-147)
  PlanBuilder pb(F);
  const uint32_t* skew = F.skew.data();
  const uint32_t m = st->m, k = st->k;
  const uint32_t mtrunc = std::min(k, m);
  LeoGroup g0{0, mtrunc, 0, 0};
  g0.n_layers = pb.ifft(mtrunc, m, skew + (m - 1), false);
  st->groups.push_back(g0);
  uint32_t i = m;
  if (m < k) {
    for (; i + m <= k; i += m) {
      LeoGroup g{i, m, (uint32_t)pb.layers.size(), 0};
      g.n_layers = pb.ifft(m, m, skew + (m - 1) + i, false);
      st->groups.push_back(g);
    }
    if (k % m) {
      LeoGroup g{i, k % m, (uint32_t)pb.layers.size(), 0};
      g.n_layers = pb.ifft(k % m, m, skew + (m - 1) + i, false);
      st->groups.push_back(g);
    }
  }
  // The groups' IFFTs as one layer list (layer j of every group, group g's
  // blocks on lanes g*m ..), for the resident encode.  All groups have the same
  // layer shapes (they depend on m only).
  bool same = st->groups.size() > 1;
  for (const LeoGroup& g : st->groups)
    same = same && g.n_layers == st->groups[0].n_layers;
  if (same) {
    st->fused_layer0 = (uint32_t)pb.layers.size();
    for (uint32_t j = 0; j < st->groups[0].n_layers; ++j) {
      const LeoLayer L0 = pb.layers[st->groups[0].layer0 + j];
      LeoLayer L{};
      L.dist = L0.dist, L.cnt = L0.cnt, L.radix = L0.radix, L.inverse = L0.inverse;
      L.lg_cnt = L0.lg_cnt, L.span_shift = L0.span_shift;
      L.block0 = (uint32_t)pb.blocks.size();
      for (size_t g = 0; g < st->groups.size(); ++g) {
        const LeoLayer Lg = pb.layers[st->groups[g].layer0 + j];
        for (uint32_t b = 0; b < Lg.n_blocks; ++b) {
          LeoBlock B = pb.blocks[Lg.block0 + b];
          B.r += (uint32_t)g * m;
          pb.blocks.push_back(B);
          ++L.n_blocks;
        }
      }
      pb.layers.push_back(L);
    }
    st->fused_layers = st->groups[0].n_layers;
  }
  st->fft_layer0 = (uint32_t)pb.layers.size();
  st->fft_layers = pb.fft(st->parity, m, skew);
  int rc = upload_plan(pb, st->enc);
  if (rc) return rc;
  if (st->d_groups.alloc(st->groups.size() * sizeof(LeoGroup)) != hipSuccess) return RSGPU_ERR_OUT_OF_MEMORY;
  if (hipMemcpy(st->d_groups.p, st->groups.data(), st->groups.size() * sizeof(LeoGroup),
                hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  std::vector<uint32_t> ex(F.exp.begin(), F.exp.begin() + F.order);
  if (st->d_log.alloc(F.order * 4) != hipSuccess || st->d_exp.alloc(F.order * 4) != hipSuccess)
    return RSGPU_ERR_OUT_OF_MEMORY;
  if (hipMemcpy(st->d_log.p, F.log.data(), F.order * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(st->d_exp.p, ex.data(), F.order * 4, hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  c.leopard = st;
  return RSGPU_OK;
}

static int launch_encode(const Codec& c, const std::vector<const uint8_t*>& ptrs, size_t len, int mode,
                         int32_t* flag, hipStream_t s, int variant) {
  LeoState& st = state(c);
  const uint64_t words = len / (st.gf16 ? 8 : 4);
  const size_t plan_bytes = plan_lds_bytes(st, st.enc, 2 * st.m);
  const uint32_t C = tile_words(st, 2 * st.m, words, variant_tile(variant), plan_bytes);
  // Every group resident in LDS when that keeps the same tile (small codecs,
  // whose tile is capped rather than LDS-bound).
  const uint32_t rrows = (uint32_t)st.groups.size() * st.m;
  const bool resident = st.fused_layers && plan_lds_bytes(st, st.enc, rrows) == plan_bytes &&
                        tile_words(st, rrows, words, variant_tile(variant), plan_bytes, 2 * st.m) == C;
  LeoArgs a{};
  size_t off = 0, bytes = 0;
  int rc = pass_args(a, ptrs, nullptr, s, &off, &bytes);
  if (rc) return rc;
  a.layers = (const LeoLayer*)st.enc.layers.p;
  a.btab = (const uint32_t*)st.enc.btab.p;
  a.groups = (const LeoGroup*)st.d_groups.p;
  a.words = words;
  a.C = C;
  a.k = st.k;
  a.parity = st.parity;
  a.m = st.m;
  a.n_groups = (uint32_t)st.groups.size();
  a.fft_layer0 = st.fft_layer0;
  a.fft_layers = st.fft_layers;
  a.flag = flag;
  a.mode = (uint32_t)mode;
  a.vec = all_aligned16(ptrs) ? 1 : 0;
  a.resident = resident ? 1 : 0;
  a.fused_layer0 = st.fused_layer0;
  a.fused_layers = st.fused_layers;
  const size_t data_lds =
      resident ? (st.gf16 && C % 4 == 0 ? (size_t)gf16_lds_dwords(rrows * C) * 4 : (size_t)rrows * C * (st.gf16 ? 8 : 4))
               : (st.gf16 && C % 4 == 0 ? (size_t)2 * gf16_lds_dwords(st.m * C) * 4
                                        : (size_t)2 * st.m * C * (st.gf16 ? 8 : 4));
  const size_t lds = place_plan(st.enc, plan_bytes, data_lds, a);
  const unsigned grid = (unsigned)((words + C - 1) / C);
  const int nt = leo_threads(st, lds, false);
  if (!st.gf16) hipLaunchKernelGGL((leo_encode_kernel<false, 512>), dim3(grid), dim3(nt), lds, s, a);
  else if (2 * lds > st.lds_limit) hipLaunchKernelGGL((leo_encode_kernel<true, 512>), dim3(grid), dim3(nt), lds, s, a);
  else hipLaunchKernelGGL((leo_encode_kernel<true, 512, 4>), dim3(grid), dim3(nt), lds, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_hip_error(e, "leo_encode_kernel");
    return RSGPU_ERR_HIP;
  }
  return release_args(off, bytes, s);
}

int leopard_encode(const Codec& c, const rsgpu_cslice* data, const rsgpu_slice* parity, size_t len,
                   hipStream_t s, int variant) {
  if (len == 0 || len % 64) return RSGPU_INCORRECT_SHARD_SIZE;  // leopard.rs:50-56
  std::vector<const uint8_t*> ptrs;
  for (size_t i = 0; i < c.k; ++i) ptrs.push_back(data[i].ptr);
  for (size_t i = 0; i < c.m; ++i) ptrs.push_back(parity[i].ptr);
  return launch_encode(c, ptrs, len, 0, nullptr, s, variant);
}

int leopard_verify(const Codec& c, const rsgpu_cslice* shards, size_t len, int32_t* d_flag,
                   hipStream_t s, int variant) {
  if (len == 0 || len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
  std::vector<const uint8_t*> ptrs;
  for (size_t i = 0; i < c.k + c.m; ++i) ptrs.push_back(shards[i].ptr);
  return launch_encode(c, ptrs, len, 1, d_flag, s, variant);
}

// Builds the decode plan of one erasure pattern (DecodePlan::new,
// leopard_gf8/decode.rs:130-152; build_decode_plan, leopard_gf16/decode.rs:507-560):
// the transform schedule, then the locator tables by a kernel on stream s.
static int build_decode_entry(const Codec& c, const std::vector<uint8_t>& present, hipStream_t s,
                              std::shared_ptr<LeoDecodeEntry>& out) {
  LeoState& st = state(c);
  const LeoField& F = st.gf16 ? field16() : field8();
  auto e = std::make_shared<LeoDecodeEntry>();
  uint32_t input_count = st.m;
  for (uint32_t i = 0; i < st.k; ++i)
    if (present[i]) input_count = st.m + i + 1;
  e->input_count = input_count;
  PlanBuilder pb(F);
  e->n_ifft = pb.ifft(input_count, st.n, F.skew.data(), true);
  e->n_fft = pb.fft(st.work_size, st.n, F.skew.data());
  int rc = upload_plan(pb, e->plan);
  if (rc) return rc;
  // erased positions of compute_error_locs
  std::vector<uint32_t> pos;
  for (uint32_t i = 0; i < st.parity; ++i)
    if (!present[st.k + i]) pos.push_back(i);
  for (uint32_t i = st.parity; i < st.m; ++i) pos.push_back(i);
  for (uint32_t i = 0; i < st.k; ++i)
    if (!present[i]) pos.push_back(st.m + i);
  const int TW = st.gf16 ? kLeoTabWords16 : kLeoTabWords8;
  if (e->pos.alloc(std::max<size_t>(1, pos.size()) * 4) != hipSuccess ||
      e->err_tabs.alloc((size_t)st.n * TW * 4) != hipSuccess ||
      e->inv_tabs.alloc((size_t)st.n * TW * 4) != hipSuccess ||
      e->err_locs.alloc((size_t)st.work_size * 4) != hipSuccess)
    return RSGPU_ERR_OUT_OF_MEMORY;
  if (!pos.empty() && hipMemcpy(e->pos.p, pos.data(), pos.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  ErrJob job{st.work_size, (uint32_t)pos.size(), st.gf16 ? 1u : 0u, F.modulus};
  hipLaunchKernelGGL(leo_errloc_kernel, dim3((st.work_size + 255) / 256), dim3(256), 0, s,
                     (const uint32_t*)st.d_log.p, (const uint32_t*)st.d_exp.p, (const uint32_t*)e->pos.p,
                     job, (uint32_t*)e->err_locs.p, (uint32_t*)e->err_tabs.p, (uint32_t*)e->inv_tabs.p);
  hipError_t he = hipGetLastError();
  if (he != hipSuccess) {
    set_hip_error(he, "leo_errloc_kernel");
    return RSGPU_ERR_HIP;
  }
  HIP_TRY(hipEventCreateWithFlags(&e->ready, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(e->ready, s));
  e->stream = s;
  out = std::move(e);
  return RSGPU_OK;
}

// Orders a launch on s after the entry's locator kernel when that kernel ran on
// another stream and is still pending.
static int wait_decode_entry(const LeoDecodeEntry& e, hipStream_t s) {
  if (e.stream == s || hipEventQuery(e.ready) == hipSuccess) return RSGPU_OK;
  HIP_TRY(hipStreamWaitEvent(s, e.ready, 0));
  return RSGPU_OK;
}

// The decode plan of a presence pattern: from the cache, from a build another
// thread has in flight, or built here.  Builds run outside the codec mutex, so a
// hit never waits behind a miss; entries evicted to make room are freed after
// the mutex is released (the reference shares GF16 plans the same way and
// dedupes builds in flight, leopard_gf16/decode.rs:586-684).
static int decode_entry(const Codec& c, const std::vector<uint8_t>& present, hipStream_t s,
                        std::shared_ptr<LeoDecodeEntry>& out) {
  LeoState& st = state(c);
  const std::string key(present.begin(), present.end());
  std::shared_ptr<LeoPending> pend;
  bool builder = false;
  {
    std::lock_guard<std::mutex> g(st.mu);
    if ((out = st.decode_cache.get(key))) return wait_decode_entry(*out, s);
    auto it = st.in_flight.find(key);
    if (it != st.in_flight.end()) {
      pend = it->second;
    } else {
      pend = std::make_shared<LeoPending>();
      st.in_flight.emplace(key, pend);
      builder = true;
    }
  }
  if (!builder) {
    std::unique_lock<std::mutex> lk(pend->m);
    pend->cv.wait(lk, [&] { return pend->done; });
    if (pend->rc) return pend->rc;
    out = pend->e;
    return wait_decode_entry(*out, s);
  }
  std::shared_ptr<LeoDecodeEntry> e;
  const int rc = build_decode_entry(c, present, s, e);
  std::vector<std::shared_ptr<LeoDecodeEntry>> evicted;  // freed after the lock below
  {
    std::lock_guard<std::mutex> g(st.mu);
    st.in_flight.erase(key);
    if (!rc) st.decode_cache.put(key, e, evicted);
  }
  {
    std::lock_guard<std::mutex> lk(pend->m);
    pend->done = true;
    pend->rc = rc;
    pend->e = e;
  }
  pend->cv.notify_all();
  if (rc) return rc;
  out = std::move(e);
  return RSGPU_OK;
}

// validate_leopard_reconstruction + reconstruct_leopard_impl (reconstruct.rs:795-1040)
int leopard_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                        const uint8_t* required, size_t n_required, hipStream_t s, int variant) {
  const size_t total = c.k + c.m;
  if (goal == 2 && n_required != c.k && n_required != total) return RSGPU_INVALID_SHARD_FLAGS;
  if (n < total) return RSGPU_TOO_FEW_SHARDS;
  if (n > total) return RSGPU_TOO_MANY_SHARDS;
  size_t len = 0, present = 0;
  bool need = false;
  for (size_t i = 0; i < n; ++i) {
    if (shards[i].present) {
      if (shards[i].len == 0) return RSGPU_EMPTY_SHARD;
      if ((len && shards[i].len != len) || shards[i].len % 64) return RSGPU_INCORRECT_SHARD_SIZE;
      len = shards[i].len;
      ++present;
    }
  }
  if (present < c.k) return RSGPU_TOO_FEW_SHARDS_PRESENT;
  std::vector<uint8_t> flags(total, 0);
  for (size_t i = 0; i < total; ++i) {
    bool req = goal == 2 ? (i < n_required && required[i]) : (goal == 0 || i < c.k);
    flags[i] = (uint8_t)((shards[i].present ? 1 : 0) | (!shards[i].present && req ? 2 : 0));
    if (!shards[i].present && req) {
      need = true;
      if (!shards[i].ptr || shards[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
    }
  }
  if (!need) return RSGPU_OK;
  if (total - present > c.m) return RSGPU_TOO_FEW_SHARDS_PRESENT;  // decode.rs:744-747
  LeoState& st = state(c);
  // Matrix variants on GF(2^8).  With exactly k shards present every input is a
  // codeword, so the generator-matrix solve over those k shards rebuilds the same
  // bytes as the FFT decoder; with more, the decoder's own matrix over all of
  // them does.
  if (variant >= 0 && variant < kNumApplyVariants && !st.gf16) {
    if (present == c.k)
      return leopard_matrix_reconstruct(c, shards, n, goal, required, n_required, s, variant);
    return leopard_decoder_matrix_reconstruct(c, shards, n, flags.data(), len, s, variant);
  }
  // GF(2^16): the decoder's own matrix over the present shards, on the vector
  // tiles of the split layout (16-byte aligned shards; else the FFT below)
  if (variant >= 0 && variant < kNumApplyVariants && st.gf16) {
    bool aligned = true;
    for (size_t i = 0; i < total; ++i)
      if (flags[i] & 3) aligned = aligned && ((uintptr_t)shards[i].ptr & 15) == 0;
    if (aligned) return leopard_decoder_matrix_reconstruct(c, shards, n, flags.data(), len, s, variant);
  }
  std::vector<uint8_t> pres(total);
  for (size_t i = 0; i < total; ++i) pres[i] = shards[i].present ? 1 : 0;
  std::shared_ptr<LeoDecodeEntry> e;
  int rc = decode_entry(c, pres, s, e);
  if (rc) return rc;
  const uint64_t words = len / (st.gf16 ? 8 : 4);
  const size_t plan_bytes = plan_lds_bytes(st, e->plan, st.n);
  // GF(2^16) decode keeps two workgroups per CU down to 4 words a lane: 1024+256
  // x 256 KiB reconstruct runs 5% faster at 4 words (two 512-thread groups) than
  // at 8 (one 1024-thread group).
  const uint32_t C = tile_words(st, st.n, words, variant_tile(variant), plan_bytes, 0, st.gf16 ? 4 : 8);
  std::vector<const uint8_t*> ptrs(total);
  for (size_t i = 0; i < total; ++i) ptrs[i] = shards[i].ptr;
  LeoArgs a{};
  size_t off = 0, bytes = 0;
  if ((rc = pass_args(a, ptrs, &flags, s, &off, &bytes))) return rc;
  a.layers = (const LeoLayer*)e->plan.layers.p;
  a.btab = (const uint32_t*)e->plan.btab.p;
  a.err_tabs = (const uint32_t*)e->err_tabs.p;
  a.inv_tabs = (const uint32_t*)e->inv_tabs.p;
  a.words = words;
  a.C = C;
  a.k = st.k;
  a.parity = st.parity;
  a.m = st.m;
  a.n = st.n;
  a.work_size = st.work_size;
  a.ifft_layer0 = 0;
  a.ifft_layers = e->n_ifft;
  a.fft_layer0 = e->n_ifft;
  a.fft_layers = e->n_fft;
  a.vec = all_aligned16(ptrs) ? 1 : 0;
  const size_t lds = place_plan(e->plan, plan_bytes,
                                st.gf16 && C % 4 == 0 ? (size_t)gf16_lds_dwords(st.n * C) * 4
                                                      : (size_t)st.n * C * (st.gf16 ? 8 : 4),
                                a);
  const unsigned grid = (unsigned)((words + C - 1) / C);
  const int nt = leo_threads(st, lds, true);
  if (!st.gf16) hipLaunchKernelGGL((leo_decode_kernel<false, 512>), dim3(grid), dim3(nt), lds, s, a);
  else if (nt == 1024) hipLaunchKernelGGL((leo_decode_kernel<true, 1024>), dim3(grid), dim3(nt), lds, s, a);
  else hipLaunchKernelGGL((leo_decode_kernel<true, 512, 4>), dim3(grid), dim3(nt), lds, s, a);
  hipError_t he = hipGetLastError();
  if (he != hipSuccess) {
    set_hip_error(he, "leo_decode_kernel");
    return RSGPU_ERR_HIP;
  }
  if ((rc = release_args(off, bytes, s))) return rc;
  for (size_t i = 0; i < total; ++i)
    if (flags[i] & 2) shards[i].present = 1;
  return RSGPU_OK;
}

}  // namespace rsgpu
