// batch.hip -- the ragged-batch kernels and their host object.
//
// Thousands of stripe operations (encode, verify, reconstruct; any mix of k, r
// and shard length) are flattened into one table of stripe descriptors plus one
// table of shard pointers in device memory.  Every stripe is cut into tiles of
// kBlock 16-byte vectors per shard; the tiles of all stripes of one element type
// form one global index space; the grid strides through it (one workgroup per
// tile by default, or a persistent grid of a few per CU).  A workgroup maps its
// tile to the stripe by a wave-uniform binary search over the stripes'
// first-tile offsets (scalar loads), then runs the same register-accumulator
// matrix apply as the single-stripe kernel with runtime k and r, software-
// pipelined so that the loads of the next input group are in flight while the
// current one computes.  One kernel instance per element type: GF(2^8) bytes,
// GF(2^16) tower pairs, GF(2^16) in Leopard's split layout.  The launch
// sequence (decode solves are enqueued when an op is added; the run is one
// kernel per element type present) can be captured in a hipGraph and replayed.
#include <algorithm>
#include <cstring>
#include <memory>

#include "batch.h"
#include "kernel_common.h"
#include "leopard.h"

namespace rsgpu {

typedef const __attribute__((address_space(4))) BatchStripe* cstripe_t;
typedef const __attribute__((address_space(4))) uint64_t* cptr_t;

// Wave-uniform position of a workgroup in the global chunk sequence: a chunk is
// one group of KB inputs of one tile of one stripe.
struct Cursor {
  uint64_t g;      // global tile
  uint64_t vtile;  // first 16-byte vector of the tile within its stripe
  uint64_t n16;
  ctab_t tables;
  uint32_t p0, stride;
  int k, R, mode, stripe, i0;
};

template <int NV>
__device__ __forceinline__ bool cursor_at(Cursor& c, uint64_t g, cstripe_t stripes, int n_stripes,
                                          uint64_t total_tiles) {
  if (g >= total_tiles) return false;
  int lo = 0, hi = n_stripes - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (stripes[mid].tile0 <= g) lo = mid;
    else hi = mid - 1;
  }
  c.g = g;
  c.stripe = lo;
  c.vtile = (g - stripes[lo].tile0) * (uint64_t)(kBlock * NV);
  c.n16 = stripes[lo].n16;
  c.tables = as_const(stripes[lo].tables);
  c.p0 = stripes[lo].ptr0;
  c.stride = stripes[lo].tab_row_stride;
  c.k = stripes[lo].k;
  c.R = stripes[lo].r;
  c.mode = stripes[lo].mode;
  c.i0 = 0;
  return true;
}

template <int KB, int NV>
__device__ __forceinline__ bool cursor_next(Cursor& c, cstripe_t stripes, int n_stripes,
                                            uint64_t total_tiles) {
  if (c.i0 + KB < c.k) {
    c.i0 += KB;
    return true;
  }
  return cursor_at<NV>(c, c.g + gridDim.x, stripes, n_stripes, total_tiles);
}

// Byte offset of 16-byte vector v of a shard: contiguous, or in the split layout
// 8 bytes of low plane at this offset and 8 of high plane 32 bytes on.
template <int ELEM>
__device__ __forceinline__ uint64_t vec_offset(uint64_t v) {
  return ELEM == kElemSplit ? ((v >> 2) << 6) + ((v & 3) << 3) : v * 16;
}

template <int ELEM>
__device__ __forceinline__ uint4 load_vec(const uint8_t* p) {
  if constexpr (ELEM == kElemSplit) return load_split(p);
  else return load16(p);
}

template <int ELEM>
__device__ __forceinline__ void store_vec(uint8_t* p, uint4 v) {
  if constexpr (ELEM == kElemSplit) store_split(p, v);
  else store16(p, v);
}

template <int KB, int NV, int ELEM>
__device__ __forceinline__ void issue(uint4 (&x)[KB][NV], const Cursor& c, cptr_t ptrs) {
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    const int i = c.i0 + b;
    const uint8_t* src = (const uint8_t*)ptrs[c.p0 + (i < c.k ? i : 0)];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const uint64_t v = c.vtile + j * kBlock + threadIdx.x;
      x[b][j] = (i < c.k && v < c.n16) ? load_vec<ELEM>(src + vec_offset<ELEM>(v)) : make_uint4(0, 0, 0, 0);
    }
  }
}

__device__ __forceinline__ uint32_t word(const uint4& v, int d) {
  return d == 0 ? v.x : d == 1 ? v.y : d == 2 ? v.z : v.w;
}

// GF(2^16): each coefficient is four byte maps [A B; C D] on the element planes
// (x' = A x ^ B c, c' = C x ^ D c; tower pairs de-interleaved into planes, split
// vectors arrive as planes), 4 * kTabWords table words per coefficient.
template <int KB, int RMAX, int NV, int ELEM>
__device__ __forceinline__ void compute16(const uint4 (&x)[KB][NV], const Cursor& c,
                                          uint32_t (&acc)[RMAX][NV][4]) {
  constexpr int W = 4 * kTabWords;
#pragma unroll
  for (int b = 0; b < KB; ++b) {
    if (c.i0 + b >= c.k) continue;  // uniform
    Sel s[NV][4];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      uint32_t w[4];
      if constexpr (ELEM == kElemTower) {
        deinterleave(x[b][j], w);
      } else {
        w[0] = x[b][j].x;
        w[1] = x[b][j].y;
        w[2] = x[b][j].z;
        w[3] = x[b][j].w;
      }
#pragma unroll
      for (int d = 0; d < 4; ++d) s[j][d] = selectors(w[d]);
    }
    ctab_t col = c.tables + (size_t)(c.i0 + b) * W;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      if (r >= c.R) continue;
      ctab_t t = col + (size_t)r * c.stride;
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          acc[r][j][d] ^= gmul(t, s[j][d]) ^ gmul(t + kTabWords, s[j][d + 2]);
          acc[r][j][d + 2] ^= gmul(t + 2 * kTabWords, s[j][d]) ^ gmul(t + 3 * kTabWords, s[j][d + 2]);
        }
    }
  }
}

// Inputs are taken in pairs so that the six byte lookups of two coefficients
// fold into the accumulator with three 3-input XORs.
template <int KB, int RMAX, int NV, int ELEM>
__device__ __forceinline__ void compute(const uint4 (&x)[KB][NV], const Cursor& c,
                                        uint32_t (&acc)[RMAX][NV][4]) {
  if constexpr (ELEM != kElemByte) {
    compute16<KB, RMAX, NV, ELEM>(x, c, acc);
    return;
  }
#pragma unroll
  for (int b = 0; b < KB; b += 2) {
    const bool has0 = c.i0 + b < c.k, has1 = c.i0 + b + 1 < c.k;  // uniform
    if (!has0) continue;
    Sel s0[NV][4], s1[NV][4];
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        s0[j][d] = selectors(word(x[b][j], d));
        s1[j][d] = selectors(word(x[b + 1][j], d));
      }
    ctab_t ta_col = c.tables + (size_t)(c.i0 + b) * kTabWords;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      if (r >= c.R) continue;
      ctab_t ta = ta_col + (size_t)r * c.stride;
      ctab_t tb = ta + kTabWords;
      if (has1) {
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            uint32_t p0 = __builtin_amdgcn_perm(ta[1], ta[0], s0[j][d].s0);
            uint32_t p1 = __builtin_amdgcn_perm(ta[3], ta[2], s0[j][d].s1);
            uint32_t p2 = __builtin_amdgcn_perm(ta[4], ta[4], s0[j][d].s2);
            uint32_t q0 = __builtin_amdgcn_perm(tb[1], tb[0], s1[j][d].s0);
            uint32_t q1 = __builtin_amdgcn_perm(tb[3], tb[2], s1[j][d].s1);
            uint32_t q2 = __builtin_amdgcn_perm(tb[4], tb[4], s1[j][d].s2);
            uint32_t t = xor3(acc[r][j][d], p0, p1);
            t = xor3(t, p2, q0);
            acc[r][j][d] = xor3(t, q1, q2);
          }
      } else {
#pragma unroll
        for (int j = 0; j < NV; ++j)
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            uint32_t p0 = __builtin_amdgcn_perm(ta[1], ta[0], s0[j][d].s0);
            uint32_t p1 = __builtin_amdgcn_perm(ta[3], ta[2], s0[j][d].s1);
            uint32_t p2 = __builtin_amdgcn_perm(ta[4], ta[4], s0[j][d].s2);
            acc[r][j][d] = xor3(acc[r][j][d], p0, p1) ^ p2;
          }
      }
    }
  }
}

// Writes (or compares) the accumulators when the chunk closes its tile.
template <int KB, int RMAX, int NV, int ELEM>
__device__ __forceinline__ void finish(const Cursor& c, uint32_t (&acc)[RMAX][NV][4], cptr_t ptrs,
                                       const BatchStripe* stripes_g) {
  if (c.i0 + KB < c.k) return;
  bool mismatch = false;
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    if (r >= c.R) continue;
    uint8_t* base = (uint8_t*)ptrs[c.p0 + c.k + r];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const uint64_t v = c.vtile + j * kBlock + threadIdx.x;
      if (v >= c.n16) continue;
      uint8_t* dst = base + vec_offset<ELEM>(v);
      uint4 o;
      if constexpr (ELEM == kElemTower) o = interleave(acc[r][j]);
      else o = make_uint4(acc[r][j][0], acc[r][j][1], acc[r][j][2], acc[r][j][3]);
      if (c.mode == kWrite) {
        store_vec<ELEM>(dst, o);
      } else {
        uint4 p = load_vec<ELEM>(dst);
        mismatch |= ((p.x ^ o.x) | (p.y ^ o.y) | (p.z ^ o.z) | (p.w ^ o.w)) != 0;
      }
    }
  }
  if (wave_leader(mismatch)) raise_flag(stripes_g[c.stripe].flag);
#pragma unroll
  for (int r = 0; r < RMAX; ++r)
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int d = 0; d < 4; ++d) acc[r][j][d] = 0;
}

// Software-pipelined grid-stride kernel: while a wave computes one chunk (KB
// inputs of its current tile) the 128-bit loads of its next chunk -- the next
// input group, or the first group of its next tile, possibly in another
// stripe -- are already in flight in the second register buffer.
template <int KB, int RMAX, int NV, int ELEM>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchStripe* stripes_g, int n_stripes,
                                                       const void* const* ptrs_g,
                                                       uint64_t total_tiles) {
  cstripe_t stripes = (cstripe_t)stripes_g;
  cptr_t ptrs = (cptr_t)ptrs_g;
  Cursor cur;
  if (!cursor_at<NV>(cur, blockIdx.x, stripes, n_stripes, total_tiles)) return;
  uint32_t acc[RMAX][NV][4];
#pragma unroll
  for (int r = 0; r < RMAX; ++r)
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int d = 0; d < 4; ++d) acc[r][j][d] = 0;
  uint4 xa[KB][NV], xb[KB][NV];
  issue<KB, NV, ELEM>(xa, cur, ptrs);
  while (true) {
    Cursor nxt = cur;
    bool more = cursor_next<KB, NV>(nxt, stripes, n_stripes, total_tiles);
    if (more) issue<KB, NV, ELEM>(xb, nxt, ptrs);
    compute<KB, RMAX, NV, ELEM>(xa, cur, acc);
    finish<KB, RMAX, NV, ELEM>(cur, acc, ptrs, stripes_g);
    if (!more) break;
    cur = nxt;
    nxt = cur;
    more = cursor_next<KB, NV>(nxt, stripes, n_stripes, total_tiles);
    if (more) issue<KB, NV, ELEM>(xa, nxt, ptrs);
    compute<KB, RMAX, NV, ELEM>(xb, cur, acc);
    finish<KB, RMAX, NV, ELEM>(cur, acc, ptrs, stripes_g);
    if (!more) break;
    cur = nxt;
  }
}

template <int KB, int RMAX>
static void launch_nv(int nv, unsigned grid, hipStream_t s, const BatchStripe* d_stripes,
                      int n_stripes, const void* const* d_ptrs, uint64_t total_tiles) {
  if (nv == 2)
    hipLaunchKernelGGL((batch_kernel<KB, RMAX, 2, kElemByte>), dim3(grid), dim3(kBlock), 0, s, d_stripes,
                       n_stripes, d_ptrs, total_tiles);
  else
    hipLaunchKernelGGL((batch_kernel<KB, RMAX, 1, kElemByte>), dim3(grid), dim3(kBlock), 0, s, d_stripes,
                       n_stripes, d_ptrs, total_tiles);
}

template <int RMAX>
static void launch_kb(int kb, int nv, unsigned grid, hipStream_t s, const BatchStripe* d_stripes,
                      int n_stripes, const void* const* d_ptrs, uint64_t total_tiles) {
  switch (kb) {
    case 4: launch_nv<4, RMAX>(nv, grid, s, d_stripes, n_stripes, d_ptrs, total_tiles); break;
    case 6: launch_nv<6, RMAX>(nv, grid, s, d_stripes, n_stripes, d_ptrs, total_tiles); break;
    case 12: launch_nv<12, RMAX>(1, grid, s, d_stripes, n_stripes, d_ptrs, total_tiles); break;
    default: launch_nv<8, RMAX>(nv, grid, s, d_stripes, n_stripes, d_ptrs, total_tiles);
  }
}

hipError_t launch_batch(int elem, const BatchStripe* d_stripes, int n_stripes, const void* const* d_ptrs,
                        uint64_t total_tiles, const BatchTune& tune, int n_cu, hipStream_t s) {
  if (total_tiles == 0 || n_stripes == 0) return hipSuccess;
  // blocks_per_cu == 0: one workgroup per tile, handed out in tile order by the
  // hardware dispatcher, so the tiles in flight stay a compact window of the
  // shards (DRAM page locality); the grid is capped so that grid x kBlock
  // threads fit 32 bits, and the kernel strides over any remainder.
  uint64_t grid = tune.blocks_per_cu > 0 ? (uint64_t)n_cu * tune.blocks_per_cu : total_tiles;
  if (grid > total_tiles) grid = total_tiles;
  const uint64_t max_grid = (UINT64_C(1) << 32) / kBlock - 1;
  if (grid > max_grid) grid = max_grid;
  if (elem == kElemTower)
    hipLaunchKernelGGL((batch_kernel<8, kMaxOut16, 1, kElemTower>), dim3((unsigned)grid), dim3(kBlock), 0, s,
                       d_stripes, n_stripes, d_ptrs, total_tiles);
  else if (elem == kElemSplit)
    hipLaunchKernelGGL((batch_kernel<8, kMaxOut16, 1, kElemSplit>), dim3((unsigned)grid), dim3(kBlock), 0, s,
                       d_stripes, n_stripes, d_ptrs, total_tiles);
  else if (tune.rmax <= 4)
    launch_kb<4>(tune.kb, tune.nv, (unsigned)grid, s, d_stripes, n_stripes, d_ptrs, total_tiles);
  else
    launch_kb<8>(tune.kb, tune.nv, (unsigned)grid, s, d_stripes, n_stripes, d_ptrs, total_tiles);
  return hipGetLastError();
}

}  // namespace rsgpu

using namespace rsgpu;

// Host object behind rsgpu_batch (declared in rsgpu.h).
struct rsgpu_batch {
  BatchPlan plan;
  BatchTune tune;
  DeviceBuffer d_stripes, d_ptrs;
  // Buffers the tables outgrew, freed with the batch: a hipFree waits for the
  // whole device, and queued launches may still read them.
  std::vector<std::unique_ptr<DeviceBuffer>> outgrown;
  // Pinned copy source of the tables, so uploads are asynchronous; `staged` is
  // recorded behind the copies out of it.
  uint8_t* h_stage = nullptr;
  size_t h_stage_bytes = 0;
  hipEvent_t staged = nullptr;
  hipStream_t stage_stream = nullptr;
  // Recorded behind every launch that reads d_stripes / d_ptrs: a later upload
  // on another stream waits for it before overwriting them in place.
  hipEvent_t ran = nullptr;
  bool dirty = true;
  int device = 0, n_cu = 256;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  // ops the batch kernels do not take (Leopard ops that their FFT serves): kept
  // and run through their own entry points on the batch's stream.
  struct Deferred {
    const rsgpu_codec* codec;
    int op;  // 0 encode, 1 verify, 2 reconstruct
    std::vector<rsgpu_slice> sl;
    std::vector<rsgpu_cslice> csl;
    std::vector<rsgpu_slot> slots;
    int32_t* flag;
    int goal;
    std::vector<int32_t> present0;  // presence as recorded when the op was added
  };
  std::vector<Deferred> deferred;
  ~rsgpu_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graph_stream) (void)hipStreamDestroy(graph_stream);
    if (staged) {
      (void)hipEventSynchronize(staged);
      (void)hipEventDestroy(staged);
    }
    if (ran) {
      (void)hipEventSynchronize(ran);
      (void)hipEventDestroy(ran);
    }
    if (h_stage) (void)hipHostFree(h_stage);
  }
};

namespace rsgpu {
// codec.hip internals used here
int batch_encode_stripe(const Codec& c, const rsgpu_slice* shards, size_t n, BatchPlan& plan);
int batch_verify_stripe(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* flag,
                        BatchPlan& plan);
int batch_reconstruct_stripe(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                             hipStream_t s, BatchPlan& plan);
// Leopard codecs: the matrix forms in the batch kernels, or
// RSGPU_ERR_UNSUPPORTED when the op runs as its own FFT launch
int batch_leopard8_encode(const Codec& c, const rsgpu_slice* shards, size_t n, BatchPlan& plan);
int batch_leopard8_verify(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* flag,
                          BatchPlan& plan);
int batch_leopard_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal, hipStream_t s,
                              BatchPlan& plan);
}  // namespace rsgpu

static void invalidate_graph(rsgpu_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void retile(rsgpu_batch* b) {
  if (b->tune.kb == 12) b->tune.nv = 1;  // the 12-input group is built for one vector per lane
  int rmax = 1;
  for (int e = 0; e < kNumElems; ++e) {
    const uint64_t tile = (uint64_t)kBlock * batch_tile_nv(e, b->tune);
    uint64_t t0 = 0;
    for (auto& s : b->plan.stripes[e]) {
      s.tile0 = t0;
      t0 += (s.n16 + tile - 1) / tile;
      if (e == kElemByte) rmax = std::max<int>(rmax, s.r);
    }
    b->plan.total_tiles[e] = t0;
  }
  b->tune.rmax = rmax <= 4 ? 4 : 8;
}

static size_t n_descriptors(const BatchPlan& p) {
  size_t n = 0;
  for (auto& v : p.stripes) n += v.size();
  return n;
}

namespace rsgpu {
int batch_reserve(rsgpu_batch* b, size_t n_stripes, size_t n_ptrs) {
  const size_t sb = std::max<size_t>(n_stripes, 1) * sizeof(BatchStripe);
  const size_t pb = std::max<size_t>(n_ptrs, 1) * sizeof(void*);
  if (b->d_stripes.bytes < sb && b->d_stripes.alloc(sb) != hipSuccess) return RSGPU_ERR_OUT_OF_MEMORY;
  if (b->d_ptrs.bytes < pb && b->d_ptrs.alloc(pb) != hipSuccess) return RSGPU_ERR_OUT_OF_MEMORY;
  if (b->h_stage_bytes < sb + pb) {
    if (b->staged && hipEventSynchronize(b->staged) != hipSuccess) return RSGPU_ERR_HIP;
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    b->h_stage = nullptr;
    b->h_stage_bytes = 0;
    if (hipHostMalloc((void**)&b->h_stage, sb + pb, hipHostMallocDefault) != hipSuccess)
      return RSGPU_ERR_OUT_OF_MEMORY;
    b->h_stage_bytes = sb + pb;
  }
  b->dirty = true;
  return RSGPU_OK;
}
}  // namespace rsgpu

static int upload(rsgpu_batch* b, hipStream_t s) {
  if (!b->dirty) {  // a launch on another stream than the upload's waits for its copies
    if (b->staged && b->stage_stream != s && hipStreamWaitEvent(s, b->staged, 0) != hipSuccess)
      return RSGPU_ERR_HIP;
    return RSGPU_OK;
  }
  retile(b);
  auto& p = b->plan;
  const size_t nd = n_descriptors(p);
  const size_t sb = std::max<size_t>(nd, 1) * sizeof(BatchStripe);
  const size_t pb = std::max<size_t>(p.ptrs.size(), 1) * sizeof(void*);
  // grow geometrically; an outgrown buffer is parked, not freed
  auto grow = [&](DeviceBuffer& d, size_t need) -> bool {
    if (d.bytes >= need) return true;
    auto old = std::make_unique<DeviceBuffer>();
    std::swap(old->p, d.p);
    std::swap(old->bytes, d.bytes);
    const size_t want = std::max(need, 2 * old->bytes);
    if (old->p) b->outgrown.push_back(std::move(old));
    return d.alloc(want) == hipSuccess;
  };
  if (!grow(b->d_stripes, sb) || !grow(b->d_ptrs, pb)) return RSGPU_ERR_OUT_OF_MEMORY;
  if (!b->staged && hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (hipEventSynchronize(b->staged) != hipSuccess) return RSGPU_ERR_HIP;  // last copies are done
  // the copies below overwrite the tables in place: order them after the
  // launches that still read them (on any stream)
  if (b->ran && hipStreamWaitEvent(s, b->ran, 0) != hipSuccess) return RSGPU_ERR_HIP;
  const size_t stage = sb + pb;
  if (b->h_stage_bytes < stage) {
    const size_t want = std::max(stage, 2 * b->h_stage_bytes);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    b->h_stage = nullptr;
    b->h_stage_bytes = 0;
    if (hipHostMalloc((void**)&b->h_stage, want, hipHostMallocDefault) != hipSuccess)
      return RSGPU_ERR_OUT_OF_MEMORY;
    b->h_stage_bytes = want;
  }
  // descriptors of the element types back to back, in kernel order
  size_t at = 0;
  for (auto& v : p.stripes) {
    std::memcpy(b->h_stage + at * sizeof(BatchStripe), v.data(), v.size() * sizeof(BatchStripe));
    at += v.size();
  }
  std::memcpy(b->h_stage + sb, p.ptrs.data(), p.ptrs.size() * sizeof(void*));
  if (nd && hipMemcpyAsync(b->d_stripes.p, b->h_stage, nd * sizeof(BatchStripe), hipMemcpyHostToDevice,
                           s) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (!p.ptrs.empty() && hipMemcpyAsync(b->d_ptrs.p, b->h_stage + sb, p.ptrs.size() * sizeof(void*),
                                        hipMemcpyHostToDevice, s) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (hipEventRecord(b->staged, s) != hipSuccess) return RSGPU_ERR_HIP;
  b->stage_stream = s;
  b->dirty = false;
  invalidate_graph(b);
  return RSGPU_OK;
}

static int run_deferred(rsgpu_batch* b, hipStream_t s) {
  for (auto& d : b->deferred) {
    int rc = RSGPU_OK;
    if (d.op == 0) rc = rsgpu_encode(d.codec, d.sl.data(), d.sl.size(), s);
    else if (d.op == 1) rc = rsgpu_verify_async(d.codec, d.csl.data(), d.csl.size(), d.flag, s);
    else {
      for (size_t i = 0; i < d.slots.size(); ++i) d.slots[i].present = d.present0[i];
      rc = d.goal == 1 ? rsgpu_reconstruct_data(d.codec, d.slots.data(), d.slots.size(), s)
                       : rsgpu_reconstruct(d.codec, d.slots.data(), d.slots.size(), s);
    }
    if (rc) return rc;
  }
  return RSGPU_OK;
}

// Records b->ran behind the launches just enqueued on s.
static int mark_ran(rsgpu_batch* b, hipStream_t s) {
  if (!b->ran && hipEventCreateWithFlags(&b->ran, hipEventDisableTiming) != hipSuccess) return RSGPU_ERR_HIP;
  if (hipEventRecord(b->ran, s) != hipSuccess) return RSGPU_ERR_HIP;
  return RSGPU_OK;
}

static int launch_main(rsgpu_batch* b, hipStream_t s) {
  auto& p = b->plan;
  hipError_t e = hipSuccess;
  size_t at = 0;
  for (int el = 0; el < kNumElems; ++el) {
    const size_t n = p.stripes[el].size();
    if (n && (e = launch_batch(el, (const BatchStripe*)b->d_stripes.p + at, (int)n,
                               (const void* const*)b->d_ptrs.p, p.total_tiles[el], b->tune, b->n_cu, s)) !=
                 hipSuccess) {
      set_hip_error(e, "batch_kernel");
      return RSGPU_ERR_HIP;
    }
    at += n;
  }
  for (auto& j : p.byte_jobs) {
    e = launch_apply_bytes(j, s);
    if (e != hipSuccess) {
      set_hip_error(e, "batch byte job");
      return RSGPU_ERR_HIP;
    }
  }
  return RSGPU_OK;
}

extern "C" {

int rsgpu_batch_new(rsgpu_batch** out) {
  if (!out) return RSGPU_ERR_INVALID_ARGUMENT;
  auto* b = new rsgpu_batch();
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) {
    delete b;
    return RSGPU_ERR_NO_DEVICE;
  }
  *out = b;
  return RSGPU_OK;
}

void rsgpu_batch_free(rsgpu_batch* b) { delete b; }

int rsgpu_batch_clear(rsgpu_batch* b) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  b->plan = BatchPlan();
  b->deferred.clear();
  b->dirty = true;
  invalidate_graph(b);
  return RSGPU_OK;
}

static bool batch_kernel_ok(const rsgpu_codec* c) { return c->c.family == RSGPU_FAMILY_CLASSIC; }

int rsgpu_batch_add_encode(rsgpu_batch* b, const rsgpu_codec* c, const rsgpu_slice* shards, size_t n) {
  if (!b || !c || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  if (c->c.family == RSGPU_FAMILY_LEOPARD_GF8) {
    const int rc = batch_leopard8_encode(c->c, shards, n, b->plan);
    if (rc != RSGPU_ERR_UNSUPPORTED) {
      if (!rc) b->dirty = true;
      return rc;
    }
  }
  if (!batch_kernel_ok(c)) {
    rsgpu_batch::Deferred d{c, 0, std::vector<rsgpu_slice>(shards, shards + n), {}, {}, nullptr, 0, {}};
    b->deferred.push_back(std::move(d));
    b->plan.n_ops++;
    invalidate_graph(b);
    return RSGPU_OK;
  }
  int rc = batch_encode_stripe(c->c, shards, n, b->plan);
  if (!rc) b->dirty = true;
  return rc;
}

int rsgpu_batch_add_verify(rsgpu_batch* b, const rsgpu_codec* c, const rsgpu_cslice* shards,
                           size_t n, int32_t* device_flag) {
  if (!b || !c || !shards || !device_flag) return RSGPU_ERR_INVALID_ARGUMENT;
  if (c->c.family == RSGPU_FAMILY_LEOPARD_GF8) {
    const int rc = batch_leopard8_verify(c->c, shards, n, device_flag, b->plan);
    if (rc != RSGPU_ERR_UNSUPPORTED) {
      if (!rc) b->dirty = true;
      return rc;
    }
  }
  if (!batch_kernel_ok(c)) {
    rsgpu_batch::Deferred d{c, 1, {}, std::vector<rsgpu_cslice>(shards, shards + n), {}, device_flag, 0, {}};
    b->deferred.push_back(std::move(d));
    b->plan.n_ops++;
    invalidate_graph(b);
    return RSGPU_OK;
  }
  int rc = batch_verify_stripe(c->c, shards, n, device_flag, b->plan);
  if (!rc) b->dirty = true;
  return rc;
}

int rsgpu_batch_add_reconstruct(rsgpu_batch* b, const rsgpu_codec* c, rsgpu_slot* shards, size_t n,
                                int32_t data_only, void* stream) {
  if (!b || !c || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  if (c->c.family != RSGPU_FAMILY_CLASSIC) {
    const int rc = batch_leopard_reconstruct(c->c, shards, n, data_only ? 1 : 0, (hipStream_t)stream, b->plan);
    if (rc != RSGPU_ERR_UNSUPPORTED) {
      if (!rc) b->dirty = true;
      return rc;
    }
  }
  if (!batch_kernel_ok(c)) {
    rsgpu_batch::Deferred d{c, 2, {}, {}, std::vector<rsgpu_slot>(shards, shards + n), nullptr,
                            data_only ? 1 : 0, {}};
    size_t present = 0;
    for (size_t i = 0; i < n; ++i) {
      d.present0.push_back(shards[i].present);
      present += shards[i].present ? 1 : 0;
    }
    // the targets read as present from here on, as after an op in the kernel;
    // the run restores the recorded presence before the call
    if (n == c->c.k + c->c.m && present >= c->c.k)
      for (size_t i = 0; i < n; ++i)
        if (!shards[i].present && (!data_only || i < c->c.k)) shards[i].present = 1;
    b->deferred.push_back(std::move(d));
    b->plan.n_ops++;
    invalidate_graph(b);
    return RSGPU_OK;
  }
  int rc = batch_reconstruct_stripe(c->c, shards, n, data_only ? 1 : 0, (hipStream_t)stream, b->plan);
  if (!rc) b->dirty = true;
  return rc;
}

int rsgpu_batch_set_tune(rsgpu_batch* b, int32_t inputs_per_group, int32_t vectors_per_lane,
                         int32_t blocks_per_cu) {
  const int kb = inputs_per_group, nv = vectors_per_lane;
  if (!b || (kb != 4 && kb != 6 && kb != 8 && kb != 12) || (nv != 1 && nv != 2) ||
      (kb == 12 && nv != 1) || blocks_per_cu < 0 || blocks_per_cu > 64)
    return RSGPU_ERR_INVALID_ARGUMENT;
  if (nv != b->tune.nv) b->dirty = true;
  b->tune.nv = nv;
  b->tune.kb = kb;
  b->tune.blocks_per_cu = blocks_per_cu;
  invalidate_graph(b);
  return RSGPU_OK;
}

int rsgpu_batch_run(rsgpu_batch* b, void* stream) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(b, s);
  if (rc) return rc;
  if ((rc = launch_main(b, s)) || (rc = mark_ran(b, s))) return rc;
  return run_deferred(b, s);
}

int rsgpu_batch_graph_launch(rsgpu_batch* b, void* stream) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(b, s);
  if (rc) return rc;
  if (!b->deferred.empty()) return rsgpu_batch_run(b, stream);  // deferred ops sync on the host
  // A batch that is one kernel launch runs it directly: a graph replay adds its own
  // start-up (64 x 10+4 x 64 KiB encode: 27 us per step replayed, 18.7 direct).
  size_t launches = b->plan.byte_jobs.size();
  for (auto& v : b->plan.stripes) launches += v.empty() ? 0 : 1;
  if (launches <= 1) {
    if ((rc = launch_main(b, s))) return rc;
    return mark_ran(b, s);
  }
  if (!b->exec) {
    // Capture on the batch's own stream (the legacy default stream cannot be
    // captured); the instantiated graph is then launched on the caller's stream.
    if (!b->graph_stream) {
      hipError_t e = hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking);
      if (e != hipSuccess) {
        set_hip_error(e, "hipStreamCreateWithFlags");
        return RSGPU_ERR_HIP;
      }
    }
    hipStream_t cs = b->graph_stream;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      set_hip_error(e, "hipStreamBeginCapture");
      return RSGPU_ERR_HIP;
    }
    rc = launch_main(b, cs);
    e = hipStreamEndCapture(cs, &b->graph);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_hip_error(e, "hipStreamEndCapture");
      return RSGPU_ERR_HIP;
    }
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_hip_error(e, "hipGraphInstantiate");
      return RSGPU_ERR_HIP;
    }
  }
  hipError_t e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) {
    set_hip_error(e, "hipGraphLaunch");
    return RSGPU_ERR_HIP;
  }
  return mark_ran(b, s);
}

int rsgpu_batch_stats(const rsgpu_batch* b, uint64_t* n_ops, uint64_t* bytes_moved,
                      uint64_t* data_bytes, uint64_t* tiles) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  if (n_ops) *n_ops = (uint64_t)b->plan.n_ops;
  if (bytes_moved) *bytes_moved = b->plan.bytes_moved;
  if (data_bytes) *data_bytes = b->plan.data_bytes;
  if (tiles) {
    uint64_t t = 0;
    for (int e = 0; e < kNumElems; ++e) {
      const uint64_t tile = (uint64_t)kBlock * batch_tile_nv(e, b->tune);
      for (auto& s : b->plan.stripes[e]) t += (s.n16 + tile - 1) / tile;
    }
    *tiles = t;
  }
  return RSGPU_OK;
}

// Exhaustive tile-shape search: every candidate is bit-exact, so the fastest
// (device time over `reps` runs) wins.  Leaves the batch tuned to the winner.
int rsgpu_batch_autotune(rsgpu_batch* b, void* stream, int32_t reps, int32_t* best_kb,
                         int32_t* best_nv, int32_t* best_blocks_per_cu, float* best_ms) {
  if (!b || reps < 1) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  static const int shapes[][2] = {{4, 1}, {6, 1}, {8, 1}, {12, 1}, {4, 2}, {6, 2}};
  static const int bpcs[] = {2, 4, 6, 8, 12, 16, 32, 0};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return RSGPU_ERR_HIP;
  float best = 1e30f;
  BatchTune winner = b->tune;
  int rc = RSGPU_OK;
  for (auto& sh : shapes)
    for (int bpc : bpcs) {
      if ((rc = rsgpu_batch_set_tune(b, sh[0], sh[1], bpc))) break;
      if ((rc = upload(b, s))) break;
      if ((rc = launch_main(b, s))) break;  // warm
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps && !rc; ++r) rc = launch_main(b, s);
      (void)hipEventRecord(e1, s);
      if (!rc) rc = mark_ran(b, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= reps;
      if (ms < best) {
        best = ms;
        winner = b->tune;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rsgpu_batch_set_tune(b, winner.kb, winner.nv, winner.blocks_per_cu);
  if (best_kb) *best_kb = winner.kb;
  if (best_nv) *best_nv = winner.nv;
  if (best_blocks_per_cu) *best_blocks_per_cu = winner.blocks_per_cu;
  if (best_ms) *best_ms = best;
  return RSGPU_OK;
}

}  // extern "C"
