// batch.hip -- the persistent ragged-batch kernel and its host object.
//
// Thousands of classic GF(2^8) stripe operations (encode, verify, reconstruct;
// any mix of k, r and shard length) are flattened into one table of stripe
// descriptors plus one table of shard pointers in device memory.  Every stripe
// is cut into tiles of kBlock * NV 16-byte vectors per shard; the tiles of all
// stripes form one global index space that a persistent grid strides through.
// A workgroup maps its tile to the stripe by a wave-uniform binary search over
// the stripes' first-tile offsets (scalar loads), then runs the same
// register-accumulator matrix apply as the single-stripe kernel with runtime
// k and r (r <= 8 outputs per descriptor).  The launch sequence (decode solves
// are enqueued when an op is added; the run is a single kernel) can be captured
// in a hipGraph and replayed.
#include <algorithm>
#include <cstring>

#include "batch.h"
#include "kernel_common.h"
#include "leopard.h"

namespace rsgpu {

typedef const __attribute__((address_space(4))) BatchStripe* cstripe_t;
typedef const __attribute__((address_space(4))) uint64_t* cptr_t;

template <int NV>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchStripe* stripes_g, int n_stripes,
                                                       const void* const* ptrs_g,
                                                       uint64_t total_tiles) {
  constexpr int KB = NV == 1 ? 16 : 8;
  cstripe_t stripes = (cstripe_t)stripes_g;
  cptr_t ptrs = (cptr_t)ptrs_g;
  const uint64_t tile = (uint64_t)kBlock * NV;
  for (uint64_t g = blockIdx.x; g < total_tiles; g += gridDim.x) {
    int lo = 0, hi = n_stripes - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (stripes[mid].tile0 <= g) lo = mid;
      else hi = mid - 1;
    }
    const uint64_t n16 = stripes[lo].n16;
    const int k = stripes[lo].k, R = stripes[lo].r, mode = stripes[lo].mode;
    const uint32_t p0 = stripes[lo].ptr0, stride = stripes[lo].tab_row_stride;
    ctab_t tables = as_const(stripes[lo].tables);
    const uint64_t t = g - stripes[lo].tile0;
    uint64_t v[NV];
    bool ok[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      v[j] = t * tile + (uint64_t)j * kBlock + threadIdx.x;
      ok[j] = v[j] < n16;
      v[j] *= 16;
    }
    uint32_t acc[kMaxOut][NV][4];
#pragma unroll
    for (int r = 0; r < kMaxOut; ++r)
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int d = 0; d < 4; ++d) acc[r][j][d] = 0;
    for (int i0 = 0; i0 < k; i0 += KB) {
      uint4 x[KB][NV];
#pragma unroll
      for (int b = 0; b < KB; ++b) {
        const uint8_t* src = (const uint8_t*)ptrs[p0 + ((i0 + b < k) ? i0 + b : 0)];
#pragma unroll
        for (int j = 0; j < NV; ++j)
          x[b][j] = (i0 + b < k && ok[j]) ? load16(src + v[j]) : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (int b = 0; b < KB; ++b) {
        if (i0 + b >= k) continue;  // uniform; keeps the loop fully unrollable
        ctab_t tcol = tables + (size_t)(i0 + b) * kTabWords;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          Sel s[4] = {selectors(x[b][j].x), selectors(x[b][j].y), selectors(x[b][j].z),
                      selectors(x[b][j].w)};
#pragma unroll
          for (int r = 0; r < kMaxOut; ++r) {
            if (r >= R) continue;
            ctab_t tc = tcol + (size_t)r * stride;
#pragma unroll
            for (int d = 0; d < 4; ++d) acc[r][j][d] ^= gmul(tc, s[d]);
          }
        }
      }
    }
    bool mismatch = false;
#pragma unroll
    for (int r = 0; r < kMaxOut; ++r) {
      if (r >= R) continue;
      uint8_t* dst = (uint8_t*)ptrs[p0 + k + r];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        if (!ok[j]) continue;
        uint4 o = make_uint4(acc[r][j][0], acc[r][j][1], acc[r][j][2], acc[r][j][3]);
        if (mode == kWrite) {
          store16(dst + v[j], o);
        } else {
          uint4 p = load16(dst + v[j]);
          mismatch |= ((p.x ^ o.x) | (p.y ^ o.y) | (p.z ^ o.z) | (p.w ^ o.w)) != 0;
        }
      }
    }
    if (mismatch) atomicOr(stripes_g[lo].flag, 1);
  }
}

hipError_t launch_batch(const BatchStripe* d_stripes, int n_stripes, const void* const* d_ptrs,
                        uint64_t total_tiles, const BatchTune& tune, int n_cu, hipStream_t s) {
  if (total_tiles == 0 || n_stripes == 0) return hipSuccess;
  uint64_t grid = (uint64_t)n_cu * tune.blocks_per_cu;
  if (grid > total_tiles) grid = total_tiles;
  if (tune.nv == 2)
    hipLaunchKernelGGL(batch_kernel<2>, dim3((unsigned)grid), dim3(kBlock), 0, s, d_stripes,
                       n_stripes, d_ptrs, total_tiles);
  else
    hipLaunchKernelGGL(batch_kernel<1>, dim3((unsigned)grid), dim3(kBlock), 0, s, d_stripes,
                       n_stripes, d_ptrs, total_tiles);
  return hipGetLastError();
}

}  // namespace rsgpu

using namespace rsgpu;

// Host object behind rsgpu_batch (declared in rsgpu.h).
struct rsgpu_batch {
  BatchPlan plan;
  BatchTune tune;
  DeviceBuffer d_stripes, d_ptrs;
  bool dirty = true;
  int device = 0, n_cu = 256;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  // ops the persistent kernel does not take (tower field, Leopard): kept and run
  // through their own entry points on the batch's stream.
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
  }
};

namespace rsgpu {
// codec.hip internals used here
int batch_encode_stripe(const Codec& c, const rsgpu_slice* shards, size_t n, BatchPlan& plan);
int batch_verify_stripe(const Codec& c, const rsgpu_cslice* shards, size_t n, int32_t* flag,
                        BatchPlan& plan);
int batch_reconstruct_stripe(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                             hipStream_t s, BatchPlan& plan);
}  // namespace rsgpu

static void invalidate_graph(rsgpu_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void retile(rsgpu_batch* b) {
  const uint64_t tile = (uint64_t)kBlock * b->tune.nv;
  uint64_t t0 = 0;
  for (auto& s : b->plan.stripes) {
    s.tile0 = t0;
    t0 += (s.n16 + tile - 1) / tile;
  }
  b->plan.total_tiles = t0;
}

static int upload(rsgpu_batch* b, hipStream_t s) {
  if (!b->dirty) return RSGPU_OK;
  retile(b);
  auto& p = b->plan;
  size_t sb = std::max<size_t>(p.stripes.size(), 1) * sizeof(BatchStripe);
  size_t pb = std::max<size_t>(p.ptrs.size(), 1) * sizeof(void*);
  if (b->d_stripes.bytes < sb && b->d_stripes.alloc(sb) != hipSuccess) return RSGPU_ERR_OUT_OF_MEMORY;
  if (b->d_ptrs.bytes < pb && b->d_ptrs.alloc(pb) != hipSuccess) return RSGPU_ERR_OUT_OF_MEMORY;
  if (!p.stripes.empty() &&
      hipMemcpyAsync(b->d_stripes.p, p.stripes.data(), p.stripes.size() * sizeof(BatchStripe),
                     hipMemcpyHostToDevice, s) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (!p.ptrs.empty() && hipMemcpyAsync(b->d_ptrs.p, p.ptrs.data(), p.ptrs.size() * sizeof(void*),
                                        hipMemcpyHostToDevice, s) != hipSuccess)
    return RSGPU_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return RSGPU_ERR_HIP;  // host vectors may change later
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

static int launch_main(rsgpu_batch* b, hipStream_t s) {
  auto& p = b->plan;
  hipError_t e = launch_batch((const BatchStripe*)b->d_stripes.p, (int)p.stripes.size(),
                              (const void* const*)b->d_ptrs.p, p.total_tiles, b->tune, b->n_cu, s);
  if (e != hipSuccess) {
    set_hip_error(e, "batch_kernel");
    return RSGPU_ERR_HIP;
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

static bool persistent_ok(const rsgpu_codec* c) {
  return c->c.family == RSGPU_FAMILY_CLASSIC && c->c.field == RSGPU_FIELD_GF8;
}

int rsgpu_batch_add_encode(rsgpu_batch* b, const rsgpu_codec* c, const rsgpu_slice* shards, size_t n) {
  if (!b || !c || !shards) return RSGPU_ERR_INVALID_ARGUMENT;
  if (!persistent_ok(c)) {
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
  if (!persistent_ok(c)) {
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
  if (!persistent_ok(c)) {
    rsgpu_batch::Deferred d{c, 2, {}, {}, std::vector<rsgpu_slot>(shards, shards + n), nullptr,
                            data_only ? 1 : 0, {}};
    for (size_t i = 0; i < n; ++i) d.present0.push_back(shards[i].present);
    b->deferred.push_back(std::move(d));
    b->plan.n_ops++;
    invalidate_graph(b);
    return RSGPU_OK;
  }
  int rc = batch_reconstruct_stripe(c->c, shards, n, data_only ? 1 : 0, (hipStream_t)stream, b->plan);
  if (!rc) b->dirty = true;
  return rc;
}

int rsgpu_batch_set_tune(rsgpu_batch* b, int32_t nv, int32_t blocks_per_cu) {
  if (!b || (nv != 1 && nv != 2) || blocks_per_cu < 1 || blocks_per_cu > 64)
    return RSGPU_ERR_INVALID_ARGUMENT;
  if (nv != b->tune.nv) b->dirty = true;
  b->tune.nv = nv;
  b->tune.blocks_per_cu = blocks_per_cu;
  invalidate_graph(b);
  return RSGPU_OK;
}

int rsgpu_batch_run(rsgpu_batch* b, void* stream) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(b, s);
  if (rc) return rc;
  if ((rc = launch_main(b, s))) return rc;
  return run_deferred(b, s);
}

int rsgpu_batch_graph_launch(rsgpu_batch* b, void* stream) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(b, s);
  if (rc) return rc;
  if (!b->deferred.empty()) return rsgpu_batch_run(b, stream);  // deferred ops sync on the host
  if (!b->exec || b->graph_stream != s) {
    invalidate_graph(b);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return RSGPU_ERR_HIP;
    rc = launch_main(b, s);
    hipError_t e = hipStreamEndCapture(s, &b->graph);
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
    b->graph_stream = s;
  }
  hipError_t e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) {
    set_hip_error(e, "hipGraphLaunch");
    return RSGPU_ERR_HIP;
  }
  return RSGPU_OK;
}

int rsgpu_batch_stats(const rsgpu_batch* b, uint64_t* n_ops, uint64_t* bytes_moved,
                      uint64_t* data_bytes, uint64_t* tiles) {
  if (!b) return RSGPU_ERR_INVALID_ARGUMENT;
  if (n_ops) *n_ops = (uint64_t)b->plan.n_ops;
  if (bytes_moved) *bytes_moved = b->plan.bytes_moved;
  if (data_bytes) *data_bytes = b->plan.data_bytes;
  if (tiles) {
    const uint64_t tile = (uint64_t)kBlock * b->tune.nv;
    uint64_t t = 0;
    for (auto& s : b->plan.stripes) t += (s.n16 + tile - 1) / tile;
    *tiles = t;
  }
  return RSGPU_OK;
}

// Exhaustive tile-shape search: every candidate is bit-exact, so the fastest
// (device time over `reps` runs) wins.  Leaves the batch tuned to the winner.
int rsgpu_batch_autotune(rsgpu_batch* b, void* stream, int32_t reps, int32_t* best_nv,
                         int32_t* best_blocks_per_cu, float* best_ms) {
  if (!b || reps < 1) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  static const int nvs[] = {1, 2};
  static const int bpcs[] = {2, 4, 8, 16, 32};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return RSGPU_ERR_HIP;
  float best = 1e30f;
  BatchTune winner = b->tune;
  int rc = RSGPU_OK;
  for (int nv : nvs)
    for (int bpc : bpcs) {
      if ((rc = rsgpu_batch_set_tune(b, nv, bpc))) break;
      if ((rc = upload(b, s))) break;
      if ((rc = launch_main(b, s))) break;  // warm
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps && !rc; ++r) rc = launch_main(b, s);
      (void)hipEventRecord(e1, s);
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
  rsgpu_batch_set_tune(b, winner.nv, winner.blocks_per_cu);
  if (best_nv) *best_nv = winner.nv;
  if (best_blocks_per_cu) *best_blocks_per_cu = winner.blocks_per_cu;
  if (best_ms) *best_ms = best;
  return RSGPU_OK;
}

}  // extern "C"
