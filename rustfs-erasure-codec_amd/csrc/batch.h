// batch.h -- many stripe operations (mixed shapes, shard sizes and operations)
// executed by one kernel launch per element type present.
#pragma once
#include "internal.h"

namespace rsgpu {

// One stripe operation as the batch kernel sees it (device memory).
struct BatchStripe {
  uint64_t n16;            // 16-byte vectors per shard
  uint64_t tile0;          // first global tile index of this stripe
  const uint32_t* tables;  // [r][k][W] byte-permute tables
  int32_t* flag;           // verify: set nonzero on mismatch
  uint32_t ptr0;           // index of in[0] in the pointer table; outputs follow the k inputs
  uint32_t tab_row_stride; // dwords between table rows
  uint16_t k, r;
  uint16_t mode, elem;     // elem: the stripe's element type (kElemByte, kElemTower, kElemSplit)
};

// Element types of the batch kernels, one kernel instance each:
//   kElemByte  GF(2^8) bytes (classic GF(2^8), Leopard GF(2^8) matrix forms);
//   kElemTower GF(2^16) tower elements as interleaved [x, c] byte pairs (classic GF(2^16));
//   kElemSplit GF(2^16) elements in the split layout of Leopard shards (each 64-byte
//              block: the low bytes of 32 elements, then their high bytes).
constexpr int kElemByte = 0, kElemTower = 1, kElemSplit = 2, kNumElems = 3;
// Outputs per descriptor of the GF(2^16) kernels (their accumulators are compiled
// for four rows).
constexpr int kMaxOut16 = 4;

// Shape of the batch kernel: inputs per pipelined load group (kb), 16-byte
// vectors per lane (nv; tiles are kBlock * nv vectors per shard), resident
// workgroups per CU of a persistent grid (0: one workgroup per tile, the
// fastest grid on every workload measured), and the accumulator rows compiled
// in (4 or 8, from the widest descriptor in the batch).
struct BatchTune {
  int kb = 8;
  int nv = 1;
  int blocks_per_cu = 0;
  int rmax = 4;
};

struct BatchPlan {
  std::vector<BatchStripe> stripes[kNumElems];  // by element type
  std::vector<const void*> ptrs;
  std::vector<std::shared_ptr<void>> keep;  // decode entries referenced by the stripes
  // tails (length % 16) and unaligned stripes run through the byte kernel
  std::vector<ByteArgs> byte_jobs;
  uint64_t total_tiles[kNumElems] = {0, 0, 0};
  uint64_t bytes_moved = 0;  // algorithmic bytes (read k, write/compare r)
  uint64_t data_bytes = 0;   // k * shard bytes
  int n_ops = 0;
};

// One batch kernel over the descriptors of one element type.  The GF(2^16)
// kernels take a fixed shape (8 inputs per load group, one vector per lane, four
// rows) and the tune's grid.
hipError_t launch_batch(int elem, const BatchStripe* d_stripes, int n_stripes, const void* const* d_ptrs,
                        uint64_t total_tiles, const BatchTune& tune, int n_cu, hipStream_t s);
// 16-byte vectors per lane of a kernel's tiles.
inline int batch_tile_nv(int elem, const BatchTune& t) { return elem == kElemByte ? t.nv : 1; }

}  // namespace rsgpu
