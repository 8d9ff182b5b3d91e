// batch.h -- many classic stripes (mixed shapes, shard sizes and operations)
// executed by one kernel launch.
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
  uint16_t mode, tower;
};

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
  std::vector<BatchStripe> stripes;
  std::vector<const void*> ptrs;
  std::vector<std::shared_ptr<void>> keep;  // decode entries referenced by the stripes
  // tails (length % 16) and unaligned stripes run through the byte kernel
  std::vector<ByteArgs> byte_jobs;
  uint64_t total_tiles = 0;
  uint64_t bytes_moved = 0;  // algorithmic bytes (read k, write/compare r)
  uint64_t data_bytes = 0;   // k * shard bytes
  int n_ops = 0;
};

hipError_t launch_batch(const BatchStripe* d_stripes, int n_stripes, const void* const* d_ptrs,
                        uint64_t total_tiles, const BatchTune& tune, int n_cu, hipStream_t s);

}  // namespace rsgpu
