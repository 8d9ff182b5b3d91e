// apply_inst.hip -- one instantiation set of the tiled matrix-apply kernel
// (apply_kernel.h): tile shape kTileShapes[APPLY_TILE], GF(2^8) elements
// (APPLY_TOWER = 0), GF(2^16) tower elements (APPLY_TOWER = 1) or split-layout
// GF(2^16) elements of Leopard shards (APPLY_TOWER = 2), all output counts and
// modes.  build.py compiles this file once per (tile, field) pair so
// the sets build in parallel; launch_apply (classic_kernels.hip) picks the set.
#include "apply_kernel.h"

#if !defined(APPLY_TILE) || !defined(APPLY_TOWER)
#error "compile with -DAPPLY_TILE=<0..4> -DAPPLY_TOWER=<0|1|2>"
#endif

namespace rsgpu {

#define RSGPU_CAT_(t, w) launch_apply_t##t##_w##w
#define RSGPU_CAT(t, w) RSGPU_CAT_(t, w)

hipError_t RSGPU_CAT(APPLY_TILE, APPLY_TOWER)(const ApplyArgs& a, int mode, int rows, hipStream_t s) {
  constexpr TileShape ts = kTileShapes[APPLY_TILE];
  return launch_apply_mode<ts.lanes, ts.nv, APPLY_TOWER>(a, mode, rows, s);
}

}  // namespace rsgpu
