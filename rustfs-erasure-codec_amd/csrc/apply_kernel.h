// apply_kernel.h -- the per-stripe r x k matrix-apply kernel, templated on its
// tile shape, and its launchers.  Instantiated by apply_inst.hip, one object per
// (tile shape, element type), so the sets compile in parallel.
//
// A workgroup of LB lanes owns a column tile of LB * NV 16-byte vectors of every
// shard and the R outputs [blockIdx.y * R, blockIdx.y * R + R).  Each lane reads
// its vectors of all k inputs with 128-bit loads, keeps its outputs'
// accumulators in VGPRs and writes every output once (encode, reconstruct), or
// compares it in registers against the stored parity without writing (verify).
// A GF(2^8) product c*x of a packed dword is three v_perm_b32 byte lookups on
// the bit fields x[2:0], x[5:3], x[7:6] (tables from make_perm_table, read into
// SGPRs), folded with 3-input XORs.  The bit-field selectors depend only on the
// input and are shared by all R outputs.
//
// Tile shapes trade parallelism for reuse: small LB and R = 1 give many
// workgroups for short shards (each re-reads the inputs, from L2), large NV and
// R = r give the fewest passes over long shards.
#pragma once
#include "kernel_common.h"

namespace rsgpu {

template <int LB, int R, int NV, int MODE, bool TOWER>
__global__ __launch_bounds__(LB) void apply_kernel(const ApplyArgs a) {
  constexpr int W = TOWER ? 4 * kTabWords : kTabWords;  // table dwords per coefficient
  // Inputs per load group: every lane issues KB*NV 128-bit loads back to back
  // before the arithmetic (64 VGPRs of input in flight for every NV).
  constexpr int KB = NV == 1 ? 16 : (NV == 2 ? 8 : 4);
  const int r0 = blockIdx.y * R;  // first output of this workgroup
  const uint64_t tile = (uint64_t)LB * NV;
  const uint64_t n_tiles = (a.n16 + tile - 1) / tile;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint64_t v[NV];
    bool ok[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      v[j] = t * tile + (uint64_t)j * LB + threadIdx.x;
      ok[j] = v[j] < a.n16;
      v[j] = (a.v_begin + v[j]) * 16;  // byte offset
    }
    uint32_t acc[R][NV][4];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int j = 0; j < NV; ++j)
#pragma unroll
        for (int d = 0; d < 4; ++d) acc[r][j][d] = 0;

    for (int i0 = 0; i0 < a.k; i0 += KB) {
      uint4 x[KB][NV];
#pragma unroll
      for (int b = 0; b < KB; ++b)
#pragma unroll
        for (int j = 0; j < NV; ++j)
          x[b][j] = (i0 + b < a.k && ok[j]) ? load16(a.in[i0 + b] + v[j]) : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int b = 0; b < KB; ++b) {
        if (i0 + b >= a.k) continue;  // uniform; `continue` keeps the loop fully unrollable
        ctab_t tcol = as_const(a.tables) + (size_t)(a.tab_col0 + i0 + b) * W +
                      (size_t)r0 * a.tab_row_stride;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          uint32_t w[4];
          if constexpr (TOWER) {
            deinterleave(x[b][j], w);
          } else {
            w[0] = x[b][j].x;
            w[1] = x[b][j].y;
            w[2] = x[b][j].z;
            w[3] = x[b][j].w;
          }
          Sel s[4];
#pragma unroll
          for (int d = 0; d < 4; ++d) s[d] = selectors(w[d]);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (r0 + r >= a.r) continue;  // last row group of an uneven split
            ctab_t tc = tcol + (size_t)r * a.tab_row_stride;
            if constexpr (TOWER) {
              // [A B; C D] per tower_coefficients: x' = A x ^ B c, c' = C x ^ D c
#pragma unroll
              for (int d = 0; d < 2; ++d) {
                acc[r][j][d] ^= gmul(tc + 0 * kTabWords, s[d]) ^ gmul(tc + 1 * kTabWords, s[d + 2]);
                acc[r][j][d + 2] ^= gmul(tc + 2 * kTabWords, s[d]) ^ gmul(tc + 3 * kTabWords, s[d + 2]);
              }
            } else {
#pragma unroll
              for (int d = 0; d < 4; ++d) acc[r][j][d] ^= gmul(tc, s[d]);
            }
          }
        }
      }
    }

    bool mismatch = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      if (!ok[j]) continue;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r0 + r >= a.r) continue;
        uint4 o;
        if constexpr (TOWER) {
          o = interleave(acc[r][j]);
        } else {
          o = make_uint4(acc[r][j][0], acc[r][j][1], acc[r][j][2], acc[r][j][3]);
        }
        uint8_t* dst = a.out[r0 + r] + v[j];
        if constexpr (MODE == kWrite) {
          store16(dst, o);
        } else if constexpr (MODE == kAccumulate) {
          uint4 p = load16(dst);
          store16(dst, make_uint4(p.x ^ o.x, p.y ^ o.y, p.z ^ o.z, p.w ^ o.w));
        } else {
          uint4 p = load16(dst);
          mismatch |= ((p.x ^ o.x) | (p.y ^ o.y) | (p.z ^ o.z) | (p.w ^ o.w)) != 0;
        }
      }
    }
    if constexpr (MODE == kVerify) {
      if (mismatch) atomicOr(a.flag, 1);
    }
  }
}

template <int LB, int R, int NV, int MODE, bool TOWER>
static hipError_t launch_apply_t(const ApplyArgs& a, hipStream_t s) {
  const uint64_t tile = (uint64_t)LB * NV;
  uint64_t tiles = (a.n16 + tile - 1) / tile;
  unsigned gx = (unsigned)(tiles < 1048576 ? (tiles ? tiles : 1) : 1048576);
  unsigned gy = (unsigned)((a.r + R - 1) / R);
  hipLaunchKernelGGL((apply_kernel<LB, R, NV, MODE, TOWER>), dim3(gx, gy), dim3(LB), 0, s, a);
  return hipGetLastError();
}

template <int LB, int NV, int MODE, bool TOWER>
static hipError_t launch_apply_rows(const ApplyArgs& a, int rows, hipStream_t s) {
  switch (rows) {
    case 1: return launch_apply_t<LB, 1, NV, MODE, TOWER>(a, s);
    case 2: return launch_apply_t<LB, 2, NV, MODE, TOWER>(a, s);
    case 3: return launch_apply_t<LB, 3, NV, MODE, TOWER>(a, s);
    case 4: return launch_apply_t<LB, 4, NV, MODE, TOWER>(a, s);
    case 5: return launch_apply_t<LB, 5, NV, MODE, TOWER>(a, s);
    case 6: return launch_apply_t<LB, 6, NV, MODE, TOWER>(a, s);
    case 7: return launch_apply_t<LB, 7, NV, MODE, TOWER>(a, s);
    case 8: return launch_apply_t<LB, 8, NV, MODE, TOWER>(a, s);
  }
  return hipErrorInvalidValue;
}

template <int LB, int NV, bool TOWER>
static hipError_t launch_apply_mode(const ApplyArgs& a, int mode, int rows, hipStream_t s) {
  switch (mode) {
    case kWrite: return launch_apply_rows<LB, NV, kWrite, TOWER>(a, rows, s);
    case kAccumulate: return launch_apply_rows<LB, NV, kAccumulate, TOWER>(a, rows, s);
    case kVerify: return launch_apply_rows<LB, NV, kVerify, TOWER>(a, rows, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace rsgpu
