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
#include <algorithm>
#include <cstddef>

#include "kernel_common.h"

namespace rsgpu {

// s_waitcnt immediate for vmcnt(n) with expcnt and lgkmcnt left open (gfx9
// encoding: vmcnt bits [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
constexpr unsigned wait_vm(unsigned n) { return (n & 15u) | ((n >> 4) << 14) | (7u << 4) | (15u << 8); }
constexpr unsigned kWaitVm16 = wait_vm(16);

// LDS bytes of one launch: the 16-byte table chunks of every (input, output)
// pair, rounded up to whole staging passes of LB chunks.
template <int LB, int R, int ELEM>
size_t apply_lds_bytes(int k) {
  const size_t chunks = (size_t)k * R * (ELEM != 0 ? 4 * kTabWords : kTabWords) / 4;
  return (chunks + LB - 1) / LB * LB * 16;
}

// ELEM: 0 GF(2^8) bytes; 1 GF(2^16) tower elements as interleaved [x, c] byte
// pairs; 2 GF(2^16) elements in the split layout of Leopard shards (each 64-byte
// block holds the low bytes of 32 elements, then their high bytes), applied with
// the same four byte maps per coefficient, low bytes as the first plane.
template <int LB, int R, int NV, int MODE, int ELEM>
__global__ __launch_bounds__(LB) void apply_kernel(const ApplyArgs a) {
  constexpr bool TOWER = ELEM != 0, SPLIT = ELEM == 2;
  constexpr int W = TOWER ? 4 * kTabWords : kTabWords;  // table dwords per coefficient
  // Inputs per load group: every lane issues KB*NV 128-bit loads back to back
  // before the arithmetic (64 VGPRs of input in flight for every NV).
  constexpr int KB = NV == 1 ? 16 : (NV == 2 ? 8 : 4);
  static_assert(KB * NV == 16, "the staging wait counts 16 shard loads");
  // [input][output][W]: the coefficient tables of this workgroup's outputs, the
  // same block layout as in global memory.  Sized by apply_lds_bytes.
  extern __shared__ __align__(16) uint32_t s_tab[];
  const int r0 = blockIdx.y * R;  // first output of this workgroup
  const int rn = a.r - r0 < R ? a.r - r0 : R;
  const uint64_t tile = (uint64_t)LB * NV;
  if constexpr (MODE == kVerify) verify_claim(a.clear_ws, a.call_id, a.flag);
  // inputs of this workgroup: all of a.k, or chunk blockIdx.z of a split over inputs
  const bool splitk = a.in_dev != nullptr;
  const int kz = splitk ? min(a.k, a.k_total - (int)blockIdx.z * a.k) : a.k;
  const int col0 = a.tab_col0 + (splitk ? (int)blockIdx.z * a.k : 0);

  // Table staging: 16-byte chunks go straight from global memory into LDS
  // (global_load_lds, no registers), issued before the shard loads so that both
  // round trips overlap.  Lane t of a pass fills chunk c0 + t; lanes past the
  // last chunk re-read it into their own (unused) slot and rows past the last
  // output re-read row rn - 1, so every address stays inside the tables.
  {
    constexpr int CW = W / 4;  // chunks per coefficient table
    const int n_chunks = kz * R * CW;
    // tables in device memory, or inline in this launch's kernel arguments
    const uint32_t* tab = a.tables ? a.tables
                                   : (const uint32_t*)((const char*)__builtin_amdgcn_kernarg_segment_ptr() +
                                                       offsetof(ApplyArgs, tab_inline));
    const int wave = (int)threadIdx.x / 64;
    for (int c0 = 0; c0 < n_chunks; c0 += LB) {
      int c = c0 + (int)threadIdx.x;
      c = c < n_chunks ? c : n_chunks - 1;
      const int blk = c / CW, part = c % CW, i = blk / R, r = blk % R;
      const int rr = r < rn ? r : rn - 1;
      const uint32_t* src =
          tab + (size_t)(r0 + rr) * a.tab_row_stride + (size_t)(col0 + i) * W + part * 4;
      __builtin_amdgcn_global_load_lds(src, s_tab + (size_t)(c0 + wave * 64) * 4, 16, 0, 0);
    }
  }
  __builtin_amdgcn_sched_barrier(0);

  // One tile per workgroup: the launch covers every tile (blockIdx.x, offset by
  // v_begin for very long shards).
  uint64_t v[NV];
  bool ok[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const uint64_t u = (uint64_t)blockIdx.x * tile + (uint64_t)j * LB + threadIdx.x;
    ok[j] = u < a.n16;
    // a lane past the end loads the last vector instead (in bounds, unused), so
    // the loads need no per-lane branch
    const uint64_t vi = a.v_begin + (ok[j] ? u : a.n16 - 1);
    // byte offset of the vector (split: of its 8 low bytes, the high bytes 32 on)
    v[j] = SPLIT ? ((vi >> 2) << 6) + ((vi & 3) << 3) : vi * 16;
  }
  uint32_t acc[R][NV][4];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int j = 0; j < NV; ++j)
#pragma unroll
      for (int d = 0; d < 4; ++d) acc[r][j][d] = 0;

  uint4 x[KB][NV];
  // Issues the loads of inputs [i0, i0 + KB): the pointers first (kernel-argument
  // reads, batched), then the shard vectors in input order.  Slots past the last
  // input re-read the last input (the same lines: cache hits, never used), so
  // every load is unconditional, the compiler counts them exactly, and input
  // b's arithmetic waits for its own loads only while the later ones fly.
  auto load_group = [&](int i0) __attribute__((always_inline)) {
    const uint8_t* p[KB];
    if (splitk) {  // this chunk's pointers from the device array, past its end the last
      const uint8_t* const* ip = a.in_dev + (size_t)blockIdx.z * a.k;
#pragma unroll
      for (int b = 0; b < KB; ++b) p[b] = ip[i0 + b < kz ? i0 + b : kz - 1];
    } else {
      const uint8_t* last = a.in[a.k - 1];
#pragma unroll
      for (int b = 0; b < KB; ++b) p[b] = a.in[i0 + b];  // i0 + b < kMaxIn: inside ApplyArgs::in
#pragma unroll
      for (int b = 0; b < KB; ++b) p[b] = i0 + b < a.k ? p[b] : last;
    }
#pragma unroll
    for (int b = 0; b < KB; ++b) {
#pragma unroll
      for (int j = 0; j < NV; ++j) x[b][j] = SPLIT ? load_split(p[b] + v[j]) : load16(p[b] + v[j]);
      __builtin_amdgcn_sched_barrier(0);  // keep the issue order
    }
  };
  // Input b's tables for all R outputs are read from LDS first (16-byte reads),
  // so their latency overlaps the selector arithmetic.  Rows past the last
  // output (an uneven row split) hold copies of row rn - 1: computed, never
  // stored, and no branch splits the outputs.
  constexpr int NT = TOWER ? 4 : 1;  // byte maps per coefficient
  auto compute_group = [&](int i0) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < KB; ++b) {
      if (i0 + b >= kz) continue;  // uniform; `continue` keeps the loop fully unrollable
      const uint32_t* tcol = s_tab + (size_t)(i0 + b) * R * W;
      uint4 tq[R][NT];
      uint32_t t4[R][NT];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          const uint32_t* t = tcol + r * W + q * kTabWords;
          tq[r][q] = *reinterpret_cast<const uint4*>(t);
          t4[r][q] = t[4];
        }
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        uint32_t w[4];
        if constexpr (TOWER && !SPLIT) {
          deinterleave(x[b][j], w);
        } else {  // split vectors arrive as planes already
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
          if constexpr (TOWER) {
            // [A B; C D] per tower_coefficients: x' = A x ^ B c, c' = C x ^ D c
#pragma unroll
            for (int d = 0; d < 2; ++d) {
              acc[r][j][d] ^= gmul(tq[r][0], t4[r][0], s[d]) ^ gmul(tq[r][1], t4[r][1], s[d + 2]);
              acc[r][j][d + 2] ^= gmul(tq[r][2], t4[r][2], s[d]) ^ gmul(tq[r][3], t4[r][3], s[d + 2]);
            }
          } else {
#pragma unroll
            for (int d = 0; d < 4; ++d) acc[r][j][d] ^= gmul(tq[r][0], t4[r][0], s[d]);
          }
        }
      }
    }
  };

  load_group(0);
  // The staging loads are older than the 16 shard loads just issued: once at
  // most 16 loads are outstanding they have landed in LDS.  The barrier then
  // publishes every wave's chunks.
  __builtin_amdgcn_s_waitcnt(kWaitVm16);
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
  compute_group(0);
  for (int i0 = KB; i0 < kz; i0 += KB) {
    load_group(i0);
    compute_group(i0);
  }

  bool mismatch = false;
  // a split over inputs writes partial rows (write mode only)
  uint8_t* outp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = r0 + (r < rn ? r : rn - 1);
    outp[r] = splitk ? a.part + ((size_t)blockIdx.z * a.r + row) * a.part_pitch : a.out[row];
  }
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    if (!ok[j]) continue;
    uint4 o[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if constexpr (TOWER && !SPLIT) {
        o[r] = interleave(acc[r][j]);
      } else {
        o[r] = make_uint4(acc[r][j][0], acc[r][j][1], acc[r][j][2], acc[r][j][3]);
      }
    }
    if constexpr (MODE == kWrite) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < rn) {
          if constexpr (SPLIT) store_split(outp[r] + v[j], o[r]);
          else store16(outp[r] + v[j], o[r]);
        }
    } else {
      // the stored outputs, all loaded before any is used (rows past rn re-read
      // row rn - 1, unused)
      uint4 p[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const uint8_t* src = outp[r] + v[j];
        p[r] = SPLIT ? load_split(src) : load16(src);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r >= rn) continue;
        if constexpr (MODE == kAccumulate) {
          const uint4 q = make_uint4(p[r].x ^ o[r].x, p[r].y ^ o[r].y, p[r].z ^ o[r].z, p[r].w ^ o[r].w);
          if constexpr (SPLIT) store_split(outp[r] + v[j], q);
          else store16(outp[r] + v[j], q);
        } else {
          mismatch |= ((p[r].x ^ o[r].x) | (p[r].y ^ o[r].y) | (p[r].z ^ o[r].z) | (p[r].w ^ o[r].w)) != 0;
        }
      }
    }
  }
  if constexpr (MODE == kVerify) {
    if (wave_leader(mismatch)) verify_report(a.clear_ws, a.call_id, a.flag);
  }
}

// The default GF(2^8) tiling: one dword of every shard per lane.  The same
// arithmetic as the 16-byte tiles, spread over four times the waves, so the
// memory, LDS and VALU latencies of one wave hide behind the others' work.
__device__ __forceinline__ uint32_t load4(const uint8_t* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
}
__device__ __forceinline__ void store4(uint8_t* p, uint32_t v) {
  __builtin_nontemporal_store(v, reinterpret_cast<uint32_t*>(p));
}

constexpr int kDwordLanes = 256;

template <int R, int MODE, int KB>
__global__ __launch_bounds__(kDwordLanes) void apply_dword_kernel(const ApplyArgs a) {
  constexpr int LB = kDwordLanes;
  constexpr int W = kTabWords;
  // KB inputs per load group (the staging wait counts KB loads): 16, or 8 / 12 for
  // calls of at most that many inputs (one group; slots past k re-read the last
  // input, so a smaller group issues fewer redundant loads)
  static_assert(KB == 8 || KB == 12 || KB == 16, "load group");
  extern __shared__ __align__(16) uint32_t s_tab[];  // [input][output][W]
  const int r0 = blockIdx.y * R;
  const int rn = a.r - r0 < R ? a.r - r0 : R;
  if constexpr (MODE == kVerify) verify_claim(a.clear_ws, a.call_id, a.flag);
  // table staging, as in apply_kernel
  {
    constexpr int CW = W / 4;
    const int n_chunks = a.k * R * CW;
    const uint32_t* tab = a.tables ? a.tables
                                   : (const uint32_t*)((const char*)__builtin_amdgcn_kernarg_segment_ptr() +
                                                       offsetof(ApplyArgs, tab_inline));
    const int wave = (int)threadIdx.x / 64;
    for (int c0 = 0; c0 < n_chunks; c0 += LB) {
      int c = c0 + (int)threadIdx.x;
      c = c < n_chunks ? c : n_chunks - 1;
      const int blk = c / CW, part = c % CW, i = blk / R, r = blk % R;
      const int rr = r < rn ? r : rn - 1;
      const uint32_t* src = tab + (size_t)(r0 + rr) * a.tab_row_stride + (size_t)(a.tab_col0 + i) * W + part * 4;
      __builtin_amdgcn_global_load_lds(src, s_tab + (size_t)(c0 + wave * 64) * 4, 16, 0, 0);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  const uint64_t n4 = a.n16 * 4;  // dwords per shard
  const uint64_t u = (uint64_t)blockIdx.x * LB + threadIdx.x;
  const bool ok = u < n4;
  const uint64_t off = (a.v_begin * 4 + (ok ? u : n4 - 1)) * 4;  // past the end: the last dword
  uint32_t acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0;
  uint32_t x[KB];
  auto load_group = [&](int i0) __attribute__((always_inline)) {
    const uint8_t* p[KB];
    const uint8_t* last = a.in[a.k - 1];
#pragma unroll
    for (int b = 0; b < KB; ++b) p[b] = a.in[i0 + b];  // i0 + b < kMaxIn: inside ApplyArgs::in
#pragma unroll
    for (int b = 0; b < KB; ++b) p[b] = i0 + b < a.k ? p[b] : last;
#pragma unroll
    for (int b = 0; b < KB; ++b) {
      x[b] = load4(p[b] + off);
      __builtin_amdgcn_sched_barrier(0);  // keep the issue order
    }
  };
  auto compute_group = [&](int i0) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < KB; ++b) {
      if (i0 + b >= a.k) continue;
      const uint32_t* tcol = s_tab + (size_t)(i0 + b) * R * W;
      uint4 tq[R];
      uint32_t t4[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        tq[r] = *reinterpret_cast<const uint4*>(tcol + r * W);
        t4[r] = tcol[r * W + 4];
      }
      const Sel sl = selectors(x[b]);
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] ^= gmul(tq[r], t4[r], sl);
    }
  };
  load_group(0);
  __builtin_amdgcn_s_waitcnt(wait_vm(KB));
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
  compute_group(0);
  for (int i0 = KB; i0 < a.k; i0 += KB) {
    load_group(i0);
    compute_group(i0);
  }
  if (!ok) return;
  bool mismatch = false;
  if constexpr (MODE == kWrite) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < rn) store4(a.out[r0 + r] + off, acc[r]);
  } else {
    // the stored outputs, all loaded before any is used (rows past rn re-read row
    // rn - 1, unused)
    uint32_t p[R];
#pragma unroll
    for (int r = 0; r < R; ++r) p[r] = load4(a.out[r0 + (r < rn ? r : rn - 1)] + off);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r >= rn) continue;
      if constexpr (MODE == kAccumulate) store4(a.out[r0 + r] + off, p[r] ^ acc[r]);
      else mismatch |= p[r] != acc[r];
    }
  }
  if constexpr (MODE == kVerify) {
    if (wave_leader(mismatch)) verify_report(a.clear_ws, a.call_id, a.flag);
  }
}

template <int R, int MODE>
static hipError_t launch_apply_dword_t(const ApplyArgs& a, hipStream_t s) {
  // Inputs per load group: one group of 8 or 12 for calls of at most that many
  // inputs, else groups of 16.  Measured per call (back to back, 1 MiB shards) on
  // MI355X against groups of 16 throughout: 8+3 encode 4.47 -> 4.03 us, verify
  // 4.03 -> 3.68, reconstruct of 3 erasures 4.40 -> 3.97; 4+2 encode 3.83 -> 3.34;
  // 10+4 encode 5.58 -> 5.47.  Exact groups of 4, 6 and 10 measured the same as
  // 8 and 12.
  const int kb = a.k <= 8 ? 8 : (a.k <= 12 ? 12 : 16);
  constexpr uint64_t kMaxTiles = 1ull << 30;  // launches of at most 2^30 workgroups
  const uint64_t n4 = a.n16 * 4;
  const uint64_t tiles = (n4 + kDwordLanes - 1) / kDwordLanes;
  const unsigned gy = (unsigned)((a.r + R - 1) / R);
  const size_t lds = apply_lds_bytes<kDwordLanes, R, 0>(a.k);
  for (uint64_t t0 = 0; t0 < tiles; t0 += kMaxTiles) {
    ApplyArgs part = a;  // v_begin and n16 stay in 16-byte vectors: kMaxTiles * 256 is a multiple of 4
    part.v_begin = a.v_begin + t0 * kDwordLanes / 4;
    part.n16 = std::min<uint64_t>(a.n16 - t0 * kDwordLanes / 4, kMaxTiles * kDwordLanes / 4);
    const unsigned gx = (unsigned)std::min<uint64_t>(tiles - t0, kMaxTiles);
    if (kb == 8) hipLaunchKernelGGL((apply_dword_kernel<R, MODE, 8>), dim3(gx, gy), dim3(kDwordLanes), lds, s, part);
    else if (kb == 12) hipLaunchKernelGGL((apply_dword_kernel<R, MODE, 12>), dim3(gx, gy), dim3(kDwordLanes), lds, s, part);
    else hipLaunchKernelGGL((apply_dword_kernel<R, MODE, 16>), dim3(gx, gy), dim3(kDwordLanes), lds, s, part);
  }
  return hipGetLastError();
}

template <int LB, int R, int NV, int MODE, int ELEM>
static hipError_t launch_apply_t(const ApplyArgs& a, hipStream_t s) {
  // one workgroup per tile; launches of at most 2^30 tiles, each shifted by v_begin
  constexpr uint64_t kMaxTiles = 1ull << 30;
  const uint64_t tile = (uint64_t)LB * NV;
  const uint64_t tiles = (a.n16 + tile - 1) / tile;
  const unsigned gy = (unsigned)((a.r + R - 1) / R);
  const size_t lds = apply_lds_bytes<LB, R, ELEM>(a.k);
  for (uint64_t t0 = 0; t0 < tiles; t0 += kMaxTiles) {
    ApplyArgs part = a;
    part.v_begin = a.v_begin + t0 * tile;
    part.n16 = std::min<uint64_t>(a.n16 - t0 * tile, kMaxTiles * tile);
    const unsigned gx = (unsigned)std::min<uint64_t>(tiles - t0, kMaxTiles);
    const unsigned gz = a.in_dev ? (unsigned)((a.k_total + a.k - 1) / a.k) : 1u;
    hipLaunchKernelGGL((apply_kernel<LB, R, NV, MODE, ELEM>), dim3(gx, gy, gz), dim3(LB), lds, s, part);
  }
  return hipGetLastError();
}

template <int LB, int NV, int MODE, int ELEM>
static hipError_t launch_apply_rows(const ApplyArgs& a, int rows, hipStream_t s) {
  switch (rows) {
    case 1: return launch_apply_t<LB, 1, NV, MODE, ELEM>(a, s);
    case 2: return launch_apply_t<LB, 2, NV, MODE, ELEM>(a, s);
    case 3: return launch_apply_t<LB, 3, NV, MODE, ELEM>(a, s);
    case 4: return launch_apply_t<LB, 4, NV, MODE, ELEM>(a, s);
    case 5: return launch_apply_t<LB, 5, NV, MODE, ELEM>(a, s);
    case 6: return launch_apply_t<LB, 6, NV, MODE, ELEM>(a, s);
    case 7: return launch_apply_t<LB, 7, NV, MODE, ELEM>(a, s);
    case 8: return launch_apply_t<LB, 8, NV, MODE, ELEM>(a, s);
  }
  return hipErrorInvalidValue;
}

template <int LB, int NV, int ELEM>
static hipError_t launch_apply_mode(const ApplyArgs& a, int mode, int rows, hipStream_t s) {
  switch (mode) {
    case kWrite: return launch_apply_rows<LB, NV, kWrite, ELEM>(a, rows, s);
    case kAccumulate: return launch_apply_rows<LB, NV, kAccumulate, ELEM>(a, rows, s);
    case kVerify: return launch_apply_rows<LB, NV, kVerify, ELEM>(a, rows, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace rsgpu
