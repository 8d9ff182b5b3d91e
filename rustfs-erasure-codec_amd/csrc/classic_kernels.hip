// classic_kernels.hip -- gfx950 kernels of the classic (matrix) Reed-Solomon
// data path: the r x k matrix apply that serves encode, verify and
// reconstruct, and the device-side solve that builds decode rows.
//
// Matrix apply: a wavefront owns a 1 KiB (NV = 1) or 2 KiB (NV = 2) column tile
// of every shard.  Each lane reads its 16-byte vector of all k inputs with
// 128-bit loads, keeps the r output accumulators in VGPRs and writes every
// output once (encode, reconstruct), or compares it in registers against the
// stored parity without writing (verify).  A GF(2^8) product c*x of a packed
// dword is three v_perm_b32 byte lookups on the bit fields x[2:0], x[5:3],
// x[7:6] (tables from make_perm_table, held in SGPRs), folded with 3-input
// XORs (v_bitop3_b32).  The bit-field selectors depend only on the input and
// are shared by all r outputs.
#include "kernel_common.h"

namespace rsgpu {


// R outputs, NV vectors per lane, TOWER selects GF(2^16) tower element pairs.
template <int R, int NV, int MODE, bool TOWER>
__global__ __launch_bounds__(kBlock) void apply_kernel(const ApplyArgs a) {
  constexpr int W = TOWER ? 4 * kTabWords : kTabWords;  // table dwords per coefficient
  // Inputs per load group: every lane issues KB*NV 128-bit loads back to back
  // before the arithmetic, so a k <= 16 stripe has all its input in flight.
  constexpr int KB = NV == 1 ? 16 : 8;
  const uint64_t tile = (uint64_t)kBlock * NV;
  const uint64_t n_tiles = (a.n16 + tile - 1) / tile;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint64_t v[NV];
    bool ok[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      v[j] = t * tile + (uint64_t)j * kBlock + threadIdx.x;
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
        ctab_t tcol = as_const(a.tables) + (size_t)(a.tab_col0 + i0 + b) * W;
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
        uint4 o;
        if constexpr (TOWER) {
          o = interleave(acc[r][j]);
        } else {
          o = make_uint4(acc[r][j][0], acc[r][j][1], acc[r][j][2], acc[r][j][3]);
        }
        uint8_t* dst = a.out[r] + v[j];
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

// Byte-granular variant for shard tails (length % 16) and unaligned views.
__device__ __forceinline__ uint8_t gmul_byte(const uint32_t* t, uint8_t x) {
  uint32_t a = t[(x & 7) >> 2] >> (8 * (x & 3));
  uint32_t b = t[2 + (((x >> 3) & 7) >> 2)] >> (8 * ((x >> 3) & 3));
  uint32_t c = t[4] >> (8 * (x >> 6));
  return (uint8_t)(a ^ b ^ c);
}

__global__ __launch_bounds__(kBlock) void apply_bytes_kernel(const ByteArgs a) {
  const int W = a.tower ? 4 * kTabWords : kTabWords;
  const uint64_t unit = a.tower ? 2 : 1;
  const uint64_t n_units = (a.end - a.begin) / unit;
  for (uint64_t u = (uint64_t)blockIdx.x * kBlock + threadIdx.x; u < n_units;
       u += (uint64_t)gridDim.x * kBlock) {
    uint64_t off = a.begin + u * unit;
    uint8_t accx[kMaxOut], accc[kMaxOut];
    for (int r = 0; r < a.r; ++r) accx[r] = accc[r] = 0;
    for (int i = 0; i < a.k; ++i) {
      const uint32_t* tcol = a.tables + (size_t)(a.tab_col0 + i) * W;
      uint8_t x = a.in[i][off];
      uint8_t c = a.tower ? a.in[i][off + 1] : 0;
      for (int r = 0; r < a.r; ++r) {
        const uint32_t* tc = tcol + (size_t)r * a.tab_row_stride;
        if (a.tower) {
          accx[r] ^= gmul_byte(tc, x) ^ gmul_byte(tc + kTabWords, c);
          accc[r] ^= gmul_byte(tc + 2 * kTabWords, x) ^ gmul_byte(tc + 3 * kTabWords, c);
        } else {
          accx[r] ^= gmul_byte(tc, x);
        }
      }
    }
    bool mismatch = false;
    for (int r = 0; r < a.r; ++r) {
      uint8_t* dst = a.out[r] + off;
      if (a.mode == kWrite) {
        dst[0] = accx[r];
        if (a.tower) dst[1] = accc[r];
      } else if (a.mode == kAccumulate) {
        dst[0] ^= accx[r];
        if (a.tower) dst[1] ^= accc[r];
      } else {
        mismatch |= dst[0] != accx[r];
        if (a.tower) mismatch |= dst[1] != accc[r];
      }
    }
    if (mismatch) atomicOr(a.flag, 1);
  }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int R, int NV, int MODE, bool TOWER>
static hipError_t launch_t(const ApplyArgs& a, hipStream_t s) {
  const uint64_t tile = (uint64_t)kBlock * NV;
  uint64_t tiles = (a.n16 + tile - 1) / tile;
  unsigned grid = (unsigned)(tiles < 1048576 ? (tiles ? tiles : 1) : 1048576);
  hipLaunchKernelGGL((apply_kernel<R, NV, MODE, TOWER>), dim3(grid), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

template <int NV, int MODE, bool TOWER>
static hipError_t launch_r(const ApplyArgs& a, hipStream_t s) {
  switch (a.r) {
    case 1: return launch_t<1, NV, MODE, TOWER>(a, s);
    case 2: return launch_t<2, NV, MODE, TOWER>(a, s);
    case 3: return launch_t<3, NV, MODE, TOWER>(a, s);
    case 4: return launch_t<4, NV, MODE, TOWER>(a, s);
    case 5: return launch_t<5, NV, MODE, TOWER>(a, s);
    case 6: return launch_t<6, NV, MODE, TOWER>(a, s);
    case 7: return launch_t<7, NV, MODE, TOWER>(a, s);
    case 8: return launch_t<8, NV, MODE, TOWER>(a, s);
  }
  return hipErrorInvalidValue;
}

template <int NV, bool TOWER>
static hipError_t launch_m(const ApplyArgs& a, int mode, hipStream_t s) {
  switch (mode) {
    case kWrite: return launch_r<NV, kWrite, TOWER>(a, s);
    case kAccumulate: return launch_r<NV, kAccumulate, TOWER>(a, s);
    case kVerify: return launch_r<NV, kVerify, TOWER>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_apply(const ApplyArgs& a, int mode, int nv, int tower, hipStream_t s) {
  if (a.n16 == 0) return hipSuccess;
  if (tower) return nv == 2 ? launch_m<2, true>(a, mode, s) : launch_m<1, true>(a, mode, s);
  return nv == 2 ? launch_m<2, false>(a, mode, s) : launch_m<1, false>(a, mode, s);
}

hipError_t launch_apply_bytes(const ByteArgs& a, hipStream_t s) {
  uint64_t units = (a.end - a.begin) / (a.tower ? 2 : 1);
  if (units == 0) return hipSuccess;
  uint64_t blocks = (units + kBlock - 1) / kBlock;
  unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(apply_bytes_kernel, dim3(grid), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, int n, hipStream_t s) {
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- decode rows
//
// For survivors valid[0..k) = [present data d_1..d_p, parity q_1..q_e] (the first
// k present shards, recovery_plan.rs:329-387) the missing data solve
//   C * D_missing = S_q ^ B * D_known,  C = P[q][missing], B = P[q][known]
// by Gauss-Jordan elimination of [C | B | I] -> [I | C^-1 B | C^-1]
// (the partial elimination of reconstruct.rs:1110-1158; with p = 0 it is the
// full inverse of reconstruct.rs:1164-1173).  Row j of [C^-1 B | C^-1] is the
// decode row of missing data j over the survivors in valid order.  A requested
// parity p gets the composed row P[p] * Dfull (recovery_plan.rs:483-515).
// Elements: GF(2^8) bytes, or tower elements (x << 8 | c).

__device__ __forceinline__ uint8_t d_mul8(const uint8_t* lg, const uint8_t* ex, uint8_t a, uint8_t b) {
  return (a == 0 || b == 0) ? 0 : ex[lg[a] + lg[b]];
}

struct DevField {
  const uint8_t* lg;
  const uint8_t* ex;
  bool tower;
  __device__ uint16_t mul(uint16_t a, uint16_t b) const {
    if (!tower) return d_mul8(lg, ex, (uint8_t)a, (uint8_t)b);
    uint8_t ax = a >> 8, ac = a & 0xff, bx = b >> 8, bc = b & 0xff;
    uint8_t o0 = d_mul8(lg, ex, ax, bx);
    uint8_t o1 = d_mul8(lg, ex, ac, bx) ^ d_mul8(lg, ex, ax, bc);
    uint8_t o2 = d_mul8(lg, ex, ac, bc);
    if (o0) {
      o1 ^= d_mul8(lg, ex, 2, o0);
      o2 ^= d_mul8(lg, ex, 128, o0);
    }
    return (uint16_t)((o1 << 8) | o2);
  }
  __device__ uint16_t inv(uint16_t a) const {
    if (!tower) return a ? ex[255 - lg[a]] : 0;
    uint8_t ax = a >> 8, ac = a & 0xff;
    if (ax == 0) return ac ? ex[255 - lg[ac]] : 0;
    uint8_t conj = ac ^ d_mul8(lg, ex, 2, ax);
    uint8_t norm = d_mul8(lg, ex, ac, conj) ^ d_mul8(lg, ex, 128, d_mul8(lg, ex, ax, ax));
    uint8_t ni = ex[255 - lg[norm]];
    return (uint16_t)((d_mul8(lg, ex, ax, ni) << 8) | d_mul8(lg, ex, conj, ni));
  }
};

constexpr int kMaxSolve = 128;  // missing data rows solvable in one workgroup

__global__ __launch_bounds__(256) void decode_rows_kernel(const uint16_t* __restrict__ matrix,
                                                          const DecodeJob job,
                                                          uint8_t* __restrict__ rows_out,
                                                          uint32_t* __restrict__ tables_out,
                                                          int32_t* __restrict__ status) {
  __shared__ uint8_t lg[256];
  __shared__ uint8_t ex[512];
  extern __shared__ uint16_t smem[];  // [e][cols] work matrix, then [n_targets][k] rows
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += blockDim.x) lg[i] = kGf8.log[i];
  for (int i = tid; i < 512; i += blockDim.x) ex[i] = kGf8.exp[i];
  const int k = job.k;
  int p = 0;  // present data among survivors (a sorted prefix of valid)
  while (p < k && job.valid[p] < k) ++p;
  const int e = k - p;  // missing data count
  const int cols = e + p + e;
  uint16_t* wm = smem;
  uint16_t* rows = smem + e * cols;
  __shared__ uint16_t missing[kMaxSolve];
  __shared__ int pivot_row;
  __syncthreads();
  DevField F{lg, ex, job.tower != 0};
  if (tid == 0) {
    int j = 0, kn = 0;
    for (int d = 0; d < k && j < e; ++d) {
      if (kn < p && job.valid[kn] == d) {
        ++kn;
      } else {
        missing[j++] = (uint16_t)d;
      }
    }
  }
  __syncthreads();
  // [C | B | I]
  for (int idx = tid; idx < e * cols; idx += blockDim.x) {
    int r = idx / cols, c = idx % cols;
    const uint16_t* prow = matrix + (size_t)job.valid[p + r] * k;
    uint16_t v;
    if (c < e) v = prow[missing[c]];
    else if (c < e + p) v = prow[job.valid[c - e]];
    else v = (c - e - p == r) ? 1 : 0;
    wm[idx] = v;
  }
  __syncthreads();
  bool singular = false;
  for (int r = 0; r < e; ++r) {
    if (tid == 0) {
      int pr = -1;
      for (int rb = r; rb < e; ++rb)
        if (wm[rb * cols + r] != 0) {
          pr = rb;
          break;
        }
      pivot_row = pr;
    }
    __syncthreads();
    int pr = pivot_row;
    if (pr < 0) {
      singular = true;
      break;
    }
    if (pr != r)
      for (int c = tid; c < cols; c += blockDim.x) {
        uint16_t t = wm[r * cols + c];
        wm[r * cols + c] = wm[pr * cols + c];
        wm[pr * cols + c] = t;
      }
    __syncthreads();
    uint16_t scale = F.inv(wm[r * cols + r]);
    __syncthreads();
    for (int c = tid; c < cols; c += blockDim.x) wm[r * cols + c] = F.mul(scale, wm[r * cols + c]);
    __syncthreads();
    // eliminate column r from every other row
    for (int idx = tid; idx < e * cols; idx += blockDim.x) {
      int rb = idx / cols, c = idx % cols;
      if (rb == r || c == r) continue;
      uint16_t f = wm[rb * cols + r];
      if (f) wm[idx] ^= F.mul(f, wm[r * cols + c]);
    }
    __syncthreads();
    for (int rb = tid; rb < e; rb += blockDim.x)
      if (rb != r) wm[rb * cols + r] = 0;
    __syncthreads();
  }
  if (singular) {
    if (tid == 0) *status = RSGPU_INVALID_CUSTOM_MATRIX;
    return;
  }
  // rows for each target
  for (int idx = tid; idx < job.n_targets * k; idx += blockDim.x) {
    int t = idx / k, col = idx % k;
    int target = job.targets[t];
    uint16_t v = 0;
    if (target < k) {
      int j = 0;
      while (j < e && missing[j] != target) ++j;
      v = wm[j * cols + e + col];  // [C^-1 B | C^-1] row j, column col
    } else {
      const uint16_t* prow = matrix + (size_t)target * k;
      if (col < p) v = prow[job.valid[col]];
      for (int j = 0; j < e; ++j) v ^= F.mul(prow[missing[j]], wm[j * cols + e + col]);
    }
    rows[idx] = v;
  }
  __syncthreads();
  const int W = job.tower ? 4 * kTabWords : kTabWords;
  for (int idx = tid; idx < job.n_targets * k; idx += blockDim.x) {
    uint16_t v = rows[idx];
    uint32_t* dst = tables_out + (size_t)idx * W;
    if (job.tower) {
      rows_out[2 * idx] = (uint8_t)(v >> 8);
      rows_out[2 * idx + 1] = (uint8_t)v;
      uint8_t cx = v >> 8, cc = v & 0xff;
      uint8_t co[4] = {(uint8_t)(cc ^ d_mul8(lg, ex, 2, cx)), cx, d_mul8(lg, ex, 128, cx), cc};
      for (int q = 0; q < 4; ++q) make_perm_table(co[q], dst + q * kTabWords);
    } else {
      rows_out[idx] = (uint8_t)v;
      make_perm_table((uint8_t)v, dst);
    }
  }
  if (tid == 0) *status = RSGPU_OK;
}

hipError_t launch_decode_rows(const uint16_t* d_matrix, const DecodeJob& job, uint8_t* d_rows,
                              uint32_t* d_tables, int32_t* d_status, hipStream_t s) {
  int p = 0;
  while (p < job.k && job.valid[p] < job.k) ++p;
  int e = job.k - p;
  if (e > kMaxSolve) return hipErrorInvalidValue;
  size_t lds = ((size_t)e * (e + p + e) + (size_t)job.n_targets * job.k) * sizeof(uint16_t);
  hipLaunchKernelGGL(decode_rows_kernel, dim3(1), dim3(256), lds, s, d_matrix, job, d_rows,
                     d_tables, d_status);
  return hipGetLastError();
}

int words_per_coef(int tower) { return tower ? 4 * kTabWords : kTabWords; }

void build_tables_host(const uint16_t* rows, int r, int k, bool tower, std::vector<uint32_t>& out) {
  const int W = words_per_coef(tower);
  out.assign((size_t)r * k * W, 0);
  for (int i = 0; i < r * k; ++i) {
    uint32_t* dst = out.data() + (size_t)i * W;
    if (tower) {
      uint8_t co[4];
      tower_coefficients(rows[i], co);
      for (int q = 0; q < 4; ++q) make_perm_table(co[q], dst + q * kTabWords);
    } else {
      make_perm_table((uint8_t)rows[i], dst);
    }
  }
}

}  // namespace rsgpu
