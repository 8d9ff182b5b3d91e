// classic_kernels.hip -- gfx950 kernels of the classic (matrix) Reed-Solomon
// data path: the GF(2^8) instantiations of the tiled r x k matrix apply
// (apply_kernel.h) that serves encode, verify and reconstruct, the byte kernel
// for tails and unaligned views, and the device-side solve that builds decode
// rows.
#include "kernel_common.h"

namespace rsgpu {

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
    if (wave_leader(mismatch)) raise_flag(a.flag);
  }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

#define RSGPU_DECL(t, w) \
  hipError_t launch_apply_t##t##_w##w(const ApplyArgs& a, int mode, int rows, hipStream_t s);
RSGPU_DECL(0, 0) RSGPU_DECL(1, 0) RSGPU_DECL(2, 0) RSGPU_DECL(3, 0) RSGPU_DECL(4, 0)
RSGPU_DECL(0, 1) RSGPU_DECL(1, 1) RSGPU_DECL(2, 1) RSGPU_DECL(3, 1) RSGPU_DECL(4, 1)
RSGPU_DECL(0, 2) RSGPU_DECL(1, 2) RSGPU_DECL(2, 2) RSGPU_DECL(3, 2) RSGPU_DECL(4, 2)
#undef RSGPU_DECL

hipError_t launch_apply_dword(const ApplyArgs& a, int mode, int rows, hipStream_t s);

using ApplyFn = hipError_t (*)(const ApplyArgs&, int, int, hipStream_t);
static const ApplyFn kApplyFns[3][kNumTiles] = {
    {launch_apply_t0_w0, launch_apply_t1_w0, launch_apply_t2_w0, launch_apply_t3_w0, launch_apply_t4_w0},
    {launch_apply_t0_w1, launch_apply_t1_w1, launch_apply_t2_w1, launch_apply_t3_w1, launch_apply_t4_w1},
    {launch_apply_t0_w2, launch_apply_t1_w2, launch_apply_t2_w2, launch_apply_t3_w2, launch_apply_t4_w2}};

hipError_t launch_apply(const ApplyArgs& a, int mode, int tile, int rows, int tower, hipStream_t s) {
  if (a.n16 == 0) return hipSuccess;
  const bool dword = tile == kTileDword && !tower;
  if ((!dword && (tile < 0 || tile >= kNumTiles)) || tower < 0 || tower > 2 || a.r < 1 || a.r > kMaxOut ||
      a.k < 1 || a.k > kMaxIn)
    return hipErrorInvalidValue;
  // the kernel stages its tables into LDS with 16-byte direct loads
  if (((uintptr_t)a.tables & 15) != 0 || (a.tab_row_stride & 3) != 0) return hipErrorInvalidValue;
  if (rows < 1 || rows > a.r) rows = a.r;
  if (dword) return launch_apply_dword(a, mode, rows, s);
  return kApplyFns[tower][tile](a, mode, rows, s);
}

hipError_t launch_apply_bytes(const ByteArgs& a, hipStream_t s) {
  uint64_t units = (a.end - a.begin) / (a.tower ? 2 : 1);
  if (units == 0) return hipSuccess;
  uint64_t blocks = (units + kBlock - 1) / kBlock;
  unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(apply_bytes_kernel, dim3(grid), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

// The fold of a split over inputs: out[row] (^)= XOR_z part[z][row], 16 bytes
// per lane per step, grid-stride over rows x len.
struct XorPartials {
  const uint8_t* part;
  uint64_t pitch, n16;
  uint8_t* out[kMaxOut];
  int32_t chunks, rows, accumulate;
};

__global__ __launch_bounds__(kBlock) void xor_partials_kernel(const XorPartials x) {
  const uint64_t total = x.n16 * (uint64_t)x.rows;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (uint64_t)gridDim.x * kBlock) {
    const int row = (int)(t / x.n16);
    const uint64_t off = (t % x.n16) * 16;
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (int z = 0; z < x.chunks; ++z) {
      const uint4 q = load16(x.part + ((size_t)z * x.rows + row) * x.pitch + off);
      acc.x ^= q.x, acc.y ^= q.y, acc.z ^= q.z, acc.w ^= q.w;
    }
    uint8_t* dst = x.out[row] + off;
    if (x.accumulate) {
      const uint4 o = load16(dst);
      acc.x ^= o.x, acc.y ^= o.y, acc.z ^= o.z, acc.w ^= o.w;
    }
    store16(dst, acc);
  }
}

hipError_t launch_xor_partials(const uint8_t* part, uint64_t pitch, int chunks, int rows,
                               uint8_t* const* out, uint64_t len, bool accumulate, hipStream_t s) {
  if (rows < 1 || rows > kMaxOut || chunks < 1 || len % 16) return hipErrorInvalidValue;
  XorPartials x{};
  x.part = part;
  x.pitch = pitch;
  x.n16 = len / 16;
  for (int r = 0; r < rows; ++r) x.out[r] = out[r];
  x.chunks = chunks;
  x.rows = rows;
  x.accumulate = accumulate ? 1 : 0;
  const uint64_t blocks = (x.n16 * rows + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(xor_partials_kernel, dim3((unsigned)std::min<uint64_t>(blocks, 8192)), dim3(kBlock), 0, s, x);
  return hipGetLastError();
}

__global__ void fill_u8_kernel(uint8_t* p, uint8_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(fill_u8_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, s, p, v, n);
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

// One decode solve by one workgroup: [C | B | I] Gauss-Jordan in `smem` ([e][cols]
// work matrix, then [n_targets][k] rows), then the rows' bytes and v_perm tables.
__device__ __forceinline__ void decode_rows_body(const uint16_t* __restrict__ matrix,
                                                 const DecodeJob& job, uint8_t* __restrict__ rows_out,
                                                 uint32_t* __restrict__ tables_out,
                                                 int32_t* __restrict__ status, uint16_t* smem) {
  __shared__ uint8_t lg[256];
  __shared__ uint8_t ex[512];
  const int tid = threadIdx.x;
  for (int i = tid; i < 256; i += blockDim.x) lg[i] = job.field ? job.field[i] : kGf8.log[i];
  for (int i = tid; i < 512; i += blockDim.x) ex[i] = job.field ? job.field[256 + i] : kGf8.exp[i];
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
    else v = (c - e - p == r) ? (uint16_t)job.one : 0;
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
  // rows for each target, over the selected survivor columns
  const int ncols = job.n_cols > 0 ? job.n_cols : k;
  for (int idx = tid; idx < job.n_targets * ncols; idx += blockDim.x) {
    int t = idx / ncols, q = idx % ncols;
    int col = job.n_cols > 0 ? job.cols[q] : q;
    int target = job.targets[t];
    uint16_t v = 0;
    if (target < k) {
      int j = 0;
      while (j < e && missing[j] != target) ++j;
      if (j < e) v = wm[j * cols + e + col];  // [C^-1 B | C^-1] row j, column col
      else v = job.valid[col] == target ? (uint16_t)job.one : (uint16_t)0;  // a surviving data shard
    } else {
      const uint16_t* prow = matrix + (size_t)target * k;
      if (col < p) v = prow[job.valid[col]];
      for (int j = 0; j < e; ++j) v ^= F.mul(prow[missing[j]], wm[j * cols + e + col]);
    }
    rows[idx] = v;
  }
  __syncthreads();
  const int W = job.tower ? 4 * kTabWords : kTabWords;
  for (int idx = tid; idx < job.n_targets * ncols; idx += blockDim.x) {
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
      make_perm_table_with([&](uint8_t a, uint8_t b) { return (uint8_t)F.mul(a, b); }, (uint8_t)v, dst);
    }
  }
  if (tid == 0) *status = RSGPU_OK;
}

__global__ __launch_bounds__(256) void decode_rows_kernel(const uint16_t* __restrict__ matrix,
                                                          const DecodeJob job,
                                                          uint8_t* __restrict__ rows_out,
                                                          uint32_t* __restrict__ tables_out,
                                                          int32_t* __restrict__ status) {
  extern __shared__ uint16_t smem[];
  decode_rows_body(matrix, job, rows_out, tables_out, status, smem);
}

// Many independent solves in one launch: workgroup b runs jobs[b].
__global__ __launch_bounds__(256) void decode_rows_multi_kernel(const uint16_t* __restrict__ matrix,
                                                                const DecodeMultiJob* __restrict__ jobs,
                                                                int32_t* __restrict__ status) {
  extern __shared__ uint16_t smem[];
  const DecodeMultiJob& J = jobs[blockIdx.x];
  decode_rows_body(matrix, J.job, J.rows, J.tables, status, smem);
}

size_t decode_rows_lds(const DecodeJob& job) {
  int p = 0;
  while (p < job.k && job.valid[p] < job.k) ++p;
  const size_t e = (size_t)(job.k - p);
  if (e > (size_t)kMaxSolve) return 0;
  const size_t ncols = job.n_cols > 0 ? (size_t)job.n_cols : (size_t)job.k;
  return (e * (e + p + e) + (size_t)job.n_targets * ncols) * sizeof(uint16_t);
}

hipError_t launch_decode_rows(const uint16_t* d_matrix, const DecodeJob& job, uint8_t* d_rows,
                              uint32_t* d_tables, int32_t* d_status, hipStream_t s) {
  const size_t lds = decode_rows_lds(job);
  if (!lds) return hipErrorInvalidValue;
  hipLaunchKernelGGL(decode_rows_kernel, dim3(1), dim3(256), lds, s, d_matrix, job, d_rows,
                     d_tables, d_status);
  return hipGetLastError();
}

hipError_t launch_decode_rows_multi(const uint16_t* d_matrix, const DecodeMultiJob* d_jobs, int n,
                                    size_t lds, int32_t* d_status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(decode_rows_multi_kernel, dim3((unsigned)n), dim3(256), lds, s, d_matrix, d_jobs,
                     d_status);
  return hipGetLastError();
}

// Writes coefficient v of output element idx: its byte(s) and v_perm tables.
__device__ __forceinline__ void put_coef(uint16_t v, size_t idx, bool tower, const DevField& F,
                                         const uint8_t* lg, const uint8_t* ex, uint8_t* rows_out,
                                         uint32_t* tables_out) {
  const int W = tower ? 4 * kTabWords : kTabWords;
  uint32_t* dst = tables_out + idx * W;
  if (tower) {
    rows_out[2 * idx] = (uint8_t)(v >> 8);
    rows_out[2 * idx + 1] = (uint8_t)v;
    uint8_t cx = v >> 8, cc = v & 0xff;
    uint8_t co[4] = {(uint8_t)(cc ^ d_mul8(lg, ex, 2, cx)), cx, d_mul8(lg, ex, 128, cx), cc};
    for (int q = 0; q < 4; ++q) make_perm_table(co[q], dst + q * kTabWords);
  } else {
    rows_out[idx] = (uint8_t)v;
    make_perm_table_with([&](uint8_t a, uint8_t b) { return (uint8_t)F.mul(a, b); }, (uint8_t)v, dst);
  }
}

// C^-1 of C[r][c] = matrix[par_surv[r]][missing[c]] (e x e) by Gauss-Jordan on
// [C | I] in LDS, written row-major to cinv.
__global__ __launch_bounds__(256) void decode_inv_kernel(const uint16_t* __restrict__ matrix,
                                                         const DecodeBigJob job,
                                                         uint16_t* __restrict__ cinv,
                                                         int32_t* __restrict__ status) {
  __shared__ uint8_t lg[256];
  __shared__ uint8_t ex[512];
  __shared__ int pivot_row;
  extern __shared__ uint16_t wm[];  // [e][2e]
  const int tid = threadIdx.x, e = job.e, cols = 2 * e;
  for (int i = tid; i < 256; i += blockDim.x) lg[i] = job.field ? job.field[i] : kGf8.log[i];
  for (int i = tid; i < 512; i += blockDim.x) ex[i] = job.field ? job.field[256 + i] : kGf8.exp[i];
  for (int idx = tid; idx < e * cols; idx += blockDim.x) {
    const int r = idx / cols, c = idx % cols;
    wm[idx] = c < e ? matrix[(size_t)job.par_surv[r] * job.k + job.missing[c]]
                    : (c - e == r ? (uint16_t)job.one : (uint16_t)0);
  }
  __syncthreads();
  DevField F{lg, ex, job.tower != 0};
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
    const int pr = pivot_row;
    if (pr < 0) {  // uniform: every thread leaves
      if (tid == 0) *status = RSGPU_INVALID_CUSTOM_MATRIX;
      return;
    }
    if (pr != r)
      for (int c = tid; c < cols; c += blockDim.x) {
        const uint16_t t = wm[r * cols + c];
        wm[r * cols + c] = wm[pr * cols + c];
        wm[pr * cols + c] = t;
      }
    __syncthreads();
    const uint16_t scale = F.inv(wm[r * cols + r]);
    __syncthreads();
    for (int c = tid; c < cols; c += blockDim.x) wm[r * cols + c] = F.mul(scale, wm[r * cols + c]);
    __syncthreads();
    for (int idx = tid; idx < e * cols; idx += blockDim.x) {
      const int rb = idx / cols, c = idx % cols;
      if (rb == r || c == r) continue;
      const uint16_t f = wm[rb * cols + r];
      if (f) wm[idx] ^= F.mul(f, wm[r * cols + c]);
    }
    __syncthreads();
    for (int rb = tid; rb < e; rb += blockDim.x)
      if (rb != r) wm[rb * cols + r] = 0;
    __syncthreads();
  }
  for (int idx = tid; idx < e * e; idx += blockDim.x) cinv[idx] = wm[(idx / e) * cols + e + idx % e];
  if (tid == 0) *status = RSGPU_OK;
}

// The same elimination with the [C | I] work matrix in device memory, for more
// missing data rows than LDS holds.  One workgroup; the pivot row is the first
// nonzero one at or below r (as in the LDS version), found in parallel.
__global__ __launch_bounds__(1024) void decode_inv_global_kernel(const uint16_t* __restrict__ matrix,
                                                                 const DecodeBigJob job, uint16_t* wm,
                                                                 uint16_t* __restrict__ cinv,
                                                                 int32_t* __restrict__ status) {
  __shared__ uint8_t lg[256];
  __shared__ uint8_t ex[512];
  __shared__ int pivot_row;
  const int tid = threadIdx.x, e = job.e;
  const size_t cols = 2 * (size_t)e;
  for (int i = tid; i < 256; i += blockDim.x) lg[i] = job.field ? job.field[i] : kGf8.log[i];
  for (int i = tid; i < 512; i += blockDim.x) ex[i] = job.field ? job.field[256 + i] : kGf8.exp[i];
  for (size_t idx = tid; idx < (size_t)e * cols; idx += blockDim.x) {
    const size_t r = idx / cols, c = idx % cols;
    wm[idx] = c < (size_t)e ? matrix[(size_t)job.par_surv[r] * job.k + job.missing[c]]
                            : (c - e == r ? (uint16_t)job.one : (uint16_t)0);
  }
  __syncthreads();
  DevField F{lg, ex, job.tower != 0};
  for (int r = 0; r < e; ++r) {
    if (tid == 0) pivot_row = e;
    __syncthreads();
    for (int rb = r + tid; rb < e; rb += blockDim.x)
      if (wm[(size_t)rb * cols + r] != 0) atomicMin(&pivot_row, rb);
    __syncthreads();
    const int pr = pivot_row;
    if (pr >= e) {  // uniform: every thread leaves
      if (tid == 0) *status = RSGPU_INVALID_CUSTOM_MATRIX;
      return;
    }
    if (pr != r)
      for (size_t c = tid; c < cols; c += blockDim.x) {
        const uint16_t t = wm[(size_t)r * cols + c];
        wm[(size_t)r * cols + c] = wm[(size_t)pr * cols + c];
        wm[(size_t)pr * cols + c] = t;
      }
    __syncthreads();
    const uint16_t scale = F.inv(wm[(size_t)r * cols + r]);
    __syncthreads();
    for (size_t c = tid; c < cols; c += blockDim.x) wm[(size_t)r * cols + c] = F.mul(scale, wm[(size_t)r * cols + c]);
    __syncthreads();
    for (size_t idx = tid; idx < (size_t)e * cols; idx += blockDim.x) {
      const size_t rb = idx / cols, c = idx % cols;
      if (rb == (size_t)r || c == (size_t)r) continue;
      const uint16_t f = wm[rb * cols + r];
      if (f) wm[idx] ^= F.mul(f, wm[(size_t)r * cols + c]);
    }
    __syncthreads();
    for (int rb = tid; rb < e; rb += blockDim.x)
      if (rb != r) wm[(size_t)rb * cols + r] = 0;
    __syncthreads();
  }
  for (size_t idx = tid; idx < (size_t)e * e; idx += blockDim.x)
    cinv[idx] = wm[(idx / e) * cols + e + idx % e];
  if (tid == 0) *status = RSGPU_OK;
}

struct U16Chunk {
  int32_t n;
  uint16_t v[1024];
};

__global__ void store_u16_kernel(uint16_t* dst, const U16Chunk c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c.n) dst[i] = c.v[i];
}

hipError_t launch_store_u16(uint16_t* dst, const uint16_t* vals, size_t n, hipStream_t s) {
  for (size_t i0 = 0; i0 < n; i0 += 1024) {
    U16Chunk c{};
    c.n = (int32_t)std::min<size_t>(1024, n - i0);
    for (int i = 0; i < c.n; ++i) c.v[i] = vals[i0 + i];
    hipLaunchKernelGGL(store_u16_kernel, dim3((c.n + 255) / 256), dim3(256), 0, s, dst + i0, c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

struct U32Chunk {
  int32_t n;
  uint32_t v[512];
};

__global__ void store_u32_kernel(uint32_t* dst, const U32Chunk c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c.n) dst[i] = c.v[i];
}

hipError_t launch_store_u32(uint32_t* dst, const uint32_t* vals, size_t n, hipStream_t s) {
  for (size_t i0 = 0; i0 < n; i0 += 512) {
    U32Chunk c{};
    c.n = (int32_t)std::min<size_t>(512, n - i0);
    for (int i = 0; i < c.n; ++i) c.v[i] = vals[i0 + i];
    hipLaunchKernelGGL(store_u32_kernel, dim3((c.n + 255) / 256), dim3(256), 0, s, dst + i0, c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

struct DecodeTargets {
  int32_t t0, n;                     // first output row of this launch, rows in it
  uint16_t target[kBigTargetChunk];  // shard index of each row
};

// Row t0 + blockIdx.y, columns blockIdx.x * 256 + threadIdx.x.  With w = u C^-1
// (u = unit row j for missing data j, else the parity row restricted to the
// missing columns), the coefficient on present data d is base(d) + sum_i w[i] *
// matrix[par_surv[i]][d] (base = the parity row's own coefficient, 0 for data
// targets) and the coefficient on parity survivor i is w[i].
__global__ __launch_bounds__(256) void decode_rows_big_kernel(const uint16_t* __restrict__ matrix,
                                                              const DecodeBigJob job,
                                                              const DecodeTargets tg,
                                                              const uint16_t* __restrict__ cinv,
                                                              uint8_t* __restrict__ rows_out,
                                                              uint32_t* __restrict__ tables_out) {
  __shared__ uint8_t lg[256];
  __shared__ uint8_t ex[512];
  extern __shared__ uint16_t w[];  // e entries
  const int tid = threadIdx.x, k = job.k, e = job.e, p = k - e;
  for (int i = tid; i < 256; i += blockDim.x) lg[i] = job.field ? job.field[i] : kGf8.log[i];
  for (int i = tid; i < 512; i += blockDim.x) ex[i] = job.field ? job.field[256 + i] : kGf8.exp[i];
  __syncthreads();
  DevField F{lg, ex, job.tower != 0};
  const int target = tg.target[blockIdx.y];
  const uint16_t* prow = matrix + (size_t)target * k;
  if (target < k) {
    int j = 0;
    while (j < e && job.missing[j] != target) ++j;  // a data target is a missing shard
    for (int i = tid; i < e; i += blockDim.x) w[i] = j < e ? cinv[j * e + i] : (uint16_t)0;
  } else {
    for (int i = tid; i < e; i += blockDim.x) {
      uint16_t v = 0;
      for (int j = 0; j < e; ++j) v ^= F.mul(prow[job.missing[j]], cinv[j * e + i]);
      w[i] = v;
    }
  }
  __syncthreads();
  const int col = blockIdx.x * blockDim.x + tid;
  if (col >= k) return;
  uint16_t v;
  if (col < p) {
    const int d = job.present[col];  // the col-th present data index
    // a data target's own column: the unit row of a surviving data shard
    v = target < k ? (d == target ? (uint16_t)job.one : (uint16_t)0) : prow[d];
    for (int i = 0; i < e; ++i) v ^= F.mul(w[i], matrix[(size_t)job.par_surv[i] * k + d]);
  } else {
    v = w[col - p];
  }
  put_coef(v, (size_t)(tg.t0 + blockIdx.y) * k + col, job.tower != 0, F, lg, ex, rows_out, tables_out);
}

hipError_t launch_decode_big(const uint16_t* d_matrix, const DecodeBigJob& job,
                             const std::vector<int>& targets, uint16_t* d_cinv, uint16_t* d_work,
                             uint8_t* d_rows, uint32_t* d_tables, int32_t* d_status, hipStream_t s) {
  if (job.e > kMaxBigSolve || (job.e > kMaxSolve && !d_work)) return hipErrorInvalidValue;
  if (job.e > kMaxSolve) {
    hipLaunchKernelGGL(decode_inv_global_kernel, dim3(1), dim3(1024), 0, s, d_matrix, job, d_work, d_cinv,
                       d_status);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  } else if (job.e > 0) {
    const size_t lds = (size_t)job.e * 2 * job.e * sizeof(uint16_t);
    hipLaunchKernelGGL(decode_inv_kernel, dim3(1), dim3(256), lds, s, d_matrix, job, d_cinv, d_status);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  for (size_t t0 = 0; t0 < targets.size(); t0 += kBigTargetChunk) {
    DecodeTargets tg{};
    tg.t0 = (int32_t)t0;
    tg.n = (int32_t)std::min<size_t>(kBigTargetChunk, targets.size() - t0);
    for (int i = 0; i < tg.n; ++i) tg.target[i] = (uint16_t)targets[t0 + i];
    const dim3 grid((unsigned)((job.k + 255) / 256), (unsigned)tg.n);
    hipLaunchKernelGGL(decode_rows_big_kernel, grid, dim3(256), (size_t)std::max(job.e, 1) * sizeof(uint16_t), s,
                       d_matrix, job, tg, d_cinv, d_rows, d_tables);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

__global__ __launch_bounds__(kBlock) void compare_rows_kernel(const CompareArgs c) {
  const uint8_t* a = c.a[blockIdx.y];
  const uint8_t* b = c.b[blockIdx.y];
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < c.len;
       i += (uint64_t)gridDim.x * kBlock)
    bad |= a[i] != b[i];
  if (wave_leader(bad)) raise_flag(c.flag);
}

hipError_t launch_compare_rows(const CompareArgs& c, hipStream_t s) {
  if (c.rows <= 0 || c.len == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((c.len + kBlock - 1) / kBlock, 1024);
  hipLaunchKernelGGL(compare_rows_kernel, dim3((unsigned)blocks, (unsigned)c.rows), dim3(kBlock), 0, s, c);
  return hipGetLastError();
}

// Pair blockIdx.y: 16-byte vectors when both ends are aligned, then the byte tail.
__global__ __launch_bounds__(kBlock) void xor_pairs_kernel(const XorPairs p) {
  const int y = blockIdx.y;
  const uint8_t* src = p.src[y];
  uint8_t* dst = p.dst[y];
  const uint64_t len = p.len[y];
  const bool vec = (((uintptr_t)src | (uintptr_t)dst) & 15) == 0;
  const uint64_t n16 = vec ? len / 16 : 0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (uint64_t i = t; i < n16; i += stride) {
    uint4 a = reinterpret_cast<const uint4*>(src)[i];
    uint4 b = reinterpret_cast<uint4*>(dst)[i];
    b.x ^= a.x;
    b.y ^= a.y;
    b.z ^= a.z;
    b.w ^= a.w;
    reinterpret_cast<uint4*>(dst)[i] = b;
  }
  for (uint64_t i = n16 * 16 + t; i < len; i += stride) dst[i] ^= src[i];
}

hipError_t launch_xor_pairs(const XorPairs& p, hipStream_t s) {
  if (p.n <= 0 || p.n > kXorPairs) return p.n == 0 ? hipSuccess : hipErrorInvalidValue;
  uint64_t longest = 0;
  for (int i = 0; i < p.n; ++i) longest = std::max<uint64_t>(longest, p.len[i]);
  if (longest == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((longest / 16 + kBlock) / kBlock, 2048);
  hipLaunchKernelGGL(xor_pairs_kernel, dim3((unsigned)blocks, (unsigned)p.n), dim3(kBlock), 0, s, p);
  return hipGetLastError();
}

int words_per_coef(int tower) { return tower ? 4 * kTabWords : kTabWords; }

void build_tables_host(const uint16_t* rows, int r, int k, bool tower, std::vector<uint32_t>& out,
                       const FieldTables* field) {
  const int W = words_per_coef(tower);
  out.assign((size_t)r * k * W, 0);
  for (int i = 0; i < r * k; ++i) {
    uint32_t* dst = out.data() + (size_t)i * W;
    if (tower) {
      uint8_t co[4];
      tower_coefficients(rows[i], co);
      for (int q = 0; q < 4; ++q) make_perm_table(co[q], dst + q * kTabWords);
    } else {
      if (field)
        make_perm_table_with([field](uint8_t a, uint8_t b) { return field->mul(a, b); }, (uint8_t)rows[i], dst);
      else
        make_perm_table((uint8_t)rows[i], dst);
    }
  }
}

}  // namespace rsgpu
