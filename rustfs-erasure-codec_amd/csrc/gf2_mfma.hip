// gf2_mfma.hip -- classic GF(2^8) encode as a GF(2) matrix product on the int8
// matrix cores (evaluation path beside the v_perm kernels).
//
// Multiplication by a constant c is linear over GF(2): bit p of c*x is the dot
// product (mod 2) of row p of an 8x8 bit matrix with the bits of x.  A stripe
// encode is therefore one (8m x 8k) bit matrix times the (8k x columns) bit
// expansion of the data, reduced mod 2.  Here the bit matrix is the A operand
// of v_mfma_i32_32x32x32_i8 (0/1 bytes, resident in registers), the data bits
// are expanded to 0/1 bytes on the fly as the B operand, and the int32
// accumulator's lowest bit is the exact GF(2) sum.  Semantics are those of
// encode_sep (src/core/encode.rs:935-998) for classic GF(2^8) codecs.
//
// Operand layout of the 32x32x32 int8 MFMA (64 lanes): lane l = i + 32*h holds
// row i of A (column i of B) and 16 K values; the K index of byte b of half h
// is the same in A and B (input 4*kk + 2*h + b/8, bit b%8), which is all the
// dot product needs.  Accumulator item 4*(i/8) + i%4 of lane j + 32*((i/4)%2)
// holds row i, column j.
//
// A wave owns 128 columns: lane j takes columns 4j..4j+3 (one dword per input)
// as four MFMA column slots u, so every input dword feeds four products and no
// LDS transpose is needed.  The four slots' sums (each <= 8k <= 248) sit in the
// four bytes of one dword without overlap, so one mask takes the parity bits of
// four columns at once.
#include <cstring>
#include <vector>

#include "gf.h"
#include "internal.h"

namespace rsgpu {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kGf2MaxK = 31;  // 8k <= 248: four column sums fit the bytes of a dword
constexpr int kGf2MaxM = 8;   // two 32-row tiles of output bits

struct Gf2Args {
  const uint8_t* in[kGf2MaxK];
  uint8_t* out[kGf2MaxM];
  const v4i* abits;  // [MT][KS][64 lanes] A operands
  uint64_t len;      // bytes per shard, a multiple of 4
  int k, m;
};

// The 4 low bits of v spread to the 4 bytes of a dword (bit q -> byte q):
// n * 0x00204081 places copies of n at bit offsets 0, 7, 14 and 21, so bit q of
// the copy at 7q lands on bit 8q, and the copies do not overlap.
__device__ __forceinline__ int spread4(uint32_t v) {
  return (int)(((v & 0xFu) * 0x00204081u) & 0x01010101u);
}

template <int MT, int KS>
__global__ __launch_bounds__(256) void gf2_mfma_encode_kernel(const Gf2Args a) {
  const int lane = threadIdx.x & 63, j = lane & 31, h = lane >> 5;
  v4i A[MT][KS];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) A[t][kk] = a.abits[(t * KS + kk) * 64 + lane];
  const uint64_t n_blocks = (a.len + 127) / 128;
  const uint64_t n_waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t blk = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < n_blocks;
       blk += n_waves) {
    const uint64_t col = blk * 128 + 4 * (uint64_t)j;
    const bool live = col < a.len;  // len % 4 == 0: a dword is wholly in or out
    uint32_t x[KS][2];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int input = 4 * kk + 2 * h + e;
        const uint8_t* p = h ? a.in[(4 * kk + 2 + e) < kGf2MaxK ? 4 * kk + 2 + e : 0]
                             : a.in[4 * kk + e];
        x[kk][e] = (live && input < a.k) ? *reinterpret_cast<const uint32_t*>(p + col) : 0u;
      }
    v16i acc[MT][4];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][u][r] = 0;
#pragma unroll
    for (int kk = 0; kk < KS; ++kk)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v4i B;
        B[0] = spread4(x[kk][0] >> (8 * u));
        B[1] = spread4(x[kk][0] >> (8 * u + 4));
        B[2] = spread4(x[kk][1] >> (8 * u));
        B[3] = spread4(x[kk][1] >> (8 * u + 4));
#pragma unroll
        for (int t = 0; t < MT; ++t)
          acc[t][u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[t][kk], B, acc[t][u], 0, 0, 0);
      }
    // Item 4*r' + q of lane (j, h) is bit 4h + q of parity 4t + r' at columns
    // col .. col + 3 (slots u = 0..3).
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int r = 4 * t + rr;
        if (r >= a.m) continue;  // uniform
        uint32_t y = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int it = 4 * rr + q;
          const uint32_t xq = (uint32_t)acc[t][0][it] | ((uint32_t)acc[t][1][it] << 8) |
                              ((uint32_t)acc[t][2][it] << 16) | ((uint32_t)acc[t][3][it] << 24);
          y |= (xq & 0x01010101u) << q;
        }
        y <<= 4 * h;
        y |= (uint32_t)__shfl_xor((int)y, 32);
        if (live && (r & 1) == h) *reinterpret_cast<uint32_t*>(a.out[r] + col) = y;
      }
  }
}

template <int MT>
static void launch_mt(int ks, unsigned grid, const Gf2Args& a, hipStream_t s) {
  switch (ks) {
#define RSGPU_GF2_KS(N) \
  case N: hipLaunchKernelGGL((gf2_mfma_encode_kernel<MT, N>), dim3(grid), dim3(256), 0, s, a); break;
    RSGPU_GF2_KS(1) RSGPU_GF2_KS(2) RSGPU_GF2_KS(3) RSGPU_GF2_KS(4)
    RSGPU_GF2_KS(5) RSGPU_GF2_KS(6) RSGPU_GF2_KS(7) RSGPU_GF2_KS(8)
#undef RSGPU_GF2_KS
  }
}

// A operands for the parity rows of a classic GF(2^8) codec: byte b of lane
// i + 32h in tile t, K step kk = bit p of coef(r, input) * x^q, where row
// 32t + i = 8r + p and (input, q) = (4kk + 2h + b/8, b%8); zero outside k x m.
static std::vector<uint8_t> gf2_operands(const Codec& c, int mt, int ks) {
  std::vector<uint8_t> bits((size_t)mt * ks * 64 * 16, 0);
  for (int t = 0; t < mt; ++t)
    for (int kk = 0; kk < ks; ++kk)
      for (int lane = 0; lane < 64; ++lane)
        for (int b = 0; b < 16; ++b) {
          const int row = 32 * t + (lane & 31), r = row / 8, p = row % 8;
          const int input = 4 * kk + 2 * (lane >> 5) + b / 8, q = b % 8;
          if (r >= (int)c.m || input >= (int)c.k) continue;
          const uint8_t coef = (uint8_t)c.matrix[(c.k + r) * c.k + input];
          bits[(((size_t)t * ks + kk) * 64 + lane) * 16 + b] = (gf8_mul(coef, (uint8_t)(1u << q)) >> p) & 1;
        }
  return bits;
}

}  // namespace rsgpu

using namespace rsgpu;

extern "C" int rsgpu_encode_sep_gf2mfma(const rsgpu_codec* h, const rsgpu_cslice* data, size_t nd,
                                        const rsgpu_slice* parity, size_t np, void* stream) {
  if (!h || (!data && nd) || (!parity && np)) return RSGPU_ERR_INVALID_ARGUMENT;
  const Codec& c = h->c;
  if (nd < c.k || nd == 0) return RSGPU_TOO_FEW_DATA_SHARDS;
  if (nd > c.k) return RSGPU_TOO_MANY_DATA_SHARDS;
  if (np < c.m || np == 0) return RSGPU_TOO_FEW_PARITY_SHARDS;
  if (np > c.m) return RSGPU_TOO_MANY_PARITY_SHARDS;
  const size_t len = data[0].len;
  if (len == 0) return RSGPU_EMPTY_SHARD;
  for (size_t i = 0; i < nd; ++i)
    if (data[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  for (size_t i = 0; i < np; ++i)
    if (parity[i].len != len) return RSGPU_INCORRECT_SHARD_SIZE;
  if (c.family != RSGPU_FAMILY_CLASSIC || c.field != RSGPU_FIELD_GF8 || c.ftab)
    return RSGPU_UNSUPPORTED_CODEC_FAMILY;
  // limits of this formulation: k <= 31, m <= 8, dword-aligned shards
  if (c.k > (size_t)kGf2MaxK || c.m > (size_t)kGf2MaxM || (len & 3)) return RSGPU_ERR_INVALID_ARGUMENT;
  Gf2Args a{};
  for (size_t i = 0; i < nd; ++i) {
    if (!data[i].ptr || ((uintptr_t)data[i].ptr & 3)) return RSGPU_ERR_INVALID_ARGUMENT;
    a.in[i] = data[i].ptr;
  }
  for (size_t i = 0; i < np; ++i) {
    if (!parity[i].ptr || ((uintptr_t)parity[i].ptr & 3)) return RSGPU_ERR_INVALID_ARGUMENT;
    a.out[i] = parity[i].ptr;
  }
  const int ks = (int)(c.k + 3) / 4, mt = (int)(c.m + 3) / 4;
  {
    std::lock_guard<std::mutex> g(c.mu);
    if (!c.d_gf2_bits.p) {
      const std::vector<uint8_t> bits = gf2_operands(c, mt, ks);
      HIP_TRY(c.d_gf2_bits.alloc(bits.size()));
      HIP_TRY(hipMemcpy(c.d_gf2_bits.p, bits.data(), bits.size(), hipMemcpyHostToDevice));
    }
  }
  a.abits = (const v4i*)c.d_gf2_bits.p;
  a.len = len;
  a.k = (int)c.k;
  a.m = (int)c.m;
  // One workgroup per 4 blocks of 128 columns (a persistent grid with the next
  // block's loads in flight measured slower: the kernel is VALU-bound).
  const uint64_t blocks = (len + 511) / 512;
  const uint64_t max_grid = (UINT64_C(1) << 32) / 256 - 1;
  const unsigned grid = (unsigned)(blocks < max_grid ? blocks : max_grid);
  hipStream_t s = (hipStream_t)stream;
  if (mt == 1) launch_mt<1>(ks, grid, a, s);
  else launch_mt<2>(ks, grid, a, s);
  HIP_TRY(hipGetLastError());
  return RSGPU_OK;
}
