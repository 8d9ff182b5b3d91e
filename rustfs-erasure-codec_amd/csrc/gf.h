// gf.h -- GF(2^8) (poly 0x11d) and GF((2^8)^2) tower arithmetic shared by host
// and device code, plus the byte-permute lookup tables the matrix-apply kernels
// consume.
//
// Field definition follows rustfs-erasure-codec build.rs:8-66 (generator
// polynomial 29 = 0x11d with the implicit x^8) and src/galois_16.rs:9-11
// (tower X^2 = 2X + 128).  Tables are generated by constexpr code at compile
// time, so host and device see the same bytes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsgpu {

struct Gf8Tables {
  uint8_t log[256];
  uint8_t exp[512];
};

constexpr Gf8Tables make_gf8_tables() {
  Gf8Tables t{};
  unsigned b = 1;
  for (unsigned l = 0; l < 255; ++l) {
    t.log[b] = (uint8_t)l;
    t.exp[l] = (uint8_t)b;
    t.exp[l + 255] = (uint8_t)b;
    b <<= 1;
    if (b >= 256) b = (b - 256) ^ 29u;
  }
  t.exp[510] = t.exp[0];
  t.exp[511] = t.exp[1];
  return t;
}

inline constexpr Gf8Tables kGf8 = make_gf8_tables();

__host__ __device__ constexpr inline uint8_t gf8_mul_c(uint8_t a, uint8_t b) {
  return (a == 0 || b == 0) ? 0 : kGf8.exp[kGf8.log[a] + kGf8.log[b]];
}

// Host-side helpers (codec construction runs on the host once per codec).
inline uint8_t gf8_mul(uint8_t a, uint8_t b) { return gf8_mul_c(a, b); }
inline uint8_t gf8_div(uint8_t a, uint8_t b) {
  if (a == 0 || b == 0) return 0;
  int lr = (int)kGf8.log[a] - (int)kGf8.log[b];
  if (lr < 0) lr += 255;
  return kGf8.exp[lr];
}
inline uint8_t gf8_exp(uint8_t a, size_t n) {
  if (n == 0) return 1;
  if (a == 0) return 0;
  return kGf8.exp[((size_t)kGf8.log[a] * (n % 255)) % 255];
}

// Tower element [x, c] stored as (x << 8) | c.
inline uint16_t gf16_mul(uint16_t a, uint16_t b) {
  uint8_t ax = a >> 8, ac = a & 0xff, bx = b >> 8, bc = b & 0xff;
  uint8_t o0 = gf8_mul(ax, bx);
  uint8_t o1 = gf8_mul(ac, bx) ^ gf8_mul(ax, bc);
  uint8_t o2 = gf8_mul(ac, bc);
  if (o0) {
    o1 ^= gf8_mul(2, o0);
    o2 ^= gf8_mul(128, o0);
  }
  return (uint16_t)((o1 << 8) | o2);
}
inline uint16_t gf16_inverse(uint16_t a) {
  uint8_t ax = a >> 8, ac = a & 0xff;
  if (ax == 0) return ac == 0 ? 0 : gf8_div(1, ac);
  uint8_t conj = ac ^ gf8_mul(2, ax);
  uint8_t norm = gf8_mul(ac, conj) ^ gf8_mul(128, gf8_mul(ax, ax));
  uint8_t ni = gf8_div(1, norm);
  return (uint16_t)((gf8_mul(ax, ni) << 8) | gf8_mul(conj, ni));
}
inline uint16_t gf16_div(uint16_t a, uint16_t b) { return gf16_mul(a, gf16_inverse(b)); }
inline uint16_t gf16_exp(uint16_t a, size_t n) {
  if (n == 0) return 1;
  if (a == 0) return 0;
  n %= 65535;
  uint16_t r = 1, base = a;
  while (n) {
    if (n & 1) r = gf16_mul(r, base);
    base = gf16_mul(base, base);
    n >>= 1;
  }
  return r;
}

// Byte-permute multiply table for one GF(2^8) coefficient c.  The product c*x
// of a byte x splits over the bit fields x[2:0], x[5:3], x[7:6]; each field
// indexes an (at most) 8-entry byte table that one v_perm_b32 reads:
//   w[0] = c*{0,1,2,3}        w[1] = c*{4,5,6,7}
//   w[2] = c*{0,8,16,24}      w[3] = c*{32,40,48,56}
//   w[4] = c*{0,64,128,192}   w[5..7] = 0 (pads the entry to 32 bytes)
constexpr int kTabWords = 8;

// The split needs only that x -> c*x is GF(2)-linear, which holds for
// multiplication by a constant in any basis, so `mul` may be any field's
// byte multiplication (the classic 0x11d field, or Leopard's).
template <class Mul>
__host__ __device__ inline void make_perm_table_with(Mul mul, uint8_t c, uint32_t* w) {
  uint32_t v[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    v[0] |= (uint32_t)mul(c, (uint8_t)i) << (8 * i);
    v[1] |= (uint32_t)mul(c, (uint8_t)(i + 4)) << (8 * i);
    v[2] |= (uint32_t)mul(c, (uint8_t)(8 * i)) << (8 * i);
    v[3] |= (uint32_t)mul(c, (uint8_t)(8 * (i + 4))) << (8 * i);
    v[4] |= (uint32_t)mul(c, (uint8_t)(64 * i)) << (8 * i);
  }
  for (int i = 0; i < 5; ++i) w[i] = v[i];
  w[5] = w[6] = w[7] = 0;
}

struct ClassicMul8 {
  __host__ __device__ uint8_t operator()(uint8_t a, uint8_t b) const { return gf8_mul_c(a, b); }
};

__host__ __device__ inline void make_perm_table(uint8_t c, uint32_t* w) {
  make_perm_table_with(ClassicMul8{}, c, w);
}

// A GF(2^8) given by log/exp tables (exp repeated over 512 entries so that
// exp[log a + log b] needs no reduction) and its multiplicative identity.
struct FieldTables {
  uint8_t log[256];
  uint8_t exp[512];
  uint8_t one;
  uint8_t mul(uint8_t a, uint8_t b) const { return (a == 0 || b == 0) ? 0 : exp[log[a] + log[b]]; }
};

// Tower multiply by e = [cx, cc] as a 2x2 GF(2^8) map on the byte planes
// (src/galois_16.rs:100-119, tower_mul_coefficients :224-231):
//   out_x = (cc ^ 2 cx) * ax ^ cx * ac,   out_c = 128 cx * ax ^ cc * ac
inline void tower_coefficients(uint16_t e, uint8_t out[4]) {
  uint8_t cx = e >> 8, cc = e & 0xff;
  out[0] = cc ^ gf8_mul(2, cx);  // x <- x
  out[1] = cx;                   // x <- c
  out[2] = gf8_mul(128, cx);     // c <- x
  out[3] = cc;                   // c <- c
}

}  // namespace rsgpu
