/*
 * leopard16.c -- oracle restatement of the Leopard GF(2^16) FFT codec of
 * rustfs-erasure-codec 9.1.0 (src/core/leopard_gf16/, a restatement of
 * klauspost/reedsolomon leopard.go).  Test infrastructure only.
 *
 * Byte shards hold GF(2^16) elements in the split layout of ops.rs:520-568: in
 * each 64-byte block, element i (0..31) = byte[i] | byte[i + 32] << 8.
 */
#include <stdlib.h>
#include <string.h>

#include "rs_oracle.h"

#define ORDER16 65536u
#define MODULUS16 65535u

static uint16_t* H_log;  /* [65536] */
static uint16_t* H_exp;  /* [131072] */
static uint16_t* H_skew; /* [65535] */
static uint16_t* H_walsh;
static int H_init = 0;

/* add_mod16 / sub_mod16 (ops.rs:42-53) */
static uint16_t add_mod16(uint16_t a, uint16_t b) {
  uint32_t sum = (uint32_t)a + b;
  return (uint16_t)(sum + (sum >> 16));
}
static uint16_t sub_mod16(uint16_t a, uint16_t b) {
  uint32_t dif = (uint32_t)a - (uint32_t)b; /* u32 wrapping_sub */
  return (uint16_t)(dif + (dif >> 16));
}

/* mul_log16 (ops.rs:12-22) */
uint16_t ora_leo16_mul_log(uint16_t a, uint16_t log_b) {
  if (a == 0) return 0;
  return H_exp[add_mod16(H_log[a], log_b)];
}

/* fwht16_mtrunc (ops.rs:301-333) */
static void fwht2(uint16_t* a, uint16_t* b) {
  uint16_t s = add_mod16(*a, *b), d = sub_mod16(*a, *b);
  *a = s;
  *b = d;
}
static void fwht16_mtrunc(uint16_t* data, unsigned m, unsigned mtrunc) {
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= m) {
    for (unsigned r = 0; r < mtrunc; r += dist4)
      for (unsigned off = r; off < r + dist; ++off) {
        fwht2(&data[off], &data[off + dist]);
        fwht2(&data[off + 2 * dist], &data[off + 3 * dist]);
        fwht2(&data[off], &data[off + 2 * dist]);
        fwht2(&data[off + dist], &data[off + 3 * dist]);
      }
    dist = dist4;
    dist4 <<= 2;
  }
}

/* init_luts16 / init_fft_skew16 (tables.rs:22-116); Cantor basis tables.rs:24-27. */
static void leo16_init(void) {
  if (H_init) return;
  static const uint16_t cantor[16] = {0x0001, 0xACCA, 0x3C0E, 0x163E, 0xC582, 0xED2E,
                                      0x914C, 0x4012, 0x6C98, 0x10D8, 0x6A72, 0xB900,
                                      0xFDB8, 0xFB34, 0xFF38, 0x991E};
  H_log = (uint16_t*)calloc(ORDER16, 2);
  H_exp = (uint16_t*)calloc(2 * ORDER16, 2);
  H_skew = (uint16_t*)calloc(MODULUS16, 2);
  H_walsh = (uint16_t*)calloc(ORDER16, 2);
  uint32_t state = 1;
  for (uint32_t i = 0; i < MODULUS16; ++i) {
    H_exp[state] = (uint16_t)i;
    state <<= 1;
    if (state >= ORDER16) state ^= 0x1002Du;
  }
  H_exp[0] = (uint16_t)MODULUS16;
  H_log[0] = 0;
  for (unsigned i = 0; i < 16; ++i) {
    unsigned width = 1u << i;
    for (unsigned j = 0; j < width; ++j) H_log[j + width] = H_log[j] ^ cantor[i];
  }
  for (unsigned i = 0; i < ORDER16; ++i) H_log[i] = H_exp[H_log[i]];
  for (unsigned i = 0; i < ORDER16; ++i) H_exp[H_log[i]] = (uint16_t)i;
  H_exp[MODULUS16] = H_exp[0];

  uint16_t temp[15];
  for (unsigned i = 1; i < 16; ++i) temp[i - 1] = (uint16_t)(1u << i);
  for (unsigned m = 0; m < 15; ++m) {
    unsigned step = 1u << (m + 1);
    H_skew[(1u << m) - 1] = 0;
    for (unsigned i = m; i < 15; ++i) {
      unsigned s = 1u << (i + 1);
      for (unsigned j = (1u << m) - 1; j < s; j += step) H_skew[j + s] = H_skew[j] ^ temp[i];
    }
    uint16_t gf_prod = ora_leo16_mul_log(temp[m], H_log[temp[m] ^ 1]);
    temp[m] = (uint16_t)(MODULUS16 - H_log[gf_prod]);
    for (unsigned i = m + 1; i < 15; ++i) {
      uint16_t sum = add_mod16(H_log[temp[i] ^ 1], temp[m]);
      temp[i] = ora_leo16_mul_log(temp[i], sum);
    }
  }
  for (unsigned i = 0; i < MODULUS16; ++i) H_skew[i] = H_log[H_skew[i]];
  for (unsigned i = 0; i < ORDER16; ++i) H_walsh[i] = H_log[i];
  H_walsh[0] = 0;
  fwht16_mtrunc(H_walsh, ORDER16, ORDER16);
  H_init = 1;
}

void ora_leo16_tables(uint16_t* fft_skew, uint16_t* log_walsh, uint16_t* log_lut,
                      uint16_t* exp_lut) {
  leo16_init();
  memcpy(fft_skew, H_skew, MODULUS16 * 2);
  memcpy(log_walsh, H_walsh, ORDER16 * 2);
  memcpy(log_lut, H_log, ORDER16 * 2);
  memcpy(exp_lut, H_exp, 2 * ORDER16 * 2);
}

/* ---- lanes of u16 elements ---- */
static void to_work(const uint8_t* src, uint16_t* dst, size_t len_bytes) {
  for (size_t b = 0; b < len_bytes; b += 64)
    for (size_t i = 0; i < 32; ++i)
      dst[b / 2 + i] = (uint16_t)(src[b + i] | ((uint16_t)src[b + i + 32] << 8));
}
static void from_work(const uint16_t* src, uint8_t* dst, size_t len_bytes) {
  for (size_t b = 0; b < len_bytes; b += 64)
    for (size_t i = 0; i < 32; ++i) {
      dst[b + i] = (uint8_t)(src[b / 2 + i] & 0xff);
      dst[b + i + 32] = (uint8_t)(src[b / 2 + i] >> 8);
    }
}
static void xor_into(uint16_t* dst, const uint16_t* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] ^= src[i];
}
static void mul_xor_into(uint16_t* dst, const uint16_t* src, uint16_t log_m, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] ^= ora_leo16_mul_log(src[i], log_m);
}
/* dit2_step16 / dit2_step_inv16 (ops.rs:382-452), log_m == 65535 -> y ^= x only */
static void fft_dit2(uint16_t* x, uint16_t* y, uint16_t log_m, size_t n) {
  if (log_m != MODULUS16) mul_xor_into(x, y, log_m, n);
  xor_into(y, x, n);
}
static void ifft_dit2(uint16_t* x, uint16_t* y, uint16_t log_m, size_t n) {
  xor_into(y, x, n);
  if (log_m != MODULUS16) mul_xor_into(x, y, log_m, n);
}
/* fft_dit4_16 / ifft_dit4_16 (ops.rs:470-518) */
static void fft_dit4(uint16_t** w, size_t i, size_t d, uint16_t m01, uint16_t m23, uint16_t m02,
                     size_t n) {
  fft_dit2(w[i], w[i + 2 * d], m02, n);
  fft_dit2(w[i + d], w[i + 3 * d], m02, n);
  fft_dit2(w[i], w[i + d], m01, n);
  fft_dit2(w[i + 2 * d], w[i + 3 * d], m23, n);
}
static void ifft_dit4(uint16_t** w, size_t i, size_t d, uint16_t m01, uint16_t m23, uint16_t m02,
                      size_t n) {
  ifft_dit2(w[i], w[i + d], m01, n);
  ifft_dit2(w[i + 2 * d], w[i + 3 * d], m23, n);
  ifft_dit2(w[i + d], w[i + 3 * d], m02, n);
  ifft_dit2(w[i], w[i + 2 * d], m02, n);
}

/* encoder IFFT (mod.rs build_ifft_dit16_plan + encode.rs:284-371) */
static void ifft_encoder(const uint8_t* const* data, unsigned mtrunc, uint16_t** work,
                         uint16_t** xor_dst, unsigned m, const uint16_t* skew, size_t len_bytes) {
  size_t n = len_bytes / 2;
  for (unsigned i = 0; i < mtrunc; ++i) to_work(data[i], work[i], len_bytes);
  for (unsigned i = mtrunc; i < m; ++i) memset(work[i], 0, n * 2);
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= m) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint16_t m01 = skew[iend], m02 = skew[iend + dist], m23 = skew[iend + 2 * dist];
      for (unsigned i = r; i < iend; ++i) ifft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist = dist4;
    dist4 <<= 2;
  }
  if (dist < m) {
    uint16_t log_m = skew[dist];
    for (unsigned i = 0; i < dist; ++i) ifft_dit2(work[i], work[i + dist], log_m, n);
  }
  if (xor_dst)
    for (unsigned i = 0; i < m; ++i) xor_into(xor_dst[i], work[i], n);
}

/* build_fft_dit16_plan (mod.rs:99-140) executed over the whole lane */
static void fft_full(uint16_t** work, unsigned mtrunc, unsigned m, size_t n) {
  unsigned dist4 = m, dist = m >> 2;
  while (dist != 0) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint16_t m01 = H_skew[iend - 1], m02 = H_skew[iend + dist - 1], m23 = H_skew[iend + 2 * dist - 1];
      for (unsigned i = r; i < iend; ++i) fft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist4 = dist;
    dist >>= 2;
  }
  if (dist4 == 2)
    for (unsigned r = 0; r < mtrunc; r += 2) fft_dit2(work[r], work[r + 1], H_skew[r], n);
}

/* build_ifft_decode_dit16_plan (mod.rs:256-330) */
static void ifft_decoder(uint16_t** work, unsigned mtrunc, unsigned m, size_t n) {
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= m) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint16_t m01 = H_skew[iend - 1], m02 = H_skew[iend + dist - 1], m23 = H_skew[iend + 2 * dist - 1];
      for (unsigned i = r; i < iend; ++i) ifft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist = dist4;
    dist4 <<= 2;
  }
  if (dist < m) {
    uint16_t log_m = H_skew[dist - 1];
    for (unsigned i = 0; i < dist; ++i) ifft_dit2(work[i], work[i + dist], log_m, n);
  }
}

static unsigned ceil_pow2(unsigned v) {
  unsigned p = 1;
  while (p < v) p <<= 1;
  return p;
}

/* encode_with_tables16 (encode.rs:109-253), driver.rs, EncodePlans (encode.rs:39-82) */
int ora_leo16_encode(int k, int parity, const uint8_t* const* data, uint8_t* const* out,
                     size_t len) {
  leo16_init();
  if (len == 0 || len % 64) return ORA_INCORRECT_SHARD_SIZE;
  unsigned m = ceil_pow2((unsigned)(parity > 0 ? parity : 1));
  if (m > MODULUS16) return ORA_TOO_MANY_SHARDS;
  unsigned mtrunc = (unsigned)k < m ? (unsigned)k : m;
  unsigned last_count = (unsigned)k % m;
  size_t n = len / 2;
  uint16_t** work = (uint16_t**)malloc(2 * m * sizeof(uint16_t*));
  for (unsigned i = 0; i < 2 * m; ++i) work[i] = (uint16_t*)malloc(n * 2);
  const uint16_t* skew = H_skew + (m - 1);
  ifft_encoder(data, mtrunc, work, NULL, m, skew, len);
  unsigned i = m;
  if (m < (unsigned)k) {
    for (; i + m <= (unsigned)k; i += m) ifft_encoder(data + i, m, work + m, work, m, skew + i, len);
    if (last_count != 0) ifft_encoder(data + i, last_count, work + m, work, m, skew + i, len);
  }
  fft_full(work, (unsigned)parity, m, n);
  for (int j = 0; j < parity; ++j) from_work(work[j], out[j], len);
  for (unsigned j = 0; j < 2 * m; ++j) free(work[j]);
  free(work);
  return ORA_OK;
}

/* compute_error_locs16 (decode.rs:177-240), dense form */
void ora_leo16_error_locs(int k, int parity, const uint8_t* present, uint16_t* err_locs) {
  leo16_init();
  unsigned m = ceil_pow2((unsigned)parity);
  unsigned work_size = m + (unsigned)k;
  memset(err_locs, 0, ORDER16 * 2);
  for (int i = 0; i < parity; ++i)
    if (!present[k + i]) err_locs[i] = 1;
  for (unsigned i = (unsigned)parity; i < m; ++i) err_locs[i] = 1;
  for (int i = 0; i < k; ++i)
    if (!present[i]) err_locs[m + i] = 1;
  fwht16_mtrunc(err_locs, ORDER16, work_size);
  for (unsigned i = 0; i < ORDER16; ++i)
    if (err_locs[i])
      err_locs[i] = (uint16_t)(((uint64_t)err_locs[i] * H_walsh[i]) % MODULUS16);
  fwht16_mtrunc(err_locs, ORDER16, ORDER16);
}

/* reconstruct_with_tables16_with_options + compute_reconstruct_chunk
 * (decode.rs:704-905), formal derivative compute_formal_derivative16. */
int ora_leo16_reconstruct(int k, int parity, uint8_t* const* shards, const uint8_t* present,
                          const uint8_t* required, size_t len) {
  leo16_init();
  int total = k + parity, npresent = 0, need = 0, missing = 0;
  for (int i = 0; i < total; ++i) {
    npresent += present[i] ? 1 : 0;
    missing += present[i] ? 0 : 1;
    if (!present[i] && required[i]) need = 1;
  }
  if (npresent < k) return ORA_TOO_FEW_SHARDS_PRESENT;
  if (!need) return ORA_OK;
  if (len == 0 || len % 64) return ORA_INCORRECT_SHARD_SIZE;
  unsigned m = ceil_pow2((unsigned)parity);
  unsigned work_size = m + (unsigned)k;
  if (work_size > ORDER16) return ORA_TOO_MANY_SHARDS;
  if (missing > parity) return ORA_TOO_FEW_SHARDS_PRESENT;
  unsigned n = ceil_pow2(work_size);
  size_t ne = len / 2;
  uint16_t* err_locs = (uint16_t*)malloc(ORDER16 * 2);
  ora_leo16_error_locs(k, parity, present, err_locs);
  unsigned input_count = m;
  for (int i = 0; i < k; ++i)
    if (present[i]) input_count = m + (unsigned)i + 1;
  uint16_t** work = (uint16_t**)malloc(n * sizeof(uint16_t*));
  for (unsigned i = 0; i < n; ++i) work[i] = (uint16_t*)calloc(ne, 2);
  uint16_t* tmp = (uint16_t*)malloc(ne * 2);
  for (int i = 0; i < k; ++i)
    if (present[i]) {
      to_work(shards[i], tmp, len);
      for (size_t e = 0; e < ne; ++e) work[m + i][e] = ora_leo16_mul_log(tmp[e], err_locs[m + i]);
    }
  for (int i = 0; i < parity; ++i)
    if (present[k + i]) {
      to_work(shards[k + i], tmp, len);
      for (size_t e = 0; e < ne; ++e) work[i][e] = ora_leo16_mul_log(tmp[e], err_locs[i]);
    }
  ifft_decoder(work, input_count, n, ne);
  for (unsigned i = 1; i < n; ++i) {
    unsigned width = ((i ^ (i - 1)) + 1) >> 1;
    for (unsigned j = 0; j < width; ++j)
      if (i + j < n) xor_into(work[i - width + j], work[i + j], ne);
  }
  fft_full(work, work_size, n, ne);
  for (int i = 0; i < total; ++i) {
    if (present[i] || !required[i]) continue;
    unsigned idx = i >= k ? (unsigned)(i - k) : m + (unsigned)i;
    uint16_t inv = (uint16_t)(MODULUS16 - err_locs[idx]);
    for (size_t e = 0; e < ne; ++e) tmp[e] = ora_leo16_mul_log(work[idx][e], inv);
    from_work(tmp, shards[i], len);
  }
  for (unsigned i = 0; i < n; ++i) free(work[i]);
  free(work);
  free(tmp);
  free(err_locs);
  return ORA_OK;
}
