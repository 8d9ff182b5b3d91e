/*
 * leopard8.c -- oracle restatement of the Leopard GF(2^8) FFT codec of
 * rustfs-erasure-codec 9.1.0 (src/core/leopard_gf8/, itself a restatement of
 * klauspost/reedsolomon leopard8.go).  Test infrastructure only.
 *
 * Work lanes are processed column by column over the whole shard: every
 * operation is byte-wise, so the reference's 32/128 KiB chunking
 * (driver.rs:18-52) does not change any output byte.
 */
#include <stdlib.h>
#include <string.h>

#include "rs_oracle.h"

#define ORDER8 256
#define MODULUS8 255

static uint8_t L_log[ORDER8], L_exp[ORDER8], L_skew[MODULUS8], L_walsh[ORDER8];
static int L_init = 0;

/* add_mod8 / sub_mod8 (ops.rs:11-24): mod-255 with end-around carry. */
static uint8_t add_mod8(uint8_t a, uint8_t b) {
  unsigned sum = (unsigned)a + b;
  return (uint8_t)(sum + (sum >> 8));
}
static uint8_t sub_mod8(uint8_t a, uint8_t b) {
  uint64_t dif = (uint64_t)a - (uint64_t)b; /* usize wrapping_sub */
  return (uint8_t)(dif + (dif >> 8));
}

/* mul_log8 (ops.rs:3-9) */
uint8_t ora_leo8_mul_log(uint8_t a, uint8_t log_b) {
  if (a == 0) return 0;
  return L_exp[add_mod8(L_log[a], log_b)];
}

/* fwht8 / fwht8_mtrunc (ops.rs:26-89): radix-4 layers over 256 points, blocks r < mtrunc. */
static void fwht2(uint8_t* a, uint8_t* b) {
  uint8_t s = add_mod8(*a, *b), d = sub_mod8(*a, *b);
  *a = s;
  *b = d;
}
static void fwht8_mtrunc(uint8_t* data, unsigned mtrunc) {
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= ORDER8) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      for (unsigned off = r; off < r + dist; ++off) {
        fwht2(&data[off], &data[off + dist]);
        fwht2(&data[off + 2 * dist], &data[off + 3 * dist]);
        fwht2(&data[off], &data[off + 2 * dist]);
        fwht2(&data[off + dist], &data[off + 3 * dist]);
      }
    }
    dist = dist4;
    dist4 <<= 2;
  }
}

/* init_luts8 / init_fft_skew8 (tables.rs:25-104); Cantor basis tables.rs:25. */
static void leo8_init(void) {
  if (L_init) return;
  static const uint8_t cantor[8] = {1, 214, 152, 146, 86, 200, 88, 230};
  unsigned state = 1;
  memset(L_exp, 0, sizeof L_exp);
  for (unsigned i = 0; i < MODULUS8; ++i) {
    L_exp[state] = (uint8_t)i;
    state <<= 1;
    if (state >= ORDER8) state ^= 0x11D;
  }
  L_exp[0] = MODULUS8;
  L_log[0] = 0;
  for (unsigned i = 0; i < 8; ++i) {
    unsigned width = 1u << i;
    for (unsigned j = 0; j < width; ++j) L_log[j + width] = L_log[j] ^ cantor[i];
  }
  for (unsigned i = 0; i < ORDER8; ++i) L_log[i] = L_exp[L_log[i]];
  for (unsigned i = 0; i < ORDER8; ++i) L_exp[L_log[i]] = (uint8_t)i;
  L_exp[MODULUS8] = L_exp[0];

  uint8_t temp[7];
  for (unsigned i = 1; i < 8; ++i) temp[i - 1] = (uint8_t)(1u << i);
  memset(L_skew, 0, sizeof L_skew);
  for (unsigned m = 0; m < 7; ++m) {
    unsigned step = 1u << (m + 1);
    L_skew[(1u << m) - 1] = 0;
    for (unsigned i = m; i < 7; ++i) {
      unsigned s = 1u << (i + 1);
      for (unsigned j = (1u << m) - 1; j < s; j += step) L_skew[j + s] = L_skew[j] ^ temp[i];
    }
    uint8_t gf_prod = ora_leo8_mul_log(temp[m], L_log[temp[m] ^ 1]);
    temp[m] = (uint8_t)(MODULUS8 - L_log[gf_prod]);
    for (unsigned i = m + 1; i < 7; ++i) {
      uint8_t sum = add_mod8(L_log[temp[i] ^ 1], temp[m]);
      temp[i] = ora_leo8_mul_log(temp[i], sum);
    }
  }
  for (unsigned i = 0; i < MODULUS8; ++i) L_skew[i] = L_log[L_skew[i]];
  for (unsigned i = 0; i < ORDER8; ++i) L_walsh[i] = L_log[i];
  L_walsh[0] = 0;
  fwht8_mtrunc(L_walsh, ORDER8);
  L_init = 1;
}

void ora_leo8_tables(uint8_t fft_skew[255], uint8_t log_walsh[256], uint8_t log_lut[256],
                     uint8_t exp_lut[256]) {
  leo8_init();
  memcpy(fft_skew, L_skew, 255);
  memcpy(log_walsh, L_walsh, 256);
  memcpy(log_lut, L_log, 256);
  memcpy(exp_lut, L_exp, 256);
}

/* ---- butterflies, ops.rs:445-586 (log_m == 255 takes the XOR-only shortcut) ---- */
static void xor_into(uint8_t* dst, const uint8_t* src, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] ^= src[i];
}
static void mul_xor_into(uint8_t* dst, const uint8_t* src, uint8_t log_m, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] ^= ora_leo8_mul_log(src[i], log_m);
}
/* fft_dit2: x ^= mul(y, log_m); y ^= x */
static void fft_dit2(uint8_t* x, uint8_t* y, uint8_t log_m, size_t n) {
  if (log_m != MODULUS8) mul_xor_into(x, y, log_m, n);
  xor_into(y, x, n);
}
/* ifft_dit2: y ^= x; x ^= mul(y, log_m) */
static void ifft_dit2(uint8_t* x, uint8_t* y, uint8_t log_m, size_t n) {
  xor_into(y, x, n);
  if (log_m != MODULUS8) mul_xor_into(x, y, log_m, n);
}
static void fft_dit4(uint8_t** w, size_t i, size_t dist, uint8_t m01, uint8_t m23, uint8_t m02,
                     size_t n) {
  fft_dit2(w[i], w[i + 2 * dist], m02, n);
  fft_dit2(w[i + dist], w[i + 3 * dist], m02, n);
  fft_dit2(w[i], w[i + dist], m01, n);
  fft_dit2(w[i + 2 * dist], w[i + 3 * dist], m23, n);
}
static void ifft_dit4(uint8_t** w, size_t i, size_t dist, uint8_t m01, uint8_t m23, uint8_t m02,
                      size_t n) {
  ifft_dit2(w[i], w[i + dist], m01, n);
  ifft_dit2(w[i + 2 * dist], w[i + 3 * dist], m23, n);
  ifft_dit2(w[i + dist], w[i + 3 * dist], m02, n);
  ifft_dit2(w[i], w[i + 2 * dist], m02, n);
}

/* Encoder IFFT (build_ifft_dit8_plan mod.rs:376-447 + ifft_dit_encoder8_with_plan
 * encode.rs:778-880): skew is fft_skew offset by (m - 1 + group offset), so the
 * stage indices are skew[i_end], skew[i_end + dist], skew[i_end + 2 dist]. */
static void ifft_encoder(const uint8_t* const* data, unsigned mtrunc, uint8_t** work,
                         uint8_t** xor_dst, unsigned m, const uint8_t* skew, size_t n) {
  for (unsigned i = 0; i < mtrunc; ++i) memcpy(work[i], data[i], n);
  for (unsigned i = mtrunc; i < m; ++i) memset(work[i], 0, n);
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= m) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint8_t m01 = skew[iend], m02 = skew[iend + dist], m23 = skew[iend + 2 * dist];
      for (unsigned i = r; i < iend; ++i) ifft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist = dist4;
    dist4 <<= 2;
  }
  if (dist < m) {
    uint8_t log_m = skew[dist];
    for (unsigned i = 0; i < dist; ++i) ifft_dit2(work[i], work[i + dist], log_m, n);
  }
  if (xor_dst)
    for (unsigned i = 0; i < m; ++i) xor_into(xor_dst[i], work[i], n);
}

/* FFT (build_fft_dit8_plan mod.rs:334-374): full fft_skew, stage indices
 * skew[i_end - 1], skew[i_end + dist - 1], skew[i_end + 2 dist - 1]. */
static void fft_full(uint8_t** work, unsigned mtrunc, unsigned m, size_t n) {
  unsigned dist4 = m, dist = m >> 2;
  while (dist != 0) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint8_t m01 = L_skew[iend - 1], m02 = L_skew[iend + dist - 1], m23 = L_skew[iend + 2 * dist - 1];
      for (unsigned i = r; i < iend; ++i) fft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist4 = dist;
    dist >>= 2;
  }
  if (dist4 == 2)
    for (unsigned r = 0; r < mtrunc; r += 2) fft_dit2(work[r], work[r + 1], L_skew[r], n);
}

/* Decoder IFFT (build_ifft_decode_dit8_plan mod.rs:455-526 + decode.rs:431-496). */
static void ifft_decoder(uint8_t** work, unsigned mtrunc, unsigned m, size_t n) {
  unsigned dist = 1, dist4 = 4;
  while (dist4 <= m) {
    for (unsigned r = 0; r < mtrunc; r += dist4) {
      unsigned iend = r + dist;
      uint8_t m01 = L_skew[iend - 1], m02 = L_skew[iend + dist - 1], m23 = L_skew[iend + 2 * dist - 1];
      for (unsigned i = r; i < iend; ++i) ifft_dit4(work, i, dist, m01, m23, m02, n);
    }
    dist = dist4;
    dist4 <<= 2;
  }
  if (dist < m) {
    uint8_t log_m = L_skew[dist - 1];
    for (unsigned i = 0; i < dist; ++i) ifft_dit2(work[i], work[i + dist], log_m, n);
  }
}

static unsigned ceil_pow2(unsigned v) {
  unsigned p = 1;
  while (p < v) p <<= 1;
  return p;
}

/* encode_with_tables (encode.rs:194-372) with the plans of EncodePlans::new
 * (encode.rs:109-147) and the driver of driver.rs:18-52. */
int ora_leo8_encode(int k, int parity, const uint8_t* const* data, uint8_t* const* out,
                    size_t len) {
  leo8_init();
  if (len == 0 || len % 64) return ORA_INCORRECT_SHARD_SIZE;
  unsigned m = ceil_pow2((unsigned)(parity > 0 ? parity : 1));
  if (m > MODULUS8) return ORA_TOO_MANY_SHARDS;
  unsigned mtrunc = (unsigned)k < m ? (unsigned)k : m;
  unsigned last_count = (unsigned)k % m;
  uint8_t** work = (uint8_t**)malloc(2 * m * sizeof(uint8_t*));
  for (unsigned i = 0; i < 2 * m; ++i) work[i] = (uint8_t*)malloc(len);
  const uint8_t* skew = L_skew + (m - 1);
  ifft_encoder(data, mtrunc, work, NULL, m, skew, len);
  unsigned i = m;
  if (m < (unsigned)k) {
    for (; i + m <= (unsigned)k; i += m)
      ifft_encoder(data + i, m, work + m, work, m, skew + i, len);
    if (last_count != 0) ifft_encoder(data + i, last_count, work + m, work, m, skew + i, len);
  }
  fft_full(work, (unsigned)parity, m, len);
  for (int j = 0; j < parity; ++j) memcpy(out[j], work[j], len);
  for (unsigned j = 0; j < 2 * m; ++j) free(work[j]);
  free(work);
  return ORA_OK;
}

/* compute_error_locs (decode.rs:68-114). */
void ora_leo8_error_locs(int k, int parity, const uint8_t* present, uint8_t err_locs[256]) {
  leo8_init();
  unsigned m = ceil_pow2((unsigned)parity);
  unsigned work_size = m + (unsigned)k;
  memset(err_locs, 0, 256);
  for (int i = 0; i < parity; ++i) err_locs[i] = present[k + i] ? 0 : 1;
  for (unsigned i = (unsigned)parity; i < m; ++i) err_locs[i] = 1;
  for (int i = 0; i < k; ++i) err_locs[m + i] = present[i] ? 0 : 1;
  fwht8_mtrunc(err_locs, work_size);
  for (unsigned i = 0; i < ORDER8; ++i)
    if (err_locs[i]) err_locs[i] = (uint8_t)(((unsigned)err_locs[i] * L_walsh[i]) % MODULUS8);
  fwht8_mtrunc(err_locs, ORDER8);
}

/* reconstruct_with_tables / reconstruct_with_resources (decode.rs:290-429, 537-551). */
int ora_leo8_reconstruct(int k, int parity, uint8_t* const* shards, const uint8_t* present,
                         const uint8_t* required, size_t len) {
  leo8_init();
  int total = k + parity, npresent = 0, need = 0;
  for (int i = 0; i < total; ++i) {
    npresent += present[i] ? 1 : 0;
    if (!present[i] && required[i]) need = 1;
  }
  if (npresent < k) return ORA_TOO_FEW_SHARDS_PRESENT;
  if (!need) return ORA_OK;
  if (len == 0 || len % 64) return ORA_INCORRECT_SHARD_SIZE;
  unsigned m = ceil_pow2((unsigned)parity);
  unsigned work_size = m + (unsigned)k;
  if (work_size > ORDER8) return ORA_TOO_MANY_SHARDS;
  unsigned n = ceil_pow2(work_size);
  uint8_t err_locs[256];
  ora_leo8_error_locs(k, parity, present, err_locs);
  unsigned input_count = m;
  for (int i = 0; i < k; ++i)
    if (present[i]) input_count = m + (unsigned)i + 1;
  uint8_t** work = (uint8_t**)malloc(n * sizeof(uint8_t*));
  for (unsigned i = 0; i < n; ++i) work[i] = (uint8_t*)calloc(len, 1);
  for (int i = 0; i < k; ++i)
    if (present[i])
      for (size_t b = 0; b < len; ++b) work[m + i][b] = ora_leo8_mul_log(shards[i][b], err_locs[m + i]);
  for (int i = 0; i < parity; ++i)
    if (present[k + i])
      for (size_t b = 0; b < len; ++b) work[i][b] = ora_leo8_mul_log(shards[k + i][b], err_locs[i]);
  ifft_decoder(work, input_count, n, len);
  /* formal derivative: lower ^= higher (decode.rs:537-551) */
  for (unsigned i = 1; i < n; ++i) {
    unsigned width = ((i ^ (i - 1)) + 1) >> 1;
    for (unsigned j = 0; j < width; ++j)
      if (i + j < n) xor_into(work[i - width + j], work[i + j], len);
  }
  fft_full(work, work_size, n, len);
  for (int i = 0; i < total; ++i) {
    if (present[i] || !required[i]) continue;
    unsigned idx = i >= k ? (unsigned)(i - k) : m + (unsigned)i;
    uint8_t inv = (uint8_t)(MODULUS8 - err_locs[idx]);
    for (size_t b = 0; b < len; ++b) shards[i][b] = ora_leo8_mul_log(work[idx][b], inv);
  }
  for (unsigned i = 0; i < n; ++i) free(work[i]);
  free(work);
  return ORA_OK;
}
