/*
 * classic.c -- oracle restatement of the classic (matrix) Reed-Solomon codec of
 * rustfs-erasure-codec 9.1.0 over GF(2^8) and the GF((2^8)^2) tower field.
 * Test infrastructure only; see rs_oracle.h.
 */
#include <stdlib.h>
#include <string.h>

#include "rs_oracle.h"

/* ------------------------------------------------------------------ GF(2^8) */

static uint8_t g_log[256];
static uint8_t g_exp[510];
static uint8_t g_mul[256][256];
static int g_init = 0;

/* build.rs:10-24 gen_log_table with GENERATING_POLYNOMIAL = 29 (build.rs:8). */
static void gen_log_table(uint8_t out[256]) {
  memset(out, 0, 256);
  unsigned b = 1;
  for (unsigned log = 0; log < 255; ++log) {
    out[b] = (uint8_t)log;
    b <<= 1;
    if (b >= 256) b = (b - 256) ^ 29u;
  }
}

/* build.rs:28-39 gen_exp_table. */
static void gen_exp_table(const uint8_t log_table[256], uint8_t out[510]) {
  memset(out, 0, 510);
  for (unsigned i = 1; i < 256; ++i) {
    unsigned log = log_table[i];
    out[log] = (uint8_t)i;
    out[log + 255] = (uint8_t)i;
  }
}

static void gf8_init(void) {
  if (g_init) return;
  gen_log_table(g_log);
  gen_exp_table(g_log, g_exp);
  /* build.rs:41-66 multiply / gen_mul_table */
  for (unsigned a = 0; a < 256; ++a)
    for (unsigned b = 0; b < 256; ++b)
      g_mul[a][b] = (a == 0 || b == 0) ? 0 : g_exp[g_log[a] + g_log[b]];
  g_init = 1;
}

void ora_gf8_tables(uint8_t log_table[256], uint8_t exp_table[510]) {
  gf8_init();
  memcpy(log_table, g_log, 256);
  memcpy(exp_table, g_exp, 510);
}

/* src/galois_8/mod.rs:150-152 */
uint8_t ora_gf8_mul(uint8_t a, uint8_t b) {
  gf8_init();
  return g_mul[a][b];
}

/* src/galois_8/mod.rs:155-166: div(a, 0) = 0. */
uint8_t ora_gf8_div(uint8_t a, uint8_t b) {
  gf8_init();
  if (a == 0 || b == 0) return 0;
  int lr = (int)g_log[a] - (int)g_log[b];
  if (lr < 0) lr += 255;
  return g_exp[lr];
}

/* src/galois_8/mod.rs:172-183: 0^0 = 1, exponent reduced mod 255. */
uint8_t ora_gf8_exp(uint8_t a, size_t n) {
  gf8_init();
  if (n == 0) return 1;
  if (a == 0) return 0;
  size_t lr = ((size_t)g_log[a] * (n % 255)) % 255;
  return g_exp[lr];
}

/* src/galois_8/scalar.rs:1-40 (mul_slice / mul_slice_xor) */
void ora_gf8_mul_slice(uint8_t c, const uint8_t* in, uint8_t* out, size_t n) {
  gf8_init();
  for (size_t i = 0; i < n; ++i) out[i] = g_mul[c][in[i]];
}

void ora_gf8_mul_slice_xor(uint8_t c, const uint8_t* in, uint8_t* out, size_t n) {
  gf8_init();
  for (size_t i = 0; i < n; ++i) out[i] ^= g_mul[c][in[i]];
}

/* ------------------------------------------------------ GF((2^8)^2) tower */
/* EXT_POLY = [1, 2, 128]: X^2 = 2X + 128 (src/galois_16.rs:9-11). */
#define T_X(e) ((uint8_t)((e) >> 8))
#define T_C(e) ((uint8_t)((e)&0xff))
#define T_MK(x, c) ((uint16_t)(((unsigned)(x) << 8) | (unsigned)(c)))

/* Mul for Element: FOIL then reduce_from (src/galois_16.rs:688-700,748-760). */
uint16_t ora_gf16_mul(uint16_t a, uint16_t b) {
  uint8_t o0 = ora_gf8_mul(T_X(a), T_X(b));
  uint8_t o1 = ora_gf8_mul(T_C(a), T_X(b)) ^ ora_gf8_mul(T_X(a), T_C(b));
  uint8_t o2 = ora_gf8_mul(T_C(a), T_C(b));
  if (o0 != 0) {
    o1 ^= ora_gf8_mul(2, o0);
    o2 ^= ora_gf8_mul(128, o0);
  }
  return T_MK(o1, o2);
}

/* inverse through the GF8 norm, zero -> zero (src/galois_16.rs:665-686). */
uint16_t ora_gf16_inverse(uint16_t a) {
  uint8_t ax = T_X(a), ac = T_C(a);
  if (ax == 0) return ac == 0 ? 0 : T_MK(0, ora_gf8_div(1, ac));
  uint8_t conj_c = ac ^ ora_gf8_mul(2, ax);
  uint8_t norm = ora_gf8_mul(ac, conj_c) ^ ora_gf8_mul(128, ora_gf8_mul(ax, ax));
  uint8_t ninv = ora_gf8_div(1, norm);
  return T_MK(ora_gf8_mul(ax, ninv), ora_gf8_mul(conj_c, ninv));
}

uint16_t ora_gf16_div(uint16_t a, uint16_t b) { return ora_gf16_mul(a, ora_gf16_inverse(b)); }

/* Element::exp (src/galois_16.rs:637-661): 0^0 = 1, exponent mod 65535. */
uint16_t ora_gf16_exp(uint16_t a, size_t n) {
  if (n == 0) return T_MK(0, 1);
  if (a == 0) return 0;
  n %= 65535;
  if (n == 0) return T_MK(0, 1);
  uint16_t result = T_MK(0, 1), base = a;
  while (n) {
    if (n & 1) result = ora_gf16_mul(result, base);
    base = ora_gf16_mul(base, base);
    n >>= 1;
  }
  return result;
}

/* ----------------------------------------------------------- field dispatch */
static uint16_t f_mul(int f, uint16_t a, uint16_t b) {
  return f == ORA_FIELD_GF8 ? ora_gf8_mul((uint8_t)a, (uint8_t)b) : ora_gf16_mul(a, b);
}
static uint16_t f_div(int f, uint16_t a, uint16_t b) {
  return f == ORA_FIELD_GF8 ? ora_gf8_div((uint8_t)a, (uint8_t)b) : ora_gf16_div(a, b);
}
static uint16_t f_exp(int f, uint16_t a, size_t n) {
  return f == ORA_FIELD_GF8 ? ora_gf8_exp((uint8_t)a, n) : ora_gf16_exp(a, n);
}
static uint16_t f_one(int f) { return f == ORA_FIELD_GF8 ? 1 : T_MK(0, 1); }
/* nth_internal: gf8 n as u8 (mod.rs:119), gf16 [n>>8, n] (galois_16.rs:63-65) */
static uint16_t f_nth(int f, size_t n) { return f == ORA_FIELD_GF8 ? (uint16_t)(uint8_t)n : (uint16_t)n; }
static size_t f_elem_bytes(int f) { return f == ORA_FIELD_GF8 ? 1 : 2; }

/* ------------------------------------------------------------- matrices */
/* Matrix::multiply (src/matrix.rs:195-214) */
int ora_matrix_multiply(int f, const uint16_t* a, int ar, int ac, const uint16_t* b, int bc,
                        uint16_t* out) {
  for (int r = 0; r < ar; ++r)
    for (int c = 0; c < bc; ++c) {
      uint16_t s = 0;
      for (int i = 0; i < ac; ++i) s ^= f_mul(f, a[r * ac + i], b[i * bc + c]);
      out[r * bc + c] = s;
    }
  return ORA_OK;
}

/* Matrix::gaussian_elim (src/matrix.rs:373-422): pivot = first nonzero below. */
static int gaussian_elim(int f, uint16_t* w, int rows, int cols) {
  for (int r = 0; r < rows; ++r) {
    if (w[r * cols + r] == 0) {
      for (int rb = r + 1; rb < rows; ++rb) {
        if (w[rb * cols + r] != 0) {
          for (int c = 0; c < cols; ++c) {
            uint16_t t = w[r * cols + c];
            w[r * cols + c] = w[rb * cols + c];
            w[rb * cols + c] = t;
          }
          break;
        }
      }
    }
    if (w[r * cols + r] == 0) return ORA_SINGULAR_MATRIX;
    if (w[r * cols + r] != f_one(f)) {
      uint16_t scale = f_div(f, f_one(f), w[r * cols + r]);
      for (int c = 0; c < cols; ++c) w[r * cols + c] = f_mul(f, scale, w[r * cols + c]);
    }
    for (int rb = r + 1; rb < rows; ++rb) {
      uint16_t scale = w[rb * cols + r];
      if (scale != 0)
        for (int c = 0; c < cols; ++c) w[rb * cols + c] ^= f_mul(f, scale, w[r * cols + c]);
    }
  }
  for (int d = 0; d < rows; ++d)
    for (int ra = 0; ra < d; ++ra) {
      uint16_t scale = w[ra * cols + d];
      if (scale != 0)
        for (int c = 0; c < cols; ++c) w[ra * cols + c] ^= f_mul(f, scale, w[d * cols + c]);
    }
  return ORA_OK;
}

/* Matrix::invert via [A | I] (src/matrix.rs:424-444). */
int ora_matrix_invert(int f, const uint16_t* a, int n, uint16_t* out) {
  uint16_t* w = (uint16_t*)calloc((size_t)n * 2 * n, sizeof(uint16_t));
  for (int r = 0; r < n; ++r) {
    memcpy(w + (size_t)r * 2 * n, a + (size_t)r * n, (size_t)n * sizeof(uint16_t));
    w[(size_t)r * 2 * n + n + r] = f_one(f);
  }
  int rc = gaussian_elim(f, w, n, 2 * n);
  if (rc == ORA_OK)
    for (int r = 0; r < n; ++r) memcpy(out + (size_t)r * n, w + (size_t)r * 2 * n + n, (size_t)n * 2);
  free(w);
  return rc;
}

/* Matrix::vandermonde (src/matrix.rs:446-459): row r = [nth(r)^0 .. nth(r)^(c-1)]. */
void ora_matrix_vandermonde(int f, int rows, int cols, uint16_t* out) {
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) out[r * cols + c] = f_exp(f, f_nth(f, (size_t)r), (size_t)c);
}

/* build_jerasure_like_matrix (src/core/codec.rs:114-197). */
static int build_jerasure(int f, int k, int total, uint16_t* vm) {
#define VM(r, c) vm[(r) * k + (c)]
  ora_matrix_vandermonde(f, total, k, vm);
  VM(0, 0) = f_one(f);
  for (int i = 1; i < k; ++i) VM(0, i) = 0;
  for (int i = 0; i < k - 1; ++i) VM(total - 1, i) = 0;
  VM(total - 1, k - 1) = f_one(f);
  for (int i = 0; i < k; ++i) {
    int r = i;
    while (r < total && VM(r, i) == 0) ++r;
    if (r != i) {
      if (r >= total) return ORA_INVALID_CUSTOM_MATRIX; /* swap_rows would panic */
      for (int c = 0; c < k; ++c) {
        uint16_t t = VM(r, c);
        VM(r, c) = VM(i, c);
        VM(i, c) = t;
      }
    } else if (VM(i, i) == 0) {
      return ORA_INVALID_CUSTOM_MATRIX;
    }
    uint16_t diag = VM(i, i);
    if (diag == 0) return ORA_INVALID_CUSTOM_MATRIX;
    uint16_t scale = diag != f_one(f) ? f_div(f, f_one(f), diag) : f_one(f);
    if (scale != f_one(f))
      for (int row = 0; row < total; ++row) VM(row, i) = f_mul(f, VM(row, i), scale);
    for (int j = 0; j < k; ++j) {
      uint16_t value = VM(i, j);
      if (j != i && value != 0)
        for (int row = 0; row < total; ++row) VM(row, j) ^= f_mul(f, value, VM(row, i));
    }
  }
  for (int j = 0; j < k; ++j) {
    uint16_t value = VM(k, j);
    if (value == 0) return ORA_INVALID_CUSTOM_MATRIX;
    if (value != f_one(f)) {
      uint16_t scale = f_div(f, f_one(f), value);
      for (int row = k; row < total; ++row) VM(row, j) = f_mul(f, VM(row, j), scale);
    }
  }
  for (int row = k + 1; row < total; ++row) {
    uint16_t value = VM(row, 0);
    if (value == 0) return ORA_INVALID_CUSTOM_MATRIX;
    if (value != f_one(f)) {
      uint16_t scale = f_div(f, f_one(f), value);
      for (int col = 0; col < k; ++col) VM(row, col) = f_mul(f, VM(row, col), scale);
    }
  }
#undef VM
  return ORA_OK;
}

/* build_matrix_with_options / build_matrix / build_cauchy_matrix / build_custom_matrix
 * (src/core/codec.rs:67-112,199-258). */
int ora_build_matrix(int f, int mode, int k, int m, const uint16_t* custom, uint16_t* out) {
  int total = k + m;
  memset(out, 0, (size_t)total * k * sizeof(uint16_t));
  if (mode == ORA_MATRIX_FAST_ONE_PARITY) {
    /* only valid when m == 1 (codec.rs:240-250) */
    for (int c = 0; c < k; ++c) {
      out[c * k + c] = f_one(f);
      out[k * k + c] = f_one(f);
    }
    return ORA_OK;
  }
  if (mode == ORA_MATRIX_VANDERMONDE) {
    uint16_t* v = (uint16_t*)malloc((size_t)total * k * sizeof(uint16_t));
    uint16_t* inv = (uint16_t*)malloc((size_t)k * k * sizeof(uint16_t));
    ora_matrix_vandermonde(f, total, k, v);
    int rc = ora_matrix_invert(f, v, k, inv); /* top k x k */
    if (rc == ORA_OK) ora_matrix_multiply(f, v, total, k, inv, k, out);
    free(v);
    free(inv);
    return rc == ORA_OK ? ORA_OK : ORA_INVALID_CUSTOM_MATRIX;
  }
  if (mode == ORA_MATRIX_CAUCHY) {
    for (int r = 0; r < total; ++r) {
      if (r < k) {
        out[r * k + r] = f_one(f);
      } else {
        for (int c = 0; c < k; ++c) {
          uint16_t den = f_nth(f, (size_t)r) ^ f_nth(f, (size_t)c);
          if (den == 0) return ORA_INVALID_CUSTOM_MATRIX;
          out[r * k + c] = f_div(f, f_one(f), den);
        }
      }
    }
    return ORA_OK;
  }
  if (mode == ORA_MATRIX_JERASURE) return build_jerasure(f, k, total, out);
  if (mode == ORA_MATRIX_CUSTOM) {
    if (!custom) return ORA_INVALID_CUSTOM_MATRIX;
    for (int r = 0; r < k; ++r) out[r * k + r] = f_one(f);
    memcpy(out + (size_t)k * k, custom, (size_t)m * k * sizeof(uint16_t));
    return ORA_OK;
  }
  return ORA_UNSUPPORTED_CODEC_FAMILY;
}

/* --------------------------------------------------------------- slices */
static uint16_t load_elem(int f, const uint8_t* p, size_t e) {
  return f == ORA_FIELD_GF8 ? p[e] : T_MK(p[2 * e], p[2 * e + 1]);
}
static void xor_elem(int f, uint8_t* p, size_t e, uint16_t v) {
  if (f == ORA_FIELD_GF8) {
    p[e] ^= (uint8_t)v;
  } else {
    p[2 * e] ^= T_X(v);
    p[2 * e + 1] ^= T_C(v);
  }
}

/* code_some_slices (src/core/encode.rs:327-602): out_j = XOR_i rows[j][i] * in_i. */
void ora_apply_rows(int f, const uint16_t* rows, int r, int k, const uint8_t* const* in,
                    uint8_t* const* out, size_t len_bytes) {
  size_t eb = f_elem_bytes(f), ne = len_bytes / eb;
  for (int j = 0; j < r; ++j) {
    memset(out[j], 0, len_bytes);
    for (int i = 0; i < k; ++i) {
      uint16_t c = rows[j * k + i];
      if (c == 0) continue;
      if (f == ORA_FIELD_GF8) {
        ora_gf8_mul_slice_xor((uint8_t)c, in[i], out[j], len_bytes);
      } else {
        for (size_t e = 0; e < ne; ++e) xor_elem(f, out[j], e, ora_gf16_mul(c, load_elem(f, in[i], e)));
      }
    }
  }
}

int ora_encode(int f, const uint16_t* matrix, int k, int m, const uint8_t* const* data,
               uint8_t* const* parity, size_t len_bytes) {
  ora_apply_rows(f, matrix + (size_t)k * k, m, k, data, parity, len_bytes);
  return ORA_OK;
}

/* check_classic_parity (src/core/verify.rs:14-33) */
int ora_verify(int f, const uint16_t* matrix, int k, int m, const uint8_t* const* shards,
               size_t len_bytes, int* ok) {
  uint8_t** buf = (uint8_t**)malloc((size_t)m * sizeof(uint8_t*));
  for (int j = 0; j < m; ++j) buf[j] = (uint8_t*)malloc(len_bytes ? len_bytes : 1);
  ora_encode(f, matrix, k, m, shards, buf, len_bytes);
  int good = 1;
  for (int j = 0; j < m; ++j)
    if (memcmp(buf[j], shards[k + j], len_bytes) != 0) good = 0;
  for (int j = 0; j < m; ++j) free(buf[j]);
  free(buf);
  *ok = good;
  return ORA_OK;
}

/* Survivors = first k present shards in index order
 * (src/core/recovery_plan.rs:329-387); decode rows = inverse of the selected
 * survivor rows of the encoding matrix (reconstruct.rs:1068-1174: all three of its
 * cases compute rows of this unique inverse). */
int ora_decode_matrix(int f, const uint16_t* matrix, int k, int m, const uint8_t* present,
                      int* valid, uint16_t* decode) {
  int nv = 0;
  for (int i = 0; i < k + m && nv < k; ++i)
    if (present[i]) valid[nv++] = i;
  if (nv < k) return ORA_TOO_FEW_SHARDS_PRESENT;
  uint16_t* sel = (uint16_t*)malloc((size_t)k * k * sizeof(uint16_t));
  for (int r = 0; r < k; ++r) memcpy(sel + (size_t)r * k, matrix + (size_t)valid[r] * k, (size_t)k * 2);
  int rc = ora_matrix_invert(f, sel, k, decode);
  free(sel);
  return rc == ORA_OK ? ORA_OK : ORA_INVALID_CUSTOM_MATRIX;
}

/* reconstruct_internal / execute_classic_plan_with (src/core/reconstruct.rs:318-416):
 * data stage then parity stage from the restored data. */
int ora_reconstruct(int f, const uint16_t* matrix, int k, int m, uint8_t* const* shards,
                    const uint8_t* present, size_t len_bytes, int data_only,
                    const uint8_t* required) {
  int total = k + m, npresent = 0, nmissing_data = 0, any_missing = 0;
  for (int i = 0; i < total; ++i) {
    npresent += present[i] ? 1 : 0;
    if (!present[i]) {
      any_missing = 1;
      if (i < k) ++nmissing_data;
    }
  }
  if (!any_missing) return ORA_OK; /* all-present early return, recovery_plan.rs:277-286 */
  if (npresent < k) return ORA_TOO_FEW_SHARDS_PRESENT;
  /* which missing shards are targets */
  uint8_t* target = (uint8_t*)calloc((size_t)total, 1);
  int ntargets = 0;
  for (int i = 0; i < total; ++i) {
    if (present[i]) continue;
    int want;
    if (required) want = required[i] != 0;
    else want = (i < k) || !data_only;
    target[i] = (uint8_t)want;
    ntargets += want;
  }
  if (ntargets == 0) {
    free(target);
    return ORA_OK;
  }
  int* valid = (int*)malloc((size_t)k * sizeof(int));
  uint16_t* decode = (uint16_t*)malloc((size_t)k * k * sizeof(uint16_t));
  int rc = ORA_OK;
  /* the full data vector: present data, or decoded rows of the survivors */
  uint8_t** data = (uint8_t**)calloc((size_t)k, sizeof(uint8_t*));
  uint8_t** tmp = (uint8_t**)calloc((size_t)k, sizeof(uint8_t*));
  if (nmissing_data > 0) {
    rc = ora_decode_matrix(f, matrix, k, m, present, valid, decode);
    if (rc != ORA_OK) goto done;
    const uint8_t** surv = (const uint8_t**)malloc((size_t)k * sizeof(uint8_t*));
    for (int c = 0; c < k; ++c) surv[c] = shards[valid[c]];
    for (int d = 0; d < k; ++d) {
      if (present[d]) continue;
      tmp[d] = (uint8_t*)malloc(len_bytes ? len_bytes : 1);
      ora_apply_rows(f, decode + (size_t)d * k, 1, k, surv, &tmp[d], len_bytes);
    }
    free(surv);
  }
  for (int d = 0; d < k; ++d) data[d] = present[d] ? shards[d] : tmp[d];
  for (int d = 0; d < k; ++d)
    if (!present[d] && target[d]) memcpy(shards[d], tmp[d], len_bytes);
  for (int p = k; p < total; ++p) {
    if (present[p] || !target[p]) continue;
    ora_apply_rows(f, matrix + (size_t)p * k, 1, k, (const uint8_t* const*)data, &shards[p], len_bytes);
  }
done:
  for (int d = 0; d < k; ++d) free(tmp[d]);
  free(tmp);
  free(data);
  free(valid);
  free(decode);
  free(target);
  return rc;
}
