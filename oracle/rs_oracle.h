/*
 * rs_oracle.h -- CPU restatement of the rustfs-erasure-codec 9.1.0 Reed-Solomon
 * data path, used ONLY as test infrastructure: the parity tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  Nothing in
 * the shipped library links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * part (golden vectors, known answers, closed forms) is listed in
 * tests/test_oracle.py; functions that no check pins are marked
 * "parity unpinned" below.
 *
 * Element conventions:
 *   - GF(2^8) elements are bytes.
 *   - GF(2^16) tower elements (galois_16) are passed as uint16_t with the
 *     X coefficient in the high byte and the constant in the low byte, i.e.
 *     element [x, c] (src/galois_16.rs:9-11) == (x << 8) | c.  In shard memory a
 *     tower element is stored as the byte pair x, c (src/galois_16.rs:19).
 *   - Matrices are row-major arrays of uint16_t in both fields.
 *
 * Error codes mirror src/errors.rs:4-21 (0 = Ok, then 1 + variant index).
 */
#ifndef RS_ORACLE_H
#define RS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORA_OK = 0,
  ORA_TOO_FEW_SHARDS = 1,
  ORA_TOO_MANY_SHARDS,
  ORA_TOO_FEW_DATA_SHARDS,
  ORA_TOO_MANY_DATA_SHARDS,
  ORA_TOO_FEW_PARITY_SHARDS,
  ORA_TOO_MANY_PARITY_SHARDS,
  ORA_TOO_FEW_BUFFER_SHARDS,
  ORA_TOO_MANY_BUFFER_SHARDS,
  ORA_INCORRECT_SHARD_SIZE,
  ORA_TOO_FEW_SHARDS_PRESENT,
  ORA_EMPTY_SHARD,
  ORA_INVALID_SHARD_FLAGS,
  ORA_INVALID_INDEX,
  ORA_INVALID_CUSTOM_MATRIX,
  ORA_UNSUPPORTED_CODEC_FAMILY,
  ORA_UNSUPPORTED_LEOPARD_PROTOTYPE,
  ORA_SINGULAR_MATRIX = 100, /* matrix.rs internal Error::SingularMatrix */
};

enum { ORA_FIELD_GF8 = 8, ORA_FIELD_GF16 = 16 };
enum { ORA_MATRIX_VANDERMONDE = 0, ORA_MATRIX_CAUCHY = 1, ORA_MATRIX_JERASURE = 2,
       ORA_MATRIX_CUSTOM = 3, ORA_MATRIX_FAST_ONE_PARITY = 4 };

/* ---- GF(2^8): build.rs:10-91, src/galois_8/mod.rs:138-193 ---- */
void ora_gf8_tables(uint8_t log_table[256], uint8_t exp_table[510]);
uint8_t ora_gf8_mul(uint8_t a, uint8_t b);
uint8_t ora_gf8_div(uint8_t a, uint8_t b);
uint8_t ora_gf8_exp(uint8_t a, size_t n);
void ora_gf8_mul_slice(uint8_t c, const uint8_t* in, uint8_t* out, size_t n);
void ora_gf8_mul_slice_xor(uint8_t c, const uint8_t* in, uint8_t* out, size_t n);

/* ---- GF((2^8)^2) tower: src/galois_16.rs:605-760 ---- */
uint16_t ora_gf16_mul(uint16_t a, uint16_t b);
uint16_t ora_gf16_div(uint16_t a, uint16_t b);
uint16_t ora_gf16_exp(uint16_t a, size_t n);
uint16_t ora_gf16_inverse(uint16_t a);

/* ---- matrices over either field: src/matrix.rs:195-459 ---- */
int ora_matrix_multiply(int field, const uint16_t* a, int ar, int ac, const uint16_t* b, int bc,
                        uint16_t* out);
int ora_matrix_invert(int field, const uint16_t* a, int n, uint16_t* out);
void ora_matrix_vandermonde(int field, int rows, int cols, uint16_t* out);
/* Encoding matrix (k+m)*k, src/core/codec.rs:67-258.  custom: m rows of k. */
int ora_build_matrix(int field, int mode, int k, int m, const uint16_t* custom, uint16_t* out);

/* ---- classic codec over byte shards (field element = 1 or 2 bytes) ---- */
/* code_some_slices: out_j = sum_i rows[j*k+i] * in_i, src/core/encode.rs:327-602 */
void ora_apply_rows(int field, const uint16_t* rows, int r, int k, const uint8_t* const* in,
                    uint8_t* const* out, size_t len_bytes);
/* encode_sep: src/core/encode.rs:950-998 (validated lengths assumed). */
int ora_encode(int field, const uint16_t* matrix, int k, int m, const uint8_t* const* data,
               uint8_t* const* parity, size_t len_bytes);
/* verify: src/core/verify.rs:14-33; *ok = 1 when all parity matches. */
int ora_verify(int field, const uint16_t* matrix, int k, int m, const uint8_t* const* shards,
               size_t len_bytes, int* ok);
/* reconstruct / reconstruct_data / reconstruct_some:
 *   src/core/recovery_plan.rs:262-517, src/core/reconstruct.rs:318-416,1068-1247.
 * shards[k+m] buffers of len_bytes; present[i] != 0 marks a present shard.
 * required == NULL: goal All (data_only == 0) or Data (data_only != 0).
 * On success every missing targeted shard is written. */
int ora_reconstruct(int field, const uint16_t* matrix, int k, int m, uint8_t* const* shards,
                    const uint8_t* present, size_t len_bytes, int data_only,
                    const uint8_t* required);
/* Survivor selection (first k present, in index order) and the missing-data rows
 * of the inverse (k x k dense decode matrix), reconstruct.rs:1068-1174. */
int ora_decode_matrix(int field, const uint16_t* matrix, int k, int m, const uint8_t* present,
                      int* valid_indices, uint16_t* decode /* k*k */);

/* ---- Leopard GF(2^8): src/core/leopard_gf8/ ---- */
void ora_leo8_tables(uint8_t fft_skew[255], uint8_t log_walsh[256], uint8_t log_lut[256],
                     uint8_t exp_lut[256]);
uint8_t ora_leo8_mul_log(uint8_t a, uint8_t log_b);
int ora_leo8_encode(int k, int m, const uint8_t* const* data, uint8_t* const* parity,
                    size_t len);
/* shards[k+m]; present flags; required[i] != 0 asks for shard i when missing. */
int ora_leo8_reconstruct(int k, int m, uint8_t* const* shards, const uint8_t* present,
                         const uint8_t* required, size_t len);
void ora_leo8_error_locs(int k, int m, const uint8_t* present, uint8_t err_locs[256]);

/* ---- Leopard GF(2^16): src/core/leopard_gf16/ ---- */
void ora_leo16_tables(uint16_t* fft_skew /*65535*/, uint16_t* log_walsh /*65536*/,
                      uint16_t* log_lut /*65536*/, uint16_t* exp_lut /*131072*/);
uint16_t ora_leo16_mul_log(uint16_t a, uint16_t log_b);
int ora_leo16_encode(int k, int m, const uint8_t* const* data, uint8_t* const* parity,
                     size_t len);
int ora_leo16_reconstruct(int k, int m, uint8_t* const* shards, const uint8_t* present,
                          const uint8_t* required, size_t len);
void ora_leo16_error_locs(int k, int m, const uint8_t* present, uint16_t* err_locs /*65536*/);

/* ---- deterministic generators: src/tests/mod.rs:37-93, benches/common/mod.rs:10,224-250 ---- */
uint64_t ora_qc_mix(uint64_t* seed);
uint64_t ora_qc_seed(const uint64_t* parts, size_t n);
void ora_qc_fill(uint8_t* out, size_t n, uint64_t seed);
void ora_corrupt_positions(size_t n, size_t total, uint64_t seed, size_t* out);
/* rand 0.10 SmallRng (Xoshiro256++ seeded by SplitMix64), u8 stream */
void ora_smallrng_fill(uint64_t seed, uint8_t* out, size_t n);
uint64_t ora_derived_seed(int op_tag, size_t k, size_t m, size_t shard_size);
/* FNV-style checksum fold, tests/golden_vectors.rs:5-18 */
uint64_t ora_checksum_shards(const uint8_t* const* shards, const size_t* lens, size_t n);

/* cpu_simd.c: the benchmark's CPU baseline (SIMD restatement of the reference's
 * CPU encode, multithreaded).  isa: 0 auto, 1 scalar, 2 AVX2 nibble tables,
 * 3 AVX-512 GFNI.  Returns the instruction set used, < 0 on bad input. */
int ora_cpu_isa(void);
int ora_cpu_encode(const uint16_t* matrix, int k, int m, const uint8_t* const* data,
                   uint8_t* const* parity, size_t len, int threads, int isa);
int ora_cpu_verify(const uint16_t* matrix, int k, int m, const uint8_t* const* shards, size_t len,
                   int threads, int isa, int* ok);
int ora_cpu_reconstruct(const uint16_t* matrix, int k, int m, uint8_t* const* shards,
                        const uint8_t* present, const uint8_t* data_rows, size_t len, int threads,
                        int isa);

#ifdef __cplusplus
}
#endif
#endif
