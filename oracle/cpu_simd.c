/*
 * cpu_simd.c -- CPU baseline for the benchmark (test infrastructure only).
 *
 * A C restatement of the reference's SIMD CPU encode, verify and reconstruct,
 * timed beside the GPU path:
 *   - coefficient rows applied chunk by chunk, every output of a chunk computed
 *     in one pass over its inputs (the GFNI row kernel of
 *     src/galois_8/x86/gfni_rows.rs:14-90: up to four outputs in registers);
 *   - GF(2^8) multiply by a constant as a GFNI affine transform with the 8x8 bit
 *     matrix of x -> c*x (GFNI_MULTIPLY_MATRICES, src/galois_8/mod.rs:33) on
 *     64-byte AVX-512 vectors, or as two 16-entry nibble lookups (MUL_LOW /
 *     MUL_HIGH, build.rs) with AVX2 byte shuffles where GFNI is absent;
 *   - chunk length from the reference's serial rule (16 KiB / 64 KiB / 256 KiB,
 *     src/core/encode.rs:604-620) and chunks spread over a pool of threads, the
 *     shape of code_some_slices_par_chunked (encode.rs:622-700) under
 *     ParallelPolicy (src/core/parallel.rs:65-127).
 * Its output is checked against the scalar oracle (classic.c) in the CPU tests.
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rs_oracle.h"

enum { ISA_AUTO = 0, ISA_SCALAR = 1, ISA_AVX2 = 2, ISA_GFNI512 = 3 };
#define MAX_K 64
#define MAX_M 32

/* One row-apply job: out_r = XOR_i coef[r][i] * in_i over `len` bytes, chunk by
 * chunk.  With `cmp` set (verify), the rows are computed into the worker's
 * scratch and compared with cmp_r instead of being written. */
typedef struct {
  int k, m, isa;
  const uint8_t* const* in;
  uint8_t* const* out;
  const uint8_t* const* cmp;
  atomic_int mismatch;
  size_t len, chunk, n_chunks;
  atomic_size_t next;
  uint8_t coef[MAX_M][MAX_K];
  uint64_t gfni[MAX_M][MAX_K];
  uint8_t lo[MAX_M][MAX_K][16], hi[MAX_M][MAX_K][16];
} Job;

/* Row i of the affine matrix (byte 7 - i) selects the input bits that feed
 * output bit i of c * x. */
static uint64_t gfni_matrix(uint8_t c) {
  uint64_t a = 0;
  for (int i = 0; i < 8; ++i) {
    uint64_t row = 0;
    for (int j = 0; j < 8; ++j)
      if ((ora_gf8_mul(c, (uint8_t)(1u << j)) >> i) & 1) row |= 1ull << j;
    a |= row << (8 * (7 - i));
  }
  return a;
}

/* Rows r0 .. r0 + rn of the input bytes [off, off + n) into dst[r] + doff. */
static void rows_scalar(const Job* J, uint8_t* const* dst, size_t doff, int r0, int rn, size_t off, size_t n) {
  for (int r = r0; r < r0 + rn; ++r) {
    uint8_t* o = dst[r] + doff;
    memset(o, 0, n);
    for (int i = 0; i < J->k; ++i) {
      const uint8_t c = J->coef[r][i];
      const uint8_t* x = J->in[i] + off;
      for (size_t p = 0; p < n; ++p) o[p] ^= ora_gf8_mul(c, x[p]);
    }
  }
}

__attribute__((target("avx512f,avx512bw,gfni"))) static void rows_gfni512(const Job* J, uint8_t* const* dst,
                                                                          size_t doff, int r0, int rn, size_t off,
                                                                          size_t n) {
  size_t p = 0;
  for (; p + 64 <= n; p += 64) {
    __m512i acc[4] = {_mm512_setzero_si512(), _mm512_setzero_si512(), _mm512_setzero_si512(),
                      _mm512_setzero_si512()};
    for (int i = 0; i < J->k; ++i) {
      const __m512i x = _mm512_loadu_si512((const void*)(J->in[i] + off + p));
      for (int r = 0; r < rn; ++r)
        acc[r] = _mm512_xor_si512(
            acc[r], _mm512_gf2p8affine_epi64_epi8(x, _mm512_set1_epi64((long long)J->gfni[r0 + r][i]), 0));
    }
    for (int r = 0; r < rn; ++r) _mm512_storeu_si512((void*)(dst[r0 + r] + doff + p), acc[r]);
  }
  if (p < n) rows_scalar(J, dst, doff + p, r0, rn, off + p, n - p);
}

__attribute__((target("avx2"))) static void rows_avx2(const Job* J, uint8_t* const* dst, size_t doff, int r0,
                                                      int rn, size_t off, size_t n) {
  const __m256i mask = _mm256_set1_epi8(0x0f);
  size_t p = 0;
  for (; p + 32 <= n; p += 32) {
    __m256i acc[4] = {_mm256_setzero_si256(), _mm256_setzero_si256(), _mm256_setzero_si256(),
                      _mm256_setzero_si256()};
    for (int i = 0; i < J->k; ++i) {
      const __m256i x = _mm256_loadu_si256((const __m256i*)(J->in[i] + off + p));
      const __m256i xl = _mm256_and_si256(x, mask);
      const __m256i xh = _mm256_and_si256(_mm256_srli_epi64(x, 4), mask);
      for (int r = 0; r < rn; ++r) {
        const __m256i tl = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)J->lo[r0 + r][i]));
        const __m256i th = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i*)J->hi[r0 + r][i]));
        acc[r] = _mm256_xor_si256(acc[r], _mm256_xor_si256(_mm256_shuffle_epi8(tl, xl),
                                                           _mm256_shuffle_epi8(th, xh)));
      }
    }
    for (int r = 0; r < rn; ++r) _mm256_storeu_si256((__m256i*)(dst[r0 + r] + doff + p), acc[r]);
  }
  if (p < n) rows_scalar(J, dst, doff + p, r0, rn, off + p, n - p);
}

/* scratch (verify): the worker's J->m rows of one chunk. */
static void do_chunk(Job* J, size_t c, uint8_t* scratch) {
  const size_t off = c * J->chunk;
  const size_t n = off + J->chunk <= J->len ? J->chunk : J->len - off;
  uint8_t* rows[MAX_M];
  uint8_t* const* dst = J->out;
  size_t doff = off;
  if (J->cmp) {
    for (int r = 0; r < J->m; ++r) rows[r] = scratch + (size_t)r * J->chunk;
    dst = rows;
    doff = 0;
  }
  for (int r0 = 0; r0 < J->m; r0 += 4) {
    const int rn = J->m - r0 < 4 ? J->m - r0 : 4;
    if (J->isa == ISA_GFNI512)
      rows_gfni512(J, dst, doff, r0, rn, off, n);
    else if (J->isa == ISA_AVX2)
      rows_avx2(J, dst, doff, r0, rn, off, n);
    else
      rows_scalar(J, dst, doff, r0, rn, off, n);
  }
  if (J->cmp)
    for (int r = 0; r < J->m; ++r)
      if (memcmp(rows[r], J->cmp[r] + off, n) != 0) atomic_store(&J->mismatch, 1);
}

static void* worker(void* arg) {
  Job* J = (Job*)arg;
  uint8_t* scratch = J->cmp ? (uint8_t*)malloc((size_t)J->m * J->chunk) : NULL;
  if (J->cmp && !scratch) {
    atomic_store(&J->mismatch, 2);
    return NULL;
  }
  for (;;) {
    size_t c = atomic_fetch_add(&J->next, 1);
    if (c >= J->n_chunks) break;
    do_chunk(J, c, scratch);
  }
  free(scratch);
  return NULL;
}

/* serial_code_chunk_len (src/core/encode.rs:610-620) */
static size_t chunk_len(size_t len) {
  if (len <= 16384) return len;
  if (len <= 65536) return 16384;
  if (len <= 4u << 20) return 65536;
  return 262144;
}

int ora_cpu_isa(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("gfni") && __builtin_cpu_supports("avx512bw")) return ISA_GFNI512;
  if (__builtin_cpu_supports("avx2")) return ISA_AVX2;
  return ISA_SCALAR;
}

/* Runs a row-apply job of r rows (coefficient bytes row-major, r x k) with
 * `threads` threads.  Returns the instruction set used (ISA_*), or a negative
 * value on bad input; with cmp, *mismatch receives 1 when any row differs. */
static int run_rows(const uint8_t* rows, int k, int r, const uint8_t* const* in, uint8_t* const* out,
                    const uint8_t* const* cmp, int* mismatch, size_t len, int threads, int isa) {
  if (k < 1 || k > MAX_K || r < 1 || r > MAX_M || threads < 1 || len == 0) return -1;
  Job* J = (Job*)calloc(1, sizeof(Job));
  if (!J) return -2;
  J->k = k;
  J->m = r;
  J->isa = isa == ISA_AUTO ? ora_cpu_isa() : isa;
  if (J->isa > ora_cpu_isa()) J->isa = ora_cpu_isa();
  J->in = in;
  J->out = out;
  J->cmp = cmp;
  atomic_init(&J->mismatch, 0);
  J->len = len;
  J->chunk = chunk_len(len);
  J->n_chunks = (len + J->chunk - 1) / J->chunk;
  atomic_init(&J->next, 0);
  for (int j = 0; j < r; ++j)
    for (int i = 0; i < k; ++i) {
      const uint8_t c = rows[(size_t)j * k + i];
      J->coef[j][i] = c;
      J->gfni[j][i] = gfni_matrix(c);
      for (int v = 0; v < 16; ++v) {
        J->lo[j][i][v] = ora_gf8_mul(c, (uint8_t)v);
        J->hi[j][i][v] = ora_gf8_mul(c, (uint8_t)(v << 4));
      }
    }
  const int used = J->isa;
  /* ParallelPolicy::decide (src/core/parallel.rs:65-127): shards below
   * PARALLEL_MIN_SHARD_BYTES (256 KiB, src/core/mod.rs:63) run serially */
  if (len < (256u << 10)) threads = 1;
  if (threads == 1) {
    worker(J);
  } else {
    pthread_t* t = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    int started = 0;
    while (t && started < threads - 1 && pthread_create(&t[started], NULL, worker, J) == 0) ++started;
    worker(J); /* the calling thread is the last worker and covers any that did not start */
    for (int i = 0; i < started; ++i) pthread_join(t[i], NULL);
    free(t);
  }
  const int mm = atomic_load(&J->mismatch);
  free(J);
  if (mm == 2) return -2;
  if (mismatch) *mismatch = mm;
  return used;
}

static void parity_rows(const uint16_t* matrix, int k, int m, uint8_t* rows) {
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < k; ++i) rows[(size_t)j * k + i] = (uint8_t)matrix[(size_t)(k + j) * k + i];
}

/* Encodes parity = matrix rows k..k+m-1 times data with `threads` threads.
 * Returns the instruction set used (ISA_*), or a negative value on bad input. */
int ora_cpu_encode(const uint16_t* matrix, int k, int m, const uint8_t* const* data,
                   uint8_t* const* parity, size_t len, int threads, int isa) {
  if (k < 1 || k > MAX_K || m < 1 || m > MAX_M) return -1;
  uint8_t rows[MAX_M * MAX_K];
  parity_rows(matrix, k, m, rows);
  return run_rows(rows, k, m, data, parity, NULL, NULL, len, threads, isa);
}

/* verify (src/core/verify.rs:14-33): the parity recomputed chunk by chunk into
 * scratch and compared with the stored parity; *ok = 1 when every byte matches. */
int ora_cpu_verify(const uint16_t* matrix, int k, int m, const uint8_t* const* shards, size_t len,
                   int threads, int isa, int* ok) {
  if (k < 1 || k > MAX_K || m < 1 || m > MAX_M || !ok) return -1;
  uint8_t rows[MAX_M * MAX_K];
  parity_rows(matrix, k, m, rows);
  int mismatch = 0;
  const int used = run_rows(rows, k, m, shards, NULL, shards + k, &mismatch, len, threads, isa);
  *ok = mismatch ? 0 : 1;
  return used;
}

/* The two stages of classic reconstruction (execute_classic_plan_with,
 * src/core/reconstruct.rs:318-416): missing data from the k survivors through
 * their decode rows, then missing parity from the completed data through the
 * encoding rows.  `data_rows` holds one row of k coefficients (over the
 * survivors, in index order) per missing data shard; the caller computes it once
 * per erasure pattern, as the reference's inversion cache does.  Returns the
 * instruction set used, or a negative value. */
int ora_cpu_reconstruct(const uint16_t* matrix, int k, int m, uint8_t* const* shards,
                        const uint8_t* present, const uint8_t* data_rows, size_t len, int threads,
                        int isa) {
  if (k < 1 || k > MAX_K || m < 1 || m > MAX_M) return -1;
  const uint8_t* surv[MAX_K];
  uint8_t* dout[MAX_M];
  uint8_t* pout[MAX_M];
  int ns = 0, nd = 0, np = 0, used = isa;
  uint8_t prow[MAX_M * MAX_K];
  for (int i = 0; i < k + m; ++i) {
    if (present[i]) {
      if (ns < k) surv[ns++] = shards[i];
    } else if (i < k) {
      if (nd >= MAX_M) return -1;
      dout[nd++] = shards[i];
    } else {
      if (np >= MAX_M) return -1;
      for (int c = 0; c < k; ++c) prow[(size_t)np * k + c] = (uint8_t)matrix[(size_t)i * k + c];
      pout[np++] = shards[i];
    }
  }
  if (ns < k) return -1;
  if (nd && (used = run_rows(data_rows, k, nd, surv, dout, NULL, NULL, len, threads, isa)) < 0) return used;
  if (np) {
    const uint8_t* data[MAX_K];
    for (int i = 0; i < k; ++i) data[i] = shards[i];
    if ((used = run_rows(prow, k, np, data, pout, NULL, NULL, len, threads, isa)) < 0) return used;
  }
  return used;
}
