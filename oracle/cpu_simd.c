/*
 * cpu_simd.c -- CPU baseline for the benchmark (test infrastructure only).
 *
 * A C restatement of the reference's SIMD CPU encode, timed beside the GPU path:
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

typedef struct {
  int k, m, isa;
  const uint8_t* const* in;
  uint8_t* const* out;
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

static void rows_scalar(const Job* J, int r0, int rn, size_t off, size_t n) {
  for (int r = r0; r < r0 + rn; ++r) {
    uint8_t* o = J->out[r] + off;
    memset(o, 0, n);
    for (int i = 0; i < J->k; ++i) {
      const uint8_t c = J->coef[r][i];
      const uint8_t* x = J->in[i] + off;
      for (size_t p = 0; p < n; ++p) o[p] ^= ora_gf8_mul(c, x[p]);
    }
  }
}

__attribute__((target("avx512f,avx512bw,gfni"))) static void rows_gfni512(const Job* J, int r0, int rn,
                                                                          size_t off, size_t n) {
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
    for (int r = 0; r < rn; ++r) _mm512_storeu_si512((void*)(J->out[r0 + r] + off + p), acc[r]);
  }
  if (p < n) rows_scalar(J, r0, rn, off + p, n - p);
}

__attribute__((target("avx2"))) static void rows_avx2(const Job* J, int r0, int rn, size_t off, size_t n) {
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
    for (int r = 0; r < rn; ++r) _mm256_storeu_si256((__m256i*)(J->out[r0 + r] + off + p), acc[r]);
  }
  if (p < n) rows_scalar(J, r0, rn, off + p, n - p);
}

static void do_chunk(const Job* J, size_t c) {
  const size_t off = c * J->chunk;
  const size_t n = off + J->chunk <= J->len ? J->chunk : J->len - off;
  for (int r0 = 0; r0 < J->m; r0 += 4) {
    const int rn = J->m - r0 < 4 ? J->m - r0 : 4;
    if (J->isa == ISA_GFNI512)
      rows_gfni512(J, r0, rn, off, n);
    else if (J->isa == ISA_AVX2)
      rows_avx2(J, r0, rn, off, n);
    else
      rows_scalar(J, r0, rn, off, n);
  }
}

static void* worker(void* arg) {
  Job* J = (Job*)arg;
  for (;;) {
    size_t c = atomic_fetch_add(&J->next, 1);
    if (c >= J->n_chunks) break;
    do_chunk(J, c);
  }
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

/* Encodes parity = matrix rows k..k+m-1 times data with `threads` threads.
 * Returns the instruction set used (ISA_*), or a negative value on bad input. */
int ora_cpu_encode(const uint16_t* matrix, int k, int m, const uint8_t* const* data,
                   uint8_t* const* parity, size_t len, int threads, int isa) {
  if (k < 1 || k > MAX_K || m < 1 || m > MAX_M || threads < 1 || len == 0) return -1;
  Job* J = (Job*)calloc(1, sizeof(Job));
  if (!J) return -2;
  J->k = k;
  J->m = m;
  J->isa = isa == ISA_AUTO ? ora_cpu_isa() : isa;
  if (J->isa > ora_cpu_isa()) J->isa = ora_cpu_isa();
  J->in = data;
  J->out = parity;
  J->len = len;
  J->chunk = chunk_len(len);
  J->n_chunks = (len + J->chunk - 1) / J->chunk;
  atomic_init(&J->next, 0);
  for (int r = 0; r < m; ++r)
    for (int i = 0; i < k; ++i) {
      const uint8_t c = (uint8_t)matrix[(size_t)(k + r) * k + i];
      J->coef[r][i] = c;
      J->gfni[r][i] = gfni_matrix(c);
      for (int v = 0; v < 16; ++v) {
        J->lo[r][i][v] = ora_gf8_mul(c, (uint8_t)v);
        J->hi[r][i][v] = ora_gf8_mul(c, (uint8_t)(v << 4));
      }
    }
  const int used = J->isa;
  if (threads == 1) {
    worker(J);
  } else {
    pthread_t* t = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    int started = 0;
    while (started < threads - 1 && pthread_create(&t[started], NULL, worker, J) == 0) ++started;
    worker(J); /* the calling thread is the last worker and covers any that did not start */
    for (int i = 0; i < started; ++i) pthread_join(t[i], NULL);
    free(t);
  }
  free(J);
  return used;
}
