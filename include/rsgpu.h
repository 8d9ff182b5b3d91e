/*
 * rsgpu.h -- C ABI of the MI355X (gfx950) Reed-Solomon data path.
 *
 * A drop-in GPU backend for the data path of rustfs-erasure-codec 9.1.0:
 * classic galois_8 / galois_16 ReedSolomon and Leopard GF8 / GF16 encode,
 * verify, reconstruct, reconstruct_data, reconstruct_some, and the block-stream
 * wrappers.  Each entry point names the reference interface it replaces.
 *
 * Conventions
 *   - Shard buffers passed to the non-"_host" calls are DEVICE pointers; the
 *     work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     the call returns once it is enqueued.  The library keeps no pointer past
 *     the completion of that work.
 *   - Every call validates its arguments completely before enqueueing any
 *     device work, in the reference's order (macros.rs:142-245 check_piece_count!
 *     then check_slices!; recovery_plan.rs:262-387 for reconstruction), so a
 *     validation error never modifies an output buffer.
 *   - Return values: 0 = Ok, 1..16 = the reference's Error variants in
 *     declaration order (src/errors.rs:4-21), >= 100 = device/runtime errors.
 *   - A codec handle is safe to share between host threads; its decode cache is
 *     guarded by a mutex (the reference keeps it behind a mutex too,
 *     src/core/inversion_cache.rs:38-48).
 */
#ifndef RSGPU_H
#define RSGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: src/errors.rs:4-21 ---- */
enum {
  RSGPU_OK = 0,
  RSGPU_TOO_FEW_SHARDS = 1,
  RSGPU_TOO_MANY_SHARDS = 2,
  RSGPU_TOO_FEW_DATA_SHARDS = 3,
  RSGPU_TOO_MANY_DATA_SHARDS = 4,
  RSGPU_TOO_FEW_PARITY_SHARDS = 5,
  RSGPU_TOO_MANY_PARITY_SHARDS = 6,
  RSGPU_TOO_FEW_BUFFER_SHARDS = 7,
  RSGPU_TOO_MANY_BUFFER_SHARDS = 8,
  RSGPU_INCORRECT_SHARD_SIZE = 9,
  RSGPU_TOO_FEW_SHARDS_PRESENT = 10,
  RSGPU_EMPTY_SHARD = 11,
  RSGPU_INVALID_SHARD_FLAGS = 12,
  RSGPU_INVALID_INDEX = 13,
  RSGPU_INVALID_CUSTOM_MATRIX = 14,
  RSGPU_UNSUPPORTED_CODEC_FAMILY = 15,
  RSGPU_UNSUPPORTED_LEOPARD_PROTOTYPE = 16,
  RSGPU_ERR_HIP = 100,             /* a HIP runtime call failed */
  RSGPU_ERR_INVALID_ARGUMENT = 101, /* NULL handle/pointer or bad enum */
  RSGPU_ERR_NO_DEVICE = 102,
  RSGPU_ERR_UNSUPPORTED = 103,      /* shape beyond this build's device limits */
  RSGPU_ERR_OUT_OF_MEMORY = 104,
  RSGPU_ERR_IO = 105,               /* a block-stream read or write callback failed */
};

/* ---- codec construction: src/core/options.rs:1-147, src/core/codec.rs:264-615 ---- */
enum { RSGPU_FIELD_GF8 = 0, RSGPU_FIELD_GF16 = 1 };                /* galois_8 / galois_16 */
enum { RSGPU_FAMILY_CLASSIC = 0, RSGPU_FAMILY_LEOPARD_GF8 = 1, RSGPU_FAMILY_LEOPARD_GF16 = 2 };
enum { RSGPU_MATRIX_VANDERMONDE = 0, RSGPU_MATRIX_CAUCHY = 1, RSGPU_MATRIX_JERASURE_LIKE = 2,
       RSGPU_MATRIX_CUSTOM = 3 };
enum { RSGPU_LEOPARD_DISABLED = 0, RSGPU_LEOPARD_AS_NEEDED = 1, RSGPU_LEOPARD_PREFER_GF16 = 2,
       RSGPU_LEOPARD_PREFER_LEOPARD = 3 };

/* CodecOptions (src/core/options.rs:99-147). */
typedef struct rsgpu_options {
  int32_t field;            /* RSGPU_FIELD_* (the ReedSolomon<F> type parameter) */
  int32_t codec_family;     /* RSGPU_FAMILY_* */
  int32_t matrix_mode;      /* RSGPU_MATRIX_* (CUSTOM only via rsgpu_codec_with_custom_matrix) */
  int32_t leopard_mode;     /* RSGPU_LEOPARD_* */
  int32_t fast_one_parity;  /* single all-ones parity row when parity == 1 */
  int32_t inversion_cache;  /* cache decode matrices (device-resident) */
  uint64_t inversion_cache_capacity;  /* 0 = derived (codec.rs:44-59) */
  uint64_t inversion_cache_max_bytes; /* 0 = 8 MiB (inversion_cache.rs:11) */
} rsgpu_options;

typedef struct rsgpu_codec rsgpu_codec;

/* CodecOptions::default() */
void rsgpu_options_default(rsgpu_options* opts);
/* ReedSolomon::new / with_options (src/core/codec.rs:264-355). */
int rsgpu_codec_new(size_t data_shards, size_t parity_shards, const rsgpu_options* opts,
                    rsgpu_codec** out);
/* ReedSolomon::with_custom_matrix (src/core/codec.rs:546-615).  `rows` holds
 * n_rows rows of `row_len` field elements (1 byte each for GF8, the byte pair
 * [x, c] for GF16); only the first parity_shards rows and data_shards columns
 * are used, as in build_custom_matrix (codec.rs:199-227). */
int rsgpu_codec_with_custom_matrix(size_t data_shards, size_t parity_shards, const uint8_t* rows,
                                   size_t n_rows, size_t row_len, const rsgpu_options* opts,
                                   rsgpu_codec** out);
void rsgpu_codec_free(rsgpu_codec* codec);
/* data_shard_count / parity_shard_count / codec_family (codec.rs:357-380). */
int rsgpu_codec_info(const rsgpu_codec* codec, size_t* data_shards, size_t* parity_shards,
                     int32_t* field, int32_t* codec_family);
/* The (k+m) x k encoding matrix, row-major, element bytes as in custom rows
 * (EncodingFormat::matrix, src/core/format.rs). */
int rsgpu_codec_matrix(const rsgpu_codec* codec, uint8_t* out, size_t out_len);

/* ---- shard views ---- */
typedef struct rsgpu_cslice { const uint8_t* ptr; size_t len; } rsgpu_cslice;
typedef struct rsgpu_slice { uint8_t* ptr; size_t len; } rsgpu_slice;
/* ShardSlot (src/lib.rs:250-304): a buffer kept across calls with a presence
 * flag.  Missing slots that reconstruction targets must have len == shard
 * length; their bytes are overwritten and `present` is set to 1. */
typedef struct rsgpu_slot { uint8_t* ptr; size_t len; int32_t present; } rsgpu_slot;

/* EncodingFormat::stable_id_bytes (src/core/format.rs:133-219, version 1): the
 * persistent identifier of the effective format (family, field, counts, matrix
 * coefficients).  *len receives the descriptor length; with out == NULL only the
 * length is returned. */
int rsgpu_codec_format_id(const rsgpu_codec* codec, uint8_t* out, size_t cap, size_t* len);
/* Shard length in bytes that ReedSolomon::split gives data_len bytes
 * (codec.rs:476-505; Leopard: leopard_aligned_shard_len, leopard.rs:75-98). */
int rsgpu_split_len(const rsgpu_codec* codec, size_t data_len, size_t* shard_len);
/* ReedSolomon::split into k device shards of rsgpu_split_len bytes (zero-padded). */
int rsgpu_split(const rsgpu_codec* codec, const uint8_t* data, size_t data_len,
                const rsgpu_slice* shards, size_t n_shards, void* stream);
/* ReedSolomon::join (codec.rs:507-538): the first min(out_len, k * len) bytes of the
 * k data shards into `out`; *written receives that count. */
int rsgpu_join(const rsgpu_codec* codec, const rsgpu_cslice* shards, size_t n_shards, size_t out_len,
               uint8_t* out, size_t* written, void* stream);

/* ---- encode: src/core/encode.rs:935-998 ---- */
/* ReedSolomon::encode: shards[0..k] data (read), shards[k..k+m] parity (written). */
int rsgpu_encode(const rsgpu_codec* codec, const rsgpu_slice* shards, size_t n_shards,
                 void* stream);
/* ReedSolomon::encode_sep */
int rsgpu_encode_sep(const rsgpu_codec* codec, const rsgpu_cslice* data, size_t n_data,
                     const rsgpu_slice* parity, size_t n_parity, void* stream);
/* Same outputs as rsgpu_encode_sep, computed as a GF(2) bit-matrix product on the
 * int8 matrix cores (bit-expanded data, parity of the int32 sums).  An evaluated
 * alternative to the VALU kernels with no counterpart in the reference: classic
 * GF(2^8) codecs with k <= 31 and m <= 8, 4-byte-aligned shards whose length is a
 * multiple of 4 (else RSGPU_ERR_INVALID_ARGUMENT); other codecs return
 * RSGPU_UNSUPPORTED_CODEC_FAMILY. */
int rsgpu_encode_sep_gf2mfma(const rsgpu_codec* codec, const rsgpu_cslice* data, size_t n_data,
                             const rsgpu_slice* parity, size_t n_parity, void* stream);

/* ---- incremental encode: src/core/encode.rs:892-1132 (classic codecs only;
 * Leopard codecs return RSGPU_UNSUPPORTED_CODEC_FAMILY) ---- */
/* ReedSolomon::encode_single: parity = c * shards[i_data] when i_data == 0, parity
 * ^= c * shards[i_data] otherwise (ShardByShard feeds shards 0..k-1 in order). */
int rsgpu_encode_single(const rsgpu_codec* codec, size_t i_data, const rsgpu_slice* shards,
                        size_t n_shards, void* stream);
/* ReedSolomon::encode_single_sep */
int rsgpu_encode_single_sep(const rsgpu_codec* codec, size_t i_data, rsgpu_cslice single,
                            const rsgpu_slice* parity, size_t n_parity, void* stream);
/* ReedSolomon::update: new_data[i].ptr == NULL marks an unchanged shard (None);
 * parity ^= c * (old ^ new) for every changed shard.  Asynchronous on `stream`,
 * nothing allocated: the gathered coefficient columns travel in the kernel
 * arguments, or (wide updates, unaligned buffers) each changed column is applied
 * from the codec's own tables. */
int rsgpu_update(const rsgpu_codec* codec, const rsgpu_cslice* old_data, size_t n_old,
                 const rsgpu_cslice* new_data, size_t n_new, const rsgpu_slice* parity,
                 size_t n_parity, void* stream);

/* ---- verify: src/core/verify.rs:14-237 ---- */
/* ReedSolomon::verify / verify_with_workspace: *ok = 1 when every parity shard
 * equals the recomputed parity.  Synchronous (waits for the stream).  Classic
 * codecs compare in registers and never write parity; Leopard codecs re-encode
 * into library-owned scratch. */
int rsgpu_verify(const rsgpu_codec* codec, const rsgpu_cslice* shards, size_t n_shards,
                 int32_t* ok, void* stream);
/* Asynchronous verify: writes 0 (match) or nonzero (mismatch) to the device
 * int32 at `device_flag` on `stream`, whatever it held before.  Classic codecs
 * of at most 64 data shards clear and report in the same launch (a per-stream
 * call id and a 16-byte library workspace per (device, stream) order the clear
 * before any report); other codecs clear with a launch of their own first. */
int rsgpu_verify_async(const rsgpu_codec* codec, const rsgpu_cslice* shards, size_t n_shards,
                       int32_t* device_flag, void* stream);
/* ReedSolomon::verify_with_buffer (verify.rs:98-118): parity is recomputed into
 * `buffer` (m device slices) and compared. */
int rsgpu_verify_with_buffer(const rsgpu_codec* codec, const rsgpu_cslice* shards,
                             size_t n_shards, const rsgpu_slice* buffer, size_t n_buffer,
                             int32_t* ok, void* stream);

/* ---- reconstruct: src/core/reconstruct.rs:539-790 ----
 * Classic codecs solve for at most 8192 missing data shards per call (GF(2^16);
 * the inversion runs in one workgroup, in LDS up to 128 rows and over a device
 * work matrix beyond); more return RSGPU_ERR_UNSUPPORTED.  Missing parity shards
 * are not limited. */
/* ReedSolomon::reconstruct (all missing shards). */
int rsgpu_reconstruct(const rsgpu_codec* codec, rsgpu_slot* shards, size_t n_shards, void* stream);
/* ReedSolomon::reconstruct_data (missing data shards only). */
int rsgpu_reconstruct_data(const rsgpu_codec* codec, rsgpu_slot* shards, size_t n_shards,
                           void* stream);
/* ReedSolomon::reconstruct_some (required flags of length k or k+m). */
int rsgpu_reconstruct_some(const rsgpu_codec* codec, rsgpu_slot* shards, size_t n_shards,
                           const uint8_t* required, size_t n_required, void* stream);
/* ---- prepared reconstruction: ReconstructWorkspace (src/core/reconstruct.rs:539-700,
 * src/core/workspace.rs:85-140).  A workspace owns the plan of one erasure layout
 * (present flags and shard length of `shards`; missing slots need no buffer here)
 * and its device decode rows, which stay valid when the codec's cache evicts or
 * is disabled.  Leopard codecs return RSGPU_UNSUPPORTED_CODEC_FAMILY. ---- */
typedef struct rsgpu_reconstruct_workspace rsgpu_reconstruct_workspace;
/* prepare_reconstruct_workspace: every missing shard. */
int rsgpu_prepare_reconstruct_workspace(const rsgpu_codec* codec, const rsgpu_slot* shards,
                                        size_t n_shards, void* stream,
                                        rsgpu_reconstruct_workspace** out);
/* prepare_reconstruct_data_workspace: missing data shards only. */
int rsgpu_prepare_reconstruct_data_workspace(const rsgpu_codec* codec, const rsgpu_slot* shards,
                                             size_t n_shards, void* stream,
                                             rsgpu_reconstruct_workspace** out);
/* prepare_reconstruct_some_workspace: the missing shards flagged in `required`
 * (k or k+m flags, else RSGPU_INVALID_SHARD_FLAGS). */
int rsgpu_prepare_reconstruct_some_workspace(const rsgpu_codec* codec, const rsgpu_slot* shards,
                                             size_t n_shards, const uint8_t* required,
                                             size_t n_required, void* stream,
                                             rsgpu_reconstruct_workspace** out);
/* reconstruct_with_workspace / reconstruct_with_workspace_par: validates counts,
 * then the codec's format against the workspace's (data/parity counts:
 * TooFew/TooMany{Data,Parity}Shards; other differences: RSGPU_INVALID_CUSTOM_MATRIX),
 * then the layout (another missing pattern: RSGPU_INVALID_SHARD_FLAGS; another
 * length: RSGPU_INCORRECT_SHARD_SIZE), then only enqueues the prepared launches on
 * `stream`.  An equivalent codec (same format, same device) may use the workspace. */
int rsgpu_reconstruct_with_workspace(const rsgpu_codec* codec, rsgpu_slot* shards, size_t n_shards,
                                     const rsgpu_reconstruct_workspace* workspace, void* stream);
/* data_targets() then parity_targets() (workspace.rs:113-122): the first
 * min(cap, n_data + n_parity) target indices into `targets`. */
int rsgpu_reconstruct_workspace_targets(const rsgpu_reconstruct_workspace* workspace, size_t* targets,
                                        size_t cap, size_t* n_data, size_t* n_parity);
void rsgpu_reconstruct_workspace_free(rsgpu_reconstruct_workspace* workspace);

/* ReedSolomon::decode_idx (src/galois_8/policy.rs:543-704): progressive recovery.
 * dst, input: k+m entries each, ptr == NULL = None (a Some of length 0 needs a
 * non-NULL ptr).  With expect_input (n_expect == k+m flags) the k expected
 * shards are the decode's survivors; each call may supply any subset of them as
 * inputs, and every non-NULL destination accumulates (XOR) its row of the decode
 * (data rows of the inverse, parity rows composed through it) over that subset,
 * so calls over a partition of the survivors sum to the reconstruction.
 * Destinations start zeroed by the caller.  With expect_input == NULL, partial
 * results merge: dst[i] ^= input[i] for every non-NULL input (lengths may differ
 * between indices).  Validation follows the reference (Leopard codecs:
 * RSGPU_UNSUPPORTED_CODEC_FAMILY; expected count other than k:
 * RSGPU_TOO_FEW_SHARDS_PRESENT), and every pair is validated before any XOR, so
 * a rejected call changes nothing.  Classic GF(2^16) codecs are accepted too
 * (the reference has the method for galois_8 only); codecs with k > 256, more
 * than 256 destinations or more than 128 missing data shards among the expected
 * return RSGPU_ERR_UNSUPPORTED.  Asynchronous on `stream`. */
int rsgpu_decode_idx(const rsgpu_codec* codec, const rsgpu_slice* dst, size_t n_dst,
                     const uint8_t* expect_input, size_t n_expect, const rsgpu_cslice* input,
                     size_t n_input, void* stream);
/* rsgpu_reconstruct (data_only = 0) or rsgpu_reconstruct_data (data_only = 1)
 * without any allocation: no hipMalloc, hipFree or allocating library call, no
 * decode-cache lookup.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  The
 * decode solve goes into a library-owned device scratch array of 2 MiB per GPU
 * (held until the process ends): 256 B + the coefficient tables (32 B per
 * coefficient, 128 B for GF(2^16)) + the rows, each rounded up to 256 B, for
 * (missing targets) x k coefficients.  Classic codecs of at most 256 shards and
 * patterns of at most 128 missing data shards;
 * larger shapes return RSGPU_ERR_UNSUPPORTED, Leopard codecs
 * RSGPU_UNSUPPORTED_CODEC_FAMILY.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams or threads, before the
 * previous call's kernels finished) race and must be serialised by the caller.
 * Same bytes and validation as the allocating entry points. */
int rsgpu_reconstruct_noalloc(const rsgpu_codec* codec, rsgpu_slot* shards, size_t n_shards,
                              int32_t data_only, void* stream);

/* ---- Leopard plug points: leopard_gf8_encode / leopard_gf16_encode
 * (src/core/leopard.rs:277-296) and reconstruct_with_tables
 * (leopard_gf8/mod.rs:546-562, leopard.rs:299-317): the FFT engines by field
 * (bits 8 or 16), data_shards and parity_shards, without a codec handle (one is
 * kept per shape and device for the process).  Shards are device pointers.
 * encode: n_data / n_parity other than data_shards / parity_shards give
 * RSGPU_TOO_FEW_SHARDS (encode.rs:201-203); the length is a non-zero multiple of
 * 64 shared by every shard.
 * reconstruct: present, inputs and outputs have data_shards + parity_shards
 * entries.  A present shard's input is read; a missing shard is rebuilt into its
 * output unless that output has length 0 (not requested).  Validation follows
 * validate_leopard_decode_buffers (src/core/reconstruct.rs:140-185). ---- */
int rsgpu_leopard_encode(int32_t bits, size_t data_shards, size_t parity_shards,
                         const rsgpu_cslice* data, size_t n_data, const rsgpu_slice* parity,
                         size_t n_parity, void* stream);
int rsgpu_leopard_reconstruct(int32_t bits, size_t data_shards, size_t parity_shards,
                              const uint8_t* present, const rsgpu_cslice* inputs,
                              const rsgpu_slice* outputs, void* stream);

/* ---- block streams: src/core/stream.rs:677-1031 ----
 * Shard streams are read and written through callbacks, one block at a time
 * (block_size clamped to [1 KiB, 16 MiB]; the reference's default is 4 MiB).
 * read(user, shard, buf, len) returns the bytes read (0 = end of stream, < 0 =
 * error); write(user, shard, buf, len) returns the bytes written (<= 0 = error).
 * Failures return RSGPU_ERR_IO (callback failure) or a codec status, with the
 * StreamError of the reference (stream.rs:107-171) in *err: kind, shard index
 * and the callback's return value.  Leopard codecs return
 * RSGPU_UNSUPPORTED_CODEC_FAMILY as in the reference.
 *
 * Pipelining: blocks are staged through two pinned host buffers.  While the
 * device copies and encodes block i, the host reads block i+1; while the device
 * works on block i+1, the host writes the outputs of block i.  Consequences for
 * stream positions (the reference, stream.rs:918-925, promises none on error):
 *   - the outputs of block i are written before an error of block i+1 (a read
 *     error, a short shard) is returned, as in the reference;
 *   - when a write of block i fails, or verify finds block i inconsistent, the
 *     readers have also consumed block i+1 (the reference stops after block i);
 *   - on success every reader is at its end, as in the reference.
 *
 * I/O mode (StreamOptions::io_mode, stream.rs:49-104, 214-226): SERIAL calls
 * every callback on the calling thread, one shard after another.  PARALLEL calls
 * the callbacks of one block's shards concurrently from library worker threads
 * (at most one call per shard at a time, and a block's calls end before the
 * next block's start); the callbacks must then be safe to call concurrently for
 * different shards, as the reference requires `Read + Send`.  AUTO (default)
 * is PARALLEL when block_size >= 4 MiB and at least 10 streams take part in that
 * direction (reads or writes), else SERIAL.  When several shards of a parallel
 * block fail, the lowest shard index is reported. */
enum { RSGPU_STREAM_OK = 0, RSGPU_STREAM_READ = 1, RSGPU_STREAM_WRITE = 2, RSGPU_STREAM_CODEC = 3 };
enum { RSGPU_STREAM_IO_AUTO = 0, RSGPU_STREAM_IO_SERIAL = 1, RSGPU_STREAM_IO_PARALLEL = 2 };
typedef struct rsgpu_stream_options {
  size_t block_size;   /* bytes per shard per block; default 4 MiB, clamped per call */
  int32_t io_mode;     /* RSGPU_STREAM_IO_* */
  int32_t reserved;    /* 0 */
} rsgpu_stream_options;
typedef struct rsgpu_stream_error {
  int32_t kind;        /* RSGPU_STREAM_* */
  int32_t shard_index;
  int64_t io_error;    /* READ/WRITE: the callback's return value; CODEC: the status */
} rsgpu_stream_error;
typedef int64_t (*rsgpu_read_cb)(void* user, size_t shard, uint8_t* buf, size_t len);
typedef int64_t (*rsgpu_write_cb)(void* user, size_t shard, const uint8_t* buf, size_t len);
/* StreamOptions::default() (stream.rs:73-80): 4 MiB blocks, AUTO.  A NULL
 * options pointer in the calls below means these defaults. */
void rsgpu_stream_options_default(rsgpu_stream_options* options);
/* ReedSolomon::encode_stream (stream.rs:677-752): readers 0..k-1 are the data
 * streams, writers 0..m-1 the parity streams (error index k + j). */
int rsgpu_encode_stream(const rsgpu_codec* codec, size_t n_data, size_t n_parity, rsgpu_read_cb read,
                        rsgpu_write_cb write, void* user, const rsgpu_stream_options* options,
                        rsgpu_stream_error* err);
/* ReedSolomon::verify_stream (stream.rs:760-812). */
int rsgpu_verify_stream(const rsgpu_codec* codec, size_t n_shards, rsgpu_read_cb read, void* user,
                        const rsgpu_stream_options* options, int32_t* ok, rsgpu_stream_error* err);
/* ReedSolomon::reconstruct_stream_from (stream.rs:948-1031): has_reader[i] /
 * has_writer[i] give the Some/None layout of the reference's two slices. */
int rsgpu_reconstruct_stream_from(const rsgpu_codec* codec, size_t n_readers, size_t n_writers,
                                  const uint8_t* has_reader, const uint8_t* has_writer,
                                  rsgpu_read_cb read, rsgpu_write_cb write, void* user,
                                  const rsgpu_stream_options* options, rsgpu_stream_error* err);
/* In-memory shard streams for the calls above, the C form of the reference's
 * std::io::Cursor streams (stream.rs:839-846): pass rsgpu_mem_stream_read /
 * rsgpu_mem_stream_write as the callbacks and a rsgpu_mem_streams as `user`.
 * A read copies from data[pos..len) and advances pos (0 at the end).  A write
 * copies into data[pos..len) and advances pos; a full buffer accepts 0 bytes,
 * which the stream calls report as a write error (the WriteZero of a Cursor
 * over a fixed slice).  Safe in PARALLEL mode: each call touches one shard. */
typedef struct rsgpu_mem_stream {
  uint8_t* data;
  size_t len;
  size_t pos;
} rsgpu_mem_stream;
typedef struct rsgpu_mem_streams {
  rsgpu_mem_stream* readers;  /* indexed by the reader's shard argument */
  rsgpu_mem_stream* writers;  /* indexed by the writer's shard argument */
} rsgpu_mem_streams;
int64_t rsgpu_mem_stream_read(void* streams, size_t shard, uint8_t* buf, size_t len);
int64_t rsgpu_mem_stream_write(void* streams, size_t shard, const uint8_t* buf, size_t len);
/* ReedSolomon::reconstruct_stream (stream.rs:839-900), the in-memory form over
 * the reference's Cursor<Vec<u8>> shards: rsgpu_vec is a byte vector (data,
 * len, cap) with a cursor position.  A shard with len != 0 is present and is
 * read from position 0 (pos is reset first, as the reference resets the
 * cursor); pos ends where the reading stopped.  A shard with len == 0 is
 * missing: the rebuilt bytes are appended at data[len..] and pos is not moved
 * (the reference writes through the cursor's Vec).  When a vector lacks room,
 * grow(user, shard, vec, need) must make vec->cap >= need (moving data, setting
 * data and cap; len and pos stay) and return 0; the first append of a shard asks
 * for the common present length when all present shards have one (the
 * reference's reserve_exact).  grow may be NULL when the caller pre-sized every
 * missing vector; a vector that cannot grow fails as a write error of that
 * shard.  In PARALLEL I/O grow is called concurrently for different shards.
 * Errors as the reference: count errors, then UnsupportedCodecFamily for
 * Leopard; no present shard returns RSGPU_OK with nothing written; fewer than
 * k present is TooFewShardsPresent; then as rsgpu_reconstruct_stream_from. */
typedef struct rsgpu_vec {
  uint8_t* data;
  size_t len;
  size_t cap;
  size_t pos;
} rsgpu_vec;
typedef int (*rsgpu_grow_cb)(void* user, size_t shard, rsgpu_vec* vec, size_t need);
int rsgpu_reconstruct_stream(const rsgpu_codec* codec, rsgpu_vec* shards, size_t n_shards, rsgpu_grow_cb grow,
                             void* user, const rsgpu_stream_options* options, rsgpu_stream_error* err);

/* ---- host-pointer variants (copy through library-owned device staging) ---- */
int rsgpu_encode_host(const rsgpu_codec* codec, uint8_t* const* shards, size_t n_shards,
                      size_t shard_len);
int rsgpu_verify_host(const rsgpu_codec* codec, const uint8_t* const* shards, size_t n_shards,
                      size_t shard_len, int32_t* ok);
int rsgpu_reconstruct_host(const rsgpu_codec* codec, uint8_t* const* shards, uint8_t* present,
                           size_t n_shards, size_t shard_len, int32_t data_only);

/* ---- the matrix-apply primitive (SelectedEncodeKernel::execute,
 * src/galois_8/codegen.rs:46-79; code_some_slices encode.rs:327-602):
 * out_j = XOR_i rows[j*k + i] * in_i over GF(2^8) (field 0) or the tower
 * field (field 1, rows as [x, c] byte pairs).  Asynchronous on `stream`: the
 * coefficient tables of up to 64 row sets stay on the device, so a repeated row
 * set only enqueues the kernels; a new one is built and uploaded first. ---- */
int rsgpu_apply_rows(int32_t field, const uint8_t* rows, size_t r, size_t k,
                     const rsgpu_cslice* in, const rsgpu_slice* out, void* stream);

/* ---- batches: many stripe operations in one kernel launch.
 * The reference runs each stripe call separately and parallelises inside it
 * (encode_sep_par encode.rs:1177, execute_classic_plan_parallel
 * reconstruct.rs:229, ParallelPolicy parallel.rs:65-264); a batch is that
 * parallel policy for many stripes at once.  Each add_* validates exactly like
 * the single-stripe call of the same name and records the op.  Classic ops run
 * in the batch kernels (one launch per element type present: GF(2^8) bytes,
 * GF(2^16) tower pairs), and so do the matrix forms of Leopard ops: GF(2^8)
 * encode, verify and reconstruct when the codec's dispatch picks a matrix
 * variant (with no dispatch state: k <= 32), GF(2^16) reconstruct when it picks
 * one (with no dispatch state: at most 4 erasures; split-layout kernel, 16-byte
 * aligned shards).  Other Leopard ops run
 * their FFT kernels after the batch kernels, on the same stream.  Reconstruct
 * ops enqueue any decode-row solve on `stream` when added and mark their targets
 * present.  Verify ops OR a nonzero value into `device_flag` on mismatch (the
 * caller zeroes it).  A batch
 * refers to each op's codec (its tables, or the codec itself for ops run through
 * their own kernels) and shards without copying them: both must stay alive until
 * the batch is cleared or freed. ---- */
typedef struct rsgpu_batch rsgpu_batch;
int rsgpu_batch_new(rsgpu_batch** out);
void rsgpu_batch_free(rsgpu_batch* batch);
int rsgpu_batch_clear(rsgpu_batch* batch);
int rsgpu_batch_add_encode(rsgpu_batch* batch, const rsgpu_codec* codec, const rsgpu_slice* shards,
                           size_t n_shards);
int rsgpu_batch_add_verify(rsgpu_batch* batch, const rsgpu_codec* codec, const rsgpu_cslice* shards,
                           size_t n_shards, int32_t* device_flag);
int rsgpu_batch_add_reconstruct(rsgpu_batch* batch, const rsgpu_codec* codec, rsgpu_slot* shards,
                                size_t n_shards, int32_t data_only, void* stream);
/* Kernel shape: inputs per pipelined load group (4, 6, 8 or 12), 16-byte vectors
 * per lane (1 or 2; 1 for 12-input groups) and resident workgroups per CU (1..64;
 * 0 launches one workgroup per tile instead of a persistent grid). */
int rsgpu_batch_set_tune(rsgpu_batch* batch, int32_t inputs_per_group, int32_t vectors_per_lane,
                         int32_t blocks_per_cu);
/* Times every kernel shape on this batch (`reps` runs each) and keeps the fastest. */
int rsgpu_batch_autotune(rsgpu_batch* batch, void* stream, int32_t reps, int32_t* inputs_per_group,
                         int32_t* vectors_per_lane, int32_t* blocks_per_cu, float* ms);
int rsgpu_batch_run(rsgpu_batch* batch, void* stream);
/* Same work, captured once into a hipGraph and replayed on later calls.  A batch
 * that is a single kernel launch is launched directly (a replay only adds its
 * start-up), and one holding ops run through their own kernels runs as
 * rsgpu_batch_run. */
int rsgpu_batch_graph_launch(rsgpu_batch* batch, void* stream);
int rsgpu_batch_stats(const rsgpu_batch* batch, uint64_t* n_ops, uint64_t* bytes_moved,
                      uint64_t* data_bytes, uint64_t* tiles);

/* ---- kernel-variant dispatch.
 * The reference picks a CPU kernel per call from the ISA and the shard shape
 * (SelectedEncodeKernel::select, src/galois_8/codegen.rs:10-44; ParallelPolicy
 * chunking, parallel.rs:65-127; Leopard prototype selection leopard.rs:191-223).
 * Here every call can run one of rsgpu_variant_count() kernel variants that all
 * produce the same bytes:
 *   kind 0: matrix apply, `lanes` per workgroup, `vectors_per_lane` 16-byte
 *           vectors per lane, `rows_per_group` outputs per workgroup (0 = all);
 *           on a Leopard GF(2^8) codec this applies its generator matrix
 *           (encode, verify, reconstruct with exactly k shards present) or, for
 *           reconstruct with more present, the FFT decoder's own matrix over
 *           every present shard (cached per erasure pattern); on a Leopard
 *           GF(2^16) codec, reconstruct of at most 32 erased shards only,
 *           through the FFT decoder's own matrix (split-layout elements, 16-byte
 *           aligned shards; others take the FFT);
 *   kind 1: Leopard FFT with a column tile of `leopard_tile_words` words.
 * A codec uses its forced variant if set, else its variant table, else the
 * default choice; an entry of -1, or one that is not valid for a call, also
 * takes the default choice.  The default choice is the default tiling, except
 * that a Leopard reconstruct of 1..4 erased shards runs its matrix form with
 * variant 12 (64-lane tiles, one output per workgroup).  Setting either while
 * calls run on the codec is not synchronised with those calls (each call reads
 * one consistent table). ---- */
int rsgpu_variant_count(void);
/* Table geometry: ops (0 encode, 1 verify, 2 reconstruct), size buckets, and the
 * bucket offset: bucket = clamp(round(2 * log2(shard_len)) - bucket0). */
int rsgpu_dispatch_dims(int32_t* n_ops, int32_t* n_buckets, int32_t* bucket0);
int rsgpu_variant_info(int32_t variant, int32_t* kind, int32_t* lanes, int32_t* vectors_per_lane,
                       int32_t* rows_per_group, int32_t* leopard_tile_words);
/* 1 if `variant` is exact for `op` with `erasures` missing shards on this codec. */
int rsgpu_variant_valid(const rsgpu_codec* codec, int32_t op, size_t erasures, int32_t variant);
/* Forces one variant for every call (-1 clears). */
int rsgpu_codec_set_variant(rsgpu_codec* codec, int32_t variant);
/* Installs table[op][e][bucket] (n_ops x n_erasures x n_buckets int16, -1 =
 * default choice; erasure counts >= n_erasures use the last row); NULL clears. */
int rsgpu_codec_set_variant_table(rsgpu_codec* codec, const int16_t* table, size_t n_erasures);
/* The variant a call with these parameters would run (-1 = default tiling). */
int rsgpu_codec_variant_for(const rsgpu_codec* codec, int32_t op, size_t erasures, size_t shard_len,
                            int32_t* variant);

/* ---- request server: HBM admission and continuous batching.
 * No counterpart in the reference (each call there runs to completion on the
 * caller's thread; stream.rs:677-960 pipelines one object).  Requests name a
 * classic codec, an op and k+m device shard pointers that are read in place; the
 * server allocates the outputs (encode: m parity shards; reconstruct: one per
 * erased shard, in shard order) from its arena.  A request is admitted, in
 * arrival order, once its footprint (k+m)*shard_len fits under hbm_budget and its
 * outputs fit in the arena; admitted requests are packed into the next idle
 * batch slot (max_inflight_batches streams, at most max_batch_ops per batch),
 * so batches overlap on the device.  Outputs stay valid until released. ---- */
typedef struct rsgpu_server rsgpu_server;
typedef struct {
  uint64_t hbm_budget;          /* admission limit on the summed footprint of admitted requests */
  uint64_t arena_bytes;         /* device memory for outputs */
  int32_t max_inflight_batches; /* batch slots (streams), 1..16 */
  int32_t max_batch_ops;        /* requests per batch */
} rsgpu_server_config;
typedef struct {
  int32_t op;                   /* 0 encode, 1 verify, 2 reconstruct */
  int32_t reserved;
  const rsgpu_codec* codec;
  size_t shard_len;
  const uint8_t* const* shards; /* k+m device pointers (erased ones are not read) */
  const uint8_t* erased;        /* reconstruct: k+m flags, nonzero = missing */
  uint64_t arrival_ns;          /* replay: arrival offset from the start of the replay */
  uint64_t tag;                 /* caller's unique id */
} rsgpu_request;
typedef struct {
  uint64_t tag;
  int32_t status;               /* 0, or the validation error of the op */
  int32_t verify_ok;            /* verify: 1 if every parity shard matched */
  uint64_t arrival_ns, admit_ns, launch_ns, done_ns;
  uint64_t footprint;
} rsgpu_completion;
typedef struct {
  uint64_t completed, batches, launched_ops, max_inflight_batches, peak_footprint, footprint;
} rsgpu_server_counters;
int rsgpu_server_new(const rsgpu_server_config* config, rsgpu_server** out);
/* Waits for in-flight batches, then frees the server and its arena. */
void rsgpu_server_free(rsgpu_server* server);
/* Queues a request arriving now. */
int rsgpu_server_submit(rsgpu_server* server, const rsgpu_request* request);
/* One loop turn: retire finished batches, admit, launch. */
int rsgpu_server_pump(rsgpu_server* server, uint32_t* batches_in_flight);
int rsgpu_server_poll(rsgpu_server* server, rsgpu_completion* out, size_t max, size_t* n);
int rsgpu_server_outputs(const rsgpu_server* server, uint64_t tag, uint8_t** outputs, size_t n);
int rsgpu_server_release(rsgpu_server* server, uint64_t tag);
/* Open-loop replay: submits each request at its arrival offset (requests sorted
 * by arrival), runs the loop until all complete; out[i] is request i's record
 * (times relative to the replay start).  keep_outputs = 0 releases on completion.
 * A request that cannot be admitted while nothing is in flight (kept outputs,
 * the replay's or earlier ones, hold the budget or the arena) completes with
 * status RSGPU_ERR_OUT_OF_MEMORY. */
int rsgpu_server_replay(rsgpu_server* server, const rsgpu_request* requests, size_t n,
                        int32_t keep_outputs, rsgpu_completion* out);
int rsgpu_server_stats(const rsgpu_server* server, rsgpu_server_counters* stats);

/* ---- the dispatcher's network: fp32 x[n_in <= 16] -> 64 ReLU -> 64 ReLU ->
 * y[n_out <= 32], trained on the device.  It predicts per-variant log-times from
 * call features so a dispatcher can fill variant tables without timing every
 * variant (the reference's selection is the fixed cascade of codegen.rs:10-44).
 * Parameters (host layout): W1[64][n_in] b1[64] W2[64][64] b2[64]
 * W3[n_out][64] b3[n_out].  x, y, t, w are device arrays, row-major per sample.
 * Loss = sum w (y - t)^2 / sum w; training takes full-batch Adam steps
 * (beta 0.9 / 0.999). ---- */
typedef struct rsgpu_mlp rsgpu_mlp;
int rsgpu_mlp_new(int32_t n_in, int32_t n_out, uint64_t seed, rsgpu_mlp** out);
void rsgpu_mlp_free(rsgpu_mlp* mlp);
int rsgpu_mlp_dims(const rsgpu_mlp* mlp, int32_t* n_in, int32_t* n_hidden, int32_t* n_out,
                   size_t* n_params);
int rsgpu_mlp_get_params(const rsgpu_mlp* mlp, float* params, size_t n_params);
/* Also resets the Adam moments. */
int rsgpu_mlp_set_params(rsgpu_mlp* mlp, const float* params, size_t n_params);
int rsgpu_mlp_forward(rsgpu_mlp* mlp, const float* x, size_t batch, float* y, void* stream);
/* Loss and its gradient (host outputs) at the current parameters. */
int rsgpu_mlp_loss_grad(rsgpu_mlp* mlp, const float* x, const float* t, const float* w, size_t batch,
                        float* loss, float* grad, size_t n_params, void* stream);
/* `steps` Adam steps with step size lr; *loss (optional) = loss before the last step. */
int rsgpu_mlp_train(rsgpu_mlp* mlp, const float* x, const float* t, const float* w, size_t batch,
                    int32_t steps, float lr, float* loss, void* stream);

/* Installs the dispatcher network's decisions as the codec's variant table: one
 * forward pass over every (op, erasures, size bucket) of this codec, argmin over
 * the variants valid for each call.  `envelope` (NULL = no limit) is int64
 * [4][6]: k_min, k_max, m_min, m_max, shard_len_min, shard_len_max of the trained
 * shapes per kind (classic GF8, classic GF16, Leopard GF8, Leopard GF16; k_max 0 =
 * none); calls outside it keep the default tiling.  The network must have 14
 * inputs (the features of rsgpu/dispatch.py) and rsgpu_variant_count() outputs,
 * or one more: that last output predicts the default tiling, which is installed
 * (-1) where it is strictly below every valid variant's prediction. */
int rsgpu_codec_install_dispatch(rsgpu_codec* codec, rsgpu_mlp* net, const int64_t* envelope,
                                 void* stream);

/* ---- decode-matrix cache statistics (ReconstructionCacheStats, codec.rs:401-426) ---- */
int rsgpu_cache_stats(const rsgpu_codec* codec, uint64_t* requests, uint64_t* hits,
                      uint64_t* misses, uint64_t* entries);

/* ---- library ---- */
const char* rsgpu_status_name(int status);
/* Binds the calling thread to `device` and builds the per-device tables. */
int rsgpu_init(int device);
/* Last HIP error text of the calling thread (empty if none). */
const char* rsgpu_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RSGPU_H */
