// leopard.h -- Leopard GF(2^8) / GF(2^16) FFT codecs on the device (leopard.hip).
#pragma once
#include "internal.h"

namespace rsgpu {

// `variant` (see kNumVariants) picks the FFT column tile; -1 or a matrix-apply
// variant keeps the default tile.  Launches are asynchronous on `s`.
int init_leopard(Codec& c);
int leopard_encode(const Codec& c, const rsgpu_cslice* data, const rsgpu_slice* parity, size_t len,
                   hipStream_t s, int variant);
// Re-encodes in registers and compares; sets *d_flag nonzero on mismatch.
int leopard_verify(const Codec& c, const rsgpu_cslice* shards, size_t len, int32_t* d_flag,
                   hipStream_t s, int variant);
// goal: 0 all, 1 data only, 2 required flags.  A matrix-apply variant on a GF(2^8)
// codec goes to leopard_matrix_reconstruct (exactly k shards present) or to
// leopard_decoder_matrix_reconstruct (more); on a GF(2^16) codec to
// leopard_decoder_matrix_reconstruct.
int leopard_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                        const uint8_t* required, size_t n_required, hipStream_t s, int variant);

// Leopard GF(2^8) byte field as log/exp tables (its Cantor-basis representation
// multiplies differently from the 0x11d field of the classic codecs).
std::shared_ptr<const FieldTables> leopard8_field_tables();

// Generator-matrix path of a Leopard GF(2^8) codec (codec.hip): the operation on
// the classic matrix-apply kernels over [I; G], where G is the codec's parity map.
int leopard_matrix_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                               const uint8_t* required, size_t n_required, hipStream_t s, int variant);

// Reconstruct as one matrix apply over the present shards, bit-identical to the
// FFT decoder (codec.hip): GF(2^8) codecs with more than k shards present, GF(2^16)
// codecs with any.  flags[i]: bit 0 present, bit 1 a target; every shard is len
// bytes.
int leopard_decoder_matrix_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, const uint8_t* flags,
                                       size_t len, hipStream_t s, int variant);

// Leopard GF(2^16) field helpers of that path: the multiplicative identity, and
// the matrix-apply tables of n coefficients for the split shard layout (four byte
// maps each, in the kernel's plane order: low->low, high->low, low->high,
// high->high; a zero coefficient gets zero maps).
uint16_t leopard16_one();
// A table of device pointers staged on stream s through the calling thread's
// pinned ring (*d its device address); release it after the launches that read
// it are enqueued.
int stage_pointer_table(const std::vector<const uint8_t*>& ptrs, hipStream_t s, const uint8_t* const** d,
                        size_t* off, size_t* bytes);
int release_pointer_table(size_t off, size_t bytes, hipStream_t s);
void leopard16_apply_tables(const uint16_t* coef, size_t n, uint32_t* out);

}  // namespace rsgpu
