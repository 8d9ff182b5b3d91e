// leopard.h -- Leopard GF(2^8) / GF(2^16) FFT codecs on the device (leopard.hip).
#pragma once
#include "internal.h"

namespace rsgpu {

int init_leopard(Codec& c);
int leopard_encode(const Codec& c, const rsgpu_cslice* data, const rsgpu_slice* parity, size_t len,
                   hipStream_t s);
// Re-encodes into library scratch and compares; sets *d_flag nonzero on mismatch.
int leopard_verify(const Codec& c, const rsgpu_cslice* shards, size_t len, int32_t* d_flag,
                   hipStream_t s);
// goal: 0 all, 1 data only, 2 required flags
int leopard_reconstruct(const Codec& c, rsgpu_slot* shards, size_t n, int goal,
                        const uint8_t* required, size_t n_required, hipStream_t s);

}  // namespace rsgpu
