// leopard.hip -- placeholder until the FFT kernels land.
#include "leopard.h"
namespace rsgpu {
int init_leopard(Codec&) { return RSGPU_OK; }
int leopard_encode(const Codec&, const rsgpu_cslice*, const rsgpu_slice*, size_t, hipStream_t) { return RSGPU_ERR_UNSUPPORTED; }
int leopard_verify(const Codec&, const rsgpu_cslice*, size_t, int32_t*, hipStream_t) { return RSGPU_ERR_UNSUPPORTED; }
int leopard_reconstruct(const Codec&, rsgpu_slot*, size_t, int, const uint8_t*, size_t, hipStream_t) { return RSGPU_ERR_UNSUPPORTED; }
}
