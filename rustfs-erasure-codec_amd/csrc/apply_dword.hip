// apply_dword.hip -- the default GF(2^8) tiling of the matrix-apply kernel
// (apply_dword_kernel in apply_kernel.h: one dword per lane), all output counts
// and modes.  launch_apply (classic_kernels.hip) selects it with kTileDword.
#include "apply_kernel.h"

namespace rsgpu {

template <int MODE>
static hipError_t dword_rows(const ApplyArgs& a, int rows, hipStream_t s) {
  switch (rows) {
    case 1: return launch_apply_dword_t<1, MODE>(a, s);
    case 2: return launch_apply_dword_t<2, MODE>(a, s);
    case 3: return launch_apply_dword_t<3, MODE>(a, s);
    case 4: return launch_apply_dword_t<4, MODE>(a, s);
    case 5: return launch_apply_dword_t<5, MODE>(a, s);
    case 6: return launch_apply_dword_t<6, MODE>(a, s);
    case 7: return launch_apply_dword_t<7, MODE>(a, s);
    case 8: return launch_apply_dword_t<8, MODE>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_apply_dword(const ApplyArgs& a, int mode, int rows, hipStream_t s) {
  switch (mode) {
    case kWrite: return dword_rows<kWrite>(a, rows, s);
    case kAccumulate: return dword_rows<kAccumulate>(a, rows, s);
    case kVerify: return dword_rows<kVerify>(a, rows, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace rsgpu
