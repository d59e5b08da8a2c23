// k_correlate instantiations for 16 px windows.
#include "correlate.hpp"

namespace piv {
hipError_t launch_correlate_16(int fmt, const PassPair *pairs, const uint32_t *begins, int npairs,
                              uint32_t total, uint32_t wpp, hipStream_t st)
{
    return launch_corr_n<16>(fmt, pairs, begins, npairs, total, wpp, st);
}
}  // namespace piv
