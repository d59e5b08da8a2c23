// k_correlate instantiations for 64 px windows.
#include "correlate.hpp"

namespace piv {
hipError_t launch_correlate_64(int fmt, const PassPair *pairs, const uint32_t *begins, int npairs,
                              uint32_t total, uint32_t wpp, hipStream_t st)
{
    return launch_corr_n<64>(fmt, pairs, begins, npairs, total, wpp, st);
}
}  // namespace piv
