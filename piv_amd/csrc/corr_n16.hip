// k_correlate instantiations for 16 px windows.
#include "correlate.hpp"

namespace piv {
hipError_t launch_correlate_16(int fmt, const PassPair *pairs, int npairs, uint32_t max_cells,
                              hipStream_t st)
{
    return launch_corr_n<16>(fmt, pairs, npairs, max_cells, st);
}
}  // namespace piv
