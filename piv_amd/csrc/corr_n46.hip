// k_correlate instantiations for 46 px windows.
#include "correlate.hpp"

namespace piv {
hipError_t launch_correlate_46(int fmt, int variant, int deform, const PassPair *pairs, int npairs,
                              uint32_t max_cells, hipStream_t st)
{
    return launch_corr_n<46>(fmt, variant, deform, pairs, npairs, max_cells, st);
}
}  // namespace piv
