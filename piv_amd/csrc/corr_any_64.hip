// k_correlate_any instantiations for windows of up to 64 px (any even size).
#include "correlate_any.hpp"

namespace piv {
hipError_t launch_correlate_any_64(int fmt, int deform, const PassPair *pairs, int npairs,
                                   uint32_t max_cells, hipStream_t st)
{
    return launch_corr_any_n<64>(fmt, deform, pairs, npairs, max_cells, st);
}
}  // namespace piv
