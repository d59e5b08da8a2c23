// k_correlate_mf (DFT-GEMM, 16-bit MFMA) instantiations for 64 px windows.
#include "correlate_mfma.hpp"

namespace piv {
hipError_t launch_correlate_mf_64(int fmt, int precision, int deform, const PassPair *pairs,
                                  int npairs, uint32_t max_cells, hipStream_t st)
{
    return launch_corr_mf_n<64>(fmt, precision, deform, pairs, npairs, max_cells, st);
}
}  // namespace piv
