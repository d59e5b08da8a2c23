// k_correlate_mf (DFT-GEMM, 16-bit MFMA) instantiations for 16 px windows.
#include "correlate_mfma.hpp"

namespace piv {
hipError_t launch_correlate_mf_16(int fmt, int precision, int deform, const PassPair *pairs,
                                  int npairs, uint32_t max_cells, hipStream_t st)
{
    return launch_corr_mf_n<16>(fmt, precision, deform, pairs, npairs, max_cells, st);
}
}  // namespace piv
