// Window sizes with an instantiated correlation kernel: the register-resident FFT kernel
// k_correlate<N> (one corr_n<N>.hip per size: every even size up to 64, the 7-smooth ones
// up to 128) and the DFT-GEMM kernel k_correlate_mf<N>
// (corrmf_n<N>.hip).  The dispatch in piv_kernels.hip is generated from these lists.
#pragma once

#define PIV_FFT_SIZES(X) \
    X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(34) X(36)  \
    X(38) X(40) X(42) X(44) X(46) X(48) X(50) X(52) X(54) X(56) X(58) X(60) X(62) X(64)      \
    X(70) X(72) X(80) X(84) X(90) X(96) X(98) X(100) X(108) X(112) X(120) X(126) X(128)
#define PIV_MF_SIZES(X) X(16) X(32) X(64)
