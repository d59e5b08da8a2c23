// Launch variants of the correlation kernel, per window size.
//
// A variant fixes the occupancy target the register allocator must meet (waves per SIMD):
// fewer waves leave more registers per lane, more waves hide more memory and LDS latency but
// may spill.  Variants differ only in register allocation and instruction scheduling, so
// every variant computes bit-identical results; the autotuner picks one per window size and
// sample format by timing them on the caller's own batch.
#pragma once

namespace piv {

constexpr int kMaxVariants = 3;

struct VariantSet {
    int count;
    int occ[kMaxVariants];  // waves per SIMD; occ[0] is the default
};

constexpr VariantSet corr_variants(int n)
{
    switch (n) {
    case 8: case 10: case 14: case 16: case 18: return {3, {4, 6, 8}};
    case 12: return {3, {6, 4, 8}};
    case 20: case 22: case 24: case 26: case 28: case 30: return {3, {4, 3, 5}};
    case 32: case 34: case 36: case 38: case 40: case 42: return {2, {3, 4}};
    case 44: case 46: return {2, {3, 2}};
    case 50: case 52: case 54: case 58: return {2, {2, 3}};
    case 60: case 62: return {1, {2, 0, 0}};  // three waves per SIMD spill (91 registers at 60)
    case 48: return {2, {2, 3}};  // three-wave workgroups of 40 KB LDS: at most 3 per SIMD
    case 56: return {2, {2, 3}};
    case 64: return {2, {2, 1}};
    default: return {1, {1, 0, 0}};  // 96 and the generic-size kernel
    }
}

}  // namespace piv
