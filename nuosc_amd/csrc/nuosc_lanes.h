// nuosc_lanes.h -- cross-lane sums of FP64 values over a 64-lane wavefront
// (gfx950): the moment reductions of the stage kernels.
//
// wave_sum16 adds 16 per-lane values over the wave as a halving tree:
// level 1 pairs lane l with l ^ 32, level 2 with l ^ 16, level 3 with l ^ 8,
// level 4 with l ^ 4; each level a lane keeps the half of its values that
// its own lane bit selects and adds the partner's copy of the same half;
// then butterflies over l ^ 2 and l ^ 1.  Lane l ends with the wave total of
// value l >> 2.  wave_sum32 adds 32 values the same way with one more halving
// level (over l ^ 2); lane l ends with the total of value l >> 1.
//
// The exchanges use no LDS and no per-lane selects:
//   l ^ 32, l ^ 16  V_PERMLANE32_SWAP / V_PERMLANE16_SWAP exchange the upper
//                   half-wave (odd rows) of one register with the lower
//                   half-wave (even rows) of the other, so the sum of the
//                   two swapped registers is already the kept half;
//   l ^ 8, l ^ 4    a DPP move restricted by bank mask to the lanes whose bit
//                   is clear (they take the partner's low half) and one to the
//                   lanes whose bit is set (the partner's high half); the
//                   other lanes keep their own value of the other half, so
//                   again one add gives the kept half (row_ror:8 is l ^ 8
//                   within a 16-lane row; row_shl:4 on banks 0, 2 and
//                   row_shr:4 on banks 1, 3 are l ^ 4);
//   l ^ 2, l ^ 1    DPP quad permutations.
// Every level adds the same two operands as the shuffle-and-select form of
// the same tree, so the totals are bitwise those of that form (IEEE addition
// is commutative).
#pragma once

#include <hip/hip_runtime.h>

namespace nuosc {

__device__ __forceinline__ unsigned long long f64_bits(double x) { return __builtin_bit_cast(unsigned long long, x); }
__device__ __forceinline__ double f64_of(unsigned lo, unsigned hi) {
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// lanes 32-63 of a <-> lanes 0-31 of b
__device__ __forceinline__ void swap_halves(double& a, double& b) {
    const unsigned long long ua = f64_bits(a), ub = f64_bits(b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = f64_of(lo[0], hi[0]);
    b = f64_of(lo[1], hi[1]);
}

// odd 16-lane rows of a <-> even rows of b
__device__ __forceinline__ void swap_rows(double& a, double& b) {
    const unsigned long long ua = f64_bits(a), ub = f64_bits(b);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    a = f64_of(lo[0], hi[0]);
    b = f64_of(lo[1], hi[1]);
}

// DPP move of a double: lanes in the enabled banks read src at the DPP
// source lane, the others keep old
template <int CTRL, int BANKS>
__device__ __forceinline__ double dpp_f64(double old, double src) {
    const unsigned long long uo = f64_bits(old), us = f64_bits(src);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)uo, (int)(unsigned)us, CTRL, 0xF, BANKS, false);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(uo >> 32), (int)(unsigned)(us >> 32), CTRL, 0xF, BANKS, false);
    return f64_of(lo, hi);
}

constexpr int kDppRowRor8 = 0x128, kDppRowShl4 = 0x104, kDppRowShr4 = 0x114;
constexpr int kDppXor2 = 0x4E, kDppXor1 = 0xB1;   // quad_perm [2,3,0,1], [1,0,3,2]

// Wave total of 16 per-lane values: lane l ends with the total of value l >> 2.
__device__ __forceinline__ double wave_sum16(const double (&a)[16]) {
    double b[8], c[4], d[2];
#pragma unroll
    for (int k = 0; k < 8; ++k) {   // l ^ 32: lanes 0-31 keep values 0-7, lanes 32-63 values 8-15
        double x = a[k], y = a[k + 8];
        swap_halves(x, y);
        b[k] = x + y;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {   // l ^ 16
        double x = b[k], y = b[k + 4];
        swap_rows(x, y);
        c[k] = x + y;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)     // l ^ 8: banks 0, 1 have the bit clear
        d[k] = dpp_f64<kDppRowRor8, 0x3>(c[k + 2], c[k]) + dpp_f64<kDppRowRor8, 0xC>(c[k], c[k + 2]);
    // l ^ 4: banks 0, 2 have the bit clear
    double e = dpp_f64<kDppRowShl4, 0x5>(d[1], d[0]) + dpp_f64<kDppRowShr4, 0xA>(d[0], d[1]);
    e = e + dpp_f64<kDppXor2, 0xF>(e, e);
    e = e + dpp_f64<kDppXor1, 0xF>(e, e);
    return e;
}

// Wave total of 32 per-lane values: lane l ends with the total of value l >> 1
// (the same exchanges, one more halving level over l ^ 2, where the kept half
// is chosen per lane).
__device__ __forceinline__ double wave_sum32(const double (&a)[32], int lane) {
    double b[16], c[8], d[4], e[2];
#pragma unroll
    for (int k = 0; k < 16; ++k) {   // l ^ 32
        double x = a[k], y = a[k + 16];
        swap_halves(x, y);
        b[k] = x + y;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {    // l ^ 16
        double x = b[k], y = b[k + 8];
        swap_rows(x, y);
        c[k] = x + y;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)      // l ^ 8
        d[k] = dpp_f64<kDppRowRor8, 0x3>(c[k + 4], c[k]) + dpp_f64<kDppRowRor8, 0xC>(c[k], c[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; ++k)      // l ^ 4
        e[k] = dpp_f64<kDppRowShl4, 0x5>(d[k + 2], d[k]) + dpp_f64<kDppRowShr4, 0xA>(d[k], d[k + 2]);
    const bool hb = lane & 2;        // l ^ 2
    const double keep = hb ? e[1] : e[0], send = hb ? e[0] : e[1];
    double f = keep + dpp_f64<kDppXor2, 0xF>(send, send);
    f = f + dpp_f64<kDppXor1, 0xF>(f, f);
    return f;
}

}  // namespace nuosc
