// ldpcp_device.h -- noise primitives shared by the run-mode and sweep kernels.
//
// Counter-based Philox4x32-10 (Salmon et al., "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) replaces the reference's MKL MT19937 banks
// (RANDOMNUMBER.cpp:157-340).  Every draw is a pure function of
// (seed, trajectory id, role, index), so a trajectory's noise does not depend
// on which thread, launch or GPU integrates it.
//
// Counter layout {index_lo, index_hi, trajectory, role}, key {seed_lo, seed_hi}:
//   role 0: step noise, one call per pair of steps (steps 2q+1, 2q+2):
//           words 0,1 -> Box-Muller pair (cos -> odd step, sin -> even step),
//           words 2,3 -> per-step flip uniforms (odd, even step)
//   role 1: initial draws, index 0: words 0,1 -> Box-Muller (cos -> V0, sin -> R0),
//           word 2 -> initial barrier state uniform
//   role 2: catch-up uniforms of the barrier hand-overs, four per call: the
//           trajectory's n-th hand-over takes word n mod 4 of index n / 4
//   role 3 + 256*array: state of barrier `id` when first touched (word 0)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ldpcp_gauss.h"

namespace ldpcp {

struct Philox4 {
    uint32_t w[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one full 32x32 -> 64-bit product per multiplier (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox4 out;
    out.w[0] = c0;
    out.w[1] = c1;
    out.w[2] = c2;
    out.w[3] = c3;
    return out;
}

// The same function where c0, c1, c3 and the key are wave-uniform and only c2
// (the trajectory) differs between lanes, as in the per-step noise (counter =
// step pair, role 0).  Then the first rounds' words are partly uniform: round 1
// leaves n2 and n3 uniform, so round 2's product M1 * c2 is a scalar multiply
// and round 3's c1 is uniform.  Rounds 1-3 are written as plain XORs with the
// uniform terms grouped (the scalar unit forms them); rounds 4-10, where every
// word is per-lane, use v_bitop3_b32.  Same result as philox4x32_10.
__device__ __forceinline__ Philox4 philox4x32_10_lane_ctr(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                       uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0, n2;
        if (r == 0) {  // c1 uniform, hi1 per lane; hi0 and c3 uniform
            n0 = hi1 ^ (c1 ^ k0);
            n2 = hi0 ^ c3 ^ k1;
        } else if (r == 1) {  // c2 uniform: hi1 uniform, c1 per lane; c3 uniform, hi0 per lane
            n0 = c1 ^ (hi1 ^ k0);
            n2 = hi0 ^ (c3 ^ k1);
        } else {  // c1 uniform; hi0, c3 per lane
            n0 = hi1 ^ (c1 ^ k0);
            n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        }
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox4 out;
    out.w[0] = c0;
    out.w[1] = c1;
    out.w[2] = c2;
    out.w[3] = c3;
    return out;
}

// (w + 1/2) / 2^32: exact in FP64, never 0 or 1.
__host__ __device__ __forceinline__ double unit_f64(uint32_t w)
{
    return ((double)w + 0.5) * 2.3283064365386963e-10;
}

// Box-Muller pair from the IEEE-operation logarithm and sine/cosine of
// ldpcp_gauss.h: the same bits on the device and in a host build.
__host__ __device__ __forceinline__ void box_muller_f64(uint32_t a, uint32_t b, double& g0, double& g1)
{
    ldpcp_box_muller(a, b, &g0, &g1);
}

// Integer form of "u < p" for u = (w + 1/2) / 2^32: w < ceil(p 2^32 - 1/2).
__host__ __device__ __forceinline__ uint64_t flip_threshold(double p)
{
    double P = ceil(p * 4294967296.0 - 0.5);
    if (P < 0.0) P = 0.0;
    if (P > 4294967296.0) P = 4294967296.0;
    return (uint64_t)P;
}

}  // namespace ldpcp
