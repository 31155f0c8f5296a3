// qcb_rng.hpp -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// SC'11) for the device heatbath and hot start.
//
// Counter scheme (the oracle restates the same one):
//   key  = { seed & 0xffffffff, seed >> 32 }
//   heatbath:  ctr = { cb_site, sweep_idx, (hs<<24)|(hit<<8)|(subgroup<<4), call }
//              hs = 2*mu + parity; call a feeds Kennedy-Pendleton attempt a with
//              (delta, r1, r2, r3) = u32(w0..w3).  Sphere draws: with m = 3*hit +
//              subgroup, the call 0x80000000 at the counter of subgroup m - (m&1)
//              feeds subgroup m with u32(w[2(m&1)]), u32(w[2(m&1)+1]) -- one call per
//              two subgroups.  u32(w) = (w + 1/2) 2^-32 is exact in double; the
//              float path takes the float nearest to it, so FP32 and FP64 lattices
//              consume identical uniforms.
//   hot start: ctr = { lex_site, mu, 1, call 0..8 } (nine Box-Muller pairs) with
//              u53(w0,w1), u53(w2,w3): top 52 bits of (hi:lo) + 2^-53, strictly in (0,1).
// Every number depends only on (seed, counter), never on launch shape or on
// other lattices in a batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qcb {

struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32) gives hi and lo together
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96 = parity)
        U4 n;
        n.x = __builtin_amdgcn_bitop3_b32(hi1, c.y, k0, 0x96);
        n.y = lo1;
        n.z = __builtin_amdgcn_bitop3_b32(hi0, c.w, k1, 0x96);
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 12) * 0x1.0p-52 + 0x1.0p-53;
}

__device__ __forceinline__ double u32d(uint32_t w) { return ((double)w + 0.5) * 0x1.0p-32; }

// The float nearest to (w + 1/2) 2^-32, i.e. (float)u32d(w), without double
// arithmetic: (w >> 8) 2^-24 and ((w & 255) + 1/2) 2^-32 are exact floats and
// the fma rounds their exact sum once.
__device__ __forceinline__ float u32f(uint32_t w) {
    const float lo = __builtin_fmaf((float)(w & 255u), 0x1.0p-32f, 0x1.0p-33f);
    return __builtin_fmaf((float)(w >> 8), 0x1.0p-24f, lo);
}

__host__ __device__ __forceinline__ uint32_t hb_ctr_z(int hs, int hit, int g) {
    return ((uint32_t)hs << 24) | ((uint32_t)hit << 8) | ((uint32_t)g << 4);
}

}  // namespace qcb
