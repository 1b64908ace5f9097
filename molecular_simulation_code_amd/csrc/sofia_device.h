// Device-side helpers shared by the kernel translation units.
#pragma once

#include "sofia_internal.h"

namespace sofia {

__device__ inline double rsq_f64(double x) {
    // v_rsq_f64 then two Newton steps y <- y + y/2 (1 - x y^2)
    double y = __builtin_amdgcn_rsq(x);
    double e = __builtin_fma(-x * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    e = __builtin_fma(-x * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    return y;
}

__device__ inline double rcp_f64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    return y;
}

template <int L, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline long long to_fix(double v) { return __double2ll_rn(v * FIX_SCALE); }
__device__ inline long long to_efix(double v) { return __double2ll_rn(v * EFIX_SCALE); }

// Add an int64 contribution to a per-replica accumulator slot.  One atomic per wave when
// every contributing lane belongs to the same replica, per-lane atomics otherwise;
// integer addition makes the total independent of that choice.
__device__ inline void wave_add(long long v, bool active, int r, long long *slot_base, int slot) {
    const uint64_t m = __ballot(active);
    if (m == 0) return;
    const int r0 = __shfl(r, __ffsll((long long)m) - 1, 64);
    const bool uniform = __all(!active || r == r0);
    if (uniform) {
        long long s = active ? v : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((threadIdx.x & 63) == 0)
            atomicAdd((unsigned long long *)&slot_base[(size_t)r0 * NE + slot], (unsigned long long)s);
    } else if (active) {
        atomicAdd((unsigned long long *)&slot_base[(size_t)r * NE + slot], (unsigned long long)v);
    }
}


__device__ inline void atomic_fix(long long *p, double v) {
    atomicAdd((unsigned long long *)p, (unsigned long long)to_fix(v));
}

// Philox4x32-10 (Salmon et al., SC'11): counter-based generator, so a random number is a
// pure function of (key, counter) -- here (seed, step, atom id) -- and does not depend on
// launch shape or on which replicas share a batch.
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)M0 * c[0];
        const uint64_t p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

// three uniforms in [-1/2, 1/2) with 24-bit resolution
__device__ inline void uniform3(uint64_t seed, uint64_t step, uint32_t atom, float u[3]) {
    uint32_t c[4] = {atom, (uint32_t)step, (uint32_t)(step >> 32), 0x5350u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = (float)(c[d] >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

inline int nblocks(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace sofia
