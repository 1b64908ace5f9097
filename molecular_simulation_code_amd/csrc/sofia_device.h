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

// Address of the copy of accumulator (r, slot) that this workgroup adds to.
__device__ inline unsigned long long *eacc_slot(long long *base, int r, int slot) {
    const size_t k = ((size_t)r * NE + slot) * ESPREAD + (blockIdx.x % ESPREAD);
    return (unsigned long long *)&base[k * ECOPY];
}

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
        if ((threadIdx.x & 63) == 0) atomicAdd(eacc_slot(slot_base, r0, slot), (unsigned long long)s);
    } else if (active) {
        atomicAdd(eacc_slot(slot_base, r, slot), (unsigned long long)v);
    }
}

__device__ inline void atomic_fix(long long *p, double v) {
    atomicAdd((unsigned long long *)p, (unsigned long long)to_fix(v));
}

// Workgroups are dispatched round-robin over the 8 XCDs.  xcd_block maps dispatch order to
// a logical block so that each XCD gets runs of XCD_RUN consecutive logical blocks (which
// then share that XCD's L2) while any contiguous range still spreads over all XCDs.  A
// bijection on [0, nblocks); the tail beyond whole rounds maps to itself.
constexpr int XCD_RUN = 16;
__device__ inline int xcd_block(int b, int nblocks) {
    const int round = 8 * XCD_RUN;
    if (b >= (nblocks / round) * round) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / XCD_RUN) * 8 + x) * XCD_RUN + (k % XCD_RUN);
}

// Minimum-image difference x[a] - x[b] of the FP64 SoA positions (x | y | z, n each).
__device__ inline void dvec_soa(const double *xd, int n, int a, int b, double box, double ibox,
                                double *o) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v = xd[(size_t)c * n + a] - xd[(size_t)c * n + b];
        o[c] = v - box * rint(v * ibox);
    }
}

// Harmonic bond a-b, E = k (r - r0)^2 (Source/mod_force_bond.f90:24-45).  f receives the
// force on a in fixed-point quanta (b gets -f); returns E in energy quanta.  Swapping a and
// b negates d exactly, so the integers do not depend on the order of the pair.
__device__ inline long long bond_term(const double *xd, int n, int a, int b, double box, double ibox,
                                      double k, double r0, long long f[3]) {
    double d[3];
    dvec_soa(xd, n, a, b, box, ibox, d);
    const double rr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dr = rr - r0;
    const double rk = k * dr;
    const double force = (rr > 0.0) ? -2.0 * rk / rr : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = to_fix(d[c] * force);
    return to_efix(rk * dr);
}

// Harmonic angle a-b-c with apex b, E = k (theta - theta0)^2
// (Source/mod_force_angle.f90:29-90).  f1 / f3 receive the forces on a / c in fixed-point
// quanta, the apex gets -(f1 + f3) (exact momentum balance); returns E in energy quanta.
// The expression is symmetric in a and c.
__device__ inline long long angle_term(const double *xd, int n, int a, int b, int c, double box,
                                       double ibox, double k, double th0, long long f1[3],
                                       long long f3[3]) {
    double d1[3], d2[3];
    dvec_soa(xd, n, a, b, box, ibox, d1);
    dvec_soa(xd, n, c, b, box, ibox, d2);
    const double rsq1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    const double rsq2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    const double r1 = sqrt(rsq1), r2 = sqrt(rsq2);
    double cs = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / (r1 * r2);
    cs = fmin(1.0, fmax(-1.0, cs));
    double sn = sqrt(1.0 - cs * cs);
    if (sn < 0.001) sn = 0.001;
    sn = 1.0 / sn;
    const double dth = acos(cs) - th0;
    const double tk = k * dth;
    const double aa = 2.0 * tk * sn;
    const double a11 = aa * cs / rsq1, a12 = -aa / (r1 * r2), a22 = aa * cs / rsq2;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        f1[q] = to_fix(-(a11 * d1[q] + a12 * d2[q]));
        f3[q] = to_fix(-(a22 * d2[q] + a12 * d1[q]));
    }
    return to_efix(tk * dth);
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
