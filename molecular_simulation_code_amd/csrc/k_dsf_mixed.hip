// Mixed-precision DSF pair kernels over the full ("nonewton") Verlet list, gfx950:
// lj_cut_coul_dsf_nonewton_NOMIC (Source/mod_force_nonbond.f90:105-196) with FP32 pair
// arithmetic and exact per-atom force sums.
//
// Positions.  Before a launch, k_fixpos writes each atom as one 16-byte record: three
// unsigned 32-bit fractions of its replica's box, X = round(x 2^32 / box) mod 2^32, with the
// atom type in the low 6 bits of X (the x resolution becomes 2^-26 box, 3e-6 A at 217 A,
// finer than a float32 coordinate there), and the charge.  The difference of two such
// words taken in 32-bit integers is already the minimum image (the box maps to 2^32
// exactly), so the pair loop has no rint/fma wrap; the integer difference is converted to
// float and scaled by box 2^-32.  Fixed point also removes the cancellation of FP32
// coordinates near the box edge: separations carry full FP32 relative precision.  One
// gather per list entry brings position, type and charge.  When every atom of a type carries
// the same charge the record drops the charge (12 bytes) and C q_i q_j comes from a
// per-type-pair table in LDS, formed as the kernel would form it from the charges.
//
// Pair math (force over r):
//   dr6i = (1/r^2)^3, 0 beyond rcut2;  F_LJ/r = dr6i (lj1 dr6i - lj2) / r^2
//   F_C/r = C qi qj g(r), g from a cubic Hermite table in r over [0.5 A, cut_coul] built
//   in FP64 from the reference's A&S 7.1.26 erfc; one extra all-zero bin past the last
//   takes every r >= cut_coul; below 0.5 A, g is evaluated directly (dsf_direct).
// Energies (energy launches only) use the matching table for E_C and lj3/lj4 for E_LJ,
// halved (full list); the virial is 0.5 sum dx^2 F/r as in the PME kernel.
//
// Exact sums.  Each pair force component is truncated to an integer number of 2^-24
// kcal/mol/A and those integers are summed in FP64, exact while a sum stays below 2^53
// quanta (5e8 kcal/mol/A).  A per-atom force therefore does not depend on the lanes per
// atom, on the list order, on the compaction below or on how replicas are batched.  Every
// multiply-add of the pair math is an explicit fmaf or kept apart, so no kernel leaves
// the fusing to the compiler and a pair rounds the same way in both kernels of this file.
//
// Loads.  The list row, the position records and the charges are read with raw buffer
// loads (32-bit offsets from a wave-uniform descriptor; an offset past the end of a buffer
// reads zero).
#include <type_traits>

#include "sofia_device.h"

namespace sofia {
namespace {

constexpr float TWO_M32 = 2.3283064365386963e-10f;   // 2^-32
constexpr int RSRC_WORD3 = 0x00020000;               // raw buffer, 32-bit dwords (gfx9)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             RSRC_WORD3);
}
__device__ __forceinline__ uint32_t ld32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ int4 ld128(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}
__device__ __forceinline__ int4 ld96(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)off, 0, 0);
    return make_int4((int)v[0], (int)v[1], (int)v[2], 0);
}
// byte offset of record j: 12 or 16 bytes per record.  The 12-byte form uses the 24-bit
// multiply (full VALU rate; a 32-bit multiply issues at a quarter of it): the host keeps
// 12-byte records to systems of fewer than 2^24 atoms.
template <bool QT>
__device__ __forceinline__ uint32_t rec_off(uint32_t j) {
    return QT ? __umul24(j, 12u) : (j << 4);
}
__device__ __forceinline__ uint32_t clamp_bytes(size_t b) {
    return b > 0xffffffffull ? 0xffffffffu : (uint32_t)b;
}

__device__ __forceinline__ float norm2_f(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}
__device__ __forceinline__ float cubic_f(const float4 &c, float t) {
    return __builtin_fmaf(t, __builtin_fmaf(t, __builtin_fmaf(t, c.w, c.z), c.y), c.x);
}
// integer separation of two fixed-point coordinates -> float (units of box 2^-32)
__device__ __forceinline__ float sep_f(int a, int b) {
    return (float)(int)((uint32_t)a - (uint32_t)b);
}

// Quantities of one pair that both kernels share.
struct PairOut {
    float force;   // F/r in kcal/mol/A^2
    float dr6i;    // LJ r^-6 (0 beyond rcut2)
    float t;       // position inside the table bin
    int kb;        // table bin (tn = the zero bin)
    float ebelow;  // Coulomb energy per C qi qj below the table start (0 elsewhere)
    bool below;    // r below the table start
};

// Constants of the direct DSF evaluation, rounded to float once per kernel (uniform, so they
// stay in scalar registers).
struct DsfDirect {
    float al, al2, pa, a1, a2, a3, a4, a5, c2, fs, es;
};
__device__ __forceinline__ DsfDirect dsf_direct_consts(const DevParams &P) {
    DsfDirect D;
    D.al = (float)P.alpha;
    D.al2 = D.al * D.al;
    D.pa = (float)(P.ewald_p * P.alpha);
    D.a1 = (float)P.a1; D.a2 = (float)P.a2; D.a3 = (float)P.a3; D.a4 = (float)P.a4; D.a5 = (float)P.a5;
    D.c2 = (float)(2.0 * P.alpha * P.my_pis_inv);
    D.fs = (float)P.f_shift;
    D.es = (float)P.e_shift;
    return D;
}

// DSF Coulomb terms evaluated directly with the A&S 7.1.26 erfc (Source/mod_force_nonbond.f90:
// 160-183), for separations below the table start: (force over r, energy) per C qi qj.
// Inlined (a call would make the kernel keep its parameters on the stack).
__device__ __forceinline__ float2 dsf_direct(float r, float dri, const DsfDirect &D) {
    const float erfcd = expf(-D.al2 * (r * r));
    const float t = 1.0f / (1.0f + D.pa * r);
    const float erfcc = t * (D.a1 + t * (D.a2 + t * (D.a3 + t * (D.a4 + t * D.a5)))) * erfcd;
    const float g = (erfcc * dri + D.c2 * erfcd + r * D.fs) * (dri * dri);
    const float e = erfcc * dri - D.es - r * D.fs;
    return make_float2(g, e);
}

// Separations below the table start (u < 0) read the zero bin when ZB (their Coulomb
// energy then comes from the direct form alone); force-only kernels leave them on whatever
// in-range bin u gives.  DIRECT: such pairs take the direct form here; otherwise the caller
// sees o.below and evaluates them in a separate pass (keeping the direct form out of its hot
// loop, where its registers would cost occupancy).
template <bool ZB, bool DIRECT>
__device__ __forceinline__ PairOut pair_dsf(float dr2, float qq, const float4 &lj,
                                            const float4 *__restrict__ s_tab, float tinv,
                                            float u0, int tn, float rcut2, const DsfDirect &D) {
    PairOut o;
    const float dri = __builtin_amdgcn_rsqf(dr2);
    const float r = dr2 * dri;
    const float dr2i = dri * dri;
    float dr6i = dr2i * dr2i * dr2i;
    if (dr2 > rcut2) dr6i = 0.0f;
    const float flj = dr6i * __builtin_fmaf(lj.x, dr6i, -lj.y);
    const float u = __builtin_fmaf(r, tinv, u0);
    // r beyond the table (u >= tn) reads the zero bin; r below its start (u < 0) takes the
    // direct form (the reference evaluates DSF at every r < cut_coul)
    o.below = u < 0.0f;
    const int kb = (int)min((uint32_t)(int)u, (uint32_t)tn);
    o.kb = (ZB && o.below) ? tn : kb;
    o.t = __builtin_amdgcn_fractf(u);
    float gc = cubic_f(s_tab[o.kb], o.t);
    o.ebelow = 0.0f;
    if (DIRECT && o.below) {
        const float2 d = dsf_direct(r, dri, D);
        gc = d.x;
        o.ebelow = d.y;
    }
    o.force = __builtin_fmaf(qq, gc, flj * dr2i);
    o.dr6i = dr6i;
    return o;
}

constexpr uint32_t TYPE_BITS = 0x3fu;   // atom type in the low bits of X (types 1..63)

// x -> round(x 2^32 / box) mod 2^32 per replica, type in the low bits of X, charge in w.
// From the FP64 engine state (xd != nullptr) or from the float positions.
__global__ void k_fixpos(int32_t n, const double *__restrict__ xd, const float4 *__restrict__ pos4,
                         const float *__restrict__ q, const int32_t *__restrict__ rep,
                         const RepInfo *__restrict__ reps, int4 *__restrict__ pf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos4[i];
    const double s = 4294967296.0 / (double)reps[rep[i]].box;
    double x, y, z;
    if (xd) {
        x = xd[i]; y = xd[(size_t)n + i]; z = xd[2 * (size_t)n + i];
    } else {
        x = p.x; y = p.y; z = p.z;
    }
    const uint32_t fx = (uint32_t)(unsigned long long)__double2ll_rn(x * s);
    const int X = (int)((fx & ~TYPE_BITS) | ((uint32_t)__float_as_int(p.w) & TYPE_BITS));
    const int Y = (int)(uint32_t)(unsigned long long)__double2ll_rn(y * s);
    const int Z = (int)(uint32_t)(unsigned long long)__double2ll_rn(z * s);
    if (q) {
        pf[i] = make_int4(X, Y, Z, __float_as_int(q[i]));
    } else {   // charges by type: 12-byte records
        int *p3 = (int *)pf + 3 * (size_t)i;
        p3[0] = X; p3[1] = Y; p3[2] = Z;
    }
}

// L lanes per atom; lane s of a group walks list entries s, s+L, ... with the loads of the
// next entry in flight during this entry's math.  Blocks are dealt to the XCDs in runs.
template <int L, bool ENERGY, bool QT>
__global__ void __launch_bounds__(1024)
k_dsf_mixed(int32_t n, const int4 *__restrict__ pf,
            const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
            const int32_t *__restrict__ nlist, const int32_t *__restrict__ numneigh, int32_t na,
            const float4 *__restrict__ ljtab, const float4 *__restrict__ ptab,
            const float4 *__restrict__ tab, DevParams P,
            double *__restrict__ fnb, float *__restrict__ ff32, long long *__restrict__ eacc) {
    // QT: charges by atom type; 12-byte records (x, y, z) and a pair table (lj1, lj2,
    // C q_i q_j) in place of the charge in the record
    extern __shared__ float4 s_dyn[];
    const int nt2 = P.ntypes * P.ntypes;
    const int tn = P.tab_n;
    float4 *s_lj = s_dyn;
    float4 *s_pt = s_dyn + nt2;
    float4 *s_tab = s_dyn + (QT ? 2 : 1) * nt2;
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) {
        s_lj[t] = ljtab[t];
        if (QT) s_pt[t] = ptab[t];
    }
    for (int t = threadIdx.x; t < (ENERGY ? 2 : 1) * (tn + 1); t += blockDim.x) s_tab[t] = tab[t];
    __syncthreads();
    constexpr uint32_t RB = QT ? 12u : 16u;   // record bytes

    const int gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = gt / L;
    const int sub = gt % L;
    const bool live = i < n;
    const int ii = live ? i : n - 1;
    // the list buffer starts at the row of the wavefront's first atom (lane 0 has the
    // smallest index), so its 32-bit offsets stay small whatever the list size
    const int iw = __builtin_amdgcn_readfirstlane(ii);
    const auto lrs = rsrc(nlist + (size_t)iw * na, clamp_bytes((size_t)(n - iw) * na * 4));
    const auto prs = rsrc(pf, clamp_bytes((size_t)n * RB));

    int4 pi;
    if (QT) {
        const int *p3 = (const int *)pf + 3 * (size_t)ii;
        pi = make_int4(p3[0], p3[1], p3[2], 0);
    } else {
        pi = pf[ii];
    }
    const int trow = ((int)((uint32_t)pi.x & TYPE_BITS) - 1) * P.ntypes - 1;   // by jtype
    const float4 *ljrow = (QT ? s_pt : s_lj) + trow;
    const float4 *erow = s_lj + trow;
    const int xi = (int)((uint32_t)pi.x & ~TYPE_BITS);
    const int r = rep[ii];
    const float sbox = reps[r].box * TWO_M32;                 // box 2^-32 (exact)
    const float s2 = sbox * sbox;
    const float fq = sbox * (float)PAIR_QUANT;                // quanta per (raw dx * F/r)
    const int nn = live ? min(numneigh[ii], na) : 0;
    const float rcut2 = (float)P.rcut2;
    const float qc = P.coulpre * __int_as_float(pi.w);
    const float tinv = P.tab_inv_h, u0 = -P.tab_r0 * P.tab_inv_h;
    const DsfDirect D = dsf_direct_consts(P);
    const float4 *s_etab = s_tab + tn + 1;

    double fxq = 0.0, fyq = 0.0, fzq = 0.0;
    double pot = 0.0, ecl = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;

    // Software pipeline over a lane's entries k, k+L, k+2L, ... in slots u = 0..3 (entry
    // k + uL of a round of 4L).  A slot reloads into its own registers right after it is
    // consumed, so its next gather (4 entries ahead) and list index (8 entries ahead) are
    // in flight for the other three slots and no register is moved while a load is
    // outstanding; scheduling barriers keep the compiler from hoisting the reloads.
    // Entries past the end of the row gather the atom's own record (a cache hit, and a real
    // atom type) and are then dropped; list reads past the row re-read its first entry
    // instead of pulling the unused tail of the row from memory.
    const uint32_t step = 4u * L;
    const uint32_t row0 = (uint32_t)(ii - iw) * (uint32_t)na * 4u;
    uint32_t off = row0 + (uint32_t)sub * 4u;
    int4 PJ[4];
    uint32_t JJ[4];
    // prologue in the loop's own load order, so the wait counts at the loop head are the
    // steady-state ones
#pragma unroll
    for (int u = 0; u < 4; ++u) JJ[u] = ld32(lrs, off + u * step);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const uint32_t j0 = (sub + u * L < nn) ? JJ[u] : (uint32_t)ii;
        PJ[u] = QT ? ld96(prs, rec_off<QT>(j0)) : ld128(prs, rec_off<QT>(j0));
        JJ[u] = ld32(lrs, off + (u + 4) * step);
        __builtin_amdgcn_sched_barrier(0);
    }
    // One list entry on record pj.  First pass (direct = false): pairs below the table start
    // are left out and flagged in lowq; second pass (direct = true): only they count, with
    // the direct form.  Either way a pair's quanta are those of a single-pass evaluation.
    bool lowq = false;
    auto entry = [&](const int4 &pj, bool valid, auto direct) {
        constexpr bool DIR = decltype(direct)::value;
        const uint32_t tj = (uint32_t)pj.x & TYPE_BITS;
        const float dx = sep_f(xi, (int)((uint32_t)pj.x & ~TYPE_BITS));
        const float dy = sep_f(pi.y, pj.y), dz = sep_f(pi.z, pj.z);
        const float dr2 = norm2_f(dx, dy, dz) * s2;
        const float4 lj = ljrow[tj];
        const float qq = QT ? lj.z : qc * __int_as_float(pj.w);
        const PairOut o = pair_dsf<ENERGY, DIR>(dr2, qq, lj, s_tab, tinv, u0, tn, rcut2, D);
        const bool take = valid && (DIR ? o.below : !o.below);
        if (!DIR) lowq = lowq || (valid && o.below);
        // an empty asm pins the force as computed on every lane, so the compiler selects
        // instead of branching around the pair math (a branch per entry costs more)
        float force = o.force;
        __asm__("" : "+v"(force));
        const float fs = take ? force * fq : 0.0f;
        fxq += (double)__builtin_truncf(dx * fs);
        fyq += (double)__builtin_truncf(dy * fs);
        fzq += (double)__builtin_truncf(dz * fs);
        if (ENERGY) {
            if (take) {
                const float4 le = erow[tj];
                pot += (double)(o.dr6i * __builtin_fmaf(o.dr6i, le.z, -le.w));
                ecl += (double)(qq * (cubic_f(s_etab[o.kb], o.t) + o.ebelow));
                const float fv = o.force * s2;
                vx += (double)(dx * dx * fv);
                vy += (double)(dy * dy * fv);
                vz += (double)(dz * dz * fv);
            }
        }
    };

    off += 8u * step;
    for (int k = sub; k < nn; k += 4 * L) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            entry(PJ[u], k + u * L < nn, std::false_type{});
            const uint32_t jg = (k + (u + 4) * L < nn) ? JJ[u] : (uint32_t)ii;
            PJ[u] = QT ? ld96(prs, rec_off<QT>(jg)) : ld128(prs, rec_off<QT>(jg));
            JJ[u] = ld32(lrs, (k + (u + 8) * L < nn) ? off + u * step : row0);
            // VALU, SALU and LDS work may cross; memory loads stay in slot order
            __builtin_amdgcn_sched_barrier(0x0086);
        }
        off += 4u * step;
    }
    // pairs closer than the table start (none in a physical system): the lane walks its
    // entries again and evaluates those directly
    if (lowq) {
        for (int k = sub; k < nn; k += L) {
            const uint32_t j = ld32(lrs, row0 + 4u * (uint32_t)k);
            entry(QT ? ld96(prs, rec_off<QT>(j)) : ld128(prs, rec_off<QT>(j)), true, std::true_type{});
        }
    }

    const double qinv = 1.0 / (double)PAIR_QUANT;
    const double Fx = group_sum<L>(fxq) * qinv, Fy = group_sum<L>(fyq) * qinv,
                 Fz = group_sum<L>(fzq) * qinv;
    const bool owner = live && sub == 0;
    if (owner) {
        fnb[i] = Fx;
        fnb[(size_t)n + i] = Fy;
        fnb[2 * (size_t)n + i] = Fz;
        if (ff32) {
            ff32[3 * (size_t)i + 0] = (float)Fx;
            ff32[3 * (size_t)i + 1] = (float)Fy;
            ff32[3 * (size_t)i + 2] = (float)Fz;
        }
    }
    if (ENERGY) {
        pot = group_sum<L>(pot);
        ecl = group_sum<L>(ecl);
        vx = group_sum<L>(vx);
        vy = group_sum<L>(vy);
        vz = group_sum<L>(vz);
        wave_add(to_efix(0.5 * pot), owner, r, eacc, E_LJ);
        wave_add(to_efix(0.5 * ecl), owner, r, eacc, E_COUL);
        wave_add(to_efix(0.5 * vx), owner, r, eacc, V_X);
        wave_add(to_efix(0.5 * vy), owner, r, eacc, V_Y);
        wave_add(to_efix(0.5 * vz), owner, r, eacc, V_Z);
    }
}

// The same pair sum with in-cutoff compaction, one atom per wavefront.
//
// Phase A walks the atom's list 64 entries at a time (loads software pipelined two chunks
// ahead), forms the integer separation and appends the entries that can contribute to a
// per-wavefront ring in LDS (raw dx, dy, dz, q_j and type_j).  Phase B takes 64 queued
// entries at a time and runs pair_dsf with every lane busy, so no lane spends a pair beyond
// the cutoff (about 42 % of a list with a 2 A skin).  The ring filter keeps a margin of
// 1e-4 in r^2 over the cutoff: entries it keeps beyond the cutoff contribute exactly zero,
// and every entry it drops contributes exactly zero in k_dsf_mixed, so the sums agree.
constexpr int QRING = 128;

template <bool ENERGY>
__global__ void __launch_bounds__(256)
k_dsf_compact(int32_t n, const int4 *__restrict__ pf,
              const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
              const int32_t *__restrict__ nlist, const int32_t *__restrict__ numneigh, int32_t na,
              const float4 *__restrict__ ljtab, const float4 *__restrict__ tab, DevParams P,
              double *__restrict__ fnb, float *__restrict__ ff32, long long *__restrict__ eacc) {
    extern __shared__ float4 s_dyn[];
    const int nt2 = P.ntypes * P.ntypes;
    const int tn = P.tab_n;
    const int ntab = (ENERGY ? 2 : 1) * (tn + 1);
    float4 *s_lj = s_dyn;
    float4 *s_tab = s_dyn + nt2;
    float4 *s_ring = s_tab + ntab;
    int *s_rtype = (int *)(s_ring + 4 * QRING);
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) s_lj[t] = ljtab[t];
    for (int t = threadIdx.x; t < ntab; t += blockDim.x) s_tab[t] = tab[t];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // uniform across the wavefront, so the buffer descriptors below stay scalar
    const int i = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + wave);
    if (i >= n) return;   // whole wavefront; no barrier follows
    float4 *ring = s_ring + wave * QRING;
    int *rtype = s_rtype + wave * QRING;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    const auto lrs = rsrc(nlist + (size_t)i * na, clamp_bytes((size_t)(n - i) * na * 4));
    const auto prs = rsrc(pf, clamp_bytes((size_t)n * 16));
    const int4 pi = pf[i];
    const float4 *ljrow = s_lj + ((int)((uint32_t)pi.x & TYPE_BITS) - 1) * P.ntypes - 1;
    const int xi = (int)((uint32_t)pi.x & ~TYPE_BITS);
    const int r = rep[i];
    const float sbox = reps[r].box * TWO_M32;
    const float s2 = sbox * sbox;
    const float fq = sbox * (float)PAIR_QUANT;
    const int nn = min(numneigh[i], na);
    const float rcut2 = (float)P.rcut2;
    const float keep2 = fmaxf(rcut2, (float)P.cut_coulsq) * 1.0001f;
    const float qc = P.coulpre * __int_as_float(pi.w);
    const float tinv = P.tab_inv_h, u0 = -P.tab_r0 * P.tab_inv_h;
    const DsfDirect D = dsf_direct_consts(P);
    const float4 *s_etab = s_tab + tn + 1;

    double fxq = 0.0, fyq = 0.0, fzq = 0.0;
    double pot = 0.0, ecl = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    int head = 0, tail = 0;

    auto consume = [&](int cnt) {
        if (lane < cnt) {
            const int slot = (head + lane) & (QRING - 1);
            const float4 e = ring[slot];
            const float4 lj = ljrow[rtype[slot]];
            const float dx = e.x, dy = e.y, dz = e.z;
            const float dr2 = norm2_f(dx, dy, dz) * s2;
            const float qq = qc * e.w;
            const PairOut o = pair_dsf<ENERGY, true>(dr2, qq, lj, s_tab, tinv, u0, tn, rcut2, D);
            const float fs = o.force * fq;
            fxq += (double)__builtin_truncf(dx * fs);
            fyq += (double)__builtin_truncf(dy * fs);
            fzq += (double)__builtin_truncf(dz * fs);
            if (ENERGY) {
                pot += (double)(o.dr6i * __builtin_fmaf(o.dr6i, lj.z, -lj.w));
                ecl += (double)(qq * (cubic_f(s_etab[o.kb], o.t) + o.ebelow));
                const float fv = o.force * s2;
                vx += (double)(dx * dx * fv);
                vy += (double)(dy * dy * fv);
                vz += (double)(dz * dz * fv);
            }
        }
        head += cnt;
    };

    if (nn > 0) {
        // two-deep pipeline: chunk c uses positions gathered two chunks earlier; the index
        // load is issued before the gathers so waiting for it leaves the gathers in flight
        const uint32_t l4 = (uint32_t)lane * 4u;
        const uint32_t j0 = ld32(lrs, l4), j1 = ld32(lrs, 256u + l4);
        uint32_t j2 = ld32(lrs, 512u + l4);
        int4 p0 = ld128(prs, j0 * 16u), p1 = ld128(prs, j1 * 16u);
        for (int base = 0; base < nn; base += 64) {
            const int k = base + lane;
            const int4 pj = p0;
            p0 = p1;
            const uint32_t j3 = ld32(lrs, (uint32_t)(base + 192) * 4u + l4);
            p1 = ld128(prs, j2 * 16u);
            j2 = j3;
            const float dx = sep_f(xi, (int)((uint32_t)pj.x & ~TYPE_BITS));
            const float dy = sep_f(pi.y, pj.y), dz = sep_f(pi.z, pj.z);
            const bool in = k < nn && norm2_f(dx, dy, dz) * s2 <= keep2;
            const uint64_t m = __ballot(in);
            if (in) {
                const int slot = (tail + __popcll(m & lt_mask)) & (QRING - 1);
                ring[slot] = make_float4(dx, dy, dz, __int_as_float(pj.w));
                rtype[slot] = (int)((uint32_t)pj.x & TYPE_BITS);
            }
            tail += __popcll(m);
            __builtin_amdgcn_wave_barrier();
            if (tail - head >= 64) consume(64);
            __builtin_amdgcn_wave_barrier();
        }
        if (tail > head) consume(tail - head);
    }

    const double qinv = 1.0 / (double)PAIR_QUANT;
    const double Fx = group_sum<64>(fxq) * qinv, Fy = group_sum<64>(fyq) * qinv,
                 Fz = group_sum<64>(fzq) * qinv;
    const bool owner = lane == 0;
    if (owner) {
        fnb[i] = Fx;
        fnb[(size_t)n + i] = Fy;
        fnb[2 * (size_t)n + i] = Fz;
        if (ff32) {
            ff32[3 * (size_t)i + 0] = (float)Fx;
            ff32[3 * (size_t)i + 1] = (float)Fy;
            ff32[3 * (size_t)i + 2] = (float)Fz;
        }
    }
    if (ENERGY) {
        pot = group_sum<64>(pot);
        ecl = group_sum<64>(ecl);
        vx = group_sum<64>(vx);
        vy = group_sum<64>(vy);
        vz = group_sum<64>(vz);
        wave_add(to_efix(0.5 * pot), owner, r, eacc, E_LJ);
        wave_add(to_efix(0.5 * ecl), owner, r, eacc, E_COUL);
        wave_add(to_efix(0.5 * vx), owner, r, eacc, V_X);
        wave_add(to_efix(0.5 * vy), owner, r, eacc, V_Y);
        wave_add(to_efix(0.5 * vz), owner, r, eacc, V_Z);
    }
}

template <int L>
void launch_lanes(sofia_ctx *c, bool energies, float *ff32, bool qt) {
    const int block = c->block;
    const long long threads = (long long)c->N * L;
    const int grid = (int)((threads + block - 1) / block);
    const int nt2 = c->dp.ntypes * c->dp.ntypes;
    const size_t lds = sizeof(float4) * ((qt ? 2 : 1) * nt2 + (energies ? 2 : 1) * (c->dp.tab_n + 1));
#define DSF_LAUNCH(E, Q)                                                                          \
    hipLaunchKernelGGL((k_dsf_mixed<L, E, Q>), dim3(grid), dim3(block), lds, c->stream, c->N,       \
                       c->pfix, c->rep, c->reps_d, c->nlist, c->numneigh, c->neigh_alloc,          \
                       c->ljtab, c->ptab, c->dsf_tab, c->dp, c->fnb, ff32, c->eacc)
    if (energies) {
        if (qt) DSF_LAUNCH(true, true); else DSF_LAUNCH(true, false);
    } else {
        if (qt) DSF_LAUNCH(false, true); else DSF_LAUNCH(false, false);
    }
#undef DSF_LAUNCH
}

void launch_compact(sofia_ctx *c, bool energies, float *ff32) {
    const int grid = (c->N + 3) / 4;
    const int nt2 = c->dp.ntypes * c->dp.ntypes;
    const size_t lds = sizeof(float4) * (nt2 + (energies ? 2 : 1) * (c->dp.tab_n + 1) + 4 * QRING) +
                       sizeof(int) * 4 * QRING;
    if (energies)
        hipLaunchKernelGGL(k_dsf_compact<true>, dim3(grid), dim3(256), lds, c->stream, c->N,
                           c->pfix, c->rep, c->reps_d, c->nlist, c->numneigh,
                           c->neigh_alloc, c->ljtab, c->dsf_tab, c->dp, c->fnb, ff32, c->eacc);
    else
        hipLaunchKernelGGL(k_dsf_compact<false>, dim3(grid), dim3(256), lds, c->stream, c->N,
                           c->pfix, c->rep, c->reps_d, c->nlist, c->numneigh,
                           c->neigh_alloc, c->ljtab, c->dsf_tab, c->dp, c->fnb, ff32, c->eacc);
}

}  // namespace

void launch_dsf_mixed(sofia_ctx *c, bool energies, float *ff32, bool from_xd) {
    // 12-byte records while every atom of a type carries the type's charge (the compacting
    // kernel reads 16-byte records)
    const bool qt = c->q_by_type && c->ptab && c->lanes != 0;
    hipLaunchKernelGGL(k_fixpos, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N,
                       from_xd ? c->a.xd : nullptr, c->a.pos4, qt ? nullptr : c->a.q, c->rep,
                       c->reps_d, c->pfix);
    switch (c->lanes) {
        case 0: launch_compact(c, energies, ff32); break;
        case 4: launch_lanes<4>(c, energies, ff32, qt); break;
        case 8: launch_lanes<8>(c, energies, ff32, qt); break;
        case 32: launch_lanes<32>(c, energies, ff32, qt); break;
        case 64: launch_lanes<64>(c, energies, ff32, qt); break;
        default: launch_lanes<16>(c, energies, ff32, qt); break;
    }
}

}  // namespace sofia
