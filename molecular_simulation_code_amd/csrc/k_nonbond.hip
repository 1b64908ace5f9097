// Nonbonded pair kernels over the full ("nonewton") Verlet list, gfx950.
//
//   mode 0  DSF, FP64 pair arithmetic        lj_cut_coul_dsf_nonewton_NOMIC/_MIC
//                                            (Source/mod_force_nonbond.f90:105-196 / 8-101)
//   mode 1  DSF, FP32 pair arithmetic with exact per-atom force sums (k_dsf_mixed.hip)
//   mode 2  Ewald real space, FP32 pair math lj_cut_coul_long_nonewton_NOMIC (:300-400)
//
// Pair expressions (Source/mod_force_nonbond.f90:153-183):
//   dr6i = (1/r^2)^3 (0 beyond rcut2);  F_LJ*r = dr6i (lj1 dr6i - lj2);  E_LJ = dr6i (lj3 dr6i - lj4)
//   pref = C qi qj / r (0 beyond cut_coulsq);  erfc(alpha r) by Abramowitz & Stegun 7.1.26
//   F_C*r = pref (erfc/r + 2 alpha/sqrt(pi) exp(-alpha^2 r^2) + r f_shift) r
//   E_C = pref (erfc - r e_shift - r^2 f_shift);  f_i += dx (F_C + F_LJ)/r^2
// Energies are halved (full list).  The virial follows the PME kernel's convention,
// 0.5 sum dx^2 F/r (the reference DSF kernel computes none).
//
// Layout: L lanes cooperate on one atom i (L = 4..64, chosen by the autotuner); lane s
// walks list entries s, s+L, ... so each group reads a contiguous run of the atom's list,
// and the L partial sums are combined with cross-lane shuffles.  The loop is software
// pipelined (the next entry's position and charge are in flight during this entry's math),
// and blocks are dealt to the XCDs in runs so that an XCD's L2 holds one region of atoms.
// Positions are the cell-sorted float4 (x, y, z, type) array, so the j gathers of
// neighbouring atoms hit the same cache lines.  No atomics on forces: every atom's sum is
// owned by one group.
#include "sofia_device.h"

namespace sofia {
namespace {

// exp(x) for the FP64 DSF Gaussian (x <= 0), table-driven (Tang's method): x = (64 n + k)
// ln2/64 + r, |r| <= ln2/128, with ln2/64 split in two parts (the first with 12 trailing zero
// bits, so n ln2/64 is exact for |x| < 44), and exp(x) = 2^n T[k] (1 + p(r)), T[k] = 2^(k/64)
// from LDS and p the degree-6 Taylor polynomial of e^r - 1 (truncation below 1e-19).  Within
// 1 ulp of the C library's exp over [-60, 0] (4.4e7 points checked on the host).
__device__ inline double exp_tab(double x, const double *__restrict__ T) {
    const double N = __builtin_rint(x * 92.33248261689366);   // 64 / ln 2
    const int m = (int)N;
    double r = __builtin_fma(N, -0.010830424696244734, x);
    r = __builtin_fma(N, -4.411764150473684e-15, r);
    double p = __builtin_fma(r, 1.0 / 720.0, 1.0 / 120.0);
    p = __builtin_fma(r, p, 1.0 / 24.0);
    p = __builtin_fma(r, p, 1.0 / 6.0);
    p = __builtin_fma(r, p, 0.5);
    p = __builtin_fma(r, p, 1.0);
    p = p * r;
    const double t = T[m & 63];
    return __builtin_ldexp(__builtin_fma(t, p, t), m >> 6);
}

template <int L, int MODE, bool ENERGY>
__global__ void __launch_bounds__(256)
k_nonbond(int32_t n, const float4 *__restrict__ pos4, const float *__restrict__ q,
          const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
          const int32_t *__restrict__ nlist, const int32_t *__restrict__ numneigh, int32_t na,
          const float4 *__restrict__ ljtab, const float4 *__restrict__ tab, DevParams P,
          double *__restrict__ fnb, float *__restrict__ ff32, long long *__restrict__ eacc) {
    // LDS, sized at launch: the LJ table (FP64 mode: as doubles, then the 64 values 2^(k/64)
    // of exp_tab)
    static_assert(MODE == 0 || MODE == 2, "mode 1 is k_dsf_mixed");
    extern __shared__ float4 s_dyn[];
    float4 *s_lj = s_dyn;
    double2 *s_lj64 = (double2 *)s_dyn;                    // 2 per type pair: (lj1, lj2), (lj3, lj4)
    const int nt2 = P.ntypes * P.ntypes;
    double *s_exp = (double *)(s_lj64 + 2 * nt2);
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) {
        const float4 v = ljtab[t];
        if (MODE == 0) {
            s_lj64[2 * t] = make_double2(v.x, v.y);
            s_lj64[2 * t + 1] = make_double2(v.z, v.w);
        } else {
            s_lj[t] = v;
        }
    }
    if (MODE == 0 && threadIdx.x < 64) s_exp[threadIdx.x] = exp2((double)threadIdx.x / 64.0);
    __syncthreads();

    const int gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = gt / L;
    const int sub = gt % L;
    const bool live = i < n;
    const int ii = live ? i : n - 1;

    const float4 xi = pos4[ii];
    const int ti = __float_as_int(xi.w) - 1;
    const int r = rep[ii];
    const float box = reps[r].box, ibox = reps[r].ibox;
    const int nn = live ? min(numneigh[ii], na) : 0;   // a row never reads past its stride
    const int32_t *lst = nlist + (size_t)ii * na;
    const float4 *ljrow = s_lj + ti * P.ntypes - 1;   // index by jtype (1-based)
    const double2 *ljrow64 = s_lj64 + 2 * (ti * P.ntypes - 1);

    double fxd = 0.0, fyd = 0.0, fzd = 0.0;
    double pot = 0.0, ecl = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;

    if (MODE == 0) {
        const double x1 = xi.x, y1 = xi.y, z1 = xi.z;
        const double bx = box, ib = ibox;
        const double qc = (double)P.coulpre * (double)q[ii];
        const double two_a_pis = (2.0 * P.alpha) * P.my_pis_inv;
        const double a2 = P.alpha * P.alpha, pa = P.ewald_p * P.alpha;
        // software pipeline: an entry's data is loaded one iteration ahead; loads past the
        // end of the row re-read its last entry (clamped index), so no branch is needed
        const int last = max(nn - 1, 0);
        const uint32_t *ulst = (const uint32_t *)lst;
        const uint32_t j0 = (nn > 0) ? ulst[min(sub, last)] : (uint32_t)ii;
        uint32_t jnn = (nn > 0) ? ulst[min(sub + L, last)] : (uint32_t)ii;
        float4 pn = pos4[j0];
        float qn = q[j0];
#pragma unroll 2
        for (int k = sub; k < nn; k += L) {
            const float4 pj = pn;
            const float qj = qn;
            pn = pos4[jnn];
            qn = q[jnn];
            jnn = ulst[min(k + 2 * L, last)];
            double dx = x1 - (double)pj.x, dy = y1 - (double)pj.y, dz = z1 - (double)pj.z;
            dx -= bx * __builtin_rint(dx * ib);
            dy -= bx * __builtin_rint(dy * ib);
            dz -= bx * __builtin_rint(dz * ib);
            const double dr2 = dx * dx + dy * dy + dz * dz;
            const double dri = rsq_f64(dr2);
            const double r = dr2 * dri;
            const double dr2i = dri * dri;
            double dr6i = dr2i * dr2i * dr2i;
            if (dr2 > P.rcut2) dr6i = 0.0;
            const int tj2 = 2 * __float_as_int(pj.w);
            const double2 l12 = ljrow64[tj2];
            const double flj = dr6i * (l12.x * dr6i - l12.y);
            double pref = qc * (double)qj * dri;
            if (dr2 > P.cut_coulsq) pref = 0.0;
            const double erfcd = exp_tab(-a2 * dr2, s_exp);
            const double t = rcp_f64(1.0 + pa * r);
            const double erfcc =
                t * (P.a1 + t * (P.a2 + t * (P.a3 + t * (P.a4 + t * P.a5)))) * erfcd;
            const double fc = pref * (erfcc * dri + two_a_pis * erfcd + r * P.f_shift) * r;
            const double force = (fc + flj) * dr2i;
            fxd += dx * force;
            fyd += dy * force;
            fzd += dz * force;
            if (ENERGY) {
                const double2 l34 = ljrow64[tj2 + 1];
                pot += dr6i * (dr6i * l34.x - l34.y);
                ecl += pref * (erfcc - r * P.e_shift - dr2 * P.f_shift);
                vx += dx * dx * force;
                vy += dy * dy * force;
                vz += dz * dz * force;
            }
        }
    } else {
        const float rcut2 = (float)P.rcut2, cut_coulsq = (float)P.cut_coulsq;
        const float qc = P.coulpre * q[ii];
        const float alpha = (float)P.gewald;
        const float nal2e = -alpha * alpha * 1.4426950408889634f;   // exp(-a^2 r^2) = 2^(..)
        const float pa = (float)P.ewald_p * alpha;
        const float c2 = (float)P.ewald_f;
        const float A1 = (float)P.a1, A2 = (float)P.a2, A3 = (float)P.a3, A4 = (float)P.a4,
                    A5 = (float)P.a5;
        // software pipeline: an entry's data is loaded one iteration ahead; loads past the
        // end of the row re-read its last entry (clamped index), so no branch is needed
        const int last = max(nn - 1, 0);
        const uint32_t *ulst = (const uint32_t *)lst;
        const uint32_t j0 = (nn > 0) ? ulst[min(sub, last)] : (uint32_t)ii;
        uint32_t jnn = (nn > 0) ? ulst[min(sub + L, last)] : (uint32_t)ii;
        float4 pn = pos4[j0];
        float qn = q[j0];
#pragma unroll 2
        for (int k = sub; k < nn; k += L) {
            const float4 pj = pn;
            const float qj = qn;
            pn = pos4[jnn];
            qn = q[jnn];
            jnn = ulst[min(k + 2 * L, last)];
            float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
            dx -= box * __builtin_rintf(dx * ibox);
            dy -= box * __builtin_rintf(dy * ibox);
            dz -= box * __builtin_rintf(dz * ibox);
            const float dr2 = dx * dx + dy * dy + dz * dz;
            const float dri = __builtin_amdgcn_rsqf(dr2);
            const float r = dr2 * dri;
            const float dr2i = dri * dri;
            float dr6i = dr2i * dr2i * dr2i;
            if (dr2 > rcut2) dr6i = 0.0f;
            const float4 lj = ljrow[__float_as_int(pj.w)];
            const float flj = dr6i * (lj.x * dr6i - lj.y);
            float pref = qc * qj * dri;
            if (dr2 > cut_coulsq) pref = 0.0f;
            const float erfcd = __builtin_amdgcn_exp2f(nal2e * dr2);
            const float t = __builtin_amdgcn_rcpf(1.0f + pa * r);
            const float erfcc = t * (A1 + t * (A2 + t * (A3 + t * (A4 + t * A5)))) * erfcd;
            const float fc = pref * (erfcc + c2 * (alpha * r) * erfcd);   // F_C * r
            const float force = (fc + flj) * dr2i;
            fxd += (double)(dx * force);
            fyd += (double)(dy * force);
            fzd += (double)(dz * force);
            if (ENERGY) {
                pot += (double)(dr6i * (dr6i * lj.z - lj.w));
                ecl += (double)(pref * erfcc);
                vx += (double)(dx * dx * force);
                vy += (double)(dy * dy * force);
                vz += (double)(dz * dz * force);
            }
        }
    }

    const double Fx = group_sum<L>(fxd), Fy = group_sum<L>(fyd), Fz = group_sum<L>(fzd);
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

template <int L, int MODE>
void launch_t(sofia_ctx *c, bool energies, float *ff32) {
    const int block = std::min(c->block, 256);   // this kernel's launch bound
    const long long threads = (long long)c->N * L;
    const int grid = (int)((threads + block - 1) / block);
    const int nt2 = c->dp.ntypes * c->dp.ntypes;
    const size_t lds = MODE == 0 ? 2 * sizeof(double2) * nt2 + 64 * sizeof(double)
                                 : sizeof(float4) * nt2;
    if (energies)
        hipLaunchKernelGGL((k_nonbond<L, MODE, true>), dim3(grid), dim3(block), lds, c->stream, c->N,
                           c->a.pos4, c->a.q, c->rep, c->reps_d, c->nlist, c->numneigh,
                           c->neigh_alloc, c->ljtab, c->dsf_tab, c->dp, c->fnb, ff32, c->eacc);
    else
        hipLaunchKernelGGL((k_nonbond<L, MODE, false>), dim3(grid), dim3(block), lds, c->stream, c->N,
                           c->a.pos4, c->a.q, c->rep, c->reps_d, c->nlist, c->numneigh,
                           c->neigh_alloc, c->ljtab, c->dsf_tab, c->dp, c->fnb, ff32, c->eacc);
}

template <int MODE>
void launch_mode(sofia_ctx *c, bool energies, float *ff32) {
    switch (c->lanes) {
        case 4: launch_t<4, MODE>(c, energies, ff32); break;
        case 8: launch_t<8, MODE>(c, energies, ff32); break;
        case 32: launch_t<32, MODE>(c, energies, ff32); break;
        case 64: launch_t<64, MODE>(c, energies, ff32); break;
        default: launch_t<16, MODE>(c, energies, ff32); break;
    }
}

}  // namespace

// Drop-in launches (positions from the float copy the caller handed over).
void launch_nonbond_out(sofia_ctx *c, int mode, bool energies, float *ff32) {
    if (mode == 0) launch_mode<0>(c, energies, ff32);
    else if (mode == 1) launch_dsf_mixed(c, energies, ff32, false);
    else launch_mode<2>(c, energies, ff32);
}

// Engine launches: the mixed kernel takes its fixed-point positions from the FP64 state.
void launch_nonbond(sofia_ctx *c, int mode, bool energies) {
    if (mode == 1) launch_dsf_mixed(c, energies, nullptr, c->engine_ready);
    else launch_nonbond_out(c, mode, energies, nullptr);
}

// Number of list entries inside the force cutoff (the algorithmic pair count of a launch).
__global__ void k_count_pairs(int32_t n, const float4 *__restrict__ pos4, const int32_t *__restrict__ rep,
                              const RepInfo *__restrict__ reps, const int32_t *__restrict__ nlist,
                              const int32_t *__restrict__ numneigh, int32_t na, float rc2,
                              unsigned long long *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt = 0;
    if (i < n) {
        const float4 xi = pos4[i];
        const RepInfo &R = reps[rep[i]];
        const int32_t *lst = nlist + (size_t)i * na;
        const int nn = min(numneigh[i], na);
        for (int k = 0; k < nn; ++k) {
            const float4 pj = pos4[lst[k]];
            float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
            dx -= R.box * __builtin_rintf(dx * R.ibox);
            dy -= R.box * __builtin_rintf(dy * R.ibox);
            dz -= R.box * __builtin_rintf(dz * R.ibox);
            cnt += (dx * dx + dy * dy + dz * dz <= rc2) ? 1 : 0;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, cnt);
}

void launch_count_pairs(sofia_ctx *c, unsigned long long *out) {
    const float rc2 = (float)(c->dp.rcut2 > c->dp.cut_coulsq ? c->dp.rcut2 : c->dp.cut_coulsq);
    hipLaunchKernelGGL(k_count_pairs, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N,
                       c->a.pos4, c->rep, c->reps_d, c->nlist, c->numneigh, c->neigh_alloc, rc2, out);
}

}  // namespace sofia
