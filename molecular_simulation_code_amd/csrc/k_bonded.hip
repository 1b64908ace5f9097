// Bonded-pair kernels: Ewald exclusion corrections, the scaled 1-4 DSF pair, and the
// harmonic bond / harmonic angle / cosine torsion / harmonic improper terms of the full MD
// step.
//
// One thread per list item; forces go to a per-atom int64 fixed-point accumulator with
// atomics (2^-32 kcal/mol/A quanta), so totals are exact and independent of thread
// order.  Energies go to per-replica int64 slots the same way.
//
// Drop-in variants read the real*4 positions and follow the reference's precision
// (real*4 coordinate difference and charge product, double for the rest); the engine
// variants read the FP64 positions of the device-resident state.
#include "sofia_device.h"

namespace sofia {
namespace {

__device__ inline double erfc_as(double t, const DevParams &P) {
    return t * (P.a1 + t * (P.a2 + t * (P.a3 + t * (P.a4 + t * P.a5))));
}

// correct_12/13/14_bonds, Source/mod_force_correction.f90:23-86 (and 112-177, 205-272)
__global__ void k_correct(int32_t items, int32_t width, int32_t colb, const int32_t *__restrict__ lst,
                          const float4 *__restrict__ pos4, const float *__restrict__ q,
                          const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                          DevParams P, double fudge, int32_t n, long long *__restrict__ fout,
                          long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items;
    double e = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    int r = 0;
    if (live) {
        const int i1 = lst[(size_t)m * width], i2 = lst[(size_t)m * width + colb];
        r = rep[i1];
        const double bx = reps[r].box, ib = reps[r].ibox;
        const float4 p1 = pos4[i1], p2 = pos4[i2];
        double dx = (double)(p1.x - p2.x), dy = (double)(p1.y - p2.y), dz = (double)(p1.z - p2.z);
        dx -= bx * rint(dx * ib);
        dy -= bx * rint(dy * ib);
        dz -= bx * rint(dz * ib);
        const double dr2 = dx * dx + dy * dy + dz * dz;
        const double dr2i = 1.0 / dr2;
        const double rr = sqrt(dr2);
        double pref = (double)((P.coulpre * q[i1]) * q[i2]) / rr;
        const double grij = P.gewald * rr;
        const double expm2 = exp(-grij * grij);
        const double t = 1.0 / (1.0 + P.ewald_p * grij);
        const double erfc = erfc_as(t, P) * expm2;
        if (dr2 > P.cut_coulsq) pref = 0.0;
        const double fc = pref * (erfc + P.ewald_f * grij * expm2) - (1.0 - fudge) * pref;
        e = pref * erfc - (1.0 - fudge) * pref;
        const double force = fc * dr2i;
        atomic_fix(&fout[i1], dx * force);
        atomic_fix(&fout[(size_t)n + i1], dy * force);
        atomic_fix(&fout[2 * (size_t)n + i1], dz * force);
        atomic_fix(&fout[i2], -dx * force);
        atomic_fix(&fout[(size_t)n + i2], -dy * force);
        atomic_fix(&fout[2 * (size_t)n + i2], -dz * force);
        vx = dx * dx * force;
        vy = dy * dy * force;
        vz = dz * dz * force;
    }
    wave_add(to_efix(e), live, r, eacc, E_CORR);
    wave_add(to_efix(vx), live, r, eacc, V_X);
    wave_add(to_efix(vy), live, r, eacc, V_Y);
    wave_add(to_efix(vz), live, r, eacc, V_Z);
}

// Term evaluators of the engine's bonded work.  Each reads the atoms of its list item, hands
// the fixed-point force on the item's atom column col (0..3) to add(col, k, q), k the
// component, and returns its energies.  The per-item kernels below add to the global
// accumulator with atomics; k_bonded_bundles adds to a workgroup's copy in LDS.  The same
// code forms every quantum either way, so both give the same integer sums.

// The 1-4 pair of dihedral_opls_DSF, Source/mod_force_dihed.f90:887-895 and 1003-1048:
// nint minimum image, 0.5-scaled DSF Coulomb inside cut_coulsq, lj14 tables inside rcut2.
// Item: a dihedral (i1, i2, i3, i4, type); the pair is (i1, i4).
template <bool F64POS, class Add>
__device__ inline void pair14_eval(const int32_t *it, const float4 *__restrict__ pos4,
                                   const double *__restrict__ xd, const float *__restrict__ q,
                                   float box, float ibox, const double4 *__restrict__ lj14,
                                   const DevParams &P, int32_t n, double &ec, double &el, Add add) {
    const int i1 = it[0], i4 = it[3];
    double d[3];
    if (F64POS) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double v = xd[(size_t)c * n + i1] - xd[(size_t)c * n + i4];
            d[c] = v - (double)box * rint(v * (double)ibox);
        }
    } else {
        const float4 p1 = pos4[i1], p4 = pos4[i4];
        const float df[3] = {p1.x - p4.x, p1.y - p4.y, p1.z - p4.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // Source/mod_minimg.f90: dx - box*nint(ibox*dx), box*nint a real*4 product
            const double v = (double)df[c];
            const float shift = box * (float)(int)round((double)ibox * v);
            d[c] = v - (double)shift;
        }
    }
    const double dr2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float4 pi1 = pos4[i1], pi4 = pos4[i4];
    const int ti = __float_as_int(pi1.w) - 1, tj = __float_as_int(pi4.w) - 1;
    double force = 0.0;
    ec = 0.0;
    el = 0.0;
    if (dr2 < P.cut_coulsq) {
        const double rr = sqrt(dr2);
        const double dr2i = 1.0 / dr2;
        const double pref = (double)((P.coulpre * q[i1]) * q[i4]) / rr;
        const double erfcd = exp(-P.alpha * P.alpha * rr * rr);
        const double t = 1.0 / (1.0 + P.ewald_p * P.alpha * rr);
        const double erfcc = erfc_as(t, P) * erfcd;
        const double fc =
            0.5 * pref * (erfcc / rr + 2.0 * P.alpha / P.my_pis * erfcd + rr * P.f_shift) * rr;
        ec = 0.5 * pref * (erfcc - rr * P.e_shift - dr2 * P.f_shift);
        double flj = 0.0;
        if (dr2 < P.rcut2) {
            const double dr6i = dr2i * dr2i * dr2i;
            const double4 c = lj14[ti * P.ntypes + tj];
            flj = dr6i * (c.x * dr6i - c.y);
            el = dr6i * (dr6i * c.z - c.w);
        }
        force = (fc + flj) * dr2i;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        add(0, c, to_fix(d[c] * force));
        add(3, c, to_fix(-d[c] * force));
    }
}

__device__ inline void dvec64(const double *xd, int n, int a, int b, double box, double ibox, double *o) {
    dvec_soa(xd, n, a, b, box, ibox, o);
}

// bond_harmonic, Source/mod_force_bond.f90:24-45: E = k (r - r0)^2.  Item (i1, i2, type).
template <class Add>
__device__ inline long long bond_eval(const int32_t *it, const double *__restrict__ xd, double box,
                                      double ibox, const double *__restrict__ kb,
                                      const double *__restrict__ r0, int32_t n, Add add) {
    const int t = it[2] - 1;
    long long f[3];
    const long long e = bond_term(xd, n, it[0], it[1], box, ibox, kb[t], r0[t], f);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        add(0, c, f[c]);
        add(1, c, -f[c]);
    }
    return e;
}

// angle_harmonic, Source/mod_force_angle.f90:29-90: E = k (theta - theta0)^2.  Item
// (i1, i2 apex, i3, type).
template <class Add>
__device__ inline long long angle_eval(const int32_t *it, const double *__restrict__ xd, double box,
                                       double ibox, const double *__restrict__ ka,
                                       const double *__restrict__ th0, int32_t n, Add add) {
    const int t = it[3] - 1;
    long long f1[3], f3[3];
    const long long e = angle_term(xd, n, it[0], it[1], it[2], box, ibox, ka[t], th0[t], f1, f3);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        add(0, k, f1[k]);
        add(1, k, -f1[k] - f3[k]);
        add(2, k, f3[k]);
    }
    return e;
}

// Torsion E = K (1 + cos(n phi - shift)) (the AMBER form of Source/mod_force_dihed_amber.f90)
// with the Blondel-Karplus gradient of phi.  For an integer multiplicity 1..6 the angle
// functions come from cos(phi), sin(phi) (dot and triple products) by the multiple-angle
// recurrence and a per-type (cos shift, sin shift): no FP64 atan2/cos/sin per torsion.
// Item (a, b, c, d, type).
template <class Add>
__device__ inline double torsion_eval(const int32_t *it, const double *__restrict__ xd, double bx,
                                      double ib, const double *__restrict__ kk,
                                      const double *__restrict__ mult,
                                      const double *__restrict__ shift_deg,
                                      const double *__restrict__ shift_cs, int32_t n, Add add) {
    const int a = it[0], b = it[1], c = it[2], d = it[3], t = it[4] - 1;
    double F[3], G[3], H[3];
    dvec64(xd, n, a, b, bx, ib, F);
    dvec64(xd, n, b, c, bx, ib, G);
    dvec64(xd, n, d, c, bx, ib, H);
    const double A[3] = {F[1] * G[2] - F[2] * G[1], F[2] * G[0] - F[0] * G[2], F[0] * G[1] - F[1] * G[0]};
    const double B[3] = {H[1] * G[2] - H[2] * G[1], H[2] * G[0] - H[0] * G[2], H[0] * G[1] - H[1] * G[0]};
    const double A2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    if (!(A2 > 0.0 && B2 > 0.0 && G2 > 0.0)) return 0.0;
    const double Gn = sqrt(G2);
    const double BxA[3] = {B[1] * A[2] - B[2] * A[1], B[2] * A[0] - B[0] * A[2], B[0] * A[1] - B[1] * A[0]};
    const double y = (BxA[0] * G[0] + BxA[1] * G[1] + BxA[2] * G[2]) / Gn;
    const double x = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    const double nn = mult[t];
    const int ni = (int)nn;
    double cnd, snd;                     // cos(n phi - shift), sin(n phi - shift)
    if ((double)ni == nn && ni >= 1 && ni <= 6) {
        const double inv = 1.0 / sqrt(A2 * B2);   // |A||B| = sqrt(x^2 + y^2)
        const double c1 = x * inv, s1 = y * inv;
        double cn = c1, sn = s1;
        for (int k = 1; k < ni; ++k) {
            const double c2 = cn * c1 - sn * s1;
            sn = sn * c1 + cn * s1;
            cn = c2;
        }
        const double cd = shift_cs[2 * t], sd = shift_cs[2 * t + 1];
        cnd = cn * cd + sn * sd;
        snd = sn * cd - cn * sd;
    } else {
        const double ph = nn * atan2(y, x) - shift_deg[t] * (3.141592653589793 / 180.0);
        cnd = cos(ph);
        snd = sin(ph);
    }
    const double e = kk[t] * (1.0 + cnd);
    const double dEdphi = -kk[t] * nn * snd;
    const double FG = F[0] * G[0] + F[1] * G[1] + F[2] * G[2];
    const double HG = H[0] * G[0] + H[1] * G[1] + H[2] * G[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double dpa = -Gn / A2 * A[k];
        const double dpd = Gn / B2 * B[k];
        const double dpb = -dpa - (FG / G2) * dpa - (HG / G2) * dpd;
        const double dpc = -dpd + (FG / G2) * dpa + (HG / G2) * dpd;
        add(0, k, to_fix(-dEdphi * dpa));
        add(1, k, to_fix(-dEdphi * dpb));
        add(2, k, to_fix(-dEdphi * dpc));
        add(3, k, to_fix(-dEdphi * dpd));
    }
    return e;
}

// improper_harmonic_lammps14, Source/mod_force_dihed.f90:1546-1724: E = K (chi - chi0)^2
// with c = cos(chi) from the bond-vector cosines c0, c1, c2 and the LAMMPS improper_harmonic
// gradient coefficients a11..a33; SMALL is the reference's real*4 literal 0.001.
// Item (i1, i2, i3, i4, type).
template <class Add>
__device__ inline double improper_eval(const int32_t *it, const double *__restrict__ xd, double bx,
                                       double ib, const double *__restrict__ kk,
                                       const double *__restrict__ chi0, int32_t n, Add add) {
    constexpr double SMALL = (double)0.001f;
    const int i1 = it[0], i2 = it[1], i3 = it[2], i4 = it[3], t = it[4] - 1;
    double v1[3], v2[3], v3[3];
    dvec64(xd, n, i1, i2, bx, ib, v1);
    dvec64(xd, n, i3, i2, bx, ib, v2);
    dvec64(xd, n, i4, i3, bx, ib, v3);
    const double ss1 = 1.0 / (v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
    const double ss2 = 1.0 / (v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
    const double ss3 = 1.0 / (v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
    const double r1 = sqrt(ss1), r2 = sqrt(ss2), r3 = sqrt(ss3);
    const double c0 = (v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2]) * r1 * r3;
    const double c1 = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) * r1 * r2;
    const double c2 = -(v3[0] * v2[0] + v3[1] * v2[1] + v3[2] * v2[2]) * r3 * r2;
    const double s1 = 1.0 / fmax(1.0 - c1 * c1, SMALL);
    const double s2 = 1.0 / fmax(1.0 - c2 * c2, SMALL);
    double s12 = sqrt(s1 * s2);
    double c = fmin(fmax((c1 * c2 + c0) * s12, -1.0), 1.0);
    const double s = fmax(sqrt(1.0 - c * c), SMALL);
    const double domega = acos(c) - chi0[t];
    double a = kk[t] * domega;
    const double e = a * domega;
    a = -a * 2.0 / s;
    c *= a;
    s12 *= a;
    const double a11 = c * ss1 * s1;
    const double a22 = -ss2 * (2.0 * c0 * s12 - c * (s1 + s2));
    const double a33 = c * ss3 * s2;
    const double a12 = -r1 * r2 * (c1 * c * s1 + c2 * s12);
    const double a13 = -r1 * r3 * s12;
    const double a23 = r2 * r3 * (c2 * c * s2 + c1 * s12);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double sx2 = a22 * v2[k] + a23 * v3[k] + a12 * v1[k];
        const double f1 = a12 * v2[k] + a13 * v3[k] + a11 * v1[k];
        const double f4 = a23 * v2[k] + a33 * v3[k] + a13 * v1[k];
        add(0, k, to_fix(f1));
        add(1, k, to_fix(-sx2 - f1));
        add(2, k, to_fix(sx2 - f4));
        add(3, k, to_fix(f4));
    }
    return e;
}

// impro_amber_force, Source/mod_force_impro_amber.f90:8-230 (cvff form):
// E = K [1 + d cos(n phi)], cos(n phi) the Chebyshev polynomial T_n(c) of c = cos(phi),
// dE/dc = K d T_n'(c); the gradient of c uses the same a11..a33 coefficients as
// improper_eval, with the bond-angle sines clamped from below by SMALL.  The energy factor
// is kept in FP64 here (the reference stores it in a real*4 variable).
template <class Add>
__device__ inline double improper_cvff_eval(const int32_t *it, const double *__restrict__ xd,
                                            double bx, double ib, const double *__restrict__ kk,
                                            const int32_t *__restrict__ dn, int32_t n, Add add) {
    constexpr double SMALL = (double)0.001f;
    const int i1 = it[0], i2 = it[1], i3 = it[2], i4 = it[3], t = it[4] - 1;
    double b1[3], b2[3], b3[3];
    dvec64(xd, n, i1, i2, bx, ib, b1);
    dvec64(xd, n, i3, i2, bx, ib, b2);
    dvec64(xd, n, i4, i3, bx, ib, b3);
    const double b1sq = b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2];
    const double b2sq = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
    const double b3sq = b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2];
    const double sb1 = 1.0 / b1sq, sb2 = 1.0 / b2sq, sb3 = 1.0 / b3sq;
    const double rb1 = sqrt(sb1), rb3 = sqrt(sb3);
    const double c0 = (b1[0] * b3[0] + b1[1] * b3[1] + b1[2] * b3[2]) * rb1 * rb3;
    const double r12c1 = 1.0 / (sqrt(b1sq) * sqrt(b2sq));
    const double c1 = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) * r12c1;
    const double r12c2 = 1.0 / (sqrt(b2sq) * sqrt(b3sq));
    const double c2 = -(b2[0] * b3[0] + b2[1] * b3[1] + b2[2] * b3[2]) * r12c2;
    const double sc1 = 1.0 / fmax(sqrt(1.0 - c1 * c1), SMALL);
    const double sc2 = 1.0 / fmax(sqrt(1.0 - c2 * c2), SMALL);
    const double s1 = sc1 * sc1, s2 = sc2 * sc2;
    double s12 = sc1 * sc2;
    double c = fmin(fmax((c0 + c1 * c2) * s12, -1.0), 1.0);
    const int d = dn[2 * t], nm = dn[2 * t + 1];
    // p = 1 + T_n(c), dp/dc = T_n'(c)
    const double rc2 = c * c;
    double p, dp;
    switch (nm) {
        case 0: p = 2.0; dp = 0.0; break;
        case 1: p = c + 1.0; dp = 1.0; break;
        case 2: p = 2.0 * rc2; dp = 4.0 * c; break;
        case 3: p = (4.0 * rc2 - 3.0) * c + 1.0; dp = 12.0 * rc2 - 3.0; break;
        case 4: p = 8.0 * (rc2 - 1.0) * rc2 + 2.0; dp = (32.0 * rc2 - 16.0) * c; break;
        case 5: p = ((16.0 * rc2 - 20.0) * rc2 + 5.0) * c + 1.0; dp = (80.0 * rc2 - 60.0) * rc2 + 5.0; break;
        default: p = ((32.0 * rc2 - 48.0) * rc2 + 18.0) * rc2; dp = (192.0 * (rc2 - 1.0) * rc2 + 36.0) * c; break;
    }
    if (d == -1) { p = 2.0 - p; dp = -dp; }
    const double e = kk[t] * p;
    const double a = kk[t] * dp;
    c *= a;
    s12 *= a;
    const double a11 = c * sb1 * s1;
    const double a22 = -sb2 * (2.0 * c0 * s12 - c * (s1 + s2));
    const double a33 = c * sb3 * s2;
    const double a12 = -r12c1 * (c1 * c * s1 + c2 * s12);
    const double a13 = -rb1 * rb3 * s12;
    const double a23 = r12c2 * (c2 * c * s2 + c1 * s12);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double sx2 = a12 * b1[k] + a22 * b2[k] + a23 * b3[k];
        const double f1 = a11 * b1[k] + a12 * b2[k] + a13 * b3[k];
        const double f4 = a13 * b1[k] + a23 * b2[k] + a33 * b3[k];
        add(0, k, to_fix(f1));
        add(1, k, to_fix(-sx2 - f1));
        add(2, k, to_fix(sx2 - f4));
        add(3, k, to_fix(f4));
    }
    return e;
}

// Global atomic sink of the per-item kernels: atom column col of item `it`.
struct AtomicSink {
    const int32_t *it;
    long long *fout;
    int32_t n;
    __device__ void operator()(int col, int k, long long q) const {
        atomicAdd((unsigned long long *)&fout[(size_t)k * n + it[col]], (unsigned long long)q);
    }
};

// ------------------------------------------------------------------ one thread per item
template <bool F64POS>
__global__ void k_pair14(int32_t items, const int32_t *__restrict__ dl, const float4 *__restrict__ pos4,
                         const double *__restrict__ xd, const float *__restrict__ q,
                         const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                         const double4 *__restrict__ lj14, DevParams P, int32_t n,
                         long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items;
    double ec = 0.0, el = 0.0;
    int r = 0;
    if (live) {
        const int32_t *it = dl + (size_t)m * 5;
        r = rep[it[0]];
        pair14_eval<F64POS>(it, pos4, xd, q, reps[r].box, reps[r].ibox, lj14, P, n, ec, el,
                            AtomicSink{it, fout, n});
    }
    wave_add(to_efix(ec), live, r, eacc, E_COUL14);
    wave_add(to_efix(el), live, r, eacc, E_POT14);
}

// Items whose type is <= 0 belong to a rigid water and are evaluated inside the SHAKE kernel.
__global__ void k_bond(int32_t items, const int32_t *__restrict__ bl, const double *__restrict__ xd,
                       const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                       const double *__restrict__ kb, const double *__restrict__ r0, int32_t n,
                       long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items && bl[(size_t)m * 3 + 2] >= 1;
    long long e = 0;
    int r = 0;
    if (live) {
        const int32_t *it = bl + (size_t)m * 3;
        r = rep[it[0]];
        e = bond_eval(it, xd, reps[r].box, reps[r].ibox, kb, r0, n, AtomicSink{it, fout, n});
    }
    wave_add(e, live, r, eacc, E_BOND);
}

// (types <= 0: rigid-water angles, evaluated inside the SHAKE kernel)
__global__ void k_angle(int32_t items, const int32_t *__restrict__ al, const double *__restrict__ xd,
                        const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                        const double *__restrict__ ka, const double *__restrict__ th0, int32_t n,
                        long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items && al[(size_t)m * 4 + 3] >= 1;
    long long e = 0;
    int r = 0;
    if (live) {
        const int32_t *it = al + (size_t)m * 4;
        r = rep[it[0]];
        e = angle_eval(it, xd, reps[r].box, reps[r].ibox, ka, th0, n, AtomicSink{it, fout, n});
    }
    wave_add(e, live, r, eacc, E_ANGLE);
}

__global__ void k_torsion(int32_t items, const int32_t *__restrict__ tl, const double *__restrict__ xd,
                          const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                          const double *__restrict__ kk, const double *__restrict__ mult,
                          const double *__restrict__ shift_deg, const double *__restrict__ shift_cs,
                          int32_t n, long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items;
    double e = 0.0;
    int r = 0;
    if (live) {
        const int32_t *it = tl + (size_t)m * 5;
        r = rep[it[0]];
        e = torsion_eval(it, xd, reps[r].box, reps[r].ibox, kk, mult, shift_deg, shift_cs, n,
                         AtomicSink{it, fout, n});
    }
    wave_add(to_efix(e), live, r, eacc, E_DIHED);
}

__global__ void k_improper(int32_t items, const int32_t *__restrict__ il, const double *__restrict__ xd,
                           const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                           const double *__restrict__ kk, const double *__restrict__ chi0,
                           int32_t n, long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items;
    double e = 0.0;
    int r = 0;
    if (live) {
        const int32_t *it = il + (size_t)m * 5;
        r = rep[it[0]];
        e = improper_eval(it, xd, reps[r].box, reps[r].ibox, kk, chi0, n, AtomicSink{it, fout, n});
    }
    wave_add(to_efix(e), live, r, eacc, E_IMPRO);
}

__global__ void k_improper_cvff(int32_t items, const int32_t *__restrict__ il,
                                const double *__restrict__ xd, const int32_t *__restrict__ rep,
                                const RepInfo *__restrict__ reps, const double *__restrict__ kk,
                                const int32_t *__restrict__ dn, int32_t n,
                                long long *__restrict__ fout, long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < items;
    double e = 0.0;
    int r = 0;
    if (live) {
        const int32_t *it = il + (size_t)m * 5;
        r = rep[it[0]];
        e = improper_cvff_eval(it, xd, reps[r].box, reps[r].ibox, kk, dn, n, AtomicSink{it, fout, n});
    }
    wave_add(to_efix(e), live, r, eacc, E_IMPRO);
}

// real*4 positions promoted to an FP64 SoA copy, the input of the bonded kernels above
__global__ void k_pos_to_xd(int32_t n, const float4 *__restrict__ pos4, double *__restrict__ xd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos4[i];
    xd[i] = p.x;
    xd[(size_t)n + i] = p.y;
    xd[2 * (size_t)n + i] = p.z;
}

// ------------------------------------------------------------------ one workgroup per bundle
// A bundle holds whole molecules (connected components of the bonded items) of one
// replica, at most BUNDLE_SLOTS atoms, its items sorted by kind so that a wavefront mostly
// runs one kind's code.  Each thread takes items, adds their forces to the bundle's atoms in
// LDS (64-bit LDS atomics), and the bundle then adds each atom's total to the global
// accumulator with a plain read-modify-write: no atom belongs to two bundles.  A term
// record is (kind << 29 | item, four 8-bit atom slots); a slot record (kind << 29 | offset)
// names a list entry that holds the slot's atom.
struct BondedArgs {
    const int32_t *lst[BK_COUNT];   // dihedrals (1-4 pairs), bonds, angles, torsions, impropers
    const float4 *pos4;
    const double *xd;
    const float *q;
    const int32_t *rep;
    const RepInfo *reps;
    const double4 *lj14;
    const double *bond_k, *bond_r0, *angle_k, *angle_th0;
    const double *tors_k, *tors_n, *tors_shift, *tors_cs;
    const double *impro_k, *impro_chi0;
    const int32_t *impro_dn;
    int32_t impro_style;
    int32_t n;
};

__global__ void __launch_bounds__(BUNDLE_THREADS)
k_bonded_bundles(const int32_t *__restrict__ tstart, const int2 *__restrict__ terms,
                 const int32_t *__restrict__ sstart, const int32_t *__restrict__ slots,
                 BondedArgs A, DevParams P, long long *__restrict__ fout,
                 long long *__restrict__ eacc) {
    __shared__ unsigned long long s_f[3 * BUNDLE_SLOTS];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int t0 = tstart[b], t1 = tstart[b + 1];
    const int s0 = sstart[b], ns = sstart[b + 1] - s0;
    for (int k = lane; k < 3 * ns; k += BUNDLE_THREADS) s_f[k] = 0ull;
    // the bundle's replica, from its first atom
    const uint32_t first = (uint32_t)slots[s0];
    const int r = A.rep[A.lst[first >> 29][first & 0x1fffffffu]];
    const float boxf = A.reps[r].box, iboxf = A.reps[r].ibox;
    const double box = boxf, ibox = iboxf;
    __syncthreads();
    // energies in quanta per item, as the per-item kernels form them
    long long ec = 0, el = 0, ed = 0, ei = 0, eb = 0, ea = 0;
    for (int t = t0 + lane; t < t1; t += BUNDLE_THREADS) {
        const int2 rec = terms[t];
        const int kind = (int)((uint32_t)rec.x >> 29), item = (int)((uint32_t)rec.x & 0x1fffffffu);
        const int32_t *it = A.lst[kind] + (size_t)item * bonded_width(kind);
        const uint32_t sl = (uint32_t)rec.y;
        auto add = [&](int col, int k, long long q) {
            atomicAdd(&s_f[3 * ((sl >> (8 * col)) & 0xffu) + k], (unsigned long long)q);
        };
        switch (kind) {
            case BK_PAIR14: {
                double c14, l14;
                pair14_eval<true>(it, A.pos4, A.xd, A.q, boxf, iboxf, A.lj14, P, A.n, c14, l14, add);
                ec += to_efix(c14);
                el += to_efix(l14);
                break;
            }
            case BK_BOND: eb += bond_eval(it, A.xd, box, ibox, A.bond_k, A.bond_r0, A.n, add); break;
            case BK_ANGLE: ea += angle_eval(it, A.xd, box, ibox, A.angle_k, A.angle_th0, A.n, add); break;
            case BK_TORSION:
                ed += to_efix(torsion_eval(it, A.xd, box, ibox, A.tors_k, A.tors_n, A.tors_shift,
                                           A.tors_cs, A.n, add));
                break;
            default:
                ei += to_efix(A.impro_style == 0
                                  ? improper_eval(it, A.xd, box, ibox, A.impro_k, A.impro_chi0, A.n, add)
                                  : improper_cvff_eval(it, A.xd, box, ibox, A.impro_k, A.impro_dn, A.n, add));
                break;
        }
    }
    __syncthreads();
    for (int s = lane; s < ns; s += BUNDLE_THREADS) {
        const uint32_t rec = (uint32_t)slots[s0 + s];
        const int atom = A.lst[rec >> 29][rec & 0x1fffffffu];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            long long *p = &fout[(size_t)k * A.n + atom];
            *p += (long long)s_f[3 * s + k];
        }
    }
    wave_add(ec, true, r, eacc, E_COUL14);
    wave_add(el, true, r, eacc, E_POT14);
    wave_add(eb, true, r, eacc, E_BOND);
    wave_add(ea, true, r, eacc, E_ANGLE);
    wave_add(ed, true, r, eacc, E_DIHED);
    wave_add(ei, true, r, eacc, E_IMPRO);
}

}  // namespace

void launch_correct(sofia_ctx *c, int which, long long *fout) {
    const int32_t *lst = which == 12 ? c->bondlist : which == 13 ? c->anglelist : c->dihedlist;
    const int items = which == 12 ? c->nb : which == 13 ? c->na : c->nd;
    const int width = which == 12 ? 3 : which == 13 ? 4 : 5;
    const int colb = which == 12 ? 1 : which == 13 ? 2 : 3;
    const double fudge = which == 12 ? c->dp.fudge12 : which == 13 ? c->dp.fudge13 : c->dp.fudge14;
    if (!items) return;
    hipLaunchKernelGGL(k_correct, dim3(nblocks(items, 256)), dim3(256), 0, c->stream, items, width,
                       colb, lst, c->a.pos4, c->a.q, c->rep, c->reps_d, c->dp, fudge, c->N, fout,
                       c->eacc);
}

void launch_pair14(sofia_ctx *c, long long *fout) {
    if (!c->nd) return;
    hipLaunchKernelGGL(k_pair14<false>, dim3(nblocks(c->nd, 256)), dim3(256), 0, c->stream, c->nd,
                       c->dihedlist, c->a.pos4, c->a.xd, c->a.q, c->rep, c->reps_d, c->lj14, c->dp,
                       c->N, fout, c->eacc);
}

// The drop-in form of the bonded terms: every bond, angle, torsion and improper of the
// uploaded lists (rigid waters included, as the reference's bond_harmonic / angle_harmonic
// loop over them) on the real*4 positions, promoted into the gather buffer's FP64 copy
// (free outside a sort), forces into fout.
void launch_bonded_dropin(sofia_ctx *c, long long *fout) {
    const int n = c->N;
    double *xd = c->b.xd;
    hipLaunchKernelGGL(k_pos_to_xd, dim3(nblocks(n, 256)), dim3(256), 0, c->stream, n, c->a.pos4, xd);
    if (c->nb && c->bond_k)
        hipLaunchKernelGGL(k_bond, dim3(nblocks(c->nb, 256)), dim3(256), 0, c->stream, c->nb,
                           c->bondlist, xd, c->rep, c->reps_d, c->bond_k, c->bond_r0, n, fout, c->eacc);
    if (c->na && c->angle_k)
        hipLaunchKernelGGL(k_angle, dim3(nblocks(c->na, 256)), dim3(256), 0, c->stream, c->na,
                           c->anglelist, xd, c->rep, c->reps_d, c->angle_k, c->angle_th0, n, fout,
                           c->eacc);
    if (c->ntor && c->tors_k)
        hipLaunchKernelGGL(k_torsion, dim3(nblocks(c->ntor, 256)), dim3(256), 0, c->stream, c->ntor,
                           c->torsions, xd, c->rep, c->reps_d, c->tors_k, c->tors_n, c->tors_shift,
                           c->tors_cs, n, fout, c->eacc);
    if (c->nimp && c->impro_style == 0)
        hipLaunchKernelGGL(k_improper, dim3(nblocks(c->nimp, 256)), dim3(256), 0, c->stream, c->nimp,
                           c->improlist, xd, c->rep, c->reps_d, c->impro_k, c->impro_chi0, n, fout,
                           c->eacc);
    else if (c->nimp)
        hipLaunchKernelGGL(k_improper_cvff, dim3(nblocks(c->nimp, 256)), dim3(256), 0, c->stream,
                           c->nimp, c->improlist, xd, c->rep, c->reps_d, c->impro_k, c->impro_dn, n,
                           fout, c->eacc);
}

void launch_bonded_engine(sofia_ctx *c) {
    const int n = c->N;
    if (c->bonded_bundles && c->nbundle > 0) {
        BondedArgs A{};
        A.lst[BK_PAIR14] = c->dihedlist;
        A.lst[BK_BOND] = c->water_split ? c->ebond : c->bondlist;
        A.lst[BK_ANGLE] = c->water_split ? c->eangle : c->anglelist;
        A.lst[BK_TORSION] = c->torsions;
        A.lst[BK_IMPRO] = c->improlist;
        A.pos4 = c->a.pos4; A.xd = c->a.xd; A.q = c->a.q; A.rep = c->rep; A.reps = c->reps_d;
        A.lj14 = c->lj14;
        A.bond_k = c->bond_k; A.bond_r0 = c->bond_r0; A.angle_k = c->angle_k; A.angle_th0 = c->angle_th0;
        A.tors_k = c->tors_k; A.tors_n = c->tors_n; A.tors_shift = c->tors_shift; A.tors_cs = c->tors_cs;
        A.impro_k = c->impro_k; A.impro_chi0 = c->impro_chi0; A.impro_dn = c->impro_dn;
        A.impro_style = c->impro_style;
        A.n = n;
        hipLaunchKernelGGL(k_bonded_bundles, dim3(c->nbundle), dim3(BUNDLE_THREADS), 0, c->stream, c->bund_tstart,
                           c->bund_terms, c->bund_sstart, c->bund_slots, A, c->dp, c->fix, c->eacc);
        return;
    }
    if (c->nd)
        hipLaunchKernelGGL(k_pair14<true>, dim3(nblocks(c->nd, 256)), dim3(256), 0, c->stream, c->nd,
                           c->dihedlist, c->a.pos4, c->a.xd, c->a.q, c->rep, c->reps_d, c->lj14,
                           c->dp, n, c->fix, c->eacc);
    // with the rigid waters fused into SHAKE, the remaining items only
    const int nb = c->water_split ? c->nbe : c->nb, na = c->water_split ? c->nae : c->na;
    const int32_t *bl = c->water_split ? c->ebond : c->bondlist;
    const int32_t *al = c->water_split ? c->eangle : c->anglelist;
    if (nb && c->bond_k)
        hipLaunchKernelGGL(k_bond, dim3(nblocks(nb, 256)), dim3(256), 0, c->stream, nb, bl, c->a.xd,
                           c->rep, c->reps_d, c->bond_k, c->bond_r0, n, c->fix, c->eacc);
    if (na && c->angle_k)
        hipLaunchKernelGGL(k_angle, dim3(nblocks(na, 256)), dim3(256), 0, c->stream, na, al, c->a.xd,
                           c->rep, c->reps_d, c->angle_k, c->angle_th0, n, c->fix, c->eacc);
    if (c->ntor && c->tors_k)
        hipLaunchKernelGGL(k_torsion, dim3(nblocks(c->ntor, 256)), dim3(256), 0, c->stream, c->ntor,
                           c->torsions, c->a.xd, c->rep, c->reps_d, c->tors_k, c->tors_n,
                           c->tors_shift, c->tors_cs, n, c->fix, c->eacc);
    if (c->nimp && c->impro_style == 0)
        hipLaunchKernelGGL(k_improper, dim3(nblocks(c->nimp, 256)), dim3(256), 0, c->stream, c->nimp,
                           c->improlist, c->a.xd, c->rep, c->reps_d, c->impro_k, c->impro_chi0, n,
                           c->fix, c->eacc);
    else if (c->nimp)
        hipLaunchKernelGGL(k_improper_cvff, dim3(nblocks(c->nimp, 256)), dim3(256), 0, c->stream,
                           c->nimp, c->improlist, c->a.xd, c->rep, c->reps_d, c->impro_k,
                           c->impro_dn, n, c->fix, c->eacc);
}

}  // namespace sofia
