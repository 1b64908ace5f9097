// Integration, wrapping, Langevin and SHAKE kernels.
//
// Two flavours:
//  * drop-in ("32"): the reference's real*4 state (ainfo positions, coords velocities and
//    forces) with its FP64 temporaries -- nve_v/nve_x (Source/mod_ensemble.f90:11-87),
//    pbc (Source/mod_pbc.f90:6-39), shake (Source/mod_shake.f90:8-913),
//    langevin_debug_intel (Source/mod_langevin.f90:15-43);
//  * engine ("64"): the same algorithms on FP64 positions and velocities, with forces read
//    as FP64 nonbond sums + int64 fixed-point sums of every other term.
#include "sofia_device.h"

namespace sofia {
namespace {

__global__ void k_zero(int32_t n3, long long *__restrict__ f, int32_t ne, long long *__restrict__ e) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) f[i] = 0;
    if (e && i < ne) e[i] = 0;
}

__global__ void k_fix_to_aos(int32_t n, const long long *__restrict__ f, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = (float)((double)f[(size_t)c * n + i] * FIX_INV);
}

__global__ void k_f64_to_aos(int32_t n, const double *__restrict__ f, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * (size_t)i + c] = (float)f[(size_t)c * n + i];
}

// ------------------------------------------------------------------ drop-in (real*4 state)
__global__ void k_nve_v32(int32_t n, float dt_partial, const float4 *__restrict__ pos4,
                          const double *__restrict__ mass, float *__restrict__ v,
                          const float *__restrict__ ff, const int32_t *__restrict__ rep,
                          double amu2e, long long *__restrict__ eacc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double k = 0.0;
    int r = 0;
    if (live) {
        r = rep[i];
        const double m = mass[__float_as_int(pos4[i].w) - 1];
        const double dtm = (double)dt_partial / m;
        float vv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vv[c] = (float)((double)v[3 * (size_t)i + c] + dtm * (double)ff[3 * (size_t)i + c]);
            v[3 * (size_t)i + c] = vv[c];
        }
        k = 0.5 * amu2e * m * (double)((vv[0] * vv[0] + vv[1] * vv[1]) + vv[2] * vv[2]);
    }
    wave_add(to_efix(k), live, r, eacc, E_KE);
}

__global__ void k_nve_x32(int32_t n, float dt, const float *__restrict__ v, float4 *__restrict__ pos4,
                          float *__restrict__ dr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos4[i];
    const float dx = dt * v[3 * (size_t)i], dy = dt * v[3 * (size_t)i + 1], dz = dt * v[3 * (size_t)i + 2];
    dr[i] += dx;
    dr[(size_t)n + i] += dy;
    dr[2 * (size_t)n + i] += dz;
    p.x += dx;
    p.y += dy;
    p.z += dz;
    pos4[i] = p;
}

// while-loop wrap into (0, box]; bounded so a non-finite coordinate cannot spin forever
__device__ inline float wrap32(float x, float box) {
    for (int it = 0; it < 64 && x > box; ++it) x = x - box;
    for (int it = 0; it < 64 && (double)x <= 0.0; ++it) x = x + box;
    return x;
}

__global__ void k_pbc32(int32_t n, float4 *__restrict__ pos4, const int32_t *__restrict__ rep,
                        const RepInfo *__restrict__ reps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float box = reps[rep[i]].box;
    float4 p = pos4[i];
    p.x = wrap32(p.x, box);
    p.y = wrap32(p.y, box);
    p.z = wrap32(p.z, box);
    pos4[i] = p;
}

// berendsen, Source/mod_ensemble.f90:167-205: position = position*mu (real*4 * real*8, stored
// real*4); dx = new - dble(old); drx = drx + dx (real*8 sum stored real*4).
__global__ void k_berendsen32(int32_t n, float4 *__restrict__ pos4, float *__restrict__ dr,
                              const int32_t *__restrict__ rep, const double *__restrict__ mu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mu[rep[i]];
    float4 p = pos4[i];
    float *c[3] = {&p.x, &p.y, &p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double old = (double)*c[d];
        const float nw = (float)(old * m);
        *c[d] = nw;
        float *r = dr + (size_t)d * n + i;
        *r = (float)((double)*r + ((double)nw - old));
    }
    pos4[i] = p;
}

__global__ void k_langevin32(int32_t n, const float4 *__restrict__ pos4, const float *__restrict__ v,
                             const int32_t *__restrict__ gid, const double *__restrict__ mass,
                             float fric, float gconst, uint64_t seed, uint64_t step,
                             float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mass[__float_as_int(pos4[i].w) - 1];
    const double g1 = -(double)fric * m, g2 = (double)gconst * sqrt(m);
    float u[3];
    uniform3(seed, step, (uint32_t)gid[i], u);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        out[3 * (size_t)i + c] = (float)(g1 * (double)v[3 * (size_t)i + c] + g2 * (double)u[c]);
}

// ------------------------------------------------------------------ SHAKE
// Constraint k joins atoms (P[k], Q[k]) of the cluster: shake2 one bond, shake3 (0,1),(0,2),
// shake4 (0,1),(0,2),(0,3), shake3angle (0,1),(0,2),(1,2).  The Lagrange multipliers solve
// |s_k + sum_l cpl[k][l] lambda_l r_l|^2 = d_k^2 by the reference's iteration: linear
// system with the quadratic terms lagged, stopped when every |delta lambda| <= tol.
struct ShakeCluster {
    int nc;          // constraints
    int P[3], Q[3];
    int idx[4];
    double im[4];
    double bl[3];
};

__device__ inline bool shake_solve(const ShakeCluster &S, const double (*r)[3], const double (*s)[3],
                                   double tol, int maxiter, double lam[3], int *iters) {
    if (S.nc == 1) {   // shake2: closed-form root of the quadratic (Source/mod_shake.f90:133-168)
        const double r01sq = r[0][0] * r[0][0] + r[0][1] * r[0][1] + r[0][2] * r[0][2];
        const double s01sq = s[0][0] * s[0][0] + s[0][1] * s[0][1] + s[0][2] * s[0][2];
        const double ims = S.im[0] + S.im[1];
        const double a = ims * ims * r01sq;
        const double b = 2.0 * ims * (s[0][0] * r[0][0] + s[0][1] * r[0][1] + s[0][2] * r[0][2]);
        const double cc = s01sq - S.bl[0] * S.bl[0];
        const double det = b * b - 4.0 * a * cc;
        if (det < 0.0) return false;
        const double l1 = (-b + sqrt(det)) / (2.0 * a), l2 = (-b - sqrt(det)) / (2.0 * a);
        lam[0] = (fabs(l1) <= fabs(l2)) ? l1 : l2;
        *iters = 1;
        return true;
    }
    const int nc = S.nc;
    double cpl[3][3], A[3][3], G[3][3], ssq[3];
    for (int k = 0; k < nc; ++k) {
        ssq[k] = s[k][0] * s[k][0] + s[k][1] * s[k][1] + s[k][2] * s[k][2];
        for (int l = 0; l < nc; ++l) {
            double c = 0.0;
            if (S.P[k] == S.P[l]) c += S.im[S.P[k]];
            if (S.P[k] == S.Q[l]) c -= S.im[S.P[k]];
            if (S.Q[k] == S.P[l]) c -= S.im[S.Q[k]];
            if (S.Q[k] == S.Q[l]) c += S.im[S.Q[k]];
            cpl[k][l] = c;
            A[k][l] = 2.0 * c * (s[k][0] * r[l][0] + s[k][1] * r[l][1] + s[k][2] * r[l][2]);
            G[k][l] = r[k][0] * r[l][0] + r[k][1] * r[l][1] + r[k][2] * r[l][2];
        }
    }
    double Ai[3][3];
    if (nc == 2) {
        const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        if (det == 0.0) return false;
        const double di = 1.0 / det;
        Ai[0][0] = A[1][1] * di; Ai[0][1] = -A[0][1] * di;
        Ai[1][0] = -A[1][0] * di; Ai[1][1] = A[0][0] * di;
    } else {
        const double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
                           A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1] -
                           A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
        if (det == 0.0) return false;
        const double di = 1.0 / det;
        Ai[0][0] = di * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
        Ai[0][1] = -di * (A[0][1] * A[2][2] - A[0][2] * A[2][1]);
        Ai[0][2] = di * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
        Ai[1][0] = -di * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
        Ai[1][1] = di * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
        Ai[1][2] = -di * (A[0][0] * A[1][2] - A[0][2] * A[1][0]);
        Ai[2][0] = di * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        Ai[2][1] = -di * (A[0][0] * A[2][1] - A[0][1] * A[2][0]);
        Ai[2][2] = di * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    }
    lam[0] = lam[1] = lam[2] = 0.0;
    int it = 0;
    bool done = false;
    while (!done && it < maxiter) {
        double bb[3], nl[3];
        for (int k = 0; k < nc; ++k) {
            double qd = 0.0;
            for (int l = 0; l < nc; ++l)
                for (int m = 0; m < nc; ++m) qd += cpl[k][l] * cpl[k][m] * G[l][m] * lam[l] * lam[m];
            bb[k] = S.bl[k] * S.bl[k] - ssq[k] - qd;
        }
        done = true;
        for (int k = 0; k < nc; ++k) {
            double v = 0.0;
            for (int l = 0; l < nc; ++l) v += Ai[k][l] * bb[l];
            nl[k] = v;
        }
        for (int k = 0; k < nc; ++k) {
            if (fabs(nl[k] - lam[k]) > tol) done = false;
            lam[k] = nl[k];
        }
        ++it;
    }
    *iters = it;
    return true;
}

__device__ inline ShakeCluster make_cluster(int num, const int32_t *at, const double *bd,
                                            const double *mass, const float4 *pos4) {
    ShakeCluster S;
    for (int k = 0; k < 4; ++k) {
        S.idx[k] = at[k];
        S.im[k] = (at[k] >= 0) ? 1.0 / mass[__float_as_int(pos4[at[k]].w) - 1] : 0.0;
    }
    for (int k = 0; k < 3; ++k) S.bl[k] = bd[k];
    if (num == 1) {
        S.nc = 1; S.P[0] = 0; S.Q[0] = 1;
    } else if (num == 3) {
        S.nc = 2; S.P[0] = 0; S.Q[0] = 1; S.P[1] = 0; S.Q[1] = 2;
    } else if (num == 4) {
        S.nc = 3; S.P[0] = 0; S.Q[0] = 1; S.P[1] = 0; S.Q[1] = 2; S.P[2] = 0; S.Q[2] = 3;
    } else {   // -3: rigid triangle, constraints 0-1, 0-2, 1-2
        S.nc = 3; S.P[0] = 0; S.Q[0] = 1; S.P[1] = 0; S.Q[1] = 2; S.P[2] = 1; S.Q[2] = 2;
    }
    return S;
}

// drop-in: real*4 positions/velocities/forces, FP64 solve (Source/mod_shake.f90:8-85, 888-913)
__global__ void k_shake32(int32_t ns, const int32_t *__restrict__ num, const int32_t *__restrict__ atoms,
                          const double *__restrict__ bond, const float4 *__restrict__ pos4,
                          const float *__restrict__ v, float *__restrict__ ff,
                          const double *__restrict__ mass, const int32_t *__restrict__ rep,
                          const RepInfo *__restrict__ reps, float dt, double dtsq, double tol,
                          int maxiter, int32_t *__restrict__ err, int32_t *__restrict__ iters_max,
                          long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < ns;
    double vir[3] = {0.0, 0.0, 0.0};
    int rr = 0;
    if (live) {
        const ShakeCluster S = make_cluster(num[m], atoms + 4 * (size_t)m, bond + 3 * (size_t)m, mass, pos4);
        rr = rep[S.idx[0]];
        const float box = reps[rr].box, ibox = reps[rr].ibox;
        double xs[4][3];
        for (int a = 0; a < 4; ++a) {
            if (S.idx[a] < 0) continue;
            const int i = S.idx[a];
            const float4 p = pos4[i];
            const float pp[3] = {p.x, p.y, p.z};
            for (int c = 0; c < 3; ++c)
                xs[a][c] = (double)(pp[c] + dt * v[3 * (size_t)i + c]) +
                           dtsq * S.im[a] * (double)ff[3 * (size_t)i + c];
        }
        double r[3][3], s[3][3];
        for (int k = 0; k < S.nc; ++k) {
            const float4 pa = pos4[S.idx[S.P[k]]], pb = pos4[S.idx[S.Q[k]]];
            const float df[3] = {pa.x - pb.x, pa.y - pb.y, pa.z - pb.z};
            for (int c = 0; c < 3; ++c) {
                const double d = (double)df[c];
                r[k][c] = d - (double)(box * (float)(int)round((double)ibox * d));
                const double e = xs[S.P[k]][c] - xs[S.Q[k]][c];
                s[k][c] = e - (double)(box * (float)(int)round((double)ibox * e));
            }
        }
        double lam[3];
        int its = 0;
        if (!shake_solve(S, r, s, tol, maxiter, lam, &its)) {
            atomicOr(err, 1);
        } else {
            atomicMax(iters_max, its);
            for (int k = 0; k < S.nc; ++k) {
                const double l = lam[k] / dtsq;
                const int a = S.idx[S.P[k]], b = S.idx[S.Q[k]];
                for (int c = 0; c < 3; ++c) {
                    ff[3 * (size_t)a + c] = (float)((double)ff[3 * (size_t)a + c] + l * r[k][c]);
                    ff[3 * (size_t)b + c] = (float)((double)ff[3 * (size_t)b + c] - l * r[k][c]);
                    vir[c] += r[k][c] * r[k][c] * l;
                }
            }
        }
    }
    wave_add(to_efix(vir[0]), live, rr, eacc, V_X);
    wave_add(to_efix(vir[1]), live, rr, eacc, V_Y);
    wave_add(to_efix(vir[2]), live, rr, eacc, V_Z);
}

// ------------------------------------------------------------------ engine (FP64 state)
__global__ void k_from32(int32_t n, const float4 *__restrict__ pos4, const float *__restrict__ v32,
                         double *__restrict__ xd, double *__restrict__ vd, double *__restrict__ drd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos4[i];
    xd[i] = p.x; xd[(size_t)n + i] = p.y; xd[2 * (size_t)n + i] = p.z;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vd[(size_t)c * n + i] = v32[3 * (size_t)i + c];
        drd[(size_t)c * n + i] = 0.0;
    }
}

__global__ void k_to32(int32_t n, const double *__restrict__ xd, const double *__restrict__ vd,
                       float4 *__restrict__ pos4, float *__restrict__ v32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos4[i];
    p.x = (float)xd[i]; p.y = (float)xd[(size_t)n + i]; p.z = (float)xd[2 * (size_t)n + i];
    pos4[i] = p;
#pragma unroll
    for (int c = 0; c < 3; ++c) v32[3 * (size_t)i + c] = (float)vd[(size_t)c * n + i];
}

// v += (dt/2 kcal2amu / m) F with F = nonbond + fixed-point rest; optional kinetic energy
__global__ void k_nve_v64(int32_t n, double dthalf, const float4 *__restrict__ pos4,
                          const double *__restrict__ mass, double *__restrict__ vd,
                          const double *__restrict__ fnb, const long long *__restrict__ fix,
                          const int32_t *__restrict__ rep, double amu2e, int ke,
                          long long *__restrict__ eacc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double k = 0.0;
    int r = 0;
    if (live) {
        r = rep[i];
        const double m = mass[__float_as_int(pos4[i].w) - 1];
        const double dtm = dthalf / m;
        double v2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = (size_t)c * n + i;
            const double f = fnb[o] + (double)fix[o] * FIX_INV;
            const double vv = vd[o] + dtm * f;
            vd[o] = vv;
            v2 += vv * vv;
        }
        k = 0.5 * amu2e * m * v2;
    }
    if (ke) wave_add(to_efix(k), live, r, eacc, E_KE);
}

// x += dt v, displacement += dt v, wrap into (0, box], refresh the float4 copy
__global__ void k_nve_x64(int32_t n, double dt, double *__restrict__ xd, const double *__restrict__ vd,
                          double *__restrict__ drd, float4 *__restrict__ pos4,
                          const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double box = reps[rep[i]].box;
    float4 p = pos4[i];
    float pf[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        const double d = dt * vd[o];
        drd[o] += d;
        double x = xd[o] + d;
        for (int it = 0; it < 64 && x > box; ++it) x -= box;
        for (int it = 0; it < 64 && x <= 0.0; ++it) x += box;
        xd[o] = x;
        pf[c] = (float)x;
    }
    p.x = pf[0]; p.y = pf[1]; p.z = pf[2];
    pos4[i] = p;
}

// Langevin force -fric m v + gconst sqrt(m) U(-1/2,1/2) added to the fixed-point sum
__global__ void k_langevin64(int32_t n, const float4 *__restrict__ pos4, const double *__restrict__ vd,
                             const int32_t *__restrict__ gid, const int32_t *__restrict__ rep,
                             const RepInfo *__restrict__ reps, const double *__restrict__ mass,
                             const uint64_t *__restrict__ step, long long *__restrict__ fix) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RepInfo &R = reps[rep[i]];
    if (R.fric == 0.0 && R.gconst == 0.0) return;
    const double m = mass[__float_as_int(pos4[i].w) - 1];
    float u[3];
    uniform3(R.seed, step[rep[i]], (uint32_t)gid[i], u);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        fix[o] += to_fix(-R.fric * m * vd[o] + R.gconst * sqrt(m) * (double)u[c]);
    }
}

// Engine SHAKE.  The solver is instantiated per cluster kind so the constraint topology is
// known at compile time (the arrays stay in registers); the arithmetic is the one of
// shake_solve.  Cluster atoms are exclusive to their cluster.
template <int KIND> __device__ constexpr int topo_nc() { return KIND == 1 ? 1 : (KIND == 3 ? 2 : 3); }
template <int KIND> __device__ constexpr int topo_na() { return KIND == 1 ? 2 : (KIND == 4 ? 4 : 3); }
template <int KIND> __device__ constexpr int topo_p(int k) { return (KIND == -3 && k == 2) ? 1 : 0; }
template <int KIND> __device__ constexpr int topo_q(int k) { return (KIND == -3 && k == 2) ? 2 : k + 1; }

template <int KIND>
__device__ inline bool shake_solve_t(const double *im, const double *bl, const double (*r)[3],
                                     const double (*s)[3], double tol, int maxiter, double *lam,
                                     int *iters) {
    constexpr int NC = topo_nc<KIND>();
    if constexpr (NC == 1) {
        const double r01sq = r[0][0] * r[0][0] + r[0][1] * r[0][1] + r[0][2] * r[0][2];
        const double s01sq = s[0][0] * s[0][0] + s[0][1] * s[0][1] + s[0][2] * s[0][2];
        const double ims = im[0] + im[1];
        const double a = ims * ims * r01sq;
        const double b = 2.0 * ims * (s[0][0] * r[0][0] + s[0][1] * r[0][1] + s[0][2] * r[0][2]);
        const double cc = s01sq - bl[0] * bl[0];
        const double det = b * b - 4.0 * a * cc;
        if (det < 0.0) return false;
        const double l1 = (-b + sqrt(det)) / (2.0 * a), l2 = (-b - sqrt(det)) / (2.0 * a);
        lam[0] = (fabs(l1) <= fabs(l2)) ? l1 : l2;
        *iters = 1;
        return true;
    } else {
        double cpl[NC][NC], A[NC][NC], G[NC][NC], ssq[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            ssq[k] = s[k][0] * s[k][0] + s[k][1] * s[k][1] + s[k][2] * s[k][2];
#pragma unroll
            for (int l = 0; l < NC; ++l) {
                double c = 0.0;
                if (topo_p<KIND>(k) == topo_p<KIND>(l)) c += im[topo_p<KIND>(k)];
                if (topo_p<KIND>(k) == topo_q<KIND>(l)) c -= im[topo_p<KIND>(k)];
                if (topo_q<KIND>(k) == topo_p<KIND>(l)) c -= im[topo_q<KIND>(k)];
                if (topo_q<KIND>(k) == topo_q<KIND>(l)) c += im[topo_q<KIND>(k)];
                cpl[k][l] = c;
                A[k][l] = 2.0 * c * (s[k][0] * r[l][0] + s[k][1] * r[l][1] + s[k][2] * r[l][2]);
                G[k][l] = r[k][0] * r[l][0] + r[k][1] * r[l][1] + r[k][2] * r[l][2];
            }
        }
        double Ai[NC][NC];
        if constexpr (NC == 2) {
            const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
            if (det == 0.0) return false;
            const double di = 1.0 / det;
            Ai[0][0] = A[1][1] * di; Ai[0][1] = -A[0][1] * di;
            Ai[1][0] = -A[1][0] * di; Ai[1][1] = A[0][0] * di;
        } else {
            const double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
                               A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1] -
                               A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
            if (det == 0.0) return false;
            const double di = 1.0 / det;
            Ai[0][0] = di * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
            Ai[0][1] = -di * (A[0][1] * A[2][2] - A[0][2] * A[2][1]);
            Ai[0][2] = di * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
            Ai[1][0] = -di * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
            Ai[1][1] = di * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
            Ai[1][2] = -di * (A[0][0] * A[1][2] - A[0][2] * A[1][0]);
            Ai[2][0] = di * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
            Ai[2][1] = -di * (A[0][0] * A[2][1] - A[0][1] * A[2][0]);
            Ai[2][2] = di * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) lam[k] = 0.0;
        int it = 0;
        bool done = false;
        while (!done && it < maxiter) {
            double bb[NC], nl[NC];
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                double qd = 0.0;
#pragma unroll
                for (int l = 0; l < NC; ++l)
#pragma unroll
                    for (int m = 0; m < NC; ++m) qd += cpl[k][l] * cpl[k][m] * G[l][m] * lam[l] * lam[m];
                bb[k] = bl[k] * bl[k] - ssq[k] - qd;
            }
            done = true;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < NC; ++l) v += Ai[k][l] * bb[l];
                nl[k] = v;
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                if (fabs(nl[k] - lam[k]) > tol) done = false;
                lam[k] = nl[k];
            }
            ++it;
        }
        *iters = it;
        return true;
    }
}

struct ShakeArgs {
    const int32_t *atoms;
    const double *bond;
    const int4 *water;        // per cluster: bond types (O-H1, O-H2), angle type, fused flag
    const float4 *pos4;
    const double *xd, *vd, *fnb;
    long long *fix;
    const double *mass;
    const int32_t *rep;
    const RepInfo *reps;
    const double *bond_k, *bond_r0, *angle_k, *angle_th0;
    int32_t n;
    double dt, dtsq, tol;
    int maxiter;
    int32_t *err, *iters_max;
};

// The two O-H bonds and the H-O-H angle of a rigid water (atoms idx = O, H1, H2), added to
// the atoms' fixed-point sums with the quanta the bond and angle kernels would add.
__device__ inline void water_bonded(const int4 w, const ShakeArgs &A, const int *idx, double box,
                                    double ibox, long long (*own)[3], long long &eb, long long &ea) {
    long long f[3], f1[3], f3[3];
    eb += bond_term(A.xd, A.n, idx[0], idx[1], box, ibox, A.bond_k[w.x - 1], A.bond_r0[w.x - 1], f);
#pragma unroll
    for (int c = 0; c < 3; ++c) { own[0][c] += f[c]; own[1][c] -= f[c]; }
    eb += bond_term(A.xd, A.n, idx[0], idx[2], box, ibox, A.bond_k[w.y - 1], A.bond_r0[w.y - 1], f);
#pragma unroll
    for (int c = 0; c < 3; ++c) { own[0][c] += f[c]; own[2][c] -= f[c]; }
    ea += angle_term(A.xd, A.n, idx[1], idx[0], idx[2], box, ibox, A.angle_k[w.z - 1],
                     A.angle_th0[w.z - 1], f1, f3);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        own[1][c] += f1[c];
        own[0][c] -= f1[c] + f3[c];
        own[2][c] += f3[c];
    }
}

// One cluster of kind KIND.  A fused rigid water (KIND -3, water flag set) also evaluates its
// two O-H bonds and its H-O-H angle here -- with the same fixed-point quanta the bond and
// angle kernels would add -- and owns its three atoms, so it writes their force sums without
// atomics.  Returns the cluster's replica; adds to the virial and energy partial sums.
template <int KIND>
__device__ inline int shake_cluster(int m, const ShakeArgs &A, double vir[3], long long &eb,
                                    long long &ea) {
    constexpr int NA = topo_na<KIND>(), NC = topo_nc<KIND>();
    int idx[NA];
    double im[NA], bl[3];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
        idx[k] = A.atoms[4 * (size_t)m + k];
        im[k] = 1.0 / A.mass[__float_as_int(A.pos4[idx[k]].w) - 1];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) bl[k] = A.bond[3 * (size_t)m + k];
    const int rr = A.rep[idx[0]];
    const double box = A.reps[rr].box, ibox = A.reps[rr].ibox;
    const int n = A.n;
    long long own[NA][3];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) own[a][c] = A.fix[(size_t)c * n + idx[a]];
    bool fused = false;
    if constexpr (KIND == -3) {
        const int4 w = A.water[m];
        fused = w.w != 0;
        if (fused) water_bonded(w, A, idx, box, ibox, own, eb, ea);
    }
    double xs[NA][3];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = (size_t)c * n + idx[a];
            const double f = A.fnb[o] + (double)own[a][c] * FIX_INV;
            xs[a][c] = A.xd[o] + A.dt * A.vd[o] + A.dtsq * im[a] * f;
        }
    double r[NC][3], s[NC][3];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int pa = topo_p<KIND>(k), qa = topo_q<KIND>(k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double d = A.xd[(size_t)c * n + idx[pa]] - A.xd[(size_t)c * n + idx[qa]];
            r[k][c] = d - box * rint(d * ibox);
            const double e = xs[pa][c] - xs[qa][c];
            s[k][c] = e - box * rint(e * ibox);
        }
    }
    double lam[NC];
    int its = 0;
    if (!shake_solve_t<KIND>(im, bl, r, s, A.tol, A.maxiter, lam, &its)) {
        atomicOr(A.err, 1);
        return rr;
    }
    atomicMax(A.iters_max, its);
    long long add[NA][3];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) add[a][c] = 0;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const double l = lam[k] / A.dtsq;
        const int pa = topo_p<KIND>(k), qa = topo_q<KIND>(k);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long long q = to_fix(l * r[k][c]);
            add[pa][c] += q;
            add[qa][c] -= q;
            vir[c] += r[k][c] * r[k][c] * l;
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            long long *p = &A.fix[(size_t)c * n + idx[a]];
            if (fused) *p = own[a][c] + add[a][c];
            else atomicAdd((unsigned long long *)p, (unsigned long long)add[a][c]);
        }
    return rr;
}

// Rigid symmetric 3-site water (a -3 cluster with |O-H1| = |O-H2| and equal H masses) by
// SETTLE (Miyamoto & Kollman, J. Comput. Chem. 13, 952, 1992): the closed-form solution of
// the constraint equations that the iterative triangle SHAKE converges to.  Positions are
// taken relative to the old O with the minimum image; the constraint force of an atom is
// m (x_settled - x_unconstrained) / dtsq, the O force is minus the sum of the H forces in
// fixed point (exact momentum), and the virial is sum_k lambda_k r_k^2, written through the
// H forces and the old O-H vectors.  The fused water bond/angle terms are evaluated first,
// as in shake_cluster<-3>.
__device__ inline int settle_water(int m, const ShakeArgs &A, double vir[3], long long &eb,
                                   long long &ea) {
    int rr;
    {
        int idx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) idx[k] = A.atoms[4 * (size_t)m + k];
        const double mo = A.mass[__float_as_int(A.pos4[idx[0]].w) - 1];
        const double mh = A.mass[__float_as_int(A.pos4[idx[1]].w) - 1];
        const double doh = A.bond[3 * (size_t)m], dhh = A.bond[3 * (size_t)m + 2];
        rr = A.rep[idx[0]];
        const double box = A.reps[rr].box, ibox = A.reps[rr].ibox;
        const int n = A.n;
        long long own[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) own[a][c] = A.fix[(size_t)c * n + idx[a]];
        const int4 w = A.water[m];
        const bool fused = w.w != 0;
        if (fused) water_bonded(w, A, idx, box, ibox, own, eb, ea);
        // old (b0, c0) and unconstrained new (x1) positions relative to the old O
        const double im[3] = {1.0 / mo, 1.0 / mh, 1.0 / mh};
        double b0[3], c0[3], x1[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t oo = (size_t)c * n + idx[0];
            const double xo = A.xd[oo];
            double d = A.xd[(size_t)c * n + idx[1]] - xo;
            b0[c] = d - box * rint(d * ibox);
            d = A.xd[(size_t)c * n + idx[2]] - xo;
            c0[c] = d - box * rint(d * ibox);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const size_t o = (size_t)c * n + idx[a];
                const double f = A.fnb[o] + (double)own[a][c] * FIX_INV;
                const double step = A.dt * A.vd[o] + A.dtsq * im[a] * f;
                const double base = a == 0 ? 0.0 : (a == 1 ? b0[c] : c0[c]);
                x1[a][c] = base + step;
            }
        }
        // canonical triangle: O at ra above the centre of mass, the H-H midpoint rb below
        const double wt = mo + 2.0 * mh, rc = 0.5 * dhh;
        const double hgt = sqrt(doh * doh - rc * rc);
        const double ra = 2.0 * mh * hgt / wt, rb = hgt - ra;
        double com[3], a1[3], b1[3], c1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            com[c] = (mo * x1[0][c] + mh * (x1[1][c] + x1[2][c])) / wt;
            a1[c] = x1[0][c] - com[c];
            b1[c] = x1[1][c] - com[c];
            c1[c] = x1[2][c] - com[c];
        }
        auto cross = [](const double *u, const double *v, double *o) {
            o[0] = u[1] * v[2] - u[2] * v[1];
            o[1] = u[2] * v[0] - u[0] * v[2];
            o[2] = u[0] * v[1] - u[1] * v[0];
        };
        double T[3][3];   // rows: the frame axes x', y', z'
        cross(b0, c0, T[2]);
        cross(a1, T[2], T[0]);
        cross(T[2], T[0], T[1]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double inv = 1.0 / sqrt(T[k][0] * T[k][0] + T[k][1] * T[k][1] + T[k][2] * T[k][2]);
            T[k][0] *= inv; T[k][1] *= inv; T[k][2] *= inv;
        }
        auto to = [&](const double *v, int k) { return T[k][0] * v[0] + T[k][1] * v[1] + T[k][2] * v[2]; };
        const double xb0 = to(b0, 0), yb0 = to(b0, 1), xc0 = to(c0, 0), yc0 = to(c0, 1);
        const double za1 = to(a1, 2);
        const double xb1 = to(b1, 0), yb1 = to(b1, 1), zb1 = to(b1, 2);
        const double xc1 = to(c1, 0), yc1 = to(c1, 1), zc1 = to(c1, 2);
        // The closed form exists only while the unconstrained step leaves the water within
        // reach of the rigid triangle: each radicand below must be non-negative (and every
        // value finite).  Otherwise the step is rejected as a failed constraint solve, as the
        // iterative solver rejects a singular one, instead of writing NaN into the state.
        const double sinphi = za1 / ra;
        const double rad_phi = 1.0 - sinphi * sinphi;
        const double cosphi = sqrt(fmax(rad_phi, 0.0));
        const double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi);
        const double rad_psi = 1.0 - sinpsi * sinpsi;
        const double cospsi = sqrt(fmax(rad_psi, 0.0));
        const double ya2 = ra * cosphi, xb2 = -rc * cospsi;
        const double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
        const double yb2 = t1 - t2, yc2 = t1 + t2;
        const double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
        const double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
        const double gamma = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
        const double ab = alpha * alpha + beta * beta;
        const double rad_ab = ab - gamma * gamma;
        const double sint = (alpha * gamma - beta * sqrt(fmax(rad_ab, 0.0))) / ab;
        const double rad_t = 1.0 - sint * sint;
        // written so that a NaN in any operand also fails
        if (!(rad_phi >= 0.0 && cosphi > 0.0 && rad_psi >= 0.0 && ab > 0.0 && rad_ab >= 0.0 &&
              rad_t >= 0.0)) {
            atomicOr(A.err, 1);
            return rr;
        }
        const double cost = sqrt(rad_t);
        const double p3[3][3] = {{-ya2 * sint, ya2 * cost, za1},
                                 {xb2 * cost - yb2 * sint, xb2 * sint + yb2 * cost, zb1},
                                 {-xb2 * cost - yc2 * sint, -xb2 * sint + yc2 * cost, zc1}};
        const double ms[3] = {mo, mh, mh};
        long long add[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            long long sum_h = 0;
#pragma unroll
            for (int a = 1; a < 3; ++a) {
                const double xnew = T[0][c] * p3[a][0] + T[1][c] * p3[a][1] + T[2][c] * p3[a][2] + com[c];
                const double f = ms[a] * (xnew - x1[a][c]) / A.dtsq;
                add[a][c] = to_fix(f);
                sum_h += add[a][c];
                vir[c] += (a == 1 ? b0[c] : c0[c]) * f;
            }
            add[0][c] = -sum_h;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                long long *p = &A.fix[(size_t)c * n + idx[a]];
                if (fused) *p = own[a][c] + add[a][c];
                else atomicAdd((unsigned long long *)p, (unsigned long long)add[a][c]);
            }
    }
    return rr;
}

// All engine constraints in one launch.  The first nb_it blocks run the iteratively solved
// clusters of kidx (kinds 1, 3, 4, -3 in that order; e1..e4 are the ends of the groups), so
// that a wavefront almost always runs a single solver.  Each wavefront of the other blocks
// owns WSPAN consecutive atoms in their current (cell-sorted) order: it queues the SETTLE
// waters whose oxygen lies there in LDS and solves them 64 at a time, so its waters come
// from a few neighbouring cells and share cache lines, with every lane busy.
constexpr int WSPAN = 512;

__global__ void __launch_bounds__(128)
k_constraints(const int32_t *__restrict__ kidx, int32_t e1, int32_t e2, int32_t e3, int32_t e4,
              int32_t nb_it, const int32_t *__restrict__ wclus, ShakeArgs A,
              long long *__restrict__ eacc) {
    if ((int)blockIdx.x < nb_it) {
        double vir[3] = {0.0, 0.0, 0.0};
        long long eb = 0, ea = 0;
        int rr = 0;
        const int t = blockIdx.x * blockDim.x + threadIdx.x;
        const bool live = t < e4;
        if (live) {
            const int m = kidx[t];
            if (t >= e3) rr = shake_cluster<-3>(m, A, vir, eb, ea);
            else if (t >= e2) rr = shake_cluster<4>(m, A, vir, eb, ea);
            else if (t >= e1) rr = shake_cluster<3>(m, A, vir, eb, ea);
            else rr = shake_cluster<1>(m, A, vir, eb, ea);
        }
        wave_add(to_efix(vir[0]), live, rr, eacc, V_X);
        wave_add(to_efix(vir[1]), live, rr, eacc, V_Y);
        wave_add(to_efix(vir[2]), live, rr, eacc, V_Z);
        wave_add(eb, live, rr, eacc, E_BOND);
        wave_add(ea, live, rr, eacc, E_ANGLE);
        return;
    }
    __shared__ int32_t s_q[2][WSPAN];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int a0 = ((blockIdx.x - nb_it) * 2 + wave) * WSPAN;
    int cnt = 0;
    for (int k = 0; k < WSPAN; k += 64) {
        const int i = a0 + k + lane;
        const int m = i < A.n ? wclus[i] : -1;
        const uint64_t b = __ballot(m >= 0);
        if (m >= 0) s_q[wave][cnt + __popcll(b & lt_mask)] = m;
        cnt += __popcll(b);
    }
    __builtin_amdgcn_wave_barrier();   // this wavefront's LDS writes precede its reads
    for (int base = 0; base < cnt; base += 64) {
        double vir[3] = {0.0, 0.0, 0.0};
        long long eb = 0, ea = 0;
        int rr = 0;
        const bool live = base + lane < cnt;
        if (live) rr = settle_water(s_q[wave][base + lane], A, vir, eb, ea);
        wave_add(to_efix(vir[0]), live, rr, eacc, V_X);
        wave_add(to_efix(vir[1]), live, rr, eacc, V_Y);
        wave_add(to_efix(vir[2]), live, rr, eacc, V_Z);
        wave_add(eb, live, rr, eacc, E_BOND);
        wave_add(ea, live, rr, eacc, E_ANGLE);
    }
}

// Plain streaming copy (16 B per thread per iteration): the HBM reference for traffic
// counters and for the achieved-bandwidth figure of the profiling helpers.
__global__ void k_copy16(size_t n16, const float4 *__restrict__ src, float4 *__restrict__ dst) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += stride) dst[k] = src[k];
}

// The force wrapper's merge: each term rounded to real*4 as its own routine returns it (the
// fixed-point sums as k_fix_to_aos rounds them), then pair + 1-4 + bonded summed in FP64 and
// stored real*4 into ff (which holds the pair term on entry).
__global__ void k_merge_forces(int32_t n, float *__restrict__ ff, const long long *__restrict__ f14,
                               const long long *__restrict__ fb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float h = (float)((double)f14[(size_t)c * n + i] * FIX_INV);
        const float b = (float)((double)fb[(size_t)c * n + i] * FIX_INV);
        ff[3 * (size_t)i + c] = (float)(((double)ff[3 * (size_t)i + c] + (double)h) + (double)b);
    }
}

// Atom types (the int bits of pos.w) outside 1..ntypes set *bad; any lane that finds one
// writes the same value, so a plain store suffices.
__global__ void k_type_range(int32_t n, const float4 *__restrict__ pos, int32_t ntypes,
                             int32_t *__restrict__ bad) {
    const int stride = gridDim.x * blockDim.x;
    bool out = false;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int t = __float_as_int(pos[k].w);
        out |= (t < 1) | (t > ntypes);
    }
    if (__ballot(out) && __lane_id() == 0) *bad = 1;
}

// The fused rigid-water bonded terms alone (steps that run no SHAKE: the minimiser).
__global__ void k_water_bonded(int32_t ns, const int32_t *__restrict__ num, ShakeArgs A,
                               long long *__restrict__ eacc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = m < ns && num[m] != 1 && num[m] != 3 && num[m] != 4 && A.water[m].w != 0;
    long long eb = 0, ea = 0;
    int rr = 0;
    if (live) {
        int idx[3];
        for (int k = 0; k < 3; ++k) idx[k] = A.atoms[4 * (size_t)m + k];
        rr = A.rep[idx[0]];
        long long own[3][3];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) own[a][c] = A.fix[(size_t)c * A.n + idx[a]];
        water_bonded(A.water[m], A, idx, A.reps[rr].box, A.reps[rr].ibox, own, eb, ea);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) A.fix[(size_t)c * A.n + idx[a]] = own[a][c];
    }
    wave_add(eb, live, rr, eacc, E_BOND);
    wave_add(ea, live, rr, eacc, E_ANGLE);
}

// Capped steepest descent: every atom moves along its force by min(gamma |F|, max_disp).
__global__ void k_minimize_step(int32_t n, double gamma, double max_disp, double *__restrict__ xd,
                                double *__restrict__ drd, const double *__restrict__ fnb,
                                const long long *__restrict__ fix, float4 *__restrict__ pos4,
                                const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double box = reps[rep[i]].box;
    double f[3], f2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        f[c] = fnb[o] + (double)fix[o] * FIX_INV;
        f2 += f[c] * f[c];
    }
    const double fn = sqrt(f2);
    const double step = (fn > 0.0) ? fmin(gamma * fn, max_disp) / fn : 0.0;
    float4 p = pos4[i];
    float pf[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        const double d = isfinite(f[c]) ? step * f[c] : 0.0;
        drd[o] += d;
        double x = xd[o] + d;
        for (int it = 0; it < 64 && x > box; ++it) x -= box;
        for (int it = 0; it < 64 && x <= 0.0; ++it) x += box;
        xd[o] = x;
        pf[c] = (float)x;
    }
    p.x = pf[0]; p.y = pf[1]; p.z = pf[2];
    pos4[i] = p;
}

__global__ void k_step_add(int32_t nrep, uint64_t *step, long long delta) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep) step[r] += (uint64_t)delta;
}

}  // namespace

void launch_zero_forces(sofia_ctx *c, long long *f, bool energies) {
    const int n3 = 3 * c->N, ne = (int)eacc_len(c->nrep);
    hipLaunchKernelGGL(k_zero, dim3(nblocks(n3 > ne ? n3 : ne, 256)), dim3(256), 0, c->stream, n3, f,
                       ne, energies ? c->eacc : nullptr);
}

void launch_fix_to_coords(sofia_ctx *c, const long long *f, float *out) {
    hipLaunchKernelGGL(k_fix_to_aos, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, f, out);
}

void launch_f64_to_coords(sofia_ctx *c, const double *f, float *out) {
    hipLaunchKernelGGL(k_f64_to_aos, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, f, out);
}

void launch_nve_v32(sofia_ctx *c, float dt_partial, const float *ff) {
    hipLaunchKernelGGL(k_nve_v32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, dt_partial,
                       c->a.pos4, c->mass, c->a.v32, ff, c->rep, c->dp.amu2e, c->eacc);
}

void launch_nve_x32(sofia_ctx *c) {
    hipLaunchKernelGGL(k_nve_x32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->dp.dt,
                       c->a.v32, c->a.pos4, c->dr32);
}

void launch_berendsen32(sofia_ctx *c, const double *mu_dev) {
    hipLaunchKernelGGL(k_berendsen32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->dr32, c->rep, mu_dev);
}

void launch_pbc32(sofia_ctx *c) {
    hipLaunchKernelGGL(k_pbc32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->rep, c->reps_d);
}

void launch_shake32(sofia_ctx *c, int iflag, float *ff, double *) {
    if (!c->ns) return;
    const float dt = c->dp.dt, k2a = c->dp.kcal2amu;
    const double dtsq = (iflag == 0) ? (double)(((0.5f * dt) * dt) * k2a) : (double)((dt * dt) * k2a);
    hipLaunchKernelGGL(k_shake32, dim3(nblocks(c->ns, 128)), dim3(128), 0, c->stream, c->ns,
                       c->shake_num, c->shake_atoms, c->shake_bond, c->a.pos4, c->a.v32, ff, c->mass,
                       c->rep, c->reps_d, dt, dtsq, c->dp.shake_tol, c->dp.shake_maxiter,
                       c->shake_err, c->shake_iters, c->eacc);
}

void launch_langevin32(sofia_ctx *c, float fric, float gconst, uint64_t seed, uint64_t step, float *out) {
    hipLaunchKernelGGL(k_langevin32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->a.v32, c->a.gid, c->mass, fric, gconst, seed, step, out);
}

void launch_engine_from32(sofia_ctx *c) {
    hipLaunchKernelGGL(k_from32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->a.v32, c->a.xd, c->a.vd, c->drd);
}

void launch_engine_to32(sofia_ctx *c) {
    hipLaunchKernelGGL(k_to32, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.xd,
                       c->a.vd, c->a.pos4, c->a.v32);
}

void launch_nve_v64(sofia_ctx *c, bool ke) {
    const double dthalf = 0.5 * (double)c->dp.dt * (double)c->dp.kcal2amu;
    hipLaunchKernelGGL(k_nve_v64, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, dthalf,
                       c->a.pos4, c->mass, c->a.vd, c->fnb, c->fix, c->rep, c->dp.amu2e, ke ? 1 : 0,
                       c->eacc);
}

void launch_nve_x64(sofia_ctx *c) {
    hipLaunchKernelGGL(k_nve_x64, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N,
                       (double)c->dp.dt, c->a.xd, c->a.vd, c->drd, c->a.pos4, c->rep, c->reps_d);
}

void launch_langevin64(sofia_ctx *c) {
    hipLaunchKernelGGL(k_langevin64, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->a.vd, c->a.gid, c->rep, c->reps_d, c->mass, c->step_d, c->fix);
}

ShakeArgs shake_args(sofia_ctx *c) {
    const double dt = c->dp.dt;
    const double dtsq = dt * dt * (double)c->dp.kcal2amu;
    ShakeArgs A;
    A.atoms = c->shake_atoms; A.bond = c->shake_bond; A.water = c->water;
    A.pos4 = c->a.pos4; A.xd = c->a.xd; A.vd = c->a.vd; A.fnb = c->fnb; A.fix = c->fix;
    A.mass = c->mass; A.rep = c->rep; A.reps = c->reps_d;
    A.bond_k = c->bond_k; A.bond_r0 = c->bond_r0; A.angle_k = c->angle_k; A.angle_th0 = c->angle_th0;
    A.n = c->N; A.dt = dt; A.dtsq = dtsq; A.tol = c->dp.shake_tol; A.maxiter = c->dp.shake_maxiter;
    A.err = c->shake_err; A.iters_max = c->shake_iters;
    return A;
}

void launch_shake64(sofia_ctx *c) {
    if (!c->ns) return;
    const int32_t *n = c->shake_kcount;
    const int e1 = n[0], e2 = e1 + n[1], e3 = e2 + n[2], e4 = e3 + n[3];
    const int nb_it = nblocks(e4, 128);
    const int nb_w = n[4] ? nblocks(c->N, 2 * WSPAN) : 0;
    if (nb_it + nb_w == 0) return;
    hipLaunchKernelGGL(k_constraints, dim3(nb_it + nb_w), dim3(128), 0, c->stream, c->shake_kidx,
                       e1, e2, e3, e4, nb_it, c->a.wclus, shake_args(c), c->eacc);
}

void launch_merge_forces(sofia_ctx *c, float *ff32, const long long *f14, const long long *fb) {
    hipLaunchKernelGGL(k_merge_forces, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, ff32,
                       f14, fb);
}

void launch_type_range(sofia_ctx *c, const float4 *pos, int32_t *bad) {
    hipLaunchKernelGGL(k_type_range, dim3(std::min(nblocks(c->N, 256), 1024)), dim3(256), 0,
                       c->stream, c->N, pos, c->p.ntypes, bad);
}

void launch_copy16(sofia_ctx *c, size_t n16, const void *src, void *dst) {
    hipLaunchKernelGGL(k_copy16, dim3(2048), dim3(256), 0, c->stream, n16, (const float4 *)src,
                       (float4 *)dst);
}

void launch_water_bonded(sofia_ctx *c) {
    if (!c->ns || !c->water_split) return;
    hipLaunchKernelGGL(k_water_bonded, dim3(nblocks(c->ns, 128)), dim3(128), 0, c->stream, c->ns,
                       c->shake_num, shake_args(c), c->eacc);
}

void launch_minimize_step(sofia_ctx *c, double gamma, double max_disp) {
    hipLaunchKernelGGL(k_minimize_step, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N, gamma,
                       max_disp, c->a.xd, c->drd, c->fnb, c->fix, c->a.pos4, c->rep, c->reps_d);
}

void launch_step_advance(sofia_ctx *c) {
    launch_step_add(c, 1);
}

void launch_step_add(sofia_ctx *c, long long delta) {
    hipLaunchKernelGGL(k_step_add, dim3(nblocks(c->nrep, 256)), dim3(256), 0, c->stream, c->nrep,
                       c->step_d, delta);
}

}  // namespace sofia
