/*
 * CPU oracle for the Sofia hot path -- TEST INFRASTRUCTURE ONLY (see sofia_oracle.h).
 *
 * A plain-C restatement of the reference Fortran routines.  Compiled with
 * -O2 -ffp-contract=off -fopenmp so that every floating-point expression is evaluated
 * in IEEE order with the same float/double promotions as the Fortran source; OpenMP is
 * used only where the reference itself opens a parallel loop over atoms, and never
 * changes a per-atom result.
 */
#include "sofia_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SPEC 12

/* sign(a, b) for real*4 / real*8 with the magnitude 1/epsilon (2^23 / 2^52) */
static inline float sgnf(float mag, float b) { return copysignf(mag, b); }
static inline double sgnd(double mag, double b) { return copysign(mag, b); }

/* Fortran nint: round half away from zero */
static inline int nint_d(double x) { return (int)lround(x); }

/* real*4 minimum image with the 2^23 rounding trick, Source/mod_neighbor.f90:181-187 */
static inline float minimg_f32(float dx, float box, float ibox) {
    const float big = 8388608.0f; /* 1/epsilon(real*4) */
    float s = dx * ibox;
    s = (s + sgnf(big, s)) - sgnf(big, dx);
    return dx - s * box;
}

/* real*8 minimum image with the 2^52 trick, Source/mod_force_nonbond.f90:144-149 */
static inline double minimg_f64(double dx, float box, float ibox) {
    const double big = 4503599627370496.0; /* 1/epsilon(real*8) */
    double s = dx * (double)ibox;
    s = (s + sgnd(big, s)) - sgnd(big, dx);
    return dx - (double)box * s;
}

/* Source/mod_minimg.f90:5-15: dx - box*nint(ibox*dx); box*nint is a real*4 product */
static inline double minimg_nint(double dx, float box, float ibox) {
    int n = nint_d((double)ibox * dx);
    float shift = box * (float)n;
    return dx - (double)shift;
}

/* ------------------------------------------------------------------------------------
 * Cell grid: check_allocation, Source/mod_neighbor.f90:1446-1451 (real*4 arithmetic);
 * vol = box**3 from Source/get_started.f90:64.
 * ---------------------------------------------------------------------------------- */
void orc_cell_grid(float box, float rcut, float cut_coul, float *rn, int32_t *ncellT,
                   int32_t *ncellD) {
    float rc = (cut_coul > rcut) ? cut_coul : rcut;
    int n = (int)(box / rc);
    float r = box / (float)n;
    float vol = (box * box) * box;
    float r3 = (r * r) * r;
    int nT = nint_d((double)(vol / r3));
    int nD = nint_d(pow((double)nT, 1.0 / 3.0));
    *rn = r;
    *ncellT = nT;
    *ncellD = nD;
}

/* neighcell_nonewton, Source/mod_neighbor.f90:1309-1378: slot 0 is the cell itself,
 * then the 26 neighbours with iz outermost and ix fastest, periodic. */
void orc_neighcell(int32_t ncellT, int32_t ncellD, int32_t *cnum) {
    int count = 0;
    for (int ic = 0; ic < ncellT; ++ic) {
        cnum[count] = ic;
        int icz = ic / (ncellD * ncellD);
        int icy = (ic - icz * ncellD * ncellD) / ncellD;
        int icx = ic - icy * ncellD - icz * ncellD * ncellD;
        for (int iz = -1; iz <= 1; ++iz)
            for (int iy = -1; iy <= 1; ++iy)
                for (int ix = -1; ix <= 1; ++ix) {
                    int zc = icz + iz, yc = icy + iy, xc = icx + ix;
                    if (zc < 0) zc += ncellD; else if (zc >= ncellD) zc -= ncellD;
                    if (yc < 0) yc += ncellD; else if (yc >= ncellD) yc -= ncellD;
                    if (xc < 0) xc += ncellD; else if (xc >= ncellD) xc -= ncellD;
                    if (iz != 0 || iy != 0 || ix != 0) {
                        ++count;
                        cnum[count] = xc + yc * ncellD + zc * ncellD * ncellD;
                    }
                }
        ++count;
    }
}

/* xyz2bin, Source/mod_neighbor.f90:1381-1424: coordinates promoted to double, scaled
 * by 1/dble(rn), truncated; x >= box goes to the last cell. */
int32_t orc_xyz2bin(float x, float y, float z, float rn, float box, int32_t ncellD) {
    double irn = 1.0 / (double)rn, bx = (double)box;
    double c[3] = {(double)x, (double)y, (double)z};
    int ix[3];
    for (int d = 0; d < 3; ++d) {
        ix[d] = (c[d] < bx) ? (int)(c[d] * irn) : ncellD - 1;
        if (ix[d] == ncellD) ix[d] = ncellD - 1;
    }
    return ix[0] + ix[1] * ncellD + ix[2] * ncellD * ncellD;
}

/* new_nlist + sort_posit_charge, Source/mod_neighbor.f90:923-1005.
 * Linked cells are built by inserting atoms 1..np at the chain head, so each chain is
 * walked in descending old index; new indices are handed out cell by cell.
 * Outputs (1-based atom ids): tt[old-1] = new, perm[new-1] = old; start/endposit per
 * cell (1-based, empty cell has endposit = start-1); atombin[new-1] = cell. */
void orc_sort_cells(int32_t np, const orc_ainfo *pos, float rn, float box, int32_t ncellT,
                    int32_t ncellD, int32_t *tt, int32_t *perm, int32_t *start,
                    int32_t *endposit, int32_t *atombin) {
    int32_t *hoc = (int32_t *)calloc((size_t)ncellT + 1, sizeof(int32_t));
    int32_t *ll = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np + 1));
    for (int i = 1; i <= np; ++i) {
        int c = orc_xyz2bin(pos[i - 1].x, pos[i - 1].y, pos[i - 1].z, rn, box, ncellD);
        ll[i] = hoc[c];
        hoc[c] = i;
    }
    int count = 1;
    for (int c = 0; c < ncellT; ++c) {
        int j = hoc[c];
        start[c] = count;
        while (j != 0) {
            tt[j - 1] = count;
            perm[count - 1] = j;
            atombin[count - 1] = c;
            j = ll[j];
            ++count;
        }
        endposit[c] = count - 1;
    }
    free(hoc);
    free(ll);
}

/* ------------------------------------------------------------------------------------
 * Verlet list, build_neighbor_nonewton_MIC / _nomic, Source/mod_neighbor.f90:90-303 and
 * 622-809.  For atom i, walk the 27 stencil cells in cnum order and the atoms of each
 * cell in index order; keep z when z != i, dr2 < rv2 (real*4 dr2 compared in double)
 * and z is not in specbond(1:num3bond(i), i).  Returns the largest count that exceeded
 * neigh_alloc (0 if none); entries past neigh_alloc are not stored.
 * ---------------------------------------------------------------------------------- */
int32_t orc_build_neighbor(int32_t np, const orc_ainfo *pos, const int32_t *atombin,
                           const int32_t *start, const int32_t *endposit, const int32_t *cnum,
                           const int32_t *num3bond, const int32_t *specbond, float box,
                           float ibox, double rv2, int32_t neigh_alloc, int32_t *nlist,
                           int32_t *numneigh) {
    int32_t maxover = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxover)
    for (int i = 1; i <= np; ++i) {
        const float x1 = pos[i - 1].x, y1 = pos[i - 1].y, z1 = pos[i - 1].z;
        const int bin = atombin[i - 1];
        const int n3 = num3bond[i - 1];
        const int32_t *bonds = specbond + (size_t)(i - 1) * SPEC;
        int32_t *out = nlist + (size_t)(i - 1) * neigh_alloc;
        int nnpt = 0;
        for (int k = 27 * bin; k <= 27 * bin + 26; ++k) {
            const int cell = cnum[k];
            for (int z = start[cell]; z <= endposit[cell]; ++z) {
                float dx = x1 - pos[z - 1].x;
                float dy = y1 - pos[z - 1].y;
                float dz = z1 - pos[z - 1].z;
                dx = minimg_f32(dx, box, ibox);
                dy = minimg_f32(dy, box, ibox);
                dz = minimg_f32(dz, box, ibox);
                float dr2 = (dx * dx + dy * dy) + dz * dz;
                if (z != i && (double)dr2 < rv2) {
                    int keep = 1;
                    for (int l = 0; l < n3; ++l)
                        if (bonds[l] == z) keep = 0;
                    if (keep) {
                        if (nnpt < neigh_alloc) out[nnpt] = z;
                        ++nnpt;
                    }
                }
            }
        }
        numneigh[i - 1] = nnpt;
        if (nnpt > neigh_alloc && nnpt > maxover) maxover = nnpt;
    }
    return maxover;
}

/* build_neighbor_n2_nonewton, Source/mod_neighbor.f90:814-861: the same distance and
 * exclusion rule over all j != i, no cell stencil. */
int32_t orc_build_neighbor_n2(int32_t np, const orc_ainfo *pos, const int32_t *num3bond,
                              const int32_t *specbond, float box, float ibox, double rv2,
                              int32_t neigh_alloc, int32_t *nlist, int32_t *numneigh) {
    int32_t maxover = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : maxover)
    for (int i = 1; i <= np; ++i) {
        const float x1 = pos[i - 1].x, y1 = pos[i - 1].y, z1 = pos[i - 1].z;
        const int n3 = num3bond[i - 1];
        const int32_t *bonds = specbond + (size_t)(i - 1) * SPEC;
        int32_t *out = nlist + (size_t)(i - 1) * neigh_alloc;
        int nnpt = 0;
        for (int j = 1; j <= np; ++j) {
            if (i == j) continue;
            float dx = minimg_f32(x1 - pos[j - 1].x, box, ibox);
            float dy = minimg_f32(y1 - pos[j - 1].y, box, ibox);
            float dz = minimg_f32(z1 - pos[j - 1].z, box, ibox);
            float dr2 = (dx * dx + dy * dy) + dz * dz;
            int excl = 0;
            for (int m = 0; m < n3; ++m)
                if (bonds[m] == j) excl = 1;
            if (!excl && (double)dr2 < rv2) {
                if (nnpt < neigh_alloc) out[nnpt] = j;
                ++nnpt;
            }
        }
        numneigh[i - 1] = nnpt;
        if (nnpt > neigh_alloc && nnpt > maxover) maxover = nnpt;
    }
    return maxover;
}

/* check_neighbor, Source/mod_neighbor.f90:895-913: rebuild when any accumulated
 * displacement has drx^2+dry^2+drz^2 >= rvrc2 (products of real*4 values, summed in
 * real*4 and compared in double). */
int32_t orc_check_neighbor(int32_t np, const float *drx, const float *dry, const float *drz,
                           double rvrc2) {
    int32_t flag = 0;
    for (int i = 0; i < np; ++i) {
        double dr2 = (double)((drx[i] * drx[i] + dry[i] * dry[i]) + drz[i] * drz[i]);
        if (dr2 >= rvrc2) flag = 1;
    }
    return flag;
}

/* ------------------------------------------------------------------------------------
 * lj_cut_coul_dsf_nonewton_NOMIC / _MIC, Source/mod_force_nonbond.f90:105-196 / 8-101.
 * FP64 pair arithmetic on real*4 inputs, full list, ff(i) overwritten with the real*4
 * rounding of the per-atom double sum, potential and e_coul halved at the end.
 * ff64 (optional) receives the unrounded per-atom sums.  virial (optional) follows the
 * convention of the PME kernel: 0.5 * sum_i sum_j dx^2 * force (Source/mod_force_nonbond.f90
 * :376-390); the reference DSF kernel itself computes no virial.
 * Energies are summed per atom and then over atoms in index order.
 * ---------------------------------------------------------------------------------- */
/* The pair loop on positions already promoted to real*8 (xyz: x, y, z per atom).  The
 * reference promotes each real*4 coordinate where it forms dx (`x1 - position(j)%x`), so on
 * promoted real*4 inputs this is the reference's arithmetic exactly. */
static void dsf_core(int32_t np, const double *xyz, const int32_t *type, const float *q,
                     const int32_t *nlist, const int32_t *numneigh, int32_t neigh_alloc,
                     const float *lj1, const float *lj2, const float *lj3, const float *lj4,
                     const orc_params *p, orc_coords *ff, double *ff64, double *potential,
                     double *e_coul, double *virial) {
    double *pe = (double *)malloc(sizeof(double) * (size_t)np * 5);
    const double two_a_pis = (2.0 * p->alpha) * p->my_pis_inv;
#pragma omp parallel for schedule(dynamic, 32)
    for (int i = 1; i <= np; ++i) {
        const double x1 = xyz[3 * (size_t)(i - 1)], y1 = xyz[3 * (size_t)(i - 1) + 1];
        const double z1 = xyz[3 * (size_t)(i - 1) + 2];
        const int itype = type[i - 1];
        const double qtmp = q[i - 1];
        const int ioffset = (itype - 1) * p->ntypes;
        const int32_t *lst = nlist + (size_t)(i - 1) * neigh_alloc;
        const int num = numneigh[i - 1];
        double ffx = 0.0, ffy = 0.0, ffz = 0.0, pot = 0.0, ec = 0.0;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        for (int j = 0; j < num; ++j) {
            const int nb = lst[j] - 1;
            double dx = x1 - xyz[3 * (size_t)nb];
            double dy = y1 - xyz[3 * (size_t)nb + 1];
            double dz = z1 - xyz[3 * (size_t)nb + 2];
            const int jtype = type[nb];
            dx = minimg_f64(dx, p->box, p->ibox);
            dy = minimg_f64(dy, p->box, p->ibox);
            dz = minimg_f64(dz, p->box, p->ibox);
            const double dr2 = (dx * dx + dy * dy) + dz * dz;
            const double dr2i = 1.0 / dr2;
            double dr6i = (dr2i * dr2i) * dr2i;
            if (dr2 > p->rcut2) dr6i = 0.0;
            const int off = ioffset + jtype - 1;
            const double forcelj = dr6i * ((double)lj1[off] * dr6i - (double)lj2[off]);
            pot += dr6i * (dr6i * (double)lj3[off] - (double)lj4[off]);
            const double r = sqrt(dr2);
            const double dri = 1.0 / r;
            double prefactor = (((double)p->coulpre * qtmp) * (double)q[nb]) * dri;
            if (dr2 > p->cut_coulsq) prefactor = 0.0;
            const double erfcd = exp(-(((p->alpha * p->alpha) * r) * r));
            const double t = 1.0 / (1.0 + (p->ewald_p * p->alpha) * r);
            const double erfcc =
                (t * (p->a1 + t * (p->a2 + t * (p->a3 + t * (p->a4 + t * p->a5))))) * erfcd;
            const double forcecoul =
                (prefactor * ((erfcc * dri + two_a_pis * erfcd) + r * p->f_shift)) * r;
            ec += prefactor * ((erfcc - r * p->e_shift) - dr2 * p->f_shift);
            const double force = (forcecoul + forcelj) * dr2i;
            ffx += dx * force;
            ffy += dy * force;
            ffz += dz * force;
            vx += (dx * dx) * force;
            vy += (dy * dy) * force;
            vz += (dz * dz) * force;
        }
        ff[i - 1].x = (float)ffx;
        ff[i - 1].y = (float)ffy;
        ff[i - 1].z = (float)ffz;
        if (ff64) {
            ff64[3 * (i - 1) + 0] = ffx;
            ff64[3 * (i - 1) + 1] = ffy;
            ff64[3 * (i - 1) + 2] = ffz;
        }
        double *e = pe + (size_t)5 * (i - 1);
        e[0] = pot; e[1] = ec; e[2] = vx; e[3] = vy; e[4] = vz;
    }
    double pot = 0.0, ec = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    for (int i = 0; i < np; ++i) {
        pot += pe[5 * i]; ec += pe[5 * i + 1];
        vx += pe[5 * i + 2]; vy += pe[5 * i + 3]; vz += pe[5 * i + 4];
    }
    free(pe);
    *potential = 0.5 * pot;
    *e_coul = 0.5 * ec;
    if (virial) { virial[0] = 0.5 * vx; virial[1] = 0.5 * vy; virial[2] = 0.5 * vz; }
}

void orc_dsf(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *nlist,
             const int32_t *numneigh, int32_t neigh_alloc, const float *lj1, const float *lj2,
             const float *lj3, const float *lj4, const orc_params *p, orc_coords *ff,
             double *ff64, double *potential, double *e_coul, double *virial) {
    double *xyz = (double *)malloc(sizeof(double) * 3 * (size_t)(np > 0 ? np : 1));
    int32_t *type = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
    for (int i = 0; i < np; ++i) {
        xyz[3 * (size_t)i] = pos[i].x;
        xyz[3 * (size_t)i + 1] = pos[i].y;
        xyz[3 * (size_t)i + 2] = pos[i].z;
        type[i] = pos[i].type;
    }
    dsf_core(np, xyz, type, q, nlist, numneigh, neigh_alloc, lj1, lj2, lj3, lj4, p, ff, ff64,
             potential, e_coul, virial);
    free(type);
    free(xyz);
}

/* The same pair sum on FP64 positions (the engine's state), for checking the engine's pair
 * kernels on their own lists and coordinates. */
void orc_dsf_x64(int32_t np, const double *xyz, const int32_t *type, const float *q,
                 const int32_t *nlist, const int32_t *numneigh, int32_t neigh_alloc,
                 const float *lj1, const float *lj2, const float *lj3, const float *lj4,
                 const orc_params *p, double *ff64, double *potential, double *e_coul,
                 double *virial) {
    orc_coords *ff = (orc_coords *)malloc(sizeof(orc_coords) * (size_t)(np > 0 ? np : 1));
    dsf_core(np, xyz, type, q, nlist, numneigh, neigh_alloc, lj1, lj2, lj3, lj4, p, ff, ff64,
             potential, e_coul, virial);
    free(ff);
}

/* lj_cut_coul_long_nonewton_NOMIC, Source/mod_force_nonbond.f90:300-400: the PME
 * real-space kernel with real*4 locals (each assignment rounds to float), erfc from the
 * A&S polynomial with gewald, double accumulators for force, energy and virial. */
/* RDF_calc, Source/mod_force_nonbond.f90:708-775: over the Verlet list (1-based ids, stride
 * neigh_alloc), every atom i of type itype_sel and neighbour j of type jtype_sel with
 * dr2 < rcut2 (FP64 distance, real*8 minimum image) adds 1 to rdf[node],
 * node = int(sqrt(dr2) / rdf_delta).  The reference has no bound on node; here nodes
 * >= nbins are not counted.  numob_rdf is incremented once per call. */
void orc_rdf_calc(int32_t np, const orc_ainfo *pos, const int32_t *nlist,
                  const int32_t *numneigh, int32_t neigh_alloc, float box, float ibox,
                  double rcut2, float rdf_delta, int32_t itype_sel, int32_t jtype_sel,
                  int32_t nbins, float *rdf, int32_t *numob_rdf) {
    *numob_rdf += 1;
    for (int32_t i = 0; i < np; ++i) {
        if (pos[i].type != itype_sel) continue;
        const double x1 = pos[i].x, y1 = pos[i].y, z1 = pos[i].z;
        const int32_t *row = nlist + (size_t)i * neigh_alloc;
        for (int32_t l = 0; l < numneigh[i]; ++l) {
            const orc_ainfo *pj = &pos[row[l] - 1];
            if (pj->type != jtype_sel) continue;
            const double dx = minimg_f64(x1 - (double)pj->x, box, ibox);
            const double dy = minimg_f64(y1 - (double)pj->y, box, ibox);
            const double dz = minimg_f64(z1 - (double)pj->z, box, ibox);
            const double dr2 = dx * dx + dy * dy + dz * dz;
            if (dr2 < rcut2) {
                const int32_t node = (int32_t)(sqrt(dr2) / (double)rdf_delta);
                if (node < nbins) rdf[node] += 1.0f;
            }
        }
    }
}

void orc_coul_long(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *nlist,
                   const int32_t *numneigh, int32_t neigh_alloc, const float *lj1,
                   const float *lj2, const float *lj3, const float *lj4, const orc_params *p,
                   orc_coords *ff, double *potential, double *e_coul, double *virial) {
    double *pe = (double *)malloc(sizeof(double) * (size_t)np * 5);
#pragma omp parallel for schedule(dynamic, 32)
    for (int i = 1; i <= np; ++i) {
        const float x1 = pos[i - 1].x, y1 = pos[i - 1].y, z1 = pos[i - 1].z;
        const int itype = pos[i - 1].type;
        const float qtmp = q[i - 1];
        const int ioffset = (itype - 1) * p->ntypes;
        const int32_t *lst = nlist + (size_t)(i - 1) * neigh_alloc;
        const int num = numneigh[i - 1];
        double ffx = 0.0, ffy = 0.0, ffz = 0.0, pot = 0.0, ec = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
        for (int j = 0; j < num; ++j) {
            const int nb = lst[j] - 1;
            float dx = minimg_f32(x1 - pos[nb].x, p->box, p->ibox);
            float dy = minimg_f32(y1 - pos[nb].y, p->box, p->ibox);
            float dz = minimg_f32(z1 - pos[nb].z, p->box, p->ibox);
            const int jtype = pos[nb].type;
            const float dr2 = (dx * dx + dy * dy) + dz * dz;
            const float dr2i = (float)(1.0 / (double)dr2);
            float dr6i = (dr2i * dr2i) * dr2i;
            if ((double)dr2 > p->rcut2) dr6i = 0.0f;
            const int off = ioffset + jtype - 1;
            const float forcelj = dr6i * (lj1[off] * dr6i - lj2[off]);
            pot += (double)(dr6i * (dr6i * lj3[off] - lj4[off]));
            const float r = sqrtf(dr2);
            float prefactor = ((p->coulpre * qtmp) * q[nb]) / r;
            const float grij = (float)(p->gewald * (double)r);
            const float expm2 = expf(-(grij * grij));
            const float t = (float)(1.0 / (1.0 + p->ewald_p * (double)grij));
            const double td = (double)t;
            const float erfc =
                (float)((td * (p->a1 + td * (p->a2 + td * (p->a3 + td * (p->a4 + td * p->a5))))) *
                        (double)expm2);
            if ((double)dr2 > p->cut_coulsq) prefactor = 0.0f;
            const float forcecoul = (float)((double)prefactor *
                                            ((double)erfc + (p->ewald_f * (double)grij) * (double)expm2));
            ec += (double)(prefactor * erfc);
            const float force = (forcecoul + forcelj) * dr2i;
            ffx += (double)(dx * force);
            ffy += (double)(dy * force);
            ffz += (double)(dz * force);
            vx += (double)((dx * dx) * force);
            vy += (double)((dy * dy) * force);
            vz += (double)((dz * dz) * force);
        }
        ff[i - 1].x = (float)ffx;
        ff[i - 1].y = (float)ffy;
        ff[i - 1].z = (float)ffz;
        double *e = pe + (size_t)5 * (i - 1);
        e[0] = pot; e[1] = ec; e[2] = vx; e[3] = vy; e[4] = vz;
    }
    double pot = 0.0, ec = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    for (int i = 0; i < np; ++i) {
        pot += pe[5 * i]; ec += pe[5 * i + 1];
        vx += pe[5 * i + 2]; vy += pe[5 * i + 3]; vz += pe[5 * i + 4];
    }
    free(pe);
    *potential = 0.5 * pot;
    *e_coul = 0.5 * ec;
    if (virial) { virial[0] = 0.5 * vx; virial[1] = 0.5 * vy; virial[2] = 0.5 * vz; }
}

/* lj_cut_coul_dsf_N2, Source/mod_force_nonbond.f90:409-502: every pair i<j with
 * dr2 < rcut2 (nint minimum image in double), fudge 0 for 1-2/1-3 partners
 * (slot <= num2bond), 0.5 for 1-4 partners, Newton's third law.  Forces are returned
 * as double sums (the reference accumulates into real*4 ff). */
void orc_dsf_n2(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *num2bond,
                const int32_t *num3bond, const int32_t *specbond, const float *lj1,
                const float *lj2, const float *lj3, const float *lj4, const orc_params *p,
                double *ff64, double *potential, double *e_coul) {
    memset(ff64, 0, sizeof(double) * 3 * (size_t)np);
    double pot = 0.0, ec = 0.0;
    const double two_a_pis = (2.0 * p->alpha) * p->my_pis_inv;
    for (int i = 1; i <= np - 1; ++i) {
        const double x1 = pos[i - 1].x, y1 = pos[i - 1].y, z1 = pos[i - 1].z;
        const int itype = pos[i - 1].type;
        const int n2 = num2bond[i - 1], n3 = num3bond[i - 1];
        const int32_t *bonds = specbond + (size_t)(i - 1) * SPEC;
        const double qtmp = q[i - 1];
        for (int j = i + 1; j <= np; ++j) {
            double dx = x1 - (double)pos[j - 1].x;
            double dy = y1 - (double)pos[j - 1].y;
            double dz = z1 - (double)pos[j - 1].z;
            dx = minimg_nint(dx, p->box, p->ibox);
            dy = minimg_nint(dy, p->box, p->ibox);
            dz = minimg_nint(dz, p->box, p->ibox);
            const double dr2 = (dx * dx + dy * dy) + dz * dz;
            if (!(dr2 < p->rcut2)) continue;
            double fudge = 1.0;
            for (int l = 1; l <= n3; ++l) {
                if (bonds[l - 1] == j && l <= n2) fudge = 0.0;
                else if (bonds[l - 1] == j && l > n2) fudge = 0.5;
            }
            const double dr2i = 1.0 / dr2;
            const double dr6i = (dr2i * dr2i) * dr2i;
            const int off = p->ntypes * (itype - 1) + pos[j - 1].type - 1;
            const double forcelj = dr6i * ((double)lj1[off] * dr6i - (double)lj2[off]);
            pot += (fudge * dr6i) * ((double)lj3[off] * dr6i - (double)lj4[off]);
            const double r = sqrt(dr2);
            const double dri = 1.0 / r;
            double prefactor = (((double)p->coulpre * qtmp) * (double)q[j - 1]) * dri;
            if (dr2 > p->cut_coulsq) prefactor = 0.0;
            const double erfcd = exp(-(((p->alpha * p->alpha) * r) * r));
            const double t = 1.0 / (1.0 + (p->ewald_p * p->alpha) * r);
            const double erfcc =
                (t * (p->a1 + t * (p->a2 + t * (p->a3 + t * (p->a4 + t * p->a5))))) * erfcd;
            const double forcecoul =
                (prefactor * ((erfcc * dri + two_a_pis * erfcd) + r * p->f_shift)) * r;
            ec += (fudge * prefactor) * ((erfcc - r * p->e_shift) - dr2 * p->f_shift);
            const double force = (fudge * (forcecoul + forcelj)) * dr2i;
            ff64[3 * (i - 1) + 0] += dx * force;
            ff64[3 * (i - 1) + 1] += dy * force;
            ff64[3 * (i - 1) + 2] += dz * force;
            ff64[3 * (j - 1) + 0] -= dx * force;
            ff64[3 * (j - 1) + 1] -= dy * force;
            ff64[3 * (j - 1) + 2] -= dz * force;
        }
    }
    *potential = pot;
    *e_coul = ec;
}

/* ------------------------------------------------------------------------------------
 * correct_12_bonds / correct_13_bonds / correct_14_bonds,
 * Source/mod_force_correction.f90:6-93 / 95-184 / 187-278.  For each bonded pair:
 * the real*4 difference promoted to double, 2^52 minimum image, prefactor from a real*4
 * charge product, A&S erfc with gewald; adds erfc part minus (1-fudge)*prefactor to ff
 * (real*4 accumulation, as ff_bond/ff_angle/ff_dihed), e_coul and the diagonal virial.
 * ia/ib are 1-based atom ids.
 * ---------------------------------------------------------------------------------- */
void orc_correct_pairs(int32_t npairs, const int32_t *ia, const int32_t *ib,
                       const orc_ainfo *pos, const float *q, const orc_params *p,
                       double fudge, orc_coords *ff, double *e_coul, double *virial) {
    double ec = *e_coul;
    double vx = virial[0], vy = virial[1], vz = virial[2];
    for (int m = 0; m < npairs; ++m) {
        const int i1 = ia[m] - 1, i2 = ib[m] - 1;
        double dx = (double)(pos[i1].x - pos[i2].x);
        double dy = (double)(pos[i1].y - pos[i2].y);
        double dz = (double)(pos[i1].z - pos[i2].z);
        dx = minimg_f64(dx, p->box, p->ibox);
        dy = minimg_f64(dy, p->box, p->ibox);
        dz = minimg_f64(dz, p->box, p->ibox);
        const double dr2 = (dx * dx + dy * dy) + dz * dz;
        const double dr2i = 1.0 / dr2;
        const double r = sqrt(dr2);
        double prefactor = (double)((p->coulpre * q[i1]) * q[i2]) / r;
        const double grij = p->gewald * r;
        const double expm2 = exp(-(grij * grij));
        const double t = 1.0 / (1.0 + p->ewald_p * grij);
        const double erfc =
            (t * (p->a1 + t * (p->a2 + t * (p->a3 + t * (p->a4 + t * p->a5))))) * expm2;
        if (dr2 > p->cut_coulsq) prefactor = 0.0;
        double forcecoul = prefactor * (erfc + (p->ewald_f * grij) * expm2);
        forcecoul = forcecoul - (1.0 - fudge) * prefactor;
        ec = ec + prefactor * erfc;
        ec = ec - (1.0 - fudge) * prefactor;
        const double force = forcecoul * dr2i;
        ff[i1].x = (float)((double)ff[i1].x + dx * force);
        ff[i1].y = (float)((double)ff[i1].y + dy * force);
        ff[i1].z = (float)((double)ff[i1].z + dz * force);
        ff[i2].x = (float)((double)ff[i2].x - dx * force);
        ff[i2].y = (float)((double)ff[i2].y - dy * force);
        ff[i2].z = (float)((double)ff[i2].z - dz * force);
        vx += (dx * dx) * force;
        vy += (dy * dy) * force;
        vz += (dz * dz) * force;
    }
    *e_coul = ec;
    virial[0] = vx; virial[1] = vy; virial[2] = vz;
}

/* The 1-4 pair inside dihedral_opls_DSF, Source/mod_force_dihed.f90:887-895, 1003-1048:
 * nint minimum image of the real*4 difference; if dr2 < cut_coulsq, 0.5 * DSF Coulomb
 * (charge product in real*4) and, if dr2 < rcut2, the lj14 tables; the pair force goes
 * to atoms 1 and 4 of each dihedlist entry by real*4 accumulation into ffh. */
void orc_pair14_dsf(int32_t nd, const int32_t *dihedlist, const orc_ainfo *pos,
                    const float *q, const double *lj14_1, const double *lj14_2,
                    const double *lj14_3, const double *lj14_4, const orc_params *p,
                    orc_coords *ffh, double *e_coul_14, double *pot_14) {
    double ec = *e_coul_14, pe = *pot_14;
    const double two_a_pis = (2.0 * p->alpha) / p->my_pis;
    for (int n = 0; n < nd; ++n) {
        const int i1 = dihedlist[5 * n + 0] - 1, i4 = dihedlist[5 * n + 3] - 1;
        double ddx = (double)(pos[i1].x - pos[i4].x);
        double ddy = (double)(pos[i1].y - pos[i4].y);
        double ddz = (double)(pos[i1].z - pos[i4].z);
        ddx = minimg_nint(ddx, p->box, p->ibox);
        ddy = minimg_nint(ddy, p->box, p->ibox);
        ddz = minimg_nint(ddz, p->box, p->ibox);
        const double dr2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
        const int it = pos[i1].type - 1, jt = pos[i4].type - 1;
        double force = 0.0;
        if (dr2 < p->cut_coulsq) {
            const double r = sqrt(dr2);
            const double dr2i = 1.0 / dr2;
            const double prefactor = (double)((p->coulpre * q[i1]) * q[i4]) / r;
            const double erfcd = exp(-(((p->alpha * p->alpha) * r) * r));
            const double t = 1.0 / (1.0 + (p->ewald_p * p->alpha) * r);
            const double erfcc =
                (t * (p->a1 + t * (p->a2 + t * (p->a3 + t * (p->a4 + t * p->a5))))) * erfcd;
            const double forcecoul =
                ((0.5 * prefactor) * ((erfcc / r + two_a_pis * erfcd) + r * p->f_shift)) * r;
            ec = ec + (0.5 * prefactor) * ((erfcc - r * p->e_shift) - dr2 * p->f_shift);
            double forcelj = 0.0;
            if (dr2 < p->rcut2) {
                const double dr6i = (dr2i * dr2i) * dr2i;
                const int off = it * p->ntypes + jt;
                forcelj = dr6i * (lj14_1[off] * dr6i - lj14_2[off]);
                pe = pe + dr6i * (dr6i * lj14_3[off] - lj14_4[off]);
            }
            force = (forcecoul + forcelj) * dr2i;
        }
        ffh[i1].x = (float)((double)ffh[i1].x + ddx * force);
        ffh[i1].y = (float)((double)ffh[i1].y + ddy * force);
        ffh[i1].z = (float)((double)ffh[i1].z + ddz * force);
        ffh[i4].x = (float)((double)ffh[i4].x - ddx * force);
        ffh[i4].y = (float)((double)ffh[i4].y - ddy * force);
        ffh[i4].z = (float)((double)ffh[i4].z - ddz * force);
    }
    *e_coul_14 = ec;
    *pot_14 = pe;
}

/* ------------------------------------------------------------------------------------
 * nve_v, Source/mod_ensemble.f90:11-40: v += (dt_partial/m) * ff (double math, real*4
 * store); ke = 0.5*amu2e*sum m v.v from the updated velocities (0.50 is real*4).
 * ---------------------------------------------------------------------------------- */
void orc_nve_v(int32_t np, float dt_partial, const orc_ainfo *pos, const double *mass,
               orc_coords *v, const orc_coords *ff, double amu2e, double *ke) {
    double k = 0.0;
    for (int i = 0; i < np; ++i) {
        const double m = mass[pos[i].type - 1];
        const double dtmass = (double)dt_partial / m;
        v[i].x = (float)((double)v[i].x + dtmass * (double)ff[i].x);
        v[i].y = (float)((double)v[i].y + dtmass * (double)ff[i].y);
        v[i].z = (float)((double)v[i].z + dtmass * (double)ff[i].z);
        k = k + m * (double)((v[i].x * v[i].x + v[i].y * v[i].y) + v[i].z * v[i].z);
    }
    *ke = (k * 0.5) * amu2e;
}

/* nve_x, Source/mod_ensemble.f90:48-87: dx = dt*v (real*4 product), drx += dx,
 * position += dx (real*4 stores). */
void orc_nve_x(int32_t np, float dt, const orc_coords *v, orc_ainfo *pos, float *drx,
               float *dry, float *drz) {
    for (int i = 0; i < np; ++i) {
        const float dx = dt * v[i].x, dy = dt * v[i].y, dz = dt * v[i].z;
        drx[i] = drx[i] + dx;
        dry[i] = dry[i] + dy;
        drz[i] = drz[i] + dz;
        pos[i].x = pos[i].x + dx;
        pos[i].y = pos[i].y + dy;
        pos[i].z = pos[i].z + dz;
    }
}

/* pbc, Source/mod_pbc.f90:6-39: wrap each coordinate into (0, box] with while loops. */
void orc_pbc(int32_t np, orc_ainfo *pos, float box) {
    for (int i = 0; i < np; ++i) {
        float *c[3] = {&pos[i].x, &pos[i].y, &pos[i].z};
        for (int d = 0; d < 3; ++d) {
            while (*c[d] > box) *c[d] = *c[d] - box;
            while ((double)*c[d] <= 0.0) *c[d] = *c[d] + box;
        }
    }
}

/* berendsen, Source/mod_ensemble.f90:167-205, with the scale factor mu given: each
 * coordinate becomes real*4(dble(x)*mu), drx/dry/drz add (new - dble(old)) in real*8 and
 * store real*4; box = real*4(box*mu), ibox = 1.0/box in real*4. */
void orc_berendsen(int32_t np, orc_ainfo *pos, float *drx, float *dry, float *drz, double mu,
                   float *box, float *ibox) {
    for (int i = 0; i < np; ++i) {
        float *c[3] = {&pos[i].x, &pos[i].y, &pos[i].z};
        float *dr[3] = {&drx[i], &dry[i], &drz[i]};
        for (int d = 0; d < 3; ++d) {
            const double cur = (double)*c[d];
            *c[d] = (float)(cur * mu);
            *dr[d] = (float)((double)*dr[d] + ((double)*c[d] - cur));
        }
    }
    *box = (float)((double)*box * mu);
    *ibox = 1.0f / *box;
}

/* langevin_debug_intel, Source/mod_langevin.f90:15-43, with the three uniform draws per
 * atom supplied by the caller: ff_lang = -fric*m*v + gconst*sqrt(m)*U(-1/2, 1/2). */
void orc_langevin(int32_t np, const orc_ainfo *pos, const orc_coords *v, const double *mass,
                  float fric, float gconst, const float *r1, const float *r2, const float *r3,
                  orc_coords *ff_lang) {
    for (int j = 0; j < np; ++j) {
        const double rmass = mass[pos[j].type - 1];
        const double gamma1 = -(double)fric * rmass;
        const double gamma2 = (double)gconst * sqrt(rmass);
        ff_lang[j].x = (float)(gamma1 * (double)v[j].x + gamma2 * (double)r1[j]);
        ff_lang[j].y = (float)(gamma1 * (double)v[j].y + gamma2 * (double)r2[j]);
        ff_lang[j].z = (float)(gamma1 * (double)v[j].z + gamma2 * (double)r3[j]);
    }
}

/* ------------------------------------------------------------------------------------
 * SHAKE, Source/mod_shake.f90:8-913.  shake_update predicts the unconstrained position
 * x' = x + dt*v + c*F/m (c = dt^2*kcal2amu, or half of it when iflag = 0); each cluster
 * then solves for its Lagrange multipliers -- closed-form quadratic for one bond
 * (shake2), linearised iteration with the quadratic terms for three atoms (shake3),
 * four atoms (shake4) and the rigid triangle (shake3angle) -- and adds lambda*r_ij/c to
 * ff (real*4 accumulation) and to the virial.  Returns 0, or -1 when a determinant is
 * negative/zero (the reference stops there).
 * ---------------------------------------------------------------------------------- */
typedef struct { double x, y, z; } d3;

static inline void mi3(double *r, float box, float ibox) {
    r[0] = minimg_nint(r[0], box, ibox);
    r[1] = minimg_nint(r[1], box, ibox);
    r[2] = minimg_nint(r[2], box, ibox);
}
static inline double dot3(const double *a, const double *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

int32_t orc_shake(int32_t iflag, int32_t np, const orc_ainfo *pos, const orc_coords *v,
                  orc_coords *ff, const double *mass, float dt, float kcal2amu, float box,
                  float ibox, int32_t nshake, const int32_t *num, const int32_t *atoms,
                  const double *bond, double tol, int32_t maxiter, double *virial,
                  int32_t *max_iter_seen) {
    /* shake_update: dtsq_factor computed in real*4 (0.5*dt*dt*kcal2amu) */
    const double dtsq = (iflag == 0) ? (double)(((0.5f * dt) * dt) * kcal2amu)
                                     : (double)((dt * dt) * kcal2amu);
    d3 *xs = (d3 *)malloc(sizeof(d3) * (size_t)np);
    for (int i = 0; i < np; ++i) {
        const double im = 1.0 / mass[pos[i].type - 1];
        xs[i].x = (double)(pos[i].x + dt * v[i].x) + (dtsq * im) * (double)ff[i].x;
        xs[i].y = (double)(pos[i].y + dt * v[i].y) + (dtsq * im) * (double)ff[i].y;
        xs[i].z = (double)(pos[i].z + dt * v[i].z) + (dtsq * im) * (double)ff[i].z;
    }
    int status = 0, maxit = 0;
    double vx = virial[0], vy = virial[1], vz = virial[2];
#define RVEC(out, a, b)                                                   \
    do {                                                                  \
        out[0] = (double)(pos[a].x - pos[b].x);                           \
        out[1] = (double)(pos[a].y - pos[b].y);                           \
        out[2] = (double)(pos[a].z - pos[b].z);                           \
        mi3(out, box, ibox);                                              \
    } while (0)
#define SVEC(out, a, b)                                                   \
    do {                                                                  \
        out[0] = xs[a].x - xs[b].x;                                       \
        out[1] = xs[a].y - xs[b].y;                                       \
        out[2] = xs[a].z - xs[b].z;                                       \
        mi3(out, box, ibox);                                              \
    } while (0)
#define ADDF(a, s, r)                                                     \
    do {                                                                  \
        ff[a].x = (float)((double)ff[a].x + (s) * (r)[0]);                \
        ff[a].y = (float)((double)ff[a].y + (s) * (r)[1]);                \
        ff[a].z = (float)((double)ff[a].z + (s) * (r)[2]);                \
    } while (0)
    for (int c = 0; c < nshake && status == 0; ++c) {
        const int n = num[c];
        const int i0 = atoms[4 * c] - 1, i1 = atoms[4 * c + 1] - 1;
        const int i2 = atoms[4 * c + 2] - 1, i3 = atoms[4 * c + 3] - 1;
        const double b1 = bond[3 * c], b2 = bond[3 * c + 1], b3 = bond[3 * c + 2];
        const double im0 = 1.0 / mass[pos[i0].type - 1];
        const double im1 = 1.0 / mass[pos[i1].type - 1];
        if (n == 1) { /* shake2, :89-187 */
            double r01[3], s01[3];
            RVEC(r01, i0, i1);
            SVEC(s01, i0, i1);
            const double r01sq = dot3(r01, r01), s01sq = dot3(s01, s01);
            const double a = ((im0 + im1) * (im0 + im1)) * r01sq;
            const double bb = (2.0 * (im0 + im1)) * dot3(s01, r01);
            const double cc = s01sq - b1 * b1;
            double determ = bb * bb - (4.0 * a) * cc;
            if (determ < 0.0) { status = -1; break; }
            const double l1 = (-bb + sqrt(determ)) / (2.0 * a);
            const double l2 = (-bb - sqrt(determ)) / (2.0 * a);
            double lam = (fabs(l1) <= fabs(l2)) ? l1 : l2;
            lam = lam / dtsq;
            ADDF(i0, lam, r01);
            ADDF(i1, -lam, r01);
            vx += (r01[0] * r01[0]) * lam; vy += (r01[1] * r01[1]) * lam; vz += (r01[2] * r01[2]) * lam;
        } else if (n == 3) { /* shake3, :194-355 */
            const double im2 = 1.0 / mass[pos[i2].type - 1];
            double r01[3], r02[3], s01[3], s02[3];
            RVEC(r01, i0, i1); RVEC(r02, i0, i2);
            SVEC(s01, i0, i1); SVEC(s02, i0, i2);
            const double s01sq = dot3(s01, s01), s02sq = dot3(s02, s02);
            const double a11 = (2.0 * (im0 + im1)) * dot3(s01, r01);
            const double a12 = (2.0 * im0) * dot3(s01, r02);
            const double a21 = (2.0 * im0) * dot3(s02, r01);
            const double a22 = (2.0 * (im0 + im2)) * dot3(s02, r02);
            const double determ = a11 * a22 - a12 * a21;
            if (determ == 0.0) { status = -1; break; }
            const double di = 1.0 / determ;
            const double a11i = a22 * di, a12i = -a12 * di, a21i = -a21 * di, a22i = a11 * di;
            const double r01sq = dot3(r01, r01), r02sq = dot3(r02, r02), r0102 = dot3(r01, r02);
            const double q1_11 = ((im0 + im1) * (im0 + im1)) * r01sq;
            const double q1_22 = (im0 * im0) * r02sq;
            const double q1_12 = ((2.0 * (im0 + im1)) * im0) * r0102;
            const double q2_22 = ((im0 + im2) * (im0 + im2)) * r02sq;
            const double q2_11 = (im0 * im0) * r01sq;
            const double q2_12 = ((2.0 * (im0 + im2)) * im0) * r0102;
            double l01 = 0.0, l02 = 0.0;
            int it = 0, done = 0;
            while (!done && it < maxiter) {
                const double qd1 = (q1_11 * l01) * l01 + (q1_22 * l02) * l02 + (q1_12 * l01) * l02;
                const double qd2 = (q2_11 * l01) * l01 + (q2_22 * l02) * l02 + (q2_12 * l01) * l02;
                const double bb1 = (b1 * b1 - s01sq) - qd1;
                const double bb2 = (b2 * b2 - s02sq) - qd2;
                const double n1 = a11i * bb1 + a12i * bb2;
                const double n2 = a21i * bb1 + a22i * bb2;
                done = 1;
                if (fabs(n1 - l01) > tol) done = 0;
                if (fabs(n2 - l02) > tol) done = 0;
                l01 = n1; l02 = n2;
                ++it;
            }
            if (it > maxit) maxit = it;
            l01 /= dtsq; l02 /= dtsq;
            ff[i0].x = (float)(((double)ff[i0].x + l01 * r01[0]) + l02 * r02[0]);
            ff[i0].y = (float)(((double)ff[i0].y + l01 * r01[1]) + l02 * r02[1]);
            ff[i0].z = (float)(((double)ff[i0].z + l01 * r01[2]) + l02 * r02[2]);
            ADDF(i1, -l01, r01);
            ADDF(i2, -l02, r02);
            vx += (r01[0] * r01[0]) * l01; vy += (r01[1] * r01[1]) * l01; vz += (r01[2] * r01[2]) * l01;
            vx += (r02[0] * r02[0]) * l02; vy += (r02[1] * r02[1]) * l02; vz += (r02[2] * r02[2]) * l02;
        } else if (n == 4 || n == -3) { /* shake4 :361-628, shake3angle :637-880 */
            const double im2 = 1.0 / mass[pos[i2].type - 1];
            const int tri = (n == -3);
            const double im3 = tri ? 0.0 : 1.0 / mass[pos[i3].type - 1];
            /* constraint k couples (p_k, q_k); shake4: (0,1),(0,2),(0,3); triangle: (0,1),(0,2),(1,2) */
            const int P[3] = {0, 0, tri ? 1 : 0};
            const int Q[3] = {1, 2, tri ? 2 : 3};
            const int idx[4] = {i0, i1, i2, tri ? i2 : i3};
            const double im[4] = {im0, im1, im2, im3};
            double r[3][3], s[3][3];
            for (int k = 0; k < 3; ++k) {
                RVEC(r[k], idx[P[k]], idx[Q[k]]);
                SVEC(s[k], idx[P[k]], idx[Q[k]]);
            }
            /* coupling coefficient between constraints k and l: d/dlambda_l of the
             * displacement of bond k = (delta(pk,pl) - delta(pk,ql)) im[pk]
             *                          - (delta(qk,pl) - delta(qk,ql)) im[qk] */
            double cpl[3][3];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    double c = 0.0;
                    if (P[k] == P[l]) c += im[P[k]];
                    if (P[k] == Q[l]) c -= im[P[k]];
                    if (Q[k] == P[l]) c -= im[Q[k]];
                    if (Q[k] == Q[l]) c += im[Q[k]];
                    cpl[k][l] = c;
                }
            double A[3][3];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) A[k][l] = (2.0 * cpl[k][l]) * dot3(s[k], r[l]);
            const double bl[3] = {b1, b2, b3};
            const double ssq[3] = {dot3(s[0], s[0]), dot3(s[1], s[1]), dot3(s[2], s[2])};
            const double determ = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
                                  A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1] -
                                  A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
            if (determ == 0.0) { status = -1; break; }
            const double di = 1.0 / determ;
            double Ai[3][3];
            Ai[0][0] = di * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
            Ai[0][1] = -di * (A[0][1] * A[2][2] - A[0][2] * A[2][1]);
            Ai[0][2] = di * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
            Ai[1][0] = -di * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
            Ai[1][1] = di * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
            Ai[1][2] = -di * (A[0][0] * A[1][2] - A[0][2] * A[1][0]);
            Ai[2][0] = di * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
            Ai[2][1] = -di * (A[0][0] * A[2][1] - A[0][1] * A[2][0]);
            Ai[2][2] = di * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
            /* quadratic terms: |sum_l cpl[k][l] lambda_l r_l|^2 */
            double G[3][3];
            for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) G[l][m] = dot3(r[l], r[m]);
            double lam[3] = {0.0, 0.0, 0.0};
            int it = 0, done = 0;
            while (!done && it < maxiter) {
                double bb[3], nl[3];
                for (int k = 0; k < 3; ++k) {
                    double qd = 0.0;
                    for (int l = 0; l < 3; ++l)
                        for (int m = 0; m < 3; ++m)
                            qd += (cpl[k][l] * cpl[k][m]) * G[l][m] * lam[l] * lam[m];
                    bb[k] = (bl[k] * bl[k] - ssq[k]) - qd;
                }
                for (int k = 0; k < 3; ++k)
                    nl[k] = (Ai[k][0] * bb[0] + Ai[k][1] * bb[1]) + Ai[k][2] * bb[2];
                done = 1;
                for (int k = 0; k < 3; ++k) {
                    if (fabs(nl[k] - lam[k]) > tol) done = 0;
                    lam[k] = nl[k];
                }
                ++it;
            }
            if (it > maxit) maxit = it;
            for (int k = 0; k < 3; ++k) {
                const double l = lam[k] / dtsq;
                ADDF(idx[P[k]], l, r[k]);
                ADDF(idx[Q[k]], -l, r[k]);
                vx += (r[k][0] * r[k][0]) * l;
                vy += (r[k][1] * r[k][1]) * l;
                vz += (r[k][2] * r[k][2]) * l;
            }
        }
    }
#undef RVEC
#undef SVEC
#undef ADDF
    free(xs);
    virial[0] = vx; virial[1] = vy; virial[2] = vz;
    if (max_iter_seen) *max_iter_seen = maxit;
    return status;
}

/* ------------------------------------------------------------------------------------
 * Bonded terms, restated from the energy expressions (FP64 throughout, nint minimum
 * image).  Forces are accumulated into a double (np, 3) array.
 *   bond:    E = k (r - r0)^2                    Source/mod_force_bond.f90:24-45
 *   angle:   E = k (theta - theta0)^2            Source/mod_force_angle.f90:29-90
 *   torsion: E = K (1 + cos(n phi - shift))      Source/mod_force_dihed_amber.f90
 * ---------------------------------------------------------------------------------- */
static void dvec(const orc_ainfo *pos, int a, int b, float box, float ibox, double *o) {
    o[0] = minimg_nint((double)(pos[a].x - pos[b].x), box, ibox);
    o[1] = minimg_nint((double)(pos[a].y - pos[b].y), box, ibox);
    o[2] = minimg_nint((double)(pos[a].z - pos[b].z), box, ibox);
}

void orc_bond_harmonic(int32_t nb, const int32_t *bondlist, const orc_ainfo *pos,
                       const double *kb, const double *r0, float box, float ibox,
                       double *f, double *e_bond) {
    double e = *e_bond;
    for (int m = 0; m < nb; ++m) {
        const int i1 = bondlist[3 * m] - 1, i2 = bondlist[3 * m + 1] - 1, t = bondlist[3 * m + 2] - 1;
        double d[3];
        dvec(pos, i1, i2, box, ibox, d);
        const double r = sqrt(dot3(d, d));
        const double dr = r - r0[t];
        const double rk = kb[t] * dr;
        e += rk * dr;
        const double force = (r > 0.0) ? -2.0 * rk / r : 0.0;
        for (int c = 0; c < 3; ++c) {
            f[3 * i1 + c] += d[c] * force;
            f[3 * i2 + c] -= d[c] * force;
        }
    }
    *e_bond = e;
}

void orc_angle_harmonic(int32_t na, const int32_t *anglelist, const orc_ainfo *pos,
                        const double *k, const double *theta0, float box, float ibox,
                        double *f, double *e_angle) {
    double e = *e_angle;
    for (int m = 0; m < na; ++m) {
        const int i1 = anglelist[4 * m] - 1, i2 = anglelist[4 * m + 1] - 1;
        const int i3 = anglelist[4 * m + 2] - 1, t = anglelist[4 * m + 3] - 1;
        double d1[3], d2[3];
        dvec(pos, i1, i2, box, ibox, d1);
        dvec(pos, i3, i2, box, ibox, d2);
        const double rsq1 = dot3(d1, d1), rsq2 = dot3(d2, d2);
        const double r1 = sqrt(rsq1), r2 = sqrt(rsq2);
        double c = dot3(d1, d2) / (r1 * r2);
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        double s = sqrt(1.0 - c * c);
        if (s < 0.001) s = 0.001;
        s = 1.0 / s;
        const double dth = acos(c) - theta0[t];
        const double tk = k[t] * dth;
        e += tk * dth;
        const double a = 2.0 * tk * s;
        const double a11 = a * c / rsq1, a12 = -a / (r1 * r2), a22 = a * c / rsq2;
        for (int q = 0; q < 3; ++q) {
            const double v1 = a11 * d1[q] + a12 * d2[q];
            const double v2 = a22 * d2[q] + a12 * d1[q];
            f[3 * i1 + q] -= v1;
            f[3 * i2 + q] += v1 + v2;
            f[3 * i3 + q] -= v2;
        }
    }
    *e_angle = e;
}

void orc_torsion_cos(int32_t nd, const int32_t *torsions, const orc_ainfo *pos,
                     const double *k, const double *mult, const double *shift_deg, float box,
                     float ibox, double *f, double *e_dihed) {
    double e = *e_dihed;
    const double deg = acos(-1.0) / 180.0;
    for (int m = 0; m < nd; ++m) {
        const int a = torsions[5 * m] - 1, b = torsions[5 * m + 1] - 1;
        const int c = torsions[5 * m + 2] - 1, d = torsions[5 * m + 3] - 1;
        const int t = torsions[5 * m + 4] - 1;
        double F[3], G[3], H[3];
        dvec(pos, a, b, box, ibox, F);   /* b1 = xa - xb */
        dvec(pos, b, c, box, ibox, G);   /* b2 = xb - xc */
        dvec(pos, d, c, box, ibox, H);   /* b3 = xd - xc */
        double A[3] = {F[1] * G[2] - F[2] * G[1], F[2] * G[0] - F[0] * G[2], F[0] * G[1] - F[1] * G[0]};
        double B[3] = {H[1] * G[2] - H[2] * G[1], H[2] * G[0] - H[0] * G[2], H[0] * G[1] - H[1] * G[0]};
        const double A2 = dot3(A, A), B2 = dot3(B, B), Gn = sqrt(dot3(G, G));
        if (A2 <= 0.0 || B2 <= 0.0 || Gn <= 0.0) continue;
        /* phi = atan2( (B x A).G/|G| , A.B ) (IUPAC sign) */
        double BxA[3] = {B[1] * A[2] - B[2] * A[1], B[2] * A[0] - B[0] * A[2], B[0] * A[1] - B[1] * A[0]};
        const double phi = atan2(dot3(BxA, G) / Gn, dot3(A, B));
        const double n = mult[t], dl = shift_deg[t] * deg;
        e += k[t] * (1.0 + cos(n * phi - dl));
        const double dEdphi = -k[t] * n * sin(n * phi - dl);
        /* Bekker / Blondel-Karplus gradient */
        const double FG = dot3(F, G), HG = dot3(H, G);
        double gA = -Gn / A2, gB = Gn / B2;
        double fa[3], fd[3], fb[3], fc[3];
        for (int q = 0; q < 3; ++q) {
            const double dpa = gA * A[q];           /* d phi / d xa */
            const double dpd = gB * B[q];           /* d phi / d xd */
            const double dpb = -dpa - (FG / (Gn * Gn)) * dpa - (HG / (Gn * Gn)) * dpd;
            const double dpc = -dpd + (FG / (Gn * Gn)) * dpa + (HG / (Gn * Gn)) * dpd;
            fa[q] = -dEdphi * dpa; fd[q] = -dEdphi * dpd;
            fb[q] = -dEdphi * dpb; fc[q] = -dEdphi * dpc;
        }
        for (int q = 0; q < 3; ++q) {
            f[3 * a + q] += fa[q]; f[3 * b + q] += fb[q];
            f[3 * c + q] += fc[q]; f[3 * d + q] += fd[q];
        }
    }
    *e_dihed = e;
}

/* improper_harmonic_lammps14, Source/mod_force_dihed.f90:1546-1724: E = K (chi - chi0)^2
 * with chi the angle between the planes (1,2,3) and (2,3,4) of improlist(1:4) (third atom
 * central), the LAMMPS improper_harmonic gradient, real*4 coordinate differences wrapped
 * by minimg (Source/mod_minimg.f90:5-15), and SMALL = 0.001 as a real*4 literal stored in
 * a double.  virial[3] gets v1.f1 + v2.f3 + (v3 + v2).f4 per component
 * (Source/mod_force_dihed.f90:1712-1717). */
void orc_improper_harmonic(int32_t ni, const int32_t *improlist, const orc_ainfo *pos,
                           const double *k, const double *chi0, float box, float ibox,
                           double *f, double *e_improper, double *virial) {
    const double small = (double)0.001f;
    double e = *e_improper;
    for (int m = 0; m < ni; ++m) {
        const int i1 = improlist[5 * m] - 1, i2 = improlist[5 * m + 1] - 1;
        const int i3 = improlist[5 * m + 2] - 1, i4 = improlist[5 * m + 3] - 1;
        const int t = improlist[5 * m + 4] - 1;
        double v1[3], v2[3], v3[3];
        dvec(pos, i1, i2, box, ibox, v1);
        dvec(pos, i3, i2, box, ibox, v2);
        dvec(pos, i4, i3, box, ibox, v3);
        const double ss1 = 1.0 / (v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
        const double ss2 = 1.0 / (v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
        const double ss3 = 1.0 / (v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
        const double r1 = sqrt(ss1), r2 = sqrt(ss2), r3 = sqrt(ss3);
        const double c0 = (v1[0] * v3[0] + v1[1] * v3[1] + v1[2] * v3[2]) * r1 * r3;
        const double c1 = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) * r1 * r2;
        const double c2 = -(v3[0] * v2[0] + v3[1] * v2[1] + v3[2] * v2[2]) * r3 * r2;
        double s1 = 1.0 - c1 * c1;
        if (s1 < small) s1 = small;
        s1 = 1.0 / s1;
        double s2 = 1.0 - c2 * c2;
        if (s2 < small) s2 = small;
        s2 = 1.0 / s2;
        double s12 = sqrt(s1 * s2);
        double c = (c1 * c2 + c0) * s12;
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        double s = sqrt(1.0 - c * c);
        if (s < small) s = small;
        const double domega = acos(c) - chi0[t];
        double a = k[t] * domega;
        e += a * domega;
        a = -a * 2.0 / s;
        c = c * a;
        s12 = s12 * a;
        const double a11 = c * ss1 * s1;
        const double a22 = -ss2 * (2.0 * c0 * s12 - c * (s1 + s2));
        const double a33 = c * ss3 * s2;
        const double a12 = -r1 * r2 * (c1 * c * s1 + c2 * s12);
        const double a13 = -r1 * r3 * s12;
        const double a23 = r2 * r3 * (c2 * c * s2 + c1 * s12);
        for (int q = 0; q < 3; ++q) {
            const double sx2 = a22 * v2[q] + a23 * v3[q] + a12 * v1[q];
            const double f1 = a12 * v2[q] + a13 * v3[q] + a11 * v1[q];
            const double f2 = -sx2 - f1;
            const double f4 = a23 * v2[q] + a33 * v3[q] + a13 * v1[q];
            const double f3 = sx2 - f4;
            f[3 * i1 + q] += f1; f[3 * i2 + q] += f2;
            f[3 * i3 + q] += f3; f[3 * i4 + q] += f4;
            if (virial) virial[q] = virial[q] + v1[q] * f1 + v2[q] * f3 + (v3[q] + v2[q]) * f4;
        }
    }
    *e_improper = e;
}

/* impro_amber_force, Source/mod_force_impro_amber.f90:8-230 (the cvff form):
 * E = K [1 + d cos(n phi)] with cos(n phi) the Chebyshev polynomial of c = cos(phi),
 * d = sign_impro = +-1 and n = multiplicity_impro in 0..6.  The routine keeps the energy
 * factor p in the module's real*4 pressure variable P (Source/global.f90:137), so p is
 * rounded to float before E += K p; pd stays double.  SMALL is the real*4 literal 0.001
 * and sqrt(1 - c^2) of the two bond angles is clamped from below by it. */
void orc_improper_cvff(int32_t ni, const int32_t *improlist, const orc_ainfo *pos,
                       const double *k, const int32_t *sign, const int32_t *mult, float box,
                       float ibox, double *f, double *e_improper, double *virial) {
    const double small = (double)0.001f;
    double e = *e_improper;
    for (int m = 0; m < ni; ++m) {
        const int i1 = improlist[5 * m] - 1, i2 = improlist[5 * m + 1] - 1;
        const int i3 = improlist[5 * m + 2] - 1, i4 = improlist[5 * m + 3] - 1;
        const int t = improlist[5 * m + 4] - 1;
        double b1[3], b2[3], b2m[3], b3[3];
        dvec(pos, i1, i2, box, ibox, b1);
        dvec(pos, i3, i2, box, ibox, b2);
        for (int q = 0; q < 3; ++q) b2m[q] = minimg_nint(-b2[q], box, ibox);
        dvec(pos, i4, i3, box, ibox, b3);
        const double sb1 = 1.0 / (b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
        const double sb2 = 1.0 / (b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        const double sb3 = 1.0 / (b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2]);
        const double rb1 = sqrt(sb1), rb3 = sqrt(sb3);
        const double c0 = (b1[0] * b3[0] + b1[1] * b3[1] + b1[2] * b3[2]) * rb1 * rb3;
        const double b1mag = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
        const double b2mag = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        const double b3mag = sqrt(b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2]);
        const double r12c1 = 1.0 / (b1mag * b2mag);
        const double c1mag = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) * r12c1;
        const double r12c2 = 1.0 / (b2mag * b3mag);
        const double c2mag = (b2m[0] * b3[0] + b2m[1] * b3[1] + b2m[2] * b3[2]) * r12c2;
        double sc1 = sqrt(1.0 - c1mag * c1mag);
        if (sc1 < small) sc1 = small;
        sc1 = 1.0 / sc1;
        double sc2 = sqrt(1.0 - c2mag * c2mag);
        if (sc2 < small) sc2 = small;
        sc2 = 1.0 / sc2;
        const double s1 = sc1 * sc1, s2 = sc2 * sc2;
        double s12 = sc1 * sc2;
        double c = (c0 + c1mag * c2mag) * s12;
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        float p = 0.0f;
        double pd = 0.0, rc2;
        switch (mult[t]) {
            case 2: p = (float)(2.0 * c * c); pd = 2.0 * c; break;
            case 3: rc2 = c * c; p = (float)((4.0 * rc2 - 3.0) * c + 1.0); pd = 6.0 * rc2 - 1.5; break;
            case 4: rc2 = c * c; p = (float)(8.0 * (rc2 - 1) * rc2 + 2.0); pd = (16.0 * rc2 - 8.0) * c; break;
            case 6:
                rc2 = c * c;
                p = (float)(((32.0 * rc2 - 48.0) * rc2 + 18.0) * rc2);
                pd = (96.0 * (rc2 - 1.0) * rc2 + 18.0) * c;
                break;
            case 1: p = (float)(c + 1.0); pd = 0.5; break;
            case 5:
                rc2 = c * c;
                p = (float)(((16.0 * rc2 - 20.0) * rc2 + 5.0) * c + 1.0);
                pd = (40.0 * rc2 - 30.0) * rc2 + 2.5;
                break;
            case 0: p = 2.0f; pd = 0.0; break;
            default: break;   /* the reference leaves p, pd as they were */
        }
        if (sign[t] == -1) {
            p = 2.0f - p;
            pd = -pd;
        }
        e = e + k[t] * (double)p;
        const double a = 2.0 * k[t] * pd;
        c = c * a;
        s12 = s12 * a;
        const double a11 = c * sb1 * s1;
        const double a22 = -sb2 * (2.0 * c0 * s12 - c * (s1 + s2));
        const double a33 = c * sb3 * s2;
        const double a12 = -r12c1 * (c1mag * c * s1 + c2mag * s12);
        const double a13 = -rb1 * rb3 * s12;
        const double a23 = r12c2 * (c2mag * c * s2 + c1mag * s12);
        for (int q = 0; q < 3; ++q) {
            const double sx2 = a12 * b1[q] + a22 * b2[q] + a23 * b3[q];
            const double f1 = a11 * b1[q] + a12 * b2[q] + a13 * b3[q];
            const double f2 = -sx2 - f1;
            const double f4 = a13 * b1[q] + a23 * b2[q] + a33 * b3[q];
            const double f3 = sx2 - f4;
            f[3 * i1 + q] += f1; f[3 * i2 + q] += f2;
            f[3 * i3 + q] += f3; f[3 * i4 + q] += f4;
            if (virial) virial[q] = virial[q] + b1[q] * f1 + b2[q] * f3 + (b3[q] + b2[q]) * f4;
        }
    }
    *e_improper = e;
}

/* ------------------------------------------------------------------------------------
 * Nucleation order parameter of get_greatestCluster (Source/mod_cluster.f90:15-83).
 * ---------------------------------------------------------------------------------- */

/* Molecule centres as xyzcenter_molecule forms them (Source/mod_cluster.f90:803-836: the
 * unweighted mean of the atoms of the molecule made whole) and pbc_molcom wraps them
 * (841-884).  A molecule is made whole around its atom of lowest global id: each atom's
 * offset from it is taken with the minimum image, rounded to a multiple of 2^-32 A and
 * summed in 64-bit integers, so the sum is exact and independent of the atom order; the
 * centre is ref + sum 2^-32 / count.  mol[i] (0-based, < 0 = not a solute molecule) and
 * gid[i] label the atoms; x is FP64 x, y, z per atom; com receives x, y, z per molecule. */
void orc_molecule_centers(int32_t n, const double *x, const int32_t *gid, const int32_t *mol,
                          int32_t nmol, float box, double *com) {
    int32_t *ref = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nmol > 0 ? nmol : 1));
    long long *sum = (long long *)calloc((size_t)3 * (nmol > 0 ? nmol : 1), sizeof(long long));
    int32_t *cnt = (int32_t *)calloc((size_t)(nmol > 0 ? nmol : 1), sizeof(int32_t));
    const double bx = (double)box, ib = 1.0 / bx;
    for (int m = 0; m < nmol; ++m) ref[m] = -1;
    for (int i = 0; i < n; ++i) {
        const int m = mol[i];
        if (m < 0 || m >= nmol) continue;
        if (ref[m] < 0 || gid[i] < gid[ref[m]]) ref[m] = i;
    }
    for (int i = 0; i < n; ++i) {
        const int m = mol[i];
        if (m < 0 || m >= nmol) continue;
        for (int d = 0; d < 3; ++d) {
            double off = x[3 * (size_t)i + d] - x[3 * (size_t)ref[m] + d];
            off = off - bx * rint(off * ib);
            sum[3 * (size_t)m + d] += llrint(off * 4294967296.0);
        }
        cnt[m] += 1;
    }
    for (int m = 0; m < nmol; ++m) {
        for (int d = 0; d < 3; ++d) {
            double c = cnt[m] ? x[3 * (size_t)ref[m] + d] +
                                    ((double)sum[3 * (size_t)m + d] * (1.0 / 4294967296.0)) / (double)cnt[m]
                              : 0.0;
            while (c > bx) c = c - bx;
            while (c < 0.0) c = c + bx;
            com[3 * (size_t)m + d] = c;
        }
    }
    free(ref); free(sum); free(cnt);
}

/* Squared centre distance as mol_build_neighbor_nosort forms it
 * (Source/mod_neighbor_cluster.f90:91-107): dx = x1 - x2, shift = -box*nint(dx*ibox) with
 * the real*4 product box*nint, dr2 = dx*dx + dy*dy + dz*dz. */
static double com_dr2(const double *a, const double *b, float box, float ibox) {
    double s[3];
    for (int d = 0; d < 3; ++d) {
        double dx = a[d] - b[d];
        const int nn = nint_d(dx * (double)ibox);
        const float shift = -(box * (float)nn);
        s[d] = dx + (double)shift;
    }
    return (s[0] * s[0] + s[1] * s[1]) + s[2] * s[2];
}

/* get_greatestCluster on given centres: denstrack(i) = number of other centres closer than
 * denscut (mol_build_neighbor_nosort); molecules with denstrack >= densmin are dense
 * (clusterCriterium_surface, Source/mod_cluster.f90:138-152); buildClusters numbers the
 * connected components of dense molecules, walking molecules in index order and flooding
 * each new one through its dense neighbours (harvestCluster, recursive in the reference,
 * an explicit stack here: the members do not depend on the walk order); the greatest
 * cluster is the first largest in that numbering (68-74); append_surface_molecules adds
 * the non-dense neighbours of its members (156-182).  Returns core + surface and sets
 * *core, *surface and, when given, denstrack[nmol] and belongsto[nmol] (cluster number,
 * 0 = none, surface molecules set to the greatest cluster as the reference does). */
int32_t orc_greatest_cluster(int32_t nmol, const double *com, float box, double denscut,
                             int32_t densmin, int32_t *core, int32_t *surface,
                             int32_t *denstrack_out, int32_t *belongsto_out) {
    const size_t nm = (size_t)(nmol > 0 ? nmol : 1);
    const float ibox = 1.0f / box;
    const double cut2 = denscut * denscut;
    int32_t *dens = (int32_t *)calloc(nm, sizeof(int32_t));
    int32_t *belong = (int32_t *)calloc(nm, sizeof(int32_t));
    int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * nm);
    int32_t *histo = (int32_t *)calloc(nm + 1, sizeof(int32_t));
    char *surf = (char *)calloc(nm, 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nmol; ++i) {
        int cnt = 0;
        for (int j = 0; j < nmol; ++j)
            if (j != i && com_dr2(com + 3 * (size_t)i, com + 3 * (size_t)j, box, ibox) < cut2) ++cnt;
        dens[i] = cnt;
    }
    int ncl = 0;
    for (int ti = 0; ti < nmol; ++ti) {
        if (belong[ti] != 0 || dens[ti] < densmin) continue;
        belong[ti] = ++ncl;
        int top = 0;
        stack[top++] = ti;
        while (top > 0) {
            const int a = stack[--top];
            for (int j = 0; j < nmol; ++j) {
                if (j == a || belong[j] != 0 || dens[j] < densmin) continue;
                if (com_dr2(com + 3 * (size_t)a, com + 3 * (size_t)j, box, ibox) < cut2) {
                    belong[j] = ncl;
                    stack[top++] = j;
                }
            }
        }
    }
    for (int i = 0; i < nmol; ++i)
        if (belong[i]) histo[belong[i]] += 1;
    int best = 0, gb = 0;
    for (int k = 1; k <= ncl; ++k)
        if (histo[k] > best) { best = histo[k]; gb = k; }
    int ns = 0;
    if (gb)
        for (int i = 0; i < nmol; ++i) {
            if (belong[i] != gb) continue;
            for (int j = 0; j < nmol; ++j)
                if (j != i && dens[j] < densmin &&
                    com_dr2(com + 3 * (size_t)i, com + 3 * (size_t)j, box, ibox) < cut2)
                    surf[j] = 1;
        }
    for (int j = 0; j < nmol; ++j)
        if (surf[j]) { ++ns; belong[j] = gb; }
    if (denstrack_out) memcpy(denstrack_out, dens, sizeof(int32_t) * (size_t)(nmol > 0 ? nmol : 0));
    if (belongsto_out) memcpy(belongsto_out, belong, sizeof(int32_t) * (size_t)(nmol > 0 ? nmol : 0));
    *core = best;
    *surface = ns;
    free(dens); free(belong); free(stack); free(histo); free(surf);
    return best + ns;
}
