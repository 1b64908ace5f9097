/*
 * CPU oracle for the Sofia hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference Fortran routines of
 * coparks2012/MOLECULAR_SIMULATION_CODE that the HIP library replaces.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
 * the checker (or the timed CPU baseline); the product path never calls it.
 *
 * Conventions follow Source/global.f90: atom ids stored in lists are 1-based, cell ids
 * are 0-based, positions are `ainfo` (x, y, z real*4 + integer type, 16 B) and forces
 * and velocities are `coords` (3 x real*4).  Every function cites the reference lines
 * it follows; evaluation order and float/double promotion follow the Fortran source
 * (compile with -ffp-contract=off so no FMA is formed).
 *
 * Parity pins (see tests/): every routine is checked against either golden vectors
 * produced by the amdflang-compiled reference routines (oracle/build_ref.sh), a
 * brute-force twin, or closed forms.
 */
#ifndef SOFIA_ORACLE_H
#define SOFIA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; int32_t type; } orc_ainfo;   /* Source/global.f90:82-85 */
typedef struct { float x, y, z; } orc_coords;                /* Source/global.f90:87-89 */

/* Scalars shared by the pair routines; precisions as in Source/global.f90:69-73,128-137 */
typedef struct {
    float box, ibox;            /* real*4 */
    float coulpre;              /* real*4 */
    int32_t ntypes;
    double rcut2, cut_coulsq;   /* real*8 (from real*4 products) */
    double rv2;
    double alpha, gewald;
    double e_shift, f_shift;
    double ewald_p, ewald_f;
    double a1, a2, a3, a4, a5;
    double my_pis, my_pis_inv;
} orc_params;

/* --- cell grid, binning, sort (Source/mod_neighbor.f90) --- */
void orc_cell_grid(float box, float rcut, float cut_coul, float *rn, int32_t *ncellT,
                   int32_t *ncellD);
void orc_neighcell(int32_t ncellT, int32_t ncellD, int32_t *cnum);
int32_t orc_xyz2bin(float x, float y, float z, float rn, float box, int32_t ncellD);
void orc_sort_cells(int32_t np, const orc_ainfo *pos, float rn, float box, int32_t ncellT,
                    int32_t ncellD, int32_t *tt, int32_t *perm, int32_t *start,
                    int32_t *endposit, int32_t *atombin);

/* --- Verlet list (Source/mod_neighbor.f90:90-303, 814-861, 895-913) --- */
int32_t orc_build_neighbor(int32_t np, const orc_ainfo *pos, const int32_t *atombin,
                           const int32_t *start, const int32_t *endposit, const int32_t *cnum,
                           const int32_t *num3bond, const int32_t *specbond, float box,
                           float ibox, double rv2, int32_t neigh_alloc, int32_t *nlist,
                           int32_t *numneigh);
int32_t orc_build_neighbor_n2(int32_t np, const orc_ainfo *pos, const int32_t *num3bond,
                              const int32_t *specbond, float box, float ibox, double rv2,
                              int32_t neigh_alloc, int32_t *nlist, int32_t *numneigh);
int32_t orc_check_neighbor(int32_t np, const float *drx, const float *dry, const float *drz,
                           double rvrc2);

/* --- nonbond pair kernels (Source/mod_force_nonbond.f90) --- */
void orc_dsf(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *nlist,
             const int32_t *numneigh, int32_t neigh_alloc, const float *lj1, const float *lj2,
             const float *lj3, const float *lj4, const orc_params *p, orc_coords *ff,
             double *ff64, double *potential, double *e_coul, double *virial);
/* RDF_calc, Source/mod_force_nonbond.f90:708-775 (bins past nbins are not counted) */
void orc_rdf_calc(int32_t np, const orc_ainfo *pos, const int32_t *nlist,
                  const int32_t *numneigh, int32_t neigh_alloc, float box, float ibox,
                  double rcut2, float rdf_delta, int32_t itype_sel, int32_t jtype_sel,
                  int32_t nbins, float *rdf, int32_t *numob_rdf);
/* orc_dsf on FP64 positions (xyz: x, y, z per atom; type per atom): the same arithmetic
 * from the promotion of the coordinates on, so on real*4-representable positions it equals
 * orc_dsf bit for bit.  ff64 receives the per-atom sums. */
void orc_dsf_x64(int32_t np, const double *xyz, const int32_t *type, const float *q,
                 const int32_t *nlist, const int32_t *numneigh, int32_t neigh_alloc,
                 const float *lj1, const float *lj2, const float *lj3, const float *lj4,
                 const orc_params *p, double *ff64, double *potential, double *e_coul,
                 double *virial);
void orc_coul_long(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *nlist,
                   const int32_t *numneigh, int32_t neigh_alloc, const float *lj1,
                   const float *lj2, const float *lj3, const float *lj4, const orc_params *p,
                   orc_coords *ff, double *potential, double *e_coul, double *virial);
void orc_dsf_n2(int32_t np, const orc_ainfo *pos, const float *q, const int32_t *num2bond,
                const int32_t *num3bond, const int32_t *specbond, const float *lj1,
                const float *lj2, const float *lj3, const float *lj4, const orc_params *p,
                double *ff64, double *potential, double *e_coul);

/* --- exclusion corrections and 1-4 pair --- */
void orc_correct_pairs(int32_t npairs, const int32_t *ia, const int32_t *ib,
                       const orc_ainfo *pos, const float *q, const orc_params *p,
                       double fudge, orc_coords *ff, double *e_coul, double *virial);
void orc_pair14_dsf(int32_t nd, const int32_t *dihedlist, const orc_ainfo *pos,
                    const float *q, const double *lj14_1, const double *lj14_2,
                    const double *lj14_3, const double *lj14_4, const orc_params *p,
                    orc_coords *ffh, double *e_coul_14, double *pot_14);

/* --- integration, wrapping, Langevin, SHAKE (FP64 math on real*4 storage) --- */
void orc_nve_v(int32_t np, float dt_partial, const orc_ainfo *pos, const double *mass,
               orc_coords *v, const orc_coords *ff, double amu2e, double *ke);
void orc_nve_x(int32_t np, float dt, const orc_coords *v, orc_ainfo *pos, float *drx,
               float *dry, float *drz);
void orc_pbc(int32_t np, orc_ainfo *pos, float box);
void orc_berendsen(int32_t np, orc_ainfo *pos, float *drx, float *dry, float *drz, double mu,
                   float *box, float *ibox);
void orc_langevin(int32_t np, const orc_ainfo *pos, const orc_coords *v, const double *mass,
                  float fric, float gconst, const float *r1, const float *r2, const float *r3,
                  orc_coords *ff_lang);
int32_t orc_shake(int32_t iflag, int32_t np, const orc_ainfo *pos, const orc_coords *v,
                  orc_coords *ff, const double *mass, float dt, float kcal2amu, float box,
                  float ibox, int32_t nshake, const int32_t *num, const int32_t *atoms,
                  const double *bond, double tolerance, int32_t maxiter, double *virial,
                  int32_t *max_iter_seen);

/* --- bonded terms (Source/mod_force_bond.f90, mod_force_angle.f90,
 *     mod_force_dihed_amber.f90 cosine series, mod_force_dihed.f90 improper) --- */
void orc_bond_harmonic(int32_t nb, const int32_t *bondlist, const orc_ainfo *pos,
                       const double *kb, const double *r0, float box, float ibox,
                       double *f, double *e_bond);
void orc_angle_harmonic(int32_t na, const int32_t *anglelist, const orc_ainfo *pos,
                        const double *k, const double *theta0, float box, float ibox,
                        double *f, double *e_angle);
void orc_torsion_cos(int32_t nd, const int32_t *torsions, const orc_ainfo *pos,
                     const double *k, const double *mult, const double *shift_deg, float box,
                     float ibox, double *f, double *e_dihed);
/* improper_harmonic_lammps14 (Source/mod_force_dihed.f90:1546-1724); improlist(5, ni) =
 * i1, i2, i3 (central), i4, type; k, chi0 (rad) per type; virial may be NULL. */
void orc_improper_harmonic(int32_t ni, const int32_t *improlist, const orc_ainfo *pos,
                           const double *k, const double *chi0, float box, float ibox,
                           double *f, double *e_improper, double *virial);
/* impro_amber_force (Source/mod_force_impro_amber.f90:8-230): E = K [1 + d cos(n phi)];
 * k, sign (d = +-1) and mult (n = 0..6) per type; virial may be NULL. */
void orc_improper_cvff(int32_t ni, const int32_t *improlist, const orc_ainfo *pos,
                       const double *k, const int32_t *sign, const int32_t *mult, float box,
                       float ibox, double *f, double *e_improper, double *virial);

/* --- nucleation order parameter: get_greatestCluster (Source/mod_cluster.f90:15-83) --- */
void orc_molecule_centers(int32_t n, const double *x, const int32_t *gid, const int32_t *mol,
                          int32_t nmol, float box, double *com);
int32_t orc_greatest_cluster(int32_t nmol, const double *com, float box, double denscut,
                             int32_t densmin, int32_t *core, int32_t *surface,
                             int32_t *denstrack, int32_t *belongsto);

#ifdef __cplusplus
}
#endif
#endif
