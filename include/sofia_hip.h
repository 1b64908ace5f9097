/*
 * sofia_hip.h -- C ABI of the MI355X (gfx950) implementation of Sofia's hot path:
 * cell binning + Verlet-list build (Source/mod_neighbor.f90), LJ + damped-shifted-force
 * pair forces (Source/mod_force_nonbond.f90), exclusion corrections
 * (Source/mod_force_correction.f90) and the scaled 1-4 pair of dihedral_opls_DSF
 * (Source/mod_force_dihed.f90:1003-1048), with FP64 velocity Verlet, PBC, Langevin and
 * SHAKE (Source/mod_ensemble.f90, mod_pbc.f90, mod_langevin.f90, mod_shake.f90).
 *
 * Plain pointers and sizes only.  Every function returns SOFIA_OK (0) or a negative
 * status; sofia_last_error() gives the message of the last failure on this thread.
 *
 * Data conventions are the reference's (Source/global.f90): per-atom arrays are in the
 * current (cell-sorted) order, atom ids stored in lists are 1-based, `sofia_ainfo` is
 * `type ainfo` (x, y, z real*4 + integer type, 16 B), `sofia_coords` is `type coords`
 * (3 x real*4), specbond is Fortran's column-major specbond(12, np).
 *
 * Two groups of entry points:
 *  (1) drop-ins for the reference subroutines, operating on one system whose state the
 *      host uploads (sofia_put_state) and reads back -- each cites the routine it replaces;
 *  (2) the device-resident engine: many independent replicas (one system = a batch of
 *      one) advanced in fused, hipGraph-captured MD segments, in FP64 or mixed precision.
 * A context is used from one host thread; work is ordered on the context's stream.
 */
#ifndef SOFIA_HIP_H
#define SOFIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SOFIA_OK = 0,
    SOFIA_E_NODEV = -1,      /* no HIP device */
    SOFIA_E_OOM = -2,        /* device allocation failed */
    SOFIA_E_ARG = -3,        /* invalid argument or size */
    SOFIA_E_OVERFLOW = -4,   /* Verlet list overflow that could not be grown */
    SOFIA_E_HIP = -5,        /* HIP runtime error */
    SOFIA_E_SHAKE = -6,      /* constraint solve failed: SHAKE determinant <= 0 (the
                                reference stops) or a rigid water beyond SETTLE's reach */
    SOFIA_E_STATE = -7       /* call out of order (e.g. no list yet) */
};

enum { SOFIA_PREC_FP64 = 0, SOFIA_PREC_MIXED = 1 };
enum { SOFIA_COUL_DSF = 1, SOFIA_COUL_EWALD = 2 };   /* coul_flag, input_files/input:26 */

typedef struct { float x, y, z; int32_t type; } sofia_ainfo;   /* Source/global.f90:82-85 */
typedef struct { float x, y, z; } sofia_coords;                /* Source/global.f90:87-89 */

/* Run parameters with the reference's precisions (Source/global.f90:69-137,
 * derived as in Source/get_started.f90:25-117). */
typedef struct {
    int32_t ntypes;
    float rcut, cut_coul, pad;          /* real*4 */
    float coulpre;                      /* real*4 332.06371 */
    float dt;                           /* fs, real*4 */
    float kcal2amu;                     /* real*4 0.0004184 */
    double rcut2, cut_coulsq;           /* from real*4 products */
    double rv2, rvrc2;                  /* (cut_coul+pad)^2, pad^2 */
    double alpha, gewald;
    double e_shift, f_shift;            /* DSF shifts */
    double ewald_p, ewald_f;
    double a1, a2, a3, a4, a5;          /* A&S 7.1.26 */
    double my_pis, my_pis_inv;
    double amu2e;
    double fudge_12, fudge_13, fudge_14_coul;
    double shake_tol;                   /* 1e-4 (real*4 literal) */
    int32_t shake_maxiter;              /* 25 */
    int32_t precision;                  /* SOFIA_PREC_* (engine) */
    int32_t coul_flag;                  /* SOFIA_COUL_* (engine) */
    int32_t neigh_alloc0;               /* initial list stride (512, Source/mod_memory.f90:137) */
} sofia_params;

typedef struct sofia_ctx sofia_ctx;

/* ---------------------------------------------------------------- lifecycle */
int sofia_last_error(char *buf, size_t len);
int sofia_device_count(int *count);
/* Free and total device memory of the current GPU (hipMemGetInfo), for callers that size a
 * batch or an admission budget by HBM.  No counterpart in the reference. */
int sofia_device_memory(size_t *free_bytes, size_t *total_bytes);
/* One system of np atoms (a batch of one). */
int sofia_ctx_create(const sofia_params *p, int32_t np, float box, sofia_ctx **out);
/* nrep replicas; natoms[r] atoms and box[r] for replica r (atoms stored back to back). */
int sofia_ctx_create_batch(const sofia_params *p, int32_t nrep, const int32_t *natoms,
                           const float *box, sofia_ctx **out);
int sofia_ctx_destroy(sofia_ctx *ctx);
int sofia_synchronize(sofia_ctx *ctx);

/* ---------------------------------------------------------------- force field / topology
 * All ids are 1-based within a replica; arrays of a batch are concatenated replica by
 * replica.  lj1..4 real*4 and lj14_1..4 real*8 are ntypes*ntypes, index
 * (itype-1)*ntypes + jtype-1 (Source/mod_readfile.f90:595-615); mass is per type. */
int sofia_set_forcefield(sofia_ctx *ctx, const float *lj1, const float *lj2, const float *lj3,
                         const float *lj4, const double *lj14_1, const double *lj14_2,
                         const double *lj14_3, const double *lj14_4, const double *mass);
/* num1bond/num2bond/num3bond and specbond(12, np) (Source/mod_buildbond_lists.f90:80-183) */
int sofia_set_exclusions(sofia_ctx *ctx, const int32_t *num1bond, const int32_t *num2bond,
                         const int32_t *num3bond, const int32_t *specbond);
/* bondlist(3,nb), anglelist(4,na), dihedlist(5,nd) as in Source/mod_buildbond_lists.f90;
 * per-replica counts in nb_rep/na_rep/nd_rep (NULL for a single system). */
int sofia_set_bonded(sofia_ctx *ctx, const int32_t *bondlist, int32_t nb, const int32_t *nb_rep,
                     const int32_t *anglelist, int32_t na, const int32_t *na_rep,
                     const int32_t *dihedlist, int32_t nd, const int32_t *nd_rep);
/* Bonded coefficients for the engine: bond (k, r0), angle (k, theta0 rad), torsion
 * (K, n, shift deg) per type; torsions(5, nt) in geometric order a-b-c-d. */
int sofia_set_bonded_coeffs(sofia_ctx *ctx, int32_t nbt, const double *bond_k,
                            const double *bond_r0, int32_t nat, const double *angle_k,
                            const double *angle_th0, int32_t ntt, const double *tors_k,
                            const double *tors_n, const double *tors_shift,
                            const int32_t *torsions, int32_t ntor, const int32_t *ntor_rep);
/* Improper dihedrals for the engine: improper_harmonic_lammps14
 * (Source/mod_force_dihed.f90:1546-1724; force wrapper call Source/mod_force.f90:542-544)
 * E = K (chi - chi0)^2, chi the angle between the planes (i1, i2, i3) and (i2, i3, i4).
 * k, chi0 (rad) per improper type (improcoeff(1:2, :), Source/global.f90:455);
 * improlist(5, ni) = i1, i2, i3 (central), i4 (1-based per replica), type
 * (Source/global.f90:450); ni_rep per-replica counts (NULL for one replica). */
int sofia_set_impropers(sofia_ctx *ctx, int32_t nit, const double *k, const double *chi0,
                        const int32_t *improlist, int32_t ni, const int32_t *ni_rep);
/* Improper dihedrals in the cvff form of impro_amber_force, which the PME force wrappers
 * call (Source/mod_force_impro_amber.f90:8-230; Source/mod_force.f90:141-143):
 * E = K [1 + d cos(n phi)] with k_impro, sign_impro (d = +-1) and multiplicity_impro
 * (n = 0..6) per type (Source/global.f90:158-159, read at Source/mod_readfile.f90:454-456).
 * Replaces a harmonic set; improlist as for sofia_set_impropers. */
int sofia_set_impropers_cvff(sofia_ctx *ctx, int32_t nit, const double *k, const int32_t *sign,
                             const int32_t *mult, const int32_t *improlist, int32_t ni,
                             const int32_t *ni_rep);
/* SHAKE clusters: num (1, 3, 4, -3), atoms(4, ns) (0 = unused), bond lengths(3, ns)
 * (Source/global.f90:351-355 shake_info). */
int sofia_set_shake(sofia_ctx *ctx, int32_t ns, const int32_t *ns_rep, const int32_t *num,
                    const int32_t *atoms, const double *bond);

/* ---------------------------------------------------------------- state transfer */
int sofia_put_state(sofia_ctx *ctx, const sofia_ainfo *pos, const float *q,
                    const sofia_coords *v, const int32_t *global_id);
int sofia_get_state(sofia_ctx *ctx, sofia_ainfo *pos, sofia_coords *v, int32_t *global_id);
/* FP64 positions/velocities of the engine (x, y, z interleaved per atom), may be NULL */
int sofia_get_state_f64(sofia_ctx *ctx, double *x, double *v);
int sofia_put_state_f64(sofia_ctx *ctx, const double *x, const double *v);
/* drx/dry/drz displacement accumulators since the last list build */
int sofia_get_displacement(sofia_ctx *ctx, float *drx, float *dry, float *drz);
int sofia_put_displacement(sofia_ctx *ctx, const float *drx, const float *dry, const float *drz);
/* Page-locked host memory for the arrays a host hands to the drop-ins (position, v, ff,
 * drx/dry/drz): transfers from and to it run at the full host-link rate instead of through
 * the runtime's staging of pageable memory.  No counterpart in the reference, whose MIC
 * offload managed its own transfer buffers (Source/mod_force.f90:446-461).  *out is NULL
 * on failure; free with sofia_host_free. */
int sofia_host_alloc(size_t bytes, void **out);
int sofia_host_free(void *p);

/* ---------------------------------------------------------------- (1) drop-ins */
/* neighbor_wrapper(step), Source/mod_neighbor.f90:8-86: check_allocation, new_nlist,
 * sort_posit_charge (the permutation is applied to the device copies of position, q, v,
 * global_id, num*bond, specbond and the bonded/SHAKE lists), then
 * build_neighbor_nonewton with grow-and-redo (Source/mod_neighbor.f90:258-298).
 * tt (optional, np) receives tt(old) = new (1-based) so the host can re-index its own
 * arrays.  Zeroes drx/dry/drz. */
int sofia_neighbor_wrapper(sofia_ctx *ctx, int32_t step, int32_t *tt);
/* The arrays the sort permutes, read back from the device already permuted, in the layout
 * the host gave them (1-based ids within each replica, 0 = empty slot, items in their
 * original order): q (np), num1bond/num2bond/num3bond (np), specbond (12 per atom),
 * bondlist (3 per bond), anglelist (4), dihedlist (5), improlist (5), SHAKE atoms (4 per
 * cluster).  Replaces, for every array the device holds, the host-side re-index after
 * neighbor_wrapper: sort_posit_charge's q, sort_specbond_numbond, sort_bondlist,
 * sort_anglelist, sort_dihedlist, sort_improlist and sort_shakeatom
 * (Source/mod_neighbor.f90:946-1168).  NULL skips an array.  A list comes back with the
 * item count the device holds (the one given to sofia_set_bonded, sofia_set_impropers or
 * sofia_set_shake); a list the host never gave the device is not written, so the host
 * re-indexes that one itself. */
int sofia_get_topology(sofia_ctx *ctx, float *q, int32_t *num1bond, int32_t *num2bond,
                       int32_t *num3bond, int32_t *specbond, int32_t *bondlist,
                       int32_t *anglelist, int32_t *dihedlist, int32_t *improlist,
                       int32_t *shake_atoms);
/* RDF_calc, Source/mod_force_nonbond.f90:708-775, on the current list and positions: for
 * every atom of type itype and list neighbour of type jtype with dr2 < rcut2 (FP64 distance,
 * real*8 minimum image), rdf[node] += 1 with node = int(sqrt(dr2) / rdf_delta), node from 0;
 * nodes >= nbins are not counted (the reference has no bound; its call is itype = jtype = 3).
 * numob_rdf (optional) is incremented.  Counts are exact while a bin stays below 2^24. */
int sofia_rdf_calc(sofia_ctx *ctx, int32_t itype, int32_t jtype, float rdf_delta, int32_t nbins,
                   float *rdf, int32_t *numob_rdf);
/* cell ranges of the last build: start/endposit (1-based, per cell 0..ncellT-1) and
 * atombin (per atom); ncellT/ncellD/rn of replica 0 */
int sofia_get_cells(sofia_ctx *ctx, int32_t *start, int32_t *endposit, int32_t *atombin,
                    int32_t *ncellT, int32_t *ncellD, float *rn);
/* nlist (np*neigh_alloc, 1-based ids, reference layout) and numneigh; NULL skips. */
int sofia_get_nlist(sofia_ctx *ctx, int32_t *nlist, int32_t *numneigh, int32_t *neigh_alloc);
/* check_neighbor(nflag, step), Source/mod_neighbor.f90:895-913 */
int sofia_check_neighbor(sofia_ctx *ctx, int32_t *nflag);
/* lj_cut_coul_dsf_nonewton_NOMIC/_MIC, Source/mod_force_nonbond.f90:105-196 / 8-101:
 * FP64 pair math over the current list; ff (np, optional) overwritten with the real*4
 * forces, ff64 (3*np, optional) with the FP64 sums; potential/e_coul are ADDED the
 * halved sums; virial (optional, 3) receives 0.5*sum dx^2 F/r. */
int sofia_lj_cut_coul_dsf(sofia_ctx *ctx, sofia_coords *ff, double *ff64, double *potential,
                          double *e_coul, double *virial);
/* Mixed-precision variant: FP32 pair math, int64 fixed-point force sums. */
int sofia_lj_cut_coul_dsf_mixed(sofia_ctx *ctx, sofia_coords *ff, double *ff64,
                                double *potential, double *e_coul, double *virial);
/* Asynchronous form of the pair-kernel drop-in, for hosts that overlap the force call with
 * their own work as the reference's MIC offload does (`offload ... signal(sig)` /
 * `offload_wait`, Source/mod_force.f90:66-156): sofia_lj_cut_coul_dsf_async enqueues the DSF
 * pair kernel (mixed != 0: the mixed-precision kernel) and the copy of its results into
 * pinned host memory and returns; sofia_nonbond_wait waits and then delivers ff (np coords,
 * overwritten), adds to *potential / *e_coul and sets virial[3], like sofia_lj_cut_coul_dsf.
 * One call may be in flight per context. */
int sofia_lj_cut_coul_dsf_async(sofia_ctx *ctx, int32_t mixed);
int sofia_nonbond_wait(sofia_ctx *ctx, sofia_coords *ff, double *potential, double *e_coul,
                       double *virial);
/* lj_cut_coul_long_nonewton_NOMIC, Source/mod_force_nonbond.f90:300-400 (PME real space) */
int sofia_lj_cut_coul_long(sofia_ctx *ctx, sofia_coords *ff, double *potential,
                           double *e_coul, double *virial);
/* correct_12_bonds / correct_13_bonds / correct_14_bonds,
 * Source/mod_force_correction.f90:6-93 / 95-184 / 187-278; which = 12, 13 or 14.
 * ff (np) receives this term's forces (ff_bond/ff_angle/ff_dihed), e_coul and virial
 * (3) are ADDED to. */
int sofia_correct_bonds(sofia_ctx *ctx, int32_t which, sofia_coords *ff, double *e_coul,
                        double *virial);
/* The 1-4 pair inside dihedral_opls_DSF, Source/mod_force_dihed.f90:1003-1048;
 * ffh (np) receives the pair forces, e_coul_14/pot_14 are ADDED to. */
int sofia_pair14_dsf(sofia_ctx *ctx, sofia_coords *ffh, double *e_coul_14, double *pot_14);
/* The bonded terms of force_wrapper_DSF_NOMIC (Source/mod_force.f90:510-554):
 * bond_harmonic (Source/mod_force_bond.f90:6-75), angle_harmonic
 * (Source/mod_force_angle.f90:10-122), the torsion series of the dihedral routine and the
 * improper term set by sofia_set_impropers / _cvff (improper_harmonic_lammps14,
 * Source/mod_force_dihed.f90:1546-1724), with the coefficients of sofia_set_bonded_coeffs,
 * on the uploaded real*4 positions: ff (np) receives their summed forces, the energies are
 * ADDED to.  Not on a context running the engine, which evaluates these itself. */
int sofia_bonded_forces(sofia_ctx *ctx, sofia_coords *ff, double *e_bond, double *e_angle,
                        double *e_dihed, double *e_improper);
/* force_wrapper_DSF_NOMIC (Source/mod_force.f90:508-551): the DSF pair forces, the 1-4 pairs
 * and the bonded terms on the current positions in one call.  ff (np, set) is the three
 * routines' real*4 outputs merged the way the host merges them, ff = real*4(dble(pair) +
 * dble(1-4) + dble(bonded)); bit for bit what sofia_lj_cut_coul_dsf, sofia_pair14_dsf and
 * sofia_bonded_forces followed by that merge give, with one copy out instead of three.
 * e (8, added to): potential, e_coul, e_coul_14, pot_14, e_bond, e_angle, e_dihed,
 * e_improper.  vir (3, set): the pair kernel's virial. */
int sofia_force_wrapper_dsf(sofia_ctx *ctx, sofia_coords *ff, double *e, double *vir);
/* nve_v(dt_partial) / nve_x / pbc(step), Source/mod_ensemble.f90:11-87,
 * Source/mod_pbc.f90:6-39, on the real*4 state with force ff (np). */
int sofia_nve_v(sofia_ctx *ctx, float dt_partial, const sofia_coords *ff, double *ke);
int sofia_nve_x(sofia_ctx *ctx);
int sofia_pbc(sofia_ctx *ctx);
/* berendsen, Source/mod_ensemble.f90:167-205, given the scale factor the host computed
 * (mu = (1 - (dt/tau)(ptarget - pressure))^(1/3), one per replica): positions scaled by mu,
 * drx/dry/drz accumulate the displacement, box = box*mu and ibox = 1.0/box in real*4
 * (box_out, optional, receives the new boxes).  The cell grid is re-derived from the new
 * box, as check_allocation does, so the next call is sofia_neighbor_wrapper.  Drop-in only
 * (the engine runs at fixed volume). */
int sofia_berendsen(sofia_ctx *ctx, const double *mu, float *box_out);
/* shake(iflag), Source/mod_shake.f90:8-85: adds constraint forces to ff (np, in/out) and
 * to virial (3).  max_iter (optional) receives the largest iteration count. */
int sofia_shake(sofia_ctx *ctx, int32_t iflag, sofia_coords *ff, double *virial,
                int32_t *max_iter);
/* langevin_debug_intel, Source/mod_langevin.f90:15-43, with a counter-based RNG:
 * ff_lang = -fric*m*v + gconst*sqrt(m)*U(-1/2,1/2), U from Philox4x32-10 keyed by
 * (seed, step, global_id). */
int sofia_langevin(sofia_ctx *ctx, float fric, float gconst, uint64_t seed, uint64_t step,
                   sofia_coords *ff_lang);

/* ---------------------------------------------------------------- (2) engine */
typedef struct {
    double temp;        /* K */
    double gamma_ps;    /* Langevin friction, 1/ps (0 = NVE) */
    uint64_t seed;      /* RNG key */
} sofia_thermostat;

/* Per-replica thermostat (nrep entries); default is NVE. */
int sofia_set_thermostat(sofia_ctx *ctx, const sofia_thermostat *t);
/* Copy the uploaded real*4 state into the FP64 engine state, build the list and compute
 * the starting forces (force_driver(0) + shake(1), Source/mod_integrate.f90:237-240). */
int sofia_md_init(sofia_ctx *ctx);
/* Advance every replica nsteps velocity-Verlet steps (Source/mod_integrate.f90:270-308):
 * nve_v, nve_x, pbc, check_neighbor, list rebuild when flagged, forces, Langevin,
 * SHAKE, nve_v.  use_graph != 0 captures one step in a hipGraph and replays it. */
int sofia_md_run(sofia_ctx *ctx, int32_t nsteps, int32_t use_graph);
/* sofia_md_run in two halves: md_launch enqueues the steps on the context's stream and
 * returns at once; md_finish waits for them, reduces the energies and reports a list
 * overflow or SHAKE failure.  Several contexts launched back to back run concurrently. */
int sofia_md_launch(sofia_ctx *ctx, int32_t nsteps, int32_t use_graph);
int sofia_md_finish(sofia_ctx *ctx);
/* Grow-and-redo (the engine's form of neighbor_wrapper's grow-and-rebuild,
 * Source/mod_neighbor.f90:258-298): md_launch keeps the segment's starting state on the
 * device; when md_finish finds that a list rebuild inside the segment overflowed the
 * stride, it restores that state, rebuilds every list with a stride 1.25x the overflowing
 * count and runs the segment again, with the same result as a run that never overflowed.
 * sofia_md_list_stats reports the current stride and how many segments were redone. */
int sofia_md_list_stats(sofia_ctx *ctx, int32_t *neigh_alloc, int32_t *redos);
/* The engine's list stride is headroom x the largest neighbour count when lists are set up
 * (default 1.5; 1.0 leaves no room, which makes overflow and redo likely). */
int sofia_set_list_headroom(sofia_ctx *ctx, double headroom);
/* MD steps captured per hipGraph (1..200, default 1): md_launch replays a k-step graph and
 * finishes with one-step graphs; k changes launch overhead only, never the trajectory. */
int sofia_set_graph_steps(sofia_ctx *ctx, int32_t steps_per_graph);
/* Continue from FP64 state uploaded with sofia_put_state_f64 (and the replica step
 * counters with sofia_set_replica_steps), e.g. a restart: rebuilds the lists and
 * re-evaluates the forces at the state, with the Langevin draw of the step that produced
 * it.  Friction and SHAKE then see the full-step velocity instead of the half-step one;
 * sofia_md_resume_forces continues bit-exactly. */
int sofia_md_resume(sofia_ctx *ctx);
/* Total force on every atom after the last step (3N doubles, the order of
 * sofia_get_state_f64): the force the next step's first half-kick uses. */
int sofia_get_forces_f64(sofia_ctx *ctx, double *f);
/* The engine's pair kernel alone (lj_cut_coul_dsf_nonewton over the engine's own lists,
 * Source/mod_force_nonbond.f90:105-196, in the context's precision) evaluated at the
 * current state, for checking the engine against the FP64 reference routine: f receives
 * 3N doubles (x, y, z per atom, device order as sofia_get_state_f64), e (optional) per
 * replica 5 doubles: LJ energy, Coulomb energy, virial x, y, z.  The forces the last step
 * carries are left as they were. */
int sofia_md_pair_forces(sofia_ctx *ctx, double *f, double *e);
/* sofia_get_nlist for replica rep alone: its natoms rows (1-based ids within the
 * replica, stride *neigh_alloc) and counts. */
int sofia_get_nlist_replica(sofia_ctx *ctx, int32_t rep, int32_t *nlist, int32_t *numneigh,
                            int32_t *neigh_alloc);
/* The engine's current knobs: pair-kernel lanes per atom, list-build waves, bonded mode,
 * bytes of the mixed pair kernel's position record (12 when charges go by atom type, 16
 * otherwise, 0 outside mixed DSF) and the Verlet skin in A.  Any pointer may be NULL. */
int sofia_get_engine_config(sofia_ctx *ctx, int32_t *lanes, int32_t *list_waves,
                            int32_t *bonded_mode, int32_t *record_bytes, double *skin);
/* sofia_md_resume, then the forces of every replica r with have[r] != 0 replaced by f
 * (3N doubles in sofia_put_state_f64's order, as returned by sofia_get_forces_f64 and
 * re-ordered with the state).  The state a step carries over is (x, v, F): F was evaluated
 * with the half-step velocity (Langevin friction, SHAKE), which the state no longer holds,
 * so carrying F keeps a moved replica on exactly the trajectory it had. */
int sofia_md_resume_forces(sofia_ctx *ctx, const double *f, const int32_t *have);
/* Capped steepest descent on the engine state (no SHAKE, no thermostat): each atom
 * moves along its force by min(gamma*|F|, max_disp) A per step; ends with a full force
 * evaluation so sofia_md_run can continue.  For relaxing generated starting structures. */
int sofia_minimize(sofia_ctx *ctx, int32_t nsteps, double gamma, double max_disp);
/* Energies per replica after the last step, SOFIA_NENERGY doubles each: potential (LJ),
 * e_coul, pot_14, e_coul_14, e_bond, e_angle, e_dihedral, ke, e_improper. */
#define SOFIA_NENERGY 9
int sofia_md_energies(sofia_ctx *ctx, double *out);
/* Number of list builds per replica since md_init (nrep ints). */
int sofia_md_rebuilds(sofia_ctx *ctx, int32_t *out);
/* Per-replica step counters (nrep uint64), the Langevin RNG key; a replica moved into a
 * new batch keeps its random stream by carrying its counter over. */
int sofia_set_replica_steps(sofia_ctx *ctx, const uint64_t *steps);
int sofia_get_replica_steps(sofia_ctx *ctx, uint64_t *steps);
/* Times `launches` back-to-back launches of the engine's pair kernel with HIP events on
 * the context stream (average ms per launch) and counts the list entries inside the
 * force cutoff (the algorithmic pairs of one launch). */
int sofia_profile_nonbond(sofia_ctx *ctx, int32_t launches, double *ms_per_launch,
                          int64_t *pairs_in_cutoff);
/* Times `launches` plain device-to-device copies of `bytes` (a multiple of 16) with the
 * library's streaming copy kernel k_copy16: the HBM reference for achieved bandwidth and
 * for scaling traffic counters (read `bytes`, write `bytes` per launch). */
int sofia_profile_copy(sofia_ctx *ctx, int64_t bytes, int32_t launches, double *ms_per_launch);
/* Runs nsteps MD steps eagerly with HIP events between the phases of a step and returns
 * the average ms per step of each phase in phase_ms[SOFIA_NPHASE]: half-kick, drift,
 * displacement check, cell sort + gather + topology remap, list build, rebuild
 * bookkeeping, force zeroing, pair kernel, bonded (1-4, bond, angle, torsion, improper), Langevin,
 * SHAKE, second half-kick.  full_rebuild != 0 makes every replica rebuild on every step
 * (the cost of one full rebuild).  Energies are not reduced on these steps. */
enum { SOFIA_NPHASE = 12 };
int sofia_profile_step(sofia_ctx *ctx, int32_t nsteps, int32_t full_rebuild, double *phase_ms);
/* Tuning knobs (autotuner): lanes per atom in the pair kernel (4..64, power of two; 0
 * selects the mixed-precision kernel that compacts the in-cutoff entries before the force
 * math, one atom per wavefront -- FP64 and Ewald modes then use 16), threads per block of
 * the mixed pair kernel (64..1024 in steps of 64; the FP64 and Ewald kernels take at most
 * 256 of them), waves per block in the list build. */
int sofia_set_tuning(sofia_ctx *ctx, int32_t lanes_per_atom, int32_t block, int32_t list_waves);
/* Engine Verlet skin in A (list radius rc + skin; default: the reference's pad).  The
 * engine's lists hold every pair inside the cutoff whatever the skin, so the skin changes
 * only speed (longer lists against more frequent rebuilds), never the trajectory.  Takes
 * effect at once on a running engine (lists are rebuilt). */
int sofia_set_skin(sofia_ctx *ctx, double skin);
/* How the engine evaluates its bonded items (1-4 pairs, bonds, angles, torsions, impropers;
 * no counterpart in the reference, which has one serial loop per kind): 0 = one kernel per
 * kind, one thread per item, forces added with atomics (the default); 1 = whole molecules
 * in bundles, one workgroup per bundle summing its forces in LDS.  Both give the same
 * integer sums, so the mode changes speed only (autotuned). */
int sofia_set_bonded_mode(sofia_ctx *ctx, int32_t mode);
/* Union-find clusters of solute molecules joined by an O-N contact < rcontact (A):
 * atoms whose type is in o_types/n_types, molecule id per atom in mol (0-based within a
 * replica, -1 = solvent).  largest[r] = largest cluster (molecules) per replica. */
int sofia_set_molecules(sofia_ctx *ctx, const int32_t *mol);
int sofia_cluster_largest(sofia_ctx *ctx, float rcontact, int32_t n_o, const int32_t *o_types,
                          int32_t n_n, const int32_t *n_types, int32_t *largest);
/* get_greatestCluster (Source/mod_cluster.f90:15-83) on the engine state, per replica, over
 * the solute molecules of sofia_set_molecules: molecule centres (the mean of the molecule's
 * atoms made whole, wrapped into the box; xyzcenter_molecule / pbc_molcom), denstrack = other
 * centres closer than denscut (mol_build_neighbor_nosort, Source/mod_neighbor_cluster.f90:
 * 56-178), dense molecules (denstrack >= densmin) joined into clusters through centre pairs
 * closer than denscut (buildClusters / harvestCluster), the first largest cluster and its
 * non-dense neighbours (append_surface_molecules).  total[r] = core[r] + surface[r], the
 * reference's return value; core and surface may be NULL. */
int sofia_cluster_greatest(sofia_ctx *ctx, double denscut, int32_t densmin, int32_t *total,
                           int32_t *core, int32_t *surface);
/* No counterpart in the reference.  sofia_cluster_greatest for one system held in device
 * memory by the caller, allocating nothing, for callers that cannot allocate (code under
 * hipGraph capture, tight serving loops).  Device inputs: x (FP64 x, y, z per atom, 3n),
 * gid (global ids, 1-based), mol (solute molecule per atom, 0..nmol-1, or -1); box in A.
 * Device output out[4]: total, core, surface, and flags (1 = the union-find gave up, 2 = a
 * molecule label of nmol or more, whose atom was left out); the
 * results are ready once `stream` (a hipStream_t, NULL = the default stream) has run the
 * call's kernels.  Working memory is a fixed device array of the library (about 2.9 MB per
 * GPU, kept until the process ends) that every call on a GPU shares: at most 32768
 * molecules per call (more is SOFIA_E_ARG), and calls that overlap on one GPU race. */
int sofia_cluster_greatest_noalloc(int32_t n, const double *x, const int32_t *gid,
                                   const int32_t *mol, int32_t nmol, float box, double denscut,
                                   int32_t densmin, int32_t *out, void *stream);

/* ---------------------------------------------------------------- files (host only) */
/* Restart file as dump_global_velocity_id_mpi writes it (Source/mod_datadump.f90:113-151):
 * box, temp, num, timeclock, then one "id x y z vx vy vz" record per atom ((i5,6f12.6);
 * ids of 100000 and more widen the id field instead of printing *****).  pos, v: AoS 3*np
 * in the caller's order, global_id the 1-based ids. */
int sofia_write_restart(const char *path, float box, float temp, int32_t num, float timeclock,
                        int32_t np, const int32_t *global_id, const float *pos, const float *v);
/* read_dump_global_id_v (Source/mod_datadump.f90:212-252): records go to the slot of their
 * global id (pos, v AoS 3*np in global-id order); an id out of 1..np, a repeated id or a
 * short file is an error. */
int sofia_read_restart(const char *path, int32_t np, float *box, float *temp, int32_t *num,
                       float *timeclock, float *pos, float *v);
/* read_in_Grofile / read_in_Grofile_box (Source/mod_GroFile.f90:89-159), parsed by column:
 * atom count into *np, coordinates (A, AoS 3*np, file order) into pos when pos != NULL (an
 * error if the count exceeds np_max), cubic box (A) into *box.  pos == NULL only sizes. */
int sofia_read_gro(const char *path, int32_t np_max, int32_t *np, float *pos, float *box);
/* GRO writer ((i5,2a5,i5,3f8.3) atom lines as in template_generator,
 * Source/mod_GroFile.f90:7-68; numbers wrap at 100000): resid per atom, resname and
 * atomname 5 chars per atom, pos in A, cubic box in A. */
int sofia_write_gro(const char *path, const char *title, int32_t np, const int32_t *resid,
                    const char *resname, const char *atomname, const float *pos, float box);

#ifdef __cplusplus
}
#endif
#endif /* SOFIA_HIP_H */
