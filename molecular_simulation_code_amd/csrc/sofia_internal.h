// Internal declarations shared by the HIP translation units of libsofia_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sofia_hip.h"

namespace sofia {

constexpr int SPEC = 12;                     // specbond slots per atom
constexpr double FIX_SCALE = 4294967296.0;   // forces: 2^32 units per kcal/mol/A
constexpr double FIX_INV = 1.0 / FIX_SCALE;
constexpr double EFIX_SCALE = 16777216.0;    // energies and virial: 2^24 units per kcal/mol
constexpr double EFIX_INV = 1.0 / EFIX_SCALE;
constexpr float PAIR_QUANT = 16777216.0f;    // 2^24: per-pair quantum of the mixed kernel
constexpr int PAIR_SHIFT = 8;                // 2^32 / 2^24

// energy / virial accumulator slots (int64 fixed point, per replica)
enum {
    E_LJ = 0, E_COUL, E_POT14, E_COUL14, E_BOND, E_ANGLE, E_DIHED, E_KE,
    V_X, V_Y, V_Z, E_CORR, E_IMPRO, NE
};
// Each (replica, slot) accumulator exists in ESPREAD copies, one 128-byte line apart, and a
// workgroup adds to copy blockIdx % ESPREAD: tens of thousands of wavefronts adding to one
// address serialise at the L2, spreading them over lines does not.  The host sums the copies
// (integers: the total does not depend on which copy a wave used).
constexpr int ESPREAD = 16;
constexpr int ECOPY = 16;                       // int64 words between copies (128 B)
constexpr int EACC_PER_SLOT = ESPREAD * ECOPY;
inline size_t eacc_len(int nrep) { return (size_t)nrep * NE * EACC_PER_SLOT; }

// Per-replica geometry and thermostat (device table).
struct RepInfo {
    float box, ibox, rn;
    int32_t ncellD, ncellT;
    int32_t atom0, natom;
    int32_t cell0;
    int32_t stencil5; // engine grid of cells >= rv/2 with a +-2 stencil (k_build_engine)
    double fric;      // friction * amu2e (kcal/mol per amu*A/fs)
    double gconst;    // noise amplitude * amu2e
    uint64_t seed;
    double chk2;      // engine rebuild threshold on |displacement|^2 (complete-list criterion)
};

// Scalars passed by value to every kernel.
struct DevParams {
    float coulpre, dt, kcal2amu;
    int32_t ntypes;
    double rcut2, cut_coulsq, rv2, rvrc2;
    double alpha, gewald, e_shift, f_shift, ewald_p, ewald_f;
    double a1, a2, a3, a4, a5, my_pis, my_pis_inv, amu2e;
    double fudge12, fudge13, fudge14;
    double shake_tol;
    int32_t shake_maxiter;
    int32_t nrep;
    int32_t natom;
    int32_t neigh_alloc;
    float tab_r0, tab_inv_h;     // DSF Coulomb table of the mixed kernel: r0 + k / inv_h
    int32_t tab_n;               // bins per table; each table has one more, all-zero bin
};

constexpr int DSF_TAB_BINS = 1024;

// Engine bonded work in bundles of whole molecules: item kinds and the widths of their lists.
enum BondedKind { BK_PAIR14 = 0, BK_BOND = 1, BK_ANGLE = 2, BK_TORSION = 3, BK_IMPRO = 4, BK_COUNT = 5 };
constexpr int BUNDLE_SLOTS = 192;   // atoms per bundle (its forces sit in LDS)
constexpr int BUNDLE_TERMS = 256;   // items per bundle when molecules are packed together
constexpr int BUNDLE_THREADS = 256; // one workgroup per bundle
__host__ __device__ inline int bonded_width(int kind) {
    return kind == BK_BOND ? 3 : (kind == BK_ANGLE ? 4 : 5);
}

// Per-atom arrays that the cell sort permutes (one set; the context keeps two).
struct AtomArrays {
    float4 *pos4 = nullptr;     // x, y, z, type (int bits)
    float *q = nullptr;
    int32_t *gid = nullptr;
    int32_t *num1 = nullptr, *num2 = nullptr, *num3 = nullptr;
    int32_t *spec = nullptr;    // [N][12] global 0-based, -1 = empty
    float *v32 = nullptr;       // AoS coords (3N)
    double *xd = nullptr;       // SoA x | y | z (3N)
    double *vd = nullptr;       // SoA (3N)
    int32_t *mol = nullptr;     // molecule id (cluster analysis), -1 = solvent
    int32_t *wclus = nullptr;   // engine: SETTLE water whose oxygen this atom is, -1 = none
};

}  // namespace sofia

struct sofia_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sofia_params p{};
    sofia::DevParams dp{};
    int32_t nrep = 0, N = 0;
    std::vector<sofia::RepInfo> reps_h;
    sofia::RepInfo *reps_d = nullptr;
    int32_t ncell_total = 0, ncell_cap = 0;

    sofia::AtomArrays a, b;     // a = current, b = gather target
    int32_t *rep = nullptr;     // replica of each atom (invariant under the sort)

    // reference-format and engine state that is not permuted as a set
    float *dr32 = nullptr;      // drx | dry | drz (3N), real*4
    double *drd = nullptr;      // FP64 displacement since the last build (engine)
    double *fnb = nullptr;      // nonbond forces, SoA (3N)
    long long *fix = nullptr;   // fixed-point accumulator of every other force, SoA (3N)
    long long *fscratch = nullptr;   // fixed-point scratch for drop-in outputs (3N)
    float *stage32 = nullptr;        // AoS real*4 staging for drop-in inputs/outputs (3N)
    long long *fix14 = nullptr;      // force wrapper: the 1-4 term's fixed-point sums (3N, lazy)

    // cluster analysis: global solute-molecule index per atom lives in AtomArrays::mol
    int32_t nmol = 0;
    int32_t *mol_rep = nullptr, *uf_parent = nullptr, *uf_size = nullptr, *uf_largest = nullptr;
    // get_greatestCluster: first molecule of each replica (nrep + 1), per-molecule centres,
    // their construction, the density counts, and per-replica results
    int32_t *gc_first = nullptr, *gc_refgid = nullptr, *gc_cnt = nullptr, *gc_dens = nullptr;
    long long *gc_sum = nullptr;
    double *gc_refx = nullptr, *gc_com = nullptr;
    unsigned long long *gc_key = nullptr;
    int32_t *gc_surface = nullptr;

    // cells
    int32_t *ccount = nullptr, *cstart = nullptr, *cnum = nullptr, *cell_rep = nullptr;
    int32_t *acell = nullptr, *aslot = nullptr, *perm = nullptr, *tt = nullptr;
    int32_t *atombin = nullptr, *sorted_ids = nullptr;

    // Verlet list
    int32_t *nlist = nullptr, *numneigh = nullptr;
    int32_t neigh_alloc = 0;
    size_t nlist_cap = 0;
    int32_t *maxcnt = nullptr;  // device scalar
    bool have_list = false;

    // flags
    int32_t *rebuild = nullptr;     // per replica
    int32_t *nbuilds = nullptr;     // per replica
    int32_t *shake_err = nullptr;   // device scalar
    int32_t *type_err = nullptr;    // device scalar: put_state's type check
    int32_t *shake_iters = nullptr; // device scalar

    // topology (global 0-based indices)
    int32_t *bondlist = nullptr, nb = 0;
    int32_t *anglelist = nullptr, na = 0;
    int32_t *dihedlist = nullptr, nd = 0;
    int32_t *torsions = nullptr, ntor = 0;
    int32_t *improlist = nullptr, nimp = 0;   // (i1, i2, i3 central, i4, type) per improper
    int32_t *shake_num = nullptr, *shake_atoms = nullptr, ns = 0;
    int32_t *topo_stage = nullptr;   // topology in the host's id layout, on its way out
    uint32_t *rdf_hist = nullptr;    // sofia_rdf_calc's counts (rdf_cap bins, lazy)
    int32_t rdf_cap = 0;
    size_t topo_stage_len = 0;
    // engine constraints: indices of the iteratively solved clusters grouped by kind 1, 3,
    // 4, -3, and the count of symmetric rigid waters (SETTLE, found through AtomArrays::wclus);
    // built when the engine sets up its lists
    int32_t *shake_kidx = nullptr;
    int32_t shake_kcount[5] = {0, 0, 0, 0, 0};
    double *shake_bond = nullptr;
    int4 *water = nullptr;       // per SHAKE cluster: rigid-water bond/angle types + fused flag
    bool water_split = false;    // water bonds/angles moved into the SHAKE kernel
    int32_t *ebond = nullptr, nbe = 0;    // engine bond / angle lists without the fused waters
    int32_t *eangle = nullptr, nae = 0;
    // engine bonded bundles (k_bonded_bundles): per bundle the ranges of its term and atom-slot
    // records; nbundle = 0 runs the per-kind kernels instead
    int32_t nbundle = 0;
    bool bonded_bundles = false;   // run the bundles (tuning choice) rather than per-kind kernels
    int32_t *bund_tstart = nullptr, *bund_sstart = nullptr, *bund_slots = nullptr;
    int2 *bund_terms = nullptr;

    // force field
    float4 *ljtab = nullptr;    // (lj1, lj2, lj3, lj4) per type pair
    double4 *lj14 = nullptr;    // (lj14_1..4) per type pair
    double *mass = nullptr;     // per type
    double *bond_k = nullptr, *bond_r0 = nullptr, *angle_k = nullptr, *angle_th0 = nullptr;
    double *tors_k = nullptr, *tors_n = nullptr, *tors_shift = nullptr;
    double *tors_cs = nullptr;  // per torsion type: cos(shift), sin(shift)
    double *impro_k = nullptr, *impro_chi0 = nullptr;   // per improper type
    int32_t *impro_dn = nullptr;   // cvff style: (d, n) per improper type
    int32_t impro_style = 0;       // 0 harmonic (lammps14), 1 cvff (amber wrapper)
    float4 *dsf_tab = nullptr;  // cubic pieces of the DSF Coulomb force / energy (mixed kernel)
    int4 *pfix = nullptr;       // mixed kernel positions: x 2^32 / box mod 2^32, type (per launch)
    // charges by atom type (every atom of a type carries the same charge): the mixed kernel
    // then reads 12-byte records and takes C q_i q_j from ptab (lj1, lj2, C q_i q_j, 0)
    bool q_by_type = false;
    float4 *ptab = nullptr;
    std::vector<float4> lj_host;   // lj1..lj4 per type pair, as uploaded
    int32_t nbt = 0, nat = 0, ntt = 0;
    bool have_ff = false;

    // energies (int64 fixed point, per replica x NE)
    long long *eacc = nullptr;
    std::vector<double> last_e;

    // engine
    bool engine_ready = false;
    bool engine_grid = false;    // cells re-sized for complete lists (first md_init/md_resume)
    int32_t n_stencil5 = 0;      // replicas on the half-width engine grid
    double skin = -1.0;          // engine skin (A); < 0: the reference's pad
    uint64_t *step_d = nullptr;  // per-replica step counters (Langevin key)
    uint64_t step_h = 0;
    // block: threads per workgroup of the mixed pair kernel.  Each workgroup stages the
    // DSF table (16 KB) and the type-pair table in LDS, so fewer, larger workgroups spend
    // less of the load path on it (1024 against 256: -4.8 % per launch on the batched mix)
    int lanes = 16, block = 1024, list_waves = 4;
    hipGraphExec_t graph = nullptr;    // one MD step
    hipGraphExec_t graphk = nullptr;   // graph_steps MD steps
    int32_t graph_steps = 1;

    // Grow-and-redo of an MD segment whose Verlet list overflowed (the engine's form of the
    // reference's grow-and-rebuild, Source/mod_neighbor.f90:258-298): the state at the start
    // of each segment is kept on the device, and a segment that overflowed is run again
    // from it with a larger list stride.
    bool redo = true;
    sofia::AtomArrays snap{};
    int32_t *snap_topo = nullptr;
    size_t snap_topo_cap = 0;
    double *snap_fnb = nullptr, *snap_drd = nullptr;
    long long *snap_fix = nullptr;
    uint64_t *snap_step = nullptr;
    int32_t *snap_nbuilds = nullptr;
    bool snap_valid = false;
    int32_t snap_nsteps = 0, snap_graph = 0;
    int32_t redos = 0;
    int32_t stride_floor = 0;          // smallest stride the next engine list set-up uses
    double headroom = 1.5;             // engine list stride: headroom x the largest count
    bool has_thermostat = false;

    // host staging (pinned) for the asynchronous pair-kernel drop-in
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    int async_mode = -1;         // pair-kernel mode in flight, -1 = none
};

namespace sofia {

// ------------------------------------------------------------------ kernel launchers
// neighbor build (k_neighbor.hip)
void launch_cell_reset(sofia_ctx *c, bool all);
void launch_bin(sofia_ctx *c);
void launch_cell_scan(sofia_ctx *c);
void launch_scatter_sort(sofia_ctx *c);
void launch_gather(sofia_ctx *c, bool engine);
void launch_remap_topology(sofia_ctx *c);
void launch_host_ids(sofia_ctx *c, const int32_t *src, int32_t *dst, size_t items, int width,
                     int nidx);
void launch_rdf(sofia_ctx *c, int32_t itype, int32_t jtype, double delta, int32_t nbins,
                uint32_t *hist);
void launch_build_list(sofia_ctx *c);
void launch_build_list_engine(sofia_ctx *c);
void launch_check_neighbor(sofia_ctx *c, bool engine);
void launch_set_rebuild(sofia_ctx *c, int32_t value);
void launch_finish_rebuild(sofia_ctx *c, bool engine);

// nonbond (k_nonbond.hip): mode 0 = DSF FP64, 1 = DSF mixed, 2 = Ewald real space (FP32)
void launch_nonbond(sofia_ctx *c, int mode, bool energies);
void launch_count_pairs(sofia_ctx *c, unsigned long long *out);

// bonded (k_bonded.hip)
void launch_correct(sofia_ctx *c, int which, long long *fout);
void launch_pair14(sofia_ctx *c, long long *fout);
void launch_bonded_engine(sofia_ctx *c);
void launch_bonded_dropin(sofia_ctx *c, long long *fout);
int build_bonded_bundles(sofia_ctx *c);   // host: molecules of the engine's bonded items
void drop_bonded_bundles(sofia_ctx *c);
void launch_dsf_mixed(sofia_ctx *c, bool energies, float *ff32, bool from_xd);

// integration (k_integrate.hip)
void launch_zero_forces(sofia_ctx *c, long long *f, bool energies);
void launch_fix_to_coords(sofia_ctx *c, const long long *f, float *out_aos);
void launch_f64_to_coords(sofia_ctx *c, const double *f, float *out_aos);
void launch_nve_v32(sofia_ctx *c, float dt_partial, const float *ff_aos);
void launch_nve_x32(sofia_ctx *c);
void launch_pbc32(sofia_ctx *c);
void launch_berendsen32(sofia_ctx *c, const double *mu_dev);
void launch_shake32(sofia_ctx *c, int iflag, float *ff_aos, double *vir_dev);
void launch_langevin32(sofia_ctx *c, float fric, float gconst, uint64_t seed, uint64_t step,
                       float *out_aos);
void launch_engine_from32(sofia_ctx *c);
void launch_engine_to32(sofia_ctx *c);
void launch_nve_v64(sofia_ctx *c, bool ke);
void launch_nve_x64(sofia_ctx *c);
void launch_langevin64(sofia_ctx *c);
void launch_shake64(sofia_ctx *c);
void launch_water_bonded(sofia_ctx *c);
void launch_copy16(sofia_ctx *c, size_t n16, const void *src, void *dst);
void launch_type_range(sofia_ctx *c, const float4 *pos, int32_t *bad);
void launch_merge_forces(sofia_ctx *c, float *ff32, const long long *f14, const long long *fb);
void launch_step_advance(sofia_ctx *c);
void launch_step_add(sofia_ctx *c, long long delta);
void launch_minimize_step(sofia_ctx *c, double gamma, double max_disp);

// cluster analysis (k_cluster.hip)
int cluster_largest(sofia_ctx *c, float rcontact, const std::vector<int> &o_types,
                    const std::vector<int> &n_types, int32_t *largest);
int cluster_greatest(sofia_ctx *c, double denscut, int32_t densmin, int32_t *total, int32_t *core,
                     int32_t *surface);
int cluster_greatest_noalloc(int32_t n, const double *x, const int32_t *gid, const int32_t *mol,
                             int32_t nmol, float box, double denscut, int32_t densmin,
                             int32_t *out, hipStream_t s);

// error helpers (sofia_api.cpp)
int fail(int code, const std::string &msg);
int check_hip(hipError_t e, const char *what);

}  // namespace sofia

#define SOFIA_HIP(call)                                                   \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return sofia::check_hip(e_, #call);         \
    } while (0)
