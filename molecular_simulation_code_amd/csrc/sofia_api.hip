// C ABI of libsofia_hip.so: context management, topology upload, the drop-in entry
// points that mirror the reference subroutines, and the device-resident MD engine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sofia_device.h"

namespace sofia {

void launch_nonbond_out(sofia_ctx *c, int mode, bool energies, float *ff32);

static thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int io_fail(int code, const std::string &msg) { return fail(code, msg); }

int check_hip(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SOFIA_E_OOM : SOFIA_E_HIP;
}

namespace {

template <typename T>
int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();   // reported here; keep it out of later status checks
        return check_hip(e, "hipMalloc");
    }
    return SOFIA_OK;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

#define TRY(x)                          \
    do {                                \
        int rc_ = (x);                  \
        if (rc_ != SOFIA_OK) return rc_; \
    } while (0)

// The captured MD step holds buffer addresses and by-value parameters (list stride, tuning,
// record layout, thermostat constants) as they were at capture; anything that changes one
// of them drops the graph, and the next graph launch captures it again.
void drop_graph(sofia_ctx *c) {
    if (!c->graph && !c->graphk) return;
    // an executable graph is destroyed only once the stream has drained: launches of it
    // may still be queued (md_launch without md_finish)
    (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    if (c->graphk) (void)hipGraphExecDestroy(c->graphk);
    c->graph = nullptr;
    c->graphk = nullptr;
}

int alloc_atoms(sofia::AtomArrays &A, int n) {
    TRY(dalloc(&A.pos4, n));
    TRY(dalloc(&A.q, n));
    TRY(dalloc(&A.gid, n));
    TRY(dalloc(&A.num1, n));
    TRY(dalloc(&A.num2, n));
    TRY(dalloc(&A.num3, n));
    TRY(dalloc(&A.spec, (size_t)n * SPEC));
    TRY(dalloc(&A.v32, (size_t)3 * n));
    TRY(dalloc(&A.xd, (size_t)3 * n));
    TRY(dalloc(&A.vd, (size_t)3 * n));
    TRY(dalloc(&A.mol, n));
    TRY(dalloc(&A.wclus, n));
    SOFIA_HIP(hipMemset(A.mol, 0xff, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(A.wclus, 0xff, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(A.spec, 0xff, sizeof(int32_t) * (size_t)n * SPEC));
    SOFIA_HIP(hipMemset(A.num1, 0, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(A.num2, 0, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(A.num3, 0, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(A.v32, 0, sizeof(float) * 3 * (size_t)n));
    return SOFIA_OK;
}

void free_atoms(sofia::AtomArrays &A) {
    dfree(A.pos4); dfree(A.q); dfree(A.gid); dfree(A.num1); dfree(A.num2); dfree(A.num3);
    dfree(A.spec); dfree(A.v32); dfree(A.xd); dfree(A.vd); dfree(A.mol); dfree(A.wclus);
}

// Every per-atom array of one set into another, ordered on stream s.
int copy_atoms(sofia::AtomArrays &dst, const sofia::AtomArrays &src, size_t n, hipStream_t s) {
    auto cp = [&](void *d, const void *p, size_t bytes) {
        return hipMemcpyAsync(d, p, bytes, hipMemcpyDeviceToDevice, s);
    };
    SOFIA_HIP(cp(dst.pos4, src.pos4, sizeof(float4) * n));
    SOFIA_HIP(cp(dst.q, src.q, sizeof(float) * n));
    SOFIA_HIP(cp(dst.gid, src.gid, sizeof(int32_t) * n));
    SOFIA_HIP(cp(dst.num1, src.num1, sizeof(int32_t) * n));
    SOFIA_HIP(cp(dst.num2, src.num2, sizeof(int32_t) * n));
    SOFIA_HIP(cp(dst.num3, src.num3, sizeof(int32_t) * n));
    SOFIA_HIP(cp(dst.spec, src.spec, sizeof(int32_t) * SPEC * n));
    SOFIA_HIP(cp(dst.v32, src.v32, sizeof(float) * 3 * n));
    SOFIA_HIP(cp(dst.xd, src.xd, sizeof(double) * 3 * n));
    SOFIA_HIP(cp(dst.vd, src.vd, sizeof(double) * 3 * n));
    SOFIA_HIP(cp(dst.mol, src.mol, sizeof(int32_t) * n));
    SOFIA_HIP(cp(dst.wclus, src.wclus, sizeof(int32_t) * n));
    return SOFIA_OK;
}

// check_allocation's grid, Source/mod_neighbor.f90:1446-1451 (real*4), vol = box**3
void cell_grid(float box, float rcut, float cut_coul, float *rn, int *ncellT, int *ncellD) {
    const float rc = (cut_coul > rcut) ? cut_coul : rcut;
    const int nn = (int)(box / rc);
    const float r = box / (float)nn;
    const float vol = (box * box) * box;
    const float r3 = (r * r) * r;
    *rn = r;
    *ncellT = (int)std::lround((double)(vol / r3));
    *ncellD = (int)std::lround(std::pow((double)*ncellT, 1.0 / 3.0));
}

// neighcell_nonewton, Source/mod_neighbor.f90:1309-1378, offset by the replica's first cell
void stencil(int ncellT, int ncellD, int cell0, int32_t *cnum) {
    int count = 0;
    for (int ic = 0; ic < ncellT; ++ic) {
        cnum[count] = cell0 + ic;
        const int icz = ic / (ncellD * ncellD);
        const int icy = (ic - icz * ncellD * ncellD) / ncellD;
        const int icx = ic - icy * ncellD - icz * ncellD * ncellD;
        for (int iz = -1; iz <= 1; ++iz)
            for (int iy = -1; iy <= 1; ++iy)
                for (int ix = -1; ix <= 1; ++ix) {
                    int zc = icz + iz, yc = icy + iy, xc = icx + ix;
                    if (zc < 0) zc += ncellD; else if (zc >= ncellD) zc -= ncellD;
                    if (yc < 0) yc += ncellD; else if (yc >= ncellD) yc -= ncellD;
                    if (xc < 0) xc += ncellD; else if (xc >= ncellD) xc -= ncellD;
                    if (iz != 0 || iy != 0 || ix != 0) {
                        ++count;
                        cnum[count] = cell0 + xc + yc * ncellD + zc * ncellD * ncellD;
                    }
                }
        ++count;
    }
}

int grow_alloc(int mx) {   // Source/mod_neighbor.f90:260-263
    int na = (int)((float)mx * 1.2f);
    while (na % 16) ++na;
    return na;
}

// The engine's list stride for a largest count mx: headroom x mx rounded up to 16 entries
// (exactly mx at headroom 1, which leaves no room at all), never below the floor a redo set.
int engine_stride(const sofia_ctx *c, int mx) {
    int na = mx;
    if (c->headroom > 1.0) na = (((int)std::ceil((double)mx * c->headroom)) + 15) / 16 * 16;
    return std::max({64, na, c->stride_floor});
}

int ensure_list(sofia_ctx *c, int na) {
    const size_t need = (size_t)c->N * (size_t)na;
    if (need > c->nlist_cap || na != c->neigh_alloc) drop_graph(c);
    if (need > c->nlist_cap) {
        dfree(c->nlist);
        TRY(dalloc(&c->nlist, need));
        c->nlist_cap = need;
    }
    c->neigh_alloc = na;
    c->dp.neigh_alloc = na;
    return SOFIA_OK;
}

// Synchronous copies on the null stream; callers synchronise the context stream first
// when kernels on it may still use the destination.
int upload(void *dst, const void *src, size_t bytes) {
    if (!bytes) return SOFIA_OK;
    SOFIA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SOFIA_OK;
}
int download(void *dst, const void *src, size_t bytes) {
    if (!bytes) return SOFIA_OK;
    SOFIA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SOFIA_OK;
}

std::vector<int32_t> atom_offsets(const sofia_ctx *c) {
    std::vector<int32_t> off(c->nrep + 1, 0);
    for (int r = 0; r < c->nrep; ++r) off[r + 1] = off[r] + c->reps_h[r].natom;
    return off;
}

// 1-based per-replica ids in `items` rows of `width` ints (first `nidx` are atom ids, 0
// = unused) -> global 0-based ids (-1 = unused), replica by replica.
int to_global_list(const sofia_ctx *c, const int32_t *src, int items, int width, int nidx,
                   const int32_t *per_rep, std::vector<int32_t> &out) {
    out.assign((size_t)items * width, 0);
    std::vector<int32_t> counts;
    if (per_rep) counts.assign(per_rep, per_rep + c->nrep);
    else if (c->nrep == 1) counts = {items};
    else return fail(SOFIA_E_ARG, "per-replica item counts are required for a batch");
    long long tot = 0;
    for (int v : counts) tot += v;
    if (tot != items) return fail(SOFIA_E_ARG, "per-replica item counts do not add up");
    int m = 0;
    for (int r = 0; r < c->nrep; ++r) {
        const int a0 = c->reps_h[r].atom0, na = c->reps_h[r].natom;
        for (int k = 0; k < counts[r]; ++k, ++m) {
            for (int w = 0; w < width; ++w) {
                const int v = src[(size_t)m * width + w];
                if (w < nidx) {
                    if (v == 0) { out[(size_t)m * width + w] = -1; continue; }
                    if (v < 1 || v > na) return fail(SOFIA_E_ARG, "atom id out of range in a bonded list");
                    out[(size_t)m * width + w] = a0 + v - 1;
                } else {
                    out[(size_t)m * width + w] = v;
                }
            }
        }
    }
    return SOFIA_OK;
}

int sync_check(sofia_ctx *c);

// Rigid waters (SHAKE kind -3 whose atoms carry exactly two bonds to the first atom and one
// angle at it, and no other bonded term) get their bonds and angle evaluated inside the
// engine's SHAKE kernel: the table `water` records the types per cluster, and the engine's
// bond and angle kernels run over copies of the lists without those items.
int unsplit_water(sofia_ctx *c) {
    if (!c->water_split) return SOFIA_OK;
    TRY(sync_check(c));
    dfree(c->ebond); dfree(c->eangle);
    c->nbe = c->nae = 0;
    if (c->water) SOFIA_HIP(hipMemset(c->water, 0, sizeof(int4) * c->ns));
    c->water_split = false;
    return SOFIA_OK;
}

int split_water(sofia_ctx *c) {
    if (c->water_split || !c->ns || !c->nb || !c->na || !c->bond_k || !c->angle_k) return SOFIA_OK;
    TRY(sync_check(c));
    const int n = c->N;
    std::vector<int32_t> b((size_t)3 * c->nb), a((size_t)4 * c->na), num(c->ns), at((size_t)4 * c->ns);
    TRY(download(b.data(), c->bondlist, sizeof(int32_t) * b.size()));
    TRY(download(a.data(), c->anglelist, sizeof(int32_t) * a.size()));
    TRY(download(num.data(), c->shake_num, sizeof(int32_t) * num.size()));
    TRY(download(at.data(), c->shake_atoms, sizeof(int32_t) * at.size()));
    std::vector<int32_t> nbond(n, 0), nang(n, 0), nother(n, 0), ncl(n, 0), bond_of(n, -1);
    for (int m = 0; m < c->nb; ++m)
        for (int k = 0; k < 2; ++k) {
            const int i = b[3 * (size_t)m + k];
            if (i >= 0) { ++nbond[i]; bond_of[i] = m; }
        }
    std::vector<int32_t> ang_at(n, -1);
    for (int m = 0; m < c->na; ++m)
        for (int k = 0; k < 3; ++k) {
            const int i = a[4 * (size_t)m + k];
            if (i >= 0) { ++nang[i]; ang_at[i] = m; }
        }
    auto touch = [&](const int32_t *lst, int items, int width, int nidx) -> int {
        if (!items) return SOFIA_OK;
        std::vector<int32_t> t((size_t)width * items);
        TRY(download(t.data(), lst, sizeof(int32_t) * t.size()));
        for (int m = 0; m < items; ++m)
            for (int k = 0; k < nidx; ++k) {
                const int i = t[(size_t)width * m + k];
                if (i >= 0) ++nother[i];
            }
        return SOFIA_OK;
    };
    TRY(touch(c->dihedlist, c->nd, 5, 4));
    TRY(touch(c->torsions, c->ntor, 5, 4));
    TRY(touch(c->improlist, c->nimp, 5, 4));
    for (int m = 0; m < c->ns; ++m)
        for (int k = 0; k < 4; ++k) {
            const int i = at[4 * (size_t)m + k];
            if (i >= 0) ++ncl[i];
        }
    std::vector<int4> w(c->ns, int4{0, 0, 0, 0});
    int fused = 0;
    for (int m = 0; m < c->ns; ++m) {
        if (num[m] != -3) continue;
        const int o = at[4 * (size_t)m], h1 = at[4 * (size_t)m + 1], h2 = at[4 * (size_t)m + 2];
        if (o < 0 || h1 < 0 || h2 < 0) continue;
        if (nbond[o] != 2 || nbond[h1] != 1 || nbond[h2] != 1) continue;
        if (nang[o] != 1 || nang[h1] != 1 || nang[h2] != 1) continue;
        if (nother[o] || nother[h1] || nother[h2] || ncl[o] != 1 || ncl[h1] != 1 || ncl[h2] != 1) continue;
        const int b1 = bond_of[h1], b2 = bond_of[h2], an = ang_at[o];
        auto pair_is = [&](int bm, int x, int y) {
            const int p = b[3 * (size_t)bm], q = b[3 * (size_t)bm + 1];
            return (p == x && q == y) || (p == y && q == x);
        };
        if (!pair_is(b1, o, h1) || !pair_is(b2, o, h2)) continue;
        const int e1 = a[4 * (size_t)an], ap = a[4 * (size_t)an + 1], e3 = a[4 * (size_t)an + 2];
        if (ap != o || !((e1 == h1 && e3 == h2) || (e1 == h2 && e3 == h1))) continue;
        const int t1 = b[3 * (size_t)b1 + 2], t2 = b[3 * (size_t)b2 + 2], ta = a[4 * (size_t)an + 3];
        if (t1 < 1 || t1 > c->nbt || t2 < 1 || t2 > c->nbt || ta < 1 || ta > c->nat) continue;
        w[m] = int4{t1, t2, ta, 1};
        b[3 * (size_t)b1 + 2] = -t1;     // marks only; the copies below leave these items out
        b[3 * (size_t)b2 + 2] = -t2;
        a[4 * (size_t)an + 3] = -ta;
        ++fused;
    }
    if (!fused) return SOFIA_OK;
    std::vector<int32_t> eb, ea;
    for (int m = 0; m < c->nb; ++m)
        if (b[3 * (size_t)m + 2] > 0) eb.insert(eb.end(), &b[3 * (size_t)m], &b[3 * (size_t)m] + 3);
    for (int m = 0; m < c->na; ++m)
        if (a[4 * (size_t)m + 3] > 0) ea.insert(ea.end(), &a[4 * (size_t)m], &a[4 * (size_t)m] + 4);
    c->nbe = (int32_t)(eb.size() / 3);
    c->nae = (int32_t)(ea.size() / 4);
    TRY(dalloc(&c->ebond, eb.size()));
    TRY(dalloc(&c->eangle, ea.size()));
    TRY(upload(c->ebond, eb.data(), sizeof(int32_t) * eb.size()));
    TRY(upload(c->eangle, ea.data(), sizeof(int32_t) * ea.size()));
    TRY(upload(c->water, w.data(), sizeof(int4) * w.size()));
    c->water_split = true;
    return SOFIA_OK;
}

void fill_devparams(sofia_ctx *c) {
    const sofia_params &p = c->p;
    sofia::DevParams &d = c->dp;
    d.coulpre = p.coulpre; d.dt = p.dt; d.kcal2amu = p.kcal2amu; d.ntypes = p.ntypes;
    d.rcut2 = p.rcut2; d.cut_coulsq = p.cut_coulsq; d.rv2 = p.rv2; d.rvrc2 = p.rvrc2;
    d.alpha = p.alpha; d.gewald = p.gewald; d.e_shift = p.e_shift; d.f_shift = p.f_shift;
    d.ewald_p = p.ewald_p; d.ewald_f = p.ewald_f;
    d.a1 = p.a1; d.a2 = p.a2; d.a3 = p.a3; d.a4 = p.a4; d.a5 = p.a5;
    d.my_pis = p.my_pis; d.my_pis_inv = p.my_pis_inv; d.amu2e = p.amu2e;
    d.fudge12 = p.fudge_12; d.fudge13 = p.fudge_13; d.fudge14 = p.fudge_14_coul;
    d.shake_tol = p.shake_tol; d.shake_maxiter = p.shake_maxiter;
    d.nrep = c->nrep; d.natom = c->N; d.neigh_alloc = c->neigh_alloc;
}

// The mixed kernel's DSF Coulomb terms as cubic pieces in r on DSF_TAB_BINS bins over
// [0.5 A, cut_coul]: g(r) = (erfc(a r)/r + 2a/sqrt(pi) exp(-a^2 r^2) + r f_shift) / r^2 (force
// over r per C qi qj) and u(r) = erfc(a r)/r - e_shift - r f_shift (energy per C qi qj), with
// the reference's A&S 7.1.26 erfc (Source/mod_force_nonbond.f90:160-183).  Hermite pieces from
// FP64 values and central-difference slopes; the interpolation error is below 1e-7 relative
// for r >= 0.8 A, under the FP32 rounding of the direct evaluation.
int build_dsf_table(sofia_ctx *c) {
    const sofia_params &p = c->p;
    const double r0 = 0.5, r1 = std::sqrt(p.cut_coulsq);
    const int nb = DSF_TAB_BINS;
    const double h = (r1 - r0) / nb;
    const double c2 = 2.0 * p.alpha * p.my_pis_inv;
    auto erfc_as = [&](double r) {
        const double t = 1.0 / (1.0 + p.ewald_p * p.alpha * r);
        return t * (p.a1 + t * (p.a2 + t * (p.a3 + t * (p.a4 + t * p.a5)))) * std::exp(-p.alpha * p.alpha * r * r);
    };
    auto g = [&](double r) {
        return (erfc_as(r) / r + c2 * std::exp(-p.alpha * p.alpha * r * r) + r * p.f_shift) / (r * r);
    };
    auto u = [&](double r) { return erfc_as(r) / r - p.e_shift - r * p.f_shift; };
    auto slope = [&](auto f, double r) { const double d = 1e-5 * r; return (f(r + d) - f(r - d)) / (2.0 * d); };
    // force pieces [0, nb], energy pieces [nb + 1, 2 nb + 1]; piece nb of each is all zero
    // and takes r >= cut_coul
    std::vector<float4> tab((size_t)2 * (nb + 1), make_float4(0.f, 0.f, 0.f, 0.f));
    for (int which = 0; which < 2; ++which)
        for (int k = 0; k < nb; ++k) {
            const double ra = r0 + k * h, rb = ra + h;
            double fa, fb, da, db;
            if (which == 0) { fa = g(ra); fb = g(rb); da = slope(g, ra) * h; db = slope(g, rb) * h; }
            else { fa = u(ra); fb = u(rb); da = slope(u, ra) * h; db = slope(u, rb) * h; }
            tab[(size_t)which * (nb + 1) + k] = make_float4((float)fa, (float)da, (float)(3.0 * (fb - fa) - 2.0 * da - db),
                                                      (float)(2.0 * (fa - fb) + da + db));
        }
    TRY(dalloc(&c->dsf_tab, tab.size()));
    TRY(upload(c->dsf_tab, tab.data(), sizeof(float4) * tab.size()));
    c->dp.tab_r0 = (float)r0;
    c->dp.tab_inv_h = (float)(1.0 / h);
    c->dp.tab_n = nb;
    return SOFIA_OK;
}

int create(const sofia_params *p, int nrep, const int32_t *natoms, const float *box, sofia_ctx **out) {
    if (!p || !natoms || !box || !out || nrep < 1) return fail(SOFIA_E_ARG, "null argument or nrep < 1");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SOFIA_E_NODEV, "no HIP device");
    if (p->ntypes < 1 || p->ntypes * p->ntypes > 1024) return fail(SOFIA_E_ARG, "ntypes must be 1..32");
    std::unique_ptr<sofia_ctx> c(new sofia_ctx());
    c->p = *p;
    c->nrep = nrep;
    SOFIA_HIP(hipGetDevice(&c->device));
    SOFIA_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int atom0 = 0, cell0 = 0;
    for (int r = 0; r < nrep; ++r) {
        if (natoms[r] < 1) return fail(SOFIA_E_ARG, "replica with no atoms");
        sofia::RepInfo R{};
        R.box = box[r];
        R.ibox = (float)(1.0 / (double)box[r]);      // ibox = 1.0d0/box, Source/get_started.f90:66
        cell_grid(R.box, p->rcut, p->cut_coul, &R.rn, &R.ncellT, &R.ncellD);
        if (R.ncellD < 3)
            return fail(SOFIA_E_ARG, "box must hold at least 3 cells of max(rcut, cut_coul) per side");
        if (R.ncellD * R.ncellD * R.ncellD != R.ncellT)
            return fail(SOFIA_E_ARG, "cell grid is not cubic (ncellD^3 != ncellT)");
        R.atom0 = atom0; R.natom = natoms[r];
        R.cell0 = cell0;
        atom0 += natoms[r];
        cell0 += R.ncellT;
        c->reps_h.push_back(R);
    }
    c->N = atom0;
    c->ncell_total = cell0;
    c->ncell_cap = cell0;
    c->last_e.assign((size_t)nrep * SOFIA_NENERGY, 0.0);
    const int n = c->N;
    TRY(alloc_atoms(c->a, n));
    TRY(alloc_atoms(c->b, n));
    TRY(dalloc(&c->rep, n));
    TRY(dalloc(&c->dr32, (size_t)3 * n));
    TRY(dalloc(&c->drd, (size_t)3 * n));
    TRY(dalloc(&c->fnb, (size_t)3 * n));
    TRY(dalloc(&c->pfix, (size_t)n));
    TRY(dalloc(&c->fix, (size_t)3 * n));
    TRY(dalloc(&c->fscratch, (size_t)3 * n));
    TRY(dalloc(&c->stage32, (size_t)3 * n));
    TRY(dalloc(&c->ccount, c->ncell_total));
    TRY(dalloc(&c->cstart, c->ncell_total));
    TRY(dalloc(&c->cnum, (size_t)27 * c->ncell_total));
    TRY(dalloc(&c->cell_rep, c->ncell_total));
    TRY(dalloc(&c->acell, n));
    TRY(dalloc(&c->aslot, n));
    TRY(dalloc(&c->perm, n));
    TRY(dalloc(&c->tt, n));
    TRY(dalloc(&c->atombin, n));
    TRY(dalloc(&c->sorted_ids, n));
    TRY(dalloc(&c->numneigh, n));
    TRY(dalloc(&c->maxcnt, 1));
    TRY(dalloc(&c->rebuild, nrep + 1));   // nrep flags, then `any`
    TRY(dalloc(&c->nbuilds, nrep));
    TRY(dalloc(&c->shake_err, 1));
    TRY(dalloc(&c->type_err, 1));
    TRY(dalloc(&c->shake_iters, 1));
    TRY(dalloc(&c->eacc, eacc_len(nrep)));
    TRY(dalloc(&c->step_d, nrep));
    TRY(dalloc(&c->reps_d, nrep));
    SOFIA_HIP(hipMemset(c->dr32, 0, sizeof(float) * 3 * (size_t)n));
    SOFIA_HIP(hipMemset(c->drd, 0, sizeof(double) * 3 * (size_t)n));
    SOFIA_HIP(hipMemset(c->fix, 0, sizeof(long long) * 3 * (size_t)n));
    SOFIA_HIP(hipMemset(c->fnb, 0, sizeof(double) * 3 * (size_t)n));
    SOFIA_HIP(hipMemset(c->rebuild, 0, sizeof(int32_t) * (nrep + 1)));
    SOFIA_HIP(hipMemset(c->nbuilds, 0, sizeof(int32_t) * nrep));
    SOFIA_HIP(hipMemset(c->numneigh, 0, sizeof(int32_t) * n));
    SOFIA_HIP(hipMemset(c->step_d, 0, sizeof(uint64_t) * nrep));
    std::vector<int32_t> rep(n), cnum((size_t)27 * c->ncell_total), crep(c->ncell_total);
    for (int r = 0; r < nrep; ++r) {
        const sofia::RepInfo &R = c->reps_h[r];
        for (int k = 0; k < R.natom; ++k) rep[R.atom0 + k] = r;
        stencil(R.ncellT, R.ncellD, R.cell0, cnum.data() + (size_t)27 * R.cell0);
        for (int k = 0; k < R.ncellT; ++k) crep[R.cell0 + k] = r;
    }
    TRY(upload(c->rep, rep.data(), sizeof(int32_t) * n));
    TRY(upload(c->cnum, cnum.data(), sizeof(int32_t) * cnum.size()));
    TRY(upload(c->cell_rep, crep.data(), sizeof(int32_t) * crep.size()));
    TRY(upload(c->reps_d, c->reps_h.data(), sizeof(sofia::RepInfo) * nrep));
    // global ids default to 1..natom per replica
    std::vector<int32_t> gid(n);
    for (int r = 0; r < nrep; ++r)
        for (int k = 0; k < c->reps_h[r].natom; ++k) gid[c->reps_h[r].atom0 + k] = k + 1;
    TRY(upload(c->a.gid, gid.data(), sizeof(int32_t) * n));
    TRY(ensure_list(c.get(), p->neigh_alloc0 > 0 ? p->neigh_alloc0 : 512));
    fill_devparams(c.get());
    TRY(build_dsf_table(c.get()));
    *out = c.release();
    return SOFIA_OK;
}

float *staging_f32(sofia_ctx *c) { return c->stage32; }

int sync_check(sofia_ctx *c) {
    SOFIA_HIP(hipStreamSynchronize(c->stream));
    SOFIA_HIP(hipGetLastError());
    return SOFIA_OK;
}

// The ESPREAD copies of every (replica, slot) accumulator summed into e[r * NE + slot].
void sum_eacc(const long long *raw, int nrep, std::vector<long long> &e) {
    e.assign((size_t)nrep * NE, 0);
    for (size_t k = 0; k < e.size(); ++k)
        for (int cp = 0; cp < ESPREAD; ++cp) e[k] += raw[(k * ESPREAD + cp) * ECOPY];
}

int read_energies(sofia_ctx *c, std::vector<long long> &e) {
    std::vector<long long> raw(eacc_len(c->nrep));
    SOFIA_HIP(hipMemcpyAsync(raw.data(), c->eacc, sizeof(long long) * raw.size(), hipMemcpyDeviceToHost,
                             c->stream));
    TRY(sync_check(c));
    sum_eacc(raw.data(), c->nrep, e);
    return SOFIA_OK;
}

double E(const std::vector<long long> &e, int r, int slot) { return (double)e[(size_t)r * NE + slot] * EFIX_INV; }

// Full sort + list build of every flagged replica; grow-and-redo on overflow.
int rebuild_now(sofia_ctx *c, bool engine, bool reference_growth) {
    launch_cell_reset(c, true);
    launch_bin(c);
    launch_cell_scan(c);
    launch_scatter_sort(c);
    launch_gather(c, engine);
    launch_remap_topology(c);
    for (int attempt = 0; attempt < 8; ++attempt) {
        SOFIA_HIP(hipMemsetAsync(c->maxcnt, 0, sizeof(int32_t), c->stream));
        launch_build_list(c);
        int32_t mx = 0;
        SOFIA_HIP(hipMemcpyAsync(&mx, c->maxcnt, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        TRY(sync_check(c));
        if (mx <= c->neigh_alloc) {
            launch_finish_rebuild(c, engine);
            c->have_list = true;
            return sync_check(c);
        }
        int na = reference_growth ? grow_alloc(mx) : engine_stride(c, mx);
        TRY(ensure_list(c, na));
    }
    return fail(SOFIA_E_OVERFLOW, "Verlet list kept overflowing");
}

}  // namespace

void drop_bonded_bundles(sofia_ctx *c) {
    dfree(c->bund_tstart); dfree(c->bund_sstart); dfree(c->bund_slots); dfree(c->bund_terms);
    c->nbundle = 0;
    drop_graph(c);
}

// The engine's bonded items (1-4 pairs, bonds, angles, torsions, impropers, as
// launch_bonded_engine walks them) grouped into molecules, the connected components of the
// atoms they touch, and the molecules of each replica packed in their current atom order
// into bundles of at most BUNDLE_SLOTS atoms and (when several share one) BUNDLE_TERMS
// items.  Items and list columns do not change when the atoms are re-sorted, so the bundles
// hold until the topology changes.  Leaves nbundle = 0 (per-kind kernels) when a molecule is
// too large or a list holds an atom index out of range.
int build_bonded_bundles(sofia_ctx *c) {
    drop_bonded_bundles(c);
    TRY(sync_check(c));
    const int n = c->N;
    struct Src { const int32_t *lst; int items; };
    Src src[BK_COUNT] = {
        {c->dihedlist, c->nd},
        {c->water_split ? c->ebond : c->bondlist, c->bond_k ? (c->water_split ? c->nbe : c->nb) : 0},
        {c->water_split ? c->eangle : c->anglelist, c->angle_k ? (c->water_split ? c->nae : c->na) : 0},
        {c->torsions, c->tors_k ? c->ntor : 0},
        {c->improlist, c->nimp},
    };
    std::vector<int32_t> lst[BK_COUNT];
    for (int k = 0; k < BK_COUNT; ++k) {
        lst[k].resize((size_t)src[k].items * bonded_width(k));
        if (src[k].items) TRY(download(lst[k].data(), src[k].lst, sizeof(int32_t) * lst[k].size()));
    }
    // the atom columns an item touches
    auto cols = [](int kind, int *out) -> int {
        if (kind == BK_PAIR14) { out[0] = 0; out[1] = 3; return 2; }
        const int m = kind == BK_BOND ? 2 : (kind == BK_ANGLE ? 3 : 4);
        for (int k = 0; k < m; ++k) out[k] = k;
        return m;
    };
    struct Term { int kind, item; };
    std::vector<Term> terms;
    for (int k = 0; k < BK_COUNT; ++k)
        for (int m = 0; m < src[k].items; ++m) {
            const int32_t *it = lst[k].data() + (size_t)m * bonded_width(k);
            if (k == BK_BOND && it[2] < 1) continue;    // rigid-water items of the SHAKE kernel
            if (k == BK_ANGLE && it[3] < 1) continue;
            if (m >= (1 << 29)) return SOFIA_OK;
            int cc[4];
            const int nc = cols(k, cc);
            for (int q = 0; q < nc; ++q)
                if (it[cc[q]] < 0 || it[cc[q]] >= n) return SOFIA_OK;
            terms.push_back({k, m});
        }
    if (terms.empty()) return SOFIA_OK;
    // molecules: union-find over the touched atoms
    std::vector<int32_t> parent(n);
    for (int i = 0; i < n; ++i) parent[i] = i;
    auto root = [&](int i) {
        while (parent[i] != i) { parent[i] = parent[parent[i]]; i = parent[i]; }
        return i;
    };
    for (const Term &t : terms) {
        const int32_t *it = lst[t.kind].data() + (size_t)t.item * bonded_width(t.kind);
        int cc[4];
        const int nc = cols(t.kind, cc);
        for (int q = 1; q < nc; ++q) {
            const int a = root(it[cc[0]]), b = root(it[cc[q]]);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    }
    // molecules in order of their smallest (current) atom index, items and atoms per molecule
    std::vector<int32_t> mol_of(n, -1), mol_min;
    std::vector<std::vector<int32_t>> mterms, matoms;
    for (size_t k = 0; k < terms.size(); ++k) {
        const Term &t = terms[k];
        const int32_t *it = lst[t.kind].data() + (size_t)t.item * bonded_width(t.kind);
        const int rt = root(it[0]);
        if (mol_of[rt] < 0) {
            mol_of[rt] = (int32_t)mterms.size();
            mterms.emplace_back();
            matoms.emplace_back();
            mol_min.push_back(rt);
        }
        mterms[mol_of[rt]].push_back((int32_t)k);
    }
    std::vector<int32_t> slot_src(n, -1);   // first (kind, offset) that names each atom
    for (size_t k = 0; k < terms.size(); ++k) {
        const Term &t = terms[k];
        const int32_t *it = lst[t.kind].data() + (size_t)t.item * bonded_width(t.kind);
        int cc[4];
        const int nc = cols(t.kind, cc);
        for (int q = 0; q < nc; ++q) {
            const int a = it[cc[q]];
            if (slot_src[a] >= 0) continue;
            const size_t off = (size_t)t.item * bonded_width(t.kind) + cc[q];
            if (off >= (1u << 29)) return SOFIA_OK;
            slot_src[a] = (int32_t)(((uint32_t)t.kind << 29) | (uint32_t)off);
            matoms[mol_of[root(a)]].push_back(a);
        }
    }
    const int nmol = (int)mterms.size();
    std::vector<int32_t> order(nmol);
    for (int m = 0; m < nmol; ++m) {
        order[m] = m;
        if ((int)matoms[m].size() > BUNDLE_SLOTS) return SOFIA_OK;
    }
    auto rep_of = [&](int atom) {
        int lo = 0, hi = c->nrep - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (c->reps_h[mid].atom0 <= atom) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    std::sort(order.begin(), order.end(), [&](int a, int b) { return mol_min[a] < mol_min[b]; });
    // pack; a bundle's items go out sorted by kind (the record's top bits)
    std::vector<int32_t> tstart{0}, sstart{0}, slots, local(n, -1);
    std::vector<int2> recs, open_recs;
    int cur_rep = -1, cur_slots = 0;
    auto close = [&]() {
        std::stable_sort(open_recs.begin(), open_recs.end(), [](const int2 &a, const int2 &b) {
            return ((uint32_t)a.x >> 29) < ((uint32_t)b.x >> 29);
        });
        recs.insert(recs.end(), open_recs.begin(), open_recs.end());
        open_recs.clear();
        tstart.push_back((int32_t)recs.size());
        sstart.push_back((int32_t)slots.size());
        cur_slots = 0;
    };
    for (int m : order) {
        const int r = rep_of(mol_min[m]);
        const int mt = (int)mterms[m].size(), ms = (int)matoms[m].size();
        if (!open_recs.empty() && (r != cur_rep || (int)open_recs.size() + mt > BUNDLE_TERMS ||
                                   cur_slots + ms > BUNDLE_SLOTS))
            close();
        cur_rep = r;
        for (int a : matoms[m]) {
            local[a] = cur_slots++;
            slots.push_back(slot_src[a]);
        }
        for (int k : mterms[m]) {
            const Term &t = terms[k];
            const int32_t *it = lst[t.kind].data() + (size_t)t.item * bonded_width(t.kind);
            int cc[4];
            const int nc = cols(t.kind, cc);
            uint32_t packed = 0;
            for (int q = 0; q < nc; ++q) packed |= (uint32_t)local[it[cc[q]]] << (8 * cc[q]);
            open_recs.push_back(int2{(int)(((uint32_t)t.kind << 29) | (uint32_t)t.item), (int)packed});
        }
    }
    if (!open_recs.empty()) close();
    const int nb = (int)tstart.size() - 1;
    TRY(dalloc(&c->bund_tstart, tstart.size()));
    TRY(dalloc(&c->bund_sstart, sstart.size()));
    TRY(dalloc(&c->bund_slots, slots.size()));
    TRY(dalloc(&c->bund_terms, recs.size()));
    TRY(upload(c->bund_tstart, tstart.data(), sizeof(int32_t) * tstart.size()));
    TRY(upload(c->bund_sstart, sstart.data(), sizeof(int32_t) * sstart.size()));
    TRY(upload(c->bund_slots, slots.data(), sizeof(int32_t) * slots.size()));
    TRY(upload(c->bund_terms, recs.data(), sizeof(int2) * recs.size()));
    c->nbundle = nb;
    return SOFIA_OK;
}

}  // namespace sofia

using namespace sofia;

namespace {
int classify_charges(sofia_ctx *c);   // the engine's charge records, defined with the engine set-up
}  // namespace

extern "C" {

int sofia_last_error(char *buf, size_t len) {
    if (!buf || !len) return SOFIA_E_ARG;
    std::snprintf(buf, len, "%s", g_last_error.c_str());
    return SOFIA_OK;
}

int sofia_device_count(int *count) {
    if (!count) return fail(SOFIA_E_ARG, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SOFIA_OK;
}

int sofia_device_memory(size_t *free_bytes, size_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(SOFIA_E_ARG, "null argument");
    SOFIA_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return SOFIA_OK;
}

int sofia_ctx_create(const sofia_params *p, int32_t np, float box, sofia_ctx **out) {
    return create(p, 1, &np, &box, out);
}

int sofia_ctx_create_batch(const sofia_params *p, int32_t nrep, const int32_t *natoms,
                           const float *box, sofia_ctx **out) {
    return create(p, nrep, natoms, box, out);
}

int sofia_ctx_destroy(sofia_ctx *c) {
    if (!c) return SOFIA_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_atoms(c->a);
    free_atoms(c->b);
    free_atoms(c->snap);
    dfree(c->snap_topo); dfree(c->snap_fnb); dfree(c->snap_drd); dfree(c->snap_fix);
    dfree(c->snap_step); dfree(c->snap_nbuilds);
    dfree(c->rep); dfree(c->dr32); dfree(c->drd); dfree(c->fnb); dfree(c->fix); dfree(c->fscratch); dfree(c->stage32);
    dfree(c->mol_rep); dfree(c->uf_parent); dfree(c->uf_size); dfree(c->uf_largest);
    dfree(c->gc_first); dfree(c->gc_refgid); dfree(c->gc_cnt); dfree(c->gc_dens); dfree(c->gc_sum);
    dfree(c->gc_refx); dfree(c->gc_com); dfree(c->gc_key); dfree(c->gc_surface);
    dfree(c->ccount); dfree(c->cstart); dfree(c->cnum); dfree(c->cell_rep); dfree(c->acell);
    dfree(c->aslot); dfree(c->perm); dfree(c->tt); dfree(c->atombin); dfree(c->sorted_ids);
    dfree(c->nlist); dfree(c->numneigh); dfree(c->maxcnt); dfree(c->rebuild); dfree(c->nbuilds);
    dfree(c->bund_tstart); dfree(c->bund_sstart); dfree(c->bund_slots); dfree(c->bund_terms);
    dfree(c->shake_err); dfree(c->shake_iters); dfree(c->water); dfree(c->dsf_tab); dfree(c->pfix); dfree(c->ebond); dfree(c->eangle); dfree(c->bondlist); dfree(c->anglelist);
    dfree(c->dihedlist); dfree(c->torsions); dfree(c->shake_num); dfree(c->shake_atoms); dfree(c->shake_kidx);
    dfree(c->ptab);
    dfree(c->shake_bond); dfree(c->ljtab); dfree(c->lj14); dfree(c->mass); dfree(c->bond_k);
    dfree(c->bond_r0); dfree(c->angle_k); dfree(c->angle_th0); dfree(c->tors_k); dfree(c->tors_n);
    dfree(c->tors_shift); dfree(c->tors_cs); dfree(c->improlist); dfree(c->impro_k);
    dfree(c->impro_chi0); dfree(c->impro_dn); dfree(c->eacc); dfree(c->step_d); dfree(c->reps_d);
    dfree(c->topo_stage); dfree(c->type_err); dfree(c->fix14); dfree(c->rdf_hist);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SOFIA_OK;
}

int sofia_synchronize(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    return sync_check(c);
}

int sofia_set_forcefield(sofia_ctx *c, const float *lj1, const float *lj2, const float *lj3,
                         const float *lj4, const double *l141, const double *l142,
                         const double *l143, const double *l144, const double *mass) {
    if (!c || !lj1 || !lj2 || !lj3 || !lj4 || !l141 || !l142 || !l143 || !l144 || !mass)
        return fail(SOFIA_E_ARG, "null argument");
    const int nt2 = c->p.ntypes * c->p.ntypes;
    std::vector<float4> lj(nt2);
    std::vector<double4> l14(nt2);
    for (int k = 0; k < nt2; ++k) {
        lj[k] = make_float4(lj1[k], lj2[k], lj3[k], lj4[k]);
        l14[k] = make_double4(l141[k], l142[k], l143[k], l144[k]);
    }
    if (!c->ljtab) {
        TRY(dalloc(&c->ljtab, nt2));
        TRY(dalloc(&c->lj14, nt2));
        TRY(dalloc(&c->mass, c->p.ntypes));
    }
    TRY(sync_check(c));   // kernels in flight may still read the tables
    TRY(upload(c->ljtab, lj.data(), sizeof(float4) * nt2));
    c->lj_host = lj;
    c->q_by_type = false;
    drop_graph(c);          // a captured step may hold the 12-byte record layout
    TRY(upload(c->lj14, l14.data(), sizeof(double4) * nt2));
    TRY(upload(c->mass, mass, sizeof(double) * c->p.ntypes));
    c->have_ff = true;
    // a running engine takes the new coefficients into its type-pair table at once
    if (c->engine_ready) TRY(classify_charges(c));
    return SOFIA_OK;
}

int sofia_set_exclusions(sofia_ctx *c, const int32_t *n1, const int32_t *n2, const int32_t *n3,
                         const int32_t *spec) {
    if (!c || !n1 || !n2 || !n3 || !spec) return fail(SOFIA_E_ARG, "null argument");
    const int n = c->N;
    std::vector<int32_t> g((size_t)n * SPEC, -1);
    for (int r = 0; r < c->nrep; ++r) {
        const auto &R = c->reps_h[r];
        for (int k = 0; k < R.natom; ++k) {
            const int i = R.atom0 + k;
            if (n3[i] < 0 || n3[i] > SPEC || n1[i] > n2[i] || n2[i] > n3[i])
                return fail(SOFIA_E_ARG, "bad num1bond/num2bond/num3bond");
            for (int l = 0; l < n3[i]; ++l) {
                const int v = spec[(size_t)i * SPEC + l];
                if (v < 1 || v > R.natom) return fail(SOFIA_E_ARG, "specbond id out of range");
                g[(size_t)i * SPEC + l] = R.atom0 + v - 1;
            }
        }
    }
    TRY(upload(c->a.num1, n1, sizeof(int32_t) * n));
    TRY(upload(c->a.num2, n2, sizeof(int32_t) * n));
    TRY(upload(c->a.num3, n3, sizeof(int32_t) * n));
    TRY(upload(c->a.spec, g.data(), sizeof(int32_t) * g.size()));
    return SOFIA_OK;
}

int sofia_set_bonded(sofia_ctx *c, const int32_t *bl, int32_t nb, const int32_t *nb_rep,
                     const int32_t *al, int32_t na, const int32_t *na_rep, const int32_t *dl,
                     int32_t nd, const int32_t *nd_rep) {
    if (!c || nb < 0 || na < 0 || nd < 0) return fail(SOFIA_E_ARG, "bad argument");
    TRY(unsplit_water(c));
    drop_bonded_bundles(c);
    std::vector<int32_t> g;
    dfree(c->bondlist); dfree(c->anglelist); dfree(c->dihedlist);
    c->nb = c->na = c->nd = 0;
    if (nb) {
        TRY(to_global_list(c, bl, nb, 3, 2, nb_rep, g));
        TRY(dalloc(&c->bondlist, g.size()));
        TRY(upload(c->bondlist, g.data(), sizeof(int32_t) * g.size()));
        c->nb = nb;
    }
    if (na) {
        TRY(to_global_list(c, al, na, 4, 3, na_rep, g));
        TRY(dalloc(&c->anglelist, g.size()));
        TRY(upload(c->anglelist, g.data(), sizeof(int32_t) * g.size()));
        c->na = na;
    }
    if (nd) {
        TRY(to_global_list(c, dl, nd, 5, 4, nd_rep, g));
        TRY(dalloc(&c->dihedlist, g.size()));
        TRY(upload(c->dihedlist, g.data(), sizeof(int32_t) * g.size()));
        c->nd = nd;
    }
    return SOFIA_OK;
}

int sofia_set_bonded_coeffs(sofia_ctx *c, int32_t nbt, const double *bk, const double *br0,
                            int32_t nat, const double *ak, const double *ath, int32_t ntt,
                            const double *tk, const double *tn, const double *ts,
                            const int32_t *tors, int32_t ntor, const int32_t *ntor_rep) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    TRY(unsplit_water(c));
    drop_bonded_bundles(c);
    dfree(c->bond_k); dfree(c->bond_r0); dfree(c->angle_k); dfree(c->angle_th0);
    dfree(c->tors_k); dfree(c->tors_n); dfree(c->tors_shift); dfree(c->tors_cs); dfree(c->torsions);
    c->ntor = 0;
    if (nbt > 0) {
        TRY(dalloc(&c->bond_k, nbt)); TRY(dalloc(&c->bond_r0, nbt));
        TRY(upload(c->bond_k, bk, sizeof(double) * nbt)); TRY(upload(c->bond_r0, br0, sizeof(double) * nbt));
    }
    if (nat > 0) {
        TRY(dalloc(&c->angle_k, nat)); TRY(dalloc(&c->angle_th0, nat));
        TRY(upload(c->angle_k, ak, sizeof(double) * nat)); TRY(upload(c->angle_th0, ath, sizeof(double) * nat));
    }
    if (ntt > 0 && ntor > 0) {
        TRY(dalloc(&c->tors_k, ntt)); TRY(dalloc(&c->tors_n, ntt)); TRY(dalloc(&c->tors_shift, ntt));
        TRY(upload(c->tors_k, tk, sizeof(double) * ntt)); TRY(upload(c->tors_n, tn, sizeof(double) * ntt));
        TRY(upload(c->tors_shift, ts, sizeof(double) * ntt));
        std::vector<double> cs((size_t)2 * ntt);
        for (int t = 0; t < ntt; ++t) {
            const double d = ts[t] * (3.141592653589793 / 180.0);
            cs[2 * (size_t)t] = std::cos(d);
            cs[2 * (size_t)t + 1] = std::sin(d);
        }
        TRY(dalloc(&c->tors_cs, cs.size()));
        TRY(upload(c->tors_cs, cs.data(), sizeof(double) * cs.size()));
        std::vector<int32_t> g;
        TRY(to_global_list(c, tors, ntor, 5, 4, ntor_rep, g));
        TRY(dalloc(&c->torsions, g.size()));
        TRY(upload(c->torsions, g.data(), sizeof(int32_t) * g.size()));
        c->ntor = ntor;
    }
    c->nbt = nbt; c->nat = nat; c->ntt = ntt;
    return SOFIA_OK;
}

namespace {
int set_impropers(sofia_ctx *c, int style, int32_t nit, const double *k, const double *chi0,
                  const int32_t *sign, const int32_t *mult, const int32_t *improlist, int32_t ni,
                  const int32_t *ni_rep) {
    if (!c || nit < 0 || ni < 0) return fail(SOFIA_E_ARG, "bad argument");
    if (ni > 0 && (nit == 0 || !k || !improlist || (style == 0 && !chi0) ||
                   (style == 1 && (!sign || !mult))))
        return fail(SOFIA_E_ARG, "impropers need their type coefficients");
    if (style == 1)
        for (int t = 0; t < nit; ++t) {
            if (sign[t] != 1 && sign[t] != -1) return fail(SOFIA_E_ARG, "cvff improper sign must be +1 or -1");
            if (mult[t] < 0 || mult[t] > 6) return fail(SOFIA_E_ARG, "cvff improper multiplicity must be 0..6");
        }
    TRY(unsplit_water(c));
    drop_bonded_bundles(c);
    dfree(c->impro_k); dfree(c->impro_chi0); dfree(c->impro_dn); dfree(c->improlist);
    c->nimp = 0;
    c->impro_style = style;
    if (!ni) return SOFIA_OK;
    for (int m = 0; m < ni; ++m) {
        const int t = improlist[5 * (size_t)m + 4];
        if (t < 1 || t > nit) return fail(SOFIA_E_ARG, "improper type out of range");
        for (int w = 0; w < 4; ++w)
            if (improlist[5 * (size_t)m + w] == 0) return fail(SOFIA_E_ARG, "improper with a missing atom");
    }
    TRY(dalloc(&c->impro_k, nit));
    TRY(upload(c->impro_k, k, sizeof(double) * nit));
    if (style == 0) {
        TRY(dalloc(&c->impro_chi0, nit));
        TRY(upload(c->impro_chi0, chi0, sizeof(double) * nit));
    } else {
        std::vector<int32_t> dn((size_t)2 * nit);
        for (int t = 0; t < nit; ++t) { dn[2 * (size_t)t] = sign[t]; dn[2 * (size_t)t + 1] = mult[t]; }
        TRY(dalloc(&c->impro_dn, dn.size()));
        TRY(upload(c->impro_dn, dn.data(), sizeof(int32_t) * dn.size()));
    }
    std::vector<int32_t> g;
    TRY(to_global_list(c, improlist, ni, 5, 4, ni_rep, g));
    TRY(dalloc(&c->improlist, g.size()));
    TRY(upload(c->improlist, g.data(), sizeof(int32_t) * g.size()));
    c->nimp = ni;
    return SOFIA_OK;
}
}  // namespace

int sofia_set_impropers(sofia_ctx *c, int32_t nit, const double *k, const double *chi0,
                        const int32_t *improlist, int32_t ni, const int32_t *ni_rep) {
    return set_impropers(c, 0, nit, k, chi0, nullptr, nullptr, improlist, ni, ni_rep);
}

int sofia_set_impropers_cvff(sofia_ctx *c, int32_t nit, const double *k, const int32_t *sign,
                             const int32_t *mult, const int32_t *improlist, int32_t ni,
                             const int32_t *ni_rep) {
    return set_impropers(c, 1, nit, k, nullptr, sign, mult, improlist, ni, ni_rep);
}

int sofia_set_shake(sofia_ctx *c, int32_t ns, const int32_t *ns_rep, const int32_t *num,
                    const int32_t *atoms, const double *bond) {
    if (!c || ns < 0) return fail(SOFIA_E_ARG, "bad argument");
    TRY(unsplit_water(c));
    drop_bonded_bundles(c);
    dfree(c->shake_num); dfree(c->shake_atoms); dfree(c->shake_bond); dfree(c->water);
    dfree(c->shake_kidx);
    for (int q = 0; q < 5; ++q) c->shake_kcount[q] = 0;
    TRY(sync_check(c));
    SOFIA_HIP(hipMemset(c->a.wclus, 0xff, sizeof(int32_t) * c->N));   // no SETTLE waters yet
    c->ns = 0;
    if (!ns) return SOFIA_OK;
    for (int k = 0; k < ns; ++k)
        if (num[k] != 1 && num[k] != 3 && num[k] != 4 && num[k] != -3)
            return fail(SOFIA_E_ARG, "SHAKE cluster type must be 1, 3, 4 or -3");
    std::vector<int32_t> g;
    TRY(to_global_list(c, atoms, ns, 4, 4, ns_rep, g));
    for (int k = 0; k < ns; ++k) {
        const int need = num[k] == 1 ? 2 : (num[k] == 4 ? 4 : 3);
        for (int a = 0; a < need; ++a)
            if (g[(size_t)k * 4 + a] < 0) return fail(SOFIA_E_ARG, "SHAKE cluster is missing an atom");
    }
    TRY(dalloc(&c->shake_num, ns));
    TRY(dalloc(&c->shake_atoms, g.size()));
    TRY(dalloc(&c->shake_bond, (size_t)3 * ns));
    TRY(upload(c->shake_num, num, sizeof(int32_t) * ns));
    TRY(upload(c->shake_atoms, g.data(), sizeof(int32_t) * g.size()));
    TRY(upload(c->shake_bond, bond, sizeof(double) * 3 * ns));
    TRY(dalloc(&c->water, ns));
    SOFIA_HIP(hipMemset(c->water, 0, sizeof(int4) * ns));
    c->ns = ns;
    return SOFIA_OK;
}

int sofia_put_state(sofia_ctx *c, const sofia_ainfo *pos, const float *q, const sofia_coords *v,
                    const int32_t *gid) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    const int n = c->N;
    TRY(sync_check(c));   // kernels in flight may still read the arrays replaced here
    if (pos && n > 0) {
        // Types within 1..ntypes, checked on the device (a host pass over the array costs
        // more than the copy) on a staging copy, the sort's gather target, which holds nothing
        // between list builds; a rejected upload leaves the state as it was.
        TRY(upload(c->b.pos4, pos, sizeof(float4) * n));
        SOFIA_HIP(hipMemsetAsync(c->type_err, 0, sizeof(int32_t), c->stream));
        launch_type_range(c, c->b.pos4, c->type_err);
        int32_t bad = 0;
        SOFIA_HIP(hipMemcpyAsync(&bad, c->type_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        TRY(sync_check(c));
        if (bad) return fail(SOFIA_E_ARG, "atom type out of range");
        launch_copy16(c, (size_t)n, c->b.pos4, c->a.pos4);
        c->q_by_type = false;   // types may have changed
        drop_graph(c);
    }
    if (q) {
        TRY(upload(c->a.q, q, sizeof(float) * n));
        c->q_by_type = false;
        drop_graph(c);
    }
    if (v) TRY(upload(c->a.v32, v, sizeof(float) * 3 * (size_t)n));
    if (gid) TRY(upload(c->a.gid, gid, sizeof(int32_t) * n));
    // a running engine re-derives its charge records from the new types and charges
    if ((pos || q) && c->engine_ready) TRY(classify_charges(c));
    return SOFIA_OK;
}

int sofia_get_state(sofia_ctx *c, sofia_ainfo *pos, sofia_coords *v, int32_t *gid) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    TRY(sync_check(c));
    const int n = c->N;
    if (pos) TRY(download(pos, c->a.pos4, sizeof(float4) * n));
    if (v) TRY(download(v, c->a.v32, sizeof(float) * 3 * (size_t)n));
    if (gid) TRY(download(gid, c->a.gid, sizeof(int32_t) * n));
    return SOFIA_OK;
}

int sofia_get_state_f64(sofia_ctx *c, double *x, double *v) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    TRY(sync_check(c));
    const int n = c->N;
    std::vector<double> t((size_t)3 * n);
    for (int w = 0; w < 2; ++w) {
        double *dst = w ? v : x;
        if (!dst) continue;
        TRY(download(t.data(), w ? c->a.vd : c->a.xd, sizeof(double) * t.size()));
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) dst[3 * (size_t)i + d] = t[(size_t)d * n + i];
    }
    return SOFIA_OK;
}

int sofia_put_state_f64(sofia_ctx *c, const double *x, const double *v) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    const int n = c->N;
    std::vector<double> t((size_t)3 * n);
    for (int w = 0; w < 2; ++w) {
        const double *src = w ? v : x;
        if (!src) continue;
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) t[(size_t)d * n + i] = src[3 * (size_t)i + d];
        TRY(upload(w ? c->a.vd : c->a.xd, t.data(), sizeof(double) * t.size()));
    }
    return SOFIA_OK;
}

int sofia_get_displacement(sofia_ctx *c, float *dx, float *dy, float *dz) {
    if (!c || !dx || !dy || !dz) return fail(SOFIA_E_ARG, "null argument");
    TRY(sync_check(c));
    const int n = c->N;
    TRY(download(dx, c->dr32, sizeof(float) * n));
    TRY(download(dy, c->dr32 + n, sizeof(float) * n));
    TRY(download(dz, c->dr32 + 2 * (size_t)n, sizeof(float) * n));
    return SOFIA_OK;
}

int sofia_put_displacement(sofia_ctx *c, const float *dx, const float *dy, const float *dz) {
    if (!c || !dx || !dy || !dz) return fail(SOFIA_E_ARG, "null argument");
    const int n = c->N;
    TRY(upload(c->dr32, dx, sizeof(float) * n));
    TRY(upload(c->dr32 + n, dy, sizeof(float) * n));
    TRY(upload(c->dr32 + 2 * (size_t)n, dz, sizeof(float) * n));
    return SOFIA_OK;
}

int sofia_host_alloc(size_t bytes, void **out) {
    if (!out) return fail(SOFIA_E_ARG, "null argument");
    *out = nullptr;
    if (bytes == 0) return fail(SOFIA_E_ARG, "zero bytes");
    void *p = nullptr;
    SOFIA_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    *out = p;
    return SOFIA_OK;
}

int sofia_host_free(void *p) {
    if (p) SOFIA_HIP(hipHostFree(p));
    return SOFIA_OK;
}

// ------------------------------------------------------------------ drop-ins
int sofia_neighbor_wrapper(sofia_ctx *c, int32_t, int32_t *tt) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    launch_set_rebuild(c, 1);
    TRY(rebuild_now(c, c->engine_ready, true));
    if (tt) {
        std::vector<int32_t> t(c->N);
        TRY(download(t.data(), c->tt, sizeof(int32_t) * c->N));
        for (int r = 0; r < c->nrep; ++r) {
            const auto &R = c->reps_h[r];
            for (int k = 0; k < R.natom; ++k) tt[R.atom0 + k] = t[R.atom0 + k] - R.atom0 + 1;
        }
    }
    return SOFIA_OK;
}

int sofia_get_topology(sofia_ctx *c, float *q, int32_t *num1bond, int32_t *num2bond,
                       int32_t *num3bond, int32_t *specbond, int32_t *bondlist, int32_t *anglelist,
                       int32_t *dihedlist, int32_t *improlist, int32_t *shake_atoms) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    TRY(sync_check(c));
    const size_t n = c->N;
    if (q) TRY(download(q, c->a.q, sizeof(float) * n));
    if (num1bond) TRY(download(num1bond, c->a.num1, sizeof(int32_t) * n));
    if (num2bond) TRY(download(num2bond, c->a.num2, sizeof(int32_t) * n));
    if (num3bond) TRY(download(num3bond, c->a.num3, sizeof(int32_t) * n));
    struct Arr {
        int32_t *dst;
        const int32_t *src;
        size_t items;
        int width, nidx;
    };
    const Arr arrs[] = {{specbond, c->a.spec, n, SPEC, SPEC},
                        {bondlist, c->bondlist, (size_t)c->nb, 3, 2},
                        {anglelist, c->anglelist, (size_t)c->na, 4, 3},
                        {dihedlist, c->dihedlist, (size_t)c->nd, 5, 4},
                        {improlist, c->improlist, (size_t)c->nimp, 5, 4},
                        {shake_atoms, c->shake_atoms, (size_t)c->ns, 4, 4}};
    size_t need = 0;
    for (const Arr &x : arrs)
        if (x.dst) need = std::max(need, x.items * x.width);
    if (need > c->topo_stage_len) {
        dfree(c->topo_stage);
        c->topo_stage_len = 0;
        TRY(dalloc(&c->topo_stage, need));
        c->topo_stage_len = need;
    }
    for (const Arr &x : arrs) {
        if (!x.dst || !x.items) continue;
        launch_host_ids(c, x.src, c->topo_stage, x.items, x.width, x.nidx);
        TRY(sync_check(c));
        TRY(download(x.dst, c->topo_stage, sizeof(int32_t) * x.items * x.width));
    }
    return SOFIA_OK;
}

int sofia_rdf_calc(sofia_ctx *c, int32_t itype, int32_t jtype, float rdf_delta, int32_t nbins,
                   float *rdf, int32_t *numob_rdf) {
    if (!c || !rdf) return fail(SOFIA_E_ARG, "null argument");
    if (nbins < 1 || !(rdf_delta > 0.0f)) return fail(SOFIA_E_ARG, "nbins and rdf_delta must be positive");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list: call sofia_neighbor_wrapper first");
    if (nbins > c->rdf_cap) {
        dfree(c->rdf_hist);
        c->rdf_cap = 0;
        TRY(dalloc(&c->rdf_hist, (size_t)nbins));
        c->rdf_cap = nbins;
    }
    SOFIA_HIP(hipMemsetAsync(c->rdf_hist, 0, sizeof(uint32_t) * nbins, c->stream));
    launch_rdf(c, itype, jtype, (double)rdf_delta, nbins, c->rdf_hist);
    std::vector<uint32_t> h(nbins);
    TRY(sync_check(c));
    TRY(download(h.data(), c->rdf_hist, sizeof(uint32_t) * nbins));
    // the reference adds 1.0 per pair to a real*4 bin; adding the count once is the same
    // while a bin stays below 2^24
    for (int k = 0; k < nbins; ++k) rdf[k] += (float)h[k];
    if (numob_rdf) *numob_rdf += 1;
    return SOFIA_OK;
}

int sofia_get_cells(sofia_ctx *c, int32_t *start, int32_t *endposit, int32_t *atombin,
                    int32_t *ncellT, int32_t *ncellD, float *rn) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no list built yet");
    TRY(sync_check(c));
    std::vector<int32_t> cs(c->ncell_total), cc(c->ncell_total), ab(c->N);
    TRY(download(cs.data(), c->cstart, sizeof(int32_t) * cs.size()));
    TRY(download(cc.data(), c->ccount, sizeof(int32_t) * cc.size()));
    TRY(download(ab.data(), c->atombin, sizeof(int32_t) * ab.size()));
    for (int r = 0; r < c->nrep; ++r) {
        const auto &R = c->reps_h[r];
        for (int k = 0; k < R.ncellT; ++k) {
            const int g = R.cell0 + k;
            if (start) start[g] = cs[g] - R.atom0 + 1;
            if (endposit) endposit[g] = cs[g] - R.atom0 + cc[g];
        }
        if (atombin)
            for (int k = 0; k < R.natom; ++k) atombin[R.atom0 + k] = ab[R.atom0 + k] - R.cell0;
    }
    if (ncellT) *ncellT = c->reps_h[0].ncellT;
    if (ncellD) *ncellD = c->reps_h[0].ncellD;
    if (rn) *rn = c->reps_h[0].rn;
    return SOFIA_OK;
}

int sofia_get_nlist(sofia_ctx *c, int32_t *nlist, int32_t *numneigh, int32_t *na) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no list built yet");
    TRY(sync_check(c));
    if (na) *na = c->neigh_alloc;
    std::vector<int32_t> nn(c->N);
    TRY(download(nn.data(), c->numneigh, sizeof(int32_t) * c->N));
    if (numneigh) std::memcpy(numneigh, nn.data(), sizeof(int32_t) * c->N);
    if (nlist) {
        const size_t total = (size_t)c->N * c->neigh_alloc;
        TRY(download(nlist, c->nlist, sizeof(int32_t) * total));
        for (int r = 0; r < c->nrep; ++r) {
            const auto &R = c->reps_h[r];
            for (int k = 0; k < R.natom; ++k) {
                const int i = R.atom0 + k;
                int32_t *row = nlist + (size_t)i * c->neigh_alloc;
                for (int m = 0; m < nn[i]; ++m) row[m] = row[m] - R.atom0 + 1;
                for (int m = nn[i]; m < c->neigh_alloc; ++m) row[m] = 0;
            }
        }
    }
    return SOFIA_OK;
}

int sofia_check_neighbor(sofia_ctx *c, int32_t *nflag) {
    if (!c || !nflag) return fail(SOFIA_E_ARG, "null argument");
    launch_set_rebuild(c, 0);
    launch_check_neighbor(c, false);
    std::vector<int32_t> f(c->nrep);
    SOFIA_HIP(hipMemcpyAsync(f.data(), c->rebuild, sizeof(int32_t) * c->nrep, hipMemcpyDeviceToHost, c->stream));
    TRY(sync_check(c));
    *nflag = 0;
    for (int v : f) *nflag |= v;
    launch_set_rebuild(c, 0);
    return sync_check(c);
}

static int nonbond_dropin(sofia_ctx *c, int mode, sofia_coords *ff, double *ff64, double *pot,
                          double *ecoul, double *vir) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list: call sofia_neighbor_wrapper first");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    launch_zero_forces(c, c->fscratch, true);
    launch_nonbond_out(c, mode, true, staging_f32(c));
    std::vector<long long> e;
    TRY(read_energies(c, e));
    const int n = c->N;
    if (ff) TRY(download(ff, staging_f32(c), sizeof(float) * 3 * (size_t)n));
    if (ff64) {
        std::vector<double> t((size_t)3 * n);
        TRY(download(t.data(), c->fnb, sizeof(double) * t.size()));
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) ff64[3 * (size_t)i + d] = t[(size_t)d * n + i];
    }
    double sp = 0, se = 0, vx = 0, vy = 0, vz = 0;
    for (int r = 0; r < c->nrep; ++r) {
        sp += E(e, r, E_LJ); se += E(e, r, E_COUL);
        vx += E(e, r, V_X); vy += E(e, r, V_Y); vz += E(e, r, V_Z);
    }
    if (pot) *pot += sp;
    if (ecoul) *ecoul += se;
    if (vir) { vir[0] = vx; vir[1] = vy; vir[2] = vz; }
    return SOFIA_OK;
}

int sofia_lj_cut_coul_dsf_async(sofia_ctx *c, int32_t mixed) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list: call sofia_neighbor_wrapper first");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    if (c->async_mode >= 0) return fail(SOFIA_E_STATE, "a pair-kernel call is already in flight");
    const size_t fbytes = sizeof(float) * 3 * (size_t)c->N, ebytes = sizeof(long long) * eacc_len(c->nrep);
    if (c->pinned_bytes < fbytes + ebytes) {
        if (c->pinned) (void)hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_bytes = 0;
        SOFIA_HIP(hipHostMalloc(&c->pinned, fbytes + ebytes, hipHostMallocDefault));
        c->pinned_bytes = fbytes + ebytes;
    }
    const int mode = mixed ? 1 : 0;
    launch_zero_forces(c, c->fscratch, true);
    launch_nonbond_out(c, mode, true, staging_f32(c));
    SOFIA_HIP(hipMemcpyAsync(c->pinned, staging_f32(c), fbytes, hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipMemcpyAsync((char *)c->pinned + fbytes, c->eacc, ebytes, hipMemcpyDeviceToHost, c->stream));
    c->async_mode = mode;
    return SOFIA_OK;
}

int sofia_nonbond_wait(sofia_ctx *c, sofia_coords *ff, double *pot, double *ec, double *vir) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (c->async_mode < 0) return fail(SOFIA_E_STATE, "no pair-kernel call in flight");
    c->async_mode = -1;
    TRY(sync_check(c));
    const size_t fbytes = sizeof(float) * 3 * (size_t)c->N;
    if (ff) std::memcpy(ff, c->pinned, fbytes);
    std::vector<long long> es;
    sum_eacc((const long long *)((const char *)c->pinned + fbytes), c->nrep, es);
    double sp = 0, se = 0, vx = 0, vy = 0, vz = 0;
    for (int r = 0; r < c->nrep; ++r) {
        const long long *o = es.data() + (size_t)r * NE;
        sp += (double)o[E_LJ] * EFIX_INV; se += (double)o[E_COUL] * EFIX_INV;
        vx += (double)o[V_X] * EFIX_INV; vy += (double)o[V_Y] * EFIX_INV; vz += (double)o[V_Z] * EFIX_INV;
    }
    if (pot) *pot += sp;
    if (ec) *ec += se;
    if (vir) { vir[0] = vx; vir[1] = vy; vir[2] = vz; }
    return SOFIA_OK;
}

int sofia_lj_cut_coul_dsf(sofia_ctx *c, sofia_coords *ff, double *ff64, double *pot, double *ec,
                          double *vir) {
    return nonbond_dropin(c, 0, ff, ff64, pot, ec, vir);
}

int sofia_lj_cut_coul_dsf_mixed(sofia_ctx *c, sofia_coords *ff, double *ff64, double *pot,
                                double *ec, double *vir) {
    return nonbond_dropin(c, 1, ff, ff64, pot, ec, vir);
}

int sofia_lj_cut_coul_long(sofia_ctx *c, sofia_coords *ff, double *pot, double *ec, double *vir) {
    return nonbond_dropin(c, 2, ff, nullptr, pot, ec, vir);
}

int sofia_correct_bonds(sofia_ctx *c, int32_t which, sofia_coords *ff, double *ec, double *vir) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (which != 12 && which != 13 && which != 14) return fail(SOFIA_E_ARG, "which must be 12, 13 or 14");
    launch_zero_forces(c, c->fscratch, true);
    launch_correct(c, which, c->fscratch);
    float *st = staging_f32(c);
    launch_fix_to_coords(c, c->fscratch, st);
    std::vector<long long> e;
    TRY(read_energies(c, e));
    if (ff) TRY(download(ff, st, sizeof(float) * 3 * (size_t)c->N));
    double s = 0, vx = 0, vy = 0, vz = 0;
    for (int r = 0; r < c->nrep; ++r) {
        s += E(e, r, E_CORR); vx += E(e, r, V_X); vy += E(e, r, V_Y); vz += E(e, r, V_Z);
    }
    if (ec) *ec += s;
    if (vir) { vir[0] += vx; vir[1] += vy; vir[2] += vz; }
    return SOFIA_OK;
}

int sofia_pair14_dsf(sofia_ctx *c, sofia_coords *ffh, double *ec14, double *pot14) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    launch_zero_forces(c, c->fscratch, true);
    launch_pair14(c, c->fscratch);
    float *st = staging_f32(c);
    launch_fix_to_coords(c, c->fscratch, st);
    std::vector<long long> e;
    TRY(read_energies(c, e));
    if (ffh) TRY(download(ffh, st, sizeof(float) * 3 * (size_t)c->N));
    double a = 0, b = 0;
    for (int r = 0; r < c->nrep; ++r) { a += E(e, r, E_COUL14); b += E(e, r, E_POT14); }
    if (ec14) *ec14 += a;
    if (pot14) *pot14 += b;
    return SOFIA_OK;
}

int sofia_bonded_forces(sofia_ctx *c, sofia_coords *ff, double *e_bond, double *e_angle,
                        double *e_dihed, double *e_improper) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_ff || !c->bond_k) return fail(SOFIA_E_STATE, "bonded coefficients not set");
    if (c->engine_ready) return fail(SOFIA_E_STATE, "a running engine evaluates its own bonded terms");
    launch_zero_forces(c, c->fscratch, true);
    launch_bonded_dropin(c, c->fscratch);
    float *st = staging_f32(c);
    launch_fix_to_coords(c, c->fscratch, st);
    std::vector<long long> e;
    TRY(read_energies(c, e));
    if (ff) TRY(download(ff, st, sizeof(float) * 3 * (size_t)c->N));
    double sb = 0, sa = 0, sd = 0, si = 0;
    for (int r = 0; r < c->nrep; ++r) {
        sb += E(e, r, E_BOND); sa += E(e, r, E_ANGLE); sd += E(e, r, E_DIHED); si += E(e, r, E_IMPRO);
    }
    if (e_bond) *e_bond += sb;
    if (e_angle) *e_angle += sa;
    if (e_dihed) *e_dihed += sd;
    if (e_improper) *e_improper += si;
    return SOFIA_OK;
}

int sofia_force_wrapper_dsf(sofia_ctx *c, sofia_coords *ff, double *e, double *vir) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list: call sofia_neighbor_wrapper first");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    if ((c->nb || c->na || c->ntor || c->nimp) && !c->bond_k)
        return fail(SOFIA_E_STATE, "bonded coefficients not set");
    if (c->engine_ready) return fail(SOFIA_E_STATE, "a running engine evaluates its own forces");
    const int n = c->N;
    if (!c->fix14) TRY(dalloc(&c->fix14, (size_t)3 * n));
    float *st = staging_f32(c);
    launch_zero_forces(c, c->fscratch, true);   // energies zeroed once: the terms use their own slots
    launch_zero_forces(c, c->fix14, false);
    launch_nonbond_out(c, 0, true, st);
    launch_pair14(c, c->fix14);
    launch_bonded_dropin(c, c->fscratch);
    launch_merge_forces(c, st, c->fix14, c->fscratch);
    std::vector<long long> es;
    TRY(read_energies(c, es));
    if (ff) TRY(download(ff, st, sizeof(float) * 3 * (size_t)n));
    // summed over replicas first, then added, as each routine of its own adds its sums
    const int slot[8] = {E_LJ, E_COUL, E_COUL14, E_POT14, E_BOND, E_ANGLE, E_DIHED, E_IMPRO};
    double sum[8] = {0, 0, 0, 0, 0, 0, 0, 0}, vx = 0, vy = 0, vz = 0;
    for (int r = 0; r < c->nrep; ++r) {
        for (int k = 0; k < 8; ++k) sum[k] += E(es, r, slot[k]);
        vx += E(es, r, V_X); vy += E(es, r, V_Y); vz += E(es, r, V_Z);
    }
    if (e)
        for (int k = 0; k < 8; ++k) e[k] += sum[k];
    if (vir) { vir[0] = vx; vir[1] = vy; vir[2] = vz; }
    return SOFIA_OK;
}

int sofia_nve_v(sofia_ctx *c, float dt_partial, const sofia_coords *ff, double *ke) {
    if (!c || !ff) return fail(SOFIA_E_ARG, "null argument");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    float *st = staging_f32(c);
    TRY(upload(st, ff, sizeof(float) * 3 * (size_t)c->N));
    SOFIA_HIP(hipMemsetAsync(c->eacc, 0, sizeof(long long) * eacc_len(c->nrep), c->stream));
    launch_nve_v32(c, dt_partial, st);
    std::vector<long long> e;
    TRY(read_energies(c, e));
    if (ke) {
        double s = 0;
        for (int r = 0; r < c->nrep; ++r) s += E(e, r, E_KE);
        *ke = s;
    }
    return SOFIA_OK;
}

int sofia_nve_x(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    launch_nve_x32(c);
    return sync_check(c);
}

namespace {
int reference_regrid(sofia_ctx *c);
}

int sofia_berendsen(sofia_ctx *c, const double *mu, float *box_out) {
    if (!c || !mu) return fail(SOFIA_E_ARG, "null argument");
    if (c->engine_ready) return fail(SOFIA_E_STATE, "berendsen is a drop-in; the engine runs at fixed volume");
    for (int r = 0; r < c->nrep; ++r)
        if (!(mu[r] > 0.5 && mu[r] < 2.0)) return fail(SOFIA_E_ARG, "box scale factor mu out of (0.5, 2)");
    double *mu_d = nullptr;
    TRY(dalloc(&mu_d, c->nrep));
    int rc = upload(mu_d, mu, sizeof(double) * c->nrep);
    if (rc == SOFIA_OK) {
        launch_berendsen32(c, mu_d);
        rc = sync_check(c);
    }
    dfree(mu_d);
    TRY(rc);
    // box = box*mu (real*4); ibox = 1.0/box (real*4 division), Source/mod_ensemble.f90:175-178;
    // the next neighbor_wrapper re-derives the grid from the new box (check_allocation)
    for (int r = 0; r < c->nrep; ++r) {
        auto &R = c->reps_h[r];
        R.box = (float)((double)R.box * mu[r]);
        R.ibox = 1.0f / R.box;
        if (box_out) box_out[r] = R.box;
    }
    TRY(reference_regrid(c));
    return SOFIA_OK;
}

int sofia_pbc(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    launch_pbc32(c);
    return sync_check(c);
}

int sofia_shake(sofia_ctx *c, int32_t iflag, sofia_coords *ff, double *vir, int32_t *max_iter) {
    if (!c || !ff) return fail(SOFIA_E_ARG, "null argument");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    float *st = staging_f32(c);
    TRY(upload(st, ff, sizeof(float) * 3 * (size_t)c->N));
    SOFIA_HIP(hipMemsetAsync(c->eacc, 0, sizeof(long long) * eacc_len(c->nrep), c->stream));
    SOFIA_HIP(hipMemsetAsync(c->shake_err, 0, sizeof(int32_t), c->stream));
    SOFIA_HIP(hipMemsetAsync(c->shake_iters, 0, sizeof(int32_t), c->stream));
    launch_shake32(c, iflag, st, nullptr);
    std::vector<long long> e;
    TRY(read_energies(c, e));
    int32_t err = 0, its = 0;
    TRY(download(&err, c->shake_err, sizeof(int32_t)));
    TRY(download(&its, c->shake_iters, sizeof(int32_t)));
    TRY(download(ff, st, sizeof(float) * 3 * (size_t)c->N));
    if (vir)
        for (int r = 0; r < c->nrep; ++r) {
            vir[0] += E(e, r, V_X); vir[1] += E(e, r, V_Y); vir[2] += E(e, r, V_Z);
        }
    if (max_iter) *max_iter = its;
    if (err) return fail(SOFIA_E_SHAKE, "SHAKE determinant <= 0");
    return SOFIA_OK;
}

int sofia_langevin(sofia_ctx *c, float fric, float gconst, uint64_t seed, uint64_t step,
                   sofia_coords *ff_lang) {
    if (!c || !ff_lang) return fail(SOFIA_E_ARG, "null argument");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    float *st = staging_f32(c);
    launch_langevin32(c, fric, gconst, seed, step, st);
    TRY(sync_check(c));
    return download(ff_lang, st, sizeof(float) * 3 * (size_t)c->N);
}

// ------------------------------------------------------------------ engine
int sofia_set_thermostat(sofia_ctx *c, const sofia_thermostat *t) {
    if (!c || !t) return fail(SOFIA_E_ARG, "null argument");
    bool any = false;
    for (int r = 0; r < c->nrep; ++r) {
        auto &R = c->reps_h[r];
        const double g = t[r].gamma_ps * 1e-3;   // 1/fs
        // Source/get_started.f90:81-84 with the friction as a parameter:
        // fric = gamma*amu2e, gconst = sqrt(12 kB kcal2amu T gamma / (dt/2)) * amu2e
        const double kB = (double)0.001987204181f;
        R.fric = g * c->p.amu2e;
        R.gconst = (g > 0.0) ? std::sqrt(12.0 * kB * (double)c->p.kcal2amu * t[r].temp * g /
                                         (0.5 * (double)c->p.dt)) * c->p.amu2e
                             : 0.0;
        R.seed = t[r].seed;
        any = any || g > 0.0;
    }
    c->has_thermostat = any;
    TRY(upload(c->reps_d, c->reps_h.data(), sizeof(sofia::RepInfo) * c->nrep));
    drop_graph(c);
    return SOFIA_OK;
}

namespace {

int nb_mode(const sofia_ctx *c) {
    if (c->p.coul_flag == SOFIA_COUL_EWALD) return 2;
    return c->p.precision == SOFIA_PREC_MIXED ? 1 : 0;
}

void force_sequence(sofia_ctx *c, bool energies) {
    launch_zero_forces(c, c->fix, energies);
    launch_nonbond(c, nb_mode(c), energies);
    launch_bonded_engine(c);
    if (c->has_thermostat) launch_langevin64(c);
    launch_shake64(c);
}

void step_sequence(sofia_ctx *c, bool energies) {
    launch_nve_v64(c, false);
    launch_nve_x64(c);
    launch_check_neighbor(c, true);
    launch_cell_reset(c, false);
    launch_bin(c);
    launch_cell_scan(c);
    launch_scatter_sort(c);
    launch_gather(c, true);
    launch_remap_topology(c);
    launch_build_list(c);
    launch_finish_rebuild(c, true);
    force_sequence(c, energies);
    launch_nve_v64(c, true);
    launch_step_advance(c);
}

int collect_energies(sofia_ctx *c) {
    std::vector<long long> e;
    TRY(read_energies(c, e));
    for (int r = 0; r < c->nrep; ++r) {
        double *o = c->last_e.data() + (size_t)r * SOFIA_NENERGY;
        o[0] = E(e, r, E_LJ); o[1] = E(e, r, E_COUL); o[2] = E(e, r, E_POT14);
        o[3] = E(e, r, E_COUL14); o[4] = E(e, r, E_BOND); o[5] = E(e, r, E_ANGLE);
        o[6] = E(e, r, E_DIHED); o[7] = E(e, r, E_KE); o[8] = E(e, r, E_IMPRO);
    }
    return SOFIA_OK;
}

int check_engine_flags(sofia_ctx *c, int32_t *overflow = nullptr) {
    int32_t err = 0, mx = 0;
    SOFIA_HIP(hipMemcpyAsync(&err, c->shake_err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipMemcpyAsync(&mx, c->maxcnt, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    TRY(sync_check(c));
    if (overflow) *overflow = mx > c->neigh_alloc ? mx : 0;
    if (mx > c->neigh_alloc) return fail(SOFIA_E_OVERFLOW, "Verlet list overflow during an MD segment");
    if (err)
        return fail(SOFIA_E_SHAKE, "constraint solve failed during an MD segment (SHAKE "
                                   "determinant <= 0, or a rigid water moved beyond SETTLE's reach)");
    return SOFIA_OK;
}

}  // namespace

namespace {

// check_allocation after a box change (Source/mod_neighbor.f90:1434-1526): the reference grid
// of every replica from its current box, per-cell arrays re-sized when the cell count grows.
int reference_regrid(sofia_ctx *c) {
    TRY(sync_check(c));
    drop_graph(c);
    int cell0 = 0;
    for (auto &R : c->reps_h) {
        cell_grid(R.box, c->p.rcut, c->p.cut_coul, &R.rn, &R.ncellT, &R.ncellD);
        if (R.ncellD < 3 || R.ncellD * R.ncellD * R.ncellD != R.ncellT)
            return fail(SOFIA_E_ARG, "box too small for the cell grid after the box change");
        R.stencil5 = 0;
        R.cell0 = cell0;
        cell0 += R.ncellT;
    }
    if (cell0 > c->ncell_cap) {
        dfree(c->ccount); dfree(c->cstart); dfree(c->cnum); dfree(c->cell_rep);
        TRY(dalloc(&c->ccount, cell0));
        TRY(dalloc(&c->cstart, cell0));
        TRY(dalloc(&c->cnum, (size_t)27 * cell0));
        TRY(dalloc(&c->cell_rep, cell0));
        SOFIA_HIP(hipMemset(c->ccount, 0, sizeof(int32_t) * cell0));
        SOFIA_HIP(hipMemset(c->cstart, 0, sizeof(int32_t) * cell0));
        c->ncell_cap = cell0;
    }
    c->ncell_total = cell0;
    c->n_stencil5 = 0;
    std::vector<int32_t> cnum((size_t)27 * cell0), crep(cell0);
    for (int r = 0; r < c->nrep; ++r) {
        const sofia::RepInfo &R = c->reps_h[r];
        stencil(R.ncellT, R.ncellD, R.cell0, cnum.data() + (size_t)27 * R.cell0);
        for (int k = 0; k < R.ncellT; ++k) crep[R.cell0 + k] = r;
    }
    TRY(upload(c->cnum, cnum.data(), sizeof(int32_t) * cnum.size()));
    TRY(upload(c->cell_rep, crep.data(), sizeof(int32_t) * crep.size()));
    TRY(upload(c->reps_d, c->reps_h.data(), sizeof(sofia::RepInfo) * c->nrep));
    c->have_list = false;   // the sort and list refer to the old grid
    return SOFIA_OK;
}

// The engine's cell grid and rebuild threshold, set once per context.
//
// The reference's grid (cells of max(rcut, cut_coul) per side) with its 27-cell stencil
// lists only pairs closer than the cell size, and its criterion (any atom displaced by
// rv - rc) lets pairs enter the cutoff unlisted, so its forces depend on when the list was
// last built.  The engine keeps every pair inside the force cutoff in the list instead:
//  * cells of at least rv/2 per side with a +-2 stencil (k_build_engine) where the box
//    holds five of them, else cells of at least rv with the 27-cell stencil where it holds
//    three, else the reference grid; every pair closer than min(rv, reach) is then listed,
//    reach being 2 cells or 1 cell;
//  * rebuild once any atom of the replica has moved by half the guaranteed skin
//    (min(rv, reach) - rc), less 1e-3 A for float rounding of positions and cell indices.
// Forces are then a function of the positions alone, which makes a replica's trajectory
// independent of its rebuild history (resume into a new batch, served alone or in a mix).
int engine_regrid(sofia_ctx *c) {
    if (c->engine_grid) return SOFIA_OK;
    TRY(sync_check(c));
    drop_graph(c);
    const double rc = std::sqrt(std::max(c->p.rcut2, c->p.cut_coulsq));
    const double rv = c->skin >= 0.0 ? rc + c->skin : std::sqrt(c->p.rv2);
    int cell0 = 0, n5 = 0;
    for (auto &R : c->reps_h) {
        // start from the reference grid every time (the skin may have changed)
        cell_grid(R.box, c->p.rcut, c->p.cut_coul, &R.rn, &R.ncellT, &R.ncellD);
        const int nh = (int)((double)R.box / (0.5 * rv));
        const int n1 = (int)((double)R.box / rv);
        double reach = R.rn;
        R.stencil5 = 0;
        if (nh >= 5 && nh <= 1024) {
            R.ncellD = nh;
            R.rn = R.box / (float)nh;
            R.stencil5 = 1;
            reach = 2.0 * (double)R.rn;
            ++n5;
        } else if (n1 >= 3 && n1 <= R.ncellD) {
            R.ncellD = n1;
            R.rn = R.box / (float)n1;
            reach = R.rn;
        }
        R.ncellT = R.ncellD * R.ncellD * R.ncellD;
        const double half = 0.5 * (std::min(rv, reach) - rc) - 1e-3;
        R.chk2 = half > 0.0 ? half * half : 0.0;
        R.cell0 = cell0;
        cell0 += R.ncellT;
    }
    if (cell0 > c->ncell_cap) {   // finer cells: re-size the per-cell arrays
        dfree(c->ccount); dfree(c->cstart); dfree(c->cnum); dfree(c->cell_rep);
        TRY(dalloc(&c->ccount, cell0));
        TRY(dalloc(&c->cstart, cell0));
        TRY(dalloc(&c->cnum, (size_t)27 * cell0));
        TRY(dalloc(&c->cell_rep, cell0));
        SOFIA_HIP(hipMemset(c->ccount, 0, sizeof(int32_t) * cell0));
        SOFIA_HIP(hipMemset(c->cstart, 0, sizeof(int32_t) * cell0));
        c->ncell_cap = cell0;
    }
    c->ncell_total = cell0;
    c->n_stencil5 = n5;
    c->dp.rv2 = rv * rv;   // list radius of both build kernels from here on
    std::vector<int32_t> cnum((size_t)27 * cell0), crep(cell0);
    for (int r = 0; r < c->nrep; ++r) {
        const sofia::RepInfo &R = c->reps_h[r];
        stencil(R.ncellT, R.ncellD, R.cell0, cnum.data() + (size_t)27 * R.cell0);
        for (int k = 0; k < R.ncellT; ++k) crep[R.cell0 + k] = r;
    }
    TRY(upload(c->cnum, cnum.data(), sizeof(int32_t) * cnum.size()));
    TRY(upload(c->cell_rep, crep.data(), sizeof(int32_t) * crep.size()));
    TRY(upload(c->reps_d, c->reps_h.data(), sizeof(sofia::RepInfo) * c->nrep));
    c->engine_grid = true;
    return SOFIA_OK;
}

// Build every replica's list with 50 % headroom (so an MD segment never has to regrow it)
// and reset the engine's counters and flags.
// Engine constraint work: the clusters solved iteratively (kinds 1, 3, 4, -3), their indices
// grouped by kind so that a wavefront almost always runs one solver, and the -3 clusters that
// are symmetric rigid waters (|O-H1| = |O-H2|, equal H masses), which SETTLE solves in closed
// form.  A SETTLE water is found through its oxygen (AtomArrays::wclus, permuted with the
// atoms), so its solver runs in atom order, next to the waters of the same cell.  Needs the
// atom types, so it runs when the engine sets up.
int classify_shake(sofia_ctx *c) {
    dfree(c->shake_kidx);
    for (int q = 0; q < 5; ++q) c->shake_kcount[q] = 0;
    if (!c->ns) return SOFIA_OK;
    TRY(sync_check(c));
    const int ns = c->ns;
    std::vector<int32_t> num(ns), at((size_t)4 * ns);
    std::vector<double> bond((size_t)3 * ns), mass(c->p.ntypes);
    std::vector<float4> pos(c->N);
    TRY(download(num.data(), c->shake_num, sizeof(int32_t) * ns));
    TRY(download(at.data(), c->shake_atoms, sizeof(int32_t) * at.size()));
    TRY(download(bond.data(), c->shake_bond, sizeof(double) * bond.size()));
    TRY(download(mass.data(), c->mass, sizeof(double) * mass.size()));
    TRY(download(pos.data(), c->a.pos4, sizeof(float4) * pos.size()));
    auto m_of = [&](int i) { return mass[__builtin_bit_cast(int, pos[i].w) - 1]; };
    std::vector<int32_t> lists[5];
    for (int k = 0; k < ns; ++k) {
        int q = num[k] == 1 ? 0 : num[k] == 3 ? 1 : num[k] == 4 ? 2 : 3;
        if (q == 3 && bond[3 * (size_t)k] == bond[3 * (size_t)k + 1] &&
            m_of(at[4 * (size_t)k + 1]) == m_of(at[4 * (size_t)k + 2]) &&
            bond[3 * (size_t)k + 2] < 2.0 * bond[3 * (size_t)k])
            q = 4;
        lists[q].push_back(k);
    }
    std::vector<int32_t> kidx;
    for (int q = 0; q < 5; ++q) {
        if (q < 4) kidx.insert(kidx.end(), lists[q].begin(), lists[q].end());
        c->shake_kcount[q] = (int32_t)lists[q].size();
    }
    if (!kidx.empty()) {
        TRY(dalloc(&c->shake_kidx, kidx.size()));
        TRY(upload(c->shake_kidx, kidx.data(), sizeof(int32_t) * kidx.size()));
    }
    std::vector<int32_t> wclus(c->N, -1);
    for (int k : lists[4]) wclus[at[4 * (size_t)k]] = k;
    return upload(c->a.wclus, wclus.data(), sizeof(int32_t) * wclus.size());
}

// Charges by atom type (and fewer than 2^24 atoms)?  Then the mixed pair kernel reads
// 12-byte position records and takes C q_i q_j from a per-type-pair table, formed in float32
// exactly as the kernel forms it from per-atom charges ((C q_i) q_j), so the forces do not
// change.
int classify_charges(sofia_ctx *c) {
    c->q_by_type = false;
    // 12-byte record offsets are formed with a 24-bit multiply: fewer than 2^24 atoms
    if (c->p.precision != SOFIA_PREC_MIXED || c->p.coul_flag == SOFIA_COUL_EWALD || c->lj_host.empty() ||
        c->N >= (1 << 24))
        return SOFIA_OK;
    TRY(sync_check(c));
    const int n = c->N, nt = c->p.ntypes;
    std::vector<float> q(n);
    std::vector<float4> pos(n);
    TRY(download(q.data(), c->a.q, sizeof(float) * n));
    TRY(download(pos.data(), c->a.pos4, sizeof(float4) * n));
    std::vector<float> qt(nt, 0.0f);
    std::vector<char> seen(nt, 0);
    for (int i = 0; i < n; ++i) {
        const int t = __builtin_bit_cast(int, pos[i].w) - 1;
        if (t < 0 || t >= nt) return SOFIA_OK;
        if (!seen[t]) {
            seen[t] = 1;
            qt[t] = q[i];
        } else if (__builtin_bit_cast(uint32_t, qt[t]) != __builtin_bit_cast(uint32_t, q[i])) {
            return SOFIA_OK;
        }
    }
    std::vector<float4> pt((size_t)nt * nt);
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            const float4 l = c->lj_host[(size_t)a * nt + b];
            volatile float qc = c->dp.coulpre * qt[a];
            volatile float qq = qc * qt[b];
            pt[(size_t)a * nt + b] = make_float4(l.x, l.y, qq, 0.0f);
        }
    if (!c->ptab) TRY(dalloc(&c->ptab, pt.size()));
    TRY(upload(c->ptab, pt.data(), sizeof(float4) * pt.size()));
    c->q_by_type = true;
    return SOFIA_OK;
}

int engine_lists(sofia_ctx *c) {
    TRY(engine_regrid(c));
    TRY(classify_charges(c));
    TRY(split_water(c));
    TRY(classify_shake(c));
    launch_set_rebuild(c, 1);
    TRY(rebuild_now(c, true, false));
    std::vector<int32_t> nn(c->N);
    TRY(download(nn.data(), c->numneigh, sizeof(int32_t) * c->N));
    int32_t mx = 0;
    for (int v : nn) mx = std::max(mx, v);
    const int want = engine_stride(c, mx);
    if (want > c->neigh_alloc) {
        TRY(ensure_list(c, want));
        launch_set_rebuild(c, 1);
        TRY(rebuild_now(c, true, false));
    }
    TRY(build_bonded_bundles(c));   // molecules packed in the order the sort just made
    SOFIA_HIP(hipMemsetAsync(c->nbuilds, 0, sizeof(int32_t) * c->nrep, c->stream));
    SOFIA_HIP(hipMemsetAsync(c->shake_err, 0, sizeof(int32_t), c->stream));
    SOFIA_HIP(hipMemsetAsync(c->maxcnt, 0, sizeof(int32_t), c->stream));
    drop_graph(c);
    return SOFIA_OK;
}

}  // namespace

int sofia_md_init(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    launch_engine_from32(c);
    c->engine_ready = true;
    TRY(engine_lists(c));
    force_sequence(c, true);
    TRY(collect_energies(c));
    return check_engine_flags(c);
}

int sofia_md_resume(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->have_ff) return fail(SOFIA_E_STATE, "force field not set");
    // FP64 state came in through sofia_put_state_f64: refresh the float copies
    launch_engine_to32(c);
    SOFIA_HIP(hipMemsetAsync(c->drd, 0, sizeof(double) * 3 * (size_t)c->N, c->stream));
    c->engine_ready = true;
    TRY(engine_lists(c));
    // the forces of the step that produced this state carried the Langevin draw of the
    // previous counter value
    launch_step_add(c, -1);
    force_sequence(c, true);
    launch_step_add(c, 1);
    TRY(collect_energies(c));
    return check_engine_flags(c);
}

int sofia_get_forces_f64(sofia_ctx *c, double *f) {
    if (!c || !f) return fail(SOFIA_E_ARG, "null argument");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    TRY(sync_check(c));
    const size_t n = (size_t)c->N;
    std::vector<double> nb(3 * n);
    std::vector<long long> fx(3 * n);
    TRY(download(nb.data(), c->fnb, sizeof(double) * nb.size()));
    TRY(download(fx.data(), c->fix, sizeof(long long) * fx.size()));
    // the sum k_nve_v64 forms (FIX_INV is a power of two, so the product is exact)
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) f[3 * i + d] = nb[d * n + i] + (double)fx[d * n + i] * FIX_INV;
    return SOFIA_OK;
}

int sofia_md_pair_forces(sofia_ctx *c, double *f, double *e) {
    if (!c || !f) return fail(SOFIA_E_ARG, "null argument");
    if (!c->engine_ready || !c->have_list) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    const size_t n = (size_t)c->N, fb = sizeof(double) * 3 * n;
    // the pair forces the last step carries are set aside and put back afterwards
    launch_zero_forces(c, c->fscratch, true);
    SOFIA_HIP(hipMemcpyAsync(c->fscratch, c->fnb, fb, hipMemcpyDeviceToDevice, c->stream));
    launch_nonbond(c, nb_mode(c), true);
    std::vector<double> nb(3 * n);
    std::vector<long long> eq;
    SOFIA_HIP(hipMemcpyAsync(nb.data(), c->fnb, fb, hipMemcpyDeviceToHost, c->stream));
    TRY(read_energies(c, eq));
    SOFIA_HIP(hipMemcpyAsync(c->fnb, c->fscratch, fb, hipMemcpyDeviceToDevice, c->stream));
    TRY(sync_check(c));
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) f[3 * i + d] = nb[d * n + i];
    if (e)
        for (int r = 0; r < c->nrep; ++r) {
            double *o = e + (size_t)r * 5;
            o[0] = E(eq, r, E_LJ); o[1] = E(eq, r, E_COUL);
            o[2] = E(eq, r, V_X); o[3] = E(eq, r, V_Y); o[4] = E(eq, r, V_Z);
        }
    return SOFIA_OK;
}

int sofia_get_nlist_replica(sofia_ctx *c, int32_t rep, int32_t *nlist, int32_t *numneigh,
                            int32_t *na) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (rep < 0 || rep >= c->nrep) return fail(SOFIA_E_ARG, "replica index out of range");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no list built yet");
    TRY(sync_check(c));
    const auto &R = c->reps_h[rep];
    const size_t stride = (size_t)c->neigh_alloc;
    if (na) *na = c->neigh_alloc;
    std::vector<int32_t> nn(R.natom);
    TRY(download(nn.data(), c->numneigh + R.atom0, sizeof(int32_t) * R.natom));
    if (numneigh) std::memcpy(numneigh, nn.data(), sizeof(int32_t) * R.natom);
    if (nlist) {
        TRY(download(nlist, c->nlist + (size_t)R.atom0 * stride, sizeof(int32_t) * R.natom * stride));
        for (int k = 0; k < R.natom; ++k) {
            int32_t *row = nlist + (size_t)k * stride;
            const int m = std::min<int>(nn[k], (int)stride);
            for (int q = 0; q < m; ++q) row[q] = row[q] - R.atom0 + 1;
            for (size_t q = m; q < stride; ++q) row[q] = 0;
        }
    }
    return SOFIA_OK;
}

int sofia_get_engine_config(sofia_ctx *c, int32_t *lanes, int32_t *list_waves,
                            int32_t *bonded_mode, int32_t *record_bytes, double *skin) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (lanes) *lanes = c->lanes;
    if (list_waves) *list_waves = c->list_waves;
    if (bonded_mode) *bonded_mode = c->bonded_bundles ? 1 : 0;
    if (record_bytes)
        *record_bytes = c->p.precision != SOFIA_PREC_MIXED || c->p.coul_flag == SOFIA_COUL_EWALD
                            ? 0
                            : ((c->q_by_type && c->ptab && c->lanes != 0) ? 12 : 16);
    if (skin) *skin = c->skin >= 0.0 ? c->skin : std::sqrt(c->p.rv2) - std::sqrt(std::max(c->p.rcut2, c->p.cut_coulsq));
    return SOFIA_OK;
}

namespace {

// Where each (replica, global id) sits in the current device order.
int gid_slots(sofia_ctx *c, std::vector<int32_t> &slot) {
    const int n = c->N;
    std::vector<int32_t> g(n);
    TRY(download(g.data(), c->a.gid, sizeof(int32_t) * n));
    slot.assign(n, -1);
    for (const auto &R : c->reps_h)
        for (int k = 0; k < R.natom; ++k) {
            const int gi = g[R.atom0 + k];
            if (gi < 1 || gi > R.natom || slot[R.atom0 + gi - 1] >= 0)
                return fail(SOFIA_E_STATE, "global ids of a replica are not a permutation of 1..natom");
            slot[R.atom0 + gi - 1] = R.atom0 + k;
        }
    return SOFIA_OK;
}

// Set the carried force of every replica r with have[r] (nullptr = all) to f, given in the
// device order described by `slot_before` (gid_slots taken when f was in that order).
int put_forces(sofia_ctx *c, const double *f, const std::vector<int32_t> &slot_before,
               const int32_t *have) {
    const int n = c->N;
    std::vector<int32_t> g(n);
    std::vector<double> nb((size_t)3 * n);
    std::vector<long long> fx((size_t)3 * n);
    TRY(download(g.data(), c->a.gid, sizeof(int32_t) * n));
    TRY(download(nb.data(), c->fnb, sizeof(double) * nb.size()));
    TRY(download(fx.data(), c->fix, sizeof(long long) * fx.size()));
    for (int r = 0; r < c->nrep; ++r) {
        if (have && !have[r]) continue;
        const auto &R = c->reps_h[r];
        for (int k = 0; k < R.natom; ++k) {
            const int j = R.atom0 + k;
            const size_t src = (size_t)slot_before[R.atom0 + g[j] - 1];
            for (int d = 0; d < 3; ++d) {
                nb[(size_t)d * n + j] = f[3 * src + d];
                fx[(size_t)d * n + j] = 0;
            }
        }
    }
    TRY(upload(c->fnb, nb.data(), sizeof(double) * nb.size()));
    TRY(upload(c->fix, fx.data(), sizeof(long long) * fx.size()));
    return SOFIA_OK;
}

}  // namespace

int sofia_md_resume_forces(sofia_ctx *c, const double *f, const int32_t *have) {
    if (!c || !f || !have) return fail(SOFIA_E_ARG, "null argument");
    TRY(sync_check(c));
    std::vector<int32_t> slot;
    TRY(gid_slots(c, slot));          // the caller's order: before the sort
    TRY(sofia_md_resume(c));
    return put_forces(c, f, slot, have);
}

namespace {

// The bonded and constraint lists that the cell sort re-indexes in place.
std::vector<std::pair<int32_t *, size_t>> sorted_lists(sofia_ctx *c) {
    return {{c->bondlist, (size_t)3 * c->nb},   {c->anglelist, (size_t)4 * c->na},
            {c->dihedlist, (size_t)5 * c->nd},  {c->torsions, (size_t)5 * c->ntor},
            {c->improlist, (size_t)5 * c->nimp}, {c->shake_atoms, (size_t)4 * c->ns},
            {c->ebond, (size_t)3 * c->nbe},      {c->eangle, (size_t)4 * c->nae}};
}

// The state an MD segment starts from (every per-atom array the sort permutes, the lists it
// re-indexes, the carried forces, displacements, step counters and build counts), copied
// on the context stream ahead of the segment's kernels.
int snapshot_take(sofia_ctx *c) {
    const size_t n = (size_t)c->N;
    if (!c->snap.pos4) {
        TRY(alloc_atoms(c->snap, c->N));
        TRY(dalloc(&c->snap_fnb, 3 * n));
        TRY(dalloc(&c->snap_drd, 3 * n));
        TRY(dalloc(&c->snap_fix, 3 * n));
        TRY(dalloc(&c->snap_step, c->nrep));
        TRY(dalloc(&c->snap_nbuilds, c->nrep));
    }
    size_t topo = 0;
    for (auto &l : sorted_lists(c)) topo += l.first ? l.second : 0;
    if (topo > c->snap_topo_cap) {
        TRY(sync_check(c));
        dfree(c->snap_topo);
        TRY(dalloc(&c->snap_topo, topo));
        c->snap_topo_cap = topo;
    }
    hipStream_t s = c->stream;
    TRY(copy_atoms(c->snap, c->a, n, s));
    size_t off = 0;
    for (auto &l : sorted_lists(c)) {
        if (!l.first || !l.second) continue;
        SOFIA_HIP(hipMemcpyAsync(c->snap_topo + off, l.first, sizeof(int32_t) * l.second,
                                 hipMemcpyDeviceToDevice, s));
        off += l.second;
    }
    SOFIA_HIP(hipMemcpyAsync(c->snap_fnb, c->fnb, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->snap_drd, c->drd, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->snap_fix, c->fix, sizeof(long long) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->snap_step, c->step_d, sizeof(uint64_t) * c->nrep, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->snap_nbuilds, c->nbuilds, sizeof(int32_t) * c->nrep,
                             hipMemcpyDeviceToDevice, s));
    c->snap_valid = true;
    return SOFIA_OK;
}

// Back to the segment's starting state, then lists rebuilt for every replica with a stride
// that holds `overflow` entries with room to spare, the carried forces following the atoms
// through the re-sort.  The engine's forces depend on positions only (complete lists, exact
// sums), so the segment run again from here is the one that would have run without the
// overflow, bit for bit.
int snapshot_restore(sofia_ctx *c, int32_t overflow) {
    const size_t n = (size_t)c->N;
    hipStream_t s = c->stream;
    TRY(sync_check(c));
    TRY(copy_atoms(c->a, c->snap, n, s));
    size_t off = 0;
    for (auto &l : sorted_lists(c)) {
        if (!l.first || !l.second) continue;
        SOFIA_HIP(hipMemcpyAsync(l.first, c->snap_topo + off, sizeof(int32_t) * l.second,
                                 hipMemcpyDeviceToDevice, s));
        off += l.second;
    }
    SOFIA_HIP(hipMemcpyAsync(c->fnb, c->snap_fnb, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->drd, c->snap_drd, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->fix, c->snap_fix, sizeof(long long) * 3 * n, hipMemcpyDeviceToDevice, s));
    SOFIA_HIP(hipMemcpyAsync(c->step_d, c->snap_step, sizeof(uint64_t) * c->nrep, hipMemcpyDeviceToDevice, s));
    TRY(sync_check(c));
    std::vector<double> f(3 * n);
    std::vector<int32_t> slot;
    TRY(sofia_get_forces_f64(c, f.data()));
    TRY(gid_slots(c, slot));
    c->stride_floor = std::max(c->stride_floor, ((overflow * 5 / 4 + 15) / 16) * 16);
    TRY(engine_lists(c));
    TRY(put_forces(c, f.data(), slot, nullptr));
    SOFIA_HIP(hipMemcpyAsync(c->nbuilds, c->snap_nbuilds, sizeof(int32_t) * c->nrep,
                             hipMemcpyDeviceToDevice, s));
    return sync_check(c);
}

int capture_steps(sofia_ctx *c, hipGraphExec_t *out, int k) {
    hipGraph_t g = nullptr;
    SOFIA_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < k; ++s) step_sequence(c, false);
    SOFIA_HIP(hipStreamEndCapture(c->stream, &g));
    // the executable graph is independent of the captured one, which goes at once
    const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        *out = nullptr;
        return check_hip(ie, "hipGraphInstantiate");
    }
    return SOFIA_OK;
}

}  // namespace

int sofia_md_launch(sofia_ctx *c, int32_t nsteps, int32_t use_graph) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    if (nsteps <= 0) return SOFIA_OK;
    if (c->redo) {
        TRY(snapshot_take(c));
        c->snap_nsteps = nsteps;
        c->snap_graph = use_graph;
    }
    const int body = nsteps - 1;   // the last step also reduces energies
    if (use_graph) {
        const int k = c->graph_steps;
        if (!c->graph) TRY(capture_steps(c, &c->graph, 1));
        if (k > 1 && body >= k && !c->graphk) TRY(capture_steps(c, &c->graphk, k));
        int s = 0;
        if (k > 1 && c->graphk)
            for (; s + k <= body; s += k) SOFIA_HIP(hipGraphLaunch(c->graphk, c->stream));
        for (; s < body; ++s) SOFIA_HIP(hipGraphLaunch(c->graph, c->stream));
    } else {
        for (int s = 0; s < body; ++s) step_sequence(c, false);
    }
    step_sequence(c, true);
    c->step_h += nsteps;
    return SOFIA_OK;
}

int sofia_md_finish(sofia_ctx *c) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    for (int attempt = 0;; ++attempt) {
        int32_t overflow = 0;
        const int rc = check_engine_flags(c, &overflow);
        if (rc == SOFIA_E_OVERFLOW && c->snap_valid && attempt < 4) {
            // grow and redo the segment (the reference grows neigh_alloc and rebuilds)
            const int32_t nsteps = c->snap_nsteps, graph = c->snap_graph;
            TRY(snapshot_restore(c, overflow));
            c->step_h -= nsteps;
            ++c->redos;
            TRY(sofia_md_launch(c, nsteps, graph));
            continue;
        }
        c->snap_valid = false;
        if (rc != SOFIA_OK) return rc;
        return collect_energies(c);
    }
}

int sofia_set_graph_steps(sofia_ctx *c, int32_t k) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (k < 1 || k > 200) return fail(SOFIA_E_ARG, "steps per graph must be 1..200");
    if (k != c->graph_steps) {
        TRY(sync_check(c));
        drop_graph(c);
        c->graph_steps = k;
    }
    return SOFIA_OK;
}

int sofia_set_list_headroom(sofia_ctx *c, double headroom) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!(headroom >= 1.0) || headroom > 4.0) return fail(SOFIA_E_ARG, "list headroom must be in [1, 4]");
    c->headroom = headroom;
    return SOFIA_OK;
}

int sofia_md_list_stats(sofia_ctx *c, int32_t *neigh_alloc, int32_t *redos) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (neigh_alloc) *neigh_alloc = c->neigh_alloc;
    if (redos) *redos = c->redos;
    return SOFIA_OK;
}

int sofia_md_run(sofia_ctx *c, int32_t nsteps, int32_t use_graph) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (nsteps <= 0) return c->engine_ready ? SOFIA_OK : fail(SOFIA_E_STATE, "call sofia_md_init first");
    TRY(sofia_md_launch(c, nsteps, use_graph));
    return sofia_md_finish(c);
}

int sofia_minimize(sofia_ctx *c, int32_t nsteps, double gamma, double max_disp) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    if (nsteps < 0 || !(gamma > 0.0) || !(max_disp > 0.0)) return fail(SOFIA_E_ARG, "bad minimizer arguments");
    const bool thermo = c->has_thermostat;
    c->has_thermostat = false;
    for (int s = 0; s < nsteps; ++s) {
        launch_zero_forces(c, c->fix, false);
        launch_nonbond(c, nb_mode(c), false);
        launch_bonded_engine(c);
        launch_water_bonded(c);
        launch_minimize_step(c, gamma, max_disp);
        launch_check_neighbor(c, true);
        launch_cell_reset(c, false);
        launch_bin(c);
        launch_cell_scan(c);
        launch_scatter_sort(c);
        launch_gather(c, true);
        launch_remap_topology(c);
        launch_build_list(c);
        launch_finish_rebuild(c, true);
    }
    c->has_thermostat = thermo;
    force_sequence(c, true);
    TRY(collect_energies(c));
    drop_graph(c);
    return check_engine_flags(c);
}

int sofia_set_replica_steps(sofia_ctx *c, const uint64_t *steps) {
    if (!c || !steps) return fail(SOFIA_E_ARG, "null argument");
    SOFIA_HIP(hipMemcpyAsync(c->step_d, steps, sizeof(uint64_t) * c->nrep, hipMemcpyHostToDevice, c->stream));
    return sync_check(c);
}

int sofia_get_replica_steps(sofia_ctx *c, uint64_t *steps) {
    if (!c || !steps) return fail(SOFIA_E_ARG, "null argument");
    SOFIA_HIP(hipMemcpyAsync(steps, c->step_d, sizeof(uint64_t) * c->nrep, hipMemcpyDeviceToHost, c->stream));
    return sync_check(c);
}

int sofia_profile_nonbond(sofia_ctx *c, int32_t launches, double *ms_per_launch,
                          int64_t *pairs_in_cutoff) {
    if (!c || launches < 1) return fail(SOFIA_E_ARG, "bad argument");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list");
    hipEvent_t t0, t1;
    SOFIA_HIP(hipEventCreate(&t0));
    SOFIA_HIP(hipEventCreate(&t1));
    SOFIA_HIP(hipEventRecord(t0, c->stream));
    for (int k = 0; k < launches; ++k) launch_nonbond(c, nb_mode(c), false);
    SOFIA_HIP(hipEventRecord(t1, c->stream));
    SOFIA_HIP(hipEventSynchronize(t1));
    float ms = 0.0f;
    SOFIA_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (ms_per_launch) *ms_per_launch = (double)ms / launches;
    if (pairs_in_cutoff) {
        unsigned long long *d = nullptr;
        TRY(dalloc(&d, 1));
        SOFIA_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream));
        launch_count_pairs(c, d);
        unsigned long long h = 0;
        SOFIA_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        TRY(sync_check(c));
        dfree(d);
        *pairs_in_cutoff = (int64_t)h;
    }
    // the profiled launches overwrote the nonbond forces with the same values; nothing to restore
    return sync_check(c);
}

int sofia_profile_copy(sofia_ctx *c, int64_t bytes, int32_t launches, double *ms_per_launch) {
    if (!c || bytes < 16 || bytes % 16 || launches < 1 || !ms_per_launch) return fail(SOFIA_E_ARG, "bad argument");
    char *a = nullptr, *b = nullptr;
    TRY(dalloc(&a, (size_t)bytes));
    int rc = dalloc(&b, (size_t)bytes);
    if (rc != SOFIA_OK) { dfree(a); return rc; }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.0f;
    if (hipMemsetAsync(a, 1, (size_t)bytes, c->stream) != hipSuccess ||
        hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) {
        rc = fail(SOFIA_E_HIP, "copy set-up failed");
    } else {
        launch_copy16(c, (size_t)bytes / 16, a, b);   // warm
        (void)hipEventRecord(t0, c->stream);
        for (int k = 0; k < launches; ++k) launch_copy16(c, (size_t)bytes / 16, a, b);
        (void)hipEventRecord(t1, c->stream);
        if (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
            rc = fail(SOFIA_E_HIP, "copy timing failed");
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipStreamSynchronize(c->stream);
    dfree(a);
    dfree(b);
    TRY(rc);
    *ms_per_launch = (double)ms / launches;
    return SOFIA_OK;
}

int sofia_md_energies(sofia_ctx *c, double *out) {
    if (!c || !out) return fail(SOFIA_E_ARG, "null argument");
    std::memcpy(out, c->last_e.data(), sizeof(double) * c->last_e.size());
    return SOFIA_OK;
}

int sofia_md_rebuilds(sofia_ctx *c, int32_t *out) {
    if (!c || !out) return fail(SOFIA_E_ARG, "null argument");
    TRY(sync_check(c));
    return download(out, c->nbuilds, sizeof(int32_t) * c->nrep);
}

int sofia_set_tuning(sofia_ctx *c, int32_t lanes, int32_t block, int32_t list_waves) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return fail(SOFIA_E_ARG, "lanes_per_atom must be 0 (compacting kernel), 4, 8, 16, 32 or 64");
    if (block < 64 || block > 1024 || block % 64) return fail(SOFIA_E_ARG, "block must be a multiple of 64 up to 1024");
    if (list_waves != 2 && list_waves != 4 && list_waves != 8 && list_waves != 16)
        return fail(SOFIA_E_ARG, "list_waves must be 2, 4, 8 or 16");
    c->lanes = lanes;
    c->block = block;
    c->list_waves = list_waves;
    drop_graph(c);
    return SOFIA_OK;
}

int sofia_profile_step(sofia_ctx *c, int32_t nsteps, int32_t full_rebuild, double *phase_ms) {
    if (!c || !phase_ms || nsteps < 1) return fail(SOFIA_E_ARG, "bad argument");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    constexpr int NP = SOFIA_NPHASE;
    hipEvent_t ev[NP + 1];
    for (int k = 0; k <= NP; ++k) SOFIA_HIP(hipEventCreate(&ev[k]));
    std::vector<double> acc(NP, 0.0);
    int rc = SOFIA_OK;
    for (int s = 0; s < nsteps && rc == SOFIA_OK; ++s) {
        int k = 0;
        auto mark = [&]() { (void)hipEventRecord(ev[k++], c->stream); };
        mark();
        launch_nve_v64(c, false); mark();
        launch_nve_x64(c); mark();
        if (full_rebuild) launch_set_rebuild(c, 1);
        launch_check_neighbor(c, true); mark();
        launch_cell_reset(c, false); launch_bin(c); launch_cell_scan(c); launch_scatter_sort(c);
        launch_gather(c, true); launch_remap_topology(c); mark();
        launch_build_list(c); mark();
        launch_finish_rebuild(c, true); mark();
        launch_zero_forces(c, c->fix, false); mark();
        launch_nonbond(c, nb_mode(c), false); mark();
        launch_bonded_engine(c); mark();
        if (c->has_thermostat) launch_langevin64(c);
        mark();
        launch_shake64(c); mark();
        launch_nve_v64(c, false); launch_step_advance(c); mark();
        rc = sync_check(c);
        for (int p = 0; p < NP && rc == SOFIA_OK; ++p) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[p], ev[p + 1]) != hipSuccess) rc = fail(SOFIA_E_HIP, "event timing");
            acc[p] += ms;
        }
    }
    for (int k = 0; k <= NP; ++k) (void)hipEventDestroy(ev[k]);
    TRY(rc);
    c->step_h += nsteps;
    for (int p = 0; p < NP; ++p) phase_ms[p] = acc[p] / nsteps;
    return check_engine_flags(c);
}

int sofia_set_bonded_mode(sofia_ctx *c, int32_t mode) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (mode != 0 && mode != 1) return fail(SOFIA_E_ARG, "bonded mode must be 0 (per-kind kernels) or 1 (molecule bundles)");
    c->bonded_bundles = mode == 1;
    drop_graph(c);
    return SOFIA_OK;
}

int sofia_set_skin(sofia_ctx *c, double skin) {
    if (!c) return fail(SOFIA_E_ARG, "null context");
    if (!(skin > 0.05) || skin > 10.0) return fail(SOFIA_E_ARG, "skin must be in (0.05, 10] A");
    c->skin = skin;
    c->engine_grid = false;
    if (c->engine_ready) {
        // the re-sort permutes the atoms but not the carried forces: take them along
        TRY(sync_check(c));
        std::vector<double> f((size_t)3 * c->N);
        std::vector<int32_t> slot;
        TRY(sofia_get_forces_f64(c, f.data()));
        TRY(gid_slots(c, slot));
        TRY(engine_lists(c));   // rebuild counters restart from zero
        TRY(put_forces(c, f.data(), slot, nullptr));
    }
    return SOFIA_OK;
}

int sofia_set_molecules(sofia_ctx *c, const int32_t *mol) {
    if (!c || !mol) return fail(SOFIA_E_ARG, "null argument");
    // per-replica molecule ids -> one global index space for the union-find
    std::vector<int32_t> g(c->N, -1), mrep;
    int base = 0;
    for (int r = 0; r < c->nrep; ++r) {
        const auto &R = c->reps_h[r];
        int mx = -1;
        for (int k = 0; k < R.natom; ++k) {
            const int v = mol[R.atom0 + k];
            if (v >= 0) {
                g[R.atom0 + k] = base + v;
                mx = std::max(mx, v);
            }
        }
        for (int k = 0; k <= mx; ++k) mrep.push_back(r);
        base += mx + 1;
    }
    TRY(sync_check(c));
    dfree(c->mol_rep); dfree(c->uf_parent); dfree(c->uf_size); dfree(c->uf_largest);
    dfree(c->gc_first); dfree(c->gc_refgid); dfree(c->gc_cnt); dfree(c->gc_dens); dfree(c->gc_sum);
    dfree(c->gc_refx); dfree(c->gc_com); dfree(c->gc_key); dfree(c->gc_surface);
    c->nmol = base;
    const size_t nm = (size_t)std::max(1, base);
    TRY(dalloc(&c->mol_rep, nm));
    TRY(dalloc(&c->uf_parent, nm));
    TRY(dalloc(&c->uf_size, nm));
    TRY(dalloc(&c->uf_largest, c->nrep + 1));   // per-replica maxima, then an error flag
    TRY(dalloc(&c->gc_first, c->nrep + 1));
    TRY(dalloc(&c->gc_refgid, nm));
    TRY(dalloc(&c->gc_cnt, nm));
    TRY(dalloc(&c->gc_dens, nm));
    TRY(dalloc(&c->gc_sum, 3 * nm));
    TRY(dalloc(&c->gc_refx, 3 * nm));
    TRY(dalloc(&c->gc_com, 3 * nm));
    TRY(dalloc(&c->gc_key, c->nrep));
    TRY(dalloc(&c->gc_surface, c->nrep));
    std::vector<int32_t> first(c->nrep + 1, base);
    for (int m = base - 1; m >= 0; --m) first[mrep[m]] = m;
    for (int r = c->nrep - 1; r >= 0; --r) first[r] = std::min(first[r], first[r + 1]);
    TRY(upload(c->gc_first, first.data(), sizeof(int32_t) * first.size()));
    if (base) TRY(upload(c->mol_rep, mrep.data(), sizeof(int32_t) * base));
    return upload(c->a.mol, g.data(), sizeof(int32_t) * c->N);
}

int sofia_cluster_largest(sofia_ctx *c, float rcontact, int32_t n_o, const int32_t *o_types,
                          int32_t n_n, const int32_t *n_types, int32_t *largest) {
    if (!c || !largest || n_o < 0 || n_n < 0) return fail(SOFIA_E_ARG, "bad argument");
    if (!c->have_list) return fail(SOFIA_E_STATE, "no Verlet list");
    if (!c->uf_parent) return fail(SOFIA_E_STATE, "call sofia_set_molecules first");
    std::vector<int> ot(o_types, o_types + n_o), nt(n_types, n_types + n_n);
    return cluster_largest(c, rcontact, ot, nt, largest);
}

int sofia_cluster_greatest(sofia_ctx *c, double denscut, int32_t densmin, int32_t *total,
                           int32_t *core, int32_t *surface) {
    if (!c || !(denscut > 0.0) || densmin < 0) return fail(SOFIA_E_ARG, "bad argument");
    if (!c->engine_ready) return fail(SOFIA_E_STATE, "call sofia_md_init first");
    if (!c->gc_first) return fail(SOFIA_E_STATE, "call sofia_set_molecules first");
    return cluster_greatest(c, denscut, densmin, total, core, surface);
}

int sofia_cluster_greatest_noalloc(int32_t n, const double *x, const int32_t *gid,
                                   const int32_t *mol, int32_t nmol, float box, double denscut,
                                   int32_t densmin, int32_t *out, void *stream) {
    return cluster_greatest_noalloc(n, x, gid, mol, nmol, box, denscut, densmin, out,
                                    (hipStream_t)stream);
}

}  // extern "C"
