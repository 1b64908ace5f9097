// Solute cluster analysis on the device: exact integer union-find over solute molecules
// joined by an O-N contact shorter than rcontact.
//
// The reference measures nucleation with a centre-of-mass density graph and a recursive
// flood fill (get_greatestCluster, Source/mod_cluster.f90:15-247).  This is the contact
// criterion of the served workload instead: the Verlet list already holds every pair
// within the cutoff, so each N atom scans its list for O atoms of other solute molecules.
// Roots are linked lowest-index-wins with compare-and-swap, so the final partition is
// unique whatever the thread order; sizes and the per-replica maximum are integer atomics.
#include <mutex>
#include <string>

#include "sofia_device.h"

namespace sofia {
namespace {

__device__ inline int ld(int32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Parents only ever point to a smaller index (union links the larger root under the
// smaller one, halving moves a node to its grandparent), so the walk strictly decreases
// and terminates even while other threads are linking.
__device__ inline int uf_find(int32_t *parent, int x) {
    for (;;) {
        const int p = ld(&parent[x]);
        if (p == x) return x;
        const int gp = ld(&parent[p]);
        if (gp != p) atomicCAS(&parent[x], p, gp);
        x = p;
    }
}

// A failed CAS means another thread linked a root first, so the loop is lock-free; the
// bound only guards against a corrupted forest, and running out of it is reported through
// err (the host turns it into an error) rather than leaving a union undone unnoticed.
__device__ inline void uf_union(int32_t *parent, int a, int b, int32_t *err) {
    for (int it = 0; it < 1 << 20; ++it) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        if (atomicCAS(&parent[a], a, b) == a) return;
    }
    atomicOr(err, 1);
}

__global__ void k_uf_init(int32_t nmol, int32_t *__restrict__ parent, int32_t *__restrict__ size) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < nmol) {
        parent[m] = m;
        size[m] = 0;
    }
}

__global__ void k_uf_contacts(int32_t n, const float4 *__restrict__ pos4, const int32_t *__restrict__ mol,
                              const int32_t *__restrict__ rep, const RepInfo *__restrict__ reps,
                              const int32_t *__restrict__ nlist, const int32_t *__restrict__ numneigh,
                              int32_t na, uint32_t nmask, uint32_t omask, float rc2,
                              int32_t *__restrict__ parent, int32_t *__restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mi = mol[i];
    if (mi < 0) return;
    const float4 pi = pos4[i];
    const int ti = __float_as_int(pi.w);
    if (ti < 1 || ti > 32 || !((nmask >> (ti - 1)) & 1u)) return;
    const RepInfo &R = reps[rep[i]];
    const int32_t *lst = nlist + (size_t)i * na;
    const int nn = min(numneigh[i], na);   // a row never reads past its stride
    for (int k = 0; k < nn; ++k) {
        const int j = lst[k];
        const int mj = mol[j];
        if (mj < 0 || mj == mi) continue;
        const float4 pj = pos4[j];
        const int tj = __float_as_int(pj.w);
        if (tj < 1 || tj > 32 || !((omask >> (tj - 1)) & 1u)) continue;
        float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= R.box * __builtin_rintf(dx * R.ibox);
        dy -= R.box * __builtin_rintf(dy * R.ibox);
        dz -= R.box * __builtin_rintf(dz * R.ibox);
        if (dx * dx + dy * dy + dz * dz < rc2) uf_union(parent, mi, mj, err);
    }
}

__global__ void k_uf_count(int32_t nmol, int32_t *__restrict__ parent, int32_t *__restrict__ size) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    atomicAdd(&size[uf_find(parent, m)], 1);
}

__global__ void k_uf_largest(int32_t nmol, const int32_t *__restrict__ parent,
                             const int32_t *__restrict__ size, const int32_t *__restrict__ mol_rep,
                             int32_t *__restrict__ largest) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol || parent[m] != m) return;
    atomicMax(&largest[mol_rep[m]], size[m]);
}

// ------------------------------------------------------------------ get_greatestCluster
// The reference's nucleation order parameter (Source/mod_cluster.f90:15-83) on the device:
// solute-molecule centres (xyzcenter_molecule, pbc_molcom), the number of other centres
// closer than denscut (mol_build_neighbor_nosort, Source/mod_neighbor_cluster.f90:56-178),
// the connected components of the dense molecules (buildClusters / harvestCluster: union-
// find here, the lowest index winning, so a component's root is its lowest-numbered
// member), the first largest component and its non-dense neighbours (append_surface_
// molecules).  Molecules of a replica are few (one per solute), so every pair of a replica
// is tested directly.  This file is compiled without FMA contraction: centres and distances
// round as the host's restatement of the same steps does.

constexpr double COM_Q = 4294967296.0;   // molecule offsets in 2^-32 A, summed exactly

// Inputs of the greatest-cluster kernels: atom positions x[i*xs + d*ds] (FP64), global ids,
// molecule per atom (-1 = none), replica per atom and per molecule (nullptr: all in
// replica 0), first molecule of each replica (nrep + 1) and the replica table.
struct GcInput {
    int32_t n, nmol, nrep;
    const double *x;
    int32_t xs, ds;
    const int32_t *gid, *mol, *rep, *mol_rep, *mol_first;
    const RepInfo *reps;
};
// Their working arrays: per molecule (refgid, cnt, dens, parent, size, sum x3, refx x3,
// com x3), per replica (key, surface) and the give-up flag of the union-find.
struct GcWork {
    int32_t *refgid, *cnt, *dens, *parent, *size;
    long long *sum;
    double *refx, *com;
    unsigned long long *key;
    int32_t *surface, *err;
};

__device__ inline int rep_of(const int32_t *r, int i) { return r ? r[i] : 0; }

__global__ void k_gc_clear(GcInput in, GcWork w) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < in.nmol) {
        w.refgid[m] = 0x7fffffff;
        w.sum[3 * (size_t)m] = 0; w.sum[3 * (size_t)m + 1] = 0; w.sum[3 * (size_t)m + 2] = 0;
        w.cnt[m] = 0;
        w.parent[m] = m;
        w.size[m] = 0;
    }
    if (m < in.nrep) {
        w.key[m] = 0;
        w.surface[m] = 0;
    }
    if (m == 0) *w.err = 0;
}

// A molecule label past nmol is reported (flag 2) and its atom left out, never indexed.
__global__ void k_com_ref(GcInput in, GcWork w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.n || in.mol[i] < 0) return;
    if (in.mol[i] >= in.nmol) {
        atomicOr(w.err, 2);
        return;
    }
    atomicMin(&w.refgid[in.mol[i]], in.gid[i]);
}

__global__ void k_com_refpos(GcInput in, GcWork w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.n || in.mol[i] < 0 || in.mol[i] >= in.nmol) return;
    const int m = in.mol[i];
    if (in.gid[i] != w.refgid[m]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) w.refx[3 * (size_t)m + d] = in.x[(size_t)i * in.xs + (size_t)d * in.ds];
}

__global__ void k_com_sum(GcInput in, GcWork w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.n || in.mol[i] < 0 || in.mol[i] >= in.nmol) return;
    const int m = in.mol[i];
    const double bx = (double)in.reps[rep_of(in.rep, i)].box, ib = 1.0 / bx;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double off = in.x[(size_t)i * in.xs + (size_t)d * in.ds] - w.refx[3 * (size_t)m + d];
        off = off - bx * rint(off * ib);
        atomicAdd((unsigned long long *)&w.sum[3 * (size_t)m + d],
                  (unsigned long long)__double2ll_rn(off * COM_Q));
    }
    atomicAdd(&w.cnt[m], 1);
}

__global__ void k_com_final(GcInput in, GcWork w) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= in.nmol) return;
    const double bx = (double)in.reps[rep_of(in.mol_rep, m)].box;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double c = w.cnt[m] ? w.refx[3 * (size_t)m + d] +
                                  ((double)w.sum[3 * (size_t)m + d] * (1.0 / COM_Q)) / (double)w.cnt[m]
                            : 0.0;
        while (c > bx) c = c - bx;      // pbc_molcom
        while (c < 0.0) c = c + bx;
        w.com[3 * (size_t)m + d] = c;
    }
}

__device__ inline double com_dr2(const double *a, const double *b, float box, float ibox) {
    double s[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double dx = a[d] - b[d];
        const int nn = (int)round(dx * (double)ibox);       // nint: half away from zero
        const float shift = -(box * (float)nn);
        s[d] = dx + (double)shift;
    }
    return (s[0] * s[0] + s[1] * s[1]) + s[2] * s[2];
}

// denstrack: other centres of the same replica closer than the cut
__global__ void k_dens_count(GcInput in, GcWork w, double cut2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.nmol) return;
    const int r = rep_of(in.mol_rep, i);
    const float box = in.reps[r].box, ibox = in.reps[r].ibox;
    const double ci[3] = {w.com[3 * (size_t)i], w.com[3 * (size_t)i + 1], w.com[3 * (size_t)i + 2]};
    int cnt = 0;
    for (int j = in.mol_first[r]; j < in.mol_first[r + 1]; ++j)
        if (j != i && com_dr2(ci, w.com + 3 * (size_t)j, box, ibox) < cut2) ++cnt;
    w.dens[i] = cnt;
}

// components of the dense molecules; each pair is linked once (from its lower index)
__global__ void k_dens_union(GcInput in, GcWork w, double cut2, int32_t densmin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= in.nmol || w.dens[i] < densmin) return;
    const int r = rep_of(in.mol_rep, i);
    const float box = in.reps[r].box, ibox = in.reps[r].ibox;
    const double ci[3] = {w.com[3 * (size_t)i], w.com[3 * (size_t)i + 1], w.com[3 * (size_t)i + 2]};
    for (int j = i + 1; j < in.mol_first[r + 1]; ++j)
        if (w.dens[j] >= densmin && com_dr2(ci, w.com + 3 * (size_t)j, box, ibox) < cut2)
            uf_union(w.parent, i, j, w.err);
}

__global__ void k_dens_size(GcInput in, GcWork w, int32_t densmin) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= in.nmol || w.dens[m] < densmin) return;
    atomicAdd(&w.size[uf_find(w.parent, m)], 1);
}

// the first largest component: largest size, then lowest root (lowest member)
__global__ void k_dens_greatest(GcInput in, GcWork w, int32_t densmin) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= in.nmol || w.dens[m] < densmin || w.parent[m] != m) return;
    const unsigned long long k = ((unsigned long long)(uint32_t)w.size[m] << 32) |
                                 (unsigned long long)(0xffffffffu - (uint32_t)m);
    atomicMax(&w.key[rep_of(in.mol_rep, m)], k);
}

// non-dense molecules closer than the cut to a member of the greatest component
__global__ void k_dens_surface(GcInput in, GcWork w, double cut2, int32_t densmin) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= in.nmol || w.dens[j] >= densmin) return;
    const int r = rep_of(in.mol_rep, j);
    const unsigned long long k = w.key[r];
    if (k == 0) return;
    const int root = (int)(0xffffffffu - (uint32_t)(k & 0xffffffffull));
    const float box = in.reps[r].box, ibox = in.reps[r].ibox;
    const double cj[3] = {w.com[3 * (size_t)j], w.com[3 * (size_t)j + 1], w.com[3 * (size_t)j + 2]};
    for (int i = in.mol_first[r]; i < in.mol_first[r + 1]; ++i) {
        if (i == j || w.dens[i] < densmin) continue;
        if (com_dr2(w.com + 3 * (size_t)i, cj, box, ibox) < cut2 && uf_find(w.parent, i) == root) {
            atomicAdd(&w.surface[r], 1);
            return;
        }
    }
}

// The whole analysis on stream s: every working array is set by the kernels themselves.
void launch_greatest(const GcInput &in, const GcWork &w, double denscut, int32_t densmin,
                     hipStream_t s) {
    const double cut2 = denscut * denscut;
    const int nc = in.nmol > in.nrep ? in.nmol : in.nrep;
    const dim3 gc(nblocks(nc > 0 ? nc : 1, 256)), gm(nblocks(in.nmol > 0 ? in.nmol : 1, 256));
    const dim3 ga(nblocks(in.n > 0 ? in.n : 1, 256)), b(256);
    hipLaunchKernelGGL(k_gc_clear, gc, b, 0, s, in, w);
    if (in.nmol <= 0) return;
    hipLaunchKernelGGL(k_com_ref, ga, b, 0, s, in, w);
    hipLaunchKernelGGL(k_com_refpos, ga, b, 0, s, in, w);
    hipLaunchKernelGGL(k_com_sum, ga, b, 0, s, in, w);
    hipLaunchKernelGGL(k_com_final, gm, b, 0, s, in, w);
    hipLaunchKernelGGL(k_dens_count, gm, b, 0, s, in, w, cut2);
    hipLaunchKernelGGL(k_dens_union, gm, b, 0, s, in, w, cut2, densmin);
    hipLaunchKernelGGL(k_dens_size, gm, b, 0, s, in, w, densmin);
    hipLaunchKernelGGL(k_dens_greatest, gm, b, 0, s, in, w, densmin);
    hipLaunchKernelGGL(k_dens_surface, gm, b, 0, s, in, w, cut2, densmin);
}

}  // namespace

int cluster_greatest(sofia_ctx *c, double denscut, int32_t densmin, int32_t *total, int32_t *core,
                     int32_t *surface) {
    const int nrep = c->nrep;
    GcInput in{c->N, c->nmol, nrep, c->a.xd, 1, c->N, c->a.gid, c->a.mol, c->rep, c->mol_rep,
               c->gc_first, c->reps_d};
    // uf_largest holds nrep maxima of the contact analysis and, after them, the error flag
    GcWork w{c->gc_refgid, c->gc_cnt, c->gc_dens, c->uf_parent, c->uf_size, c->gc_sum,
             c->gc_refx, c->gc_com, c->gc_key, c->gc_surface, c->uf_largest + nrep};
    launch_greatest(in, w, denscut, densmin, c->stream);
    std::vector<unsigned long long> k(nrep);
    std::vector<int32_t> s(nrep), e(1);
    SOFIA_HIP(hipMemcpyAsync(k.data(), w.key, sizeof(unsigned long long) * nrep, hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipMemcpyAsync(s.data(), w.surface, sizeof(int32_t) * nrep, hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipMemcpyAsync(e.data(), w.err, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipStreamSynchronize(c->stream));
    if (e[0] & 2) return fail(SOFIA_E_STATE, "an atom's molecule label is nmol or more");
    if (e[0]) return fail(SOFIA_E_STATE, "union-find gave up on a union (corrupted forest)");
    for (int r = 0; r < nrep; ++r) {
        const int32_t cs = (int32_t)(k[r] >> 32);
        if (core) core[r] = cs;
        if (surface) surface[r] = s[r];
        if (total) total[r] = cs + s[r];
    }
    return SOFIA_OK;
}

int cluster_largest(sofia_ctx *c, float rcontact, const std::vector<int> &o_types,
                    const std::vector<int> &n_types, int32_t *largest) {
    uint32_t om = 0, nm = 0;
    for (int t : o_types) {
        if (t < 1 || t > 32) return fail(SOFIA_E_ARG, "O type out of range 1..32");
        om |= 1u << (t - 1);
    }
    for (int t : n_types) {
        if (t < 1 || t > 32) return fail(SOFIA_E_ARG, "N type out of range 1..32");
        nm |= 1u << (t - 1);
    }
    if (rcontact * rcontact >= (float)c->dp.rcut2)
        return fail(SOFIA_E_ARG, "contact distance must be inside the pair cutoff");
    // uf_largest holds nrep maxima and, after them, the give-up flag of uf_union
    int32_t *err = c->uf_largest + c->nrep;
    SOFIA_HIP(hipMemsetAsync(c->uf_largest, 0, sizeof(int32_t) * (c->nrep + 1), c->stream));
    if (c->nmol > 0) {
        hipLaunchKernelGGL(k_uf_init, dim3(nblocks(c->nmol, 256)), dim3(256), 0, c->stream, c->nmol,
                           c->uf_parent, c->uf_size);
        hipLaunchKernelGGL(k_uf_contacts, dim3(nblocks(c->N, 256)), dim3(256), 0, c->stream, c->N,
                           c->a.pos4, c->a.mol, c->rep, c->reps_d, c->nlist, c->numneigh,
                           c->neigh_alloc, nm, om, rcontact * rcontact, c->uf_parent, err);
        hipLaunchKernelGGL(k_uf_count, dim3(nblocks(c->nmol, 256)), dim3(256), 0, c->stream, c->nmol,
                           c->uf_parent, c->uf_size);
        hipLaunchKernelGGL(k_uf_largest, dim3(nblocks(c->nmol, 256)), dim3(256), 0, c->stream,
                           c->nmol, c->uf_parent, c->uf_size, c->mol_rep, c->uf_largest);
    }
    std::vector<int32_t> out(c->nrep + 1);
    SOFIA_HIP(hipMemcpyAsync(out.data(), c->uf_largest, sizeof(int32_t) * (c->nrep + 1),
                             hipMemcpyDeviceToHost, c->stream));
    SOFIA_HIP(hipStreamSynchronize(c->stream));
    if (out[c->nrep]) return fail(SOFIA_E_STATE, "union-find gave up on a union (corrupted forest)");
    for (int r = 0; r < c->nrep; ++r) largest[r] = out[r];
    return SOFIA_OK;
}

// ------------------------------------------------------------------ allocation-free variant
// get_greatestCluster for one system whose positions, ids and molecule labels the caller
// holds on the device, for callers that cannot allocate (graph capture, serving loops).
// Every working array lives in `scratch`, a fixed file-scope device array; each GPU holds
// its own copy until the process ends, and host code reaches it only through the address
// hipGetSymbolAddress returns on that device.  Calls that overlap on one GPU share it and
// race.
namespace {

constexpr int GC_MAX_MOL = 32768;
constexpr size_t GC_ALIGN = 256;
constexpr size_t gc_round(size_t b) { return (b + GC_ALIGN - 1) / GC_ALIGN * GC_ALIGN; }
// layout: RepInfo, molecule range, per-replica key / surface / flag, then per molecule
constexpr size_t GC_OFF_REP = 0;
constexpr size_t GC_OFF_FIRST = GC_OFF_REP + gc_round(sizeof(RepInfo));
constexpr size_t GC_OFF_KEY = GC_OFF_FIRST + gc_round(2 * sizeof(int32_t));
constexpr size_t GC_OFF_SURF = GC_OFF_KEY + gc_round(sizeof(unsigned long long));
constexpr size_t GC_OFF_ERR = GC_OFF_SURF + gc_round(sizeof(int32_t));
constexpr size_t GC_OFF_I32 = GC_OFF_ERR + gc_round(sizeof(int32_t));     // 5 int32 arrays
constexpr size_t GC_OFF_SUM = GC_OFF_I32 + gc_round(5 * sizeof(int32_t) * GC_MAX_MOL);
constexpr size_t GC_OFF_REFX = GC_OFF_SUM + gc_round(3 * sizeof(long long) * GC_MAX_MOL);
constexpr size_t GC_OFF_COM = GC_OFF_REFX + gc_round(3 * sizeof(double) * GC_MAX_MOL);
constexpr size_t GC_SCRATCH_BYTES = GC_OFF_COM + gc_round(3 * sizeof(double) * GC_MAX_MOL);

__device__ __align__(256) unsigned char scratch[GC_SCRATCH_BYTES];

constexpr int GC_MAX_DEV = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[GC_MAX_DEV] = {};

// scratch's address on the current device, looked up once per device
int scratch_on_device(unsigned char **out) {
    int dev = 0;
    SOFIA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= GC_MAX_DEV) return fail(SOFIA_E_NODEV, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        SOFIA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(SOFIA_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return SOFIA_OK;
}

// the one-replica table and molecule range, written on the device (no host copies)
__global__ void k_gc_setup(RepInfo *rep, int32_t *first, float box, int32_t nmol) {
    if (threadIdx.x != 0) return;
    RepInfo R{};
    R.box = box;
    R.ibox = (float)(1.0 / (double)box);
    *rep = R;
    first[0] = 0;
    first[1] = nmol;
}

// out = total, core, surface, union-find give-up flag
__global__ void k_gc_out(const unsigned long long *key, const int32_t *surface, const int32_t *err,
                         int32_t *out) {
    if (threadIdx.x != 0) return;
    const int32_t core = (int32_t)(*key >> 32);
    out[0] = core + *surface;
    out[1] = core;
    out[2] = *surface;
    out[3] = *err;
}

}  // namespace

int cluster_greatest_noalloc(int32_t n, const double *x, const int32_t *gid, const int32_t *mol,
                             int32_t nmol, float box, double denscut, int32_t densmin,
                             int32_t *out, hipStream_t s) {
    if (n < 0 || nmol < 0 || !out || !(denscut > 0.0) || densmin < 0 || !(box > 0.0f))
        return fail(SOFIA_E_ARG, "bad argument");
    if (n > 0 && (!x || !gid || !mol)) return fail(SOFIA_E_ARG, "null device array");
    if (nmol > GC_MAX_MOL)
        return fail(SOFIA_E_ARG, "more molecules than the allocation-free scratch holds (" +
                                     std::to_string(GC_MAX_MOL) + ")");
    unsigned char *p = nullptr;
    const int rc = scratch_on_device(&p);
    if (rc != SOFIA_OK) return rc;
    RepInfo *rep = (RepInfo *)(p + GC_OFF_REP);
    int32_t *first = (int32_t *)(p + GC_OFF_FIRST);
    int32_t *i32 = (int32_t *)(p + GC_OFF_I32);
    GcInput in{n, nmol, 1, x, 3, 1, gid, mol, nullptr, nullptr, first, rep};
    GcWork w{i32, i32 + GC_MAX_MOL, i32 + 2 * GC_MAX_MOL, i32 + 3 * GC_MAX_MOL, i32 + 4 * GC_MAX_MOL,
             (long long *)(p + GC_OFF_SUM), (double *)(p + GC_OFF_REFX), (double *)(p + GC_OFF_COM),
             (unsigned long long *)(p + GC_OFF_KEY), (int32_t *)(p + GC_OFF_SURF),
             (int32_t *)(p + GC_OFF_ERR)};
    hipLaunchKernelGGL(k_gc_setup, dim3(1), dim3(64), 0, s, rep, first, box, nmol);
    launch_greatest(in, w, denscut, densmin, s);
    hipLaunchKernelGGL(k_gc_out, dim3(1), dim3(64), 0, s, w.key, w.surface, w.err, out);
    SOFIA_HIP(hipGetLastError());
    return SOFIA_OK;
}

}  // namespace sofia
