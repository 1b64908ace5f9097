// Solute cluster analysis on the device: exact integer union-find over solute molecules
// joined by an O-N contact shorter than rcontact.
//
// The reference measures nucleation with a centre-of-mass density graph and a recursive
// flood fill (get_greatestCluster, Source/mod_cluster.f90:15-247).  This is the contact
// criterion of the served workload instead: the Verlet list already holds every pair
// within the cutoff, so each N atom scans its list for O atoms of other solute molecules.
// Roots are linked lowest-index-wins with compare-and-swap, so the final partition is
// unique whatever the thread order; sizes and the per-replica maximum are integer atomics.
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

}  // namespace

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

}  // namespace sofia
