// Cell binning, cell sort, permutation gather and Verlet-list build for gfx950.
//
// Semantics follow Source/mod_neighbor.f90 exactly (bit-exact list parity):
//   xyz2bin        1381-1424  cell = int(dble(x) * (1/dble(rn))), x >= box -> last cell
//   new_nlist +    923-1005   cells in id order; inside a cell, descending old index
//   sort_posit_charge          (the linked-list chain order)
//   build_neighbor_nonewton 90-303 / 622-809: stencil cells in cnum order, atoms of each
//                  cell in index order, real*4 minimum image with round-to-nearest,
//                  dr2 = (dx*dx + dy*dy) + dz*dz in real*4, kept if z != i,
//                  dble(dr2) < rv2 and z not in specbond(1:num3bond(i), i).
// This file is compiled with -ffp-contract=off so no FMA changes the real*4 distance.
//
// GPU design: one workgroup per cell stages the 27 stencil cells (x, y, z, index as a
// float4) in LDS in stencil order; each wavefront takes one i atom at a time and sweeps
// the staged candidates 64 at a time, compacting hits with a ballot + popcount prefix so
// the output keeps the reference's order.  Several replicas share one launch: every
// cell belongs to one replica and only replicas flagged for a rebuild do work.
#include <algorithm>

#include "sofia_internal.h"

namespace sofia {
namespace {

constexpr int LIST_CAP = 4096;   // staged candidates per cell (64 KiB of LDS)

__device__ inline int xyz2bin_dev(float4 p, const RepInfo &R) {
    const double irn = 1.0 / (double)R.rn;
    const double bx = (double)R.box;
    const double c[3] = {(double)p.x, (double)p.y, (double)p.z};
    int ix[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int v = (c[d] < bx) ? (int)(c[d] * irn) : R.ncellD - 1;
        // the reference clamps ncellD back into range; negative inputs are not produced
        // by pbc() but are clamped here so a bad coordinate cannot index out of bounds
        v = v >= R.ncellD ? R.ncellD - 1 : (v < 0 ? 0 : v);
        ix[d] = v;
    }
    return ix[0] + ix[1] * R.ncellD + ix[2] * R.ncellD * R.ncellD;
}

// real*4 minimum image: (s + sign(2^23,s)) - sign(2^23,dx) is round-to-nearest-even of
// s = dx*ibox for |s| < 2^22, i.e. rintf.
__device__ inline float mimg32(float dx, float box, float ibox) {
    const float s = __builtin_rintf(dx * ibox);
    return dx - s * box;
}

// The list-maintenance kernels below run every step and do work only for the replicas
// flagged in `rebuild`; `any` (the word after the nrep flags) is set whenever one is, so a
// step that rebuilds nothing ends each of them after one read.  They loop over their items
// with a fixed grid (at most GRID_CAP workgroups), which keeps such a step's launches short.
constexpr int GRID_CAP = 2048;

__global__ void k_cell_reset(int32_t ncell, const int32_t *__restrict__ cell_rep,
                             const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any,
                             int32_t *__restrict__ ccount) {
    if (!*any) return;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
        if (rebuild[cell_rep[c]]) ccount[c] = 0;
}

__global__ void k_bin(int32_t n, const float4 *__restrict__ pos4, const int32_t *__restrict__ rep,
                      const RepInfo *__restrict__ reps, const int32_t *__restrict__ rebuild,
                      const int32_t *__restrict__ any, int32_t *__restrict__ ccount,
                      int32_t *__restrict__ acell, int32_t *__restrict__ aslot) {
    if (!*any) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int r = rep[i];
        // perm and tt of a replica that rebuilds nothing are stale: every reader (scatter,
        // cell sort, gather, copy-back, topology remap) skips such replicas
        if (!rebuild[r]) continue;
        const RepInfo R = reps[r];
        const int c = R.cell0 + xyz2bin_dev(pos4[i], R);
        acell[i] = c;
        aslot[i] = atomicAdd(&ccount[c], 1);
    }
}

// One workgroup per replica: exclusive scan of its cell counts into absolute starts.
__global__ void k_cell_scan(const RepInfo *__restrict__ reps, const int32_t *__restrict__ rebuild,
                            const int32_t *__restrict__ ccount, int32_t *__restrict__ cstart) {
    const int r = blockIdx.x;
    if (!rebuild[r]) return;
    const RepInfo R = reps[r];
    __shared__ int32_t s[1024];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = R.atom0;
    __syncthreads();
    for (int base = 0; base < R.ncellT; base += 1024) {
        const int c = base + threadIdx.x;
        const int v = (c < R.ncellT) ? ccount[R.cell0 + c] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {          // Hillis-Steele inclusive scan
            const int add = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (c < R.ncellT) cstart[R.cell0 + c] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
}

__global__ void k_scatter(int32_t n, const int32_t *__restrict__ rep,
                          const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any,
                          const int32_t *__restrict__ acell, const int32_t *__restrict__ aslot,
                          const int32_t *__restrict__ cstart, int32_t *__restrict__ sorted_ids) {
    if (!*any) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (rebuild[rep[i]]) sorted_ids[cstart[acell[i]] + aslot[i]] = i;
}

// One workgroup per cell at a time: order the cell's atoms by descending old index (the
// chain order of the reference's linked cells) by ranking, which is deterministic.
__global__ void k_cell_sort(int32_t ncell, const int32_t *__restrict__ cell_rep,
                            const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any,
                            const int32_t *__restrict__ ccount, const int32_t *__restrict__ cstart,
                            const int32_t *__restrict__ sorted_ids, int32_t *__restrict__ perm,
                            int32_t *__restrict__ tt, int32_t *__restrict__ atombin) {
    if (!*any) return;
    constexpr int CAP = 2048;
    __shared__ int32_t ids[CAP];
    for (int c = blockIdx.x; c < ncell; c += gridDim.x) {   // uniform over the workgroup
        if (!rebuild[cell_rep[c]]) continue;
        const int n = ccount[c], base = cstart[c];
        const bool staged = n <= CAP;
        if (staged)
            for (int e = threadIdx.x; e < n; e += blockDim.x) ids[e] = sorted_ids[base + e];
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int me = staged ? ids[e] : sorted_ids[base + e];
            int rank = 0;
            for (int f = 0; f < n; ++f) {
                const int other = staged ? ids[f] : sorted_ids[base + f];
                rank += other > me;
            }
            perm[base + rank] = me;
            tt[me] = base + rank;
            atombin[base + rank] = c;
        }
        __syncthreads();   // ids is refilled for the next cell
    }
}

// Gather every permuted per-atom array from `src` into `dst` (flagged replicas only),
// remapping specbond entries through tt; FP64 engine state too when `engine`.
__global__ void k_gather(int32_t n, const int32_t *__restrict__ rep,
                         const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any,
                         const int32_t *__restrict__ perm, const int32_t *__restrict__ tt,
                         AtomArrays src, AtomArrays dst, int engine) {
    if (!*any) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (!rebuild[rep[k]]) continue;
        const int o = perm[k];
        dst.pos4[k] = src.pos4[o];
        dst.q[k] = src.q[o];
        dst.gid[k] = src.gid[o];
        dst.num1[k] = src.num1[o];
        dst.num2[k] = src.num2[o];
        dst.num3[k] = src.num3[o];
        dst.mol[k] = src.mol[o];
        dst.wclus[k] = src.wclus[o];
#pragma unroll
        for (int l = 0; l < SPEC; ++l) {
            const int s = src.spec[(size_t)o * SPEC + l];
            dst.spec[(size_t)k * SPEC + l] = (s >= 0) ? tt[s] : -1;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) dst.v32[3 * (size_t)k + d] = src.v32[3 * (size_t)o + d];
        if (engine) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dst.xd[(size_t)d * n + k] = src.xd[(size_t)d * n + o];
                dst.vd[(size_t)d * n + k] = src.vd[(size_t)d * n + o];
            }
        }
    }
}

// Copy the gathered arrays back (flagged replicas only) so device pointers stay fixed,
// which keeps captured graphs valid; zero the displacement accumulators.
__global__ void k_copyback(int32_t n, const int32_t *__restrict__ rep,
                           const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any,
                           AtomArrays src, AtomArrays dst, float *__restrict__ dr32,
                           double *__restrict__ drd, int engine) {
    if (!*any) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (!rebuild[rep[k]]) continue;
        dst.pos4[k] = src.pos4[k];
        dst.q[k] = src.q[k];
        dst.gid[k] = src.gid[k];
        dst.num1[k] = src.num1[k];
        dst.num2[k] = src.num2[k];
        dst.num3[k] = src.num3[k];
        dst.mol[k] = src.mol[k];
        dst.wclus[k] = src.wclus[k];
#pragma unroll
        for (int l = 0; l < SPEC; ++l) dst.spec[(size_t)k * SPEC + l] = src.spec[(size_t)k * SPEC + l];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            dst.v32[3 * (size_t)k + d] = src.v32[3 * (size_t)k + d];
            dr32[(size_t)d * n + k] = 0.0f;
        }
        if (engine) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                dst.xd[(size_t)d * n + k] = src.xd[(size_t)d * n + k];
                dst.vd[(size_t)d * n + k] = src.vd[(size_t)d * n + k];
                drd[(size_t)d * n + k] = 0.0;
            }
        }
    }
}

// Every bonded list in one launch: item t of the concatenation falls in list L of S, whose
// items are `width` ints with the first `nidx` atom ids (-1 = unused).  Items of replicas
// that rebuild nothing this step are skipped (their atoms keep their places).
struct RemapSet {
    int32_t *lst[8];
    int32_t end[8];           // cumulative item counts
    int32_t width[8], nidx[8];
    int32_t nlists;
};

__global__ void k_remap_all(RemapSet S, const int32_t *__restrict__ tt, const int32_t *__restrict__ rep,
                            const int32_t *__restrict__ rebuild, const int32_t *__restrict__ any) {
    if (!*any || S.nlists == 0) return;
    const int total = S.end[S.nlists - 1];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int L = 0;
        while (t >= S.end[L]) ++L;
        const int m = t - (L ? S.end[L - 1] : 0);
        int32_t *item = S.lst[L] + (size_t)m * S.width[L];
        const int a0 = item[0];
        if (a0 < 0 || !rebuild[rep[a0]]) continue;
        for (int k = 0; k < S.nidx[L]; ++k) {
            const int v = item[k];
            if (v >= 0) item[k] = tt[v];
        }
    }
}

// A device topology array in the host's layout: in the first nidx columns of each item a
// global 0-based atom id (-1 = empty) becomes the 1-based id within its replica (0 = empty);
// the other columns (types) are copied.
__global__ void k_host_ids(size_t count, int32_t width, int32_t nidx, const int32_t *__restrict__ src,
                           int32_t *__restrict__ dst, const int32_t *__restrict__ rep,
                           const RepInfo *__restrict__ reps) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count;
         e += (size_t)gridDim.x * blockDim.x) {
        const int v = src[e];
        const bool is_id = (int)(e % (size_t)width) < nidx;
        dst[e] = !is_id ? v : (v < 0 ? 0 : v - reps[rep[v]].atom0 + 1);
    }
}

// RDF_calc (Source/mod_force_nonbond.f90:708-775): one wavefront per atom.  An atom of type
// itype walks its list row; each neighbour of type jtype with dr2 < rcut2 (FP64 distance, the
// real*8 minimum image, no contraction in this file) adds 1 to bin int(sqrt(dr2) / delta).
// The counts are integers, so the histogram does not depend on the order of the additions.
__global__ void k_rdf(int32_t n, const float4 *__restrict__ pos4, const int32_t *__restrict__ rep,
                      const RepInfo *__restrict__ reps, const int32_t *__restrict__ nlist,
                      const int32_t *__restrict__ numneigh, int32_t na, double rcut2,
                      double delta, int32_t itype, int32_t jtype, int32_t nbins,
                      uint32_t *__restrict__ hist) {
    const int i = blockIdx.x * (blockDim.x / 64) + (int)(threadIdx.x / 64);
    if (i >= n) return;
    const float4 xi = pos4[i];
    if (__float_as_int(xi.w) != itype) return;   // uniform across the wavefront
    const RepInfo R = reps[rep[i]];
    const double bx = R.box, ib = R.ibox;
    const double x1 = xi.x, y1 = xi.y, z1 = xi.z;
    const int nn = min(numneigh[i], na);
    const int32_t *row = nlist + (size_t)i * na;
    for (int k = (int)(threadIdx.x & 63); k < nn; k += 64) {
        const float4 pj = pos4[row[k]];
        if (__float_as_int(pj.w) != jtype) continue;
        double dx = x1 - (double)pj.x, dy = y1 - (double)pj.y, dz = z1 - (double)pj.z;
        dx = dx - bx * __builtin_rint(dx * ib);
        dy = dy - bx * __builtin_rint(dy * ib);
        dz = dz - bx * __builtin_rint(dz * ib);
        const double dr2 = dx * dx + dy * dy + dz * dz;
        if (dr2 < rcut2) {
            const int node = (int)(sqrt(dr2) / delta);
            if (node < nbins) atomicAdd(&hist[node], 1u);
        }
    }
}

template <int WAVES>
__global__ void __launch_bounds__(WAVES * 64)
k_build_list(const float4 *__restrict__ pos4, const int32_t *__restrict__ num3,
             const int32_t *__restrict__ spec, const int32_t *__restrict__ cstart,
             const int32_t *__restrict__ ccount, const int32_t *__restrict__ cnum,
             const int32_t *__restrict__ cell_rep, const RepInfo *__restrict__ reps,
             const int32_t *__restrict__ rebuild, double rv2, int32_t na,
             int32_t *__restrict__ nlist, int32_t *__restrict__ numneigh,
             int32_t *__restrict__ maxcnt) {
    const int c = blockIdx.x;
    const int r = cell_rep[c];
    if (!rebuild[r] || reps[r].stencil5) return;
    const float box = reps[r].box, ibox = reps[r].ibox;

    __shared__ int32_t s_cell[27];
    __shared__ int32_t s_start[27];
    __shared__ int32_t s_off[28];
    __shared__ float4 s_j[LIST_CAP];

    const int tid = threadIdx.x;
    if (tid < 27) {
        const int nc = cnum[27 * c + tid];
        s_cell[tid] = nc;
        s_start[tid] = cstart[nc];
    }
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int k = 0; k < 27; ++k) {
            s_off[k] = acc;
            acc += ccount[s_cell[k]];
        }
        s_off[27] = acc;
    }
    __syncthreads();
    const int total = s_off[27];
    const bool staged = total <= LIST_CAP;
    if (staged) {
        for (int k = 0; k < 27; ++k) {
            const int cnt = s_off[k + 1] - s_off[k], st = s_start[k];
            for (int t = tid; t < cnt; t += WAVES * 64) {
                float4 p = pos4[st + t];
                p.w = __int_as_float(st + t);
                s_j[s_off[k] + t] = p;
            }
        }
    }
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int i0 = cstart[c], ni = ccount[c];
    for (int ii = wave; ii < ni; ii += WAVES) {
        const int i = i0 + ii;
        const float4 xi = pos4[i];
        const int n3 = num3[i];
        int ex[SPEC];
        int exmin = 0x7fffffff, exmax = -1;
#pragma unroll
        for (int l = 0; l < SPEC; ++l) {
            ex[l] = (l < n3) ? spec[(size_t)i * SPEC + l] : -1;
            if (l < n3) {
                exmin = min(exmin, ex[l]);
                exmax = max(exmax, ex[l]);
            }
        }
        int cnt = 0;
        int32_t *out = nlist + (size_t)i * na;
        for (int base = 0; base < total; base += 64) {
            const int t = base + lane;
            bool pred = false;
            int z = -1;
            if (t < total) {
                float4 pj;
                if (staged) {
                    pj = s_j[t];
                    z = __float_as_int(pj.w);
                } else {
                    int k = 0;
                    while (k < 26 && t >= s_off[k + 1]) ++k;
                    z = s_start[k] + (t - s_off[k]);
                    pj = pos4[z];
                }
                const float dx = mimg32(xi.x - pj.x, box, ibox);
                const float dy = mimg32(xi.y - pj.y, box, ibox);
                const float dz = mimg32(xi.z - pj.z, box, ibox);
                const float dr2 = (dx * dx + dy * dy) + dz * dz;
                pred = (z != i) && ((double)dr2 < rv2);
            }
            const bool maybe = pred && z >= exmin && z <= exmax;
            if (__any(maybe)) {
                if (maybe) {
#pragma unroll
                    for (int l = 0; l < SPEC; ++l)
                        if (ex[l] == z) pred = false;
                }
            }
            const uint64_t mask = __ballot(pred);
            const int pos = cnt + __popcll(mask & lt_mask);
            if (pred && pos < na) out[pos] = z;
            cnt += __popcll(mask);
        }
        if (lane == 0) {
            numneigh[i] = cnt;
            if (cnt > na) atomicMax(maxcnt, cnt);
        }
    }
}

__global__ void k_check32(int32_t n, const int32_t *__restrict__ rep, const float *__restrict__ dr,
                          double rvrc2, int32_t *__restrict__ rebuild, int32_t *__restrict__ any) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = dr[i], y = dr[(size_t)n + i], z = dr[2 * (size_t)n + i];
    const double d2 = (double)((x * x + y * y) + z * z);
    // the first atom to flag its replica also raises `any`
    if (d2 >= rvrc2 && atomicOr(&rebuild[rep[i]], 1) == 0) atomicOr(any, 1);
}

// Engine criterion: per-replica threshold that keeps every pair inside the force cutoff in
// the list (half the usable skin), so forces do not depend on when the list was built.
__global__ void k_check64(int32_t n, const int32_t *__restrict__ rep, const double *__restrict__ dr,
                          const RepInfo *__restrict__ reps, int32_t *__restrict__ rebuild,
                          int32_t *__restrict__ any) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = rep[i];
    const double x = dr[i], y = dr[(size_t)n + i], z = dr[2 * (size_t)n + i];
    if ((x * x + y * y) + z * z >= reps[r].chk2 && atomicOr(&rebuild[r], 1) == 0) atomicOr(any, 1);
}

// rebuild[0..nrep) and `any` = rebuild[nrep] to value
__global__ void k_set_flags(int32_t nrep, int32_t *__restrict__ rebuild, int32_t value) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= nrep) rebuild[r] = value;
}

__global__ void k_finish(int32_t nrep, int32_t *__restrict__ rebuild, int32_t *__restrict__ nbuilds) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrep && rebuild[r]) {
        nbuilds[r] += 1;
        rebuild[r] = 0;
    }
    if (r == nrep) rebuild[r] = 0;   // any
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

}  // namespace

// grid of the looping list-maintenance kernels
inline int capped(int n, int b) { return std::max(1, std::min(nblk(n, b), GRID_CAP)); }
inline const int32_t *any_flag(const sofia_ctx *c) { return c->rebuild + c->nrep; }

void launch_cell_reset(sofia_ctx *c, bool) {
    hipLaunchKernelGGL(k_cell_reset, dim3(capped(c->ncell_total, 256)), dim3(256), 0, c->stream,
                       c->ncell_total, c->cell_rep, c->rebuild, any_flag(c), c->ccount);
}

void launch_bin(sofia_ctx *c) {
    hipLaunchKernelGGL(k_bin, dim3(capped(c->N, 256)), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->rep, c->reps_d, c->rebuild, any_flag(c), c->ccount, c->acell, c->aslot);
}

void launch_cell_scan(sofia_ctx *c) {
    hipLaunchKernelGGL(k_cell_scan, dim3(c->nrep), dim3(1024), 0, c->stream, c->reps_d, c->rebuild,
                       c->ccount, c->cstart);
}

void launch_scatter_sort(sofia_ctx *c) {
    hipLaunchKernelGGL(k_scatter, dim3(capped(c->N, 256)), dim3(256), 0, c->stream, c->N, c->rep,
                       c->rebuild, any_flag(c), c->acell, c->aslot, c->cstart, c->sorted_ids);
    hipLaunchKernelGGL(k_cell_sort, dim3(capped(c->ncell_total, 1)), dim3(256), 0, c->stream,
                       c->ncell_total, c->cell_rep, c->rebuild, any_flag(c), c->ccount, c->cstart,
                       c->sorted_ids, c->perm, c->tt, c->atombin);
}

void launch_gather(sofia_ctx *c, bool engine) {
    hipLaunchKernelGGL(k_gather, dim3(capped(c->N, 256)), dim3(256), 0, c->stream, c->N, c->rep,
                       c->rebuild, any_flag(c), c->perm, c->tt, c->a, c->b, engine ? 1 : 0);
    hipLaunchKernelGGL(k_copyback, dim3(capped(c->N, 256)), dim3(256), 0, c->stream, c->N, c->rep,
                       c->rebuild, any_flag(c), c->b, c->a, c->dr32, c->drd, engine ? 1 : 0);
}

void launch_remap_topology(sofia_ctx *c) {
    RemapSet S{};
    int total = 0;
    auto add = [&](int32_t *lst, int items, int width, int nidx) {
        if (!items) return;
        S.lst[S.nlists] = lst;
        total += items;
        S.end[S.nlists] = total;
        S.width[S.nlists] = width;
        S.nidx[S.nlists] = nidx;
        ++S.nlists;
    };
    add(c->bondlist, c->nb, 3, 2);
    add(c->anglelist, c->na, 4, 3);
    add(c->dihedlist, c->nd, 5, 4);
    add(c->torsions, c->ntor, 5, 4);
    add(c->improlist, c->nimp, 5, 4);
    add(c->shake_atoms, c->ns, 4, 4);
    add(c->ebond, c->nbe, 3, 2);
    add(c->eangle, c->nae, 4, 3);
    if (total)
        hipLaunchKernelGGL(k_remap_all, dim3(capped(total, 256)), dim3(256), 0, c->stream, S, c->tt,
                           c->rep, c->rebuild, any_flag(c));
}

void launch_rdf(sofia_ctx *c, int32_t itype, int32_t jtype, double delta, int32_t nbins,
                uint32_t *hist) {
    hipLaunchKernelGGL(k_rdf, dim3((c->N + 3) / 4), dim3(256), 0, c->stream, c->N, c->a.pos4,
                       c->rep, c->reps_d, c->nlist, c->numneigh, c->neigh_alloc, c->dp.rcut2,
                       delta, itype, jtype, nbins, hist);
}

void launch_host_ids(sofia_ctx *c, const int32_t *src, int32_t *dst, size_t items, int width,
                     int nidx) {
    const size_t count = items * (size_t)width;
    if (!count) return;
    const int grid = (int)std::min<size_t>((count + 255) / 256, GRID_CAP);
    hipLaunchKernelGGL(k_host_ids, dim3(grid), dim3(256), 0, c->stream, count, width, nidx, src,
                       dst, c->rep, c->reps_d);
}

void launch_build_list(sofia_ctx *c) {
    if (c->n_stencil5 > 0) launch_build_list_engine(c);
    if (c->n_stencil5 == c->nrep) return;
    const int w = c->list_waves;
#define LB(W)                                                                                    \
    hipLaunchKernelGGL(k_build_list<W>, dim3(c->ncell_total), dim3(W * 64), 0, c->stream,       \
                       c->a.pos4, c->a.num3, c->a.spec, c->cstart, c->ccount, c->cnum,            \
                       c->cell_rep, c->reps_d, c->rebuild, c->dp.rv2, c->neigh_alloc, c->nlist,   \
                       c->numneigh, c->maxcnt)
    if (w <= 2) LB(2);
    else if (w <= 4) LB(4);
    else if (w <= 8) LB(8);
    else LB(16);
#undef LB
}

void launch_check_neighbor(sofia_ctx *c, bool engine) {
    if (engine)
        hipLaunchKernelGGL(k_check64, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->N, c->rep,
                           c->drd, c->reps_d, c->rebuild, c->rebuild + c->nrep);
    else
        hipLaunchKernelGGL(k_check32, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->N, c->rep,
                           c->dr32, c->dp.rvrc2, c->rebuild, c->rebuild + c->nrep);
}

void launch_set_rebuild(sofia_ctx *c, int32_t value) {
    hipLaunchKernelGGL(k_set_flags, dim3(nblk(c->nrep + 1, 256)), dim3(256), 0, c->stream, c->nrep,
                       c->rebuild, value);
}

void launch_finish_rebuild(sofia_ctx *c, bool) {
    hipLaunchKernelGGL(k_finish, dim3(nblk(c->nrep + 1, 256)), dim3(256), 0, c->stream, c->nrep,
                       c->rebuild, c->nbuilds);
}

}  // namespace sofia
