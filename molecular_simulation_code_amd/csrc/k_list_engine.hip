// Verlet-list build of the device-resident engine, on half-width cells.
//
// The engine's lists only have to hold every pair closer than rv (order and exact
// boundary rounding are free, unlike the drop-in build in k_neighbor.hip).  That allows a
// cheaper search:
//  * cells are at least rv/2 wide, so a pair closer than rv lies within +-2 cells on each
//    axis: 125 stencil cells, 15.6 rv^3 of candidate volume instead of 27 rv^3;
//  * one workgroup owns one cell and stages its stencil in LDS in (dz, dy, dx) order with
//    the periodic image shift already added, so a candidate costs 3 subtractions and a
//    squared norm (no minimum image) and the 5 cells of a (dy, dz) row are contiguous;
//  * for each atom, 25 lanes bound the 25 rows at once: a row whose y-z rectangle is
//    farther than rv is skipped, and of the others only the x-cells meeting the chord
//    [x - w, x + w], w = sqrt(rv^2 - d_yz^2), are swept -- about 45 % of the stencil;
//  * sweeps go 64 candidates at a time, hits compacted with a ballot + popcount prefix.
// Culling distances carry a 0.01 A margin, far above the float rounding of positions, cell
// indices and image shifts, so culling never drops a pair the distance test would keep.
// A stencil larger than the LDS capacity is staged in pieces; per-atom counts then carry
// over in numneigh (each atom stays with the same wavefront).
#include <algorithm>
#include <cmath>

#include "sofia_device.h"

namespace sofia {
namespace {

constexpr int ECAP = 4096;            // staged candidates per pass (64 KiB of LDS)
constexpr float CULL_MARGIN = 0.01f;  // A

// Two workgroups fit a CU's LDS; the second bound (waves per SIMD) keeps the register
// budget that lets both run (64 VGPRs at 16 waves per workgroup).
template <int WAVES>
__global__ void __launch_bounds__(WAVES * 64, WAVES * 2 / 4)
k_build_engine(int32_t ncell, const float4 *__restrict__ pos4, const int32_t *__restrict__ num3,
               const int32_t *__restrict__ spec, const int32_t *__restrict__ cstart,
               const int32_t *__restrict__ ccount, const int32_t *__restrict__ cell_rep,
               const RepInfo *__restrict__ reps, const int32_t *__restrict__ rebuild,
               const int32_t *__restrict__ any, float rv, float rv2, int32_t na,
               int32_t *__restrict__ nlist, int32_t *__restrict__ numneigh,
               int32_t *__restrict__ maxcnt) {
    if (!*any) return;   // no replica rebuilds this step
    __shared__ int32_t s_src[128];
    __shared__ int32_t s_cnt[128];
    __shared__ int32_t s_off[129];
    __shared__ float4 s_j[ECAP];

    // a fixed grid walks the cells; slot s and s + gridDim.x stay on one XCD, and
    // xcd_block gives each XCD runs of neighbouring cells
    for (int slot = blockIdx.x; slot < ncell; slot += gridDim.x) {
        const int c = xcd_block(slot, ncell);
        const int r = cell_rep[c];
        if (!rebuild[r]) continue;
        const RepInfo R = reps[r];
        if (!R.stencil5 || ccount[c] == 0) continue;
        const int n = R.ncellD, lc = c - R.cell0;
        const int cx = lc % n, cy = (lc / n) % n, cz = lc / (n * n);
        const float box = R.box, rn = R.rn;

        const int tid = threadIdx.x;
        if (tid < 128) {
            int cnt = 0, src = 0;
            if (tid < 125) {
                int x = cx + tid % 5 - 2, y = cy + (tid / 5) % 5 - 2, z = cz + tid / 25 - 2;
                x += (x < 0) ? n : ((x >= n) ? -n : 0);     // n >= 5: one wrap at most
                y += (y < 0) ? n : ((y >= n) ? -n : 0);
                z += (z < 0) ? n : ((z >= n) ? -n : 0);
                const int nc = R.cell0 + x + (y + z * n) * n;
                src = cstart[nc];
                cnt = ccount[nc];
            }
            s_src[tid] = src;
            s_cnt[tid] = cnt;
        }
        __syncthreads();
        if (tid < 64) {   // exclusive scan of the 128 counts, two per lane
            const int a = s_cnt[2 * tid], b = s_cnt[2 * tid + 1];
            int v = a + b;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o, 64);
                if (tid >= o) v += u;
            }
            const int ex = v - a - b;
            s_off[2 * tid] = ex;
            s_off[2 * tid + 1] = ex + a;
            if (tid == 63) s_off[128] = v;
        }
        __syncthreads();
        const int total = s_off[125];

        // the wave index is uniform: the atoms a wave takes, and their exclusions, stay scalar
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
        const int i0 = cstart[c], ni = ccount[c];
        const float rvm = rv + CULL_MARGIN, rvm2 = rvm * rvm;
        const float irn = 1.0f / rn;

        for (int g0 = 0; g0 < total || g0 == 0; g0 += ECAP) {
            const int g1 = min(total, g0 + ECAP);
            if (g0 > 0) __syncthreads();   // the previous pass is done with s_j
            // a wave stages whole (dy, dz) rows: the row's 5 cells are found with 4 compares
            // against its (uniform) cell offsets, and the y / z image shifts are uniform
            for (int row = wave; row < 25; row += WAVES) {
                const int k0 = 5 * row;
                const int r0 = max(s_off[k0], g0), r1 = min(s_off[k0 + 5], g1);
                if (r0 >= r1) continue;
                const int ay = cy + row % 5 - 2, az = cz + row / 5 - 2;
                const float sy = (ay < 0) ? -box : ((ay >= n) ? box : 0.0f);
                const float sz = (az < 0) ? -box : ((az >= n) ? box : 0.0f);
                int b[4], d[5];
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    if (q > 0) b[q - 1] = s_off[k0 + q];
                    d[q] = s_src[k0 + q] - s_off[k0 + q];   // staged index -> atom id
                }
                for (int f = r0 + lane; f < r1; f += 64) {
                    const int dk = (int)(f >= b[0]) + (int)(f >= b[1]) + (int)(f >= b[2]) + (int)(f >= b[3]);
                    const int id = f + (dk == 0 ? d[0] : dk == 1 ? d[1] : dk == 2 ? d[2] : dk == 3 ? d[3] : d[4]);
                    const int ax = cx + dk - 2;
                    float4 p = pos4[id];
                    p.x += (ax < 0) ? -box : ((ax >= n) ? box : 0.0f);
                    p.y += sy;
                    p.z += sz;
                    p.w = __int_as_float(id);
                    s_j[f - g0] = p;
                }
            }
            __syncthreads();

            for (int ii = wave; ii < ni; ii += WAVES) {
                const int i = i0 + ii;
                const float4 xi = pos4[i];
                const int n3 = num3[i];
                // the atom's exclusions (wave-uniform: scalar loads, held in scalar registers)
                int ex[SPEC];
                int exmin = 0x7fffffff, exmax = -1;
#pragma unroll
                for (int l = 0; l < SPEC; ++l) {
                    const int v = spec[(size_t)i * SPEC + l];
                    ex[l] = (l < n3) ? v : -1;
                    exmin = (l < n3) ? min(exmin, v) : exmin;
                    exmax = (l < n3) ? max(exmax, v) : exmax;
                }
                // row bounds: lane k < 25 takes row (dy, dz) = (k % 5 - 2, k / 5 - 2)
                int rlo = 0, rhi = 0;
                if (lane < 25) {
                    const float ylo = (float)(cy + lane % 5 - 2) * rn, zlo = (float)(cz + lane / 5 - 2) * rn;
                    const float ey = fmaxf(0.0f, fmaxf(ylo - xi.y, xi.y - (ylo + rn)) - CULL_MARGIN);
                    const float ez = fmaxf(0.0f, fmaxf(zlo - xi.z, xi.z - (zlo + rn)) - CULL_MARGIN);
                    const float dyz = ey * ey + ez * ez;
                    if (dyz < rvm2) {
                        const float w = sqrtf(rvm2 - dyz);
                        const int a = min(2, max(-2, (int)floorf((xi.x - w) * irn) - cx));
                        const int b = min(2, max(-2, (int)floorf((xi.x + w) * irn) - cx));
                        rlo = max(s_off[lane * 5 + a + 2], g0);
                        rhi = min(s_off[lane * 5 + b + 3], g1);
                    }
                }
                uint64_t rows = __ballot(rhi > rlo);
                int cnt = (g0 == 0) ? 0 : numneigh[i];
                int32_t *out = nlist + (size_t)i * na;
                // test one staged candidate per lane; hits appended in staged order.  The tests
                // combine without short-circuit (no branch per condition); the exclusion
                // compares run only for a chunk where some lane falls in [exmin, exmax]
                // (measured cheaper than compares bounded by the atom's exclusion count).
                auto sweep = [&](const float4 pj, const bool valid) {
                    const int z = __float_as_int(pj.w);
                    const float dx = xi.x - pj.x, dy = xi.y - pj.y, dz = xi.z - pj.z;
                    bool pred = valid & (dx * dx + dy * dy + dz * dz < rv2) & (z != i);
                    if (__ballot(pred & (z >= exmin) & (z <= exmax))) {
#pragma unroll
                        for (int l = 0; l < SPEC; ++l) pred = pred & (ex[l] != z);
                    }
                    const uint64_t mask = __ballot(pred);
                    const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi(
                                              (uint32_t)(mask >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                    if (pred & (pos < na)) out[pos] = z;
                    cnt += __popcll(mask);
                };
                while (rows) {
                    const int k = __builtin_ctzll(rows);
                    rows &= rows - 1;
                    const int lo = __builtin_amdgcn_readlane(rlo, k) - g0;
                    const int hi = __builtin_amdgcn_readlane(rhi, k) - g0;
                    // two chunks per iteration, both LDS reads issued before either is used;
                    // indices past the row re-read its last slot and are masked off
                    for (int base = lo; base < hi; base += 128) {
                        const int t0 = base + lane, t1 = base + 64 + lane;
                        const float4 p0 = s_j[min(t0, hi - 1)];
                        const float4 p1 = s_j[min(t1, hi - 1)];
                        sweep(p0, t0 < hi);
                        if (base + 64 < hi) sweep(p1, t1 < hi);
                    }
                }
                if (lane == 0) {
                    numneigh[i] = cnt;
                    if (g1 >= total && cnt > na) atomicMax(maxcnt, cnt);
                }
            }
            if (g1 >= total) break;
        }
        __syncthreads();   // the shared arrays are refilled for the next cell
    }
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

}  // namespace

void launch_build_list_engine(sofia_ctx *c) {
    const float rv2 = (float)c->dp.rv2;
    const float rv = std::sqrt(rv2);
    // at most 4 workgroups per CU fit (64 KiB of staged candidates each); a multiple of 8
    const int grid = std::min(c->ncell_total, 2048);
#define LBE(W)                                                                                   \
    hipLaunchKernelGGL(k_build_engine<W>, dim3(grid), dim3(W * 64), 0, c->stream,                \
                       c->ncell_total, c->a.pos4, c->a.num3, c->a.spec, c->cstart, c->ccount,    \
                       c->cell_rep, c->reps_d, c->rebuild, c->rebuild + c->nrep, rv, rv2,        \
                       c->neigh_alloc, c->nlist, c->numneigh, c->maxcnt)
    if (c->list_waves >= 16) LBE(16);
    else if (c->list_waves >= 8) LBE(8);
    else LBE(4);
#undef LBE
}

}  // namespace sofia
