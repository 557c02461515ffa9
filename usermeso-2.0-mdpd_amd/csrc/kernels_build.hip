// Neighbor-rebuild phase of the mdpd355 engine: periodic wrap, cell keys,
// counting sort by cell + per-particle rank of the sub-bin key inside its
// cell, gather into the other buffer, per-particle neighbor-row build, commit.
//
// Replaces, for one or many periodic boxes at once:
//   MesoAtom::sort_local            src/USER-MESO/atom_meso.cu:271-387
//   MesoNeighbor::binning_meso      src/USER-MESO/neighbor_meso.cu:390-683
//   MesoNeighbor::full_bin_meso     src/USER-MESO/neigh_build_meso.cu:17-127,286-449
//   Domain::pbc + MesoComm::borders (on-device periodic images instead of ghosts)
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace mdpd {

#pragma clang fp contract(off)

// --------------------------------------------------------------------------
// K1: wrap into the box (Domain::pbc in FP32), reference sort key, cell slot,
// arrival rank inside the cell.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_keys(PoolDev P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    if (i >= d->part_base + d->n) return;

    const int cur = d->sel, cura = d->sela;
    float4 p = P.pos[cur][i];
    int img = P.image[cura][i];
    float c[3] = {p.x, p.y, p.z};
    int im[3] = {(img & 1023) - 512, ((img >> 10) & 1023) - 512, ((img >> 20) & 1023) - 512};
    for (int a = 0; a < 3; ++a) {
        const float flo = d->lo[a], fhi = d->hi[a], prd = d->L[a];
        float v = c[a];
        if (v < flo) { v = v + prd; im[a] -= 1; }
        if (v >= fhi) { v = v - prd; v = fmaxf(v, flo); im[a] += 1; }
        c[a] = v;
    }
    P.pos[cur][i] = make_float4(c[0], c[1], c[2], p.w);
    P.image[cura][i] = ((im[0] + 512) & 1023) | (((im[1] + 512) & 1023) << 10) | (((im[2] + 512) & 1023) << 20);

    int b[3];
    uint32_t sub[3];
    for (int a = 0; a < 3; ++a) {
        // bin: (int)((x - lo) * bininv + 1.0), FP64, clamped to the interior
        const double t = ((double)c[a] - d->lo_d[a]) * d->bininv_d[a];
        const double u = t + 1.0;
        int bb = (int)u;
        bb = bb > d->mbin[a] - 2 ? d->mbin[a] - 2 : bb;
        bb = bb < 1 ? 1 : bb;
        b[a] = bb;
        // sub-bin: (x - (bin-1) * binsize) * 16 * bininv (atom_meso.cu:298-300)
        const double off = (double)(uint32_t)(bb - 1) * d->binsize_d[a];
        const double rel = (double)c[a] - off;
        const double sc = rel * (16.0 * d->bininv_d[a]);
        int ss = (int)sc;
        ss = ss > 15 ? 15 : ss;
        ss = ss < 0 ? 0 : ss;
        sub[a] = (uint32_t)ss;
    }
    const uint32_t z2 = (uint32_t)morton3(sub[0], sub[1], sub[2]);
    if (P.key) {
        const uint64_t z1 = morton3((uint32_t)b[0], (uint32_t)b[1], (uint32_t)b[2]);
        P.key[i] = (z1 << d->l1_shift) | z2;
    }
    const int lin = (b[0] - 1) + d->nb[0] * ((b[1] - 1) + d->nb[1] * (b[2] - 1));
    const int gc = d->cell_base + P.cell_rank[d->cell_base + lin];
    P.pcell[i] = gc;
    P.psub[i] = (int)z2;
    P.pslot[i] = atomicAdd(P.cell_count + gc, 1);
}

// --------------------------------------------------------------------------
// K2: exclusive scan of cell counts in Morton-rank order, one workgroup per
// system; clears the counts for the next rebuild.
// --------------------------------------------------------------------------
__device__ inline int wave_incl_scan(int x, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(1024) k_scan(PoolDev P)
{
    const int s = blockIdx.x;
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    // chunks of 8192 counts: coalesced load into LDS, 8 consecutive counts per
    // thread, wave scan + scan of the 16 wave totals, coalesced store
    constexpr int PER = 8, CH = 1024 * PER;
    __shared__ int buf[CH];
    __shared__ int wex[16];
    __shared__ int carry;
    const int nc = d->ncell, base = d->cell_base, pb = d->part_base;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = pb;
    __syncthreads();
    for (int c0 = 0; c0 < nc; c0 += CH) {
        for (int k = t; k < CH; k += 1024) {
            const int c = c0 + k;
            int v = 0;
            if (c < nc) { v = P.cell_count[base + c]; P.cell_count[base + c] = 0; }
            buf[k] = v;
        }
        __syncthreads();
        int v[PER], sum = 0;
        for (int k = 0; k < PER; ++k) { v[k] = buf[t * PER + k]; sum += v[k]; }
        const int incl = wave_incl_scan(sum, lane);
        if (lane == 63) wex[w] = incl;
        __syncthreads();
        if (t == 0) {
            int acc = 0;
            for (int k = 0; k < 16; ++k) { const int y = wex[k]; wex[k] = acc; acc += y; }
        }
        __syncthreads();
        int run = carry + wex[w] + incl - sum;
        for (int k = 0; k < PER; ++k) { buf[t * PER + k] = run; run += v[k]; }
        __syncthreads();
        for (int k = t; k < CH; k += 1024) {
            const int c = c0 + k;
            if (c < nc) P.cell_start[base + c] = buf[k];
        }
        if (t == 1023) carry = run;
        __syncthreads();
    }
    if (t == 0) P.cell_start[base + nc] = pb + d->n;
}

// --------------------------------------------------------------------------
// K3: scatter particle indices into their cell segments (arrival order).
// --------------------------------------------------------------------------
// K3: scatter particle indices into their cell segments (arrival order), the
// sub-bin key beside each (cl is free once k_scan has used it).
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_scatter(PoolDev P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    if (i >= d->part_base + d->n) return;
    const int f = P.cell_start[P.pcell[i]] + P.pslot[i];
    P.tmp[f] = i;
    P.cl[f] = P.psub[i];
}

// --------------------------------------------------------------------------
// K4: one thread per particle, by its arrival slot f in its cell's segment:
// its rank in the cell by (sub-bin key, previous index) -- the stable sort of
// the 2-level Morton key -- gives its sorted slot; every per-particle array
// moves there in the other buffer, and the sorted slot records its cell (in
// pslot, free after k_scatter) for the neighbor-row build.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sortgather(PoolDev P)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nact) return;
    const int s = P.wave_sys[f >> 6];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    if (f >= d->part_base + d->n) return;
    const int idx = P.tmp[f];
    const int key = P.cl[f];
    const int cell = P.pcell[idx];
    const int start = P.cell_start[cell], end = P.cell_start[cell + 1];
    int rank = 0;
    for (int g = start; g < end; ++g) {
        const int kg = P.cl[g], jg = P.tmp[g];
        rank += (kg < key) || (kg == key && jg < idx);
    }
    const int dst = start + rank;
    const int cur = d->sel, cura = d->sela;
    const float4 p = P.pos[cur][idx];
    float4 v = P.vel[cur][idx];
    const int tg = P.tag[cura][idx];
    if (d->phase == PH_FRESH) {
        // a newly loaded system: signatures of its first force pass
        const uint32_t sg = d->style == STYLE_MDPD ? sig_mdpd(tg, d->seed_cur)
                                                   : sig_dpd(tg, d->seed_cur, v.x, v.y, v.z);
        v.w = __uint_as_float(sg);
    }
    P.pos[cur ^ 1][dst] = p;
    P.vel[cur ^ 1][dst] = v;
    P.xhold[dst] = p;
    P.tag[cura ^ 1][dst] = tg;
    P.mask[cura ^ 1][dst] = P.mask[cura][idx];
    P.origin[cura ^ 1][dst] = P.origin[cura][idx];
    P.image[cura ^ 1][dst] = P.image[cura][idx];
    P.phi[cura ^ 1][dst] = P.phi[cura][idx];
    P.pslot[dst] = cell;
    if (P.key) P.key[(size_t)P.cap + dst] = P.key[idx];
}

// --------------------------------------------------------------------------
// K5: one lane per centre particle (sorted order): its row lists every
// (j, image) of the 27 stencil cells of its cell -- stencil cell k has offset
// (k%3-1, k/3%3-1, k/9-1), periodic images shift by +-L -- in stencil order,
// then j ascending, with (float) r^2 <= (float)(rc+skin)^2 evaluated without
// contraction (neigh_build_meso.cu:86-93), itself excluded.  The set is the
// reference's (neigh_build_meso.cu:58-118); each lane appends to its own row:
// entry p of particle i at tile*maxnbr*64 + p*64 + i%64.
// relist: rows of the current, already sorted buffer (the row width grew);
// otherwise of the data k_sortgather just wrote.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_nlist(PoolDev P, int relist)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active || (relist ? d->stale != 0 : d->stale == 0)) return;
    if (i >= d->part_base + d->n) return;
    const float4 *pos = P.pos[relist ? d->sel : d->sel ^ 1];
    const int xyz = P.cell_xyz[P.pslot[i]];
    const int cx = xyz & 1023, cy = (xyz >> 10) & 1023, cz = (xyz >> 20) & 1023;
    const int nb0 = d->nb[0], nb1 = d->nb[1], nb2 = d->nb[2], cb = d->cell_base;
    const float4 pi = pos[i];
    const float rc2 = d->rc2_list;
    const int maxnbr = P.maxnbr;
    int *row = P.nlist + (size_t)(i >> 6) * maxnbr * WAVE + (i & 63);
    int cnt = 0;
    auto test = [&](int j, const float4 &q, float shx, float shy, float shz, uint32_t code) {
        const float xj = q.x + shx, yj = q.y + shy, zj = q.z + shz;
        const float dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        if (r2 <= rc2 && !(j == i && code == (uint32_t)IMG_ZERO)) {
            if (cnt < maxnbr) row[(size_t)cnt * WAVE] = (int)((uint32_t)j | (code << IMG_SHIFT));
            ++cnt;
        }
    };
    // one plane of the stencil (9 cells, fixed dz) at a time: the 9 rank
    // look-ups, then the 18 offset look-ups, are issued together
    for (int kz = 0; kz < 3; ++kz) {
        int nz = cz + kz - 1, sz = 0;
        if (nz < 0) { nz += nb2; sz = -1; } else if (nz >= nb2) { nz -= nb2; sz = 1; }
        int ns[9], sxa[9], sya[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            int nx = cx + q % 3 - 1, ny = cy + q / 3 - 1, sx = 0, sy = 0;
            if (nx < 0) { nx += nb0; sx = -1; } else if (nx >= nb0) { nx -= nb0; sx = 1; }
            if (ny < 0) { ny += nb1; sy = -1; } else if (ny >= nb1) { ny -= nb1; sy = 1; }
            sxa[q] = sx;
            sya[q] = sy;
            ns[q] = P.cell_rank[cb + nx + nb0 * (ny + nb1 * nz)];
        }
        int st[9], en[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            st[q] = P.cell_start[cb + ns[q]];
            en[q] = P.cell_start[cb + ns[q] + 1];
        }
        const float shz = (float)sz * d->L[2];
        for (int q = 0; q < 9; ++q) {
            const float shx = (float)sxa[q] * d->L[0], shy = (float)sya[q] * d->L[1];
            const uint32_t code = (uint32_t)((sxa[q] + 1) + 3 * (sya[q] + 1) + 9 * (sz + 1));
            // four candidates' positions in flight, tested in order
            int j = st[q];
            const int e = en[q];
            for (; j + 4 <= e; j += 4) {
                const float4 q0 = pos[j], q1 = pos[j + 1], q2 = pos[j + 2], q3 = pos[j + 3];
                test(j, q0, shx, shy, shz, code);
                test(j + 1, q1, shx, shy, shz, code);
                test(j + 2, q2, shx, shy, shz, code);
                test(j + 3, q3, shx, shy, shz, code);
            }
            for (; j < e; ++j) test(j, pos[j], shx, shy, shz, code);
        }
    }
    P.ncount[i] = cnt;
    if (cnt > maxnbr) atomicOr(&d->overflow, 1);
}

// overflow flags of the sorted systems cleared before their rows are rebuilt
// in place (k_nlist sets them again if the rows are still too narrow)
__global__ void k_clear_overflow(PoolDev P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.nsys) return;
    SysDev *d = P.sys + s;
    if (d->active && !d->stale) d->overflow = 0;
}

// --------------------------------------------------------------------------
// Commit after a rebuild: the sorted data in the other buffers becomes current.
// --------------------------------------------------------------------------
__global__ void k_commit_rebuild(PoolDev P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.nsys) return;
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    d->sel ^= 1;
    d->sela ^= 1;
    d->stale = 0;
    d->ago = 0;
}

// --------------------------------------------------------------------------
// Signatures of the current step (setup only; in the step loop the fused
// force epilogue writes them for the following step).
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_signatures(PoolDev P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active || i >= d->part_base + d->n) return;
    const uint32_t st = premix<64>(d->seed, (uint32_t)d->step);
    float4 v = P.vel[d->sel][i];
    const int tag = P.tag[d->sela][i];
    const uint32_t sg = d->style == STYLE_MDPD ? sig_mdpd(tag, st) : sig_dpd(tag, st, v.x, v.y, v.z);
    v.w = __uint_as_float(sg);
    P.vel[d->sel][i] = v;
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
static inline int nblk(int n, int b) { return (n + b - 1) / b; }
hipError_t launch_rebuild(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    const dim3 g(nblk(P.nact, 256)), b(256);
    hipLaunchKernelGGL(k_keys, g, b, 0, st, P);
    hipLaunchKernelGGL(k_scan, dim3(P.nsys), dim3(1024), 0, st, P);
    hipLaunchKernelGGL(k_scatter, g, b, 0, st, P);
    hipLaunchKernelGGL(k_sortgather, g, b, 0, st, P);
    hipLaunchKernelGGL(k_nlist, g, b, 0, st, P, 0);
    return hipGetLastError();
}

hipError_t launch_nlist_only(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nlist, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P, 0);
    return hipGetLastError();
}

hipError_t launch_relist(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_overflow, dim3(nblk(P.nsys, 64)), dim3(64), 0, st, P);
    hipLaunchKernelGGL(k_nlist, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P, 1);
    return hipGetLastError();
}

hipError_t launch_commit_rebuild(const PoolDev &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_commit_rebuild, dim3(nblk(P.nsys, 64)), dim3(64), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_signatures(const PoolDev &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_signatures, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P);
    return hipGetLastError();
}

}  // namespace mdpd
