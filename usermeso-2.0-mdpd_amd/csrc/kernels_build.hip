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

// The neighbor-row build of a pool walks a worklist of the cells to list
// (k_scan_apply appends them); one system (the plain path) keeps one wave
// per cell slot: its slots are few and the list would only add its own cost.
__host__ __device__ inline bool nlist_listed(const PoolDev &P) { return P.nsys > 1; }

// --------------------------------------------------------------------------
// K1: wrap into the box (Domain::pbc in FP32), reference sort key, sub-cell
// slot (cell slot * SUBCELLS + top bits of the sub-bin key), arrival rank
// inside the sub-cell.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_keys(PoolDev P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && P.work_count) *P.work_count = 0;        // this rebuild's worklist starts empty
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    // the wave stays whole (the slot's run below is found across lanes);
    // lanes past the system's last particle only take part in that
    const bool valid = i < d->part_base + d->n;
    int gc = -1;
    if (valid) {
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
    gc = (d->cell_base + P.cell_rank[d->cell_base + lin]) * SUBCELLS + (int)(z2 >> SUBCELL_SHIFT);
    P.pcell[i] = gc;
    P.psub[i] = (int)z2;
    }
    // arrival slots: the particles are still nearly in the last build's
    // order, so neighboring lanes mostly share a sub-cell; each run of equal
    // sub-cells takes its slots with one atomic (the order of arrival inside
    // a sub-cell does not matter: the rank in k_sortgather decides)
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(gc, 1, 64);
    const bool head = valid && (lane == 0 || prev != gc);
    const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
    const uint64_t below = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);    // lanes 0..lane
    const uint64_t hb = heads & below;
    const int h = hb ? 63 - __clzll(hb) : 0;                                // this lane's run head
    const uint64_t later = heads & ~below;
    const int nvalid = __popcll(__builtin_amdgcn_ballot_w64(valid));
    const int next = later ? min(__ffsll((long long)later) - 1, nvalid) : nvalid;
    int base = 0;
    if (head) base = atomicAdd(P.cell_count + gc, next - lane);
    base = __shfl(base, h, 64);
    if (valid) P.pslot[i] = base + (lane - h);
}

// --------------------------------------------------------------------------
// K2: exclusive scan of sub-cell counts in (cell Morton rank, sub-cell)
// order, per system, from its first particle slot; clears the counts for the
// next rebuild.  Chunks of SCAN_CH counts, three launches without waits:
//   k_scan_sums   per chunk: its total, parked in the chunk's first start slot
//   k_scan_chunks per system: exclusive scan of the chunk totals in place,
//                 and the end entry
//   k_scan_apply  per chunk: exclusive scan of its counts from its offset;
//                 in a pool also each cell's particle range by linear cell
//                 id and the worklist of occupied cells for the row build
// Grid (G, nsys): workgroup g of a system takes chunks g, g + G, ...
// --------------------------------------------------------------------------
constexpr int SCAN_BLK = 256, SCAN_PER = 16, SCAN_CH = SCAN_BLK * SCAN_PER;

// exclusive scan of one value per thread over the workgroup; returns the
// workgroup total through *total
__device__ inline int block_excl_scan(int x, int *wsum, int *total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int incl = wave_incl_scan(x, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int k = 0; k < SCAN_BLK / 64; ++k) {
        const int y = wsum[k];
        before += k < w ? y : 0;
        all += y;
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

__global__ void __launch_bounds__(SCAN_BLK) k_scan_sums(PoolDev P)
{
    const SysDev *d = P.sys + blockIdx.y;
    if (!d->active || !d->stale) return;
    __shared__ int wsum[SCAN_BLK / 64];
    const int nc = d->ncell * SUBCELLS, base = d->cell_base * SUBCELLS;
    for (int c0 = blockIdx.x * SCAN_CH; c0 < nc; c0 += gridDim.x * SCAN_CH) {
        // clamped loads, all in flight; entries past the end count 0
        int sum = 0;
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            const int c = c0 + u * SCAN_BLK + (int)threadIdx.x;
            const int v = P.cell_count[base + min(c, nc - 1)];
            sum += c < nc ? v : 0;
        }
        int total;
        block_excl_scan(sum, wsum, &total);
        if (threadIdx.x == 0) P.cell_start[base + c0] = total;
    }
}

__global__ void __launch_bounds__(SCAN_BLK) k_scan_chunks(PoolDev P)
{
    const SysDev *d = P.sys + blockIdx.x;
    if (!d->active || !d->stale) return;
    __shared__ int wsum[SCAN_BLK / 64];
    const int nc = d->ncell * SUBCELLS, base = d->cell_base * SUBCELLS;
    const int nch = (nc + SCAN_CH - 1) / SCAN_CH;
    int carry = d->part_base;
    for (int k0 = 0; k0 < nch; k0 += SCAN_BLK) {
        const int k = k0 + threadIdx.x;
        const int v = k < nch ? P.cell_start[base + k * SCAN_CH] : 0;
        int total;
        const int ex = block_excl_scan(v, wsum, &total);
        if (k < nch) P.cell_start[base + k * SCAN_CH] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) P.cell_start[base + nc] = d->part_base + d->n;
}

__global__ void __launch_bounds__(SCAN_BLK) k_scan_apply(PoolDev P)
{
    const SysDev *d = P.sys + blockIdx.y;
    if (!d->active || !d->stale) return;
    __shared__ int buf[SCAN_CH];
    __shared__ int wsum[SCAN_BLK / 64];
    __shared__ int s_wl[SCAN_BLK / 64 + 1];                // worklist: wave counts, the chunk's base
    const int nc = d->ncell * SUBCELLS, base = d->cell_base * SUBCELLS;
    const int t = threadIdx.x;
    for (int c0 = blockIdx.x * SCAN_CH; c0 < nc; c0 += gridDim.x * SCAN_CH) {
        const int off = P.cell_start[base + c0];     // the chunk's offset
        // coalesced load into LDS (counts cleared), SCAN_PER consecutive
        // counts per thread, workgroup scan, coalesced store
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            const int c = c0 + u * SCAN_BLK + t;
            const int v = P.cell_count[base + min(c, nc - 1)];
            buf[u * SCAN_BLK + t] = c < nc ? v : 0;
        }
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            const int c = c0 + u * SCAN_BLK + t;
            if (c < nc) P.cell_count[base + c] = 0;
        }
        __syncthreads();
        int v[SCAN_PER], sum = 0;
        for (int k = 0; k < SCAN_PER; ++k) { v[k] = buf[t * SCAN_PER + k]; sum += v[k]; }
        static_assert(SCAN_PER % SUBCELLS == 0 && SCAN_CH % SUBCELLS == 0, "whole cells per thread");
        constexpr int CPT = SCAN_PER / SUBCELLS;                 // cells per thread
        int cst[CPT];                                            // first particle of each
        int total;
        int run = off + block_excl_scan(sum, wsum, &total);
#pragma unroll
        for (int k = 0; k < SCAN_PER; ++k) {
            if (k % SUBCELLS == 0) cst[k / SUBCELLS] = run;
            buf[t * SCAN_PER + k] = run;
            run += v[k];
        }
        // pools: every cell's particle range by linear cell id, and the cells
        // holding particles appended to the neighbor-row worklist with their
        // system, coordinates and range, in cell-slot order within the chunk
        // (one atomic per workgroup and chunk: same-address atomics
        // serialize); counts past the system's end are zero and not listed
        if (nlist_listed(P)) {
            const int cell0 = d->cell_base + (c0 + t * SCAN_PER) / SUBCELLS;
            int xyz[CPT], full = 0;
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                const bool in = c0 + t * SCAN_PER + q * SUBCELLS < nc;
                int m = 0;
#pragma unroll
                for (int u = 0; u < SUBCELLS; ++u) m += v[q * SUBCELLS + u];
                xyz[q] = in ? P.cell_xyz[cell0 + q] : 0;
                if (in) {
                    const int lin = (xyz[q] & 1023) + d->nb[0] * (((xyz[q] >> 10) & 1023) + d->nb[1] * ((xyz[q] >> 20) & 1023));
                    P.cell_lse[d->cell_base + lin] = make_int2(cst[q], cst[q] + m);
                }
                full |= (m > 0 ? 1 : 0) << q;
            }
            const int cnt = __popc(full);
            const int lane = t & 63, wv = t >> 6;
            const int incl = wave_incl_scan(cnt, lane);
            if (lane == 63) s_wl[wv] = incl;
            __syncthreads();
            if (t == 0) {
                int tot = 0;
                for (int k = 0; k < SCAN_BLK / 64; ++k) tot += s_wl[k];
                s_wl[SCAN_BLK / 64] = tot > 0 ? atomicAdd(P.work_count, tot) : 0;
            }
            __syncthreads();
            int o = s_wl[SCAN_BLK / 64] + incl - cnt;
            for (int k = 0; k < wv; ++k) o += s_wl[k];
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                if (full >> q & 1) {
                    int m = 0;
#pragma unroll
                    for (int u = 0; u < SUBCELLS; ++u) m += v[q * SUBCELLS + u];
                    P.cell_work[o++] = make_int4((int)blockIdx.y, xyz[q], cst[q], cst[q] + m);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SCAN_PER; ++u) {
            const int k = u * SCAN_BLK + t;
            if (c0 + k < nc) P.cell_start[base + c0 + k] = buf[k];
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// K3: scatter particle indices into their sub-cell segments (arrival order),
// the sub-bin key beside each (cl is free once k_scan has used it).
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
// K4: one thread per particle, by its previous slot i: its rank in its
// sub-cell by (sub-bin key, previous index) -- with the sub-cells in key
// order this is the stable sort of the 2-level Morton key -- gives its sorted
// slot; every per-particle array moves there in the other buffer (loads in
// the previous order, issued while the rank is found; stores scattered), and
// the sorted slot records its sub-cell (in pslot, free after k_scatter) for
// the neighbor-row build.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sortgather(PoolDev P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nact) return;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    if (i >= d->part_base + d->n) return;
    const int cell = P.pcell[i];
    const int key = P.psub[i];
    const int cur = d->sel, cura = d->sela;
    const float4 p = P.pos[cur][i];
    float4 v = P.vel[cur][i];
    const int tg = P.tag[cura][i];
    const int mk = P.mask[cura][i], og = P.origin[cura][i], im = P.image[cura][i];
    const float ph = P.phi[cura][i];
    const int start = P.cell_start[cell], end = P.cell_start[cell + 1];
    // four members' keys in flight per round trip (loads clamped to the
    // segment, members past its end not counted)
    int rank = 0;
    for (int g0 = start; g0 < end; g0 += 4) {
        int kg[4], jg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int g = min(g0 + u, end - 1);
            kg[u] = P.cl[g];
            jg[u] = P.tmp[g];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            rank += g0 + u < end && ((kg[u] < key) || (kg[u] == key && jg[u] < i));
    }
    const int dst = start + rank;
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
    P.mask[cura ^ 1][dst] = mk;
    P.origin[cura ^ 1][dst] = og;
    P.image[cura ^ 1][dst] = im;
    P.phi[cura ^ 1][dst] = ph;
    P.pslot[dst] = cell;
    if (P.key) P.key[(size_t)P.cap + dst] = P.key[i];
}

// --------------------------------------------------------------------------
// K5: neighbor rows, one wave per cell.  A centre's row lists every
// (j, image) of the 27 stencil cells of its cell -- stencil cell k has offset
// (k%3-1, k/3%3-1, k/9-1), periodic images shift by +-L -- in stencil order,
// then j ascending, with (float) r^2 <= (float)(rc+skin)^2, itself excluded.
// r^2 = fma(dz, dz, fma(dy, dy, dx dx)): the reference's
// `dx * dx + dy * dy + dz * dz` (neigh_build_meso.cu:86-93) as its -O3 build
// (nvcc, floating-point contraction on by default) evaluates it.  The set is
// the reference's (neigh_build_meso.cu:58-118).
//
// All centres of a cell share the stencil, so the wave walks the stencil's
// candidates 64 at a time, one per lane (coalesced loads, image shift applied
// once), NL_ITER iterations held in registers, and tests them against one
// centre after another (the centre's position is uniform).  The lanes that
// pass append in lane order (ballot + mbcnt) to an LDS copy of the centre's
// row, which keeps the row order; the copy is stored to the chunked row
// layout once per centre and batch.  The workgroup is one wave.
// Rows are built from the positions of the last build (xhold), which the
// cells were sorted by: right after the re-sort they equal the sorted
// positions; with relist (the row width grew) every sorted system gets back
// exactly the rows of its last build, whatever steps it has taken since, so
// the rebuild trigger (displacement from xhold) keeps its guarantee.
// relist selects the systems not flagged stale, otherwise the stale ones.
//
// A pool's rebuild walks the worklist k_scan_apply made: the cells holding
// particles of the systems re-sorted this pass, each with its system,
// coordinates and particle range; a listed cell reads its 27 neighbor ranges
// from the linear range table.  A pool's cell slots are
// mostly empty (a droplet fills a quarter of its box) or belong to systems
// not re-sorted, and dispatching one workgroup per cell slot costs ~0.2 ns
// each whether it works or exits (136 us per pass for 650,000 slots), so a
// fixed grid of NL_GRID waves takes the list instead, XCD x the x-th eighth
// of it in strides, so that cells next to each other in Morton order share
// that XCD's L2 (batched droplets: rebuild phase 0.379 -> 0.278 ms per pass).
// A relist covers every cell slot, one wave per slot.
// --------------------------------------------------------------------------
constexpr int NL_ITER = 7;             // candidate tiles per batch: 448 candidates (a cell's stencil averages ~370)
constexpr int NL_WAVES = 8;            // waves per SIMD of the worklist walk (63 VGPRs; measured
                                       // faster than seven, which needs no SGPR spills to VGPR lanes)
constexpr int NL_GRID = 65536;         // its waves (measured against 7,168 and 16,384: more
                                       // waves balance uneven cells better; 65,536 cost 14 us
                                       // to dispatch when no cell is listed)

// LISTED: e = (system, packed coordinates, first particle, end) of a listed
// cell, whose neighbor cells' ranges come from the linear table; otherwise
// e.x = a cell slot, looked up through its system and the Morton rank table
template <bool LISTED>
__device__ __attribute__((always_inline)) void nlist_cell(const PoolDev &P, int4 e, int relist, int *s_row)
{
    int s, xyz, c0, c1, cs = e.x;
    if (LISTED) {
        s = e.x; xyz = e.y; c0 = e.z; c1 = e.w;
    } else {
        s = P.cell_sys[cs];
        if (s < 0) return;
    }
    SysDev *d = P.sys + s;
    const int cb = d->cell_base;
    if (!LISTED) {
        if (!d->active || (relist ? d->stale != 0 : d->stale == 0)) return;
        if (cs < cb || cs >= cb + d->ncell) return;       // the end slot holds no cell
        c0 = P.cell_start[cs * SUBCELLS];
        c1 = P.cell_start[(cs + 1) * SUBCELLS];
        if (c0 >= c1) return;
    }
    const int lane = threadIdx.x;
    const float4 *pos = P.xhold;
    // stencil table: lane k < 27 holds cell k's first particle, its length,
    // image code and shift
    int kst = 0, klen = 0;
    uint32_t kcode = IMG_ZERO;
    float ksx = 0.f, ksy = 0.f, ksz = 0.f;
    if (lane < 27) {
        if (!LISTED) xyz = P.cell_xyz[cs];
        const int cx = xyz & 1023, cy = (xyz >> 10) & 1023, cz = (xyz >> 20) & 1023;
        const int nb0 = d->nb[0], nb1 = d->nb[1], nb2 = d->nb[2];
        int nx = cx + lane % 3 - 1, ny = cy + lane / 3 % 3 - 1, nz = cz + lane / 9 - 1;
        int sx = 0, sy = 0, sz = 0;
        if (nx < 0) { nx += nb0; sx = -1; } else if (nx >= nb0) { nx -= nb0; sx = 1; }
        if (ny < 0) { ny += nb1; sy = -1; } else if (ny >= nb1) { ny -= nb1; sy = 1; }
        if (nz < 0) { nz += nb2; sz = -1; } else if (nz >= nb2) { nz -= nb2; sz = 1; }
        const int lin = nx + nb0 * (ny + nb1 * nz);
        if (LISTED) {
            const int2 r = P.cell_lse[cb + lin];
            kst = r.x;
            klen = r.y - r.x;
        } else {
            const int rank = P.cell_rank[cb + lin];
            kst = P.cell_start[(cb + rank) * SUBCELLS];
            klen = P.cell_start[(cb + rank + 1) * SUBCELLS] - kst;
        }
        kcode = (uint32_t)((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1));
        ksx = (float)sx * d->L[0];
        ksy = (float)sy * d->L[1];
        ksz = (float)sz * d->L[2];
    }
    const int kend = wave_incl_scan(klen, lane);
    const int kpre = kend - klen;                           // candidates before cell k
    const int ntot = __builtin_amdgcn_readlane(kend, 26);
    const float rc2 = d->rc2_list;
    const int maxnbr = P.maxnbr;
    for (int cg = c0; cg < c1; cg += WAVE) {
        const int ng = min(WAVE, c1 - cg);
        int cntv = 0;                                       // lane l: row length of centre cg + l
        const float4 pg = pos[min(cg + lane, c1 - 1)];      // lane l: position of centre cg + l
        for (int tb = 0; tb < ntot; tb += NL_ITER * WAVE) {
            // candidates tb + 64 m + lane: stencil cell by binary search over
            // the exclusive prefix (the last cell starting at or before t)
            float qx[NL_ITER], qy[NL_ITER], qz[NL_ITER];
            int qe[NL_ITER];
            const int nm = min(NL_ITER, (ntot - tb + WAVE - 1) / WAVE);
#pragma unroll
            for (int m = 0; m < NL_ITER; ++m) {
                if (m >= nm) break;
                const int t = tb + m * WAVE + lane;
                int k = 0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1) {
                    const int kk = min(k + step, 26);
                    if (__shfl(kpre, kk) <= t) k = kk;
                }
                // every shuffle with the whole wave active (a lane reads a
                // source lane's register only while that lane is active)
                const int st_k = __shfl(kst, k), pre_k = __shfl(kpre, k);
                const uint32_t code = (uint32_t)__shfl((int)kcode, k);
                const float sx_k = __shfl(ksx, k), sy_k = __shfl(ksy, k), sz_k = __shfl(ksz, k);
                const bool valid = t < ntot;
                const int j = valid ? st_k + (t - pre_k) : c0;
                const float4 q = pos[j];
                qx[m] = valid ? q.x + sx_k : __builtin_inff();
                qy[m] = valid ? q.y + sy_k : __builtin_inff();
                qz[m] = valid ? q.z + sz_k : __builtin_inff();
                qe[m] = (int)((uint32_t)j | (code << IMG_SHIFT));
            }
            for (int c = cg; c < cg + ng; ++c) {
                const int cl = c - cg;
                const float pcx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pg.x), cl));
                const float pcy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pg.y), cl));
                const float pcz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pg.z), cl));
                const int q0 = __builtin_amdgcn_readlane(cntv, cl);     // row length before this batch
                int n = 0;                                  // hits of this batch, staged in LDS
#pragma unroll
                for (int m = 0; m < NL_ITER; ++m) {
                    if (m >= nm) break;
                    const float dx = pcx - qx[m], dy = pcy - qy[m], dz = pcz - qz[m];
                    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                    const uint64_t mask = __builtin_amdgcn_ballot_w64(r2 <= rc2);
                    if (mask) {
                        if (__builtin_amdgcn_inverse_ballot_w64(mask))
                            s_row[n + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] = qe[m];
                        n += __popcll(mask);
                    }
                }
                // this batch's entries of centre c: rows [q0, ...) in the
                // chunked layout, entries past the row width dropped (the
                // count marks the overflow).  The centre itself is among the
                // hits once in its whole row (stencil cell 13, the only one
                // with a zero image shift, distance 0): that entry is left
                // out and the entries after it move up by one.
                __syncthreads();
                const int selfe = (int)((uint32_t)c | (IMG_ZERO << IMG_SHIFT));
                int *row = P.nlist + nbr_at(c, 0, maxnbr);
                int shift = 0;
                for (int k0 = 0; k0 < n; k0 += WAVE) {
                    const int k = k0 + lane;
                    const int v = k < n ? s_row[k] : 0;
                    const uint64_t sb = __builtin_amdgcn_ballot_w64(k < n && v == selfe);
                    const int ps = sb ? (int)__builtin_ctzll(sb) : WAVE;   // lane holding the centre
                    const uint32_t q = (uint32_t)(q0 + k - shift - (lane > ps ? 1 : 0));
                    if (k < n && lane != ps && q < (uint32_t)maxnbr)
                        row[(q & ~(uint32_t)(NBR_CHUNK - 1)) * WAVE + (q & (uint32_t)(NBR_CHUNK - 1))] = v;
                    if (sb) shift = 1;
                }
                __syncthreads();
                if (lane == cl) cntv = q0 + n - shift;
            }
        }
        if (lane < ng) P.ncount[cg + lane] = cntv;
        if (__ballot(lane < ng && cntv > maxnbr) && lane == 0) atomicOr(&d->overflow, 1);
    }
}

// one wave per cell slot (a relist, or one system): eight waves per SIMD
// (seven candidate tiles in registers, no spills), measured faster than seven
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(8, 8))) k_nlist(PoolDev P, int relist, int ncells)
{
    __shared__ int s_row[NL_ITER * WAVE];                   // one centre's entries of one batch
    const int cs = xcd_logical(blockIdx.x, ncells);        // cell slot (Morton rank order)
    if (cs < ncells) nlist_cell<false>(P, make_int4(cs, 0, 0, 0), relist, s_row);
}

// Commit after a rebuild: the sorted data in the other buffers becomes current.
__device__ inline void commit_rebuild_sys(const PoolDev &P, int s)
{
    SysDev *d = P.sys + s;
    if (!d->active || !d->stale) return;
    d->sel ^= 1;
    d->sela ^= 1;
    d->stale = 0;
    d->ago = 0;
    d->builds += 1;
}

// the worklist walk; it also commits the rebuild of every re-sorted system
// (the row build reads only their geometry, and the next kernel sees the
// flipped buffers), which saves a launch per pass
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(NL_WAVES, NL_WAVES))) k_nlist_listed(PoolDev P)
{
    __shared__ int s_row[NL_ITER * WAVE];
    for (int s = blockIdx.x * WAVE + threadIdx.x; s < P.nsys; s += gridDim.x * WAVE) commit_rebuild_sys(P, s);
    // gridDim.x is a multiple of NXCD; workgroup b runs on XCD b % NXCD
    const int n = *P.work_count;
    const int x = blockIdx.x % NXCD, per = gridDim.x / NXCD;
    const int hi = (int)((long long)n * (x + 1) / NXCD);
    for (int w = (int)((long long)n * x / NXCD) + blockIdx.x / NXCD; w < hi; w += per)
        nlist_cell<true>(P, P.cell_work[w], 0, s_row);
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
// Commit after a rebuild with one wave per cell slot (the plain path)
// --------------------------------------------------------------------------
__global__ void k_commit_rebuild(PoolDev P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < P.nsys) commit_rebuild_sys(P, s);
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
static void launch_nlist(const PoolDev &P, int relist, hipStream_t st)
{
    if (P.ncact <= 0) return;
    if (relist || !nlist_listed(P))
        hipLaunchKernelGGL(k_nlist, dim3(xcd_grid(P.ncact)), dim3(WAVE), 0, st, P, relist, P.ncact);
    else
        hipLaunchKernelGGL(k_nlist_listed, dim3(std::min(NL_GRID, xcd_grid(P.ncact))), dim3(WAVE), 0, st, P);
}

hipError_t launch_rebuild(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    const dim3 g(nblk(P.nact, 256)), b(256);
    hipLaunchKernelGGL(k_keys, g, b, 0, st, P);
    // workgroups per system for the chunked scan: as many as the largest
    // system has chunks, up to 64 (smaller systems' extra workgroups exit)
    const long long nsub = P.maxsub > 0 ? (long long)P.maxsub : (long long)P.ncellslots * SUBCELLS;
    const int nch = (int)((nsub + SCAN_CH - 1) / SCAN_CH);
    const int gs = std::max(1, std::min(64, nch));
    hipLaunchKernelGGL(k_scan_sums, dim3(gs, P.nsys), dim3(SCAN_BLK), 0, st, P);
    hipLaunchKernelGGL(k_scan_chunks, dim3(P.nsys), dim3(SCAN_BLK), 0, st, P);
    hipLaunchKernelGGL(k_scan_apply, dim3(gs, P.nsys), dim3(SCAN_BLK), 0, st, P);
    hipLaunchKernelGGL(k_scatter, g, b, 0, st, P);
    hipLaunchKernelGGL(k_sortgather, g, b, 0, st, P);
    launch_nlist(P, 0, st);
    return hipGetLastError();
}

hipError_t launch_nlist_only(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    launch_nlist(P, 0, st);
    return hipGetLastError();
}

hipError_t launch_relist(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_overflow, dim3(nblk(P.nsys, 64)), dim3(64), 0, st, P);
    launch_nlist(P, 1, st);
    return hipGetLastError();
}

hipError_t launch_commit_rebuild(const PoolDev &P, hipStream_t st)
{
    // after a worklist walk the commit has been made by the walk itself
    if (nlist_listed(P) && P.ncact > 0 && P.nact > 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit_rebuild, dim3(nblk(P.nsys, 64)), dim3(64), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_signatures(const PoolDev &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_signatures, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P);
    return hipGetLastError();
}

}  // namespace mdpd
