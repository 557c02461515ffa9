// Neighbor-rebuild phase of the mdpd355 engine: periodic wrap, cell keys,
// counting sort by cell + LDS per-cell sort of the sub-bin key, gather into the
// other buffer, wave64 ballot/prefix neighbor-list build, commit.
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
// system; clears the counts for the next rebuild.  The system's non-empty
// cells are compacted (staged in cl[part_base ...], at most n of them) and
// appended to the pool's work list wl with one atomic per system, so the
// per-cell kernels visit occupied cells of stale systems only.
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
    __shared__ int wex[16], wne[16];
    __shared__ int carry, ne_carry, ne_base;
    const int nc = d->ncell, base = d->cell_base, pb = d->part_base;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) { carry = pb; ne_carry = 0; }
    __syncthreads();
    for (int c0 = 0; c0 < nc; c0 += CH) {
        for (int k = t; k < CH; k += 1024) {
            const int c = c0 + k;
            int v = 0;
            if (c < nc) { v = P.cell_count[base + c]; P.cell_count[base + c] = 0; }
            buf[k] = v;
        }
        __syncthreads();
        int v[PER], sum = 0, nne = 0;
        for (int k = 0; k < PER; ++k) { v[k] = buf[t * PER + k]; sum += v[k]; nne += v[k] > 0; }
        const int incl = wave_incl_scan(sum, lane);
        const int nincl = wave_incl_scan(nne, lane);
        if (lane == 63) { wex[w] = incl; wne[w] = nincl; }
        __syncthreads();
        if (t == 0) {
            int acc = 0, acn = 0;
            for (int k = 0; k < 16; ++k) {
                const int y = wex[k], z = wne[k];
                wex[k] = acc; wne[k] = acn;
                acc += y; acn += z;
            }
        }
        __syncthreads();
        int run = carry + wex[w] + incl - sum;
        int nrun = ne_carry + wne[w] + nincl - nne;
        for (int k = 0; k < PER; ++k) {
            buf[t * PER + k] = run;
            run += v[k];
            if (v[k] > 0) P.cl[pb + nrun++] = base + c0 + t * PER + k;
        }
        __syncthreads();
        for (int k = t; k < CH; k += 1024) {
            const int c = c0 + k;
            if (c < nc) P.cell_start[base + c] = buf[k];
        }
        if (t == 1023) { carry = run; ne_carry = nrun; }
        __syncthreads();
    }
    if (t == 0) {
        P.cell_start[base + nc] = pb + d->n;
        d->nne = ne_carry;
        ne_base = ne_carry > 0 ? atomicAdd(P.work_total, ne_carry) : 0;
    }
    __syncthreads();
    const int b0 = ne_base, m = min(ne_carry, P.cap - b0);
    for (int k = t; k < m; k += 1024) P.wl[b0 + k] = P.cl[pb + k];
}

// --------------------------------------------------------------------------
// Work list of every occupied cell of the active, sorted systems, from the
// cell offsets of their last re-sort (neighbor rows rebuilt in place after the
// row width grew).  Stale systems are left out: their next re-sort lists them,
// and a system admitted but not yet sorted has no valid offsets.  One
// workgroup per system.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_relist_cells(PoolDev P)
{
    SysDev *d = P.sys + blockIdx.x;
    if (!d->active || d->stale) return;
    __shared__ int s_base;
    const int nc = d->ncell, base = d->cell_base;
    if (threadIdx.x == 0) {
        s_base = d->nne > 0 ? atomicAdd(P.work_total, d->nne) : 0;
        d->overflow = 0;                 // k_nlist sets it again if the rows are still too narrow
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    __shared__ int s_off;
    if (threadIdx.x == 0) s_off = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nc; c0 += 256) {
        const int c = c0 + threadIdx.x;
        const bool ne = c < nc && P.cell_start[base + c + 1] > P.cell_start[base + c];
        const uint64_t bal = __ballot(ne);
        int wbase = 0;
        if (lane == 0 && bal) wbase = atomicAdd(&s_off, __popcll(bal));
        wbase = __shfl(wbase, 0, 64);
        if (ne) {
            const int p = wbase + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (p < d->nne && s_base + p < P.cap) P.wl[s_base + p] = base + c;
        }
    }
}

// --------------------------------------------------------------------------
// K3: scatter particle indices into their cell segments (arrival order).
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
    P.tmp[P.cell_start[P.pcell[i]] + P.pslot[i]] = i;
}

// --------------------------------------------------------------------------
// K4: one wave per cell: order the cell's particles by (sub-bin key, previous
// index) -- the stable sort of the 2-level Morton key -- and gather every
// per-particle array into the other buffer.
// --------------------------------------------------------------------------
constexpr int K4_LDS = 128;

__global__ void __launch_bounds__(256) k_sortgather(PoolDev P)
{
    __shared__ int s_key[4][K4_LDS];
    __shared__ int s_idx[4][K4_LDS];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // grid-stride over the work list: occupied cells of the stale systems
    const int total = min(*P.work_total, P.cap);   // at most one entry per particle slot
    for (int item = blockIdx.x * 4 + wv; item < total; item += gridDim.x * 4) {
        const int slot = P.wl[item];
        SysDev *d = P.sys + P.cell_sys[slot];
        const int start = P.cell_start[slot], end = P.cell_start[slot + 1];
        const int m = end - start;
        if (m <= 0) continue;
        const int cur = d->sel, cura = d->sela;
        const bool fresh = d->phase == PH_FRESH;
        const bool in_lds = m <= K4_LDS;
        if (in_lds) {
            for (int e = lane; e < m; e += WAVE) {
                const int idx = P.tmp[start + e];
                s_idx[wv][e] = idx;
                s_key[wv][e] = P.psub[idx];
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int e = lane; e < m; e += WAVE) {
            const int idx = in_lds ? s_idx[wv][e] : P.tmp[start + e];
            const int key = in_lds ? s_key[wv][e] : P.psub[idx];
            int rank = 0;
            for (int f = 0; f < m; ++f) {
                const int kf = in_lds ? s_key[wv][f] : P.psub[P.tmp[start + f]];
                const int jf = in_lds ? s_idx[wv][f] : P.tmp[start + f];
                rank += (kf < key) || (kf == key && jf < idx);
            }
            const int dst = start + rank;
            const float4 p = P.pos[cur][idx];
            float4 v = P.vel[cur][idx];
            const int tg = P.tag[cura][idx];
            if (fresh) {
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
            if (P.key) P.key[(size_t)P.cap + dst] = P.key[idx];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// --------------------------------------------------------------------------
// K5: one wave per cell: neighbor rows for the cell's particles.  Lanes take
// candidates from the 27 stencil cells (periodic images), every centre is
// tested with one ballot, and hits are appended at count + mbcnt(ballot) --
// the wave64 form of gpu_build_neighbor_list (neigh_build_meso.cu:58-118).
// Hit test: (float) r^2 <= (float)(rc+skin)^2 without contraction
// (neigh_build_meso.cu:86-93).  Row layout: tile (i/64), entry p at
// tile*maxnbr*64 + p*64 + i%64.
// --------------------------------------------------------------------------
constexpr int K5_CH = 16;          // centres per LDS chunk (per wave)
constexpr int K5_WAVES = 4;        // independent waves (cells) per workgroup

// per-wave LDS words: stencil start/prefix/code, stencil shifts (3 each), rows
// of the chunk's centres (odd stride: conflict-free column reads).  Centre
// coordinates and row counts live in lane c of the wave and are broadcast with
// readlane.
static inline __host__ __device__ int nlist_wave_words(int maxnbr)
{
    return 32 * 3 + 32 * 3 + K5_CH * (maxnbr + 1);
}

static inline size_t nlist_lds_bytes(int maxnbr) { return (size_t)K5_WAVES * nlist_wave_words(maxnbr) * 4; }

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(WAVE * K5_WAVES) k_nlist(PoolDev P, int relist)
{
    extern __shared__ int smem[];
    const int maxnbr = P.maxnbr;
    const int stride = maxnbr + 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int *base_w = smem + wv * nlist_wave_words(maxnbr);
    int *s_start = base_w;                             // 32
    int *s_pre = s_start + 32;                         // 32
    int *s_code = s_pre + 32;                          // 32
    float *s_sh = (float *)(s_code + 32);              // 32 * 3
    int *s_rows = (int *)(s_sh + 96);                  // K5_CH * stride
    // grid-stride over the work list (occupied cells of the systems to list)
    const int total = min(*P.work_total, P.cap);   // at most one entry per particle slot
    for (int item = blockIdx.x * K5_WAVES + wv; item < total; item += gridDim.x * K5_WAVES) {
        const int slot = P.wl[item];
        SysDev *d = P.sys + P.cell_sys[slot];
        const int start = P.cell_start[slot], end = P.cell_start[slot + 1];
        const int m = end - start;
        // a re-sort lists the sorted data k_sortgather wrote to the other
        // buffer; a relist (rows grown in place) lists the current buffer
        const float4 *pos = P.pos[relist ? d->sel : d->sel ^ 1];

        // stencil: lane k < 27 describes neighbor cell k
        const int xyz = P.cell_xyz[slot];
        const int cx = xyz & 1023, cy = (xyz >> 10) & 1023, cz = (xyz >> 20) & 1023;
        int len = 0;
        if (lane < 27) {
            const int ox = lane % 3 - 1, oy = (lane / 3) % 3 - 1, oz = lane / 9 - 1;
            int nx = cx + ox, ny = cy + oy, nz = cz + oz, sx = 0, sy = 0, sz = 0;
            if (nx < 0) { nx += d->nb[0]; sx = -1; } else if (nx >= d->nb[0]) { nx -= d->nb[0]; sx = 1; }
            if (ny < 0) { ny += d->nb[1]; sy = -1; } else if (ny >= d->nb[1]) { ny -= d->nb[1]; sy = 1; }
            if (nz < 0) { nz += d->nb[2]; sz = -1; } else if (nz >= d->nb[2]) { nz -= d->nb[2]; sz = 1; }
            const int lin = nx + d->nb[0] * (ny + d->nb[1] * nz);
            const int ns = d->cell_base + P.cell_rank[d->cell_base + lin];
            const int st = P.cell_start[ns];
            len = P.cell_start[ns + 1] - st;
            s_start[lane] = st;
            s_code[lane] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
            s_sh[lane * 3 + 0] = (float)sx * d->L[0];
            s_sh[lane * 3 + 1] = (float)sy * d->L[1];
            s_sh[lane * 3 + 2] = (float)sz * d->L[2];
        }
        int incl = len;
        for (int off = 1; off < 32; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane < 27) s_pre[lane] = incl - len;
        const int total = __shfl(incl, 26, 64);
        wave_sync();

        const float rc2 = d->rc2_list;
        int any_over = 0;
        for (int c0 = 0; c0 < m; c0 += K5_CH) {
            const int mc = min(K5_CH, m - c0);
            float myx = 0.f, myy = 0.f, myz = 0.f;
            int mycnt = 0;
            if (lane < mc) {
                const float4 q = pos[start + c0 + lane];
                myx = q.x; myy = q.y; myz = q.z;
            }
            // lanes take the flattened candidates of the 27 stencil cells, 64
            // at a time; the next batch's index, image code and position are
            // fetched before the current batch is tested, so the gather's
            // latency overlaps the centre loop
            auto fetch = [&](int b0, int &j, int &code, int &k, float4 &q) {
                const int mi = b0 + lane;
                j = 0;
                code = IMG_ZERO;
                k = 0;
                q = make_float4(0.f, 0.f, 0.f, 0.f);
                if (mi < total) {
                    // largest k with pre[k] <= mi (pre is non-decreasing)
                    int lo = 0, hi = 26;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_pre[mid] <= mi) lo = mid; else hi = mid - 1;
                    }
                    k = lo;
                    j = s_start[k] + (mi - s_pre[k]);
                    code = s_code[k];
                    q = pos[j];
                }
            };
            int jn, cn, kn;
            float4 qn;
            fetch(0, jn, cn, kn, qn);
            for (int b0 = 0; b0 < total; b0 += WAVE) {
                const bool valid = b0 + lane < total;
                const int j = jn, code = cn, k = kn;
                const float4 pj = qn;
                if (b0 + WAVE < total) fetch(b0 + WAVE, jn, cn, kn, qn);
                float xj = 0.f, yj = 0.f, zj = 0.f;
                if (valid) {
                    xj = pj.x + s_sh[k * 3 + 0];
                    yj = pj.y + s_sh[k * 3 + 1];
                    zj = pj.z + s_sh[k * 3 + 2];
                }
                const int entry = (int)((uint32_t)j | ((uint32_t)code << IMG_SHIFT));
                for (int c = 0; c < mc; ++c) {
                    // centre c lives in lane c: broadcast through scalar registers
                    const float xc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myx), c));
                    const float yc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myy), c));
                    const float zc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myz), c));
                    const float dx = xc - xj, dy = yc - yj, dz = zc - zj;
                    const float r2 = (dx * dx + dy * dy) + dz * dz;
                    const int ic = start + c0 + c;
                    const bool hit = valid && r2 <= rc2 && !(j == ic && code == IMG_ZERO);
                    const uint64_t bal = __ballot(hit);
                    if (bal) {
                        const int n0 = __builtin_amdgcn_readlane(mycnt, c);
                        const int p = n0 + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                        if (hit && p < maxnbr) s_rows[c * stride + p] = entry;
                        if (lane == c) mycnt += __popcll(bal);
                    }
                }
            }
            wave_sync();
            // counts, then rows: lane = centre (lane & 15) and row offset (lane >> 4)
            int cntc = 0;
            if (lane < mc) {
                cntc = mycnt;
                P.ncount[start + c0 + lane] = cntc;
                any_over |= cntc > maxnbr;
            }
            int rmax = min(cntc, maxnbr);
            for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off, 64));
            const int c = lane & (K5_CH - 1);
            const int myc = c < mc ? min(__shfl(mycnt, c, 64), maxnbr) : 0;
            const int ic = start + c0 + c;
            const size_t base = (size_t)(ic >> 6) * maxnbr * WAVE + (ic & 63);
            for (int p0 = 0; p0 < rmax; p0 += WAVE / K5_CH) {
                const int p = p0 + lane / K5_CH;
                if (p < myc) P.nlist[base + (size_t)p * WAVE] = s_rows[c * stride + p];
            }
            wave_sync();
        }
        if (__ballot(any_over) && lane == 0) atomicOr(&d->overflow, 1);
        wave_sync();
    }
}

// --------------------------------------------------------------------------
// Commit after a rebuild: the sorted data in the other buffers becomes current.
// --------------------------------------------------------------------------
__global__ void k_commit_rebuild(PoolDev P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) *P.work_total = 0;       // the work list has been consumed
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
// cell-parallel kernels loop over the work list (at most one entry per
// particle); up to 2048 workgroups of 4 waves cover the chip
constexpr int CELL_GRID_MAX = 2048;
static inline int cell_grid(const PoolDev &P, int waves_per_block)
{
    return std::max(1, std::min(nblk(P.nact, waves_per_block * 8), CELL_GRID_MAX));
}

hipError_t launch_rebuild(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P);
    hipLaunchKernelGGL(k_scan, dim3(P.nsys), dim3(1024), 0, st, P);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(P.nact, 256)), dim3(256), 0, st, P);
    hipLaunchKernelGGL(k_sortgather, dim3(cell_grid(P, 4)), dim3(256), 0, st, P);
    hipLaunchKernelGGL(k_nlist, dim3(cell_grid(P, K5_WAVES)), dim3(WAVE * K5_WAVES), nlist_lds_bytes(P.maxnbr), st, P, 0);
    return hipGetLastError();
}

hipError_t launch_nlist_only(const PoolDev &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_nlist, dim3(cell_grid(P, K5_WAVES)), dim3(WAVE * K5_WAVES), nlist_lds_bytes(P.maxnbr), st, P, 0);
    return hipGetLastError();
}

hipError_t launch_relist(const PoolDev &P, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    hipLaunchKernelGGL(k_relist_cells, dim3(P.nsys), dim3(256), 0, st, P);
    hipLaunchKernelGGL(k_nlist, dim3(cell_grid(P, K5_WAVES)), dim3(WAVE * K5_WAVES), nlist_lds_bytes(P.maxnbr), st, P, 1);
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
