// Per-step passes of the mdpd355 engine: local density, initial half-kick,
// and the pair-force pass with velocity-Verlet fused into its epilogue.
//
// Replaces, per step (ModifiedVerlet::run, src/USER-MESO/mvv_meso.cu:321-564):
//   FixArbitraryBoundaryRho::pre_force / gpu_compute_rho   fix_aboundary_rho_meso.cu:119-212
//   MesoPairMDPD::compute / gpu_mdpd                        pair_mdpd_meso.cu:93-293
//   MesoPairDPD::compute / gpu_dpd                          pair_dpd_meso.cu:91-266
//   FixNVEMeso::initial_integrate / final_integrate         fix_nve_meso.cu:63-199
//   MesoFixAddForce::post_force                             fix_addforce_meso.cu:72-108
//   dp2sp_merged signatures                                 atom_vec_mdpd_atomic_meso.cu:140-206
//
// Each step the first pass (density for mDPD, a filter pass for DPD) walks the
// full neighbor rows (cutoff + skin) and writes, per particle, the compact row
// of entries inside the pair cutoff; the force pass then walks only those.
// Compact rows are staged per lane in LDS and stored as whole coalesced rows.
//
// Launch shape: the first and force passes run one workgroup per 64-particle
// tile with SPLIT waves sharing the tile's rows; systems are padded to 256
// slots; workgroup ids are remapped so each of the 8 XCDs sweeps one
// contiguous particle range.
#include "common.hpp"
#include "kernels.hpp"

#include <type_traits>

namespace mdpd {

constexpr int BLK = 256;

// Per-tile "a particle moved past skin/2" flags: each workgroup ORs its lanes
// in LDS and stores one word per 64-particle tile it covers (plain stores, no
// atomics: same-address device-scope atomics serialize across the XCDs and
// cost ~50 ns each).  k_commit_step reduces them per system after the pass.
__device__ inline void tile_flags(int *moved, int tile0, int ntile, int flag)
{
    __shared__ int s_or[4];
    if (threadIdx.x < 4) s_or[threadIdx.x] = 0;
    __syncthreads();
    if (flag) s_or[(threadIdx.x >> 6) % ntile] = 1;
    __syncthreads();
    if (threadIdx.x < ntile) moved[tile0 + threadIdx.x] = s_or[threadIdx.x];
}

// a - (b + s L) with a single rounding.  Adding s L to the neighbor's
// coordinate first rounds at the magnitude of L (half an ulp of L: 4.8e-7 at
// L = 8), an error in r that the density weight (r_d - r)^2 amplifies for a
// neighbor near r_d.  Here the difference a - b and its exact rounding error
// (Knuth's TwoSum) come first; a - b lies within r_c + skin of s L, so for
// L >= 2 (r_c + skin) subtracting s L is exact (Sterbenz) and the error term
// is added back last, giving the correctly rounded displacement.  With s = 0
// the result equals a - b bit for bit.
__device__ inline float image_diff(float a, float b, float sl)
{
    const float d = a - b;
    const float bv = d - a;
    const float e = (a - (d - bv)) + (-b - bv);
    return (d - sl) + e;
}

// displacement r_i - (r_j + s L) of a neighbor entry with image code `code`
__device__ inline void pair_delta(const float4 &pi, const float4 &q, uint32_t code, const SysDev *d, float &dx,
                                  float &dy, float &dz)
{
    if (code == IMG_ZERO) {
        dx = pi.x - q.x; dy = pi.y - q.y; dz = pi.z - q.z;
    } else {
        int sx, sy, sz;
        image_decode(code, sx, sy, sz);
        dx = image_diff(pi.x, q.x, (float)sx * d->L[0]);
        dy = image_diff(pi.y, q.y, (float)sy * d->L[1]);
        dz = image_diff(pi.z, q.z, (float)sz * d->L[2]);
    }
}

// r^2 with the fusion spelled out, so the first pass's compaction test and
// the force pass's cutoff test see the same bits
__device__ inline float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

__device__ inline float pair_rsq(const float4 &pi, const float4 &q, uint32_t code, const SysDev *d)
{
    float dx, dy, dz;
    pair_delta(pi, q, code, d, dx, dy, dz);
    return dist2(dx, dy, dz);
}

// --------------------------------------------------------------------------
// first pass: density rho_i = m_i * sum_j w(r_ij) over 1e-20 <= r^2 < rd^2
// (no self term; fix_aboundary_rho_meso.cu:131-164), and the compact rows of
// entries with r^2 < rc^2 for the force pass.  WITH_RHO = false is the DPD
// filter pass (no density).
//
// One workgroup per 64-particle tile; wave w of SPLIT walks the chunks of
// row entries 8k..8k+7 with k = w mod SPLIT and writes its compact entries
// to compact rows t * SPLIT + w, so the force pass can give the same rows to
// the same wave.
// Partial densities are summed in LDS in wave order (deterministic).
// --------------------------------------------------------------------------
template <bool WITH_RHO, int SPLIT>
__global__ void __launch_bounds__(WAVE * SPLIT) k_first(PoolDev P, int ntiles)
{
    constexpr int STAGE = 32 / SPLIT;   // LDS-staged compact entries per lane (measured best)
    __shared__ int s_c[SPLIT][STAGE * WAVE];
    __shared__ float s_sum[SPLIT][WAVE];
    const int tile = xcd_logical(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = tile * WAVE + lane;
    const int s = P.wave_sys[tile];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || d->overflow) return;    // an overflowed system waits for wider rows
    const bool valid = i < d->part_base + d->n;
    const float4 *pos = P.pos[d->sel];
    const int maxnbr = P.maxnbr;
    const size_t tbase = (size_t)tile * maxnbr * WAVE + lane;
    const int *row = P.nlist + nbr_at(i, 0, maxnbr);
    int *crow = P.clist + tbase;
    int cnt = 0;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        pi = pos[i];
        cnt = min(P.ncount[i], maxnbr);
    }
    const float h = d->rd, h2 = h * h;
    const float rc2 = d->rc2_max;
    const bool lucy = d->lucy;
    float sum = 0.f;
    int nc = 0;
    // ROOM: the staging area is known to hold this entry (no spill check)
    auto visit = [&](uint32_t e, const float4 &q, auto room) {
        const float rsq = pair_rsq(pi, q, e >> IMG_SHIFT, d);
        if (WITH_RHO && rsq < h2 && rsq >= 1e-20f) {
            const float r = __builtin_amdgcn_sqrtf(rsq);
            const float hm = h - r;
            sum += lucy ? (h + 3.f * r) * hm * hm * hm : hm * hm;
        }
        if (rsq < rc2 && rsq >= 1e-20f) {
            if constexpr (decltype(room)::value) {
                s_c[w][nc * WAVE + lane] = (int)e;
            } else {
                if (nc < STAGE) s_c[w][nc * WAVE + lane] = (int)e;
                else crow[(size_t)(nc * SPLIT + w) * WAVE] = (int)e;
            }
            ++nc;
        }
    };
    // a chunk of eight entries (two 16-byte loads) and their eight position
    // gathers in flight per lane, the next chunk's entries loading meanwhile;
    // wave w takes chunks w, w + SPLIT, ... and visits them in row order.
    // Entries past the row's end are neither gathered nor visited.
    // With one wave per tile (SPLIT 1) the LDS staging, not registers, limits
    // occupancy, so a wave takes U = 2 chunks per iteration: sixteen gathers
    // in flight, the next two chunks' entries loading meanwhile.
    constexpr int U = SPLIT == 1 ? 2 : 1;
    const int nch = (cnt + NBR_CHUNK - 1) / NBR_CHUNK;
    const int4 *row4 = (const int4 *)row;
    int4 ea[U], eb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        ea[u] = make_int4(0, 0, 0, 0);
        eb[u] = ea[u];
        const int cu = w + u * SPLIT;
        if (cu < nch) {
            ea[u] = row4[(size_t)cu * (NBR_CHUNK * WAVE / 4)];
            eb[u] = row4[(size_t)cu * (NBR_CHUNK * WAVE / 4) + 1];
        }
    }
    for (int c = w; c < nch; c += SPLIT * U) {
        uint32_t e[U * NBR_CHUNK];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            e[u * NBR_CHUNK + 0] = (uint32_t)ea[u].x; e[u * NBR_CHUNK + 1] = (uint32_t)ea[u].y;
            e[u * NBR_CHUNK + 2] = (uint32_t)ea[u].z; e[u * NBR_CHUNK + 3] = (uint32_t)ea[u].w;
            e[u * NBR_CHUNK + 4] = (uint32_t)eb[u].x; e[u * NBR_CHUNK + 5] = (uint32_t)eb[u].y;
            e[u * NBR_CHUNK + 6] = (uint32_t)eb[u].z; e[u * NBR_CHUNK + 7] = (uint32_t)eb[u].w;
            const int cn = c + (U + u) * SPLIT;
            if (cn < nch) {
                ea[u] = row4[(size_t)cn * (NBR_CHUNK * WAVE / 4)];
                eb[u] = row4[(size_t)cn * (NBR_CHUNK * WAVE / 4) + 1];
            }
        }
        float4 q[U * NBR_CHUNK];
        // when every active lane's row covers all of this iteration's
        // entries and its staging area has room for all of them, the
        // per-entry row-end and staging checks are skipped (same result)
        const bool full = (c + (U - 1) * SPLIT + 1) * NBR_CHUNK <= cnt && nc + U * NBR_CHUNK <= STAGE;
        if (__builtin_amdgcn_ballot_w64(!full) == 0) {
#pragma unroll
            for (int k = 0; k < U * NBR_CHUNK; ++k) q[k] = pos[e[k] & IDX_MASK];
#pragma unroll
            for (int k = 0; k < U * NBR_CHUNK; ++k) visit(e[k], q[k], std::true_type{});
        } else {
#pragma unroll
            for (int k = 0; k < U * NBR_CHUNK; ++k) {
                const int p = (c + (k / NBR_CHUNK) * SPLIT) * NBR_CHUNK + k % NBR_CHUNK;
                q[k] = pos[(p < cnt ? e[k] : (uint32_t)i) & IDX_MASK];
            }
#pragma unroll
            for (int k = 0; k < U * NBR_CHUNK; ++k) {
                const int p = (c + (k / NBR_CHUNK) * SPLIT) * NBR_CHUNK + k % NBR_CHUNK;
                if (p < cnt) visit(e[k], q[k], std::false_type{});
            }
        }
    }
    // coalesced store of the staged rows
    int ncmax = min(nc, STAGE);
    for (int off = 32; off > 0; off >>= 1) ncmax = max(ncmax, __shfl_xor(ncmax, off, 64));
    for (int t = 0; t < ncmax; ++t)
        if (t < nc) crow[(size_t)(t * SPLIT + w) * WAVE] = s_c[w][t * WAVE + lane];
    P.ccount[ccount_at(tile, w, lane)] = nc;
    if (WITH_RHO) {
        s_sum[w][lane] = sum;
        __syncthreads();
        if (w == 0 && valid) {
            float tot = s_sum[0][lane];
            for (int k = 1; k < SPLIT; ++k) tot += s_sum[k][lane];
            const int mk = P.mask[d->sela][i];
            const int ti = __float_as_int(pi.w);
            const float r = (mk & d->rho_bit) ? P.mass[d->mass_base + ti] * d->rho_factor * tot : 0.f;
            P.rho[i] = r;
            // the force pass of a one-type system gathers a neighbor's
            // position and density with one 16-byte load
            P.pr[i] = make_float4(pi.x, pi.y, pi.z, r);
        }
    } else if (w == 0 && valid && d->style == STYLE_MDPD) {
        // mDPD without the density fix: zero density (P.rho keeps its zeros)
        P.pr[i] = make_float4(pi.x, pi.y, pi.z, 0.f);
    }
}

// --------------------------------------------------------------------------
// initial half-kick + drift from the stored force (first step of a run)
// --------------------------------------------------------------------------
__device__ inline uint32_t next_signature(const SysDev *d, int tag, float vx, float vy, float vz)
{
    return d->style == STYLE_MDPD ? sig_mdpd(tag, d->seed_next) : sig_dpd(tag, d->seed_next, vx, vy, vz);
}

__device__ inline void commit_advance(SysDev *d, bool moved)
{
    d->sel ^= 1;
    d->step += 1;
    d->seed_next = premix<64>(d->seed, (uint32_t)(d->step + 1));
    d->stale = moved ? 1 : 0;
}

__global__ void __launch_bounds__(BLK) k_initial(PoolDev P, int nlogical)
{
    const int lb = xcd_logical(blockIdx.x, nlogical);
    if (lb >= nlogical) return;
    const int i = lb * BLK + threadIdx.x;
    const int s = P.wave_sys[(lb * BLK) >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active) return;
    int moved = 0;
    if (i < d->part_base + d->n) {
        const int cur = d->sel;
        float4 p = P.pos[cur][i];
        float4 v = P.vel[cur][i];
        const float4 f = P.frc[i];
        const int mk = P.mask[d->sela][i];
        if (mk & d->nve_bit) {
            const float dtfm = d->dtf * P.minv[d->mass_base + __float_as_int(p.w)];
            v.x += dtfm * f.x; v.y += dtfm * f.y; v.z += dtfm * f.z;
            p.x += d->dt * v.x; p.y += d->dt * v.y; p.z += d->dt * v.z;
        }
        v.w = __uint_as_float(next_signature(d, P.tag[d->sela][i], v.x, v.y, v.z));
        P.pos[cur ^ 1][i] = p;
        P.vel[cur ^ 1][i] = v;
        const float4 h = P.xhold[i];
        const float ex = p.x - h.x, ey = p.y - h.y, ez = p.z - h.z;
        moved = ex * ex + ey * ey + ez * ez > d->trig2;
    }
    tile_flags(P.moved, lb * (BLK / WAVE), BLK / WAVE, moved);
}

// --------------------------------------------------------------------------
// pair force (+ fused velocity-Verlet) over the compact rows.  One workgroup
// per 64-particle tile; wave w of SPLIT takes compact rows t * SPLIT + w (the
// ones the first pass wrote for it); partial sums meet in LDS and wave 0 adds
// them in wave order, then runs the integration epilogue.
// --------------------------------------------------------------------------
struct PairC {
    float rc, rc2, rcinv, rd, rdinv, a, b, gamma, sigma, expw;
};

__device__ inline PairC load_coef(const float *tab, int idx)
{
    const float *c = tab + idx * NCOEF;
    PairC q;
    q.rc = c[C_RC]; q.rc2 = c[C_RC2]; q.rcinv = c[C_RCINV]; q.rd = c[C_RD]; q.rdinv = c[C_RDINV];
    q.a = c[C_A]; q.b = c[C_B]; q.gamma = c[C_GAMMA]; q.sigma = c[C_SIGMA]; q.expw = c[C_EXPW];
    return q;
}

template <bool EV>
struct Acc {
    static constexpr int N = EV ? 10 : 3;
    float v[N];
    __device__ Acc() { for (int k = 0; k < N; ++k) v[k] = 0.f; }
};

// TEST = false: the entry is known to pass the cutoff test (a one-type
// system's compact row holds exactly the entries with
// 1e-20 <= r^2 < rc^2, r^2 computed from the same positions by dist2, and
// the system's largest cutoff is its only one)
template <int STYLE, bool EV, bool TEST = true>
__device__ inline void pair_term(Acc<EV> &a, const float4 &pi, const float4 &vi, float rho_i, const float4 &q,
                                 const float4 &vj, float rho_j, uint32_t code, const PairC &c, float dtis,
                                 const SysDev *d)
{
    float dx, dy, dz;
    pair_delta(pi, q, code, d, dx, dy, dz);
    const float rsq = dist2(dx, dy, dz);
    const bool in = !TEST || (rsq < c.rc2 && rsq >= 1e-20f);
    const float rinv = __builtin_amdgcn_rsqf(rsq);
    const float r = rsq * rinv;
    const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
    const float dot = dx * dvx + dy * dvy + dz * dvz;
    const float wc = 1.f - r * c.rcinv;
    const float xi = pair_gaussian(__float_as_uint(vi.w), __float_as_uint(vj.w));
    // sigma / sqrt(dt): uniform for a one-type system away from walls
    const float sdt = c.sigma * dtis;
    float fpair, en = 0.f;
    if (STYLE == STYLE_MDPD) {
        // pair_mdpd_meso.cu:156-191: F = [A wc + B (rho_i + rho_j) wr
        //   - gamma wc^2 (dx.dv)/r + sigma wc xi / sqrt(dt)] / r, evaluated as
        // [wc (A + sigma' xi - gamma wc (dx.dv)/r) + B (rho_i + rho_j) wr] / r
        const float wcr = fmaxf(1.f - r * c.rdinv, 0.f);
        const float rr = rho_i + rho_j;
        float t = fmaf(sdt, xi, c.a);
        t = fmaf(-(c.gamma * wc * rinv), dot, t);
        fpair = fmaf(c.b * rr, wcr, t * wc) * rinv;
        if (EV) en = 0.5f * c.a * c.rc * wc * wc + 0.5f * c.b * c.rd * rr * wcr * wcr;
    } else {
        // pair_dpd_meso.cu:145-175: F = [a0 wc - gamma wr^2 (dx.dv)/r
        //   + sigma wr xi / sqrt(dt)] / r, evaluated as
        // [a0 wc + wr (sigma' xi - gamma wr (dx.dv)/r)] / r
        const float wr = c.expw == 1.f ? wc : powf(wc, c.expw);
        const float t = fmaf(-(c.gamma * wr * rinv), dot, sdt * xi);
        fpair = fmaf(c.a, wc, wr * t) * rinv;
        if (EV) en = 0.5f * c.a * c.rc * wc * wc;
    }
    fpair = in ? fpair : 0.f;
    a.v[0] += dx * fpair; a.v[1] += dy * fpair; a.v[2] += dz * fpair;
    if (EV) {
        a.v[3] += in ? en : 0.f;
        a.v[4] += dx * dx * fpair; a.v[5] += dy * dy * fpair; a.v[6] += dz * dz * fpair;
        a.v[7] += dx * dy * fpair; a.v[8] += dx * dz * fpair; a.v[9] += dy * dz * fpair;
    }
}

template <int STYLE, int MODE, bool UNI, int SPLIT>
__global__ void __launch_bounds__(WAVE * SPLIT) k_force(PoolDev P, int tile0, int ntiles)
{
    // HARVEST re-evaluates the last pass of a finished pooled system from the
    // state that pass read, which its commit left in the other buffer; it
    // adds energy and virial and changes no state
    constexpr bool EV = MODE != MODE_FUSED && MODE != MODE_PHASED;
    constexpr bool STEP = MODE != MODE_SETUP && MODE != MODE_HARVEST;
    constexpr int NA = Acc<EV>::N;
    __shared__ float s_acc[SPLIT > 1 ? SPLIT - 1 : 1][NA][WAVE];
    // lane-balanced pair loop (see below): one wave per tile, one type
    constexpr bool BAL = SPLIT == 1 && UNI;
    constexpr int TAIL_CAP = BAL ? 192 : 1;
    __shared__ float4 s_pi[BAL ? WAVE : 1], s_vi[BAL ? WAVE : 1];
    __shared__ float s_ri[BAL ? WAVE : 1];
    __shared__ float s_tail[TAIL_CAP * NA];
    const int lt = xcd_logical(blockIdx.x, ntiles);
    if (lt >= ntiles) return;
    const int tile = tile0 + lt;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = tile * WAVE + lane;
    const int s = P.wave_sys[tile];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active && MODE != MODE_HARVEST) return;
    if (UNI != (d->ntypes == 1)) return;     // host launches the matching variant
    // a system waiting for wider neighbor rows is skipped by the stepping
    // passes (no state change until the host has regrown the rows), and a
    // pool may mix styles: each launch takes the systems of its style
    if (STEP && d->overflow) return;
    if (MODE == MODE_PHASED && d->style != STYLE) return;
    if (MODE == MODE_HARVEST && d->split_last != SPLIT) return;   // the split of its last pass
    const bool valid = i < d->part_base + d->n;
    const int cur = MODE == MODE_HARVEST ? d->sel ^ 1 : d->sel;
    const float4 *pos = P.pos[cur];
    const float4 *vel = P.vel[cur];
    const int *maskv = P.mask[d->sela];
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    float rho_i = 0.f, phi_i = 0.f;
    int ti = 0, cnt = 0, mk = 0;
    if (valid) {
        pi = pos[i];
        vi = vel[i];
        ti = __float_as_int(pi.w);
        mk = maskv[i];
        if (STYLE == STYLE_MDPD) rho_i = P.rho[i];
        cnt = P.ccount[ccount_at(tile, w, lane)];
    }
    // arbitrary boundary: a mobile particle scales gamma / sigma of its pairs
    // with wall particles by its wall fraction (pair_mdpd_meso.cu:169-185)
    const bool arb_i = STYLE == STYLE_MDPD && d->pair_arb && (mk & d->pair_mobile_bit);
    if (arb_i) phi_i = P.phi[d->sela][i];
    const int wall_bit = d->pair_wall_bit;
    auto wall_scale = [&](PairC &c, uint32_t j) {
        if (arb_i && (maskv[j] & wall_bit)) {
            const float ratio = wall_ratio(phi_i, c.rc);
            c.gamma *= ratio;
            c.sigma *= __builtin_amdgcn_sqrtf(ratio);
        }
    };
    const float dtis = d->dtinvsqrt;
    const float *coef = P.coef + d->coef_base;
    const int nt = d->ntypes;
    PairC cu;
    if (UNI) cu = load_coef(coef, 0);
    const int maxnbr = P.maxnbr;
    const int *crow = P.clist + (size_t)tile * maxnbr * WAVE + lane;
    Acc<EV> a;
    // pairs [t, end) of this lane's compact row, four gathers in flight (one
    // type) or two
    // neighbor j's position (q.w: its type bits, used by several types) and
    // density; one type of mDPD takes both from the first pass's x, y, z,
    // rho record with one 16-byte load
    auto nbr = [&](uint32_t j, float4 &q, float &rj) {
        if (UNI && STYLE == STYLE_MDPD) {
            q = P.pr[j];
            rj = q.w;
        } else {
            q = pos[j];
            rj = STYLE == STYLE_MDPD ? P.rho[j] : 0.f;
        }
    };
    auto walk = [&](int t, int end) {
        if (UNI) {
            for (; t + 4 <= end; t += 4) {
                uint32_t e[4], j[4];
                float4 q[4], vv[4];
                float rj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    e[u] = (uint32_t)crow[(size_t)((t + u) * SPLIT + w) * WAVE];
                    j[u] = e[u] & IDX_MASK;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    nbr(j[u], q[u], rj[u]);
                    vv[u] = vel[j[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    PairC c = cu;
                    wall_scale(c, j[u]);
                    pair_term<STYLE, EV, !UNI>(a, pi, vi, rho_i, q[u], vv[u], rj[u], e[u] >> IMG_SHIFT, c, dtis, d);
                }
            }
        }
        for (; t + 2 <= end; t += 2) {
            const uint32_t e0 = (uint32_t)crow[(size_t)(t * SPLIT + w) * WAVE];
            const uint32_t e1 = (uint32_t)crow[(size_t)((t + 1) * SPLIT + w) * WAVE];
            const uint32_t j0 = e0 & IDX_MASK, j1 = e1 & IDX_MASK;
            float4 q0, q1;
            float r0, r1;
            nbr(j0, q0, r0);
            nbr(j1, q1, r1);
            const float4 w0 = vel[j0], w1 = vel[j1];
            PairC c0 = UNI ? cu : load_coef(coef, ti * nt + __float_as_int(q0.w));
            PairC c1 = UNI ? cu : load_coef(coef, ti * nt + __float_as_int(q1.w));
            wall_scale(c0, j0);
            wall_scale(c1, j1);
            pair_term<STYLE, EV, !UNI>(a, pi, vi, rho_i, q0, w0, r0, e0 >> IMG_SHIFT, c0, dtis, d);
            pair_term<STYLE, EV, !UNI>(a, pi, vi, rho_i, q1, w1, r1, e1 >> IMG_SHIFT, c1, dtis, d);
        }
        if (t < end) {
            const uint32_t e0 = (uint32_t)crow[(size_t)(t * SPLIT + w) * WAVE];
            const uint32_t j0 = e0 & IDX_MASK;
            float4 q0;
            float r0;
            nbr(j0, q0, r0);
            PairC c0 = UNI ? cu : load_coef(coef, ti * nt + __float_as_int(q0.w));
            wall_scale(c0, j0);
            pair_term<STYLE, EV, !UNI>(a, pi, vi, rho_i, q0, vel[j0], r0, e0 >> IMG_SHIFT, c0, dtis, d);
        }
    };
    // Lane balance (one wave per tile, one type, no wall scaling): a wave
    // runs as long as its longest row, ~1.5x its mean row.  The first M
    // entries of every row (M: the mean row, raised until the rest fits
    // TAIL_CAP) are walked by their own lane; the entries past M, the tails,
    // laid out one after another in particle order, are dealt to all 64
    // lanes.  A tail pair's terms go to the LDS slot of its position, and
    // each particle adds its slots, in row order, after its own M entries.
    bool balanced = false;
    if constexpr (BAL) balanced = !(STYLE == STYLE_MDPD && d->pair_arb);
    if (balanced) {
        const int tot = wave_sum(cnt);
        int M = tot >> 6;
        int tail = wave_sum(max(cnt - M, 0));
        for (int it = 0; tail > TAIL_CAP && it < 4; ++it) {
            M += (tail - TAIL_CAP + WAVE - 1) / WAVE;
            tail = wave_sum(max(cnt - M, 0));
        }
        if (tail > TAIL_CAP) {          // rows too uneven: the longest row sets M
            M = cnt;
            for (int off = 32; off > 0; off >>= 1) M = max(M, __shfl_xor(M, off, 64));
            tail = 0;
        }
        walk(0, min(cnt, M));
        if (tail > 0) {
            const int sur = max(cnt - M, 0);
            const int first = wave_incl_scan(sur, lane) - sur;     // this particle's first tail slot
            s_pi[lane] = pi;
            s_vi[lane] = vi;
            s_ri[lane] = rho_i;
            __syncthreads();
            const int *ctile = P.clist + (size_t)tile * maxnbr * WAVE;
            for (int f0 = 0; f0 < tail; f0 += WAVE) {
                const int f = f0 + lane;
                // owner: the last particle whose first slot is <= f (the
                // particles before it with no tail share its first slot)
                int p = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1)
                    if (__shfl(first, p + step) <= f) p += step;
                const int t = M + f - __shfl(first, p);
                const bool act = f < tail;
                // a lane past the last tail evaluates its own particle with
                // itself (r = 0: terms not finite) and stores nothing
                const uint32_t e = act ? (uint32_t)ctile[(size_t)t * WAVE + p]
                                       : ((uint32_t)i | ((uint32_t)IMG_ZERO << IMG_SHIFT));
                const uint32_t j = e & IDX_MASK;
                float4 q;
                float rj;
                nbr(j, q, rj);
                Acc<EV> b;
                pair_term<STYLE, EV, !UNI>(b, s_pi[p], s_vi[p], s_ri[p], q, vel[j], rj, e >> IMG_SHIFT, cu, dtis, d);
                if (act)
                    for (int k = 0; k < NA; ++k) s_tail[k * TAIL_CAP + f] = b.v[k];
            }
            __syncthreads();
            for (int f = first; f < first + sur; ++f)
                for (int k = 0; k < NA; ++k) a.v[k] += s_tail[k * TAIL_CAP + f];
        }
    } else {
        walk(0, cnt);
    }
    if (SPLIT > 1) {
        if (w > 0)
            for (int k = 0; k < NA; ++k) s_acc[w - 1][k][lane] = a.v[k];
        __syncthreads();
        if (w == 0)
            for (int q = 0; q < SPLIT - 1; ++q)
                for (int k = 0; k < NA; ++k) a.v[k] += s_acc[q][k][lane];
    }
    int moved = 0;
    if (w == 0 && valid) {
        float fx = a.v[0], fy = a.v[1], fz = a.v[2];
        if (mk & d->add_bit) { fx += d->add[0]; fy += d->add[1]; fz += d->add[2]; }
        if (EV) {
            P.frc[i] = make_float4(fx, fy, fz, 0.5f * a.v[3]);
            const size_t cap = (size_t)P.cap;
            P.vir[i] = 0.5f * a.v[4]; P.vir[cap + i] = 0.5f * a.v[5]; P.vir[2 * cap + i] = 0.5f * a.v[6];
            P.vir[3 * cap + i] = 0.5f * a.v[7]; P.vir[4 * cap + i] = 0.5f * a.v[8]; P.vir[5 * cap + i] = 0.5f * a.v[9];
        }
        if (STEP) {
            // which half-kicks this pass applies: the pooled engine decides per
            // system, the plain engine per launch
            const int ph = MODE == MODE_PHASED ? d->phase : (MODE == MODE_FINAL ? (int)PH_LAST : (int)PH_RUN);
            float vx = vi.x, vy = vi.y, vz = vi.z;
            float4 pn = pi;
            if ((mk & d->nve_bit) && ph != PH_FRESH) {
                const float dtfm = d->dtf * P.minv[d->mass_base + ti];
                vx += dtfm * fx; vy += dtfm * fy; vz += dtfm * fz;             // final_integrate
            }
            if (ph == PH_LAST) {
                P.pos[cur ^ 1][i] = pi;
                P.vel[cur ^ 1][i] = make_float4(vx, vy, vz, vi.w);
                P.vend[i] = make_float4(vx, vy, vz, 0.f);
                if (MODE == MODE_PHASED) P.frc[i] = make_float4(fx, fy, fz, 0.f);
            } else {
                if (MODE == MODE_FUSED_OUT) P.vend[i] = make_float4(vx, vy, vz, 0.f);
                if (mk & d->nve_bit) {
                    const float dtfm = d->dtf * P.minv[d->mass_base + ti];
                    vx += dtfm * fx; vy += dtfm * fy; vz += dtfm * fz;         // next initial_integrate
                    pn.x += d->dt * vx; pn.y += d->dt * vy; pn.z += d->dt * vz;
                }
                const uint32_t sg = next_signature(d, P.tag[d->sela][i], vx, vy, vz);
                P.pos[cur ^ 1][i] = pn;
                P.vel[cur ^ 1][i] = make_float4(vx, vy, vz, __uint_as_float(sg));
                const float4 h = P.xhold[i];
                const float ex = pn.x - h.x, ey = pn.y - h.y, ez = pn.z - h.z;
                moved = ex * ex + ey * ey + ez * ez > d->trig2;
            }
        }
    }
    if (MODE == MODE_FUSED || MODE == MODE_FUSED_OUT || MODE == MODE_PHASED) tile_flags(P.moved, tile, 1, moved);
}

// --------------------------------------------------------------------------
// Wall branch of aboundary/rho/meso after the initial integrate
// (FixArbitraryBoundaryRho::post_integrate / gpu_aboundary,
// fix_aboundary_rho_meso.cu:220-317): wall fraction phi_i and wall normal of
// each mobile particle from the wall particles within cut_phi (every entry of
// its neighbor row, positions after the drift), bounce-back when
// phi_i > phi_c (v <- -v + 2 max(0, v.n) n, x <- x - v dtf/m + dt v_new),
// then the tile's displacement flag from the final positions.  Works on the
// buffer the pass just wrote (sel ^ 1, before the step commit); one wave per
// 64-particle tile.  `pooled`: apply only to systems whose pass integrated.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(WAVE) k_aboundary(PoolDev P, int ntiles, int pooled)
{
    const int tile = xcd_logical(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int s = P.wave_sys[tile];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || !d->fix_arb || d->overflow) return;
    if (pooled && d->phase == PH_LAST) return;
    const int lane = threadIdx.x;
    const int i = tile * WAVE + lane;
    const bool valid = i < d->part_base + d->n;
    const int sa = d->sela;
    float4 *pos = P.pos[d->sel ^ 1];
    float4 *vel = P.vel[d->sel ^ 1];
    const int *mask = P.mask[sa];
    int moved = 0;
    if (valid) {
        float4 p = pos[i];
        // the rows do not change between builds: the first wall pass after a
        // build records which of them hold a wall particle, and the later
        // passes skip the others (no wall neighbor: phi stays 0, no bounce)
        const bool fresh = d->ago == 0;
        if ((mask[i] & d->fix_mobile_bit) && (fresh || P.wall_near[i])) {
            const int maxnbr = P.maxnbr;
            const int cnt = min(P.ncount[i], maxnbr);
            int near = 0;
            const int4 *row4 = (const int4 *)(P.nlist + nbr_at(i, 0, maxnbr));
            const float c = d->cut_phi, c2 = c * c;
            const int wbit = d->fix_wall_bit;
            float phi = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
            // the row in chunks of eight entries: the chunk's eight masks in
            // flight, then the positions of its wall neighbors, the next
            // chunk's entries loading meanwhile; terms added in row order
            const int nch = (cnt + NBR_CHUNK - 1) / NBR_CHUNK;
            int4 ea = make_int4(0, 0, 0, 0), eb = ea;
            if (nch > 0) {
                ea = row4[0];
                eb = row4[1];
            }
            for (int ch = 0; ch < nch; ++ch) {
                const uint32_t e[NBR_CHUNK] = {(uint32_t)ea.x, (uint32_t)ea.y, (uint32_t)ea.z, (uint32_t)ea.w,
                                               (uint32_t)eb.x, (uint32_t)eb.y, (uint32_t)eb.z, (uint32_t)eb.w};
                if (ch + 1 < nch) {
                    ea = row4[(size_t)(ch + 1) * (NBR_CHUNK * WAVE / 4)];
                    eb = row4[(size_t)(ch + 1) * (NBR_CHUNK * WAVE / 4) + 1];
                }
                bool wall[NBR_CHUNK];
#pragma unroll
                for (int k = 0; k < NBR_CHUNK; ++k) {
                    const bool in = ch * NBR_CHUNK + k < cnt;
                    wall[k] = in && (mask[(in ? e[k] : (uint32_t)i) & IDX_MASK] & wbit);
                    near |= wall[k] ? 1 : 0;
                }
                float4 q[NBR_CHUNK];
#pragma unroll
                for (int k = 0; k < NBR_CHUNK; ++k)
                    if (wall[k]) q[k] = pos[e[k] & IDX_MASK];
#pragma unroll
                for (int k = 0; k < NBR_CHUNK; ++k) {
                    if (!wall[k]) continue;
                    float dx, dy, dz;
                    pair_delta(p, q[k], e[k] >> IMG_SHIFT, d, dx, dy, dz);
                    const float rsq = dx * dx + dy * dy + dz * dz;
                    if (!(rsq < c2)) continue;
                    const float rinv = __builtin_amdgcn_rsqf(rsq);
                    const float r = fmaxf(rsq * rinv, 1e-10f);
                    const float dd = c - r;
                    phi += (c + 3.f * r) * dd * dd * dd * d->phi_factor;
                    const float g = r * dd * dd * d->dw_factor * rinv;
                    nx -= g * dx; ny -= g * dy; nz -= g * dz;
                }
            }
            if (fresh) P.wall_near[i] = near;
            P.phi[sa][i] = phi;
            if (phi > d->phi_c) {
                const float inv = __builtin_amdgcn_rsqf(nx * nx + ny * ny + nz * nz);
                nx *= inv; ny *= inv; nz *= inv;
                float4 v = vel[i];
                const float dtfm = d->dtf * P.minv[d->mass_base + __float_as_int(p.w)];
                const float ox = p.x - v.x * dtfm, oy = p.y - v.y * dtfm, oz = p.z - v.z * dtfm;
                const float tmp = fmaxf(0.f, v.x * nx + v.y * ny + v.z * nz);
                v.x = -v.x + 2.f * tmp * nx;
                v.y = -v.y + 2.f * tmp * ny;
                v.z = -v.z + 2.f * tmp * nz;
                p.x = ox + d->dt * v.x;
                p.y = oy + d->dt * v.y;
                p.z = oz + d->dt * v.z;
                if (d->style == STYLE_DPD)      // the DPD signature reads v
                    v.w = __uint_as_float(sig_dpd(P.tag[sa][i], d->seed_next, v.x, v.y, v.z));
                pos[i] = p;
                vel[i] = v;
            }
        }
        const float4 h = P.xhold[i];
        const float ex = p.x - h.x, ey = p.y - h.y, ez = p.z - h.z;
        moved = ex * ex + ey * ey + ez * ez > d->trig2;
    }
    const uint64_t any = __ballot(moved);
    if (lane == 0) P.moved[tile] = any != 0 ? 1 : 0;
}

// --------------------------------------------------------------------------
// Step commit, one workgroup per system, after a pass that wrote the other
// buffer: ADVANCE (initial kick or fused step) flips the buffer, advances the
// step, precomputes the next TEA seed and ORs the tiles' moved flags into
// `stale`; FINAL (end-of-run half-kick) only flips the buffer.  DECIDE (plain
// engine) and PHASED (pooled engine) make the reneighbor decision of
// Neighbor::decide (src/neighbor.cpp:1216-1290) on the device: after `ago`
// steps since the last build, rebuild when ago >= delay, ago % every == 0 and
// (no check, or a particle moved more than skin/2 since the last build);
// PHASED also follows each system's lifecycle.  A system whose rows
// overflowed was skipped by the pass and is not advanced.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_commit_step(PoolDev P, int advance, int split)
{
    __shared__ int s_any;
    SysDev *d = P.sys + blockIdx.x;
    if (!d->active) return;
    if (d->overflow) {
        // skipped by the passes: no state change until the rows are regrown
        if (advance == COMMIT_PHASED && threadIdx.x == 0)
            P.status[blockIdx.x] = (d->job << 3) | (d->stale ? 1 : 0) | 4;
        return;
    }
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    if (advance) {
        const int t0 = d->part_base / WAVE, nt = d->npad / WAVE;
        int any = 0;
        // four tile flags in flight per thread, 1,024 threads: a 200k-particle
        // system's ~3,000 tiles in one round (0.4% on the 193,600-particle slab)
        for (int t = threadIdx.x; t < nt; t += 4 * (int)blockDim.x) {
            int f[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = t + u * (int)blockDim.x;
                f[u] = k < nt ? P.moved[t0 + k] : 0;
            }
            any |= f[0] | f[1] | f[2] | f[3];
        }
        if (any) s_any = 1;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (advance == COMMIT_PHASED) {
            // pooled engine: per-system lifecycle
            d->split_last = split;
            if (d->phase == PH_LAST) {
                d->sel ^= 1;
                d->phase = PH_DONE;
                d->active = 0;
            } else {
                commit_advance(d, false);
                const int ago = ++d->ago;
                const bool due = ago >= d->n_delay && ago % d->n_every == 0;
                d->stale = due && (!d->n_check || s_any != 0) ? 1 : 0;
                d->steps_left -= 1;
                d->phase = d->steps_left <= 0 ? PH_LAST : PH_RUN;
            }
            P.status[blockIdx.x] = (d->job << 3) | (d->stale ? 1 : 0) | (d->phase == PH_DONE ? 2 : 0);
        } else if (advance == COMMIT_DECIDE) {
            commit_advance(d, false);
            const int ago = ++d->ago;
            const bool due = ago >= d->n_delay && ago % d->n_every == 0;
            d->stale = due && (!d->n_check || s_any != 0) ? 1 : 0;
        } else if (advance == COMMIT_ADVANCE) {
            commit_advance(d, s_any != 0);
        } else {
            d->sel ^= 1;
        }
    }
}

// --------------------------------------------------------------------------
// Finished pooled system: its current state, the last force evaluation and
// its descriptor packed into one block at dst (x, v, f | tag, mask, origin,
// image | rho, virial | SysDev, each array npad long), so the host copies one
// range.  Runs after the HARVEST force pass, which reads the other buffer.
// The system's tiles and cell slots are unmapped here (no kernel after it
// finds the system through them); the host frees the ranges once the copy
// has landed.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_harvest(PoolDev P, int s, int pb, int npad, char *dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const SysDev *d = P.sys + s;
    if (blockIdx.x == 0) {
        const int *src = (const int *)d;
        int *o = (int *)(dst + (size_t)npad * HARVEST_SLOT_BYTES);
        for (int w = threadIdx.x; w < (int)(sizeof(SysDev) / 4); w += blockDim.x) o[w] = src[w];
    }
    for (int c = k, nc = d->ncell + 1; c < nc; c += gridDim.x * blockDim.x) P.cell_sys[d->cell_base + c] = -1;
    if (k < npad / WAVE) P.wave_sys[pb / WAVE + k] = -1;
    if (k >= npad) return;
    const int i = pb + k, cur = d->sel, cura = d->sela;
    const size_t n = npad;
    float4 *o4 = (float4 *)dst;
    o4[k] = P.pos[cur][i];
    o4[n + k] = P.vel[cur][i];
    o4[2 * n + k] = P.frc[i];
    int *oi = (int *)(o4 + 3 * n);
    oi[k] = P.tag[cura][i];
    oi[n + k] = P.mask[cura][i];
    oi[2 * n + k] = P.origin[cura][i];
    oi[3 * n + k] = P.image[cura][i];
    float *of = (float *)(oi + 4 * n);
    of[k] = P.rho[i];
#pragma unroll
    for (int c = 0; c < 6; ++c) of[(c + 1) * n + k] = P.vir[(size_t)c * P.cap + i];
}

// --------------------------------------------------------------------------
// stateless helpers (parity tests)
// --------------------------------------------------------------------------
__global__ void k_pair_gaussian(const uint32_t *a, const uint32_t *b, float *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = pair_gaussian(a[i], b[i]);
}

__global__ void k_premix(const uint32_t *a, const uint32_t *b, uint32_t *out, int n, int rounds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v0 = a[i], v1 = b[i];
    tea_rounds(v0, v1, rounds);
    out[i] = v0 ^ v1;
}

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
template <int SPLIT>
static void launch_first_t(const PoolDev &P, bool with_rho, int nt, hipStream_t st)
{
    if (with_rho)
        hipLaunchKernelGGL((k_first<true, SPLIT>), dim3(xcd_grid(nt)), dim3(WAVE * SPLIT), 0, st, P, nt);
    else
        hipLaunchKernelGGL((k_first<false, SPLIT>), dim3(xcd_grid(nt)), dim3(WAVE * SPLIT), 0, st, P, nt);
}

hipError_t launch_first(const PoolDev &P, bool with_rho, int split, hipStream_t st)
{
    const int nt = P.nact / WAVE;
    if (nt == 0) return hipSuccess;
    if (P.maxnbr % NBR_ALIGN) return hipErrorInvalidValue;   // a split wave's compact rows must fit
    switch (split) {
    case 1: launch_first_t<1>(P, with_rho, nt, st); break;
    case 2: launch_first_t<2>(P, with_rho, nt, st); break;
    case 4: launch_first_t<4>(P, with_rho, nt, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_commit_step(const PoolDev &P, int kind, int split, hipStream_t st)
{
    if (P.nsys == 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit_step, dim3(P.nsys), dim3(1024), 0, st, P, kind, split);
    return hipGetLastError();
}

hipError_t launch_aboundary(const PoolDev &P, int pooled, hipStream_t st)
{
    const int nt = P.nact / WAVE;
    if (nt == 0) return hipSuccess;
    hipLaunchKernelGGL(k_aboundary, dim3(xcd_grid(nt)), dim3(WAVE), 0, st, P, nt, pooled);
    return hipGetLastError();
}

hipError_t launch_harvest(const PoolDev &P, int s, int part_base, int npad, void *dst, size_t room, hipStream_t st)
{
    if (npad <= 0) return hipSuccess;
    if (s < 0 || s >= P.nsys || part_base < 0 || part_base + npad > P.cap || !dst || harvest_bytes(npad) > room)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_harvest, dim3((npad + 255) / 256), dim3(256), 0, st, P, s, part_base, npad, (char *)dst);
    return hipGetLastError();
}

hipError_t launch_initial(const PoolDev &P, hipStream_t st)
{
    const int nl = P.nact / BLK;
    if (nl == 0) return hipSuccess;
    hipLaunchKernelGGL(k_initial, dim3(xcd_grid(nl)), dim3(BLK), 0, st, P, nl);
    return hipGetLastError();
}

template <int STYLE, int MODE, int SPLIT>
static void launch_force_t(const PoolDev &P, bool uni, int t0, int nt, hipStream_t st)
{
    if (uni)
        hipLaunchKernelGGL((k_force<STYLE, MODE, true, SPLIT>), dim3(xcd_grid(nt)), dim3(WAVE * SPLIT), 0, st, P, t0, nt);
    else
        hipLaunchKernelGGL((k_force<STYLE, MODE, false, SPLIT>), dim3(xcd_grid(nt)), dim3(WAVE * SPLIT), 0, st, P, t0, nt);
}

template <int STYLE, int MODE>
static hipError_t launch_force_s(const PoolDev &P, bool uni, int split, int t0, int nt, hipStream_t st)
{
    switch (split) {
    case 1: launch_force_t<STYLE, MODE, 1>(P, uni, t0, nt, st); break;
    case 2: launch_force_t<STYLE, MODE, 2>(P, uni, t0, nt, st); break;
    case 4: launch_force_t<STYLE, MODE, 4>(P, uni, t0, nt, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_force(const PoolDev &P, int style, int mode, bool uni, int split, hipStream_t st)
{
    return launch_force_range(P, style, mode, uni, split, 0, P.nact / WAVE, st);
}

hipError_t launch_force_range(const PoolDev &P, int style, int mode, bool uni, int split, int t0, int nt,
                              hipStream_t st)
{
    if (nt <= 0) return hipSuccess;
    if (t0 < 0 || t0 + nt > P.cap / WAVE) return hipErrorInvalidValue;
#define MDPD_FORCE_CASE(S, M) \
    if (style == S && mode == M) return launch_force_s<S, M>(P, uni, split, t0, nt, st);
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_SETUP)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FUSED)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FUSED_OUT)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FINAL)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_PHASED)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_HARVEST)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_SETUP)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FUSED)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FUSED_OUT)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FINAL)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_PHASED)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_HARVEST)
#undef MDPD_FORCE_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_pair_gaussian(const uint32_t *a, const uint32_t *b, float *out, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_gaussian, dim3((n + 255) / 256), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_premix(const uint32_t *a, const uint32_t *b, uint32_t *out, int n, int rounds, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_premix, dim3((n + 255) / 256), dim3(256), 0, st, a, b, out, n, rounds);
    return hipGetLastError();
}

}  // namespace mdpd
