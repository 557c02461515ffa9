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
// Launch shape: one thread per particle, 256-thread workgroups; each system
// occupies whole workgroups (systems are padded to 256 slots), and workgroup
// ids are remapped so each of the 8 XCDs sweeps one contiguous particle range.
#include "common.hpp"
#include "kernels.hpp"

namespace mdpd {

constexpr int BLK = 256;
constexpr int NXCD = 8;

// hardware dispatches workgroup b to XCD b % 8; give XCD x the contiguous
// logical range [x*per, (x+1)*per)
__device__ inline int xcd_logical(int b, int nlogical)
{
    const int per = (nlogical + NXCD - 1) / NXCD;
    return (b % NXCD) * per + b / NXCD;
}

static inline int xcd_grid(int nlogical) { return ((nlogical + NXCD - 1) / NXCD) * NXCD; }

// Last-workgroup commit.  Every workgroup of a system calls this once after
// finishing all of its reads and writes; the last one to arrive applies
// `commit` with no other workgroup of that system still running.
template <typename F>
__device__ inline void block_commit(SysDev *d, int pending_or, F commit)
{
    __shared__ int s_or;
    if (threadIdx.x == 0) s_or = 0;
    __syncthreads();
    if (pending_or) atomicOr(&s_or, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_or) atomicOr(&d->stale_pending, 1);
        __threadfence();
        const int nblk = d->npad / BLK;
        const int old = atomicAdd(&d->done, 1);
        if (old == nblk - 1) {
            __threadfence();
            commit();
            d->done = 0;
            __threadfence();
        }
    }
}

__device__ inline void apply_image(uint32_t code, const SysDev *d, float &x, float &y, float &z)
{
    int sx, sy, sz;
    image_decode(code, sx, sy, sz);
    x += (float)sx * d->L[0];
    y += (float)sy * d->L[1];
    z += (float)sz * d->L[2];
}

// --------------------------------------------------------------------------
// density pass: rho_i = m_i * sum_j w(r_ij), 1e-20 <= r^2 < rd^2, no self
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_density(PoolDev P, int nlogical)
{
    const int lb = xcd_logical(blockIdx.x, nlogical);
    if (lb >= nlogical) return;
    const int i = lb * BLK + threadIdx.x;
    const int s = P.wave_sys[i >> 6];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || i >= d->part_base + d->n) return;
    const int cur = d->sel;
    const float4 *pos = P.pos[cur];
    const float4 pi = pos[i];
    const int mk = P.mask[d->sela][i];
    if (!(mk & d->rho_bit)) { P.rho[i] = 0.f; return; }
    const int maxnbr = P.maxnbr;
    const int cnt = min(P.ncount[i], maxnbr);
    const int *row = P.nlist + (size_t)(i >> 6) * maxnbr * WAVE + (i & 63);
    const float h = d->rd, h2 = h * h;
    float sum = 0.f;
    if (d->lucy) {
        for (int p = 0; p < cnt; ++p) {
            const uint32_t e = (uint32_t)row[(size_t)p * WAVE];
            const uint32_t j = e & IDX_MASK, code = e >> IMG_SHIFT;
            const float4 q = pos[j];
            float xj = q.x, yj = q.y, zj = q.z;
            if (code != IMG_ZERO) apply_image(code, d, xj, yj, zj);
            const float dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
            const float rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < h2 && rsq >= 1e-20f) {
                const float r = __fsqrt_rn(rsq);
                const float hm = h - r;
                sum += (h + 3.f * r) * hm * hm * hm;
            }
        }
    } else {
        for (int p = 0; p < cnt; ++p) {
            const uint32_t e = (uint32_t)row[(size_t)p * WAVE];
            const uint32_t j = e & IDX_MASK, code = e >> IMG_SHIFT;
            const float4 q = pos[j];
            float xj = q.x, yj = q.y, zj = q.z;
            if (code != IMG_ZERO) apply_image(code, d, xj, yj, zj);
            const float dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
            const float rsq = dx * dx + dy * dy + dz * dz;
            if (rsq < h2 && rsq >= 1e-20f) {
                const float r = __fsqrt_rn(rsq);
                const float hm = h - r;
                sum += hm * hm;
            }
        }
    }
    const int ti = __float_as_int(pi.w);
    P.rho[i] = P.mass[d->mass_base + ti] * d->rho_factor * sum;
}

// --------------------------------------------------------------------------
// initial half-kick + drift from the stored force (first step of a run)
// --------------------------------------------------------------------------
__device__ inline uint32_t next_signature(const SysDev *d, int tag, float vx, float vy, float vz)
{
    return d->style == STYLE_MDPD ? sig_mdpd(tag, d->seed_next) : sig_dpd(tag, d->seed_next, vx, vy, vz);
}

__device__ inline void commit_advance(SysDev *d)
{
    d->sel ^= 1;
    d->step += 1;
    d->seed_next = premix<64>(d->seed, (uint32_t)(d->step + 1));
    d->stale = d->stale_pending;
    d->stale_pending = 0;
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
    block_commit(d, moved, [&]() { commit_advance(d); });
}

// --------------------------------------------------------------------------
// pair force (+ fused velocity-Verlet)
// --------------------------------------------------------------------------
struct Pair6 {
    float rc2, rcinv, rdinv, a, b, gamma, sigma, expw;
};

__device__ inline Pair6 load_coef(const float *tab, int idx)
{
    const float *c = tab + idx * NCOEF;
    Pair6 q;
    q.rc2 = c[C_RC2]; q.rcinv = c[C_RCINV]; q.rdinv = c[C_RDINV];
    q.a = c[C_A]; q.b = c[C_B]; q.gamma = c[C_GAMMA]; q.sigma = c[C_SIGMA]; q.expw = c[C_EXPW];
    return q;
}

template <int STYLE, int MODE, bool UNI>
__global__ void __launch_bounds__(BLK) k_force(PoolDev P, int nlogical)
{
    constexpr bool EV = MODE != MODE_FUSED;
    const int lb = xcd_logical(blockIdx.x, nlogical);
    if (lb >= nlogical) return;
    const int i = lb * BLK + threadIdx.x;
    const int s = P.wave_sys[(lb * BLK) >> 6];
    if (s < 0) return;
    SysDev *d = P.sys + s;
    if (!d->active) return;
    if (UNI != (d->ntypes == 1)) return;     // host launches the matching variant
    int moved = 0;
    if (i < d->part_base + d->n) {
        const int cur = d->sel;
        const float4 *pos = P.pos[cur];
        const float4 *vel = P.vel[cur];
        const float4 pi = pos[i];
        const float4 vi = vel[i];
        const uint32_t sig_i = __float_as_uint(vi.w);
        const int ti = __float_as_int(pi.w);
        const float rho_i = STYLE == STYLE_MDPD ? P.rho[i] : 0.f;
        const float dtis = d->dtinvsqrt;
        const float *coef = P.coef + d->coef_base;
        const int nt = d->ntypes;
        Pair6 cu;
        if (UNI) cu = load_coef(coef, 0);
        const int maxnbr = P.maxnbr;
        const int cnt = min(P.ncount[i], maxnbr);
        const int *row = P.nlist + (size_t)(i >> 6) * maxnbr * WAVE + (i & 63);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        float en = 0.f, vxx = 0.f, vyy = 0.f, vzz = 0.f, vxy = 0.f, vxz = 0.f, vyz = 0.f;
        for (int p = 0; p < cnt; ++p) {
            const uint32_t e = (uint32_t)row[(size_t)p * WAVE];
            const uint32_t j = e & IDX_MASK, code = e >> IMG_SHIFT;
            const float4 q = pos[j];
            float xj = q.x, yj = q.y, zj = q.z;
            if (code != IMG_ZERO) apply_image(code, d, xj, yj, zj);
            const float dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
            const float rsq = dx * dx + dy * dy + dz * dz;
            const Pair6 c = UNI ? cu : load_coef(coef, ti * nt + __float_as_int(q.w));
            if (rsq < c.rc2 && rsq >= 1e-20f) {
                const float4 vj = vel[j];
                const float rinv = __builtin_amdgcn_rsqf(rsq);
                const float r = rsq * rinv;
                const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
                const float dot = dx * dvx + dy * dvy + dz * dvz;
                const float wc = 1.f - r * c.rcinv;
                const float xi = pair_gaussian(sig_i, __float_as_uint(vj.w));
                float fpair;
                if (STYLE == STYLE_MDPD) {
                    // pair_mdpd_meso.cu:156-191
                    const float wcr = fmaxf(1.f - r * c.rdinv, 0.f);
                    const float rr = rho_i + P.rho[j];
                    fpair = (c.a * wc + c.b * rr * wcr) - c.gamma * wc * wc * dot * rinv + c.sigma * wc * xi * dtis;
                    fpair *= rinv;
                    if (EV) en += 0.5f * c.a * (1.f / c.rcinv) * wc * wc + 0.5f * c.b * (1.f / c.rdinv) * rr * wcr * wcr;
                } else {
                    // pair_dpd_meso.cu:145-175
                    const float wr = c.expw == 1.f ? wc : powf(wc, c.expw);
                    fpair = c.a * wc - c.gamma * wr * wr * dot * rinv + c.sigma * wr * xi * dtis;
                    fpair *= rinv;
                    if (EV) en += 0.5f * c.a * (1.f / c.rcinv) * wc * wc;
                }
                fx += dx * fpair; fy += dy * fpair; fz += dz * fpair;
                if (EV) {
                    vxx += dx * dx * fpair; vyy += dy * dy * fpair; vzz += dz * dz * fpair;
                    vxy += dx * dy * fpair; vxz += dx * dz * fpair; vyz += dy * dz * fpair;
                }
            }
        }
        const int mk = P.mask[d->sela][i];
        if (mk & d->add_bit) { fx += d->add[0]; fy += d->add[1]; fz += d->add[2]; }
        if (EV) {
            P.frc[i] = make_float4(fx, fy, fz, 0.5f * en);
            const size_t cap = (size_t)P.cap;
            P.vir[i] = 0.5f * vxx; P.vir[cap + i] = 0.5f * vyy; P.vir[2 * cap + i] = 0.5f * vzz;
            P.vir[3 * cap + i] = 0.5f * vxy; P.vir[4 * cap + i] = 0.5f * vxz; P.vir[5 * cap + i] = 0.5f * vyz;
        }
        if (MODE != MODE_SETUP) {
            float vx = vi.x, vy = vi.y, vz = vi.z;
            float4 pn = pi;
            if (mk & d->nve_bit) {
                const float dtfm = d->dtf * P.minv[d->mass_base + ti];
                vx += dtfm * fx; vy += dtfm * fy; vz += dtfm * fz;             // final_integrate
            }
            if (MODE == MODE_FINAL) {
                P.pos[cur ^ 1][i] = pi;
                P.vel[cur ^ 1][i] = make_float4(vx, vy, vz, vi.w);
                P.vend[i] = make_float4(vx, vy, vz, 0.f);
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
    if (MODE == MODE_FUSED || MODE == MODE_FUSED_OUT)
        block_commit(d, moved, [&]() { commit_advance(d); });
    else if (MODE == MODE_FINAL)
        block_commit(d, 0, [&]() { d->sel ^= 1; });
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
hipError_t launch_density(const PoolDev &P, hipStream_t st)
{
    const int nl = P.cap / BLK;
    if (nl == 0) return hipSuccess;
    hipLaunchKernelGGL(k_density, dim3(xcd_grid(nl)), dim3(BLK), 0, st, P, nl);
    return hipGetLastError();
}

hipError_t launch_initial(const PoolDev &P, hipStream_t st)
{
    const int nl = P.cap / BLK;
    if (nl == 0) return hipSuccess;
    hipLaunchKernelGGL(k_initial, dim3(xcd_grid(nl)), dim3(BLK), 0, st, P, nl);
    return hipGetLastError();
}

template <int STYLE, int MODE>
static void launch_force_t(const PoolDev &P, bool uni, int nl, hipStream_t st)
{
    if (uni)
        hipLaunchKernelGGL((k_force<STYLE, MODE, true>), dim3(xcd_grid(nl)), dim3(BLK), 0, st, P, nl);
    else
        hipLaunchKernelGGL((k_force<STYLE, MODE, false>), dim3(xcd_grid(nl)), dim3(BLK), 0, st, P, nl);
}

hipError_t launch_force(const PoolDev &P, int style, int mode, bool uni, hipStream_t st)
{
    const int nl = P.cap / BLK;
    if (nl == 0) return hipSuccess;
#define MDPD_FORCE_CASE(S, M) \
    if (style == S && mode == M) { launch_force_t<S, M>(P, uni, nl, st); return hipGetLastError(); }
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_SETUP)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FUSED)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FUSED_OUT)
    MDPD_FORCE_CASE(STYLE_MDPD, MODE_FINAL)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_SETUP)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FUSED)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FUSED_OUT)
    MDPD_FORCE_CASE(STYLE_DPD, MODE_FINAL)
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
