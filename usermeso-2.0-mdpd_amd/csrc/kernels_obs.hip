// Device observables of the mdpd355 engine: pair-distance histogram
// (fix rdf/fast/meso) and binned number density along an axis
// (fix density/meso).  Both count into LDS per workgroup and add the
// workgroup's bins to the global histogram once (one atomic per bin per
// workgroup).  Slot nbin of a histogram counts the samples taken; a step
// whose rows overflowed (the step is repeated once they are widened) is
// not sampled.  k_rdf and k_profile sample one system (the plain path);
// k_obs_pool samples each pooled system by its own descriptors.
//
// Replaces:
//   MesoFixRDFFast::post_force / gpu_calc_rdf        fix_rdf_fast_meso.cu:102-175
//   MesoFixDensityProf::compute / gpu_density_profile fix_density_profile_meso.cpp:118-190
#include "common.hpp"
#include "kernels.hpp"

namespace mdpd {

constexpr int OBS_BLK = 256;

// pairs (i in group gi, j in group gj) of particle i's neighbor row with
// r < rc, at the positions of the current buffer, counted into the LDS bins
// h; bin = floor(r * nbin / rc).  Every pair appears from both sides, as
// with the reference's full list.
__device__ inline void rdf_count(const PoolDev &P, const SysDev *d, int i, unsigned *h, int nbin, float rc, int gi,
                                 int gj)
{
    const int *mask = P.mask[d->sela];
    if (!(mask[i] & gi)) return;
    const float4 *pos = P.pos[d->sel];
    const float4 pi = pos[i];
    const int maxnbr = P.maxnbr;
    const int cnt = min(P.ncount[i], maxnbr);
    const int *row = P.nlist + nbr_at(i, 0, maxnbr);
    const float rc2 = rc * rc, binv = (float)nbin / rc;
    for (int q = 0; q < cnt; ++q) {
        const uint32_t e = (uint32_t)row[(size_t)(q & ~(NBR_CHUNK - 1)) * WAVE + (q & (NBR_CHUNK - 1))];
        const uint32_t j = e & IDX_MASK;
        if (!(mask[j] & gj)) continue;
        const float4 pj = pos[j];
        float xj = pj.x, yj = pj.y, zj = pj.z;
        const uint32_t code = e >> IMG_SHIFT;
        if (code != IMG_ZERO) {
            int sx, sy, sz;
            image_decode(code, sx, sy, sz);
            xj += (float)sx * d->L[0];
            yj += (float)sy * d->L[1];
            zj += (float)sz * d->L[2];
        }
        const float dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
        const float rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < rc2) {
            const float r = rsq > 0.f ? rsq * __builtin_amdgcn_rsqf(rsq) : 0.f;   // coincident pair: bin 0
            const int b = min((int)floorf(r * binv), nbin - 1);
            atomicAdd(&h[b], 1u);
        }
    }
}

__global__ void __launch_bounds__(OBS_BLK) k_rdf(PoolDev P, unsigned *hist, int nbin, float rc, int gi, int gj)
{
    if (P.sys->overflow) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) hist[nbin] += 1u;
    __shared__ unsigned h[OBS_MAXBIN];
    for (int b = threadIdx.x; b < nbin; b += OBS_BLK) h[b] = 0u;
    __syncthreads();
    const int i = blockIdx.x * OBS_BLK + threadIdx.x;
    if (i < P.nact) {
        const int s = P.wave_sys[i >> 6];
        const SysDev *d = s >= 0 ? P.sys + s : nullptr;
        if (d && d->active && i < d->part_base + d->n) rdf_count(P, d, i, h, nbin, rc, gi, gj);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbin; b += OBS_BLK)
        if (h[b]) atomicAdd(hist + b, h[b]);
}

// particles of group g binned by coordinate `axis` of the current buffer:
// bin = floor((x - lo) * nbin / L); out-of-range coordinates (a particle
// past the box between re-sorts) are not counted
// particle i of group g counted into LDS bin floor((x - lo) * nbin / L)
__device__ inline void profile_count(const PoolDev &P, const SysDev *d, int i, unsigned *h, int nbin, int axis, int g)
{
    if (!(P.mask[d->sela][i] & g)) return;
    const float4 p = P.pos[d->sel][i];
    const float c = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
    const int b = (int)floorf((c - d->lo[axis]) * ((float)nbin / d->L[axis]));
    if (b >= 0 && b < nbin) atomicAdd(&h[b], 1u);
}

__global__ void __launch_bounds__(OBS_BLK) k_profile(PoolDev P, unsigned *count, int nbin, int axis, int g)
{
    if (P.sys->overflow) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) count[nbin] += 1u;
    __shared__ unsigned h[OBS_MAXBIN];
    for (int b = threadIdx.x; b < nbin; b += OBS_BLK) h[b] = 0u;
    __syncthreads();
    const int i = blockIdx.x * OBS_BLK + threadIdx.x;
    if (i < P.nact) {
        const int s = P.wave_sys[i >> 6];
        const SysDev *d = s >= 0 ? P.sys + s : nullptr;
        if (d && d->active && i < d->part_base + d->n) profile_count(P, d, i, h, nbin, axis, g);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbin; b += OBS_BLK)
        if (h[b]) atomicAdd(count + b, h[b]);
}

// Pooled systems: a workgroup covers 256 slots of one system (particle
// ranges and their padding are multiples of 256) and samples the step this
// pass evaluated (PH_RUN and PH_LAST passes: the step is d->step until the
// commit, and x(t) stays in the current buffer; the setup pass of a fresh
// system is not sampled, as a run does not sample its setup): each rdf fix
// whose `every` divides the step, and the density fix when the step lies in
// one of its windows.  The workgroup at the system's first slot counts the
// sample.
__global__ void __launch_bounds__(OBS_BLK) k_obs_pool(PoolDev P)
{
    __shared__ unsigned h[OBS_MAXBIN];
    const int i0 = blockIdx.x * OBS_BLK;
    if (i0 >= P.nact) return;
    const int s = P.wave_sys[i0 >> 6];
    if (s < 0) return;
    const SysDev *d = P.sys + s;
    if (!d->active || d->overflow || (d->nrdf <= 0 && d->ndens <= 0)) return;
    if (d->phase != PH_RUN && d->phase != PH_LAST) return;
    const int i = i0 + threadIdx.x;
    const bool mine = i < d->part_base + d->n && P.wave_sys[i >> 6] == s;
    for (int k = 0; k < d->nrdf && k < POOL_RDF; ++k) {
        if (d->step % d->rdf_every[k]) continue;                // uniform over the workgroup
        const int nbin = min(d->rdf_nbin[k], OBS_MAXBIN);
        unsigned *hist = d->rdf_hist[k];
        if (i0 == d->part_base && threadIdx.x == 0) hist[nbin] += 1u;
        for (int b = threadIdx.x; b < nbin; b += OBS_BLK) h[b] = 0u;
        __syncthreads();
        if (mine) rdf_count(P, d, i, h, nbin, d->rdf_rc[k], d->rdf_gi[k], d->rdf_gj[k]);
        __syncthreads();
        for (int b = threadIdx.x; b < nbin; b += OBS_BLK)
            if (h[b]) atomicAdd(hist + b, h[b]);
        __syncthreads();
    }
    if (d->ndens > 0) {
        const long long t = d->step, ev = d->dens_every;
        const long long m = (t + ev / 2) / ev * ev;                // the nearest window centre
        const long long w = (m - d->dens_m0) / ev;
        if (t - m >= -d->dens_window && t - m < d->dens_window && m >= d->dens_m0 && w < d->dens_nwin) {
            const int nbin = min(d->dens_nbin, OBS_MAXBIN);
            unsigned *hist = d->dens_hist + w * (nbin + 1);
            if (i0 == d->part_base && threadIdx.x == 0) hist[nbin] += 1u;
            for (int b = threadIdx.x; b < nbin; b += OBS_BLK) h[b] = 0u;
            __syncthreads();
            if (mine) profile_count(P, d, i, h, nbin, d->dens_along, d->dens_g);
            __syncthreads();
            for (int b = threadIdx.x; b < nbin; b += OBS_BLK)
                if (h[b]) atomicAdd(hist + b, h[b]);
        }
    }
}

hipError_t launch_rdf(const PoolDev &P, unsigned *hist, int nbin, float rc, int gi, int gj, hipStream_t st)
{
    if (nbin < 1 || nbin > OBS_MAXBIN) return hipErrorInvalidValue;
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rdf, dim3((P.nact + OBS_BLK - 1) / OBS_BLK), dim3(OBS_BLK), 0, st, P, hist, nbin, rc, gi, gj);
    return hipGetLastError();
}

hipError_t launch_obs_pool(const PoolDev &P, hipStream_t st)
{
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_obs_pool, dim3((P.nact + OBS_BLK - 1) / OBS_BLK), dim3(OBS_BLK), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_profile(const PoolDev &P, unsigned *count, int nbin, int axis, int g, hipStream_t st)
{
    if (nbin < 1 || nbin > OBS_MAXBIN || axis < 0 || axis > 2) return hipErrorInvalidValue;
    if (P.nact == 0) return hipSuccess;
    hipLaunchKernelGGL(k_profile, dim3((P.nact + OBS_BLK - 1) / OBS_BLK), dim3(OBS_BLK), 0, st, P, count, nbin, axis, g);
    return hipGetLastError();
}

}  // namespace mdpd
