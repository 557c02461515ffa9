// Host-side launchers for the mdpd355 kernels.
#pragma once

#include "common.hpp"

namespace mdpd {

enum ForceMode : int {
    MODE_SETUP = 0,      // forces (+energy/virial) of the current state, no state change
    MODE_FUSED = 1,      // force, final half-kick, next initial half-kick + drift
    MODE_FUSED_OUT = 2,  // MODE_FUSED and also write f, e, virial, end-of-step v
    MODE_FINAL = 3,      // force and final half-kick only (end of a run)
    MODE_PHASED = 4,     // pooled engine: each system follows its own SysPhase
    MODE_HARVEST = 5,    // pooled engine, finished system: f, e, virial of its last pass
};

// step commit after a pass (k_commit_step)
enum CommitKind : int {
    COMMIT_FINAL = 0,    // end-of-run half-kick: flip the buffer only
    COMMIT_ADVANCE = 1,  // flip, step + 1, next seed; stale = a particle moved past skin/2
    COMMIT_PHASED = 2,   // pooled engine: per-system lifecycle + reneighbor decision
    COMMIT_DECIDE = 3,   // plain engine: COMMIT_ADVANCE with the reneighbor decision
};

// rebuild phase: wrap, keys, sort, gather, neighbor rows (stale systems only)
hipError_t launch_rebuild(const PoolDev &P, hipStream_t st);
hipError_t launch_nlist_only(const PoolDev &P, hipStream_t st);
// neighbor rows of every sorted active system rebuilt in place (row width grew)
hipError_t launch_relist(const PoolDev &P, hipStream_t st);
hipError_t launch_commit_rebuild(const PoolDev &P, hipStream_t st);
hipError_t launch_signatures(const PoolDev &P, hipStream_t st);

// per step
// first pass of a step: density (mDPD with the density fix) and compact rows
hipError_t launch_first(const PoolDev &P, bool with_rho, int split, hipStream_t st);
hipError_t launch_initial(const PoolDev &P, hipStream_t st);
// after launch_initial / a fused force pass (advance) or a final pass (!advance)
// kind: CommitKind; split = waves per tile of the pass
hipError_t launch_commit_step(const PoolDev &P, int kind, int split, hipStream_t st);
hipError_t launch_force(const PoolDev &P, int style, int mode, bool uni, int split, hipStream_t st);
// the same over the 64-particle tiles [tile0, tile0 + ntiles) only
hipError_t launch_force_range(const PoolDev &P, int style, int mode, bool uni, int split, int tile0, int ntiles,
                              hipStream_t st);
// wall branch of aboundary/rho/meso (wall fraction, bounce-back, displacement
// flags) on the buffer a pass just wrote; pooled: only systems that integrated
hipError_t launch_aboundary(const PoolDev &P, int pooled, hipStream_t st);
// pooled engine, finished system s (npad slots from part_base): its results
// packed into dst (room bytes, at least harvest_bytes(npad)) in this layout:
//   float4 x[npad] | float4 v[npad] | float4 f[npad] (w: pair energy) |
//   int tag[npad], mask[npad], origin[npad], image[npad] | float rho[npad] |
//   float virial[6][npad] | SysDev
constexpr size_t HARVEST_SLOT_BYTES = 3 * 16 + 4 * 4 + 4 + 6 * 4;
inline size_t harvest_bytes(size_t npad) { return npad * HARVEST_SLOT_BYTES + sizeof(SysDev); }
hipError_t launch_harvest(const PoolDev &P, int s, int part_base, int npad, void *dst, size_t room, hipStream_t st);

// observables (accumulate into device histograms of nbin + 1 slots, slot nbin
// counting the samples; a system whose rows overflowed is not sampled)
hipError_t launch_rdf(const PoolDev &P, unsigned *hist, int nbin, float rc, int gi, int gj, hipStream_t st);
hipError_t launch_profile(const PoolDev &P, unsigned *count, int nbin, int axis, int g, hipStream_t st);
// rdf/fast/meso and density/meso of every pooled system that has them and
// samples this pass
hipError_t launch_obs_pool(const PoolDev &P, hipStream_t st);

// stateless helpers for parity tests
hipError_t launch_pair_gaussian(const uint32_t *a, const uint32_t *b, float *out, int n, hipStream_t st);
hipError_t launch_premix(const uint32_t *a, const uint32_t *b, uint32_t *out, int n, int rounds, hipStream_t st);

}  // namespace mdpd
