// Shared definitions for the mdpd355 engine: the per-system descriptor the
// kernels read, particle-array bundles, and the device-side integer/float
// helpers (TEA, signatures, Morton interleave, the FP32 pair Gaussian).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdpd {

constexpr int WAVE = 64;
constexpr int MAXTYPES = 8;
constexpr int POOL_RDF = 2;        // rdf/fast/meso fixes per pooled job
constexpr int OBS_MAXBIN = 4096;   // bins of an observable histogram
constexpr int POOL_DENS_WORDS = 8192;   // density windows x (bins + 1) per pooled job
constexpr int NCOEF = 10;          // per type pair, see Coef
constexpr int IMG_SHIFT = 27;      // neighbor entry = j | image << 27
constexpr uint32_t IDX_MASK = (1u << IMG_SHIFT) - 1u;
constexpr int IMG_ZERO = 13;       // (0,0,0) image: (sx+1) + 3(sy+1) + 9(sz+1)

// Neighbor rows are stored per 64-particle tile in chunks of NBR_CHUNK
// entries per particle: entry q of tile lane l is at
//   tile * maxnbr * 64 + (q / 8) * 512 + l * 8 + q % 8.
// Consecutive entries of one particle fill whole 32-byte sectors (the row
// builder appends a run of a particle's entries with one store), and a
// reader loads four entries of its particle with one 16-byte load.  maxnbr
// is always a multiple of NBR_ALIGN = 32 = NBR_CHUNK x the largest split: the
// first pass gives wave w of SPLIT the chunks w, w + SPLIT, ..., so a wave
// holds at most maxnbr / SPLIT entries and its compact row t * SPLIT + w
// stays below maxnbr.
constexpr int NBR_CHUNK = 8;
constexpr int NBR_ALIGN = 32;

// Compact-row counts of a step: MAXSPLIT slots of 64 per tile whatever the
// split of the pass, so passes at different splits never write each other's
// tiles' counts (a finished pooled system keeps the counts of its last pass
// until its harvest pass reads them).
constexpr int MAXSPLIT = 4;
__host__ __device__ inline size_t ccount_at(int tile, int w, int lane)
{
    return ((size_t)tile * MAXSPLIT + w) * WAVE + lane;
}

__host__ __device__ inline size_t nbr_at(int i, int q, int maxnbr)
{
    return (size_t)(i >> 6) * maxnbr * WAVE + (size_t)(q & ~(NBR_CHUNK - 1)) * WAVE +
           (size_t)(i & 63) * NBR_CHUNK + (q & (NBR_CHUNK - 1));
}

// Each cell is split into SUBCELLS sub-cells by the top bits of the sub-bin
// Morton key (bit 10: upper half in y, bit 11: upper half in z).  The sort
// key orders a cell's particles by that code first, so every sub-cell is a
// contiguous segment of the cell and the counting sort runs per sub-cell
// (fewer same-address atomics, a short rank loop per particle).
constexpr int SUBCELLS = 4;
constexpr int SUBCELL_SHIFT = 10;

// The hardware dispatches workgroup b to XCD b % 8 (each XCD has its own
// L2); xcd_logical gives XCD x the contiguous logical range
// [x*per, (x+1)*per) of a grid of xcd_grid(n) workgroups.
constexpr int NXCD = 8;
__device__ inline int xcd_logical(int b, int nlogical)
{
    const int per = (nlogical + NXCD - 1) / NXCD;
    return (b % NXCD) * per + b / NXCD;
}
static inline int xcd_grid(int nlogical) { return ((nlogical + NXCD - 1) / NXCD) * NXCD; }

// inclusive prefix sum over the 64 lanes of a wave (every lane active)
__device__ inline int wave_incl_scan(int x, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// sum over the 64 lanes of a wave, in every lane (every lane active)
__device__ inline int wave_sum(int x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

enum Style : int { STYLE_MDPD = 0, STYLE_DPD = 1 };

// Lifecycle of a system in a pooled (batched / served) launch.  A job of K
// steps makes K+1 force passes: FRESH (setup forces + first half-kick and
// drift), K-1 RUN passes (fused step), one LAST pass (final half-kick only).
enum SysPhase : int { PH_FRESH = 0, PH_RUN = 1, PH_LAST = 2, PH_DONE = 3 };

// Per type pair coefficients in FP32, laid out as NCOEF floats.
// mDPD (pair_mdpd_meso.cu:70-91): rc, 1/rc, rd, 1/rd, A, B, gamma, sigma.
// DPD  (pair_dpd_meso.cu):        rc, 1/rc, -, -,   a0, -, gamma, sigma, expw.
enum Coef : int { C_RC = 0, C_RCINV, C_RD, C_RDINV, C_A, C_B, C_GAMMA, C_SIGMA, C_EXPW, C_RC2 };

// One simulated box.  Static fields are written by the host before a
// launch; dynamic fields are written only by the commit points named below.
struct SysDev {
    // geometry (FP32 copies for the per-step kernels)
    float lo[3], hi[3], L[3];
    // bin geometry in FP64 for the sort-key formula (neighbor_meso.cu:837-853)
    double lo_d[3], binsize_d[3], bininv_d[3];
    int mbin[3];                 // bins incl. ghost layer
    int nb[3];                   // interior cells per dim (= mbin - 2)
    int ncell;                   // nb[0]*nb[1]*nb[2]
    int cell_base;               // first cell slot in pooled cell arrays (ncell+1 slots)
    int part_base;               // first particle slot (multiple of 64)
    int n;                       // particles
    int npad;                    // n rounded up to 64
    int l1_shift;                // 12
    float rc2_list;              // (float)((rc+skin)^2)
    float rc2_max;               // largest pair cutoff squared (compaction test)
    float trig2;                 // (skin/2)^2
    float dt, dtf, dtinvsqrt;
    uint32_t seed;
    int style;
    int ntypes;
    int coef_base;               // offset (in floats) into the pooled coefficient table
    int mass_base;               // offset into the pooled per-type 1/m, m tables
    int nve_bit, rho_bit, add_bit;
    float add[3];
    float rd, rho_factor;
    int lucy;
    int maxnbr;                  // neighbor row width for this pool
    // dynamic state
    int sel;                     // buffer holding x, v (flips every step and at re-sort)
    int sela;                    // buffer holding tag, mask, origin, image (flips at re-sort)
    int stale;                   // neighbor list must be rebuilt before the next force pass
    int stale_pending;           // set during a step by the displacement check
    int overflow;                // a neighbor row or cell exceeded its capacity
    int active;                  // 0: slot idle (batched / served engine)
    int steps_left;              // pooled engine: steps still to run
    long long step;              // ntimestep of the next / current force evaluation
    uint32_t seed_next;          // premix_TEA<64>(seed, step + 1)
    uint32_t seed_cur;           // premix_TEA<64>(seed, step): signatures of a fresh system
    int phase;                   // pooled engine: PH_FRESH / PH_RUN / PH_LAST / PH_DONE
    int job;                     // pooled engine: caller's job id of this slot
    int split_last;              // pooled engine: waves per tile of the system's last pass
    // arbitrary boundary: wall-scaled gamma / sigma of mobile-wall pairs
    // (pair_style mdpd/meso ... mobile wall) and wall fraction + bounce-back
    // (fix aboundary/rho/meso ... wall G mobile G)
    int pair_arb, pair_mobile_bit, pair_wall_bit;
    int fix_arb, fix_mobile_bit, fix_wall_bit;
    // pooled engine: the job's rdf/fast/meso fixes, sampled after the force
    // pass of every step that is a multiple of rdf_every; rdf_hist[k] holds
    // rdf_nbin[k] bins and then the sample count
    int nrdf;
    int rdf_nbin[POOL_RDF], rdf_every[POOL_RDF], rdf_gi[POOL_RDF], rdf_gj[POOL_RDF];
    float rdf_rc[POOL_RDF];
    unsigned *rdf_hist[POOL_RDF];
    // pooled engine: the job's density/meso fix; window w (centre dens_m0 +
    // w * dens_every) keeps dens_nbin counts and then its sample count at
    // dens_hist + w * (dens_nbin + 1)
    int ndens, dens_nbin, dens_every, dens_window, dens_along, dens_g, dens_nwin;
    long long dens_m0;
    unsigned *dens_hist;
    float cut_phi, phi_c, phi_factor, dw_factor;
    // pooled engine: Neighbor::decide on the device (src/neighbor.cpp:1216-1290)
    int ago;                     // steps since the last build
    int n_every, n_delay, n_check;
    int builds;                  // neighbor builds committed (k_commit_rebuild)
};

// Device arrays of one pool (a pool packs one or more systems).
struct PoolDev {
    float4 *pos[2];      // x, y, z, bits(type-1)
    float4 *vel[2];      // vx, vy, vz, bits(signature)
    int *tag[2], *mask[2], *origin[2], *image[2];
    float4 *xhold;       // positions at the last build
    float *rho;
    float4 *pr;          // x, y, z and rho of the current step (written by the mDPD first pass)
    float *phi[2];       // wall fraction of the step (flips with tag / mask at re-sort)
    float4 *frc;         // fx, fy, fz, e_pair   (output steps only)
    float *vir;          // 6 * cap, SoA blocks  (output steps only)
    float4 *vend;        // end-of-step velocity  (output steps only)
    // sort / cell scratch
    int *pcell;          // global sub-cell slot per particle (cell slot * SUBCELLS + sub-cell)
    int *pslot;          // arrival rank within the sub-cell; after the re-sort: sub-cell slot of each sorted slot
    int *psub;           // 12-bit sub-bin Morton code
    int *tmp;            // particle index per sorted slot (pre-order)
    unsigned long long *key;  // optional: reference sort key per particle (debug)
    int *cell_count;     // per sub-cell slot
    int *cell_start;     // per sub-cell slot, global particle index; slot ncell * SUBCELLS = end
    int *cell_xyz;       // packed interior coords of the cell at this Morton rank
    int *cell_rank;      // Morton rank of linear cell id (per system, cell_base + lin)
    int *cell_sys;       // system of each cell slot
    int4 *cell_work;     // pools: the cells holding particles of the systems re-sorted this pass, as
                         // (system, packed coordinates, first particle, end) (k_scan_apply appends)
    int *work_count;     // [0]: entries of cell_work (k_keys clears it)
    int2 *cell_lse;      // pools: particle range [start, end) of each cell by linear cell id (cell_base + lin)
    int *nlist;         // (cap/64) tiles * maxnbr * 64
    int *ncount;         // per particle
    int *clist;          // per step: entries of nlist within the pair cutoff, same layout
    int *ccount;         // per particle
    int *wave_sys;       // system of each 64-particle wave
    int *moved;          // per 64-particle tile: a particle moved past skin/2 this step
    int *status;         // pooled engine, per system: job << 3 | 1 stale | 2 done | 4 overflow
    int *cl;             // per particle slot, in cell-segment order: sub-bin key (re-sort scratch)
    int *wall_near;      // per particle: its neighbor row holds a wall particle (wall branch; set by
                         // the first wall pass after each build, valid until the next)
    SysDev *sys;
    float *coef;         // pooled coefficient tables
    float *minv;         // pooled per-type 1/m
    float *mass;         // pooled per-type m
    int cap;             // particle slots (multiple of 64): array extent and strides
    int nact;            // slots [0, nact) hold systems: extent of the launches (<= cap)
    int ncellslots;
    int ncact;           // cell slots [0, ncact) hold systems: extent of per-cell launches
    int nsys;            // descriptor slots [0, nsys) hold systems: extent of per-system launches
    int maxsub;          // sub-cell slots of the largest system (0: all cell slots, one system)
    int maxnbr;
};

// ----------------------------------------------------------------------------
// TEA and friends: src/USER-MESO/math_meso.h:178-201,435-463
// ----------------------------------------------------------------------------

__host__ __device__ inline void tea_rounds(uint32_t &v0, uint32_t &v1, int rounds)
{
    uint32_t sum = 0;
    for (int r = 0; r < rounds; ++r) {
        sum += 0x9E3779B9u;
        v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xC8013EA4u);
        v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7E95761Eu);
    }
}

template <int N>
__host__ __device__ inline uint32_t premix(uint32_t v0, uint32_t v1)
{
    uint32_t sum = 0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        sum += 0x9E3779B9u;
        v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xC8013EA4u);
        v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7E95761Eu);
    }
    return v0 ^ v1;
}

__host__ inline uint32_t premix_host(uint32_t v0, uint32_t v1, int rounds)
{
    tea_rounds(v0, v1, rounds);
    return v0 ^ v1;
}

__device__ inline uint32_t brev32(uint32_t x) { return __builtin_bitreverse32(x); }

// bit i of x -> bit 3i (x < 2^10 here: coordinates and 11-bit mantissa slices)
__host__ __device__ inline uint64_t spread3(uint32_t x)
{
    uint64_t v = x & 0x1FFFFFu;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ inline uint64_t morton3(uint32_t i, uint32_t j, uint32_t k)
{
    return spread3(i) | (spread3(j) << 1) | (spread3(k) << 2);
}

__device__ inline uint32_t mantissa_bits(float u, float v, float w)
{
    uint32_t i = (__float_as_uint(u) & 0x7FF000u) >> 12;
    uint32_t j = (__float_as_uint(v) & 0x7FF000u) >> 12;
    uint32_t k = (__float_as_uint(w) & 0x7FF000u) >> 12;
    return (uint32_t)morton3(i, j, k);
}

// mDPD signature: atom_vec_mdpd_atomic_meso.cu:166
__device__ inline uint32_t sig_mdpd(int tag, uint32_t seed_t) { return premix<32>(brev32((uint32_t)tag), seed_t); }
// DPD signature: atom_vec_meso.cu:164
__device__ inline uint32_t sig_dpd(int tag, uint32_t seed_t, float vx, float vy, float vz)
{
    return seed_t ^ premix<16>(brev32((uint32_t)tag), mantissa_bits(vx, vy, vz));
}

// ----------------------------------------------------------------------------
// FP32 pair Gaussian from the four-round TEA bits (math_meso.h:465-473).
// cos(pi x) and ln u are evaluated in FP32 from the mathematics:
//  * cos(pi x) = -sin(pi t), t = x - 1/2 in [-1/2, 1/2), with the Taylor
//    series of sin to t^11 (truncation < 6e-8);
//  * ln u for u = v1 / 2^32: the log2 of u rounded to FP32, corrected to
//    first order by u's rounding error from the exact integer complement
//    2^32 - v1 (u near 1 keeps its precision).
// ----------------------------------------------------------------------------

__device__ inline float sinpi_half(float t)
{
    // sin(pi t), |t| <= 1/2; c_k = (-1)^k pi^(2k+1) / (2k+1)!
    const float t2 = t * t;
    float s = -0.0073704309457143504f;        // -pi^11/11!
    s = fmaf(s, t2, 0.082145886611128228f);   //  pi^9/9!
    s = fmaf(s, t2, -0.59926452932079207f);   // -pi^7/7!
    s = fmaf(s, t2, 2.5501640398773455f);     //  pi^5/5!
    s = fmaf(s, t2, -5.1677127800499703f);    // -pi^3/3!
    s = fmaf(s, t2, 3.1415926535897932f);     //  pi
    return s * t;
}

// -2 ln u for u = w / 2^32, 1 <= w < 2^32, with one hardware log2 and no
// branches.  u_f = fl(w) 2^-32 is a normal float (the power-of-two scale is
// exact) and ln u = ln u_f + ln(u / u_f) ~ ln u_f + (u - u_f), where
// u - u_f = (1 - u_f) - c 2^-32 with the exact integer complement
// c = 2^32 - w.  For u_f >= 1/2, 1 - u_f is exact (Sterbenz) and the
// neglected factor 1/u_f - 1 <= 1 - u_f scales a term of at most 2^-25, so
// u near 1, where ln u ~ -(1 - u), keeps its relative precision; below 1/2,
// |ln u| > 0.69 and the correction stays below 2^-24.
__device__ inline float neg2_ln_u32(uint32_t w)
{
    const float uf = (float)w * 2.3283064365386963e-10f;                       // w 2^-32
    // -2 (u - u_f) = 2 u_f - 2 + c 2^-31
    const float e = fmaf((float)(0u - w), 4.6566128730773926e-10f, fmaf(uf, 2.0f, -2.0f));
    return fmaf(__builtin_amdgcn_logf(uf), -1.3862943611198906f, e);           // -2 ln 2
}

__device__ inline float pair_gaussian(uint32_t si, uint32_t sj)
{
    uint32_t v0 = si > sj ? si : sj;
    uint32_t v1 = si > sj ? sj : si;
    tea_rounds(v0, v1, 4);
    // x = (v0 & 0x7FFFFFFF) 2^-31 ; t = x - 1/2 exactly from the integer
    const int ti = (int)(v0 & 0x7FFFFFFFu) - (1 << 30);
    const float t = (float)ti * 4.656612873077392578125e-10f;
    const float s = sinpi_half(t);                 // cos(pi x) = -sin(pi t)
    const float f = (v0 & 0x80000000u) ? -s : s;   // times +1 / -1 by the top bit
    const float r = __builtin_amdgcn_sqrtf(neg2_ln_u32(v1 > 1u ? v1 : 1u));
    return fminf(fmaxf(r * f, -4.0f), 4.0f);
}

// gamma / sigma scale of a mobile-wall pair from the mobile particle's wall
// fraction phi and the pair cutoff rcw (pair_mdpd_meso.cu:171-182):
// h = max(1 - (2.088 phi^3 + 1.478 phi)^(1/4), 0.025) rcw with phi <= 1/2,
// ratio = 1 + 0.187 (rcw/h - 1) - 0.093 (1 - h/rcw)^3
__device__ inline float wall_ratio(float phi, float rcw)
{
    const float ph = fminf(phi, 0.5f);
    const float q4 = 2.088f * ph * ph * ph + 1.478f * ph;
    float h = 1.0f - __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(q4));
    h = fmaxf(h, 0.025f) * rcw;
    const float q = 1.0f - h / rcw;
    return 1.0f + 0.187f * (rcw / h - 1.0f) - 0.093f * q * q * q;
}

// image code -> shift multipliers
__device__ inline void image_decode(uint32_t code, int &sx, int &sy, int &sz)
{
    int c = (int)code;
    sz = (c * 57) >> 9;          // c / 9 for c < 27
    int r = c - 9 * sz;
    sy = (r * 11) >> 5;          // r / 3 for r < 9
    sx = r - 3 * sy;
    sx -= 1; sy -= 1; sz -= 1;
}

}  // namespace mdpd
