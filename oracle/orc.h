/*
 * oracle/orc.h -- CPU restatement (FP64 unless stated) of the USER-MESO
 * per-timestep mDPD/DPD path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the engine links, loads or calls this
 * code; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, as the checker.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the USER-MESO reference tree).
 */
#ifndef MDPD355_ORACLE_H
#define MDPD355_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXTYPES 8

/* ---------------- integer pieces (bit-exact) ---------------- */

/* TEA rounds with an arbitrary key (the reference fixes the key, see
 * orc_premix); used to pin the round structure against the published TEA
 * test vector.  src/USER-MESO/math_meso.h:449-455 */
void orc_tea_key(uint32_t *v0, uint32_t *v1, int rounds, const uint32_t key[4]);
/* __TEA_core<N> with the author's key; src/USER-MESO/math_meso.h:443-455 */
void orc_tea(uint32_t *v0, uint32_t *v1, int rounds);
/* premix_TEA<N>(v0, v1) = v0 ^ v1 after N rounds; math_meso.h:459-463 */
uint32_t orc_premix(uint32_t v0, uint32_t v1, int rounds);
/* CUDA __brev: bit reversal of a 32-bit word */
uint32_t orc_brev32(uint32_t x);
/* bit_space3 / interleave3 / morton_encode; math_meso.h:178-201 */
uint64_t orc_interleave3(uint32_t i, uint32_t j, uint32_t k);
/* __mantissa: interleave of FP32 bits 12..22 of (u, v, w); math_meso.h:435-441 */
uint32_t orc_mantissa(float u, float v, float w);
/* seed_now() = premix_TEA<64>(seed, ntimestep); pair_mdpd_meso.cu:295-297,
 * pair_dpd_meso.cu:268-270 */
uint32_t orc_seed_now(uint32_t seed, int64_t ntimestep);
/* mDPD particle signature: premix_TEA<32>(brev(tag), seed_t);
 * atom_vec_mdpd_atomic_meso.cu:166 */
uint32_t orc_sig_mdpd(int32_t tag, uint32_t seed_t);
/* DPD particle signature: seed_t ^ premix_TEA<16>(brev(tag), mantissa(v));
 * atom_vec_meso.cu:164 */
uint32_t orc_sig_dpd(int32_t tag, uint32_t seed_t, float vx, float vy, float vz);
/* the two TEA words a pair's Gaussian is built from: (v0, v1) = TEA^4(max,min)
 * of the two signatures (unsigned compare); math_meso.h:465-469,
 * pair_mdpd_meso.cu:163 */
void orc_pair_bits(uint32_t si, uint32_t sj, uint32_t *v0, uint32_t *v1);

/* ---------------- pairwise Gaussian ---------------- */

/* gaussian_TEA<4> with the author's own polynomials (__cospi, __log2u);
 * math_meso.h:370-423,465-473.  Newton-iteration rcp/rsqrt are replaced by
 * correctly rounded division / sqrt (the reference's are within 1 ulp). */
double orc_gaussian_poly(uint32_t si, uint32_t sj);
/* same bits, textbook libm cos/log/sqrt: pins the polynomial transcription */
double orc_gaussian_libm(uint32_t si, uint32_t sj);
/* author's polynomials alone (for closed-form checks) */
double orc_cospi_poly(double x);
double orc_log2u_poly(uint32_t x);

/* ---------------- bins, keys, sort ---------------- */

typedef struct {
    double lo[3], hi[3];
    double cutneigh;          /* rc + skin */
    int    mbin[3];           /* bins incl. one ghost layer per side */
    int    nbin[3];           /* interior bins = mbin - 2 */
    double binsize[3], bininv[3];
    int    l1_width, l1_shift, l2_width;
} orc_bins;

/* MesoNeighbor::setup_bins; src/USER-MESO/neighbor_meso.cu:771-853 */
void orc_setup_bins(const double lo[3], const double hi[3], double cutneigh, orc_bins *b);
/* bin index per dimension (ghost offset +1, clamped to the interior after the
 * on-device wrap); neighbor_meso.cu:414-416, atom_meso.cu:295-297 */
void orc_bin_of(const orc_bins *b, float x, float y, float z, int bin[3]);
/* 2-level Morton sort key (border bit 0); atom_meso.cu:292-305,346-365 */
uint64_t orc_sort_key(const orc_bins *b, float x, float y, float z);
/* std::stable_sort of (key, index) pairs = LSD radix SortPlan result;
 * sort_meso.h:21-116,364-406.  perm[k] = index of the k-th smallest key. */
void orc_stable_sort_u64(int n, const uint64_t *key, int *perm);
/* Domain::pbc in FP32: src/domain.cpp (wrap into [lo, hi)) */
void orc_wrap_fp32(int n, float *x, const double lo[3], const double hi[3]);

/* ---------------- neighbor sets (FP32 test, bit-exact) ---------------- */

/* Brute force over all j and the 27 periodic images: j is a neighbor of i iff
 * (j, image) != (i, 0) and the FP32 expression fma(dz, dz, fma(dy, dy, dx*dx))
 * <= (float)(cut*cut) (the reference's -O3 build contracts dx*dx+dy*dy+dz*dz
 * into that), images shifted in FP32 by (float)(hi-lo);
 * neigh_build_meso.cu:86-93.  Entries are encoded j | image << 27 with
 * image = (sx+1) + 3(sy+1) + 9(sz+1).  Writes counts[n]; list rows of width
 * maxnbr, each sorted ascending.  Returns the largest count (overflow if it
 * exceeds maxnbr: rows are truncated). */
int orc_neighbors_fp32(int n, const float *x, const double lo[3], const double hi[3],
                       double cut, int maxnbr, int *counts, int *list);

/* ---------------- physics (FP64) ---------------- */

typedef struct {
    int    ntypes;
    /* per type pair (0-based types): mDPD uses cut, cut_r, A, B, gamma, sigma;
     * DPD uses cut, a0, gamma, sigma, expw */
    double cut[ORC_MAXTYPES][ORC_MAXTYPES];
    double cut_r[ORC_MAXTYPES][ORC_MAXTYPES];
    double A[ORC_MAXTYPES][ORC_MAXTYPES];
    double B[ORC_MAXTYPES][ORC_MAXTYPES];
    double gamma[ORC_MAXTYPES][ORC_MAXTYPES];
    double sigma[ORC_MAXTYPES][ORC_MAXTYPES];
    double expw[ORC_MAXTYPES][ORC_MAXTYPES];
    double mass[ORC_MAXTYPES];
    double dt;
    double lo[3], hi[3];
} orc_params;

/* local density (quadratic weight 15/(2 pi rd^5)(rd-r)^2, self excluded,
 * times m_i); fix_aboundary_rho_meso.cu:88-90,119-165.  lucy != 0 selects
 * 105/(16 pi rd^7)(rd+3r)(rd-r)^3 (the commented alternative, :89,157).
 * Only particles with mask & groupbit get a value; others get 0. */
void orc_density(int n, const float *x, const int *type, const int *mask,
                 const orc_params *p, double rd, int lucy, int groupbit, double *rho);

/* mDPD pair force, energy and virial; pair_mdpd_meso.cu:93-228.
 * x, v: FP32 positions / velocities (3n), rho: densities, sig: signatures,
 * f: 3n out, e: n out (may be NULL), vir: 6n out xx,yy,zz,xy,xz,yz (may be
 * NULL). */
void orc_force_mdpd(int n, const float *x, const float *v, const double *rho,
                    const uint32_t *sig, const int *type, const orc_params *p,
                    double *f, double *e, double *vir);
/* DPD pair force; pair_dpd_meso.cu:91-205 */
void orc_force_dpd(int n, const float *x, const float *v, const uint32_t *sig,
                   const int *type, const orc_params *p,
                   double *f, double *e, double *vir);

/* ---------------- arbitrary boundary (aboundary/rho/meso wall branch) ---------------- */

typedef struct {
    /* pair_style mdpd/meso T rc seed <mobile group> <wall group>:
     * wall-scaled gamma / sigma for mobile-wall pairs; pair_mdpd_meso.cu:169-185 */
    int    pair_on, pair_mobile_bit, pair_wall_bit;
    /* fix aboundary/rho/meso ... cut_phi c phi_c p rho_wall w wall G mobile G:
     * wall fraction and bounce-back after initial integrate;
     * fix_aboundary_rho_meso.cu:93-94,220-287 */
    int    fix_on, fix_mobile_bit, fix_wall_bit;
    double cut_phi, phi_c, rho_wall;
} orc_arb;

/* ratio scaling gamma (x ratio) and sigma (x sqrt(ratio)) of a mobile-wall
 * pair from the mobile particle's wall fraction phi and the pair cutoff rcw;
 * pair_mdpd_meso.cu:171-182 */
double orc_wall_ratio(double phi, double rcw);
/* wall fraction phi_i and unit wall normal of every particle in the fix's
 * mobile group (0 / zero vector for others), from the wall-group particles j
 * within cut_phi (all pairs, minimum image); then bounce-back of the mobile
 * particles with phi_i > phi_c: v <- -v + 2 max(0, v.n) n,
 * x <- (x - v dtf/m) + dt v_new (fix_aboundary_rho_meso.cu:269-285, with its
 * dtf/m).  x, v FP64 in/out (3n); phi (n) and nw (3n, may be NULL) out. */
void orc_aboundary(int n, double *x, double *v, const int *type, const int *mask,
                   const orc_params *p, const orc_arb *arb, double *phi, double *nw);
/* mDPD force with the wall scaling (phi of each particle, masks); otherwise
 * as orc_force_mdpd */
void orc_force_mdpd_arb(int n, const float *x, const float *v, const double *rho,
                        const uint32_t *sig, const int *type, const int *mask,
                        const double *phi, const orc_params *p, const orc_arb *arb,
                        double *f, double *e, double *vir);

/* ---------------- step loop (FP64 state) ---------------- */

typedef struct {
    int      style;          /* 0 = mDPD, 1 = DPD */
    uint32_t seed;
    int64_t  ntimestep;
    double   skin;
    double   cut_rho;        /* mDPD density cutoff */
    int      lucy;
    int      rho_groupbit, nve_groupbit, addforce_groupbit;
    double   addforce[3];
    int      nthreads;       /* OpenMP threads (0 = runtime default) */
    const orc_arb *arb;      /* arbitrary boundary (NULL: off) */
    double  *phi;            /* n, wall fractions carried between steps (needed with arb) */
} orc_run_cfg;

/* ModifiedVerlet::setup + run(nsteps) with neigh_modify every 1 delay 0
 * check yes; mvv_meso.cu:203-283,309-572, fix_nve_meso.cu:63-95,157-178,
 * fix_addforce_meso.cu:72-108; with cfg->arb the wall branch runs after
 * initial integrate (post_integrate) and scales the force.  x, v (3n) are FP64 in/out, f out; the whole
 * state and all arithmetic are FP64 (the DPD signature reads the FP32
 * rounding of v, as dp2sp_merged does).  Returns the number of neighbor
 * builds including the setup build. */
int orc_run(int n, double *x, double *v, double *f, double *rho, const int *tag,
            const int *type, const int *mask, const orc_params *p,
            orc_run_cfg *cfg, int nsteps);

/* ---------------- CPU classic DPD (LAMMPS pair_style dpd) ---------------- */

/* Marsaglia RNG, src/random_mars.cpp:25-116 */
typedef struct { double u[98]; double c, cd, cm; int i97, j97; int save; double second; } orc_ranmars;
void   orc_ranmars_init(orc_ranmars *r, int seed);
double orc_ranmars_uniform(orc_ranmars *r);
double orc_ranmars_gaussian(orc_ranmars *r);

/* LAMMPS pair_style dpd + fix nve + Verlet, half neighbor list, newton on;
 * src/pair_dpd.cpp:63-160, src/fix_nve.cpp:64-130, src/verlet.cpp:212+.
 * nthreads <= 1: one serial Marsaglia stream (pair_dpd.cpp:198); nthreads > 1:
 * OpenMP threads with a force array and a stream RanMars(seed + t) each,
 * summed after the pair loop (the neighbor build and integration also use
 * nthreads; 0 leaves the OpenMP default for the build).  Thermo
 * accumulators: sum of T and P over steps sampled every `sample_every` steps
 * after `equil` steps; g(r) histogram with nrdf bins over [0, rc).  Returns
 * number of rebuilds. */
int orc_lammps_dpd_run(int n, double *x, double *v, const double lo[3], const double hi[3],
                       double a0, double gamma, double temperature, double rc,
                       double skin, double dt, int seed, int nsteps, int equil,
                       int sample_every, double *t_sum, double *p_sum, int *nsamples,
                       int nrdf, double *rdf_hist, int nthreads);

/* observables; compute_temp_meso.cu:49-101 (dof = 3N - 3),
 * compute_pressure_meso.cu:115-155 */
double orc_temperature(int n, const double *v, const int *type, const orc_params *p);
double orc_pressure(int n, const double *v, const double *vir, const int *type,
                    const orc_params *p);

/* ---------------- array helpers (test harness) ---------------- */
void orc_premix_n(int n, const uint32_t *a, const uint32_t *b, int rounds, uint32_t *out);
void orc_gaussian_n(int n, const uint32_t *a, const uint32_t *b, int libm, double *out);
void orc_sig_mdpd_n(int n, const int32_t *tag, uint32_t seed_t, uint32_t *out);
void orc_sig_dpd_n(int n, const int32_t *tag, uint32_t seed_t, const float *v, uint32_t *out);
void orc_sort_keys_n(int n, const float *x, const double lo[3], const double hi[3], double cutneigh,
                     uint64_t *key);

#ifdef __cplusplus
}
#endif
#endif
