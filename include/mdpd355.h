/*
 * mdpd355.h -- C ABI of the MI355X engine for USER-MESO's per-timestep
 * mDPD / DPD particle path (pair_style mdpd/meso and dpd/meso, fix nve/meso,
 * fix aboundary/rho/meso, fix addforce/meso, run_style mvv/meso).
 *
 * The boundary keeps the shape of the reference's own embedding interface,
 * the LAMMPS library API (src/library.h:27-42), restricted to the styles of
 * this path: a handle is opened, driven with input-script commands, filled
 * with atoms, run, and read back.  All functions return an MDPD_* status;
 * the text of the last error of a handle is available from mdpd_last_error.
 * No torch or HIP types cross the boundary: plain pointers and sizes only.
 * The handle owns all device memory; host buffers passed in are copied.
 * One host thread drives a handle at a time; distinct handles are
 * independent.
 */
#ifndef MDPD355_H
#define MDPD355_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MDPD_OK = 0,
    MDPD_E_ARG = 1,         /* bad argument or script syntax (error->all in the reference) */
    MDPD_E_STATE = 2,       /* call out of order (e.g. run before pair_coeff) */
    MDPD_E_HIP = 3,         /* HIP runtime failure */
    MDPD_E_OVERFLOW = 4,    /* a device table could not hold the system */
    MDPD_E_NOMEM = 5,       /* allocation failed */
    MDPD_E_UNSUPPORTED = 6  /* command or style outside this path */
};

typedef struct mdpd_system mdpd_system;

/* ---- lifecycle -------------------------------------------------------- */

/* Open a handle on HIP device `device`.  Replaces lammps_open
 * (src/library.h:27-28) + MesoDevice::init (src/USER-MESO/engine_meso.cu:38-95). */
int mdpd_open(int device, mdpd_system **out);
/* Release the handle and its device memory.  lammps_close (src/library.h:29) */
void mdpd_close(mdpd_system *sys);
/* Text of the last error on this handle ("" if none).  The reference raises
 * error->all / SIGABRT instead (src/USER-MESO/util_meso.h:100-101). */
const char *mdpd_last_error(const mdpd_system *sys);
const char *mdpd_version(void);

/* ---- input-script surface --------------------------------------------- */

/* Execute one input-script line.  lammps_command (src/library.h:31) ->
 * Input::execute_command (src/input.cpp:542-579).  Accepted: units lj,
 * dimension 3, boundary p p p, atom_style mdpd/atomic/meso <rho_est> |
 * dpd/atomic/meso, mass, group ID type|id ..., pair_style mdpd/meso T rc seed,
 * pair_style dpd/meso rc seed, pair_coeff, fix ID G nve/meso,
 * fix ID G aboundary/rho/meso cut_rho rd [weight quadratic|lucy]
 *   [cut_phi c phi_c p rho_wall w wall G mobile G],
 * pair_style mdpd/meso T rc seed <mobile group> <wall group>,
 * fix ID G addforce/meso fx fy fz, fix ID G rdf/fast/meso output F nbin N
 * [every M] [other G] [pi 3.1415|exact] (pi: an option of this library, the
 * constant of the shell volumes; 3.1415, the reference's, by default),
 * fix ID G density/meso output F along x|y|z nbin N
 * every M window W, read_data F, write_data F, write_restart F,
 * read_restart F (checkpoint: every atom as the record of
 * AtomVecMDPDAtomic::pack_restart, atom_vec_mdpd_atomic_meso.cu:548-570, in
 * this library's own header with step, dt, box, masses, groups and pair
 * settings; fixes and run_style are given again after read_restart),
 * velocity G set vx vy vz,
 * group ID union G..., unfix, timestep, neighbor skin bin,
 * neigh_modify every/delay/check, run_style mvv/meso, reset_timestep,
 * thermo, run N.  Comments (#) and blank lines are
 * ignored; echo/log/print/newton/comm_modify/atom_modify are accepted and
 * ignored. */
int mdpd_command(mdpd_system *sys, const char *line);
/* Execute every line of a text buffer.  lammps_file (src/library.h:30) */
int mdpd_commands_string(mdpd_system *sys, const char *text);

/* ---- box and atoms ------------------------------------------------------ */

/* create_box: orthogonal periodic box with ntypes atom types */
int mdpd_create_box(mdpd_system *sys, int ntypes, const double lo[3], const double hi[3]);
/* Append n atoms: tag (NULL: next ids), type 1..ntypes, mask (NULL: group
 * "all" only), x and v as n*3 doubles (v may be NULL).  Positions and
 * velocities are stored in FP32 on the device.  Atoms are only accepted
 * before the first setup/run, or after mdpd_clear_atoms. */
int mdpd_create_atoms(mdpd_system *sys, int n, const int *tag, const int *type,
                      const int *mask, const double *x, const double *v);
int mdpd_clear_atoms(mdpd_system *sys);
/* lammps_get_natoms (src/library.h:40) */
int mdpd_get_natoms(const mdpd_system *sys);
/* bitmask of a group name (0 if unknown) */
int mdpd_group_bit(const mdpd_system *sys, const char *name);

/* ---- run ---------------------------------------------------------------- */

/* Verlet setup: wrap, sort, neighbor build, density and forces of the
 * current state (ModifiedVerlet::setup, src/USER-MESO/mvv_meso.cu:203-283). */
int mdpd_setup(mdpd_system *sys);
/* run N (ModifiedVerlet::run, src/USER-MESO/mvv_meso.cu:309-572); performs
 * setup first if needed.  Rebuilds follow neigh_modify every/delay/check
 * with an on-device displacement test. */
int mdpd_run(mdpd_system *sys, long long nsteps);
long long mdpd_get_timestep(const mdpd_system *sys);
/* number of neighbor-list builds so far */
long long mdpd_get_nbuild(const mdpd_system *sys);

/* ---- results ------------------------------------------------------------- */

/* lammps_gather_atoms (src/library.h:41): copy a per-atom quantity in the
 * order atoms were created.  type 0 = int, 1 = double; count values per atom.
 * Names: "x" (1,3) "v" (1,3) "f" (1,3) "rho" (1,1) "phi" (1,1: wall
 * fraction of the arbitrary-boundary branch) "e_pair" (1,1)
 * "virial" (1,6: xx yy zz xy xz yz) "tag" (0,1) "type" (0,1) "mask" (0,1)
 * "image" (0,3).  f, e_pair and virial are those of the last force pass of
 * setup or of the last step of a run. */
int mdpd_gather_atoms(mdpd_system *sys, const char *name, int type, int count, void *data);
/* lammps_scatter_atoms (src/library.h:42) for "x" and "v" (type 1, count 3):
 * overwrite positions / velocities in creation order; the next run repeats
 * setup. */
int mdpd_scatter_atoms(mdpd_system *sys, const char *name, int type, int count, const void *data);
/* thermo of the last force pass: out[0] temperature (compute temp/meso,
 * dof = 3N-3), out[1] pressure (compute pressure/meso), out[2] pair energy,
 * out[3] kinetic energy.  compute_temp_meso.cu:49-101,
 * compute_pressure_meso.cu:115-155 */
int mdpd_thermo(mdpd_system *sys, double out[4]);

/* Current result of an observable fix, per bin: x = bin centre, value =
 * g(r) for fix rdf/fast/meso (the reference's normalization with 3.1415,
 * fix_rdf_fast_meso.cu:177-220; sampled after the force pass every `every`
 * steps, fix_rdf_fast_meso.cu:140-175) or the mean count per bin of the
 * current window for fix density/meso (fix_density_profile_meso.cpp:153-215).
 * *nbin receives the bin count; with n = 0 only *nbin is written.  The fixes
 * also write their output files as the reference does: rdf when the fix ends
 * (unfix, redefinition, mdpd_close), density at the end of each window as
 * <output>.<step>. */
int mdpd_fix_result(mdpd_system *sys, const char *fix_id, int n, double *x, double *value, int *nbin);

/* ---- tuning ----------------------------------------------------------------
 * Launch-shape knobs of the per-step passes.  No counterpart in the reference,
 * whose launch sizes come from occupancy tables (src/USER-MESO/occupancy_meso.cu).
 * "split": waves sharing one 64-particle tile's neighbor rows in the density
 * and force passes (1, 2 or 4; default 2).  Takes effect at the next pass;
 * results differ from another split only by FP32 summation order.
 * "graph" 1/0: replay the fused steps of a run from a captured hipGraph
 * (default 1) or launch their kernels one by one; results are identical. */
int mdpd_tune(mdpd_system *sys, const char *knob, int value);

/* ---- kernel timing ---------------------------------------------------------
 * HIP events recorded around every launch on the handle's stream, folded into
 * per-phase totals at the end of each setup / run.  Phases: "rebuild"
 * (wrap, keys, sort, gather, neighbor rows), "density", "force" (pair force
 * with the fused integration), "initial" (first half-kick of a run), "wall"
 * (wall fractions and bounce-back of fix aboundary/rho/meso's wall branch).  The
 * reference's counterpart is MesoTimer::device_stamp and the -profile switch
 * (src/USER-MESO/timer_meso.cu:14-19, src/lammps.cpp:189-191). */
int mdpd_profile(mdpd_system *sys, int enable);          /* also resets the totals */
int mdpd_profile_get(mdpd_system *sys, const char *phase, double *total_ms, long long *launches);

/* ---- parity hooks: the engine's internal tables --------------------------
 * These expose what the reference keeps in meso_atom / MesoNeighList device
 * arrays (atom_meso.cu:271-387, neigh_build_meso.cu:17-127), in the engine's
 * sorted order, for bit-exact comparison with the formulas. */

/* Sorted order of the last build: origin[k] = creation index of the particle
 * in sorted slot k; key[k] = its 2-level Morton sort key (border bit 0);
 * xw = its wrapped FP32 position at that build (3 per slot).  Any pointer may
 * be NULL. */
int mdpd_debug_sorted(mdpd_system *sys, int *origin, unsigned long long *key, float *xw);
/* Neighbor rows in sorted order: counts[k], rows[k*maxnbr + p] (entry =
 * sorted index j | image << 27, image = (sx+1)+3(sy+1)+9(sz+1)).  *maxnbr
 * receives the row width; rows may be NULL to query it. */
int mdpd_debug_neighbors(mdpd_system *sys, int *maxnbr, int *counts, int *rows);
/* TEA signatures of the next force pass, in creation order. */
int mdpd_debug_signatures(mdpd_system *sys, unsigned int *sig);

/* Stateless device evaluation (no handle): n pair Gaussians
 * gaussian_TEA<4>(a > b, a, b) (math_meso.h:465-473) in FP32, and n
 * premix_TEA<rounds>(a, b) values (math_meso.h:459-463). */
int mdpd_tea_gaussian(int device, int n, const unsigned int *a, const unsigned int *b, float *out);
int mdpd_tea_premix(int device, int n, const unsigned int *a, const unsigned int *b, int rounds,
                    unsigned int *out);

/* mdpd_tea_gaussian without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  Inputs and outputs are staged in a
 * fixed device array that each GPU holds until the process ends
 * (3 x MDPD_TEA_SCRATCH_MAX words, 12 MiB); n > MDPD_TEA_SCRATCH_MAX returns
 * MDPD_E_OVERFLOW before any device work.  All calls on one GPU share that
 * array: calls that overlap on the same GPU race. */
#define MDPD_TEA_SCRATCH_MAX (1 << 20)
int mdpd_tea_gaussian_noalloc(int device, int n, const unsigned int *a, const unsigned int *b, float *out);

/* ---- pooled engine (batched and served runs) -------------------------------
 * Many independent systems ("jobs") packed into one set of device arrays and
 * advanced together, one launch sequence per pass for the whole pool.  The
 * reference's closest counterpart is its ensemble mode, which runs
 * independent cases as separate MPI communicators (src/main.cpp:49-117), and
 * each job follows ModifiedVerlet::run (src/USER-MESO/mvv_meso.cu:309-572)
 * exactly as mdpd_run does on a handle of its own, with the reneighbor
 * decision (neigh_modify every/delay/check, src/neighbor.cpp:1216-1290) made
 * per job on the device.
 *
 * A job is submitted from a configured handle (script state + atoms are
 * copied; the handle stays the caller's and is not changed), is admitted when
 * its particles and cells fit the pool's HBM budget (in arrival order; a job
 * that does not fit yet lets later jobs that do fit pass it, at most 64
 * times before it goes first),
 * keeps its state resident while it runs, and leaves when done; its results
 * stay readable until mdpd_pool_release (92 bytes per particle slot of host
 * memory, pinned while finished jobs hold less than 1 GiB of it).  One host
 * thread drives a pool. */
typedef struct mdpd_pool mdpd_pool;

enum { MDPD_JOB_QUEUED = 0, MDPD_JOB_RUNNING = 1, MDPD_JOB_DONE = 2 };

/* hbm_budget_bytes: device memory the pool may hold (its particle capacity is
 * derived from it at the first admission); max_systems: jobs resident at once. */
int mdpd_pool_open(int device, double hbm_budget_bytes, int max_systems, mdpd_pool **out);
void mdpd_pool_close(mdpd_pool *pool);
const char *mdpd_pool_last_error(const mdpd_pool *pool);
/* Queue `sys` for nsteps steps (like `run nsteps` on it); *job = its id. */
int mdpd_pool_submit(mdpd_pool *pool, mdpd_system *sys, long long nsteps, int *job);
/* Drive the device: admit queued jobs, run passes, harvest finished jobs.
 * Returns after max_passes passes (0: no limit), as soon as a job finished
 * when stop_on_done, or when no job is left; *ndone = jobs finished. */
int mdpd_pool_advance(mdpd_pool *pool, long long max_passes, int stop_on_done, int *ndone);
/* Run until every submitted job is done. */
int mdpd_pool_run(mdpd_pool *pool);
/* Admit the queued jobs that fit now and wait until their state is resident
 * on the device (no pass is run); *admitted = jobs admitted. */
int mdpd_pool_admit(mdpd_pool *pool, int *admitted);
int mdpd_pool_job_state(mdpd_pool *pool, int job, int *state);
/* Ids of jobs finished since the last call (up to max). */
int mdpd_pool_done(mdpd_pool *pool, int *jobs, int max, int *count);
/* Per-atom results of a finished job, creation order, as mdpd_gather_atoms:
 * x, v, f (type 1, count 3), e_pair, rho (1, 1), virial (1, 6), tag, type,
 * mask (0, 1), image (0, 3).  f, e_pair and virial are those of the last
 * force evaluation, as after `run` on a handle. */
int mdpd_pool_gather(mdpd_pool *pool, int job, const char *name, int type, int count, void *data);
/* {T, P, pe, ke} of a finished job, as mdpd_thermo. */
int mdpd_pool_thermo(mdpd_pool *pool, int job, double out[4]);

/* Result of a finished job's observable fix.  No counterpart in the
 * reference (LAMMPS runs one system).  fix rdf/fast/meso: g(r) as
 * mdpd_fix_result returns it for a handle that ran the job (same bins,
 * counts and normalization; sampled after the force pass of every step that
 * is a multiple of `every`, fix_rdf_fast_meso.cu:140-175).  fix density/meso:
 * the profile of the job's last window (see mdpd_pool_fix_windows).  A job
 * takes at most two rdf fixes and one density fix whose windows x (nbin + 1)
 * stay within 8,192 words; pool jobs write no output files.  The counts take
 * 64 KiB of device memory per resident-job slot, outside the HBM budget,
 * from the first job that has such a fix. */
int mdpd_pool_fix_result(mdpd_pool *pool, int job, const char *fix_id, int n, double *x, double *value, int *nbin);

/* Every window of a finished job's fix density/meso: centre[w] is the
 * window's centre step (the step in the name of the file a handle writes,
 * fix_density_profile_meso.cpp:192-215) and value[w * nbin + b] the mean
 * count in bin b over the window's samples.  *nwin and *nbin receive the
 * sizes; with max_windows = 0 only they are written.  No counterpart in the
 * reference. */
int mdpd_pool_fix_windows(mdpd_pool *pool, int job, const char *fix_id, int max_windows, long long *centre,
                          double *value, int *nwin, int *nbin);
/* Atoms, final timestep, and submit / admit / done times (seconds on
 * mdpd_clock) of a job. */
int mdpd_pool_job_info(mdpd_pool *pool, int job, long long *natoms, long long *timestep, double times[3]);
int mdpd_pool_release(mdpd_pool *pool, int job);
/* "split" 0 (autotune on the first passes, once more after 256 passes and
 * whenever the resident load has doubled or halved since, at most every 256
 * passes; the default), 1, 2 or 4;
 * "graph" 1/0: replay each pass as a captured hipGraph (default 1);
 * "poll": passes between status reads (default 8);
 * "hold" MiB: pinned host memory finished, unreleased jobs may keep their
 * results in (default 1024); past it a finished job's results move to
 * pageable memory;
 * "profile" N: every N-th batch of passes is launched without the graph with
 * HIP events between its phases (0: off, the default; also resets the totals). */
int mdpd_pool_tune(mdpd_pool *pool, const char *knob, int value);
/* Statistics: passes, particle_steps (of finished jobs), split, capacity
 * (particle slots), maxnbr, bytes (device bytes held), regrows, running,
 * queued, used_slots, tune_ms_1 / tune_ms_2 / tune_ms_4 (ms of the first and
 * force passes per pass, the split-dependent part the autotuner compares),
 * prof_passes and prof_{rebuild,first,force,commit}_ms (sampled totals),
 * held_pinned (pinned host bytes holding finished, unreleased results). */
int mdpd_pool_stat(mdpd_pool *pool, const char *name, double *value);
/* Monotonic clock of the job times, in seconds. */
double mdpd_clock(void);

#ifdef __cplusplus
}
#endif
#endif
