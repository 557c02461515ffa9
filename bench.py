"""mDPD throughput benchmark on one MI355X: batched / served vs plain.

Default run (one JSON line; progress on stderr):

* batched (`value`): 256 independent periodic boxes, each a spherical mDPD
  droplet of N particles, N ~ LogNormal(ln 8000, 0.5) clipped to [2000, 40000],
  uniform in a sphere of radius R = (3N / (4 pi 6.05))^(1/3) at density 6.05 in
  a cube of side 2.5 R; Maxwell-Boltzmann velocities at kT = 1; A = -40,
  B = 25, r_c = 1, r_d = 0.75, gamma = 4.5, dt = 0.01, quadratic density
  weight, skin 0.3, neigh_modify every 1 delay 0 check yes; 2000 steps per
  job (`--job-steps`).  All jobs are admitted to one pool (inputs resident in
  HBM) before the clock starts; the timed region is the whole batched run
  (autotuned split, hipGraph-replayed passes, harvest of every job's
  results); `value` = total particle-steps / its wall time.  The plain path
  (one system per handle, its fused steps replayed from a hipGraph with the
  reneighbor decision on the device) runs a size-stratified sample of the
  same jobs for the same 2000 steps, timed per job after its setup.
* served: Poisson arrivals at 50 jobs/s for 120 s, each job a droplet or a
  slab (50/50) with N log-uniform in [1000, 200000], 2000 steps, admitted by
  the pool's HBM budget while others run; throughput and p50/p99 latency from
  arrival to result; a size-stratified sample of the same jobs on the plain
  path.
* slab: the plain liquid slab (box 40 x 40 x 20 with N = 96,800, and
  40 x 80 x 20 with N = 193,600; slab z in [5, 15) at density 6.05, Lucy
  weight), 5000 steps each, with the force-kernel and whole-step fractions of
  HBM peak from a separate profiled run.
* CPU baselines (GPU box host): the FP64 transcription of LAMMPS pair_style
  dpd + fix nve on the classic DPD fluid (L = 22, N = 32,000, a = 25,
  gamma = 4.5, dt = 0.04), on all physical cores and on one thread, and the
  FP64 mDPD oracle on the median droplet, each over a bounded sample.

`--steps K` and `--warmup W` keep the benchmark contract: W untimed passes
on a warm-up pool, and `ms_per_step` = the timed region / K; the jobs run
2000 steps whatever K is.

Other workloads (`--workload`): batched, served, slab, slab-literal, pore
(quasi-2D pore flow, 200 x 100 x 4, frozen disc walls, body force; ~480,000
particles), dpd (the classic DPD fluid on the GPU beside its CPU baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "usermeso-2.0-mdpd_amd"))

METRIC = "mDPD particle-steps/s, batched/served vs plain, 1 GPU; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
RHO, RC, RD, SKIN = 6.05, 1.0, 0.75, 0.3
NLIST = RHO * 4.0 / 3.0 * np.pi * (RC + SKIN) ** 3          # 55.7 entries per row
B_FORCE = 12 + 12 + 4 + 4 + 4 + 12 + 12 + 4 * NLIST          # bytes per particle, force + fused VV pass
B_DENSITY = 12 + 4 + 4 * NLIST                               # bytes per particle, density pass
B_STEP = B_FORCE + B_DENSITY                                 # 522 B per particle-step
# compulsory bytes of one neighbor build per particle: read the state
# (x, v, tag, mask, origin, image, phi: 52 B), write it sorted (52 B) and the
# build positions (16 B), write the row (4 B per entry) and its count
B_REBUILD = 52 + 52 + 16 + 4 * NLIST + 4                     # 347 B per particle-build
# HBM bytes per particle moved by the pooled force pass, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (medians over the passes of 64
# droplets, counters scaled by a 1 GiB copy: x2.0 fetch, x1.0 write)
T_FORCE_POOL = 174.3 + 32.1
GB = float(1 << 30)

SLABS = {
    "slab": dict(box=(40.0, 80.0, 20.0), slab=(5.0, 15.0)),
    "slab-literal": dict(box=(40.0, 40.0, 20.0), slab=(5.0, 15.0)),
}

DECK = """
    neighbor {skin} bin
    neigh_modify every 1 delay 0 check yes
    mass 1 1.0
    pair_style mdpd/meso 1.0 {rc} {seed}
    pair_coeff * * -40 25 4.5 {rc} {rd}
    fix rho all aboundary/rho/meso cut_rho {rd} {weight}
    fix int all nve/meso
    timestep 0.01
    run_style mvv/meso
"""

T_START = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def _mb(rng, n):
    v = rng.normal(0.0, 1.0, (n, 3))
    v -= v.mean(axis=0)
    return v


def _f32(a):
    return a.astype(np.float32).astype(np.float64)


def make_slab(box, slab, seed=20240601):
    lx, ly, lz = box
    n = int(round(RHO * lx * ly * (slab[1] - slab[0])))
    rng = np.random.default_rng(seed)
    x = np.empty((n, 3))
    x[:, 0] = rng.uniform(0, lx, n)
    x[:, 1] = rng.uniform(0, ly, n)
    x[:, 2] = rng.uniform(slab[0], slab[1], n)
    return _f32(x), _f32(_mb(rng, n))


def make_droplet(n, seed):
    """n particles uniform in a sphere at density 6.05; cube side 2.5 R."""
    rng = np.random.default_rng(seed)
    R = (3.0 * n / (4.0 * np.pi * RHO)) ** (1.0 / 3.0)
    L = 2.5 * R
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = R * rng.uniform(0, 1, n) ** (1.0 / 3.0)
    return (L, L, L), _f32(L / 2 + d * r[:, None]), _f32(_mb(rng, n))


def make_slab_job(n, seed):
    """n particles in a slab of thickness 10 (middle half of a box 20 high)."""
    rng = np.random.default_rng(seed)
    side = np.sqrt(n / (RHO * 10.0))
    box = (side, side, 20.0)
    x = np.empty((n, 3))
    x[:, 0] = rng.uniform(0, side, n)
    x[:, 1] = rng.uniform(0, side, n)
    x[:, 2] = rng.uniform(5.0, 15.0, n)
    return box, _f32(x), _f32(_mb(rng, n))


def make_pore(seed=20240603, lx=200.0, ly=100.0, lz=4.0, cover=0.40, rho=6.0):
    """Quasi-2D pore medium: seeded non-overlapping discs (radius U(4, 10),
    periodic in x and y) placed until they cover 40% of the x-y area, extruded
    in z; particles uniform at density 6, type 1 (frozen wall) inside a disc,
    type 2 (fluid) elsewhere."""
    rng = np.random.default_rng(seed)
    discs, area = [], 0.0
    while area < cover * lx * ly:
        r = rng.uniform(4.0, 10.0)
        c = rng.uniform((0, 0), (lx, ly))
        ok = True
        for (q, rq) in discs:
            d = c - q
            d -= np.array([lx, ly]) * np.round(d / np.array([lx, ly]))
            if np.hypot(*d) < r + rq:
                ok = False
                break
        if ok:
            discs.append((c, r))
            area += np.pi * r * r
    n = int(round(rho * lx * ly * lz))
    x = rng.uniform(0, 1, (n, 3)) * np.array([lx, ly, lz])
    wall = np.zeros(n, bool)
    for (q, r) in discs:
        d = x[:, :2] - q
        d -= np.array([lx, ly]) * np.round(d / np.array([lx, ly]))
        wall |= (d * d).sum(axis=1) < r * r
    types = np.where(wall, 1, 2).astype(np.int32)
    v = _mb(rng, n)
    v[wall] = 0.0
    v[~wall] -= v[~wall].mean(axis=0)
    return (lx, ly, lz), _f32(x), _f32(v), types, len(discs)


def batch_sizes(jobs, seed):
    rng = np.random.default_rng(seed)
    return np.clip(np.round(np.exp(rng.normal(np.log(8000.0), 0.5, jobs))), 2000, 40000).astype(int)


def system(box, x, v, seed, lucy=False):
    import mdpd355
    s = mdpd355.System(0)
    s.command(f"atom_style mdpd/atomic/meso {RHO}")
    s.create_box(1, (0, 0, 0), box)
    s.create_atoms(np.ones(len(x), np.int32), x, v)
    s.commands(DECK.format(skin=SKIN, rc=RC, rd=RD, seed=seed, weight="weight lucy" if lucy else ""))
    return s


# ---------------------------------------------------------------------------
# plain path: one handle per system
# ---------------------------------------------------------------------------
PLAIN_GRAPH = [True]          # --no-graph: plain runs launch every step's kernels directly


def plain_job(box, x, v, steps, seed, lucy=False, split=None):
    """Wall time of `run steps` after setup (the run returns after the
    stream has drained)."""
    s = system(box, x, v, seed, lucy)
    if split:
        s.tune("split", split)
    if not PLAIN_GRAPH[0]:
        s.tune("graph", 0)
    s.setup()
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    out = {"dt": dt, "nbuild": s.nbuild, "temp": s.thermo()["temp"]}
    s.close()
    return out


def plain_phases(box, x, v, steps, seed, lucy=False, split=None):
    """Per-phase kernel times (HIP events around every launch) of a separate
    run: a profiled run launches each step's kernels one by one."""
    s = system(box, x, v, seed, lucy)
    if split:
        s.tune("split", split)
    s.setup()
    s.profile(True)
    s.run(steps)
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    s.close()
    return prof


def plain_best_split(box, x, v, seed, steps=200, lucy=False):
    """Launch shape of the plain path for one system: the fastest of 1, 2, 4
    waves per tile over a short run."""
    best, times = None, {}
    for sp in (1, 2, 4):
        times[sp] = plain_job(box, x, v, steps, seed, lucy=lucy, split=sp)["dt"]
        if best is None or times[sp] < times[best]:
            best = sp
    return best, times


def plain_sample(sizes, spec, steps, k, split):
    """Plain runs of k jobs spread evenly over the size-sorted job list;
    spec(i) -> (box, x, v, seed) of job i."""
    order = np.argsort(sizes, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, min(k, len(order))).round().astype(int)]
    work = secs = 0.0
    for i in pick:
        box, x, v, seed = spec(int(i))
        r = plain_job(box, x, v, steps, seed, split=split)
        work += len(x) * steps
        secs += r["dt"]
    return work / secs, [int(sizes[i]) for i in pick], secs


# ---------------------------------------------------------------------------
# batched: one pool, all jobs resident
# ---------------------------------------------------------------------------
POOL_STATS = ("passes", "split", "capacity", "maxnbr", "bytes", "regrows", "prof_passes", "prof_particles",
              "prof_rebuild_ms", "prof_first_ms", "prof_force_ms", "prof_commit_ms", "tune_ms_1", "tune_ms_2",
              "tune_ms_4", "particle_steps", "particle_builds", "captures")


def batched(specs, job_steps, warmup, split, budget, graph=True):
    import mdpd355
    # warm-up: a small pool of the first jobs for `warmup` passes (loads
    # every kernel, builds a graph); not timed
    if warmup > 0:
        w = mdpd355.Pool(budget, 16)
        if not graph:
            w.tune("graph", 0)
        for box, x, v, seed in specs[:8]:
            s = system(box, x, v, seed)
            w.submit(s, warmup)
            s.close()
        w.run()
        w.close()
    p = mdpd355.Pool(budget, max(len(specs), 1))
    p.tune("split", split or 0)
    p.tune("profile", 16)
    if not graph:
        p.tune("graph", 0)
    ids = []
    for box, x, v, seed in specs:
        s = system(box, x, v, seed)
        ids.append(p.submit(s, job_steps))
        s.close()
    admitted = p.admit()
    assert admitted == len(specs), f"only {admitted} of {len(specs)} jobs fit the pool"
    log(f"batched: {len(specs)} jobs resident, running {job_steps} steps each")
    t0 = time.perf_counter()
    p.run()
    dt = time.perf_counter() - t0
    st = {k: p.stat(k) for k in POOL_STATS}
    temps = [p.thermo(j)["temp"] for j in ids[:: max(1, len(ids) // 16)]]
    p.close()
    return dt, st, temps


def sampled(st):
    npp = max(st["prof_passes"], 1.0)
    return {k: st[f"prof_{k}_ms"] / npp for k in ("rebuild", "first", "force", "commit")}


def rebuild_record(st, wall_s):
    """The rebuild phase apart: its sampled share of a pass, how many
    particle-builds the jobs made per particle-step, and its algorithmic rate."""
    ms = sampled(st)
    per_pass = sum(ms.values())
    t_rebuild = ms["rebuild"] * 1e-3 * st["passes"]
    bytes_ = st["particle_builds"] * B_REBUILD
    ach = bytes_ / t_rebuild / 1e9 if t_rebuild > 0 else None
    return {
        "ms_per_pass": ms["rebuild"], "share_of_pass_kernel_time": ms["rebuild"] / per_pass if per_pass else None,
        "particle_builds_per_particle_step": st["particle_builds"] / max(st["particle_steps"], 1.0),
        "bytes_per_particle_build": B_REBUILD,
        "achieved_GBps": ach, "frac": ach / HBM_PEAK_GBS if ach else None,
        "note": "rebuild kernels run every pass and exit for systems that are not stale; their time over all "
                "passes against the compulsory bytes of the builds the jobs made",
    }


def force_roofline(st, kernel, traffic_per_particle=None):
    """k_force of the pool pass: algorithmic bytes of the particles the sampled
    passes evaluated over the force phase's event time in those passes."""
    t = st["prof_force_ms"] * 1e-3
    parts = st["prof_particles"]
    if t <= 0 or parts <= 0:
        return None
    achieved = B_FORCE * parts / t / 1e9
    frac = achieved / HBM_PEAK_GBS
    if frac > 1.0:
        raise RuntimeError(f"force roofline above peak ({achieved:.0f} GB/s): the sampled passes are not "
                           f"what their particle count says")
    per = parts / max(st["prof_passes"], 1.0)
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": frac, "traffic": None if traffic_per_particle is None else traffic_per_particle * per,
            "bytes_per_particle": B_FORCE, "particles_per_launch": per,
            "avg_launch_us": st["prof_force_ms"] / max(st["prof_passes"], 1.0) * 1e3,
            "sampled_passes": st["prof_passes"]}


# ---------------------------------------------------------------------------
# served: Poisson arrivals into a pool with an HBM budget
# ---------------------------------------------------------------------------
def served(seconds, rate, steps, budget, seed, max_wall, split=0, graph=True):
    import mdpd355
    rng = np.random.default_rng(seed)
    arr = []
    t = rng.exponential(1.0 / rate)
    while t < seconds:
        arr.append(t)
        t += rng.exponential(1.0 / rate)
    kinds = rng.random(len(arr)) < 0.5
    sizes = np.round(np.exp(rng.uniform(np.log(1000.0), np.log(200000.0), len(arr)))).astype(int)

    def spec(i):
        box, x, v = make_droplet(sizes[i], 7000 + i) if kinds[i] else make_slab_job(sizes[i], 7000 + i)
        return box, x.astype(np.float32), v.astype(np.float32)
    # the clients make their synthetic particles ahead of their arrival
    # times (four producer threads, in arrival order); building the system
    # and submitting it (the engine's API) is timed
    from concurrent.futures import ThreadPoolExecutor
    log(f"served: {len(arr)} jobs ({int(sizes.sum())} particles) made by 4 client threads")
    producers = ThreadPoolExecutor(4)
    inputs = [producers.submit(spec, i) for i in range(len(arr))]
    if arr:
        inputs[min(len(arr), 64) - 1].result()          # a head start before the clock
    p = mdpd355.Pool(budget, 1024)
    p.tune("poll", 4)
    p.tune("profile", 16)
    p.tune("split", split)                     # 0: autotuned
    if not graph:
        p.tune("graph", 0)
    meta, lat, work = {}, [], 0.0
    occ = []                                   # (used slot fraction, running jobs, queued jobs) per loop
    t_adv = 0.0                                # wall time inside advance()
    log(f"served: {seconds:g} s of arrivals at {rate:g}/s, {steps} steps per job")
    t0 = mdpd355.clock()
    next_log = 15.0
    truncated = False
    i = 0
    while True:
        now = mdpd355.clock() - t0
        while i < len(arr) and arr[i] <= now:
            box, x, v = inputs[i].result()
            inputs[i] = None
            s = system(box, x.astype(np.float64), v.astype(np.float64), 9000 + i)
            meta[p.submit(s, steps)] = (t0 + arr[i], len(x))
            s.close()
            i += 1
        if p.stat("running") + p.stat("queued") > 0:
            ta = mdpd355.clock()
            p.advance(max_passes=8)
            t_adv += mdpd355.clock() - ta
            occ.append((p.stat("used_slots") / max(p.stat("capacity"), 1.0), p.stat("running"), p.stat("queued")))
        elif i < len(arr):
            time.sleep(min(0.002, max(0.0, arr[i] - now)))
        for j in p.done():
            info = p.info(j)
            a, n = meta.pop(j)
            lat.append(info["t_done"] - a)
            work += n * steps
            p.release(j)
        if i >= len(arr) and not meta:
            break
        if now > next_log:
            log(f"served: t={now:.0f}s arrived {i}/{len(arr)} done {len(lat)} running {p.stat('running'):.0f} "
                f"queued {p.stat('queued'):.0f}")
            next_log += 15.0
        if now > max_wall:
            truncated = True
            log(f"served: stopping at the {max_wall:g} s wall limit with {len(meta)} jobs unfinished")
            break
    wall = mdpd355.clock() - t0
    for f in inputs[i:]:
        f.cancel()
    producers.shutdown(wait=True)
    st = {k: p.stat(k) for k in POOL_STATS}
    p.close()
    lat = np.array(lat)
    out = {
        "arrivals": len(arr), "completed": int(len(lat)), "truncated_at_wall_limit": truncated,
        "window_s": seconds, "rate_per_s": rate, "steps_per_job": steps, "budget_GB": budget / GB,
        "wall_s": wall, "value": work / wall if wall > 0 else 0.0,
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "offered_particle_steps_per_s": float(sizes.sum()) * steps / seconds if len(arr) else 0.0,
        "mean_used_slot_fraction": float(np.mean([o[0] for o in occ])) if occ else None,
        "mean_running_jobs": float(np.mean([o[1] for o in occ])) if occ else None,
        "mean_queued_jobs": float(np.mean([o[2] for o in occ])) if occ else None,
        "advance_wall_fraction": t_adv / wall if wall > 0 else None,
        "passes": st["passes"], "ms_per_pass": wall * 1e3 / max(st["passes"], 1.0),
        "mean_particles_per_pass": st["particle_steps"] / max(st["passes"], 1.0),
        "graph_captures": st["captures"], "split": st["split"],
        "sampled_ms_per_pass": sampled(st),
        "rebuild": rebuild_record(st, wall),
    }
    roof = force_roofline(st, "k_force<mDPD, MODE_PHASED> (served pool pass)")
    if roof:
        out["force_kernel"] = {k: roof[k] for k in ("achieved", "frac", "avg_launch_us", "particles_per_launch")}
    def spec64(i):
        box, x, v = spec(i)
        return box, x.astype(np.float64), v.astype(np.float64), 9000 + i
    return out, sizes, spec64


# ---------------------------------------------------------------------------
# CPU baselines (host of the GPU box)
# ---------------------------------------------------------------------------
def host_cpus():
    """CPUs this process may run on, their physical cores, model, SMT."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    smt = None
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    phys = len(cores)
    if omp and omp.isdigit():
        phys = min(phys, int(omp))      # the box's CPU share for this process
    return {"nproc": len(cpus), "physical_cores": phys, "model": model, "smt_active": smt,
            "omp_num_threads_env": omp}


def cpu_c1(seconds, threads):
    """FP64 LAMMPS pair_style dpd + fix nve transcription on the classic DPD
    fluid (config 1), `threads` OpenMP threads, a bounded number of steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    L, n = 22.0, 32000
    rng = np.random.default_rng(20241203)
    x = _f32(rng.uniform(0, L, (n, 3)))
    v = _f32(_mb(rng, n))
    kw = dict(dt=0.04, seed=4321, nthreads=threads)
    t0 = time.perf_counter()
    orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=3, **kw)
    per = (time.perf_counter() - t0) / 3
    steps = int(max(3, min(1000, seconds / per)))
    t0 = time.perf_counter()
    out = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=steps, **kw)
    t = time.perf_counter() - t0
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 transcription of LAMMPS pair_style dpd + fix nve (half list, newton on, Marsaglia "
                      f"streams) on the config-1 DPD fluid L=22 N={n} a=25 gamma=4.5 dt=0.04, {steps} of its "
                      f"1000 steps from the random start incl. neighbor builds, {t:.1f} s wall, {threads} "
                      f"thread(s)", "builds": out["nbuild"]}


def cpu_mdpd(box, x, v, lucy, label, seconds, threads, seed=4242):
    """The FP64 mDPD oracle step (OpenMP) on one of the benchmark's systems."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    p = orc.params((0, 0, 0), box, 0.01, cut=RC, cut_r=RD, A=-40.0, B=25.0, gamma=4.5,
                   sigma=np.sqrt(9.0), expw=1.0)
    n = len(x)
    args = (np.arange(1, n + 1), np.zeros(n, int), np.ones(n, int), p)
    kw = dict(style=0, seed=seed, skin=SKIN, cut_rho=RD, lucy=lucy, nthreads=threads)
    t0 = time.perf_counter()
    orc.run(x, v, *args, 0, **kw)                         # setup only
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.run(x, v, *args, 2, **kw)
    per = max((time.perf_counter() - t0 - t_setup) / 2, 1e-6)
    steps = int(max(2, min(6000, seconds / per)))
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    t = time.perf_counter() - t0 - t_setup
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 mDPD oracle step (OpenMP, {threads} threads) on {label} N={n}, {steps} steps after "
                      f"setup, {t:.1f} s wall (setup {t_setup:.1f} s excluded); the reference has no CPU mDPD"}


def cpu_baselines(a, droplet=None):
    hw = host_cpus()
    cores = max(1, hw["physical_cores"])
    log(f"cpu: config-1 DPD on {cores} cores and 1 thread ({hw['model']}, nproc {hw['nproc']})")
    allc = cpu_c1(a.cpu_seconds, cores)
    one = cpu_c1(a.cpu_seconds, 1)
    out = dict(allc)
    out["host"] = hw
    out["one_thread"] = {k: one[k] for k in ("value", "cores", "sample")}
    if droplet is not None:
        box, x, v, label = droplet
        log("cpu: mDPD oracle on the median droplet")
        out["mdpd_oracle"] = cpu_mdpd(box, x, v, False, label, a.cpu_seconds, cores)
    return out


# ---------------------------------------------------------------------------
# slab records (plain path)
# ---------------------------------------------------------------------------
def slab_record(name, steps, split=None, profile_steps=200):
    w = SLABS[name]
    box = w["box"]
    x, v = make_slab(box, w["slab"])
    n = len(x)
    if split is None:
        split, _ = plain_best_split(box, x, v, 4242, steps=100, lucy=True)
    plain_job(box, x, v, 20, 4242, lucy=True, split=split)      # warm-up
    r = plain_job(box, x, v, steps, 4242, lucy=True, split=split)
    value = n * steps / r["dt"]
    prof = plain_phases(box, x, v, profile_steps, 4242, lucy=True, split=split)
    f_ms, f_n = prof["force"]
    t_force = f_ms / max(f_n, 1) * 1e-3
    fk = B_FORCE * n / t_force / 1e9
    return {
        "workload": f"plain mDPD liquid slab: box {box[0]:g}x{box[1]:g}x{box[2]:g}, slab z in [5,15) at rho 6.05, "
                    f"N={n}, A=-40 B=25 rc=1 rd=0.75 gamma=4.5 kT=1 dt=0.01, Lucy weight, skin 0.3, {steps} steps",
        "particles": n, "steps": steps, "value": value, "split": split, "builds": r["nbuild"],
        "step_algorithmic_GBps": value * B_STEP / 1e9, "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
        "force_kernel_GBps": fk, "force_kernel_frac": fk / HBM_PEAK_GBS, "force_kernel_us": t_force * 1e6,
        "phase_ms_per_step": {k: ms / max(profile_steps, 1) for k, (ms, _) in prof.items()},
        "final_temp": r["temp"],
    }


def base_line(a, value, ms_per_step, config, roof, cpu):
    return {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic", "config": config, "roofline": roof,
        "cpu_baseline": cpu,
    }


def run_headline(a, legs):
    sizes = batch_sizes(a.jobs, a.seed)
    specs = [(*make_droplet(int(n), a.seed + 1 + k), 5000 + k) for k, n in enumerate(sizes)]
    total = int(sizes.sum())
    config = {"workload": f"batched mDPD droplets: {a.jobs} periodic boxes, N ~ LogNormal(ln 8000, 0.5) clipped "
                          f"[2000, 40000] (total {total}), droplet at rho 6.05 in a cube of side 2.5 R, A=-40 "
                          f"B=25 rc=1 rd=0.75 gamma=4.5 kT=1 dt=0.01, quadratic weight, skin 0.3, "
                          f"{a.job_steps} steps each, one pool (hipGraph-replayed passes, autotuned split)",
              "mode": "batched", "jobs": a.jobs, "particles": total, "steps_per_job": a.job_steps}
    value, roof, cpu, ms_per_step = None, None, None, None
    if "batched" in legs:
        dt, st, temps = batched(specs, a.job_steps, a.warmup, a.split, a.budget_gb * GB, graph=not a.no_graph)
        value = total * a.job_steps / dt
        ms_per_step = dt * 1e3 / a.steps
        log(f"batched: {value:.3e} particle-steps/s over {dt:.2f} s")
        med = int(np.argsort(sizes)[len(sizes) // 2])
        box, x, v, seed = specs[med]
        psplit, ptimes, pv, pns, ps = None, {}, None, [], 0.0
        if a.plain_sample > 0:
            psplit, ptimes = plain_best_split(box, x, v, seed)
            pv, pns, ps = plain_sample(sizes, lambda i: specs[i], a.job_steps, a.plain_sample, psplit)
        log(f"plain: {pv:.3e} particle-steps/s (split {psplit})" if pv else "plain: skipped")
        config.update({
            "timed_region_s": dt, "ms_per_step_definition": "timed region / --steps (the region is the whole "
                                                            "batch of 2000-step jobs)",
            "passes": st["passes"], "ms_per_pass": dt * 1e3 / max(st["passes"], 1.0),
            "plain_value": pv, "batched_vs_plain": value / pv if pv else None,
            "plain_sample": f"{len(pns)} of the same jobs (sizes {pns}) on one handle each, {a.job_steps} steps, "
                            f"fused steps replayed from a hipGraph with the reneighbor decision on the device, "
                            f"split {psplit} (fastest over 200 steps of the median job: "
                            f"{ {k: round(t, 4) for k, t in ptimes.items()} } s), {ps:.1f} s",
            "split": st["split"], "pool_capacity": st["capacity"], "maxnbr": st["maxnbr"],
            "pool_GB": st["bytes"] / GB, "regrows": st["regrows"],
            "tune_first_force_ms_per_pass": {"1": st["tune_ms_1"], "2": st["tune_ms_2"], "4": st["tune_ms_4"]},
            "sampled_ms_per_pass": sampled(st), "sampled_passes": st["prof_passes"],
            "rebuild": rebuild_record(st, dt),
            "step_algorithmic_GBps": value * B_STEP / 1e9,
            "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
            "final_temp_sample": temps,
        })
        roof = force_roofline(st, "k_force<mDPD, MODE_PHASED> (batched pool pass)", T_FORCE_POOL)
    if "slab" in legs:
        config["slab"] = []
        for name in ("slab-literal", "slab"):
            log(f"slab: {name}")
            config["slab"].append(slab_record(name, a.slab_steps))
    if "wall" in legs:
        log("wall")
        config["wall"] = wall_record(a.job_steps)
    if "obs" in legs:
        log("observables")
        config["observables"] = observables_record(a.job_steps)
    if "cpu" in legs:
        med = int(np.argsort(sizes)[len(sizes) // 2])
        box, x, v, _ = specs[med]
        cpu = cpu_baselines(a, (box, x, v, "the median droplet of the batch"))
    if "served" in legs:
        srv, ssizes, sspec = served(a.served_seconds, a.served_rate, a.job_steps, a.served_budget_gb * GB,
                                    a.seed + 99, a.served_max_wall, split=a.served_split, graph=not a.no_graph)
        log(f"served: {srv['value']:.3e} particle-steps/s, p50 {srv['p50_latency_s']} s, p99 {srv['p99_latency_s']} s")
        if a.plain_sample > 0:
            k = max(4, a.plain_sample // 2)
            order = np.argsort(ssizes, kind="stable")
            mid = int(order[len(order) // 2])
            box, x, v, seed = sspec(mid)
            ssplit, _ = plain_best_split(box, x, v, seed)
            sv, sns, ss = plain_sample(ssizes, sspec, a.job_steps, k, ssplit)
            srv.update({"plain_value": sv, "plain_sample_sizes": sns, "plain_split": ssplit,
                        "served_vs_plain": srv["value"] / sv})
        config["served"] = srv
        if value is None:
            value = srv["value"]
            ms_per_step = srv["wall_s"] * 1e3 / a.steps
            config["mode"] = "served"
    print(json.dumps(base_line(a, value, ms_per_step, config, roof, cpu)), flush=True)


def run_pore(a):
    import mdpd355
    box, x, v, types, ndisc = make_pore()
    n = len(x)
    nwall = int((types == 1).sum())

    def handle():
        s = mdpd355.System(0)
        s.command("atom_style mdpd/atomic/meso 6.0")
        s.create_box(2, (0, 0, 0), box)
        s.create_atoms(types, x, v)
        s.commands(f"""
            neighbor {SKIN} bin
            neigh_modify every 1 delay 0 check yes
            mass * 1.0
            group fluid type 2
            pair_style mdpd/meso 1.0 {RC} 4242
            pair_coeff 1 1 -40 25 4.5 {RC} {RD}
            pair_coeff 1 2 -40 0 4.5 {RC} {RD}
            pair_coeff 2 2 -40 25 4.5 {RC} {RD}
            fix rho all aboundary/rho/meso cut_rho {RD}
            fix push fluid addforce/meso 0.02 0 0
            fix int fluid nve/meso
            timestep 0.01
            run_style mvv/meso
        """)
        if a.split:
            s.tune("split", a.split)
        return s

    if a.warmup > 0:
        s = handle()
        s.run(a.warmup)
        s.close()
    s = handle()
    s.setup()
    t0 = time.perf_counter()
    s.run(a.job_steps)
    dt = time.perf_counter() - t0
    vx = s.gather("v")[types == 2, 0].mean()
    nb = s.nbuild
    s.close()
    s = handle()
    s.setup()
    s.profile(True)
    s.run(200)
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    s.close()
    value = n * a.job_steps / dt
    f_ms, f_n = prof["force"]
    t_force = f_ms / max(f_n, 1) * 1e-3
    config = {
        "workload": f"plain mDPD quasi-2D pore flow: box 200x100x4, {ndisc} seeded non-overlapping discs "
                    f"(radius U(4,10)) covering 40% of the area, frozen wall particles (type 1, outside the NVE "
                    f"group, in the density) and fluid (type 2) at density 6, N={n} ({nwall} wall), A=-40, "
                    f"B=25 (solid-fluid B=0), gamma=4.5, rc=1, rd=0.75, g_x=0.02 on the fluid, dt=0.01, "
                    f"skin 0.3, one system per launch, {a.job_steps} steps",
        "mode": "plain", "particles": n, "wall_particles": nwall, "builds": nb, "timed_region_s": dt,
        "phase_ms_per_step": {k: ms / 200 for k, (ms, _) in prof.items()},
        "fluid_mean_vx_end": float(vx),
        "step_algorithmic_GBps": value * B_STEP / 1e9,
    }
    achieved = B_FORCE * n / t_force / 1e9
    roof = {"bound": "hbm", "kernel": "k_force<mDPD, fused velocity-Verlet, 2 types>", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_particle": B_FORCE, "particles_per_launch": n, "avg_launch_us": t_force * 1e6}
    print(json.dumps(base_line(a, value, dt / a.steps * 1e3, config, roof, None)))


# ---------------------------------------------------------------------------
# wall: the arbitrary-boundary branch of fix aboundary/rho/meso (wall
# fractions, wall-scaled pair forces, bounce-back) on the plain path
# ---------------------------------------------------------------------------
WALL_BOX = (40.0, 12.0, 40.0)
WALL_DECK = """
    neighbor {skin} bin
    neigh_modify every 1 delay 0 check yes
    mass * 1.0
    group fluid type 1
    group wall type 2
    pair_style mdpd/meso 1.0 {rc} 2654868 fluid wall
    pair_coeff 1 1 -40 25 18.0 {rc} {rd}
    pair_coeff 1 2 -20 25 18.0 {rc} {rd}
    pair_coeff 2 2 -40 25 18.0 {rc} {rd}
    fix BC all aboundary/rho/meso cut_rho {rd} phi_c 0.5 cut_phi 1.0 rho_wall 8.0 wall wall mobile fluid
    fix int fluid nve/meso
    timestep 0.01
    run_style mvv/meso
"""


def make_channel(seed=20240605):
    """Slit channel: frozen wall particles (type 2) on a simple cubic lattice of
    spacing 1/2 (density 8) in y in [0, 1) and [11, 12), fluid (type 1) at
    density 6.05 uniform in y in [1, 11) with Maxwell velocities."""
    lx, ly, lz = WALL_BOX
    rng = np.random.default_rng(seed)
    g = lambda lo, hi: np.arange(lo, hi, 0.5) + 0.25
    walls = np.stack(np.meshgrid(g(0.0, lx), np.r_[g(0.0, 1.0), g(ly - 1.0, ly)], g(0.0, lz), indexing="ij"),
                     -1).reshape(-1, 3)
    nf = int(round(RHO * lx * (ly - 2.0) * lz))
    fluid = np.c_[rng.uniform(0, lx, nf), rng.uniform(1.0, ly - 1.0, nf), rng.uniform(0, lz, nf)]
    v = np.zeros((nf + len(walls), 3))
    v[:nf] = _mb(rng, nf)
    types = np.r_[np.ones(nf, np.int32), np.full(len(walls), 2, np.int32)]
    return _f32(np.r_[fluid, walls]), _f32(v), types, nf


def wall_record(steps, profile_steps=200, split=None):
    import mdpd355
    x, v, types, nf = make_channel()
    n = len(x)

    def handle(sp):
        s = mdpd355.System(0)
        s.command("atom_style mdpd/atomic/meso 8.0")
        s.create_box(2, (0, 0, 0), WALL_BOX)
        s.create_atoms(types, x, v)
        s.commands(WALL_DECK.format(skin=SKIN, rc=RC, rd=RD))
        if sp:
            s.tune("split", sp)
        if not PLAIN_GRAPH[0]:
            s.tune("graph", 0)
        return s

    def timed_run(sp, k):
        s = handle(sp)
        s.setup()
        t0 = time.perf_counter()
        s.run(k)
        dt = time.perf_counter() - t0
        out = (dt, s.nbuild, s.thermo()["temp"])
        s.close()
        return out

    if split is None:
        times = {sp: timed_run(sp, 100)[0] for sp in (1, 2, 4)}
        split = min(times, key=times.get)
    timed_run(split, 20)                                            # warm-up
    dt, nb, temp = timed_run(split, steps)
    s = handle(split)
    s.setup()
    origin = s.debug_sorted()[0]
    counts = np.asarray(s.debug_neighbors()[0])[:n]
    nrow = float(np.mean(counts[np.asarray(origin)[:n] < nf]))        # rows of the fluid particles
    s.profile(True)
    s.run(profile_steps)
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial", "wall")}
    s.close()
    value = n * steps / dt
    w_ms, w_n = prof["wall"]
    t_wall = w_ms / max(w_n, 1) * 1e-3
    # the wall pass per fluid particle: its full row (4 B per entry), position,
    # mask and build position read, wall fraction written
    b_wall = 4.0 * nrow + 16 + 4 + 16 + 4
    return {
        "workload": f"plain mDPD slit channel with the arbitrary-boundary branch: box "
                    f"{WALL_BOX[0]:g}x{WALL_BOX[1]:g}x{WALL_BOX[2]:g}, frozen lattice walls (density 8) in y in "
                    f"[0,1) and [11,12), fluid at density 6.05 between, N={n} ({n - nf} wall), A11=A22=-40 "
                    f"A12=-20 B=25 gamma=18 rc=1 rd=0.75, phi_c 0.5, cut_phi 1, rho_wall 8, dt=0.01, skin 0.3, "
                    f"{steps} steps",
        "particles": n, "wall_particles": n - nf, "steps": steps, "value": value, "timed_region_s": dt,
        "split": split, "builds": nb,
        "phase_ms_per_step": {k: ms / max(profile_steps, 1) for k, (ms, _) in prof.items()},
        "wall_pass_us": t_wall * 1e6, "mean_row_entries_fluid": nrow, "wall_pass_bytes_per_fluid_particle": b_wall,
        "wall_pass_GBps": b_wall * nf / t_wall / 1e9 if t_wall > 0 else None,
        "final_temp": temp,
    }


def observables_record(steps):
    """The device observables (fix rdf/fast/meso every 10 steps, fix
    density/meso every 100 steps over a 10-step window) on the 96,800-particle
    slab: the same run with and without them."""
    import tempfile
    w = SLABS["slab-literal"]
    box = w["box"]
    x, v = make_slab(box, w["slab"])
    n = len(x)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for tag, extra in (("without", ""),
                           ("with", f"fix g all rdf/fast/meso output {tmp}/rdf.out nbin 100 every 10\n"
                                    f"fix p all density/meso output {tmp}/prof along z nbin 40 every 100 window 10")):
            best = None
            for rep in range(2):
                s = system(box, x, v, 4242, lucy=True)
                s.tune("split", 4)
                if extra:
                    s.commands(extra)
                s.setup()
                t0 = time.perf_counter()
                s.run(steps)
                dt = time.perf_counter() - t0
                s.close()
                best = dt if best is None else min(best, dt)
            out[tag] = best
    return {
        "workload": f"plain mDPD liquid slab 40x40x20 (N={n}, Lucy weight, split 4), {steps} steps, with and "
                    f"without fix rdf/fast/meso (100 bins, every 10 steps) and fix density/meso (40 bins along z, "
                    f"every 100 steps, 10-step window)",
        "particles": n, "steps": steps,
        "value_without": n * steps / out["without"], "value_with": n * steps / out["with"],
        "overhead_us_per_step": (out["with"] - out["without"]) / steps * 1e6,
    }


def run_wall(a):
    r = wall_record(a.job_steps)
    log(f"wall: {r['value']:.3e} particle-steps/s, wall pass {r['wall_pass_us']:.1f} us")
    achieved = r["wall_pass_GBps"]
    roof = {"bound": "hbm", "kernel": "k_aboundary (wall fractions and bounce-back)", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": None, "bytes_per_particle": r["wall_pass_bytes_per_fluid_particle"],
            "particles_per_launch": r["particles"] - r["wall_particles"], "avg_launch_us": r["wall_pass_us"]}
    config = dict(r, mode="plain")
    print(json.dumps(base_line(a, r["value"], r["timed_region_s"] * 1e3 / a.steps, config, roof, None)))


def run_dpd(a):
    """Config 1: the classic DPD fluid (L = 22, N = 32,000, a = 25, gamma = 4.5,
    dt = 0.04) on the GPU (dpd/meso, sigma = 3) beside the CPU LAMMPS
    pair_style dpd + fix nve transcription."""
    import mdpd355
    L, n = 22.0, 32000
    rng = np.random.default_rng(20241203)
    x = _f32(rng.uniform(0, L, (n, 3)))
    v = _f32(_mb(rng, n))

    def handle():
        s = mdpd355.System(0)
        s.command("atom_style dpd/atomic/meso")
        s.create_box(1, (0, 0, 0), (L, L, L))
        s.create_atoms(np.ones(n, np.int32), x, v)
        s.commands("""
            neighbor 0.3 bin
            neigh_modify every 1 delay 0 check yes
            mass 1 1.0
            pair_style dpd/meso 1.0 4242
            pair_coeff * * 25.0 4.5 3.0 1.0 1.0
            fix int all nve/meso
            timestep 0.04
            run_style mvv/meso
        """)
        return s

    s = handle()
    s.run(a.warmup)
    t0 = time.perf_counter()
    s.run(1000)
    dt = time.perf_counter() - t0
    th = s.thermo()
    s.close()
    s = handle()
    s.setup()
    s.profile(True)
    s.run(200)
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    s.close()
    value = n * 1000 / dt
    cpu = None if a.no_cpu else cpu_baselines(a)
    b_dpd = 56 + 4 * 3.0 * 4.0 / 3.0 * np.pi * 1.3**3
    f_ms, f_n = prof["force"]
    config = {
        "workload": f"plain DPD fluid (config 1): periodic cube L=22, N={n} uniform, a=25 gamma=4.5 sigma=3 "
                    f"rc=1 dt=0.04, skin 0.3, dpd/meso on one handle, 1000 steps",
        "mode": "plain", "particles": n, "final_temp": th["temp"], "final_press": th["press"],
        "timed_region_s": dt, "phase_ms_per_step": {k: ms / 200 for k, (ms, _) in prof.items()},
    }
    roof = {"bound": "hbm", "kernel": "k_force<DPD, fused velocity-Verlet>",
            "achieved": b_dpd * n / (f_ms / max(f_n, 1) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "traffic": None, "bytes_per_particle": b_dpd}
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    print(json.dumps(base_line(a, value, dt / a.steps * 1e3, config, roof, cpu)))


def run_slab(a):
    rec = slab_record(a.workload, a.slab_steps, split=a.split)
    w = SLABS[a.workload]
    cpu = None
    if not a.no_cpu:
        x, v = make_slab(w["box"], w["slab"])
        cpu = cpu_mdpd(w["box"], x, v, True, "the same slab", a.cpu_seconds, max(1, host_cpus()["physical_cores"]))
    roof = {"bound": "hbm", "kernel": "k_force<mDPD, fused velocity-Verlet>", "achieved": rec["force_kernel_GBps"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rec["force_kernel_frac"], "traffic": None,
            "bytes_per_particle": B_FORCE, "particles_per_launch": rec["particles"],
            "avg_launch_us": rec["force_kernel_us"]}
    config = dict(rec)
    config["mode"] = "plain"
    print(json.dumps(base_line(a, rec["value"], rec["particles"] * rec["steps"] / rec["value"] * 1e3 / a.steps,
                               config, roof, cpu)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="K of the contract: ms_per_step = timed region / K")
    ap.add_argument("--warmup", type=int, default=5, help="untimed passes of a warm-up pool")
    ap.add_argument("--job-steps", type=int, default=2000, help="steps of every batched / served job")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "batched", "served", "pore", "dpd", "wall"] + sorted(SLABS))
    ap.add_argument("--legs", default="batched,slab,wall,obs,cpu,served",
                    help="legs of the headline workload (batched, slab, wall, obs, cpu, served)")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20241202)
    ap.add_argument("--budget-gb", type=float, default=16.0, help="HBM budget of the batched pool")
    ap.add_argument("--plain-sample", type=int, default=16, help="jobs of the batch timed on the plain path")
    ap.add_argument("--served-seconds", type=float, default=120.0, help="arrival window of the served leg")
    ap.add_argument("--served-rate", type=float, default=50.0)
    ap.add_argument("--served-budget-gb", type=float, default=4.0)
    ap.add_argument("--served-max-wall", type=float, default=300.0,
                    help="the served leg stops here (reported as truncated) so that the run keeps its time limit")
    ap.add_argument("--served-split", type=int, default=0, choices=[0, 1, 2, 4],
                    help="waves per tile of the served pool (0: autotuned)")
    ap.add_argument("--slab-steps", type=int, default=5000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU work per baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each pass's (pools) and step's (plain runs) kernels directly instead of replaying a "
                         "hipGraph")
    ap.add_argument("--split", type=int, default=None, help="waves per tile (default: autotuned)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device with a plain copy kernel (a known-size "
                         "reference for HBM byte counters under rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    PLAIN_GRAPH[0] = not a.no_graph
    if a.calibrate_copy_mib > 0:
        import torch
        src = torch.ones(a.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {a.calibrate_copy_mib} MiB", file=sys.stderr)
        del src, dst
    if a.workload == "headline":
        legs = set(a.legs.split(","))
        if a.no_cpu:
            legs.discard("cpu")
        run_headline(a, legs)
    elif a.workload == "batched":
        run_headline(a, {"batched"} | (set() if a.no_cpu else {"cpu"}))
    elif a.workload == "served":
        run_headline(a, {"served"})
    elif a.workload == "pore":
        run_pore(a)
    elif a.workload == "wall":
        run_wall(a)
    elif a.workload == "dpd":
        run_dpd(a)
    else:
        run_slab(a)


if __name__ == "__main__":
    main()
