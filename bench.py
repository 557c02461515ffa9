"""mDPD throughput benchmark on one MI355X: batched / served vs plain.

Default workload ("batched"): 256 independent periodic boxes, each a
spherical mDPD droplet of N particles, N ~ LogNormal(ln 8000, 0.5) clipped to
[2000, 40000], uniform in a sphere of radius R = (3N / (4 pi 6.05))^(1/3) at
density 6.05 in a cube of side 2.5 R; Maxwell-Boltzmann velocities at kT = 1;
A = -40, B = 25, r_c = 1, r_d = 0.75, gamma = 4.5, dt = 0.01, quadratic
density weight, skin 0.3, neigh_modify every 1 delay 0 check yes; `--steps`
(2000) steps per job.  All jobs are admitted to one pool (inputs resident in
HBM) before the timed region; the timed region is the whole batched run
(autotuned split, hipGraph-replayed passes, harvest of every job's results).
`value` is total particle-steps / wall time.  The plain path (one system per
handle, one launch sequence per system and step) runs a stratified sample of
the same jobs for the same number of steps, timed per job after its setup.

The served leg (`--served-seconds`, 0 to skip) submits jobs with Poisson
arrivals (`--served-rate` per second): each a droplet or a slab (50/50), N
log-uniform in [1000, 200000], `--steps` steps, admitted by the pool's HBM
budget while others run; it reports throughput and p50/p99 latency from
arrival to result.  The jobs' synthetic particles are made before the clock
starts (client data); building each system and submitting it is timed.

Other workloads on the plain path (one handle): `slab` (box 40 x 80 x 20,
slab z in [5, 15) at density 6.05, N = 193,600, Lucy weight; `slab-literal`:
40 x 40 x 20, N = 96,800), `pore` (quasi-2D pore flow, 200 x 100 x 4, frozen
disc walls, body force; ~480,000 particles) and `dpd` (classic DPD fluid,
L = 22, N = 32,000, beside the CPU LAMMPS pair_style dpd transcription).

Prints one JSON line.
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
# HBM bytes per particle moved by the pooled force pass, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (medians over the passes of 64
# droplets, counters scaled by a 1 GiB copy: x2.0 fetch, x1.0 write):
# 181.4 read + 32.1 written
T_FORCE_POOL = 181.4 + 32.1
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
def plain_job(box, x, v, steps, seed, lucy=False, split=None, profile=False):
    s = system(box, x, v, seed, lucy)
    if split:
        s.tune("split", split)
    s.setup()
    if profile:
        s.profile(True)
    t0 = time.perf_counter()
    s.run(steps)                      # returns after the stream has drained
    dt = time.perf_counter() - t0
    out = {"dt": dt, "nbuild": s.nbuild}
    if profile:
        out["prof"] = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
        out["thermo"] = s.thermo()
    s.close()
    return out


def plain_sample(sizes, spec, steps, k):
    """Plain runs of k jobs spread evenly over the size-sorted job list;
    spec(i) -> (box, x, v, seed) of job i."""
    order = np.argsort(sizes, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, min(k, len(order))).round().astype(int)]
    work = secs = 0.0
    for i in pick:
        box, x, v, seed = spec(int(i))
        r = plain_job(box, x, v, steps, seed)
        work += len(x) * steps
        secs += r["dt"]
    return work / secs, [int(sizes[i]) for i in pick], secs


# ---------------------------------------------------------------------------
# batched: one pool, all jobs resident
# ---------------------------------------------------------------------------
def batched(specs, steps, warmup, split, budget, graph=True):
    import mdpd355
    # warm-up: a small pool of the first jobs (loads every kernel, builds a graph)
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
        ids.append(p.submit(s, steps))
        s.close()
    admitted = p.admit()
    assert admitted == len(specs), f"only {admitted} of {len(specs)} jobs fit the pool"
    t0 = time.perf_counter()
    p.run()
    dt = time.perf_counter() - t0
    st = {k: p.stat(k) for k in ("passes", "split", "capacity", "maxnbr", "bytes", "regrows", "prof_passes",
                                 "prof_rebuild_ms", "prof_first_ms", "prof_force_ms", "prof_commit_ms",
                                 "tune_ms_1", "tune_ms_2", "tune_ms_4")}
    temps = [p.thermo(j)["temp"] for j in ids[:: max(1, len(ids) // 16)]]
    p.close()
    return dt, st, temps


# ---------------------------------------------------------------------------
# served: Poisson arrivals into a pool with an HBM budget
# ---------------------------------------------------------------------------
def served(seconds, rate, steps, budget, seed, max_wall=240.0, split=0, graph=True):
    import mdpd355
    rng = np.random.default_rng(seed)
    arr = []
    t = rng.exponential(1.0 / rate)
    while t < seconds:
        arr.append(t)
        t += rng.exponential(1.0 / rate)
    kinds = rng.random(len(arr)) < 0.5
    sizes = np.round(np.exp(rng.uniform(np.log(1000.0), np.log(200000.0), len(arr)))).astype(int)
    spec = lambda i: (make_droplet(sizes[i], 7000 + i) if kinds[i] else make_slab_job(sizes[i], 7000 + i))
    # the clients' inputs exist before they arrive: the synthetic particles
    # are made before the clock starts; building the system and submitting
    # it (the engine's API) is timed
    inputs = [spec(i) for i in range(len(arr))]
    p = mdpd355.Pool(budget, 1024)
    p.tune("poll", 4)
    p.tune("profile", 16)
    p.tune("split", split)                     # 0: autotuned
    if not graph:
        p.tune("graph", 0)
    meta, lat, work = {}, [], 0.0
    occ = []                                   # (used slot fraction, running jobs, queued jobs) per loop
    t_adv = 0.0                                # wall time inside advance()
    t0 = mdpd355.clock()
    i = 0
    while True:
        now = mdpd355.clock() - t0
        while i < len(arr) and arr[i] <= now:
            box, x, v = inputs[i]
            inputs[i] = None
            s = system(box, x, v, 9000 + i)
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
        if mdpd355.clock() - t0 > max_wall:
            break
    wall = mdpd355.clock() - t0
    st = {k: p.stat(k) for k in ("passes", "particle_steps", "captures", "split", "prof_passes", "prof_rebuild_ms",
                                 "prof_first_ms", "prof_force_ms", "prof_commit_ms")}
    p.close()
    lat = np.array(lat)
    npp = max(st["prof_passes"], 1.0)
    out = {
        "arrivals": len(arr), "completed": int(len(lat)), "window_s": seconds, "rate_per_s": rate,
        "steps_per_job": steps, "budget_GB": budget / GB, "wall_s": wall,
        "value": work / wall if wall > 0 else 0.0,
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
        "sampled_ms_per_pass": {k: st[f"prof_{k}_ms"] / npp for k in ("rebuild", "first", "force", "commit")},
    }
    return out, sizes, (lambda i: (*spec(i), 9000 + i))


# ---------------------------------------------------------------------------
# CPU baseline: the oracle's FP64 OpenMP step on one job, bounded sample
# ---------------------------------------------------------------------------
def cpu_baseline(box, x, v, lucy, label, target_s=12.0, seed=4242):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    threads = min(16, os.cpu_count() or 1)
    p = orc.params((0, 0, 0), box, 0.01, cut=RC, cut_r=RD, A=-40.0, B=25.0, gamma=4.5,
                   sigma=np.sqrt(9.0), expw=1.0)
    n = len(x)
    args = (np.arange(1, n + 1), np.zeros(n, int), np.ones(n, int), p)
    kw = dict(style=0, seed=seed, skin=SKIN, cut_rho=RD, lucy=lucy, nthreads=threads)
    t0 = time.perf_counter()
    orc.run(x, v, *args, 0, **kw)                         # setup only
    t_setup = time.perf_counter() - t0
    steps = 2
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    per = max((time.perf_counter() - t0 - t_setup) / steps, 1e-6)
    steps = int(max(2, min(6000, target_s / per)))
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    t = time.perf_counter() - t0 - t_setup
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 mDPD step (OpenMP, {threads} threads) on {label} N={n}, {steps} steps "
                      f"after setup, {t:.1f} s wall (setup {t_setup:.1f} s excluded)"}


def roofline(t_force_s, particles, kernel, traffic_per_particle=None):
    achieved = B_FORCE * particles / t_force_s / 1e9
    traffic = None if traffic_per_particle is None else traffic_per_particle * particles
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_particle": B_FORCE,
            "particles_per_launch": particles, "avg_launch_us": t_force_s * 1e6}


def base_line(a, value, ms_per_step, config, roof, cpu):
    return {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic", "config": config, "roofline": roof,
        "cpu_baseline": cpu,
    }


def run_batched(a):
    sizes = batch_sizes(a.jobs, a.seed)
    specs = [(*make_droplet(int(n), a.seed + 1 + k), 5000 + k) for k, n in enumerate(sizes)]
    total = int(sizes.sum())
    dt, st, temps = batched(specs, a.steps, a.warmup, a.split, a.budget_gb * GB, graph=not a.no_graph)
    value = total * a.steps / dt
    plain_value, plain_ns, plain_s = (plain_sample(sizes, lambda i: specs[i], a.steps, a.plain_sample)
                                      if a.plain_sample > 0 else (None, [], 0.0))
    srv = None
    if a.served_seconds > 0:
        srv, ssizes, sspec = served(a.served_seconds, a.served_rate, a.steps, a.served_budget_gb * GB, a.seed + 99,
                                    split=a.served_split, graph=not a.no_graph)
        if a.plain_sample > 0:
            sv, sns, ss = plain_sample(ssizes, sspec, a.steps, max(4, a.plain_sample // 2))
            srv.update({"plain_value": sv, "plain_sample_sizes": sns, "served_vs_plain": srv["value"] / sv})
    t_force = st["prof_force_ms"] / max(st["prof_passes"], 1) * 1e-3
    config = {
        "workload": f"batched mDPD droplets: {a.jobs} periodic boxes, N ~ LogNormal(ln 8000, 0.5) clipped "
                    f"[2000, 40000] (total {total}), droplet at rho 6.05 in a cube of side 2.5 R, A=-40 B=25 "
                    f"rc=1 rd=0.75 gamma=4.5 kT=1 dt=0.01, quadratic weight, skin 0.3, {a.steps} steps each, "
                    f"one pool (hipGraph-replayed passes, autotuned split)",
        "mode": "batched", "jobs": a.jobs, "particles": total,
        "plain_value": plain_value, "batched_vs_plain": value / plain_value if plain_value else None,
        "plain_sample": f"{len(plain_ns)} of the same jobs (sizes {plain_ns}) on one handle each, "
                        f"{a.steps} steps, {plain_s:.1f} s",
        "passes": st["passes"], "split": st["split"], "pool_capacity": st["capacity"], "maxnbr": st["maxnbr"],
        "pool_GB": st["bytes"] / GB, "regrows": st["regrows"],
        "tune_first_force_ms_per_pass": {"1": st["tune_ms_1"], "2": st["tune_ms_2"], "4": st["tune_ms_4"]},
        "sampled_ms_per_pass": {k: st[f"prof_{k}_ms"] / max(st["prof_passes"], 1)
                                for k in ("rebuild", "first", "force", "commit")},
        "sampled_passes": st["prof_passes"],
        "step_algorithmic_GBps": value * B_STEP / 1e9,
        "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
        "final_temp_sample": temps,
        "served": srv,
    }
    roof = roofline(t_force, total, "k_force<mDPD, MODE_PHASED> (pool pass, all systems)", T_FORCE_POOL)
    cpu = None
    if not a.no_cpu:
        med = int(np.argsort(sizes)[len(sizes) // 2])
        box, x, v, _ = specs[med]
        cpu = cpu_baseline(box, x, v, False, "the median droplet of the batch")
    print(json.dumps(base_line(a, value, dt / max(st["passes"], 1) * 1e3, config, roof, cpu)))


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
    s.profile(True)
    t0 = time.perf_counter()
    s.run(a.steps)
    dt = time.perf_counter() - t0
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    vx = s.gather("v")[types == 2, 0].mean()
    nb = s.nbuild
    s.close()
    value = n * a.steps / dt
    f_ms, f_n = prof["force"]
    config = {
        "workload": f"plain mDPD quasi-2D pore flow: box 200x100x4, {ndisc} seeded non-overlapping discs "
                    f"(radius U(4,10)) covering 40% of the area, frozen wall particles (type 1, outside the NVE "
                    f"group, in the density) and fluid (type 2) at density 6, N={n} ({nwall} wall), A=-40, "
                    f"B=25 (solid-fluid B=0), gamma=4.5, rc=1, rd=0.75, g_x=0.02 on the fluid, dt=0.01, "
                    f"skin 0.3, one system per launch",
        "mode": "plain", "particles": n, "wall_particles": nwall, "builds": nb,
        "phase_ms_per_step": {k: ms / max(a.steps, 1) for k, (ms, _) in prof.items()},
        "fluid_mean_vx_end": float(vx),
        "step_algorithmic_GBps": value * B_STEP / 1e9,
    }
    roof = roofline(f_ms / max(f_n, 1) * 1e-3, n, "k_force<mDPD, fused velocity-Verlet, 2 types>")
    print(json.dumps(base_line(a, value, dt / a.steps * 1e3, config, roof, None)))


def run_dpd(a):
    """Config 1: the classic DPD fluid (L = 22, N = 32,000, a = 25, gamma = 4.5,
    dt = 0.04) on the GPU (dpd/meso, sigma = 3) beside the CPU LAMMPS
    pair_style dpd + fix nve transcription (serial: one Marsaglia stream)."""
    import mdpd355
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    L, n = 22.0, 32000
    rng = np.random.default_rng(a.seed)
    x = _f32(rng.uniform(0, L, (n, 3)))
    v = _f32(_mb(rng, n))
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
    s.run(a.warmup)
    s.profile(True)
    t0 = time.perf_counter()
    s.run(a.steps)
    dt = time.perf_counter() - t0
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    th = s.thermo()
    s.close()
    value = n * a.steps / dt
    cpu = None
    if not a.no_cpu:
        steps = 2
        t0 = time.perf_counter()
        orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=steps, dt=0.04, seed=4321)
        per = (time.perf_counter() - t0) / steps
        steps = int(max(2, min(1000, 12.0 / per)))
        t0 = time.perf_counter()
        out = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=steps, dt=0.04, seed=4321)
        t = time.perf_counter() - t0
        cpu = {"value": n * steps / t, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"LAMMPS pair_style dpd + fix nve transcription (FP64, half list, serial RanMars) on "
                         f"the same N={n} fluid, {steps} steps incl. the initial build, {t:.1f} s wall"}
    b_dpd = 56 + 4 * 3.0 * 4.0 / 3.0 * np.pi * 1.3**3
    f_ms, f_n = prof["force"]
    config = {
        "workload": f"plain DPD fluid (config 1): periodic cube L=22, N={n} uniform, a=25 gamma=4.5 sigma=3 "
                    f"rc=1 dt=0.04, skin 0.3, dpd/meso on one handle",
        "mode": "plain", "particles": n, "final_temp": th["temp"], "final_press": th["press"],
        "phase_ms_per_step": {k: ms / max(a.steps, 1) for k, (ms, _) in prof.items()},
    }
    roof = {"bound": "hbm", "kernel": "k_force<DPD, fused velocity-Verlet>",
            "achieved": b_dpd * n / (f_ms / max(f_n, 1) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "traffic": None, "bytes_per_particle": b_dpd}
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    print(json.dumps(base_line(a, value, dt / a.steps * 1e3, config, roof, cpu)))


def run_slab(a):
    w = SLABS[a.workload]
    box = w["box"]
    x, v = make_slab(box, w["slab"])
    n = len(x)
    if a.warmup > 0:
        plain_job(box, x, v, a.warmup, 4242, lucy=True, split=a.split)
    r = plain_job(box, x, v, a.steps, 4242, lucy=True, split=a.split, profile=True)
    dt, prof = r["dt"], r["prof"]
    value = n * a.steps / dt
    f_ms, f_n = prof["force"]
    config = {
        "workload": f"plain mDPD liquid slab ({a.workload}): box {box[0]:g}x{box[1]:g}x{box[2]:g}, slab z in "
                    f"[5,15) at rho 6.05, N={n}, A=-40 B=25 rc=1 rd=0.75 gamma=4.5 kT=1 dt=0.01, Lucy weight, "
                    f"skin 0.3, one system per launch",
        "mode": "plain", "particles": n, "rebuilds_in_timed_region": r["nbuild"],
        "phase_ms_per_step": {k: ms / max(a.steps, 1) for k, (ms, _) in prof.items()},
        "step_algorithmic_GBps": value * B_STEP / 1e9,
        "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
        "final_temp": r["thermo"]["temp"],
    }
    roof = roofline(f_ms / max(f_n, 1) * 1e-3, n, "k_force<mDPD, fused velocity-Verlet>")
    cpu = None if a.no_cpu else cpu_baseline(box, x, v, True, "the same slab")
    print(json.dumps(base_line(a, value, dt / a.steps * 1e3, config, roof, cpu)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="steps per job")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="batched", choices=["batched", "pore", "dpd"] + sorted(SLABS))
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20241202)
    ap.add_argument("--budget-gb", type=float, default=16.0, help="HBM budget of the batched pool")
    ap.add_argument("--plain-sample", type=int, default=16, help="jobs of the batch timed on the plain path")
    ap.add_argument("--served-seconds", type=float, default=10.0, help="arrival window of the served leg (0: skip)")
    ap.add_argument("--served-rate", type=float, default=50.0)
    ap.add_argument("--served-budget-gb", type=float, default=4.0)
    ap.add_argument("--served-split", type=int, default=0, choices=[0, 1, 2, 4],
                    help="waves per tile of the served pool (0: autotuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched and served pools: launch each pass's kernels directly instead of replaying a hipGraph")
    ap.add_argument("--split", type=int, default=None, help="waves per tile (default: autotuned / 2)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device with a plain copy kernel (a known-size "
                         "reference for HBM byte counters under rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy_mib > 0:
        import torch
        src = torch.ones(a.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {a.calibrate_copy_mib} MiB", file=sys.stderr)
        del src, dst
    if a.workload == "batched":
        run_batched(a)
    elif a.workload == "pore":
        run_pore(a)
    elif a.workload == "dpd":
        run_dpd(a)
    else:
        run_slab(a)


if __name__ == "__main__":
    main()
