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
arrival to result.

`--workload slab` runs the plain mDPD liquid slab instead (box 40 x 80 x 20,
slab z in [5, 15) at density 6.05, N = 193,600, Lucy weight; `slab-literal`:
40 x 40 x 20, N = 96,800) on one handle.

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
def batched(specs, steps, warmup, split, budget):
    import mdpd355
    # warm-up: a small pool of the first jobs (loads every kernel, builds a graph)
    if warmup > 0:
        w = mdpd355.Pool(budget, 16)
        for box, x, v, seed in specs[:8]:
            s = system(box, x, v, seed)
            w.submit(s, warmup)
            s.close()
        w.run()
        w.close()
    p = mdpd355.Pool(budget, max(len(specs), 1))
    p.tune("split", split or 0)
    p.tune("profile", 16)
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
def served(seconds, rate, steps, budget, seed, max_wall=240.0):
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
    p = mdpd355.Pool(budget, 1024)
    p.tune("split", 2)
    p.tune("poll", 4)
    meta, lat, work = {}, [], 0.0
    t0 = mdpd355.clock()
    i = 0
    while True:
        now = mdpd355.clock() - t0
        while i < len(arr) and arr[i] <= now:
            box, x, v = spec(i)
            s = system(box, x, v, 9000 + i)
            meta[p.submit(s, steps)] = (t0 + arr[i], len(x))
            s.close()
            i += 1
        if p.stat("running") + p.stat("queued") > 0:
            p.advance(max_passes=8)
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
    p.close()
    lat = np.array(lat)
    out = {
        "arrivals": len(arr), "completed": int(len(lat)), "window_s": seconds, "rate_per_s": rate,
        "steps_per_job": steps, "budget_GB": budget / GB, "wall_s": wall,
        "value": work / wall if wall > 0 else 0.0,
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "offered_particle_steps_per_s": float(sizes.sum()) * steps / seconds if len(arr) else 0.0,
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
    steps = int(max(2, min(2000, target_s / per)))
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    t = time.perf_counter() - t0 - t_setup
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 mDPD step (OpenMP, {threads} threads) on {label} N={n}, {steps} steps "
                      f"after setup, {t:.1f} s wall (setup {t_setup:.1f} s excluded)"}


def roofline(t_force_s, particles, kernel):
    achieved = B_FORCE * particles / t_force_s / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_particle": B_FORCE,
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
    dt, st, temps = batched(specs, a.steps, a.warmup, a.split, a.budget_gb * GB)
    value = total * a.steps / dt
    plain_value, plain_ns, plain_s = (plain_sample(sizes, lambda i: specs[i], a.steps, a.plain_sample)
                                      if a.plain_sample > 0 else (None, [], 0.0))
    srv = None
    if a.served_seconds > 0:
        srv, ssizes, sspec = served(a.served_seconds, a.served_rate, a.steps, a.served_budget_gb * GB, a.seed + 99)
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
        "tune_ms_per_pass": {"1": st["tune_ms_1"], "2": st["tune_ms_2"], "4": st["tune_ms_4"]},
        "sampled_ms_per_pass": {k: st[f"prof_{k}_ms"] / max(st["prof_passes"], 1)
                                for k in ("rebuild", "first", "force", "commit")},
        "sampled_passes": st["prof_passes"],
        "step_algorithmic_GBps": value * B_STEP / 1e9,
        "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
        "final_temp_sample": temps,
        "served": srv,
    }
    roof = roofline(t_force, total, "k_force<mDPD, MODE_PHASED> (pool pass, all systems)")
    cpu = None
    if not a.no_cpu:
        med = int(np.argsort(sizes)[len(sizes) // 2])
        box, x, v, _ = specs[med]
        cpu = cpu_baseline(box, x, v, False, "the median droplet of the batch")
    print(json.dumps(base_line(a, value, dt / max(st["passes"], 1) * 1e3, config, roof, cpu)))


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
    ap.add_argument("--workload", default="batched", choices=["batched"] + sorted(SLABS))
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20241202)
    ap.add_argument("--budget-gb", type=float, default=16.0, help="HBM budget of the batched pool")
    ap.add_argument("--plain-sample", type=int, default=16, help="jobs of the batch timed on the plain path")
    ap.add_argument("--served-seconds", type=float, default=10.0, help="arrival window of the served leg (0: skip)")
    ap.add_argument("--served-rate", type=float, default=50.0)
    ap.add_argument("--served-budget-gb", type=float, default=4.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--split", type=int, default=None, help="waves per tile (default: autotuned / 2)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "batched":
        run_batched(a)
    else:
        run_slab(a)


if __name__ == "__main__":
    main()
