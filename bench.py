"""mDPD throughput benchmark on one MI355X.

Workload (default): the plain mDPD liquid slab -- periodic box 40 x 80 x 20,
fluid slab filling the middle half in z (z in [5, 15)) at density 6.05,
N = 193,600 particles at seeded uniform positions, Maxwell-Boltzmann
velocities at kT = 1, A = -40, B = 25, r_c = 1, r_d = 0.75, gamma = 4.5,
dt = 0.01, Lucy density weight, skin 0.3, one system per launch.
`--workload slab-literal` runs the 40 x 40 x 20 box (N = 96,800).

A step is one full mDPD timestep of the system (density pass, force pass
with fused velocity-Verlet, neighbor rebuilds when triggered).  The timed
region runs `--steps` steps after `--warmup` untimed ones; inputs are resident
in HBM throughout.  Prints one JSON line.
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

WORKLOADS = {
    "slab": dict(box=(40.0, 80.0, 20.0), slab=(5.0, 15.0)),
    "slab-literal": dict(box=(40.0, 40.0, 20.0), slab=(5.0, 15.0)),
}


def make_slab(box, slab, seed=20240601):
    lx, ly, lz = box
    n = int(round(RHO * lx * ly * (slab[1] - slab[0])))
    rng = np.random.default_rng(seed)
    x = np.empty((n, 3))
    x[:, 0] = rng.uniform(0, lx, n)
    x[:, 1] = rng.uniform(0, ly, n)
    x[:, 2] = rng.uniform(slab[0], slab[1], n)
    v = rng.normal(0.0, 1.0, (n, 3))
    v -= v.mean(axis=0)
    return x.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64)


DECK = """
    neighbor {skin} bin
    neigh_modify every 1 delay 0 check yes
    mass 1 1.0
    pair_style mdpd/meso 1.0 {rc} {seed}
    pair_coeff * * -40 25 4.5 {rc} {rd}
    fix rho all aboundary/rho/meso cut_rho {rd} weight lucy
    fix int all nve/meso
    timestep 0.01
    run_style mvv/meso
"""


def gpu_run(box, x, v, steps, warmup, seed=4242, split=None):
    import mdpd355
    s = mdpd355.System(0)
    if split:
        s.tune("split", split)
    s.command(f"atom_style mdpd/atomic/meso {RHO}")
    s.create_box(1, (0, 0, 0), box)
    s.create_atoms(np.ones(len(x), np.int32), x, v)
    s.commands(DECK.format(skin=SKIN, rc=RC, rd=RD, seed=seed))
    s.setup()
    s.run(warmup)
    nb0 = s.nbuild
    s.profile(True)
    t0 = time.perf_counter()
    s.run(steps)                      # returns after the stream has drained
    dt = time.perf_counter() - t0
    prof = {ph: s.profile_get(ph) for ph in ("force", "density", "rebuild", "initial")}
    nbuild = s.nbuild - nb0
    thermo = s.thermo()
    s.close()
    return dt, prof, nbuild, thermo


def cpu_baseline(box, x, v, target_s=12.0, seed=4242):
    """Oracle ("port"): FP64 OpenMP restatement of the same step, bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import orc
    threads = min(16, os.cpu_count() or 1)
    p = orc.params((0, 0, 0), box, 0.01, cut=RC, cut_r=RD, A=-40.0, B=25.0, gamma=4.5,
                   sigma=np.sqrt(9.0), expw=1.0)
    n = len(x)
    args = (np.arange(1, n + 1), np.zeros(n, int), np.ones(n, int), p)
    kw = dict(style=0, seed=seed, skin=SKIN, cut_rho=RD, lucy=True, nthreads=threads)
    t0 = time.perf_counter()
    orc.run(x, v, *args, 0, **kw)                         # setup only
    t_setup = time.perf_counter() - t0
    steps = 2
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    t_two = time.perf_counter() - t0 - t_setup
    per = max(t_two / steps, 1e-6)
    steps = int(max(2, min(200, target_s / per)))
    t0 = time.perf_counter()
    orc.run(x, v, *args, steps, **kw)
    t = time.perf_counter() - t0 - t_setup
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 mDPD step (OpenMP), same slab N={n}, {steps} steps after setup, "
                      f"{t:.1f} s wall (setup {t_setup:.1f} s excluded)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--workload", default="slab", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--split", type=int, default=None, help="waves per tile in the per-step passes")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[a.workload]
    box = w["box"]
    x, v = make_slab(box, w["slab"])
    n = len(x)
    dt, prof, nbuild, thermo = gpu_run(box, x, v, a.steps, a.warmup, split=a.split)
    value = n * a.steps / dt
    f_ms, f_n = prof["force"]
    t_force = f_ms / max(f_n, 1) * 1e-3
    achieved = B_FORCE * n / t_force / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"plain mDPD liquid slab ({a.workload}): box {box[0]:g}x{box[1]:g}x{box[2]:g}, "
                        f"slab z in [5,15) at rho 6.05, N={n}, A=-40 B=25 rc=1 rd=0.75 gamma=4.5 kT=1 "
                        f"dt=0.01, Lucy weight, skin 0.3, one system per launch",
            "mode": "plain",
            "particles": n,
            "rebuilds_in_timed_region": nbuild,
            "phase_ms_per_step": {k: ms / max(a.steps, 1) for k, (ms, _) in prof.items()},
            "step_algorithmic_GBps": value * B_STEP / 1e9,
            "step_hbm_frac": value * B_STEP / 1e9 / HBM_PEAK_GBS,
            "final_temp": thermo["temp"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_force<mDPD, fused velocity-Verlet>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_particle": B_FORCE,
            "avg_launch_us": t_force * 1e6,
        },
        "cpu_baseline": None if a.no_cpu else cpu_baseline(box, x, v),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
