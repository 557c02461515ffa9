"""Statistical parity: the engine's dpd/meso run against the CPU LAMMPS
pair_style dpd + fix nve run (the oracle's transcription of
src/pair_dpd.cpp:63-160, src/fix_nve.cpp:64-130, RanMars
src/random_mars.cpp:25-116) on the same state.

The two use different random streams (TEA pair Gaussians vs the serial
Marsaglia generator), so trajectories differ; the ensemble averages must
agree: <T>, <P> and the radial distribution g(r) over r < r_c, sampled at the
same steps.  Tolerances are a few standard errors of 50 correlated samples
of a 2187-particle fluid: |dT| <= 0.02, |dP| <= 1.5 %, |dg(r)| <= 0.05 per bin
of width 0.02 for r >= 0.2.
"""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu

L, RHO, A0, GAMMA, DT = 9.0, 3.0, 25.0, 4.5, 0.04
EQUIL, STEPS, EVERY, NRDF = 150, 400, 5, 50


def rdf_counts(x, L, rc=1.0, nbins=NRDF):
    """pairs i < j with minimum-image distance r < rc, binned over [0, rc)"""
    n = len(x)
    h = np.zeros(nbins)
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(axis=1))
        r = r[r < rc]
        np.add.at(h, (r / rc * nbins).astype(int), 1.0)
    return h


def g_of_r(hist, nsamples, n, L, rc=1.0, nbins=NRDF):
    edges = np.linspace(0, rc, nbins + 1)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    ideal = 0.5 * n * (n / L**3) * shell * nsamples
    return hist / ideal, 0.5 * (edges[1:] + edges[:-1])


def test_dpd_matches_cpu_lammps_dpd(engine, orc):
    n = int(RHO * L**3)
    x, v = uniform_box(n, L, 77)
    cpu = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), a0=A0, gamma=GAMMA, dt=DT, seed=4321,
                             nsteps=STEPS, equil=EQUIL, sample_every=EVERY, nrdf=NRDF)

    s = engine.System()
    s.command("atom_style dpd/atomic/meso")
    s.create_box(1, (0, 0, 0), (L, L, L))
    s.create_atoms(np.ones(n, np.int32), x, v)
    s.commands(f"""
        neighbor 0.3 bin
        neigh_modify every 1 delay 0 check yes
        mass 1 1.0
        pair_style dpd/meso 1.0 98765
        pair_coeff * * {A0} {GAMMA} {np.sqrt(2 * GAMMA)} 1.0 1.0
        fix int all nve/meso
        timestep {DT}
        run_style mvv/meso
    """)
    s.run(EQUIL)
    ts, ps, hist = [], [], np.zeros(NRDF)
    for _ in range((STEPS - EQUIL) // EVERY):
        s.run(EVERY)
        th = s.thermo()
        ts.append(th["temp"])
        ps.append(th["press"])
        hist += rdf_counts(s.gather("x"), L)
    T, P = np.mean(ts), np.mean(ps)
    g_gpu, r = g_of_r(hist, len(ts), n, L)
    g_cpu, _ = g_of_r(cpu["rdf_hist"], cpu["nsamples"], n, L)
    sel = r >= 0.2
    dg = np.max(np.abs(g_gpu[sel] - g_cpu[sel]))
    print(f"GPU T {T:.4f} P {P:.3f} | CPU T {cpu['T']:.4f} P {cpu['P']:.3f} | max |dg| {dg:.4f} "
          f"(samples {len(ts)} / {cpu['nsamples']})")
    assert cpu["nsamples"] == len(ts)
    assert abs(T - cpu["T"]) <= 0.02
    assert abs(P - cpu["P"]) <= 0.015 * abs(cpu["P"])
    assert dg <= 0.05
