"""The workload shapes of the benchmark configurations, small enough for the
FP64 oracle, on the GPU against the oracle:

* liquid slab with vacuum (Lucy and quadratic density weights): setup
  densities and forces, and a 5-step run;
* quasi-2D pore flow: frozen disc walls outside the integration group
  (fix_nve_meso.cu:83-92), body force on the fluid (fix_addforce_meso.cu:72-108),
  solid-fluid B = 0 (examples/v2.5/quasi_2d_pores/in.quasi_2d_pores:55-90):
  one step at the parity bound and a 10-step run;
* droplets of log-normal sizes as jobs of one pool (batched mode), each
  against its own oracle run;
* a served mix: a droplet job running when a slab job arrives, both against
  their oracle runs.

Bounds: setup and one-step quantities at the parity bounds of
test_gpu_parity (1e-5); after several steps the FP32 and FP64 trajectories
separate slowly (chaotic dynamics from identical random streams), so
positions are held to 1e-4 absolute and velocities and forces to 1e-4 per
particle (relative to max(|q_i|, rms q)).
"""
import numpy as np
import pytest

from test_gpu_parity import RD, SEED, assert_rho_close, assert_vec_close, mdpd_system

pytestmark = pytest.mark.gpu

GB = float(1 << 30)
RHO = 6.05


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def maxwell(rng, n):
    v = rng.normal(0.0, 1.0, (n, 3))
    return f32(v - v.mean(axis=0))


def slab(lx, ly, lz, z0, z1, seed):
    rng = np.random.default_rng(seed)
    n = int(round(RHO * lx * ly * (z1 - z0)))
    x = np.c_[rng.uniform(0, lx, n), rng.uniform(0, ly, n), rng.uniform(z0, z1, n)]
    return np.array([lx, ly, lz]), f32(x), maxwell(rng, n)


def droplet(n, seed):
    rng = np.random.default_rng(seed)
    R = (3.0 * n / (4.0 * np.pi * RHO)) ** (1.0 / 3.0)
    L = 2.5 * R
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    x = L / 2 + d * (R * rng.uniform(0, 1, n) ** (1.0 / 3.0))[:, None]
    return np.array([L, L, L]), f32(x), maxwell(rng, n)


def params(orc, box, **kw):
    base = dict(cut=1.0, cut_r=RD, A=-40.0, B=25.0, gamma=4.5, sigma=3.0, expw=1.0)
    base.update(kw)
    return orc.params((0, 0, 0), box, 0.01, **base)


def oracle_run(orc, box, x, v, nsteps, seed, lucy=False, **kw):
    n = len(x)
    p = params(orc, box)
    return orc.run(x, v, np.arange(1, n + 1), np.zeros(n, int), np.ones(n, int), p, nsteps, style=0, seed=seed,
                   lucy=lucy, **kw)


def assert_track(box, got_x, got_v, got_f, ref, what, tol=1e-4):
    xo, vo, fo = ref[0], ref[1], ref[2]
    d = got_x - xo
    d -= box * np.round(d / box)
    print(f"{what}: max |dx| {np.abs(d).max():.2e}")
    assert np.abs(d).max() < tol, what
    assert_vec_close(got_v, vo, tol=tol, what=f"{what} v")
    assert_vec_close(got_f, fo, tol=tol, what=f"{what} f")


def slab_system(engine, box, x, v, lucy):
    return mdpd_system(engine, x, v, box, lucy=lucy)


# ---------------------------------------------------------------- slab

@pytest.mark.parametrize("lucy", [False, True])
def test_slab_with_vacuum_matches_oracle(engine, orc, lucy):
    box, x, v = slab(10.0, 10.0, 12.0, 3.0, 9.0, 81)
    n = len(x)
    s = slab_system(engine, box, x, v, lucy)
    s.setup()
    xc, vc = s.gather("x"), s.gather("v")
    p = params(orc, box)
    t0 = np.zeros(n, int)
    rho = orc.density(xc, t0, np.ones(n, int), p, RD, lucy=lucy)
    assert_rho_close(orc, s.gather("rho"), rho, xc, box, lucy=lucy, what=f"slab density lucy={lucy}")
    sig = orc.sig_mdpd(np.arange(1, n + 1), orc.seed_now(SEED, 0))
    f, e, vir = orc.force_mdpd(xc, vc, rho, sig, t0, p)
    assert_vec_close(s.gather("f"), f, what=f"slab force lucy={lucy}")
    assert_vec_close(s.gather("virial"), vir, what=f"slab virial lucy={lucy}")
    s.run(5)
    ref = oracle_run(orc, box, xc, vc, 5, SEED, lucy=lucy)
    assert_track(box, s.gather("x"), s.gather("v"), s.gather("f"), ref, f"slab 5 steps lucy={lucy}")


# ---------------------------------------------------------------- pore

def pore(seed):
    """12 x 12 x 4 box, two frozen discs (radius 2.5) of wall particles
    (type 2) in a fluid (type 1), both at density 6."""
    rng = np.random.default_rng(seed)
    box = np.array([12.0, 12.0, 4.0])
    n = int(6.0 * box.prod())
    x = rng.uniform(0, 1, (n, 3)) * box
    wall = np.zeros(n, bool)
    for c in ((3.0, 3.0), (8.5, 8.0)):
        wall |= ((x[:, :2] - c) ** 2).sum(axis=1) < 2.5**2
    types = np.where(wall, 2, 1).astype(np.int32)
    v = rng.normal(0, 1, (n, 3))
    v[wall] = 0.0
    v[~wall] -= v[~wall].mean(axis=0)
    return box, f32(x), f32(v), types


def pore_system(engine, box, x, v, types):
    s = mdpd_system(engine, x, v, box, types=types, ntypes=2, coeff="1 1 -40 25 4.5 1.0 0.75",
                    extra="pair_coeff 1 2 -40 0 4.5 1.0 0.75\npair_coeff 2 2 -40 25 4.5 1.0 0.75")
    s.commands("""
        group fluid type 1
        group wall type 2
        unfix int
        fix int fluid nve/meso
        fix push fluid addforce/meso 0.02 0.0 0.0
    """)
    return s


def test_frozen_walls_and_body_force(engine, orc):
    box, x, v, types = pore(83)
    n = len(x)
    wall = types == 2
    s = pore_system(engine, box, x, v, types)
    s.setup()
    x0, v0, mask = s.gather("x"), s.gather("v"), s.gather("mask")
    fbit = s.group_bit("fluid")
    assert np.array_equal((mask & fbit) != 0, ~wall)
    B = [[25.0, 0.0], [0.0, 25.0]]
    p = orc.params((0, 0, 0), box, 0.01, ntypes=2, mass=(1.0, 1.0), cut=1.0, cut_r=RD, A=-40.0, B=B,
                   gamma=4.5, sigma=3.0, expw=1.0)
    kw = dict(style=0, seed=SEED, nve_groupbit=fbit, addforce_groupbit=fbit, addforce=(0.02, 0.0, 0.0))
    # setup forces: the body force is in f of the fluid particles only
    ref0 = orc.run(x0, v0, np.arange(1, n + 1), types - 1, mask, p, 0, **kw)
    assert_vec_close(s.gather("f"), ref0[2], what="pore setup force (with body force)")
    s.run(1)
    ref1 = orc.run(x0, v0, np.arange(1, n + 1), types - 1, mask, p, 1, **kw)
    assert np.array_equal(s.gather("x")[wall], x0[wall]) and np.all(s.gather("v")[wall] == 0)
    d = s.gather("x") - ref1[0]
    d -= box * np.round(d / box)
    assert np.abs(d).max() < 5e-6
    assert_vec_close(s.gather("v"), ref1[1], what="pore v after one step")
    assert_vec_close(s.gather("f"), ref1[2], what="pore f after one step")
    s.run(9)
    ref10 = orc.run(x0, v0, np.arange(1, n + 1), types - 1, mask, p, 10, **kw)
    assert_track(box, s.gather("x"), s.gather("v"), s.gather("f"), ref10, "pore 10 steps")
    assert np.array_equal(s.gather("x")[wall], x0[wall])


# ---------------------------------------------------------------- batched

def pool_job(engine, box, x, v, seed, lucy=False):
    return mdpd_system(engine, x, v, box, seed=seed, lucy=lucy)


def test_droplet_jobs_in_a_pool_match_their_oracle_runs(engine, orc):
    rng = np.random.default_rng(85)
    sizes = np.clip(np.round(np.exp(rng.normal(np.log(3000.0), 0.5, 5))), 1000, 8000).astype(int)
    specs = [droplet(int(m), 900 + k) for k, m in enumerate(sizes)]
    nsteps = [5, 4, 5, 3, 5]
    p = engine.Pool(0.5 * GB, 8)
    ids = [p.submit(pool_job(engine, box, x, v, 2000 + k), k2)
           for k, ((box, x, v), k2) in enumerate(zip(specs, nsteps))]
    p.run()
    for k, (j, (box, x, v)) in enumerate(zip(ids, specs)):
        ref = oracle_run(orc, box, x, v, nsteps[k], 2000 + k)
        assert p.info(j)["timestep"] == nsteps[k]
        assert_track(box, p.gather(j, "x"), p.gather(j, "v"), p.gather(j, "f"), ref, f"droplet job {k} (N={len(x)})")


# ---------------------------------------------------------------- served

def test_served_slab_and_droplet_match_their_oracle_runs(engine, orc):
    # a droplet job is on the device when a slab job arrives; each result
    # equals its own oracle run
    dbox, dx, dv = droplet(2500, 91)
    sbox, sx, sv = slab(9.0, 9.0, 14.0, 4.0, 10.0, 92)
    p = engine.Pool(0.5 * GB, 8)
    p.tune("poll", 2)
    jd = p.submit(pool_job(engine, dbox, dx, dv, 3001), 5)
    p.advance(max_passes=2)
    assert p.state(jd) == engine.JOB_RUNNING
    js = p.submit(pool_job(engine, sbox, sx, sv, 3002, lucy=True), 4)
    p.run()
    ref_d = oracle_run(orc, dbox, dx, dv, 5, 3001)
    ref_s = oracle_run(orc, sbox, sx, sv, 4, 3002, lucy=True)
    assert_track(dbox, p.gather(jd, "x"), p.gather(jd, "v"), p.gather(jd, "f"), ref_d, "served droplet")
    assert_track(sbox, p.gather(js, "x"), p.gather(js, "v"), p.gather(js, "f"), ref_s, "served slab")
    info_d, info_s = p.info(jd), p.info(js)
    assert info_s["t_admit"] < info_d["t_done"]         # the slab ran while the droplet did
