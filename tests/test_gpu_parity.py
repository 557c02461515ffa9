"""Parity of the HIP path (through the C ABI) against the oracle.

Bit-exact: TEA words, signatures, sort keys, sorted order, neighbor sets.
Floating point (FP32 engine vs FP64 oracle on the same FP32 state):
  densities     |rho_i - rho_ref_i| <= 1e-5 |rho_ref_i| + m k_i max|w'| dr, with
                k_i the neighbors within r_d and dr = 8 ulp(1) = 4.8e-7 the FP32
                error bound of a pair distance r < 1 (correctly rounded
                displacement, r^2, sqrt): a lone neighbor near r_d has a tiny
                weight whose relative error FP32 cannot bound
  forces        ||F - F_ref|| / ||F_ref|| <= 1e-5 over the system, and per
                particle |dF_i| <= 1e-5 max(|F_ref_i|, rms|F_ref|)
  energy/virial same metric as forces
"""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu

SEED = 123456
RD = 0.75
IDX = (1 << 27) - 1


def assert_vec_close(got, ref, tol=1e-5, what=""):
    got = np.asarray(got, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    glob = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    err = np.linalg.norm(got - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    rms = np.sqrt(np.mean(mag**2))
    worst = np.max(err / np.maximum(mag, rms))
    print(f"{what}: global rel {glob:.3e}, worst per-particle {worst:.3e}")
    assert glob <= tol, (what, glob)
    assert worst <= tol, (what, worst)


DR_FP32 = 8.0 * 2.0**-24          # FP32 error bound of a pair distance r < 1


def rho_floor(orc, x, L, rd=RD, lucy=False, mass=1.0):
    """Absolute floor of the density comparison per particle: each neighbor
    within r_d adds m max|dw/dr| dr, max|dw/dr| = 2 f r_d for the quadratic
    weight f (r_d - r)^2 and 16/9 f_L r_d^3 for Lucy's f_L (r_d + 3r)(r_d - r)^3
    (the maximum of 12 f_L r (r_d - r)^2, at r = r_d / 3)."""
    counts, _ = orc.neighbors(x, (0, 0, 0), np.broadcast_to(L, (3,)), rd * 1.001)
    if lucy:
        slope = 16.0 / 9.0 * 105.0 / (16.0 * np.pi * rd**7) * rd**3
    else:
        slope = 2.0 * 15.0 / (2.0 * np.pi * rd**5) * rd
    return mass * counts * slope * DR_FP32


def assert_rho_close(orc, got, ref, x, L, lucy=False, what="density"):
    floor = rho_floor(orc, x, L, lucy=lucy)
    err = np.abs(np.asarray(got, np.float64) - ref)
    rel = err / np.maximum(np.abs(ref), 1e-300)
    print(f"{what}: worst rel {rel[ref > 0].max() if np.any(ref > 0) else 0:.3e}, "
          f"worst vs bound {np.max(err / (1e-5 * np.abs(ref) + floor + 1e-30)):.3e}")
    assert np.all(err <= 1e-5 * np.abs(ref) + floor), what


def mdpd_system(engine, x, v, L, *, types=None, ntypes=1, lucy=False, rho_est=6.05, skin=0.3,
                seed=SEED, extra="", coeff="* * -40 25 4.5 1.0 0.75"):
    n = len(x)
    s = engine.System()
    s.commands(f"units lj\natom_style mdpd/atomic/meso {rho_est}")
    s.create_box(ntypes, (0, 0, 0), np.broadcast_to(L, (3,)))
    s.create_atoms(np.ones(n, np.int32) if types is None else types, x, v)
    s.commands(f"""
        neighbor {skin} bin
        neigh_modify every 1 delay 0 check yes
        mass * 1.0
        pair_style mdpd/meso 1.0 1.0 {seed}
        pair_coeff {coeff}
        fix rho all aboundary/rho/meso cut_rho {RD} {'weight lucy' if lucy else ''}
        fix int all nve/meso
        timestep 0.01
        run_style mvv/meso
        {extra}
    """)
    return s


def dpd_system(engine, x, v, L, seed=SEED, expw=1.0):
    n = len(x)
    s = engine.System()
    s.command("atom_style dpd/atomic/meso")
    s.create_box(1, (0, 0, 0), (L, L, L))
    s.create_atoms(np.ones(n, np.int32), x, v)
    s.commands(f"""
        neighbor 0.3 bin
        neigh_modify every 1 delay 0 check yes
        mass 1 1.0
        pair_style dpd/meso 1.0 {seed}
        pair_coeff * * 25.0 4.5 3.0 {expw} 1.0
        fix int all nve/meso
        timestep 0.04
        run_style mvv/meso
    """)
    return s


def oracle_params(orc, L, dt=0.01, **kw):
    base = dict(cut=1.0, cut_r=RD, A=-40.0, B=25.0, gamma=4.5, sigma=np.sqrt(2 * 4.5), expw=1.0)
    base.update(kw)
    return orc.params((0, 0, 0), np.broadcast_to(L, (3,)), dt, **base)


# ---------------------------------------------------------------- stateless

@pytest.mark.parametrize("rounds", [4, 16, 32, 64])
def test_tea_premix_bit_exact(engine, orc, rounds):
    rng = np.random.default_rng(rounds)
    a = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(engine.tea_premix(a, b, rounds), orc.premix(a, b, rounds))


def test_pair_gaussian_fp32(engine, orc):
    rng = np.random.default_rng(2)
    a = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    g = engine.tea_gaussian(a, b).astype(np.float64)
    ref = orc.gaussian(a, b)
    err = np.abs(g - ref)
    print("pair Gaussian max abs err", err.max())
    assert err.max() < 4e-6
    assert np.array_equal(g, engine.tea_gaussian(b, a).astype(np.float64))


# ---------------------------------------------------------------- setup tables

@pytest.fixture(scope="module")
def setup_case(engine, orc):
    L = 8.0
    x, v = uniform_box(2400, L, 101)
    x[:7] += np.array([L, -L, 0.0])          # some particles start outside the box
    s = mdpd_system(engine, x, v, L)
    s.setup()
    return s, L, x, v


def test_signatures_bit_exact(setup_case, orc):
    s, L, x, v = setup_case
    want = orc.sig_mdpd(np.arange(1, len(x) + 1), orc.seed_now(SEED, 0))
    assert np.array_equal(s.debug_signatures(), want)


def test_keys_and_sorted_order_bit_exact(setup_case, orc):
    s, L, x, v = setup_case
    origin, keys, xw = s.debug_sorted()
    xc = orc.wrap(x.astype(np.float32), (0, 0, 0), (L, L, L))     # creation order, wrapped
    assert np.array_equal(xw, xc[origin])                          # wrap is bit-exact
    ref_keys = orc.sort_keys(xc, (0, 0, 0), (L, L, L), 1.3)
    assert np.array_equal(keys, ref_keys[origin])
    assert np.array_equal(origin, orc.stable_sort(ref_keys))


def _neighbor_sets(origin, counts, rows):
    out = [None] * len(origin)
    for k in range(len(origin)):
        e = rows[k, :counts[k]].astype(np.int64)
        j = origin[e & IDX]
        out[origin[k]] = set(zip(j.tolist(), (e >> 27).tolist()))
    return out


def test_neighbor_sets_bit_exact(setup_case, orc):
    s, L, x, v = setup_case
    origin, _, xw = s.debug_sorted()
    counts, rows = s.debug_neighbors()
    assert counts.max() <= rows.shape[1]
    got = _neighbor_sets(origin, counts, rows)
    xc = np.empty_like(xw)
    xc[origin] = xw
    rc, rrows = orc.neighbors(xc, (0, 0, 0), (L, L, L), 1.3)
    for i in range(len(x)):
        ref = {(int(e) & IDX, int(e) >> 27) for e in rrows[i]}
        assert got[i] == ref, i


def _assert_stencil_order(xw, counts, rows, L, cut):
    """Each row lists the 27 stencil cells in order k = (ox+1) + 3(oy+1) +
    9(oz+1) and, inside a cell, sorted indices ascending: the order of a row
    built over whole cells, whatever sub-cells the build skipped."""
    nb = int(L * (1.0 / cut))
    bininv = 1.0 / (L / nb)
    cell = np.clip((xw.astype(np.float64) * bininv + 1.0).astype(np.int64), 1, nb) - 1
    n = len(xw)
    for i in range(n):
        e = rows[i, :counts[i]].astype(np.int64) & 0xFFFFFFFF
        j, code = e & IDX, e >> 27
        shift = np.stack([code % 3 - 1, code // 3 % 3 - 1, code // 9 - 1], axis=1)
        off = cell[j] + shift * nb - cell[i]
        assert np.all(np.abs(off) <= 1), i
        k = (off[:, 0] + 1) + 3 * (off[:, 1] + 1) + 9 * (off[:, 2] + 1)
        key = k * n + j
        assert np.all(np.diff(key) > 0), i


@pytest.mark.parametrize("L", [8.0, 6.5])
def test_neighbor_rows_in_stencil_order(engine, orc, L):
    # L = 8: bins of 4/3 > rc + skin, the build skips far sub-cells;
    # L = 6.5: bins of exactly rc + skin, every sub-cell is visited
    x, v = uniform_box(int(6.05 * L**3), L, 7)
    bs = L / int(L / 1.3)
    x[:40, 1] = (np.arange(40) % (2 * int(L / 1.3))) * bs / 2     # on half-cell boundaries
    x[40:80, 2] = (np.arange(40) % (2 * int(L / 1.3))) * bs / 2
    s = mdpd_system(engine, x, v, L)
    s.setup()
    origin, _, xw = s.debug_sorted()
    counts, rows = s.debug_neighbors()
    got = _neighbor_sets(origin, counts, rows)
    xc = np.empty_like(xw)
    xc[origin] = xw
    _, rrows = orc.neighbors(xc, (0, 0, 0), (L, L, L), 1.3)
    for i in range(len(x)):
        assert got[i] == {(int(e) & IDX, int(e) >> 27) for e in rrows[i]}, i
    _assert_stencil_order(xw, counts, rows, L, 1.3)


def test_density_parity(setup_case, orc):
    s, L, x, v = setup_case
    xc = s.gather("x")
    p = oracle_params(orc, L)
    ref = orc.density(xc, np.zeros(len(x), int), np.ones(len(x), int), p, RD)
    assert_rho_close(orc, s.gather("rho"), ref, xc, L)


def test_force_energy_virial_parity(setup_case, orc):
    s, L, x, v = setup_case
    xc, vc = s.gather("x"), s.gather("v")
    p = oracle_params(orc, L)
    t0 = np.zeros(len(x), int)
    rho = orc.density(xc, t0, np.ones(len(x), int), p, RD)
    sig = orc.sig_mdpd(np.arange(1, len(x) + 1), orc.seed_now(SEED, 0))
    f, e, vir = orc.force_mdpd(xc, vc, rho, sig, t0, p)
    assert_vec_close(s.gather("f"), f, what="mDPD force")
    assert_vec_close(s.gather("virial"), vir, what="mDPD virial")
    ge = s.gather("e_pair")
    assert np.abs(ge - e).max() <= 1e-5 * np.sqrt(np.mean(e**2))
    fsum = np.abs(s.gather("f").sum(axis=0)).max()
    assert fsum < 1e-5 * np.abs(f).sum() / len(f) * np.sqrt(len(f))


def test_lucy_density(engine, orc):
    L = 7.0
    x, v = uniform_box(1500, L, 7)
    s = mdpd_system(engine, x, v, L, lucy=True)
    s.setup()
    p = oracle_params(orc, L)
    xc = s.gather("x")
    ref = orc.density(xc, np.zeros(len(x), int), np.ones(len(x), int), p, RD, lucy=True)
    assert_rho_close(orc, s.gather("rho"), ref, xc, L, lucy=True, what="Lucy density")


def test_two_types_forces(engine, orc):
    L = 7.0
    x, v = uniform_box(1800, L, 17)
    types = np.where(np.arange(len(x)) % 3 == 0, 2, 1).astype(np.int32)
    s = mdpd_system(engine, x, v, L, types=types, ntypes=2, coeff="1 1 -40 25 4.5 1.0 0.75",
                    extra="pair_coeff 1 2 -30 25 4.5 1.0 0.75\npair_coeff 2 2 -20 20 2.0 0.9 0.7")
    s.setup()
    A = [[-40, -30], [-30, -20]]
    B = [[25, 25], [25, 20]]
    g = np.array([[4.5, 4.5], [4.5, 2.0]])
    p = orc.params((0, 0, 0), (L, L, L), 0.01, ntypes=2, mass=(1.0, 1.0), cut=[[1, 1], [1, 0.9]],
                   cut_r=[[RD, RD], [RD, 0.7]], A=A, B=B, gamma=g, sigma=np.sqrt(2 * g), expw=1.0)
    xc, vc = s.gather("x"), s.gather("v")
    t0 = types - 1
    rho = orc.density(xc, t0, np.ones(len(x), int), p, RD)
    sig = orc.sig_mdpd(np.arange(1, len(x) + 1), orc.seed_now(SEED, 0))
    f, _, _ = orc.force_mdpd(xc, vc, rho, sig, t0, p)
    assert_vec_close(s.gather("f"), f, what="2-type force")


# ---------------------------------------------------------------- DPD

@pytest.mark.parametrize("expw", [1.0, 0.5])
def test_dpd_signatures_and_forces(engine, orc, expw):
    L = 8.0
    x, v = uniform_box(1536, L, 23)
    s = dpd_system(engine, x, v, L, expw=expw)
    s.setup()
    xc, vc = s.gather("x"), s.gather("v")
    st = orc.seed_now(SEED, 0)
    sig = orc.sig_dpd(np.arange(1, len(x) + 1), st, vc)
    assert np.array_equal(s.debug_signatures(), sig)
    p = orc.params((0, 0, 0), (L, L, L), 0.04, cut=1.0, A=25.0, gamma=4.5, sigma=3.0, expw=expw)
    f, e, vir = orc.force_dpd(xc, vc, sig, np.zeros(len(x), int), p)
    assert_vec_close(s.gather("f"), f, what=f"DPD force expw={expw}")
    assert_vec_close(s.gather("virial"), vir, what="DPD virial")


# ---------------------------------------------------------------- stepping

def test_one_step_matches_oracle(engine, orc):
    L = 8.0
    x, v = uniform_box(2048, L, 31)
    s = mdpd_system(engine, x, v, L)
    s.setup()
    x0, v0 = s.gather("x"), s.gather("v")
    s.run(1)
    p = oracle_params(orc, L)
    xo, vo, fo, rho, nb = orc.run(x0, v0, np.arange(1, len(x) + 1), np.zeros(len(x), int),
                                  np.ones(len(x), int), p, 1, style=0, seed=SEED, ntimestep=0)
    assert s.timestep == 1
    dx = s.gather("x") - xo
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 5e-6
    assert_vec_close(s.gather("v"), vo, what="v after one step")
    assert_vec_close(s.gather("f"), fo, what="f after one step")


def test_rebuilds_keep_tables_exact(engine, orc):
    L = 7.5
    x, v = uniform_box(1900, L, 41)
    s = mdpd_system(engine, x, v, L)
    s.run(25)
    assert s.nbuild >= 3
    origin, keys, xw = s.debug_sorted()
    ref_keys = orc.sort_keys(xw, (0, 0, 0), (L, L, L), 1.3)
    assert np.array_equal(keys, ref_keys)
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0) if len(keys) > 1 else True
    counts, rows = s.debug_neighbors()
    got = _neighbor_sets(origin, counts, rows)
    xc = np.empty_like(xw)
    xc[origin] = xw
    _, rrows = orc.neighbors(xc, (0, 0, 0), (L, L, L), 1.3)
    for i in range(len(x)):
        assert got[i] == {(int(e) & IDX, int(e) >> 27) for e in rrows[i]}, i
    # forces of the final state against the oracle
    xc2, vc2 = s.gather("x"), s.gather("v")
    p = oracle_params(orc, L)
    t0 = np.zeros(len(x), int)
    rho = orc.density(xc2, t0, np.ones(len(x), int), p, RD)
    assert_rho_close(orc, s.gather("rho"), rho, xc2, L, what="density after 25 steps")


def test_short_trajectory_tracks_oracle(engine, orc):
    # FP32 vs FP64 trajectories separate slowly at first: after 10 steps the
    # positions still agree to ~1e-4
    L = 8.0
    x, v = uniform_box(2048, L, 37)
    s = mdpd_system(engine, x, v, L)
    s.setup()
    x0, v0 = s.gather("x"), s.gather("v")
    s.run(10)
    p = oracle_params(orc, L)
    xo, vo, *_ = orc.run(x0, v0, np.arange(1, len(x) + 1), np.zeros(len(x), int), np.ones(len(x), int),
                         p, 10, style=0, seed=SEED, ntimestep=0)
    dx = s.gather("x") - xo
    dx -= L * np.round(dx / L)
    print("max |dx| after 10 steps", np.abs(dx).max())
    assert np.abs(dx).max() < 1e-3


def test_momentum_conserved_over_run(engine):
    L = 8.0
    x, v = uniform_box(2400, L, 47)
    s = mdpd_system(engine, x, v, L)
    s.run(50)
    pm = s.gather("v").sum(axis=0)
    print("net momentum after 50 steps", pm)
    assert np.abs(pm).max() < 2e-3


def test_thermo_matches_oracle_trajectory(engine, orc):
    # compute temp/meso and pressure/meso of the device state after 20 steps
    # against the FP64 oracle run from the same start (random start: T is far
    # from kT while the overlaps relax, which is the physics, not an error)
    L = 8.0
    x, v = uniform_box(2400, L, 53)
    s = mdpd_system(engine, x, v, L)
    s.setup()
    x0, v0 = s.gather("x"), s.gather("v")
    s.run(20)
    th = s.thermo()
    p = oracle_params(orc, L)
    n = len(x)
    t0, ones = np.zeros(n, int), np.ones(n, int)
    xo, vo, fo, rho, _ = orc.run(x0, v0, np.arange(1, n + 1), t0, ones, p, 20, style=0, seed=SEED)
    sig = orc.sig_mdpd(np.arange(1, n + 1), orc.seed_now(SEED, 20))
    # the step-20 force pass saw the half-step velocity (before the final kick)
    vhalf = vo - 0.5 * 0.01 * fo
    _, e, vir = orc.force_mdpd(xo, vhalf, rho, sig, t0, p)
    T = orc.lib().orc_temperature(n, vo.ctypes.data_as(orc.ctypes.c_void_p), t0.astype(np.int32).ctypes.data_as(orc.ctypes.c_void_p), orc.ctypes.byref(p))
    P = orc.lib().orc_pressure(n, vo.ctypes.data_as(orc.ctypes.c_void_p), vir.ctypes.data_as(orc.ctypes.c_void_p),
                               t0.astype(np.int32).ctypes.data_as(orc.ctypes.c_void_p), orc.ctypes.byref(p))
    print("thermo", th, "oracle T", T, "P", P)
    assert th["temp"] == pytest.approx(T, rel=1e-4)
    assert th["press"] == pytest.approx(P, rel=1e-4)
    assert th["pe"] == pytest.approx(e.sum(), rel=1e-4)


def test_run_continuation_is_bitwise(engine):
    # run 30 + run 30 == run 60: the end-of-run half-kick followed by the next
    # run's first half-kick performs the same FP32 operations as the fused step
    L = 8.0
    x, v = uniform_box(2400, L, 59)
    a = mdpd_system(engine, x, v, L)
    a.run(30)
    a.run(30)
    b = mdpd_system(engine, x, v, L)
    b.run(60)
    assert a.timestep == b.timestep == 60
    assert np.array_equal(a.gather("x"), b.gather("x"))
    assert np.array_equal(a.gather("v"), b.gather("v"))
    assert np.array_equal(a.gather("f"), b.gather("f"))


def test_run_continuation_keeps_the_rebuild_schedule(engine):
    # neigh_modify every 4 delay 10: the steps since the last build carry
    # over from one run to the next (each step's commit counts them on the
    # device), so run 23 + run 37 rebuilds exactly when run 60 does and ends
    # bit for bit where run 60 ends
    L = 8.0
    x, v = uniform_box(2400, L, 67)
    deck = "neigh_modify every 4 delay 10 check yes"
    a = mdpd_system(engine, x, v, L, extra=deck)
    a.run(23)
    a.run(37)
    b = mdpd_system(engine, x, v, L, extra=deck)
    b.run(60)
    assert a.nbuild == b.nbuild and b.nbuild >= 3
    for k in ("x", "v", "f"):
        assert np.array_equal(a.gather(k), b.gather(k)), k


def test_replayed_steps_equal_launched_steps(engine):
    # the fused steps of a run replayed from a captured hipGraph (batches of
    # steps between readbacks of the device step counter) and the same
    # kernels launched one by one give identical results
    L = 8.0
    x, v = uniform_box(2400, L, 69)
    a = mdpd_system(engine, x, v, L)
    a.run(150)
    b = mdpd_system(engine, x, v, L)
    b.tune("graph", 0)
    b.run(150)
    assert a.nbuild == b.nbuild
    for k in ("x", "v", "f", "rho", "virial"):
        assert np.array_equal(a.gather(k), b.gather(k)), k


def test_rows_widened_mid_run_equal_wide_rows(engine):
    # two streams (density 3) run into each other while the mDPD fluid
    # condenses: local densities double and rows sized for density 3
    # overflow in a rebuild inside a batch of replayed steps.  The step that
    # found them is skipped on the device, the rows are widened and re-listed
    # from the positions of that build, and the step runs again: the run
    # ends bit for bit where a run with wide rows from the start ends
    L, n = 8.0, 1536
    rng = np.random.default_rng(71)
    x = rng.uniform(0, L, (n, 3))
    v = rng.normal(0, 0.3, (n, 3))
    v[:, 0] += np.where(x[:, 0] < L / 2, 4.0, -4.0)
    x, v = x.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64)
    a = mdpd_system(engine, x, v, L, rho_est=3.0)
    a.setup()
    w0 = a.debug_neighbors()[1].shape[1]
    a.run(80)
    w1 = a.debug_neighbors()[1].shape[1]
    b = mdpd_system(engine, x, v, L, rho_est=16.0)
    b.run(80)
    print(f"row width {w0} -> {w1}; builds {a.nbuild}")
    assert w1 > w0
    assert a.nbuild == b.nbuild
    for k in ("x", "v", "f", "rho"):
        assert np.array_equal(a.gather(k), b.gather(k)), k


def test_observable_fixes_count_what_the_host_counts(engine, tmp_path):
    # rdf/fast/meso and density/meso sample the device state; one sample of
    # each is recounted on the host from the gathered positions (bins may
    # differ by a count where a distance falls on a bin edge in FP32)
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 21)
    s = mdpd_system(engine, x, v, L, extra=f"""
        fix g all rdf/fast/meso output {tmp_path}/rdf.out nbin 20 every 5
        fix ge all rdf/fast/meso output {tmp_path}/rdf_pi.out nbin 20 every 5 pi exact
        fix p all density/meso output {tmp_path}/prof along z nbin 6 every 1000 window 10
    """)
    s.run(5)                       # rdf samples step 5; density samples steps 1..5 (window [-10, 10) of m = 0)
    xr = s.gather("x")
    _, g = s.fix_result("g")
    # the same counts over shells computed with pi
    np.testing.assert_allclose(s.fix_result("ge")[1], g * 3.1415 / np.pi, rtol=1e-12)
    with pytest.raises(RuntimeError, match="pi must be"):
        s.command(f"fix gx all rdf/fast/meso output {tmp_path}/x.out nbin 20 pi 3.14")
    d = xr[:, None, :] - xr[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))[~np.eye(n, dtype=bool)]
    h = np.histogram(r[r < 1.0], bins=20, range=(0, 1.0))[0]
    shell = 4.0 / 3.0 * 3.1415 * ((np.arange(1, 21) * 0.05) ** 3 - (np.arange(20) * 0.05) ** 3)
    g_host = h / n / shell / (n / L**3)
    tol = 2.0 / n / shell / (n / L**3)
    assert np.all(np.abs(g - g_host) <= tol + 1e-12), np.abs(g - g_host) / tol
    zc, prof = s.fix_result("p")
    assert np.allclose(zc, (np.arange(6) + 0.5) * 1.0)
    # each frame counts every particle inside the box once (as the reference,
    # a particle past a face between re-sorts is not counted)
    assert 0.95 * n < prof.sum() <= n
    s.command("unfix g")
    out = np.loadtxt(tmp_path / "rdf.out")
    assert out.shape == (20, 2) and np.allclose(out[:, 1], g, rtol=1e-12)


def _read_dump(path):
    frames, lines, k = {}, open(path).read().split("\n"), 0
    while k < len(lines) and lines[k].startswith("ITEM: TIMESTEP"):
        t = int(lines[k + 1]); n = int(lines[k + 3]); cols = lines[k + 8].split()[2:]
        rows = np.array([[float(c) for c in l.split()] for l in lines[k + 9:k + 9 + n]])
        frames[t] = dict(zip(cols, rows.T))
        k += 9 + n
    return frames


def test_observables_between_replayed_steps(engine, tmp_path):
    # with observable fixes active, the steps on which none of them samples
    # or writes are replayed from the captured step graph: samples, files and
    # trajectory equal a run that launches every step's kernels
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 23)
    res = []
    for graph in (1, 0):
        d = tmp_path / f"graph{graph}"
        d.mkdir()
        s = mdpd_system(engine, x, v, L, extra=f"""
            fix g all rdf/fast/meso output {d}/rdf.out nbin 20 every 7
            fix p all density/meso output {d}/prof along z nbin 6 every 50 window 3
        """)
        s.tune("graph", graph)
        s.run(120)
        res.append((s.fix_result("g")[1], s.fix_result("p")[1], s.gather("x"), s.gather("v"),
                    np.loadtxt(d / "prof.000000100")))
        s.close()
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)


def test_observables_beyond_the_kept_step_graphs(engine, tmp_path):
    # five rdf fixes sampling every 2, 3, 5, 7 and 11 steps give more sample
    # patterns than step graphs are kept for, so some sampling steps launch
    # their kernels directly; an unfix between two runs drops the captured
    # graphs.  Everything equals a run that launches every step's kernels.
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 29)
    every = (2, 3, 5, 7, 11)
    res = []
    for graph in (1, 0):
        d = tmp_path / f"graph{graph}"
        d.mkdir()
        fixes = "\n".join(f"fix g{e} all rdf/fast/meso output {d}/rdf{e}.out nbin 16 every {e}" for e in every)
        s = mdpd_system(engine, x, v, L, extra=fixes + f"\nfix p all density/meso output {d}/prof along y nbin 5 every 20 window 2")
        s.tune("graph", graph)
        s.run(70)
        s.command("unfix g3")
        s.run(80)
        res.append([s.fix_result(f"g{e}")[1] for e in every if e != 3] +
                   [s.fix_result("p")[1], s.gather("x"), s.gather("v"), np.loadtxt(d / "prof.000000140")])
        s.close()
        res[-1].append(np.loadtxt(d / "rdf3.out"))
    for a, b in zip(res[0], res[1]):
        np.testing.assert_array_equal(a, b)


def test_dump_custom_frames_match_the_state_of_their_step(engine, tmp_path):
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 31)
    f = tmp_path / "traj.dump"
    s = mdpd_system(engine, x, v, L, extra=f"dump d all custom 5 {f} id type x y z vx vy vz fx fy fz")
    s.run(10)
    fr = _read_dump(f)
    assert sorted(fr) == [0, 5, 10]
    assert np.array_equal(fr[0]["id"], np.arange(1, n + 1))
    # step 0: the setup state; step 10: the end of the run
    np.testing.assert_allclose(np.c_[fr[0]["x"], fr[0]["y"], fr[0]["z"]], x, rtol=1e-7, atol=1e-7)
    for key, name in (("x", "x"), ("vx", "v"), ("fx", "f")):
        got = np.c_[fr[10][key], fr[10][key.replace("x", "y", 1) if key != "x" else "y"],
                    fr[10][key.replace("x", "z", 1) if key != "x" else "z"]]
        np.testing.assert_allclose(got, s.gather(name), rtol=1e-7, atol=1e-7)
    # step 5 (written from inside the fused step loop) equals a run of 5
    r = mdpd_system(engine, x, v, L)
    r.run(5)
    np.testing.assert_allclose(np.c_[fr[5]["x"], fr[5]["y"], fr[5]["z"]], r.gather("x"), rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(np.c_[fr[5]["vx"], fr[5]["vy"], fr[5]["vz"]], r.gather("v"), rtol=1e-7, atol=1e-7)
