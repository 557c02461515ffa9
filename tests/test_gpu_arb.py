"""Arbitrary-boundary branch (pair_style mdpd/meso ... mobile wall +
fix aboundary/rho/meso ... wall G mobile G) against the oracle.

A slit channel: frozen wall particles (type 2, density 8) fill |y| > 4 of a
periodic box, fluid (type 1) fills the channel and starts up to 0.2 inside
the wall faces; the contact-angle deck's coefficients (A11 = A22 = -40,
A12 = -20, B = 25, gamma = 18, rc = 1, rd = 0.75; phi_c = 0.5, cut_phi = 1,
rho_wall = 8).

* Wall-scaled force on one state: after a step, the engine's forces equal the
  FP64 oracle's on the engine's own FP32 state (positions, half-step
  velocities, densities, wall fractions, signatures) within 1e-5 relative,
  globally and per particle -- the bound of the plain force test.
* Trajectory: positions after 1 and 5 steps within 1e-5 relative per
  particle; velocities within 1e-5 globally and 1e-4 per particle (a bounced
  particle deep in a wall sees wall-scaled gamma up to 8x and sigma up to
  2.9x, and the half-kick of those forces amplifies FP32 rounding); wall
  fractions after one step within 1e-4.  Bounce decisions of particles
  within 1e-4 of phi_c are excused.
* A longer run against a lattice wall (no holes) keeps the fluid at the wall
  faces.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_vec_close

pytestmark = pytest.mark.gpu

LX, LY, LZ = 10.0, 10.0, 6.0
DECK = """
    neighbor 0.3 bin
    neigh_modify every 1 delay 0 check yes
    mass * 1.0
    group fluid type 1
    group wall type 2
    pair_style mdpd/meso 1.0 1.0 2654868 fluid wall
    pair_coeff 1 1 -40 25 18.0 1.0 0.75
    pair_coeff 1 2 -20 25 18.0 1.0 0.75
    pair_coeff 2 2 -40 25 18.0 1.0 0.75
    fix BC all aboundary/rho/meso cut_rho 0.75 phi_c 0.5 cut_phi 1.0 rho_wall 8.0 wall wall mobile fluid
    fix int fluid nve/meso
    timestep 0.01
    run_style mvv/meso
"""


def channel(seed=3, lattice=False):
    rng = np.random.default_rng(seed)
    if lattice:
        # simple cubic lattice of spacing 1/2 (density 8) in y in [0, 1) U [9, 10)
        g = lambda L: np.arange(int(round(L / 0.5))) * 0.5 + 0.25
        ys = np.r_[g(1.0), 9.0 + g(1.0)]
        walls = np.array([[a, b, c] for a in g(LX) for b in ys for c in g(LZ)])
    else:
        nw = int(8.0 * LX * 2.0 * LZ)
        wy = rng.uniform(0, 2.0, nw)
        wy = np.where(wy < 1.0, wy, wy + 8.0)             # y in [0, 1) or [9, 10)
        walls = np.c_[rng.uniform(0, LX, nw), wy, rng.uniform(0, LZ, nw)]
    nw = len(walls)
    nf = int(3.0 * LX * 8.0 * LZ)
    fluid = np.c_[rng.uniform(0, LX, nf), rng.uniform(0.8, 9.2, nf), rng.uniform(0, LZ, nf)]
    x = np.r_[fluid, walls].astype(np.float32).astype(np.float64)
    v = np.zeros_like(x)
    v[:nf] = rng.normal(0, 1, (nf, 3))
    v[:nf] -= v[:nf].mean(axis=0)
    v = v.astype(np.float32).astype(np.float64)
    types = np.r_[np.ones(nf, np.int32), np.full(nw, 2, np.int32)]
    return x, v, types, nf


def engine_run(engine, x, v, types, nsteps):
    s = engine.System()
    s.command("atom_style mdpd/atomic/meso 8.0")
    s.create_box(2, (0, 0, 0), (LX, LY, LZ))
    s.create_atoms(types, x, v)
    s.commands(DECK)
    s.run(nsteps)
    return s


def oracle_run(orc, x, v, types, nsteps):
    n = len(x)
    mask = np.where(types == 2, 1 | 4, 1 | 2)              # all=1, fluid=2, wall=4
    p = orc.params((0, 0, 0), (LX, LY, LZ), 0.01, mass=(1.0, 1.0), ntypes=2, cut=1.0, cut_r=0.75,
                   A=np.array([[-40.0, -20.0], [-20.0, -40.0]]), B=25.0, gamma=18.0, sigma=6.0)
    a = orc.arb(2, 4, cut_phi=1.0, phi_c=0.5, rho_wall=8.0)
    phi = np.zeros(n)
    xo, vo, f, rho, nb = orc.run(x, v, np.arange(1, n + 1), types - 1, mask, p, nsteps, style=0,
                                 seed=2654868, skin=0.3, cut_rho=0.75, rho_groupbit=1, nve_groupbit=2,
                                 arb=a, phi=phi)
    return xo, vo, phi


def test_wall_scaled_force_on_one_state(engine, orc):
    x, v, types, nf = channel()
    s = engine_run(engine, x, v, types, 1)
    n = len(x)
    gx, gv, gf = s.gather("x"), s.gather("v"), s.gather("f")
    rho, phi = s.gather("rho"), s.gather("phi")
    fluid = types == 1
    vh = gv.copy()
    vh[fluid] -= 0.005 * gf[fluid]                     # the force pass read v(1/2) = v(1) - dt/2 f / m
    mask = np.where(types == 2, 1 | 4, 1 | 2)
    p = orc.params((0, 0, 0), (LX, LY, LZ), 0.01, mass=(1.0, 1.0), ntypes=2, cut=1.0, cut_r=0.75,
                   A=np.array([[-40.0, -20.0], [-20.0, -40.0]]), B=25.0, gamma=18.0, sigma=6.0)
    a = orc.arb(2, 4, cut_phi=1.0, phi_c=0.5, rho_wall=8.0)
    sig = orc.sig_mdpd(np.arange(1, n + 1), orc.seed_now(2654868, 1))
    fr, _, _ = orc.force_mdpd_arb(gx, vh.astype(np.float32), rho, sig, types - 1, mask, phi, p, a)
    scaled = fluid & (phi > 0)
    print(f"{int(scaled.sum())} fluid particles with wall-scaled pairs, max phi {phi.max():.3f}")
    assert scaled.sum() > 50
    assert_vec_close(gf, fr, 1e-5, "wall-scaled force")


@pytest.mark.parametrize("nsteps", [1, 5])
def test_channel_steps_match_oracle(engine, orc, nsteps):
    x, v, types, nf = channel()
    s = engine_run(engine, x, v, types, nsteps)
    gx, gv, gphi = s.gather("x"), s.gather("v"), s.gather("phi")
    ox, ov, ophi = oracle_run(orc, x, v, types, nsteps)
    keep = np.abs(ophi - 0.5) >= 1e-4                     # bounce decisions on the threshold excused
    print(f"{nsteps} steps: {int((ophi[:nf] > 0.5).sum())} fluid particles past phi_c, "
          f"{int((~keep).sum())} excused")
    if nsteps == 1:
        assert (ophi[:nf] > 0.5).sum() >= 10                # the bounce-back is exercised
        assert np.max(np.abs(gphi - ophi)) < 1e-4
    for got, ref, what, per in (((gx - x), (ox - x), "displacement", 1e-5), (gv, ov, "velocity", 1e-4)):
        rms = np.sqrt(np.mean(np.sum(ref[keep] ** 2, axis=1)))
        err = np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), rms)
        glob = np.linalg.norm((got - ref)[keep]) / np.linalg.norm(ref[keep])
        w = int(np.argmax(np.where(keep, err, -1)))
        print(f"{what} after {nsteps}: global {glob:.3e}, worst {err[keep].max():.3e} "
              f"(particle {w}, phi {ophi[w]:.4f}, y {x[w, 1]:.3f})")
        assert glob <= 1e-5
        assert err[keep].max() <= per
    assert np.array_equal(gx[nf:], x[nf:])                  # frozen walls


def test_fluid_stays_out_of_the_walls(engine):
    x, v, types, nf = channel(4, lattice=True)
    s = engine_run(engine, x, v, types, 400)
    y = s.gather("x")[:nf, 1] % LY
    phi = s.gather("phi")[:nf]
    print(f"fluid y range {y.min():.3f} .. {y.max():.3f}, max phi {phi.max():.3f}")
    # bounce-back (phi > 1/2, i.e. past the face of the lattice wall at
    # y = 1 and y = 9) keeps the fluid at the faces
    assert np.sum((y < 0.7) | (y > 9.3)) == 0
    assert phi.max() < 0.6
    assert np.all(np.isfinite(s.gather("v")))
