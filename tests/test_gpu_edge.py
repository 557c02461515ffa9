"""Edge cases of box size and particle count against the oracle.

* Boxes one, two and three cells wide (cell side >= r_c + skin = 1.3): a cell
  meets itself in its own 27-cell stencil under different periodic images, so
  a neighbor can appear with two image codes.  Neighbor sets bit-exact,
  densities and forces at the bounds of the main parity tests (1e-5, with
  the FP32 pair-distance floor for densities).
* One particle: its images are farther than r_c + skin, it has no neighbors,
  no force, and flies straight (x = x0 + v t).
* Two particles: the pair force against the oracle.
* A partial last tile (N = 65: one full tile of 64 and one particle).
* Particles exactly on the box faces, just below L, and a period or more
  outside, in a sparse box: wrap, sort keys and order bit-exact against the
  oracle, densities and forces as above.
* A droplet whose surface crosses a face, plus isolated pairs and triples
  across the faces: densities made of one to three weights, some through an
  image, and a 5-step run against the oracle.
"""
import numpy as np
import pytest

from conftest import uniform_box
from test_gpu_parity import (IDX, RD, SEED, _neighbor_sets, assert_rho_close, assert_vec_close, mdpd_system,
                             oracle_params)

pytestmark = pytest.mark.gpu


def _check_against_oracle(engine, orc, s, L, n, what):
    origin, _, xw = s.debug_sorted()
    counts, rows = s.debug_neighbors()
    got = _neighbor_sets(origin, counts, rows)
    xc = np.empty_like(xw)
    xc[origin] = xw
    _, rrows = orc.neighbors(xc, (0, 0, 0), (L, L, L), 1.3)
    repeated = 0
    for i in range(n):
        ref = {(int(e) & IDX, int(e) >> 27) for e in rrows[i]}
        assert got[i] == ref, (what, i)
        js = [j for j, _ in ref]
        repeated += len(js) - len(set(js))
    xg, vg = s.gather("x"), s.gather("v")
    p = oracle_params(orc, L)
    t0 = np.zeros(n, int)
    rho = orc.density(xg, t0, np.ones(n, int), p, RD)
    assert_rho_close(orc, s.gather("rho"), rho, xg, L, what=f"density {what}")
    sig = orc.sig_mdpd(np.arange(1, n + 1), orc.seed_now(SEED, 0))
    f, _, _ = orc.force_mdpd(xg, vg, rho, sig, t0, p)
    return f, repeated


@pytest.mark.parametrize("L,n", [(1.5, 20), (2.0, 48), (2.7, 65), (3.9, 360)])
def test_small_periodic_boxes_match_oracle(engine, orc, L, n):
    x, v = uniform_box(n, L, 31)
    s = mdpd_system(engine, x, v, L)
    s.setup()
    f, repeated = _check_against_oracle(engine, orc, s, L, n, f"L={L}")
    print(f"L={L} n={n}: {repeated} neighbors listed under more than one image")
    if L < 2.6:
        assert repeated > 0          # the multi-image case is exercised
    assert_vec_close(s.gather("f"), f, what=f"force L={L}")


def test_one_particle_flies_free(engine):
    L = 1.4
    x0 = np.array([[0.3, 0.7, 1.1]])
    v0 = np.array([[0.1, -0.2, 0.3]])
    s = mdpd_system(engine, x0, v0, L)
    s.setup()
    counts, _ = s.debug_neighbors()
    assert counts[0] == 0
    s.run(5)
    assert np.all(s.gather("f") == 0)
    np.testing.assert_allclose(s.gather("x"), x0 + 5 * 0.01 * v0, atol=1e-6)
    np.testing.assert_allclose(s.gather("v"), v0.astype(np.float32), atol=1e-7)


def test_two_particles_match_oracle(engine, orc):
    L = 2.8
    x = np.array([[0.4, 1.0, 1.2], [0.9, 1.3, 0.9]])
    v = np.array([[0.5, 0.1, -0.2], [-0.3, 0.2, 0.4]])
    s = mdpd_system(engine, x, v, L)
    s.setup()
    f, _ = _check_against_oracle(engine, orc, s, L, 2, "two particles")
    assert np.linalg.norm(f[0]) > 0
    assert_vec_close(s.gather("f"), f, what="two-particle force")
    np.testing.assert_allclose(s.gather("f").sum(axis=0), 0, atol=1e-5)


def test_coincident_particles_are_left_out(engine, orc):
    # particles at exactly the same position: each lists the other (a
    # neighbor is left out of its own row by identity, not by distance), and
    # the density and force passes leave the pair out (r^2 < 1e-20,
    # fix_aboundary_rho_meso.cu:155, pair_mdpd_meso.cu:140); a short run
    # separates them with finite forces
    L, n = 4.0, 400
    x, v = uniform_box(n, L, 57)
    for a, b in ((0, 1), (10, 11), (20, 21), (30, 31), (30, 32)):
        x[b] = x[a]
    s = mdpd_system(engine, x, v, L)
    s.setup()
    f, _ = _check_against_oracle(engine, orc, s, L, n, "coincident")
    assert_vec_close(s.gather("f"), f, what="coincident pairs: force")
    s.run(5)
    for k in ("x", "v", "f", "rho"):
        assert np.isfinite(s.gather(k)).all(), k


def test_particles_on_the_box_faces(engine, orc):
    # positions exactly on the faces, the float just below L, tiny negatives
    # that round onto L after the wrap, and whole periods outside: wrap, keys,
    # sorted order bit-exact; neighbor sets, densities, forces as above.  The
    # box is sparse (1.2 per unit volume), so many densities are made of one
    # or two weights, some across a face: the displacement through an image
    # must carry no rounding at the magnitude of L
    L, n = 8.0, 600
    x, v = uniform_box(n, L, 41)
    faces = [0.0, L, float(np.nextafter(np.float32(L), np.float32(0))), -1e-7, L + 1e-7, 2 * L, -L, -0.0]
    for k, val in enumerate(faces):
        x[k, 0] = val
        x[20 + k, 1] = val
        x[40 + k, 2] = val
    s = mdpd_system(engine, x, v, L)
    s.setup()
    origin, keys, xw = s.debug_sorted()
    xc = orc.wrap(x.astype(np.float32), (0, 0, 0), (L, L, L))
    assert np.array_equal(xw, xc[origin])
    ref_keys = orc.sort_keys(xc, (0, 0, 0), (L, L, L), 1.3)
    assert np.array_equal(keys, ref_keys[origin])
    assert np.array_equal(origin, orc.stable_sort(ref_keys))
    f, _ = _check_against_oracle(engine, orc, s, L, n, "faces")
    assert_vec_close(s.gather("f"), f, what="force with particles on the faces")


def test_droplet_surface_across_a_face(engine, orc):
    # a liquid droplet (rho 6.05) centred on the x face of a periodic box, so
    # its surface particles, with one to three neighbors inside r_d, meet
    # their neighbors through the image; plus a few isolated pairs and
    # triples straddling the faces at distances just inside r_d
    L, R = 10.0, 3.2
    rng = np.random.default_rng(61)
    n = int(6.05 * 4.0 / 3.0 * np.pi * R**3)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    xd = np.array([0.2, 5.0, 5.0]) + d * (R * rng.uniform(0, 1, n) ** (1.0 / 3.0))[:, None]
    few = [[0.03, 1.0, 1.0], [L - 0.69, 1.02, 1.05],                      # a pair across x, r ~ 0.72
           [2.0, 0.02, 8.0], [2.1, L - 0.6, 8.05], [2.3, 0.5, 8.0],       # a triple across y
           [7.0, 8.0, L - 0.01], [7.2, 8.1, 0.7], [6.9, 8.3, 0.05]]      # a triple across z
    x = np.vstack([xd, few])
    x = x.astype(np.float32).astype(np.float64)
    v = rng.normal(0, 1, x.shape).astype(np.float32).astype(np.float64)
    s = mdpd_system(engine, x, v, L)
    s.setup()
    f, _ = _check_against_oracle(engine, orc, s, L, len(x), "droplet across a face")
    assert_vec_close(s.gather("f"), f, what="force, droplet across a face")
    # the run keeps tracking the oracle through several re-sorts
    x0, v0 = s.gather("x"), s.gather("v")
    s.run(5)
    p = oracle_params(orc, L)
    m = len(x)
    xo, vo, fo, rho, _ = orc.run(x0, v0, np.arange(1, m + 1), np.zeros(m, int), np.ones(m, int), p, 5, style=0,
                                 seed=SEED, ntimestep=0)
    dx = s.gather("x") - xo
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 1e-4
    assert_vec_close(s.gather("v"), vo, tol=1e-4, what="v after 5 steps, droplet across a face")
