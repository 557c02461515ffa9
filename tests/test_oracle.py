"""Pins for the CPU oracle (oracle/orc.c).

Each test checks an oracle function against something it does not share code
with: published known-answer vectors, closed forms, textbook routines, an
independent implementation (numpy / scipy), or a conservation law.
"""
import math

import numpy as np
import pytest
from scipy import spatial, stats

from conftest import uniform_box


# ---------------------------------------------------------------- TEA / RNG

def test_tea_published_vector(orc):
    # TEA (Wheeler & Needham 1994), 32 cycles, all-zero key and block:
    # ciphertext 41ea3a0a 94baa940.  Pins delta, shifts, operand order and
    # the sum update of the round used by math_meso.h:449-455.
    assert orc.tea_key(0, 0, 32, [0, 0, 0, 0]) == (0x41EA3A0A, 0x94BAA940)


def test_tea_is_a_permutation(orc):
    # a Feistel network is invertible: decrypt with the textbook inverse round
    rng = np.random.default_rng(7)
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    for _ in range(50):
        v0, v1 = (int(t) for t in rng.integers(0, 2**32, 2, dtype=np.uint64))
        a, b = orc.tea_key(v0, v1, 4, key)
        s = (0x9E3779B9 * 4) & 0xFFFFFFFF
        for _r in range(4):
            b = (b - ((((a << 4) + key[2]) ^ (a + s) ^ ((a >> 5) + key[3])) & 0xFFFFFFFF)) & 0xFFFFFFFF
            a = (a - ((((b << 4) + key[0]) ^ (b + s) ^ ((b >> 5) + key[1])) & 0xFFFFFFFF)) & 0xFFFFFFFF
            s = (s - 0x9E3779B9) & 0xFFFFFFFF
        assert (a, b) == (v0, v1)


def test_premix_is_xor_of_tea_words(orc):
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    for (a, b, r) in [(1, 2, 64), (0xDEADBEEF, 12345, 32), (7, 0, 16)]:
        v0, v1 = orc.tea_key(a, b, r, key)
        assert int(orc.premix(np.array([a]), np.array([b]), r)[0]) == v0 ^ v1


def test_ranmars_published_vector(orc):
    # Marsaglia & Zaman universal generator self-test: seeds ij=1802, kl=9373,
    # after 20000 draws the next six times 2^24 are these integers.
    # RanMars(seed): ij = (seed-1)/30082, kl = (seed-1) - 30082 ij
    # (src/random_mars.cpp:35-36); its constructor draws once.
    r = orc.ranmars(1802 * 30082 + 9373 + 1)
    orc.ranmars_uniform(r, 19999)
    got = orc.ranmars_uniform(r, 6) * 4096.0 * 4096.0
    assert list(got) == [6533892.0, 14220222.0, 7275067.0, 6172232.0, 8354498.0, 10633180.0]


def test_ranmars_gaussian_moments(orc):
    r = orc.ranmars(12345)
    g = orc.ranmars_gaussian(r, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01


# ---------------------------------------------------------------- bits

def test_interleave3_bit_positions(orc):
    L = orc.lib()
    for b in range(21):
        assert L.orc_interleave3(1 << b, 0, 0) == 1 << (3 * b)
        assert L.orc_interleave3(0, 1 << b, 0) == 1 << (3 * b + 1)
        assert L.orc_interleave3(0, 0, 1 << b) == 1 << (3 * b + 2)
    assert L.orc_interleave3(3, 3, 3) == 63


def test_brev32(orc):
    L = orc.lib()
    assert L.orc_brev32(1) == 0x80000000
    assert L.orc_brev32(0x0000FFFF) == 0xFFFF0000
    rng = np.random.default_rng(3)
    for x in rng.integers(0, 2**32, 100, dtype=np.uint64):
        x = int(x)
        assert L.orc_brev32(x) == int(format(x, "032b")[::-1], 2)


def test_mantissa_bits_independent(orc):
    L = orc.lib()
    rng = np.random.default_rng(4)
    v = rng.normal(0, 2, (200, 3)).astype(np.float32)
    bits = v.view(np.uint32)
    for k in range(200):
        fields = [(int(bits[k, d]) >> 12) & 0x7FF for d in range(3)]
        want = 0
        for bit in range(11):
            for d in range(3):
                want |= ((fields[d] >> bit) & 1) << (3 * bit + d)
        want &= 0xFFFFFFFF
        assert L.orc_mantissa(float(v[k, 0]), float(v[k, 1]), float(v[k, 2])) == want


# ---------------------------------------------------------------- Gaussian

def test_polynomials_against_libm(orc):
    L = orc.lib()
    xs = np.linspace(0, 1, 2001)
    err = max(abs(L.orc_cospi_poly(float(x)) - math.cos(math.pi * x)) for x in xs)
    assert err < 2e-10                      # author's bound 1.104e-10 relative
    rng = np.random.default_rng(5)
    for u in list(rng.integers(1, 2**32, 2000, dtype=np.uint64)) + [1, 2, 2**31, 2**32 - 1]:
        u = int(u)
        want = math.log2(u) - 32.0
        assert abs(L.orc_log2u_poly(u) - want) <= 5e-12 * max(1.0, abs(want)) + 1e-15


def test_gaussian_poly_vs_libm_and_symmetry(orc):
    rng = np.random.default_rng(6)
    a = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 100000, dtype=np.uint64).astype(np.uint32)
    g1 = orc.gaussian(a, b)
    g2 = orc.gaussian(a, b, libm=True)
    assert np.max(np.abs(g1 - g2)) < 1e-8
    assert np.array_equal(g1, orc.gaussian(b, a))         # xi_ij = xi_ji
    assert np.max(np.abs(g1)) <= 4.0


def test_gaussian_distribution(orc):
    rng = np.random.default_rng(8)
    a = rng.integers(0, 2**32, 400000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 400000, dtype=np.uint64).astype(np.uint32)
    g = orc.gaussian(a, b)
    assert abs(g.mean()) < 0.006
    assert abs(g.var() - 1.0) < 0.008
    assert stats.kstest(g, "norm").pvalue > 1e-3


# ---------------------------------------------------------------- keys and sort

def test_sort_key_geometry(orc):
    # bin part of the key is the Morton code of the containing cell (ghost
    # offset +1), sub part the Morton code of the 16^3 sub-cell: both from
    # plain floor() geometry for a box at the origin
    lo, hi, cut = (0, 0, 0), (7.9, 6.5, 9.1), 1.3
    x, _ = uniform_box(3000, np.subtract(hi, lo), 11, lo)
    keys = orc.sort_keys(x, lo, hi, cut)
    nb = np.floor(np.subtract(hi, lo) / cut).astype(int)
    bs = np.subtract(hi, lo) / nb
    L = orc.lib()
    for k in range(0, 3000, 7):
        c = np.floor(x[k] / bs).astype(int)
        s = np.floor((x[k] - c * bs) / bs * 16).astype(int)
        if np.any(np.abs(x[k] / bs - np.round(x[k] / bs)) < 1e-6):
            continue
        want = (L.orc_interleave3(*(int(t) + 1 for t in c)) << 12) | L.orc_interleave3(*(int(t) for t in s))
        assert int(keys[k]) == want


def test_stable_sort_matches_numpy(orc):
    rng = np.random.default_rng(12)
    keys = rng.integers(0, 50, 5000).astype(np.uint64)       # many ties
    assert np.array_equal(orc.stable_sort(keys), np.argsort(keys, kind="stable"))


def test_wrap_into_box(orc):
    lo, hi = (0.0, -2.0, 1.0), (5.0, 3.0, 4.0)
    x = np.array([[-0.1, 3.2, 0.5], [5.0, -2.0, 4.0], [4.999, 0.0, 1.0], [-1e-9, 2.99999, 3.9]], dtype=np.float32)
    w = orc.wrap(x, lo, hi)
    assert np.all(w >= np.float32(lo)) and np.all(w < np.float32(hi))
    assert np.allclose(w[0], [4.9, -1.8, 3.5], atol=1e-5)


# ---------------------------------------------------------------- neighbor sets

def test_neighbor_sets_vs_kdtree(orc):
    L = 6.0
    x, _ = uniform_box(1500, L, 21)
    cut = 1.3
    counts, rows = orc.neighbors(x, (0, 0, 0), (L, L, L), cut)
    tree = spatial.cKDTree(x, boxsize=L)
    pairs = tree.query_pairs(cut, output_type="ndarray")
    ref = [set() for _ in range(len(x))]
    for i, j in pairs:
        ref[i].add(int(j)); ref[j].add(int(i))
    mism = 0
    for i in range(len(x)):
        got = {int(e) & ((1 << 27) - 1) for e in rows[i]}
        if got != ref[i]:
            # allowed only where FP32 and FP64 disagree at the cutoff itself
            for j in got ^ ref[i]:
                d = x[i] - x[j]
                d -= L * np.round(d / L)
                assert abs(np.linalg.norm(d) - cut) < 1e-5
                mism += 1
    assert mism < 5


# ---------------------------------------------------------------- density

RD = 0.75


def _p(orc, L, **kw):
    base = dict(cut=1.0, cut_r=RD, A=-40.0, B=25.0, gamma=4.5, sigma=3.0, expw=1.0)
    base.update(kw)
    return orc.params((0, 0, 0), (L, L, L), 0.01, **base)


def test_density_two_particles_closed_form(orc):
    L = 10.0
    p = _p(orc, L, mass=(1.7,))
    for r in (0.1, 0.3, 0.5, 0.7499, 0.76, 2.0):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], dtype=np.float32)
        rho = orc.density(x, [0, 0], [1, 1], p, RD)
        rr = float(np.float32(5.0 + r) - np.float32(5.0))
        want = 1.7 * 15.0 / (2 * math.pi * RD**5) * (RD - rr) ** 2 if rr < RD else 0.0
        assert rho == pytest.approx([want, want], rel=1e-12, abs=1e-14)
        lucy = orc.density(x, [0, 0], [1, 1], p, RD, lucy=True)
        want_l = 1.7 * 105.0 / (16 * math.pi * RD**7) * (RD + 3 * rr) * (RD - rr) ** 3 if rr < RD else 0.0
        assert lucy == pytest.approx([want_l, want_l], rel=1e-12, abs=1e-14)


def test_density_single_particle_and_group(orc):
    p = _p(orc, 5.0)
    x = np.array([[1.0, 1.0, 1.0], [1.2, 1.0, 1.0]], dtype=np.float32)
    assert orc.density(x[:1], [0], [1], p, RD)[0] == 0.0      # no self term
    rho = orc.density(x, [0, 0], [1, 2], p, RD, groupbit=1)
    assert rho[1] == 0.0 and rho[0] > 0


def test_density_lattice_shell_sum(orc):
    # simple cubic lattice, spacing a: rho = sum over lattice vectors R != 0
    # with |R| < rd of w(|R|), enumerated directly
    a, n = 0.4, 10
    L = a * n
    g = np.arange(n) * a
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    p = _p(orc, L)
    rho = orc.density(x, np.zeros(len(x), int), np.ones(len(x), int), p, RD)
    want = 0.0
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                r = a * math.sqrt(i * i + j * j + k * k)
                if 0 < r < RD:
                    want += 15.0 / (2 * math.pi * RD**5) * (RD - r) ** 2
    assert np.allclose(rho, want, rtol=1e-6)


# ---------------------------------------------------------------- forces

def test_mdpd_two_particle_conservative(orc):
    L = 10.0
    p = _p(orc, L, gamma=0.0, sigma=0.0)
    for r in (0.2, 0.6, 0.9):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], dtype=np.float32)
        rr = float(np.float32(5.0 + r) - np.float32(5.0))
        rho = np.array([3.0, 5.0])
        f, e, vir = orc.force_mdpd(x, np.zeros((2, 3)), rho, [1, 2], [0, 0], p)
        wc, wr = 1 - rr, max(1 - rr / RD, 0.0)
        fmag = -40.0 * wc + 25.0 * 8.0 * wr       # along -dx (j sits at +x)
        assert f[0, 0] == pytest.approx(-fmag, rel=1e-12)
        assert f[1, 0] == pytest.approx(fmag, rel=1e-12)
        assert np.allclose(f[:, 1:], 0)
        e_pair = 0.5 * (0.5 * -40.0 * 1.0 * wc * wc + 0.5 * 25.0 * RD * 8.0 * wr * wr)
        assert e == pytest.approx([e_pair, e_pair], rel=1e-12)
        assert vir[:, 0].sum() == pytest.approx(rr * fmag, rel=1e-12)   # sum r.F over the pair


def test_mdpd_momentum_conservation(orc):
    L = 6.0
    x, v = uniform_box(1300, L, 31)
    p = _p(orc, L)
    t0 = np.zeros(len(x), int)
    rho = orc.density(x, t0, np.ones(len(x), int), p, RD)
    sig = orc.sig_mdpd(np.arange(1, len(x) + 1), orc.seed_now(99, 5))
    f, _, _ = orc.force_mdpd(x, v, rho, sig, t0, p)
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * np.abs(f).sum()


def test_mdpd_with_b_zero_is_dpd(orc):
    L = 6.0
    x, v = uniform_box(1000, L, 41)
    t0 = np.zeros(len(x), int)
    p = _p(orc, L, B=0.0)
    sig = orc.sig_mdpd(np.arange(1, len(x) + 1), 777)
    f1, _, _ = orc.force_mdpd(x, v, np.zeros(len(x)), sig, t0, p)
    f2, _, _ = orc.force_dpd(x, v, sig, t0, p)
    assert np.allclose(f1, f2, rtol=1e-10, atol=1e-10)


def test_dpd_energy_is_potential_of_force(orc):
    # with gamma = sigma = 0: F = -dE/dr for E = sum_pairs 0.5 a0 rc wc^2
    L = 10.0
    p = _p(orc, L, gamma=0.0, sigma=0.0, A=25.0)
    h = 1e-4
    for r in (0.3, 0.8):
        def energy(rv):
            x = np.array([[5.0, 5.0, 5.0], [5.0 + rv, 5.0, 5.0]])
            _, e, _ = orc.force_dpd(x, np.zeros((2, 3)), [1, 2], [0, 0], p)
            return e.sum()
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        f, _, _ = orc.force_dpd(x, np.zeros((2, 3)), [1, 2], [0, 0], p)
        dE = (energy(r + h) - energy(r - h)) / (2 * h)
        assert f[1, 0] == pytest.approx(-dE, rel=2e-3)


def test_dpd_weight_exponent_two_particles(orc):
    # dpd/meso with w_r = w_c^expw (pair_dpd_meso.cu:153): for two particles
    # on the x axis with a relative velocity, F_i = [a0 w_c - gamma w_r^2
    # (dx.dv)/r + sigma w_r xi / sqrt(dt)] dx / r, written out by hand here
    # with xi from the (separately pinned) pair Gaussian of the signatures
    L, dt, a0, g, sg = 10.0, 0.04, 25.0, 4.5, 3.0
    sig = [0x12345678, 0x9ABCDEF0]
    xi = float(orc.gaussian([sig[0]], [sig[1]])[0])
    for expw in (1.0, 0.5, 0.25):
        p = orc.params((0, 0, 0), (L, L, L), dt, cut=1.0, A=a0, gamma=g, sigma=sg, expw=expw)
        for r, dv in ((0.3, 0.7), (0.85, -1.2)):
            x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], dtype=np.float32)
            v = np.array([[dv, 0.0, 0.0], [0.0, 0.0, 0.0]], dtype=np.float32)
            rr = float(np.float32(5.0 + r) - np.float32(5.0))
            dvx = float(np.float32(dv))
            wc = 1.0 - rr
            wr = wc ** expw
            dx = -rr                                   # x_i - x_j
            fpair = (a0 * wc - g * wr * wr * (dx * dvx) / rr + sg * wr * xi / math.sqrt(dt)) / rr
            f, e, _ = orc.force_dpd(x, v, sig, [0, 0], p)
            assert f[0, 0] == pytest.approx(fpair * dx, rel=1e-12), (expw, r)
            assert f[1, 0] == pytest.approx(-fpair * dx, rel=1e-12), (expw, r)
            assert np.allclose(f[:, 1:], 0)
            assert e.sum() == pytest.approx(0.5 * a0 * wc * wc, rel=1e-12)   # two halves of 0.5 a0 rc wc^2


# ---------------------------------------------------------------- observables

def test_temperature_hand_values(orc):
    # compute temp/meso (compute_temp_meso.cu:49-101): sum m v^2 / dof with
    # dof = 3N - 3, here for four particles of two masses
    L = 10.0
    p = orc.params((0, 0, 0), (L, L, L), 0.01, mass=(1.0, 2.0), ntypes=2, cut=1.0)
    v = [[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, -1.0], [1.0, 1.0, 1.0]]
    t = [0, 1, 0, 1]
    # 1*1 + 2*4 + 1*1 + 2*3 = 16 over dof 9
    assert orc.temperature(v, t, p) == pytest.approx(16.0 / 9.0, rel=1e-15)
    assert orc.temperature([[3.0, 0.0, 0.0]], [0], p) == 0.0     # one particle: no dof


def test_pressure_ideal_and_pair_virial(orc):
    # compute pressure/meso (compute_pressure_meso.cu:115-155):
    # P = (dof T + sum of virial traces) / (3 V)
    L = 10.0
    V = L**3
    p = orc.params((0, 0, 0), (L, L, L), 0.01, mass=(1.0, 2.0), ntypes=2, cut=1.0)
    v = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, -1.0], [1.0, 1.0, 1.0]])
    t = [0, 1, 0, 1]
    # no virial: P V = sum m v^2 / 3 (the kinetic part, whatever dof is)
    assert orc.pressure(v, np.zeros((4, 6)), t, p) == pytest.approx(16.0 / 3.0 / V, rel=1e-15)
    # at rest, a conservative mDPD pair at distance r on the x axis: the
    # virial is r F(r) with the hand-evaluated pair force
    pm = _p(orc, L, gamma=0.0, sigma=0.0)
    r = 0.5
    x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], dtype=np.float32)
    rho = np.array([4.0, 6.0])
    _, _, vir = orc.force_mdpd(x, np.zeros((2, 3)), rho, [1, 2], [0, 0], pm)
    rr = float(np.float32(5.0 + r) - np.float32(5.0))
    fmag = -40.0 * (1 - rr) + 25.0 * 10.0 * max(1 - rr / RD, 0.0)
    assert orc.pressure(np.zeros((2, 3)), vir, [0, 0], pm) == pytest.approx(rr * fmag / 3.0 / V, rel=1e-12)


# ---------------------------------------------------------------- integrator

def test_free_flight_and_parabola(orc):
    L = 50.0
    x = np.array([[10.0, 10.0, 10.0], [30.0, 30.0, 30.0]])
    v = np.array([[0.5, -0.25, 1.0], [0.0, 0.0, 0.0]])
    p = _p(orc, L, A=0.0, B=0.0, gamma=0.0, sigma=0.0)
    g = (0.02, 0.0, -0.01)
    xo, vo, *_ = orc.run(x, v, [1, 2], [0, 0], [1, 1], p, 100, addforce_groupbit=1, addforce=g, skin=0.3)
    t = 100 * 0.01
    want = x + v * t + 0.5 * np.array(g) * t * t
    assert np.allclose(xo, want, atol=1e-12)
    assert np.allclose(vo, v + np.array(g) * t, atol=1e-12)
    # groups (fix_nve_meso.cu:83-92, fix_addforce_meso.cu:88-92): both feel
    # the body force of group 4; particle 1 is integrated (group 2) and
    # follows the parabola, particle 2 is outside the integration group and
    # stays where it is with its velocity unchanged
    xo, vo, *_ = orc.run(x, v + [0.0, 0.0, 0.3], [1, 2], [0, 0], [1 | 2 | 4, 1 | 4], p, 100, nve_groupbit=2,
                         addforce_groupbit=4, addforce=g, skin=0.3)
    assert np.allclose(xo[0], x[0] + (v[0] + [0.0, 0.0, 0.3]) * t + 0.5 * np.array(g) * t * t, atol=1e-12)
    assert np.array_equal(xo[1], x[1]) and np.array_equal(vo[1], [0.0, 0.0, 0.3])


# ---------------------------------------------------------------- CPU DPD statistics

def test_lammps_dpd_statistics(orc):
    # rho = 3, a = 25, gamma = 4.5, dt = 0.04: <T> = kT = 1 up to the
    # known few-percent velocity-Verlet heating at dt = 0.04; the excess
    # pressure must lie between the mean-field bound 0.1047 a rho^2
    # (g(r) = 1: (2 pi / 3) rho^2 a * 1/20) and a clearly correlated fluid.
    # The Groot-Warren fit alpha = 0.101 is the large-density limit; at rho = 3
    # this run gives alpha ~ 0.092 (p ~ 23.8), see the GPU/CPU comparison for
    # the statistical parity proper.
    L = 9.0
    n = int(3 * L**3)
    x, v = uniform_box(n, L, 51)
    out = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), a0=25.0, gamma=4.5, dt=0.04, seed=4321,
                             nsteps=400, equil=150, sample_every=5)
    assert out["T"] == pytest.approx(1.0, abs=0.05)
    alpha = (out["P"] - 3.0 * out["T"]) / (25.0 * 9.0)
    assert 0.085 < alpha < 0.1047


def test_lammps_dpd_threads(orc):
    # one thread is the serial run bit for bit (one stream, same order);
    # four threads draw from four streams and give the same statistics
    L = 9.0
    n = int(3 * L**3)
    x, v = uniform_box(n, L, 53)
    kw = dict(a0=25.0, gamma=4.5, dt=0.04, seed=4321)
    a = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=30, **kw)
    b = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=30, nthreads=1, **kw)
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    out = orc.lammps_dpd_run(x, v, (0, 0, 0), (L, L, L), nsteps=400, equil=150, sample_every=5, nthreads=4, **kw)
    assert out["T"] == pytest.approx(1.0, abs=0.05)
    alpha = (out["P"] - 3.0 * out["T"]) / (25.0 * 9.0)
    assert 0.085 < alpha < 0.1047


# ---------------------------------------------------------------- arbitrary boundary

RHO_W, CPHI = 8.0, 1.0


def _lucy3d(r, c=CPHI, rho_w=RHO_W):
    # 3-D Lucy kernel normalized over the sphere of radius c, divided by the
    # wall density: integral over a filled sphere at density rho_w is 1
    return 105.0 / (16.0 * np.pi * c**7 * rho_w) * (c + 3 * r) * (c - r) ** 3


def test_wall_ratio_closed_forms(orc):
    # phi = 0: h = rcw, no scaling
    assert orc.wall_ratio(0.0, 1.0) == pytest.approx(1.0, abs=1e-15)
    # 2.088/8 + 1.478/2 = 1 exactly: h = 0 at phi = 1/2, clamped to 0.025 rcw
    top = 1 + 0.187 * (1 / 0.025 - 1) - 0.093 * 0.975**3
    for phi in (0.5, 0.7, 3.0):
        assert orc.wall_ratio(phi, 1.0) == pytest.approx(top, rel=1e-14)
    # phi = 1/4 by hand: (2.088/64 + 1.478/4)^(1/4) = 0.402125^(1/4)
    h = 1 - 0.402125 ** 0.25
    assert orc.wall_ratio(0.25, 1.0) == pytest.approx(1 + 0.187 * (1 / h - 1) - 0.093 * (1 - h) ** 3, rel=1e-12)
    # scale invariance in the cutoff and monotone growth towards the wall
    assert orc.wall_ratio(0.25, 0.8) == pytest.approx(orc.wall_ratio(0.25, 1.0), rel=1e-12)
    # (h reaches the 0.025 clamp just below phi = 1/2: flat from there)
    r = [orc.wall_ratio(p, 1.0) for p in np.linspace(0, 0.46, 47)]
    assert np.all(np.diff(r) > 0)
    assert orc.wall_ratio(0.49, 1.0) == pytest.approx(top, rel=1e-14)


def test_wall_scaled_pair_force_two_particles(orc):
    # pair_mdpd_meso.cu:169-185: for a mobile i and a wall j, gamma scales by
    # ratio(phi_i) and sigma by sqrt(ratio); a wall i with a mobile j and a
    # mobile-mobile pair are not scaled.  Hand-evaluated with A = B = 0 on
    # two particles on the x axis (the dissipative and random terms apart)
    L, dt, g, sg, phi_i = 10.0, 0.01, 4.5, 3.0, 0.3
    ratio = orc.wall_ratio(phi_i, 1.0)
    sig = [0x01234567, 0x89ABCDEF]
    xi = float(orc.gaussian([sig[0]], [sig[1]])[0])
    r, dv = 0.6, 0.8
    x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], dtype=np.float32)
    v = np.array([[dv, 0.0, 0.0], [0.0, 0.0, 0.0]], dtype=np.float32)
    rr = float(np.float32(5.0 + r) - np.float32(5.0))
    dvx = float(np.float32(dv))
    wc, dx = 1.0 - rr, -rr
    a = orc.arb(2, 4, pair=True, fix=False)
    phi = np.array([phi_i, 0.0])
    for gam, sgm, scale in ((g, 0.0, ratio), (0.0, sg, math.sqrt(ratio))):
        p = orc.params((0, 0, 0), (L, L, L), dt, mass=(1.0, 1.0), ntypes=2, cut=1.0, cut_r=RD, A=0.0, B=0.0,
                       gamma=gam, sigma=sgm)
        base = (-gam * wc * wc * dx * dvx / rr + sgm * wc * xi / math.sqrt(dt)) / rr * dx
        for mask, want in (([1 | 2, 1 | 4], base * scale),      # mobile i, wall j: scaled
                           ([1 | 4, 1 | 2], base),              # wall i, mobile j: not scaled
                           ([1 | 2, 1 | 2], base)):             # mobile pair: not scaled
            f, _, _ = orc.force_mdpd_arb(x, v, np.zeros(2), sig, [0, 1], mask, phi, p, a)
            assert f[0, 0] == pytest.approx(want, rel=1e-12), (gam, sgm, mask)


def _arb_state(orc, x, types, L):
    n = len(x)
    mask = np.where(types == 2, 1 | 4, 1 | 2)          # all=1, fluid=2, wall=4
    p = orc.params((0, 0, 0), L, 0.01, mass=(1.0, 1.0), ntypes=2, cut=1.0, cut_r=0.75, A=-40.0, B=25.0,
                   gamma=18.0, sigma=6.0)
    a = orc.arb(2, 4, cut_phi=CPHI, phi_c=0.5, rho_wall=RHO_W)
    return mask, p, a


def test_wall_fraction_single_wall_particle(orc):
    L = (10.0, 10.0, 10.0)
    for r in (0.1, 0.37, 0.8, 0.999, 1.2):
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r * 0.6, 5.0 - r * 0.8, 5.0]])
        types = np.array([0, 1])                           # 0-based: fluid, wall
        mask, p, a = _arb_state(orc, x, types + 1, L)
        v = np.zeros((2, 3))
        xo, vo, phi, nw = orc.aboundary(x, v, types, mask, p, a)
        want = _lucy3d(r) if r < CPHI else 0.0
        assert phi[0] == pytest.approx(want, rel=1e-12, abs=1e-300)
        assert phi[1] == 0.0                                # walls are not mobile
        if r < CPHI:
            u = nw[0] / np.linalg.norm(nw[0])
            assert np.allclose(u, [-0.6, 0.8, 0.0], atol=1e-12)   # from the wall towards the particle
        assert np.array_equal(xo, x) and np.array_equal(vo, v)   # phi < phi_c: no bounce


def test_wall_fraction_of_a_half_space(orc):
    # walls fill the slab z in [8, 12) as a simple cubic lattice of spacing
    # 1/2 (density 8, one lattice point per cell of the slab); probes on the
    # slab face z = 8 see half of the kernel (phi -> 1/2), probes 2 deep see
    # all of it (phi -> 1), up to the lattice-sum error of a smooth kernel
    rng = np.random.default_rng(5)
    Lx, Lz = 12.0, 12.0
    g = np.arange(24) * 0.5 + 0.25
    walls = np.array([[a, b, c] for a in g for b in g for c in 8.25 + np.arange(8) * 0.5])
    nw_ = len(walls)
    probes = np.c_[rng.uniform(0, Lx, (400, 2)), np.full(400, 8.0)]
    deep = np.c_[rng.uniform(0, Lx, (400, 2)), np.full(400, 10.0)]
    x = np.r_[walls, probes, deep]
    types = np.r_[np.ones(nw_, int), np.zeros(800, int)]
    mask, p, a = _arb_state(orc, x, types + 1, (Lx, Lx, Lz))
    a.phi_c = 10.0                                          # no bounce: phi only
    _, _, phi, nw = orc.aboundary(x, np.zeros_like(x), types, mask, p, a)
    surf, inner = phi[nw_:nw_ + 400], phi[nw_ + 400:]
    assert abs(surf.mean() - 0.5) < 0.005
    assert abs(inner.mean() - 1.0) < 0.01
    # the normal at the surface points out of the wall (-z here)
    assert np.mean(nw[nw_:nw_ + 400, 2] < 0) > 0.95


def test_bounce_back(orc):
    # a fluid particle 0.1 inside the face (z = 5) of a three-layer lattice
    # wall, centred over a lattice point: phi > 1/2 and the normal is +z
    g = np.arange(-3, 4) * 0.5
    layer = np.array([[5 + gx, 5 + gy] for gx in g for gy in g])
    walls = np.concatenate([np.c_[layer, np.full(len(layer), z)] for z in (4.75, 4.25, 3.75)])
    for v0, want in (((0.0, 0.0, -1.0), (0.0, 0.0, 1.0)),      # into the wall: reversed
                     ((1.0, 0.5, 1.0), (-1.0, -0.5, 1.0))):   # leaving: tangential reversed
        x = np.r_[[[5.0, 5.0, 4.9]], walls]
        v = np.zeros_like(x)
        v[0] = v0
        types = np.r_[0, np.ones(len(walls), int)]
        mask, p, a = _arb_state(orc, x, types + 1, (10.0, 10.0, 10.0))
        xo, vo, phi, nw = orc.aboundary(x, v, types, mask, p, a)
        assert phi[0] > 0.5
        assert np.allclose(nw[0] / np.linalg.norm(nw[0]), [0, 0, 1], atol=1e-12)
        assert np.allclose(vo[0], want, atol=1e-12)
        # x <- (x - v dtf/m) + dt v_new, with the reference's dtf = dt/2
        assert np.allclose(xo[0], x[0] - np.array(v0) * 0.005 + 0.01 * np.array(want), atol=1e-12)
        assert np.array_equal(xo[1:], x[1:])
