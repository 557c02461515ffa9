"""End-to-end regression against the reference's static contact-angle example
(examples/v2.5/static_contact_angle_in_a_slit_nano_channel/in.gpu.static_contact):
9,280 atoms from its LAMMPS_init_noimage.dat (fluid type 1 in a slit between
equilibrated wall particles of type 2), mDPD with the arbitrary-boundary
branch (pair_style ... group_fluid1 group_solid2, fix aboundary/rho/meso with
wall fraction and bounce-back), read through read_data.

The deck's fluid density profile along y (10 bins, counts per bin, time
averaged) is compared with the reference's own outputs for two wall
wettabilities A_12 = -20, -30 and -35 (GPU run out.profile.a_*, CPU run
out.cpu.a*; the two references agree to about 1%).  The engine runs the
5,000-step pre-run, then averages the profile over 30,000 steps (the
reference: 100,000); bins inside the channel must agree with the reference
GPU profile within 3%, and the wall bins stay empty.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "contact_angle.npz")


def write_data(path, g):
    lo, hi = g["lo"], g["hi"]
    with open(path, "w") as f:
        f.write("LAMMPS data file (static contact angle initial state)\n\n")
        f.write(f"{len(g['ids'])} atoms\n2 atom types\n\n")
        for a, d in zip("xyz", range(3)):
            f.write(f"{lo[d]:.16e} {hi[d]:.16e} {a}lo {a}hi\n")
        f.write("\nMasses\n\n1 1\n2 1\n\nAtoms\n\n")
        for i, t, x in zip(g["ids"], g["types"], g["x"]):
            f.write(f"{i} {t} {x[0]:.16e} {x[1]:.16e} {x[2]:.16e}\n")


def deck(a12):
    return f"""
        units           lj
        dimension       3
        boundary        p p p
        neighbor        0.3 bin
        neigh_modify    every 1 delay 0 check yes
        atom_style      mdpd/atomic/meso 3
        read_data       {{data}}
        group           group_fluid1 type 1
        group           group_solid2 type 2
        pair_style      mdpd/meso 1.0 1.0 123456 group_fluid1 group_solid2
        pair_coeff      1 1 -40 25 18.0 1.0 0.75
        pair_coeff      1 2 {a12} 25 18.0 1.0 0.75
        pair_coeff      2 2 -40 25 18.0 1.0 0.75
        mass            1 1
        mass            2 1
        timestep        0.01
        fix             BC all aboundary/rho/meso cut_rho 0.75 phi_c 0.5 cut_phi 1.0 rho_wall 8.0 wall group_solid2 mobile group_fluid1
        fix             int_f group_fluid1 nve/meso
        run_style       mvv/meso
    """


@pytest.mark.parametrize("a12", [20, 30, 35])
def test_contact_angle_profile_matches_reference(engine, tmp_path, a12):
    g = np.load(FIX)
    data = tmp_path / "init.data"
    write_data(data, g)
    s = engine.System()
    s.commands(deck(-a12).replace("{data}", str(data)))
    fluid = g["types"] == 1
    s.run(5000)
    lo, hi = g["lo"][1], g["hi"][1]
    prof = np.zeros(10)
    nsamp = 0
    for _ in range(300):
        s.run(100)
        y = s.gather("x")[fluid, 1]
        y = lo + np.mod(y - lo, hi - lo)
        prof += np.histogram(y, bins=10, range=(lo, hi))[0]
        nsamp += 1
    prof /= nsamp
    ref = g[f"gpu_a{a12}"]
    cpu = g[f"cpu_a{a12}"]
    print(f"A12=-{a12}\n engine {np.round(prof, 1)}\n ref GPU {np.round(ref, 1)}\n ref CPU {np.round(cpu, 1)}")
    inner = slice(1, 9)
    rel = np.abs(prof[inner] - ref[inner]) / ref[inner]
    assert np.all(rel < 0.03), rel
    assert prof[0] < 1.0 and prof[9] < 1.0
    assert abs(prof.sum() - fluid.sum()) < 1e-6
