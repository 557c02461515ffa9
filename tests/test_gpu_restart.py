"""Checkpoint / resume through write_restart and read_restart.

* A setup that follows a run (any configuration change between two runs,
  here `reset_timestep`) continues from the state the run left: positions,
  velocities, image flags and wall fractions are those of the run's last
  step, not of the configuration before the run.
* read_restart of a file written after N steps, followed by M steps, equals
  bit for bit the run that wrote the file continued by a re-setup at step N
  and the same M steps (both set up from the same state), for mDPD, DPD and
  the arbitrary-boundary channel whose wall fractions phi the force pass
  reads (phi is carried in the record, as AtomVecMDPDAtomic::pack_restart
  stores it).
* The record columns hold the run's own x, v, rho, phi, tag, type, mask and
  packed image.
* A pool job submitted from a restarted system equals the plain run of it.
"""
import numpy as np
import pytest

from conftest import uniform_box
from test_gpu_arb import DECK, LX, LY, LZ, channel
from test_gpu_parity import dpd_system, mdpd_system

pytestmark = pytest.mark.gpu

REC = 13


def records(path, n):
    raw = path.read_bytes()                    # the header is not a whole number of doubles
    return np.frombuffer(raw[len(raw) - 8 * REC * n:], np.float64).reshape(n, REC)


def state(s):
    return {k: s.gather(k) for k in ("x", "v", "f", "image")}


def assert_same(a, b, what):
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def test_setup_after_a_run_continues_from_the_run(engine):
    L = 8.0
    x0, v0 = uniform_box(2400, L, 71)
    s = mdpd_system(engine, x0, v0, L)
    s.run(20)
    x, v, im = s.gather("x"), s.gather("v"), s.gather("image")
    assert not np.array_equal(x, x0)
    s.command("reset_timestep 20")
    s.run(0)                                   # a new setup from the host copy
    # the setup's sort wraps the few positions the run left just outside the
    # box and counts the crossing in the image flags; every other position is
    # unchanged to the bit
    x2, im2 = s.gather("x"), s.gather("image")
    same = (im2 == im).all(axis=1)
    assert same.sum() >= len(x) - 32
    np.testing.assert_array_equal(x2[same], x[same])
    np.testing.assert_allclose(x2 + L * im2, x + L * im, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(s.gather("v"), v)
    assert s.timestep == 20


MDPD_AFTER = """
    neighbor 0.3 bin
    neigh_modify every 1 delay 0 check yes
    fix rho all aboundary/rho/meso cut_rho 0.75
    fix int all nve/meso
    run_style mvv/meso
"""
DPD_AFTER = """
    neighbor 0.3 bin
    neigh_modify every 1 delay 0 check yes
    fix int all nve/meso
    run_style mvv/meso
"""


@pytest.mark.parametrize("style", ["mdpd", "dpd"])
def test_restart_resumes_bitwise(engine, tmp_path, style):
    L = 8.0
    x0, v0 = uniform_box(2400, L, 73)
    a = mdpd_system(engine, x0, v0, L) if style == "mdpd" else dpd_system(engine, x0, v0, L)
    a.run(20)
    path = tmp_path / "state.restart"
    a.command(f"write_restart {path}")
    rec = records(path, 2400)
    np.testing.assert_array_equal(rec[:, 0], REC)
    np.testing.assert_array_equal(rec[:, 1:4], a.gather("x"))
    np.testing.assert_array_equal(rec[:, 4:7], a.gather("v"))
    if style == "mdpd":
        np.testing.assert_array_equal(rec[:, 7], a.gather("rho"))
    np.testing.assert_array_equal(rec[:, 9], a.gather("tag"))
    np.testing.assert_array_equal(rec[:, 10], a.gather("type"))
    np.testing.assert_array_equal(rec[:, 11], a.gather("mask"))
    im = a.gather("image")
    np.testing.assert_array_equal(rec[:, 12], (im[:, 0] + 512) + ((im[:, 1] + 512) << 10) + ((im[:, 2] + 512) << 20))
    a.command("reset_timestep 20")
    a.run(25)

    b = engine.System()
    b.command(f"read_restart {path}")
    assert b.timestep == 20 and b.natoms == 2400
    b.commands(MDPD_AFTER if style == "mdpd" else DPD_AFTER)
    b.run(25)
    assert b.timestep == a.timestep == 45
    assert_same(state(b), state(a), f"{style} restart")


def test_restart_carries_wall_fractions(engine, tmp_path):
    x, v, types, nf = channel(seed=5)
    n = len(x)
    a = engine.System()
    a.command("atom_style mdpd/atomic/meso 8.0")
    a.create_box(2, (0, 0, 0), (LX, LY, LZ))
    a.create_atoms(types, x, v)
    a.commands(DECK)
    a.run(10)
    phi = a.gather("phi")
    assert phi.max() > 0.5                     # some fluid is inside a wall
    path = tmp_path / "channel.restart"
    a.command(f"write_restart {path}")
    np.testing.assert_array_equal(records(path, n)[:, 8], phi)
    a.command("reset_timestep 10")
    a.run(10)

    b = engine.System()
    b.command(f"read_restart {path}")
    assert b.group_bit("fluid") == a.group_bit("fluid") and b.group_bit("wall") == a.group_bit("wall")
    b.commands("""
        neighbor 0.3 bin
        neigh_modify every 1 delay 0 check yes
        fix BC all aboundary/rho/meso cut_rho 0.75 phi_c 0.5 cut_phi 1.0 rho_wall 8.0 wall wall mobile fluid
        fix int fluid nve/meso
        run_style mvv/meso
    """)
    b.run(10)
    assert_same(state(b), state(a), "channel restart")
    np.testing.assert_array_equal(b.gather("phi"), a.gather("phi"))


def test_pool_job_from_a_restarted_system(engine, tmp_path):
    x, v, types, nf = channel(seed=6)
    a = engine.System()
    a.command("atom_style mdpd/atomic/meso 8.0")
    a.create_box(2, (0, 0, 0), (LX, LY, LZ))
    a.create_atoms(types, x, v)
    a.commands(DECK)
    a.run(8)
    path = tmp_path / "c.restart"
    a.command(f"write_restart {path}")

    def restarted():
        s = engine.System()
        s.command(f"read_restart {path}")
        s.commands("""
            neighbor 0.3 bin
            neigh_modify every 1 delay 0 check yes
            fix BC all aboundary/rho/meso cut_rho 0.75 phi_c 0.5 cut_phi 1.0 rho_wall 8.0 wall wall mobile fluid
            fix int fluid nve/meso
            run_style mvv/meso
        """)
        return s

    plain = restarted()
    plain.tune("split", 2)
    plain.run(12)
    p = engine.Pool(0.5 * float(1 << 30), 4)
    p.tune("split", 2)
    j = p.submit(restarted(), 12)
    p.run()
    assert p.state(j) == engine.JOB_DONE
    for k in ("x", "v", "f", "image"):
        np.testing.assert_array_equal(p.gather(j, k), plain.gather(k), err_msg=k)
    assert p.info(j)["timestep"] == plain.timestep == 20
