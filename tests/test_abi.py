"""C ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and the input-script surface validates commands the
way the reference styles do (error paths return codes with messages)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import REPO


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(mdpd_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_exports_every_declared_symbol(engine):
    lib = ctypes.CDLL(engine.library_path())
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_version(engine):
    assert b"mdpd355" in engine.load().mdpd_version()


def _deck(s, L=8.0):
    s.commands("""
        units lj
        dimension 3
        boundary p p p
        atom_style mdpd/atomic/meso 6.05
    """)
    s.create_box(1, (0, 0, 0), (L, L, L))


def test_script_surface_accepts_reference_deck(engine):
    s = engine.System()
    _deck(s)
    s.commands("""
        neighbor 0.3 bin
        neigh_modify every 1 delay 0 check yes
        mass 1 1.0
        pair_style mdpd/meso 1.0 1.0 123456
        pair_coeff * * -40 25 4.5 1.0 0.75       # cutA=1.0, cutB=0.75
        fix BC all aboundary/rho/meso cut_rho 0.75
        timestep 0.01
        fix int_f all nve/meso
        run_style mvv/meso
        thermo 100
    """)
    s.close()


@pytest.mark.parametrize("line,code", [
    ("pair_coeff * * -40 25 4.5 0.7 0.75", 1),            # cutA < cutB (pair_mdpd_meso.cu:346)
    ("pair_coeff * * -40 25 4.5 1.0", 1),                 # wrong arity (:332)
    ("pair_coeff 2 2 -40 25 4.5 1.0 0.75", 1),            # type out of range
    ("pair_style mdpd/meso 1.0 1.0", 1),                  # too few args (:301)
    ("fix r all aboundary/rho/meso", 1),                  # no cut_rho (:98-99)
    ("fix r nogroup nve/meso", 1),                        # unknown group
    ("pair_style lj/cut 2.5", 6),                         # outside this engine
    ("run_style verlet", 6),
    ("frobnicate 1 2", 6),
])
def test_script_errors(engine, line, code):
    s = engine.System()
    _deck(s)
    s.command("pair_style mdpd/meso 1.0 1.0 123456")
    with pytest.raises(engine.MdpdError) as ei:
        s.command(line)
    assert ei.value.code == code
    assert str(ei.value)
    s.close()


def test_atom_style_requires_density_estimate(engine):
    s = engine.System()
    with pytest.raises(engine.MdpdError) as ei:
        s.command("atom_style mdpd/atomic/meso")
    assert "estimation" in str(ei.value)


def test_pair_coeff_before_style_and_missing_coeffs(engine):
    s = engine.System()
    _deck(s)
    with pytest.raises(engine.MdpdError) as ei:
        s.command("pair_coeff * * -40 25 4.5 1.0 0.75")
    assert ei.value.code == 2
    x = np.random.default_rng(0).uniform(0, 8, (100, 3))
    s.create_atoms(np.ones(100, np.int32), x)
    s.commands("mass 1 1.0\npair_style mdpd/meso 1.0 1.0 5\nrun_style mvv/meso")
    with pytest.raises(engine.MdpdError) as ei:
        s.setup()
    assert "pair coeffs" in str(ei.value)


def test_groups_and_gather_before_setup(engine):
    s = engine.System()
    s.command("atom_style dpd/atomic/meso")
    s.create_box(2, (0, 0, 0), (5, 5, 5))
    x = np.random.default_rng(1).uniform(0, 5, (10, 3))
    t = np.array([1, 2] * 5, np.int32)
    s.create_atoms(t, x)
    s.command("group walls type 2")
    bit = s.group_bit("walls")
    assert bit == 2
    m = s.gather("mask")
    assert np.array_equal(m, np.where(t == 2, 1 | bit, 1))
    assert np.allclose(s.gather("x"), x)
    assert np.array_equal(s.gather("tag"), np.arange(1, 11))


def test_open_does_not_need_a_device(engine):
    # parsing works without touching HIP; setup is where the device is needed
    s = engine.System(device=0)
    s.command("units lj")
    s.close()


def _job(engine, n=500, L=8.0, seed=1):
    s = engine.System()
    _deck(s, L)
    rng = np.random.default_rng(seed)
    s.create_atoms(np.ones(n, np.int32), rng.uniform(0, L, (n, 3)))
    s.commands("""
        mass 1 1.0
        pair_style mdpd/meso 1.0 1.0 123456
        pair_coeff * * -40 25 4.5 1.0 0.75
        fix rho all aboundary/rho/meso cut_rho 0.75
        fix int all nve/meso
        timestep 0.01
        run_style mvv/meso
    """)
    return s


def test_pool_queue_and_errors_without_a_device(engine):
    # submission validates and copies the system on the host; the device is
    # first touched when the pool is advanced
    p = engine.Pool(64e6, 4)
    j = p.submit(_job(engine), 10)
    assert p.state(j) == engine.JOB_QUEUED
    assert p.stat("queued") == 1 and p.stat("running") == 0
    # 64 MB at ~330 + 8*maxnbr bytes per slot; rows are a multiple of 32 wide
    # (a split first pass keeps each wave's compact rows inside the tile)
    assert p.stat("maxnbr") > 0 and p.stat("maxnbr") % 32 == 0
    assert 40000 < p.stat("capacity") < 64e6 / 1000
    with pytest.raises(engine.MdpdError) as e:
        p.gather(j, "x")
    assert e.value.code == 2
    with pytest.raises(engine.MdpdError) as e:
        p.release(j)
    assert e.value.code == 2
    with pytest.raises(engine.MdpdError) as e:
        p.submit(_job(engine), 0)
    assert e.value.code == 1
    # a job larger than the whole budget is refused at submission
    small = engine.Pool(2e6, 4)
    with pytest.raises(engine.MdpdError) as e:
        small.submit(_job(engine, n=4000, L=9.0), 10)
    assert e.value.code == 5
    # an incomplete script is refused before it reaches the queue
    s = engine.System()
    _deck(s)
    s.create_atoms(np.ones(10, np.int32), np.random.default_rng(0).uniform(0, 8, (10, 3)))
    s.command("run_style mvv/meso")
    with pytest.raises(engine.MdpdError) as e:
        p.submit(s, 10)
    assert e.value.code == 2
    with pytest.raises(engine.MdpdError):
        p.tune("split", 3)
    with pytest.raises(engine.MdpdError):
        p.stat("nonsense")


def test_read_data_atoms_masses_velocities(engine, tmp_path):
    # LAMMPS data file layout (src/read_data.cpp; Atoms lines of
    # atom_style mdpd/atomic/meso: id type x y z [ix iy iz])
    f = tmp_path / "in.data"
    f.write_text("""LAMMPS data file via write_data

4 atoms
2 atom types

-3.0 3.0 xlo xhi
-1.0 1.0 ylo yhi
0.0 2.0 zlo zhi

Masses

1 1.0
2 2.5

Atoms # mdpd/atomic/meso

10 2 -2.5 0.5 1.0
3 1 0.0 0.0 0.5 0 0 1
7 1 1.5 -0.5 1.5
1 2 2.9 0.9 0.1

Velocities

3 0.5 0.0 0.0
10 -1.0 2.0 3.0
1 0 0 0
7 0.1 0.2 0.3
""")
    s = engine.System()
    s.command("atom_style mdpd/atomic/meso 3")
    s.command(f"read_data {f}")
    assert s.natoms == 4
    assert list(s.gather("tag")) == [10, 3, 7, 1]
    assert list(s.gather("type")) == [2, 1, 1, 2]
    np.testing.assert_array_equal(s.gather("x")[1], [0.0, 0.0, 0.5])
    np.testing.assert_array_equal(s.gather("v")[0], [-1.0, 2.0, 3.0])
    s.commands("group heavy type 2\ngroup both union heavy all\nvelocity heavy set 0 NULL 0")
    np.testing.assert_array_equal(s.gather("v")[0], [0.0, 2.0, 0.0])
    assert s.group_bit("both") != 0
    # the box is defined by the file: a second read_data is refused
    with pytest.raises(engine.MdpdError) as e:
        s.command(f"read_data {f}")
    assert e.value.code == 2
    t = engine.System()
    t.command("atom_style mdpd/atomic/meso 3")
    with pytest.raises(engine.MdpdError) as e:
        t.command(f"read_data {tmp_path / 'missing.data'}")
    assert e.value.code == 1
    # write_data -> read_data round trip: wrapped positions with image flags
    out = tmp_path / "out.data"
    s.command(f"write_data {out}")
    r = engine.System()
    r.command("atom_style mdpd/atomic/meso 3")
    r.command(f"read_data {out}")
    for name in ("tag", "type", "x", "v"):
        np.testing.assert_array_equal(r.gather(name), s.gather(name))
    line3 = next(l for l in out.read_text().splitlines() if l.startswith("3 1 "))
    assert line3.endswith(" 0 0 1")                          # its image flags survive


def test_every_gather_name_reaches_the_engine(engine):
    # each per-atom name of the binding is known to the engine: before setup
    # the engine answers with its own state error, not an unknown-name error
    s = engine.System()
    _deck(s)
    s.create_atoms(np.ones(3, np.int32), np.zeros((3, 3)))
    for name in engine._FIELDS:
        if name in ("x", "v", "tag", "type", "mask"):
            s.gather(name)
            continue
        with pytest.raises(engine.MdpdError) as e:
            s.gather(name)
        assert e.value.code == 2 and "only available after setup" in str(e.value)


def test_equal_style_variables(engine):
    # variable NAME equal FORMULA; ${NAME} and $x substitute its value
    s = engine.System()
    s.commands("""
        variable    cut_att   equal 1.00
        variable    skin      equal 0.3*${cut_att}
        variable    x         equal -(2^3 - 1)/2 + ${skin}
        units       lj
        atom_style  mdpd/atomic/meso 3
        neighbor    ${skin} bin
    """)
    s.create_box(1, (0, 0, 0), (5, 5, 5))
    s.command("mass 1 ${cut_att}")
    with pytest.raises(engine.MdpdError) as e:
        s.command("neighbor ${undefined} bin")
    assert e.value.code == 1 and "illegal variable" in str(e.value)
    with pytest.raises(engine.MdpdError):
        s.command("variable bad equal 1+(2")
        s.command("mass 1 ${bad}")
    with pytest.raises(engine.MdpdError) as e:
        s.command("variable s string abc")
    assert e.value.code == 6


def _restart_source(engine):
    s = engine.System()
    s.command("atom_style mdpd/atomic/meso 7.5")
    s.create_box(2, (-1, 0, 0), (4, 5, 6))
    rng = np.random.default_rng(4)
    x = rng.uniform((-1, 0, 0), (4, 5, 6), (30, 3))
    x[3] = (9.5, -2.0, 6.5)                      # outside the box: write_data wraps it
    v = rng.normal(size=(30, 3))
    s.create_atoms(np.array([1, 2] * 15, np.int32), x, v)
    s.commands("""
        mass 1 1.0
        mass 2 2.5
        group fluid type 1
        group wall type 2
        pair_style mdpd/meso 1.5 1.0 777 fluid wall
        pair_coeff 1 1 -40 25 18.0 1.0 0.75
        pair_coeff 1 2 -20 25 18.0 1.0 0.75
        timestep 0.02
        reset_timestep 300
    """)
    return s, x, v


def test_restart_round_trip_without_a_device(engine, tmp_path):
    # write_restart / read_restart keep box, masses, groups, pair settings
    # and coefficients, step, dt and the atoms; a restored system writes the
    # same file again byte for byte
    import struct
    s, x, v = _restart_source(engine)
    path = tmp_path / "a.restart"
    s.command(f"write_restart {path}")
    r = engine.System()
    r.command(f"read_restart {path}")
    assert r.natoms == 30 and r.timestep == 300
    for name in ("tag", "type", "mask", "x", "v"):
        np.testing.assert_array_equal(r.gather(name), s.gather(name))
    assert (r.group_bit("fluid"), r.group_bit("wall")) == (2, 4)
    r.command(f"write_restart {tmp_path / 'b.restart'}")
    assert (tmp_path / "b.restart").read_bytes() == path.read_bytes()
    s.command(f"write_data {tmp_path / 's.data'}")
    r.command(f"write_data {tmp_path / 'r.data'}")
    assert (tmp_path / "s.data").read_text() == (tmp_path / "r.data").read_text()

    # the layout, read here on its own
    raw = path.read_bytes()
    assert raw[:8] == b"MDPD355R"
    assert struct.unpack_from("<iii", raw, 8) == (1, 0x01020304, 0)   # version, byte order, mdpd/atomic/meso
    assert struct.unpack_from("<dqd", raw, 20) == (7.5, 300, 0.02)
    assert struct.unpack_from("<i6d", raw, 44) == (2, -1.0, 0.0, 0.0, 4.0, 5.0, 6.0)
    assert struct.unpack_from("<idid", raw, 96) == (1, 1.0, 1, 2.5)
    off = 120
    groups = []
    (ng,) = struct.unpack_from("<i", raw, off)
    off += 4
    for _ in range(ng):
        bit, ln = struct.unpack_from("<ii", raw, off)
        groups.append((raw[off + 8:off + 8 + ln].decode(), bit))
        off += 8 + ln
    assert groups == [("all", 1), ("fluid", 2), ("wall", 4)]
    assert struct.unpack_from("<iddiii", raw, off) == (0, 1.5, 1.0, 777, 2, 4)
    off += 4 + 16 + 12
    coeffs = []
    for _ in range(3):                          # (1,1) (1,2) (2,2)
        (flag,) = struct.unpack_from("<i", raw, off)
        off += 4
        coeffs.append(struct.unpack_from("<7d", raw, off) if flag else None)
        off += 56 if flag else 0
    assert coeffs[0][:3] == (-40.0, 25.0, 18.0) and coeffs[0][5:] == (1.0, 0.75)
    assert coeffs[1][0] == -20.0 and coeffs[2] is None
    (natoms,) = struct.unpack_from("<q", raw, off)
    off += 8
    assert natoms == 30 and len(raw) == off + 30 * 13 * 8
    rec = np.frombuffer(raw, np.float64, offset=off).reshape(30, 13)
    np.testing.assert_array_equal(rec[:, 0], 13)
    np.testing.assert_array_equal(rec[:, 1:4], x)
    np.testing.assert_array_equal(rec[:, 4:7], v)
    np.testing.assert_array_equal(rec[:, 7:9], 0)                   # rho, phi: not set up yet
    np.testing.assert_array_equal(rec[:, 9], np.arange(1, 31))
    np.testing.assert_array_equal(rec[:, 10], [1, 2] * 15)
    np.testing.assert_array_equal(rec[:, 11], [1 | 2, 1 | 4] * 15)
    np.testing.assert_array_equal(rec[:, 12], 512 | 512 << 10 | 512 << 20)   # image (0, 0, 0)


def test_restart_errors(engine, tmp_path):
    s, _, _ = _restart_source(engine)
    path = tmp_path / "a.restart"
    s.command(f"write_restart {path}")
    with pytest.raises(engine.MdpdError) as e:
        s.command(f"read_restart {path}")                  # the box is already defined
    assert e.value.code == 2
    raw = path.read_bytes()
    cases = {
        "missing.restart": None,
        "short.restart": raw[:-8],
        "trailing.restart": raw + b"\0",
        "magic.restart": b"LAMMPS R" + raw[8:],
        "type.restart": raw[:len(raw) - 30 * 104 + 80] + np.float64(3.0).tobytes() + raw[len(raw) - 30 * 104 + 88:],
        "tag.restart": raw[:len(raw) - 30 * 104 + 72] + np.float64(0.5).tobytes() + raw[len(raw) - 30 * 104 + 80:],
        # an atom count far beyond the records that follow it
        "count.restart": raw[:len(raw) - 30 * 104 - 8] + np.int64(1 << 30).tobytes() + raw[len(raw) - 30 * 104:],
    }
    for name, data in cases.items():
        p = tmp_path / name
        if data is not None:
            p.write_bytes(data)
        t = engine.System()
        with pytest.raises(engine.MdpdError) as e:
            t.command(f"read_restart {p}")
        assert e.value.code == 1, name
        assert t.natoms == 0, name                          # nothing installed from a bad file
    t = engine.System()
    with pytest.raises(engine.MdpdError) as e:
        t.command(f"write_restart {tmp_path / 'x.restart'}")
    assert e.value.code == 2


def test_system_without_atoms_is_refused_before_the_device(engine):
    # a complete script with no atoms: setup and pool submission stop at
    # validation with a state error, before any device work
    s = engine.System()
    _deck(s)
    s.commands("""
        mass 1 1.0
        pair_style mdpd/meso 1.0 1.0 5
        pair_coeff * * -40 25 4.5 1.0 0.75
        fix int all nve/meso
        run_style mvv/meso
    """)
    assert s.natoms == 0
    with pytest.raises(engine.MdpdError) as e:
        s.setup()
    assert e.value.code == 2 and "no atoms" in str(e.value)
    with pytest.raises(engine.MdpdError) as e:
        s.run(5)
    assert e.value.code == 2
    p = engine.Pool(64e6, 4)
    with pytest.raises(engine.MdpdError) as e:
        p.submit(s, 10)
    assert e.value.code == 2 and "no atoms" in str(e.value)
    assert p.stat("queued") == 0
