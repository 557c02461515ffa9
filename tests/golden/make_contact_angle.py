"""Write tests/golden/contact_angle.npz from the reference's static
contact-angle example (examples/v2.5/static_contact_angle_in_a_slit_nano_channel):
the initial state LAMMPS_init_noimage.dat (9,280 atoms: fluid type 1, wall
type 2) and the fluid density profiles along y (10 bins) that the reference
printed for A_12 = -20, -30, -35 -- the GPU runs (out.profile.a_*) and the
CPU runs (CPU_simulations_to_match/out.cpu.a*).  Data only: no reference code
is run or copied.

Usage: python tests/golden/make_contact_angle.py [reference root]
"""
import os
import sys

import numpy as np

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
EX = os.path.join(ROOT, "examples/v2.5/static_contact_angle_in_a_slit_nano_channel")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "contact_angle.npz")


def read_atoms(path):
    lines = open(path).read().splitlines()
    box = {}
    for l in lines[:15]:
        f = l.split()
        if len(f) == 4 and f[2] in ("xlo", "ylo", "zlo"):
            box[f[2][0]] = (float(f[0]), float(f[1]))
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("Atoms"))
    rows = []
    for l in lines[k + 1:]:
        f = l.split()
        if not f:
            if rows:
                break
            continue
        rows.append(f)
    ids = np.array([int(r[0]) for r in rows], dtype=np.int32)
    types = np.array([int(r[1]) for r in rows], dtype=np.int32)
    x = np.array([[float(c) for c in r[2:5]] for r in rows])
    lo = np.array([box[a][0] for a in "xyz"])
    hi = np.array([box[a][1] for a in "xyz"])
    return ids, types, x, lo, hi


def read_gpu_profile(path):
    a = np.loadtxt(path)
    return a[:, 1]                      # counts per bin, bins in order along y


def read_cpu_profile(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    rows = [r for r in rows if len(r) == 4]
    return np.array([float(r[2]) for r in rows])


def main():
    ids, types, x, lo, hi = read_atoms(os.path.join(EX, "LAMMPS_init_noimage.dat"))
    out = dict(ids=ids, types=types, x=x, lo=lo, hi=hi)
    for a in (20, 30, 35):
        out[f"gpu_a{a}"] = read_gpu_profile(os.path.join(EX, f"out.profile.a_-{a}"))
        out[f"cpu_a{a}"] = read_cpu_profile(os.path.join(EX, "CPU_simulations_to_match", f"out.cpu.a-{a}"))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(ids)} atoms, box {lo} .. {hi}")
    for a in (20, 30, 35):
        print(a, out[f"gpu_a{a}"], out[f"cpu_a{a}"])


if __name__ == "__main__":
    main()
