"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The engine never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
MAXT = 8

_lib = None


def build():
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _declare(_lib)
    return _lib


class Params(ctypes.Structure):
    _fields_ = [("ntypes", ctypes.c_int)] + [
        (nm, (ctypes.c_double * MAXT) * MAXT) for nm in ("cut", "cut_r", "A", "B", "gamma", "sigma", "expw")
    ] + [("mass", ctypes.c_double * MAXT), ("dt", ctypes.c_double),
         ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3)]


class RunCfg(ctypes.Structure):
    _fields_ = [("style", ctypes.c_int), ("seed", ctypes.c_uint32), ("ntimestep", ctypes.c_int64),
                ("skin", ctypes.c_double), ("cut_rho", ctypes.c_double), ("lucy", ctypes.c_int),
                ("rho_groupbit", ctypes.c_int), ("nve_groupbit", ctypes.c_int),
                ("addforce_groupbit", ctypes.c_int), ("addforce", ctypes.c_double * 3),
                ("nthreads", ctypes.c_int), ("arb", ctypes.c_void_p), ("phi", ctypes.c_void_p)]


class Arb(ctypes.Structure):
    _fields_ = [("pair_on", ctypes.c_int), ("pair_mobile_bit", ctypes.c_int), ("pair_wall_bit", ctypes.c_int),
                ("fix_on", ctypes.c_int), ("fix_mobile_bit", ctypes.c_int), ("fix_wall_bit", ctypes.c_int),
                ("cut_phi", ctypes.c_double), ("phi_c", ctypes.c_double), ("rho_wall", ctypes.c_double)]


class RanMars(ctypes.Structure):
    _fields_ = [("u", ctypes.c_double * 98), ("c", ctypes.c_double), ("cd", ctypes.c_double),
                ("cm", ctypes.c_double), ("i97", ctypes.c_int), ("j97", ctypes.c_int),
                ("save", ctypes.c_int), ("second", ctypes.c_double)]


def _declare(L):
    P = ctypes.c_void_p
    U32, I, D = ctypes.c_uint32, ctypes.c_int, ctypes.c_double
    sig = {
        "orc_tea_key": (None, [P, P, I, P]),
        "orc_premix": (U32, [U32, U32, I]),
        "orc_brev32": (U32, [U32]),
        "orc_interleave3": (ctypes.c_uint64, [U32, U32, U32]),
        "orc_mantissa": (U32, [ctypes.c_float, ctypes.c_float, ctypes.c_float]),
        "orc_seed_now": (U32, [U32, ctypes.c_int64]),
        "orc_sig_mdpd": (U32, [ctypes.c_int32, U32]),
        "orc_sig_dpd": (U32, [ctypes.c_int32, U32, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
        "orc_pair_bits": (None, [U32, U32, P, P]),
        "orc_gaussian_poly": (D, [U32, U32]),
        "orc_gaussian_libm": (D, [U32, U32]),
        "orc_cospi_poly": (D, [D]),
        "orc_log2u_poly": (D, [U32]),
        "orc_stable_sort_u64": (None, [I, P, P]),
        "orc_wrap_fp32": (None, [I, P, P, P]),
        "orc_neighbors_fp32": (I, [I, P, P, P, D, I, P, P]),
        "orc_density": (None, [I, P, P, P, P, D, I, I, P]),
        "orc_force_mdpd": (None, [I, P, P, P, P, P, P, P, P, P]),
        "orc_force_dpd": (None, [I, P, P, P, P, P, P, P, P]),
        "orc_wall_ratio": (D, [D, D]),
        "orc_aboundary": (None, [I, P, P, P, P, P, P, P, P]),
        "orc_force_mdpd_arb": (None, [I, P, P, P, P, P, P, P, P, P, P, P, P]),
        "orc_run": (I, [I, P, P, P, P, P, P, P, P, P, I]),
        "orc_ranmars_init": (None, [P, I]),
        "orc_ranmars_uniform": (D, [P]),
        "orc_ranmars_gaussian": (D, [P]),
        "orc_lammps_dpd_run": (I, [I, P, P, P, P, D, D, D, D, D, D, I, I, I, I, P, P, P, I, P, I]),
        "orc_temperature": (D, [I, P, P, P]),
        "orc_pressure": (D, [I, P, P, P, P]),
        "orc_premix_n": (None, [I, P, P, I, P]),
        "orc_gaussian_n": (None, [I, P, P, I, P]),
        "orc_sig_mdpd_n": (None, [I, P, U32, P]),
        "orc_sig_dpd_n": (None, [I, P, U32, P, P]),
        "orc_sort_keys_n": (None, [I, P, P, P, D, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _d3(v):
    return np.ascontiguousarray(v, dtype=np.float64).reshape(3)


# ---------------------------------------------------------------- integers

def tea_key(v0, v1, rounds, key):
    a, b = ctypes.c_uint32(v0), ctypes.c_uint32(v1)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    lib().orc_tea_key(ctypes.byref(a), ctypes.byref(b), rounds, _p(k))
    return a.value, b.value


def premix(a, b, rounds):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.empty_like(a)
    lib().orc_premix_n(a.size, _p(a), _p(b), rounds, _p(out))
    return out


def seed_now(seed, step):
    return lib().orc_seed_now(seed & 0xFFFFFFFF, step)


def sig_mdpd(tag, seed_t):
    tag = np.ascontiguousarray(tag, dtype=np.int32)
    out = np.empty(tag.size, dtype=np.uint32)
    lib().orc_sig_mdpd_n(tag.size, _p(tag), seed_t & 0xFFFFFFFF, _p(out))
    return out


def sig_dpd(tag, seed_t, v):
    tag = np.ascontiguousarray(tag, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    out = np.empty(tag.size, dtype=np.uint32)
    lib().orc_sig_dpd_n(tag.size, _p(tag), seed_t & 0xFFFFFFFF, _p(v), _p(out))
    return out


def gaussian(a, b, libm=False):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.empty(a.size, dtype=np.float64)
    lib().orc_gaussian_n(a.size, _p(a), _p(b), int(libm), _p(out))
    return out


def sort_keys(x, lo, hi, cutneigh):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    out = np.empty(x.shape[0], dtype=np.uint64)
    lib().orc_sort_keys_n(x.shape[0], _p(x), _p(_d3(lo)), _p(_d3(hi)), float(cutneigh), _p(out))
    return out


def stable_sort(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.empty(keys.size, dtype=np.int32)
    lib().orc_stable_sort_u64(keys.size, _p(keys), _p(perm))
    return perm


def wrap(x, lo, hi):
    x = np.array(x, dtype=np.float32).reshape(-1, 3)
    lib().orc_wrap_fp32(x.shape[0], _p(x), _p(_d3(lo)), _p(_d3(hi)))
    return x


def neighbors(x, lo, hi, cut, maxnbr=256):
    """FP32 neighbor sets: returns (counts, list of sorted entry arrays)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    counts = np.empty(n, dtype=np.int32)
    rows = np.full((n, maxnbr), -1, dtype=np.int32)
    mx = lib().orc_neighbors_fp32(n, _p(x), _p(_d3(lo)), _p(_d3(hi)), float(cut), maxnbr, _p(counts), _p(rows))
    if mx > maxnbr:
        return neighbors(x, lo, hi, cut, maxnbr=mx)
    return counts, [rows[i, :counts[i]].copy() for i in range(n)]


# ---------------------------------------------------------------- physics

def params(lo, hi, dt, mass=(1.0,), ntypes=1, **pair):
    """pair: dict of per-pair matrices (ntypes x ntypes) or scalars for cut, cut_r, A, B, gamma, sigma, expw."""
    p = Params()
    p.ntypes = ntypes
    for nm in ("cut", "cut_r", "A", "B", "gamma", "sigma", "expw"):
        val = pair.get(nm, 0.0)
        m = np.broadcast_to(np.asarray(val, dtype=np.float64), (ntypes, ntypes))
        arr = getattr(p, nm)
        for a in range(ntypes):
            for b in range(ntypes):
                arr[a][b] = float(m[a, b])
    for t in range(ntypes):
        p.mass[t] = float(mass[t] if len(mass) > t else mass[0])
    p.dt = float(dt)
    for d in range(3):
        p.lo[d] = float(lo[d])
        p.hi[d] = float(hi[d])
    return p


def density(x, types0, mask, p, rd, lucy=False, groupbit=1):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    t = np.ascontiguousarray(types0, dtype=np.int32)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    rho = np.empty(n, dtype=np.float64)
    lib().orc_density(n, _p(x), _p(t), _p(m), ctypes.byref(p), float(rd), int(lucy), int(groupbit), _p(rho))
    return rho


def force_mdpd(x, v, rho, sig, types0, p):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    f = np.empty((n, 3)); e = np.empty(n); vir = np.empty((n, 6))
    lib().orc_force_mdpd(n, _p(x), _p(v), _p(rho), _p(sig), _p(t), ctypes.byref(p), _p(f), _p(e), _p(vir))
    return f, e, vir


def force_dpd(x, v, sig, types0, p):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    f = np.empty((n, 3)); e = np.empty(n); vir = np.empty((n, 6))
    lib().orc_force_dpd(n, _p(x), _p(v), _p(sig), _p(t), ctypes.byref(p), _p(f), _p(e), _p(vir))
    return f, e, vir


def temperature(v, types0, p):
    """compute temp/meso of FP64 velocities (dof = 3N - 3)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    return lib().orc_temperature(v.shape[0], _p(v), _p(t), ctypes.byref(p))


def pressure(v, vir, types0, p):
    """compute pressure/meso from velocities and per-atom virials (n x 6)."""
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(vir, dtype=np.float64).reshape(-1, 6)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    return lib().orc_pressure(v.shape[0], _p(v), _p(w), _p(t), ctypes.byref(p))


def arb(mobile_bit, wall_bit, cut_phi=1.0, phi_c=0.5, rho_wall=8.0, pair=True, fix=True):
    """Arbitrary-boundary settings: the same mobile / wall groups for the pair
    style's wall scaling and for the fix's wall fraction + bounce-back."""
    a = Arb()
    a.pair_on, a.pair_mobile_bit, a.pair_wall_bit = int(pair), mobile_bit, wall_bit
    a.fix_on, a.fix_mobile_bit, a.fix_wall_bit = int(fix), mobile_bit, wall_bit
    a.cut_phi, a.phi_c, a.rho_wall = cut_phi, phi_c, rho_wall
    return a


def wall_ratio(phi, rcw=1.0):
    return lib().orc_wall_ratio(float(phi), float(rcw))


def aboundary(x, v, types0, mask, p, a):
    """wall fraction + bounce-back of the mobile group; returns (x, v, phi, nw)."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    t = np.ascontiguousarray(types0, dtype=np.int32)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    phi = np.empty(n); nw = np.empty((n, 3))
    lib().orc_aboundary(n, _p(x), _p(v), _p(t), _p(m), ctypes.byref(p), ctypes.byref(a), _p(phi), _p(nw))
    return x, v, phi, nw


def force_mdpd_arb(x, v, rho, sig, types0, mask, phi, p, a):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    f = np.empty((n, 3)); e = np.empty(n); vir = np.empty((n, 6))
    lib().orc_force_mdpd_arb(n, _p(x), _p(v), _p(rho), _p(sig), _p(t), _p(m), _p(phi), ctypes.byref(p),
                             ctypes.byref(a), _p(f), _p(e), _p(vir))
    return f, e, vir


def run(x, v, tag, types0, mask, p, nsteps, style=0, seed=1, ntimestep=0, skin=0.3, cut_rho=0.75,
        lucy=False, rho_groupbit=1, nve_groupbit=1, addforce_groupbit=0, addforce=(0, 0, 0), nthreads=0,
        arb=None, phi=None):
    """FP64 setup + run; returns (x, v, f, rho, nbuild).  With arb (see arb()),
    the wall branch runs each step; phi (n, in/out) carries the wall fractions."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    f = np.empty((n, 3)); rho = np.zeros(n)
    cfg = RunCfg()
    cfg.style, cfg.seed, cfg.ntimestep = style, seed & 0xFFFFFFFF, ntimestep
    cfg.skin, cfg.cut_rho, cfg.lucy = skin, cut_rho, int(lucy)
    cfg.rho_groupbit, cfg.nve_groupbit, cfg.addforce_groupbit = rho_groupbit, nve_groupbit, addforce_groupbit
    for d in range(3):
        cfg.addforce[d] = addforce[d]
    cfg.nthreads = nthreads
    if arb is not None:
        if phi is None or phi.dtype != np.float64 or phi.shape != (n,):
            raise ValueError("arb needs phi: float64 array of n")
        cfg.arb = ctypes.cast(ctypes.pointer(arb), ctypes.c_void_p)
        cfg.phi = phi.ctypes.data_as(ctypes.c_void_p)
    tag = np.ascontiguousarray(tag, dtype=np.int32)
    t = np.ascontiguousarray(types0, dtype=np.int32)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    nb = lib().orc_run(n, _p(x), _p(v), _p(f), _p(rho), _p(tag), _p(t), _p(m), ctypes.byref(p),
                       ctypes.byref(cfg), int(nsteps))
    return x, v, f, rho, nb


def ranmars(seed):
    r = RanMars()
    lib().orc_ranmars_init(ctypes.byref(r), int(seed))
    return r


def ranmars_uniform(r, k=1):
    L = lib()
    return np.array([L.orc_ranmars_uniform(ctypes.byref(r)) for _ in range(k)])


def ranmars_gaussian(r, k=1):
    L = lib()
    return np.array([L.orc_ranmars_gaussian(ctypes.byref(r)) for _ in range(k)])


def lammps_dpd_run(x, v, lo, hi, a0=25.0, gamma=4.5, temperature=1.0, rc=1.0, skin=0.3, dt=0.04,
                   seed=12345, nsteps=100, equil=0, sample_every=10, nrdf=50, nthreads=0):
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    ts, ps = ctypes.c_double(), ctypes.c_double()
    ns = ctypes.c_int()
    hist = np.zeros(nrdf)
    nb = lib().orc_lammps_dpd_run(n, _p(x), _p(v), _p(_d3(lo)), _p(_d3(hi)), a0, gamma, temperature, rc,
                                  skin, dt, int(seed), int(nsteps), int(equil), int(sample_every),
                                  ctypes.byref(ts), ctypes.byref(ps), ctypes.byref(ns), nrdf, _p(hist),
                                  int(nthreads))
    k = max(ns.value, 1)
    return {"x": x, "v": v, "T": ts.value / k, "P": ps.value / k, "nsamples": ns.value,
            "rdf_hist": hist, "nbuild": nb}
