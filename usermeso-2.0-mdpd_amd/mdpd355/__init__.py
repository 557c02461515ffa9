"""mdpd355: Python host mirror of the engine's C ABI (include/mdpd355.h).

The engine is driven the way USER-MESO is driven through the LAMMPS library
interface (src/library.h: lammps_open / lammps_command / lammps_gather_atoms):

    import mdpd355
    s = mdpd355.System()
    s.commands('''
        units lj
        atom_style mdpd/atomic/meso 6.05
        ''')
    s.create_box(1, (0, 0, 0), (10, 10, 10))
    s.create_atoms(types, x, v)
    s.commands('''
        mass 1 1.0
        pair_style mdpd/meso 1.0 1.0 123456
        pair_coeff * * -40 25 4.5 1.0 0.75
        fix rho all aboundary/rho/meso cut_rho 0.75
        fix int all nve/meso
        timestep 0.01
        run_style mvv/meso
        ''')
    s.run(100)
    x = s.gather("x")

Errors raised by the engine become MdpdError with the engine's message.  The
shared library is required: importing this module on a machine without the
built libmdpd355.so raises immediately (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

__all__ = ["System", "Pool", "MdpdError", "tea_gaussian", "tea_premix", "clock", "library_path", "load"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBNAME = "libmdpd355.so"
_lib = None

_STATUS = {
    0: "MDPD_OK", 1: "MDPD_E_ARG", 2: "MDPD_E_STATE", 3: "MDPD_E_HIP",
    4: "MDPD_E_OVERFLOW", 5: "MDPD_E_NOMEM", 6: "MDPD_E_UNSUPPORTED",
}


class MdpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


def library_path():
    return os.path.join(_HERE, _LIBNAME)


def load():
    """Load libmdpd355.so (built in-tree by build.py); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build the engine first (python usermeso-2.0-mdpd_amd/build.py)")
    lib = ctypes.CDLL(path)
    P, I, D, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    sig = {
        "mdpd_open": (I, [I, ctypes.POINTER(P)]),
        "mdpd_close": (None, [P]),
        "mdpd_last_error": (ctypes.c_char_p, [P]),
        "mdpd_version": (ctypes.c_char_p, []),
        "mdpd_command": (I, [P, ctypes.c_char_p]),
        "mdpd_commands_string": (I, [P, ctypes.c_char_p]),
        "mdpd_create_box": (I, [P, I, ctypes.POINTER(D), ctypes.POINTER(D)]),
        "mdpd_create_atoms": (I, [P, I, P, P, P, P, P]),
        "mdpd_clear_atoms": (I, [P]),
        "mdpd_get_natoms": (I, [P]),
        "mdpd_group_bit": (I, [P, ctypes.c_char_p]),
        "mdpd_setup": (I, [P]),
        "mdpd_run": (I, [P, LL]),
        "mdpd_get_timestep": (LL, [P]),
        "mdpd_get_nbuild": (LL, [P]),
        "mdpd_gather_atoms": (I, [P, ctypes.c_char_p, I, I, P]),
        "mdpd_scatter_atoms": (I, [P, ctypes.c_char_p, I, I, P]),
        "mdpd_thermo": (I, [P, ctypes.POINTER(D)]),
        "mdpd_tune": (I, [P, ctypes.c_char_p, I]),
        "mdpd_fix_result": (I, [P, ctypes.c_char_p, I, P, P, ctypes.POINTER(I)]),
        "mdpd_profile": (I, [P, I]),
        "mdpd_profile_get": (I, [P, ctypes.c_char_p, ctypes.POINTER(D), ctypes.POINTER(LL)]),
        "mdpd_debug_sorted": (I, [P, P, P, P]),
        "mdpd_debug_neighbors": (I, [P, ctypes.POINTER(I), P, P]),
        "mdpd_debug_signatures": (I, [P, P]),
        "mdpd_tea_gaussian": (I, [I, I, P, P, P]),
        "mdpd_tea_premix": (I, [I, I, P, P, I, P]),
        "mdpd_pool_open": (I, [I, D, I, ctypes.POINTER(P)]),
        "mdpd_pool_close": (None, [P]),
        "mdpd_pool_last_error": (ctypes.c_char_p, [P]),
        "mdpd_pool_submit": (I, [P, P, LL, ctypes.POINTER(I)]),
        "mdpd_pool_advance": (I, [P, LL, I, ctypes.POINTER(I)]),
        "mdpd_pool_run": (I, [P]),
        "mdpd_pool_admit": (I, [P, ctypes.POINTER(I)]),
        "mdpd_pool_job_state": (I, [P, I, ctypes.POINTER(I)]),
        "mdpd_pool_done": (I, [P, P, I, ctypes.POINTER(I)]),
        "mdpd_pool_gather": (I, [P, I, ctypes.c_char_p, I, I, P]),
        "mdpd_pool_thermo": (I, [P, I, ctypes.POINTER(D)]),
        "mdpd_pool_fix_result": (I, [P, I, ctypes.c_char_p, I, P, P, ctypes.POINTER(I)]),
        "mdpd_pool_fix_windows": (I, [P, I, ctypes.c_char_p, I, P, P, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "mdpd_pool_job_info": (I, [P, I, ctypes.POINTER(LL), ctypes.POINTER(LL), ctypes.POINTER(D)]),
        "mdpd_pool_release": (I, [P, I]),
        "mdpd_pool_tune": (I, [P, ctypes.c_char_p, I]),
        "mdpd_pool_stat": (I, [P, ctypes.c_char_p, ctypes.POINTER(D)]),
        "mdpd_clock": (D, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


_FIELDS = {  # name -> (type, count, dtype)
    "x": (1, 3, np.float64), "v": (1, 3, np.float64), "f": (1, 3, np.float64),
    "rho": (1, 1, np.float64), "phi": (1, 1, np.float64), "e_pair": (1, 1, np.float64),
    "virial": (1, 6, np.float64),
    "tag": (0, 1, np.int32), "type": (0, 1, np.int32), "mask": (0, 1, np.int32), "image": (0, 3, np.int32),
}


class System:
    """One engine handle (the counterpart of one LAMMPS instance running the
    USER-MESO mDPD/DPD styles)."""

    def __init__(self, device=0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.mdpd_open(int(device), ctypes.byref(h))
        if rc != 0:
            raise MdpdError(rc, f"mdpd_open(device={device}) failed")
        self._h = h

    # -- plumbing --
    def _check(self, rc):
        if rc != 0:
            raise MdpdError(rc, self._lib.mdpd_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mdpd_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- script surface --
    def command(self, line):
        self._check(self._lib.mdpd_command(self._h, line.encode()))

    def commands(self, text):
        self._check(self._lib.mdpd_commands_string(self._h, text.encode()))

    def create_box(self, ntypes, lo, hi):
        lo = (ctypes.c_double * 3)(*lo)
        hi = (ctypes.c_double * 3)(*hi)
        self._check(self._lib.mdpd_create_box(self._h, int(ntypes), lo, hi))

    def create_atoms(self, types, x, v=None, tag=None, mask=None):
        types = np.ascontiguousarray(types, dtype=np.int32)
        n = types.shape[0]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        tag = None if tag is None else np.ascontiguousarray(tag, dtype=np.int32)
        mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        self._check(self._lib.mdpd_create_atoms(self._h, n, _ptr(tag), _ptr(types), _ptr(mask), _ptr(x), _ptr(v)))

    def clear_atoms(self):
        self._check(self._lib.mdpd_clear_atoms(self._h))

    @property
    def natoms(self):
        return self._lib.mdpd_get_natoms(self._h)

    def group_bit(self, name):
        return self._lib.mdpd_group_bit(self._h, name.encode())

    # -- run --
    def setup(self):
        self._check(self._lib.mdpd_setup(self._h))

    def run(self, nsteps):
        self._check(self._lib.mdpd_run(self._h, int(nsteps)))

    @property
    def timestep(self):
        return self._lib.mdpd_get_timestep(self._h)

    @property
    def nbuild(self):
        return self._lib.mdpd_get_nbuild(self._h)

    # -- results --
    def gather(self, name):
        t, c, dt = _FIELDS[name]
        out = np.empty((self.natoms, c), dtype=dt)
        self._check(self._lib.mdpd_gather_atoms(self._h, name.encode(), t, c, _ptr(out)))
        return out[:, 0].copy() if c == 1 else out

    def scatter(self, name, data):
        data = np.ascontiguousarray(data, dtype=np.float64).reshape(self.natoms, 3)
        self._check(self._lib.mdpd_scatter_atoms(self._h, name.encode(), 1, 3, _ptr(data)))

    def thermo(self):
        out = (ctypes.c_double * 4)()
        self._check(self._lib.mdpd_thermo(self._h, out))
        return {"temp": out[0], "press": out[1], "pe": out[2], "ke": out[3]}

    def fix_result(self, fix_id):
        """(bin centres, values) of an observable fix (rdf/fast/meso, density/meso)."""
        nb = ctypes.c_int()
        self._check(self._lib.mdpd_fix_result(self._h, fix_id.encode(), 0, None, None, ctypes.byref(nb)))
        x = np.empty(nb.value)
        v = np.empty(nb.value)
        self._check(self._lib.mdpd_fix_result(self._h, fix_id.encode(), nb.value, _ptr(x), _ptr(v), ctypes.byref(nb)))
        return x, v

    # -- tuning / kernel timing --
    def tune(self, knob, value):
        self._check(self._lib.mdpd_tune(self._h, knob.encode(), int(value)))

    def profile(self, enable=True):
        self._check(self._lib.mdpd_profile(self._h, int(bool(enable))))

    def profile_get(self, phase):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(self._lib.mdpd_profile_get(self._h, phase.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- parity hooks --
    def debug_sorted(self):
        n = self.natoms
        origin = np.empty(n, dtype=np.int32)
        key = np.empty(n, dtype=np.uint64)
        xw = np.empty((n, 3), dtype=np.float32)
        self._check(self._lib.mdpd_debug_sorted(self._h, _ptr(origin), _ptr(key), _ptr(xw)))
        return origin, key, xw

    def debug_neighbors(self):
        n = self.natoms
        m = ctypes.c_int(0)
        self._check(self._lib.mdpd_debug_neighbors(self._h, ctypes.byref(m), None, None))
        counts = np.empty(n, dtype=np.int32)
        rows = np.empty((n, m.value), dtype=np.int32)
        self._check(self._lib.mdpd_debug_neighbors(self._h, ctypes.byref(m), _ptr(counts), _ptr(rows)))
        return counts, rows

    def debug_signatures(self):
        out = np.empty(self.natoms, dtype=np.uint32)
        self._check(self._lib.mdpd_debug_signatures(self._h, _ptr(out)))
        return out


JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2


class Pool:
    """Pooled engine: many independent systems advanced together, one launch
    sequence per pass (batched runs and served jobs).  Jobs are submitted from
    configured System handles, whose state and atoms are copied."""

    def __init__(self, hbm_budget_bytes, max_systems, device=0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.mdpd_pool_open(int(device), float(hbm_budget_bytes), int(max_systems), ctypes.byref(h))
        if rc != 0:
            raise MdpdError(rc, "mdpd_pool_open failed")
        self._h = h
        self._natoms = {}

    def _check(self, rc):
        if rc != 0:
            raise MdpdError(rc, self._lib.mdpd_pool_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mdpd_pool_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, system, nsteps):
        j = ctypes.c_int()
        self._check(self._lib.mdpd_pool_submit(self._h, system._h, int(nsteps), ctypes.byref(j)))
        self._natoms[j.value] = system.natoms
        return j.value

    def advance(self, max_passes=0, stop_on_done=False):
        d = ctypes.c_int()
        self._check(self._lib.mdpd_pool_advance(self._h, int(max_passes), int(bool(stop_on_done)), ctypes.byref(d)))
        return d.value

    def run(self):
        self._check(self._lib.mdpd_pool_run(self._h))

    def admit(self):
        a = ctypes.c_int()
        self._check(self._lib.mdpd_pool_admit(self._h, ctypes.byref(a)))
        return a.value

    def state(self, job):
        st = ctypes.c_int()
        self._check(self._lib.mdpd_pool_job_state(self._h, int(job), ctypes.byref(st)))
        return st.value

    def done(self, max_jobs=4096):
        ids = np.empty(max_jobs, dtype=np.int32)
        c = ctypes.c_int()
        self._check(self._lib.mdpd_pool_done(self._h, _ptr(ids), int(max_jobs), ctypes.byref(c)))
        return ids[: c.value].tolist()

    def gather(self, job, name):
        t, c, dt = _FIELDS[name]
        out = np.empty((self._natoms[job], c), dtype=dt)
        self._check(self._lib.mdpd_pool_gather(self._h, int(job), name.encode(), t, c, _ptr(out)))
        return out[:, 0].copy() if c == 1 else out

    def thermo(self, job):
        out = (ctypes.c_double * 4)()
        self._check(self._lib.mdpd_pool_thermo(self._h, int(job), out))
        return {"temp": out[0], "press": out[1], "pe": out[2], "ke": out[3]}

    def fix_result(self, job, fix_id):
        """(bin centres, values) of a finished job's rdf/fast/meso fix (g(r)) or
        density/meso fix (the profile of its last window)."""
        nb = ctypes.c_int()
        self._check(self._lib.mdpd_pool_fix_result(self._h, int(job), fix_id.encode(), 0, None, None, ctypes.byref(nb)))
        x = np.empty(nb.value)
        v = np.empty(nb.value)
        self._check(self._lib.mdpd_pool_fix_result(self._h, int(job), fix_id.encode(), nb.value, _ptr(x), _ptr(v),
                                                   ctypes.byref(nb)))
        return x, v

    def fix_windows(self, job, fix_id):
        """(centre steps, profiles [window, bin]) of a finished job's density/meso fix."""
        nw, nb = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.mdpd_pool_fix_windows(self._h, int(job), fix_id.encode(), 0, None, None,
                                                    ctypes.byref(nw), ctypes.byref(nb)))
        c = np.zeros(nw.value, np.int64)
        v = np.zeros((nw.value, nb.value))
        if nw.value:
            self._check(self._lib.mdpd_pool_fix_windows(self._h, int(job), fix_id.encode(), nw.value, _ptr(c), _ptr(v),
                                                        ctypes.byref(nw), ctypes.byref(nb)))
        return c, v

    def info(self, job):
        n = ctypes.c_longlong()
        ts = ctypes.c_longlong()
        t = (ctypes.c_double * 3)()
        self._check(self._lib.mdpd_pool_job_info(self._h, int(job), ctypes.byref(n), ctypes.byref(ts), t))
        return {"natoms": n.value, "timestep": ts.value, "t_submit": t[0], "t_admit": t[1], "t_done": t[2]}

    def release(self, job):
        self._check(self._lib.mdpd_pool_release(self._h, int(job)))
        self._natoms.pop(job, None)

    def tune(self, knob, value):
        self._check(self._lib.mdpd_pool_tune(self._h, knob.encode(), int(value)))

    def stat(self, name):
        v = ctypes.c_double()
        self._check(self._lib.mdpd_pool_stat(self._h, name.encode(), ctypes.byref(v)))
        return v.value


def clock():
    """The engine's monotonic clock (seconds), the time base of job times."""
    return load().mdpd_clock()


def tea_gaussian(a, b, device=0):
    """FP32 pair Gaussians gaussian_TEA<4>(a > b, a, b) evaluated on the device."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.empty(a.shape[0], dtype=np.float32)
    rc = lib.mdpd_tea_gaussian(int(device), a.shape[0], _ptr(a), _ptr(b), _ptr(out))
    if rc != 0:
        raise MdpdError(rc, "mdpd_tea_gaussian failed")
    return out


def tea_premix(a, b, rounds, device=0):
    """premix_TEA<rounds>(a, b) evaluated on the device."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.empty(a.shape[0], dtype=np.uint32)
    rc = lib.mdpd_tea_premix(int(device), a.shape[0], _ptr(a), _ptr(b), int(rounds), _ptr(out))
    if rc != 0:
        raise MdpdError(rc, "mdpd_tea_premix failed")
    return out
