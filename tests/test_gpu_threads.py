"""Independent contexts on separate host threads.

Each handle and each pool owns its stream, graphs and device memory, and one
host thread drives it; contexts driven from different threads at the same
time give the results of the same work done one after another.
"""
import threading

import numpy as np
import pytest

from conftest import uniform_box
from test_gpu_parity import mdpd_system

pytestmark = pytest.mark.gpu

GB = float(1 << 30)
CASES = [(1296, 6.0, 41, 40), (2400, 8.0, 42, 30)]     # n, box edge, seed, steps
POOL_CASE = (1728, 7.0, 43, 25)


def _handle_run(engine, case):
    n, L, seed, steps = case
    x, v = uniform_box(n, L, seed)
    s = mdpd_system(engine, x, v, L)
    s.run(steps)
    out = (s.gather("x"), s.gather("v"), s.thermo()["press"])
    s.close()
    return out


def _pool_run(engine, case):
    n, L, seed, steps = case
    x, v = uniform_box(n, L, seed)
    p = engine.Pool(0.25 * GB, 4)
    p.tune("split", 2)                  # no timing-based choice: it would follow the other threads' load
    j = p.submit(mdpd_system(engine, x, v, L), steps)
    p.run()
    out = (p.gather(j, "x"), p.gather(j, "v"), p.thermo(j)["press"])
    p.close()
    return out


def test_contexts_on_separate_threads_equal_sequential_runs(engine):
    work = [lambda c=c: _handle_run(engine, c) for c in CASES] + [lambda: _pool_run(engine, POOL_CASE)]
    want = [w() for w in work]
    got, errors = [None] * len(work), []

    def drive(k):
        try:
            got[k] = work[k]()
        except Exception as e:          # reported below with the thread's case
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=drive, args=(k,)) for k in range(len(work))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    for k, (a, b) in enumerate(zip(want, got)):
        np.testing.assert_array_equal(a[0], b[0], err_msg=f"x, context {k}")
        np.testing.assert_array_equal(a[1], b[1], err_msg=f"v, context {k}")
        assert a[2] == b[2], (k, a[2], b[2])
