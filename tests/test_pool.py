"""Pooled engine (batched / served runs) through the C ABI.

A job in the pool runs the same kernels, in the same order, as a plain run
of the same system (`run N` on its own handle): the per-system results are
bit-identical to the plain run and to the same job run alone in a pool, also
when it is admitted late into the slot and ranges a finished job freed, and
when the pool widens its neighbor rows mid-run.  Only a different launch
split (waves per tile) may change results, by FP32 summation order; that case
is held to 1e-5 relative.
"""
import numpy as np
import pytest

from conftest import uniform_box
from test_gpu_parity import assert_vec_close, dpd_system, mdpd_system

pytestmark = pytest.mark.gpu

GB = float(1 << 30)


def droplet(engine, n, seed, rho=6.05, extra="", rho_est=6.05, ntypes=1):
    """n particles uniform in a sphere at density rho, cube side 2.5 R."""
    rng = np.random.default_rng(seed)
    R = (3.0 * n / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
    L = 2.5 * R
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = R * rng.uniform(0, 1, n) ** (1.0 / 3.0)
    x = (L / 2 + d * r[:, None]).astype(np.float32).astype(np.float64)
    v = rng.normal(0, 1, (n, 3))
    v = (v - v.mean(axis=0)).astype(np.float32).astype(np.float64)
    types = None if ntypes == 1 else rng.integers(1, ntypes + 1, n).astype(np.int32)
    coeff = "* * -40 25 4.5 1.0 0.75"
    if ntypes == 2:
        coeff = "1 1 -40 25 4.5 1.0 0.75\npair_coeff 1 2 -30 25 4.5 1.0 0.75\npair_coeff 2 2 -40 25 4.5 1.0 0.75"
    return mdpd_system(engine, x, v, np.array([L, L, L]), types=types, ntypes=ntypes, extra=extra,
                       rho_est=rho_est, seed=1000 + seed, coeff=coeff)


def plain_result(s, nsteps):
    s.tune("split", 2)
    s.run(nsteps)
    return {k: s.gather(k) for k in ("x", "v", "f", "image", "e_pair", "virial")}, s.thermo(), s.timestep


def pool_results(engine, systems, nsteps, budget=0.5 * GB, max_systems=16, split=2, poll=4):
    p = engine.Pool(budget, max_systems)
    p.tune("split", split)
    p.tune("poll", poll)
    ids = [p.submit(s, k) for s, k in zip(systems, nsteps)]
    p.run()
    out = []
    for j in ids:
        assert p.state(j) == engine.JOB_DONE
        r = {k: p.gather(j, k) for k in ("x", "v", "f", "image", "e_pair", "virial")}
        out.append((r, p.thermo(j), p.info(j)["timestep"]))
    return out, p


def assert_same(a, b, what):
    for k in a[0]:
        np.testing.assert_array_equal(a[0][k], b[0][k], err_msg=f"{what}: {k}")
    assert a[2] == b[2], (what, a[2], b[2])
    for k in a[1]:
        assert a[1][k] == b[1][k], (what, k, a[1][k], b[1][k])


def test_single_job_equals_plain_run(engine):
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 11)
    ref = plain_result(mdpd_system(engine, x, v, L), 25)
    got, p = pool_results(engine, [mdpd_system(engine, x, v, L)], [25])
    assert_same(got[0], ref, "single mDPD job")
    assert p.stat("passes") >= 26


def test_results_moved_to_pageable_memory_are_unchanged(engine):
    # with no pinned memory allowed for finished jobs ("hold" 0) every result
    # block is copied to pageable memory before it is read back
    sizes = [(1296, 6.0, 11, 25), (700, 5.0, 12, 18)]
    def run(hold):
        p = engine.Pool(0.5 * GB, 16)
        p.tune("split", 2)
        p.tune("poll", 4)
        p.tune("hold", hold)
        ids = []
        for n, L, seed, k in sizes:
            x, v = uniform_box(n, L, seed)
            ids.append(p.submit(mdpd_system(engine, x, v, L), k))
        p.run()
        out = [({k: p.gather(j, k) for k in ("x", "v", "f", "image", "e_pair", "virial", "rho", "tag", "type", "mask")},
                p.thermo(j), p.info(j)["timestep"]) for j in ids]
        # finished jobs hold pinned memory only below the limit, none after release
        assert (p.stat("held_pinned") > 0) == (hold > 0)
        for j in ids:
            p.release(j)
        assert p.stat("held_pinned") == 0
        p.close()
        return out
    pinned, pageable = run(1024), run(0)
    assert all(np.isfinite(a[0]["x"]).all() for a in pinned)
    for a, b in zip(pinned, pageable):
        assert_same(a, b, "hold 0")


def test_mixed_pool_equals_each_alone(engine):
    # mDPD droplets of several sizes, one with two types and a body force,
    # a DPD cube; different run lengths
    systems = lambda: [
        droplet(engine, 2000, 1),
        droplet(engine, 5000, 2, ntypes=2, extra="fix push all addforce/meso 0.05 0 0"),
        dpd_system(engine, *uniform_box(3000, 10.0, 3), 10.0),
        droplet(engine, 800, 4),
    ]
    nsteps = [30, 17, 24, 9]
    alone = [pool_results(engine, [s], [k])[0][0] for s, k in zip(systems(), nsteps)]
    mixed, p = pool_results(engine, systems(), nsteps)
    for k, (a, b) in enumerate(zip(mixed, alone)):
        assert_same(a, b, f"job {k}")
    # and each equals a plain run of the same system
    plain = plain_result(systems()[1], nsteps[1])
    assert_same(mixed[1], plain, "two-type droplet vs plain")


def test_queued_jobs_reuse_freed_ranges(engine):
    # a budget that holds about two of these jobs: the rest wait in the queue
    # and are admitted into the slot / ranges of finished ones
    mk = lambda: [droplet(engine, 3000 + 500 * k, 20 + k) for k in range(6)]
    nsteps = [12, 20, 8, 15, 10, 6]
    alone = [pool_results(engine, [s], [k])[0][0] for s, k in zip(mk(), nsteps)]
    small, p = pool_results(engine, mk(), nsteps, budget=9.0e6, max_systems=3)
    assert p.stat("capacity") < 3 * 6000
    for k, (a, b) in enumerate(zip(small, alone)):
        assert_same(a, b, f"queued job {k}")


def test_narrow_rows_regrow_in_place(engine):
    # rho_est 0.5 sizes the rows for ~9 neighbors: the first build overflows,
    # the pool widens its rows (no re-sort) and the results do not change
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 12)
    ref = plain_result(mdpd_system(engine, x, v, L), 15)
    got, p = pool_results(engine, [mdpd_system(engine, x, v, L, rho_est=0.5)], [15])
    assert p.stat("regrows") >= 1
    assert_same(got[0], ref, "regrown rows")


def test_autotuned_split_within_tolerance(engine):
    n, L = 1296, 6.0
    x, v = uniform_box(n, L, 13)
    ref = plain_result(mdpd_system(engine, x, v, L), 5)
    got, p = pool_results(engine, [mdpd_system(engine, x, v, L)], [5], split=0)
    assert p.stat("split") in (1, 2, 4)
    assert p.stat("tune_ms_2") > 0
    assert_vec_close(got[0][0]["x"] - x, ref[0]["x"] - x, 1e-5, "autotuned displacement")
    assert_vec_close(got[0][0]["f"], ref[0]["f"], 1e-5, "autotuned force")


def test_served_submission_while_running(engine):
    # jobs submitted while others are on the device (continuous batching):
    # results equal each job alone
    mk = [lambda k=k: droplet(engine, 1500 + 700 * k, 40 + k) for k in range(5)]
    nsteps = [40, 10, 25, 5, 30]
    alone = [pool_results(engine, [m()], [k])[0][0] for m, k in zip(mk, nsteps)]
    p = engine.Pool(0.5 * GB, 8)
    p.tune("split", 2)
    p.tune("poll", 2)
    ids = [p.submit(mk[0](), nsteps[0])]
    for k in range(1, 5):
        p.advance(max_passes=6)
        ids.append(p.submit(mk[k](), nsteps[k]))
    p.run()
    for k, j in enumerate(ids):
        r = ({q: p.gather(j, q) for q in ("x", "v", "f", "image", "e_pair", "virial")}, p.thermo(j),
             p.info(j)["timestep"])
        assert_same(r, alone[k], f"served job {k}")
    info = [p.info(j) for j in ids]
    assert all(i["t_submit"] <= i["t_admit"] <= i["t_done"] for i in info)
    assert sorted(p.done()) == sorted(ids)


def test_backfill_admits_past_a_job_that_does_not_fit(engine):
    # job b does not fit beside the running job a; the later, smaller job c
    # does and is admitted past b; b runs when a's range is free.  Every
    # result equals the job run alone.
    p = engine.Pool(9.0e6, 4)
    p.tune("split", 2)
    p.tune("poll", 4)
    pad = lambda n: (n + 255) // 256 * 256
    na, nc = 3000, 200
    ja = p.submit(droplet(engine, na, 40), 20)
    cap = int(p.stat("capacity"))
    nb = cap - pad(na) + 256                      # fits alone, not beside a
    assert pad(na) + pad(nb) > cap >= pad(nb) and pad(na) + pad(nc) <= cap
    jb = p.submit(droplet(engine, nb, 41), 10)
    jc = p.submit(droplet(engine, nc, 42), 15)
    p.admit()
    assert (p.state(ja), p.state(jb), p.state(jc)) == (engine.JOB_RUNNING, engine.JOB_QUEUED, engine.JOB_RUNNING)
    p.run()
    for j, (n, seed, k) in zip((ja, jb, jc), ((na, 40, 20), (nb, 41, 10), (nc, 42, 15))):
        assert p.state(j) == engine.JOB_DONE
        got = ({q: p.gather(j, q) for q in ("x", "v", "f", "image", "e_pair", "virial")}, p.thermo(j),
               p.info(j)["timestep"])
        alone = pool_results(engine, [droplet(engine, n, seed)], [k])[0][0]
        assert_same(got, alone, f"job of {n} atoms")


def test_regrow_relists_moved_systems_from_their_last_build(engine):
    # job a runs for a while; job b, submitted later, is far denser than the
    # pool's rows allow, so its first build overflows and the pool widens the
    # rows and re-lists every sorted system -- a among them, after steps taken
    # since its last build.  a gets back the rows of its last build: its
    # results equal a run that never overflowed.
    a_sys = lambda: droplet(engine, 3000, 50)
    alone = pool_results(engine, [a_sys()], [40])[0][0]
    p = engine.Pool(0.5 * GB, 4)
    p.tune("split", 2)
    p.tune("poll", 4)
    ja = p.submit(a_sys(), 40)
    p.advance(max_passes=17)
    assert p.state(ja) == engine.JOB_RUNNING
    n, L = 1500, 3.4                               # density ~38
    x, v = uniform_box(n, L, 51)
    jb = p.submit(mdpd_system(engine, x, v, L), 3)
    p.run()
    assert p.stat("regrows") >= 1
    assert p.state(jb) == engine.JOB_DONE
    got = ({k: p.gather(ja, k) for k in ("x", "v", "f", "image", "e_pair", "virial")}, p.thermo(ja),
           p.info(ja)["timestep"])
    assert_same(got, alone, "job re-listed after steps")


def test_many_resident_jobs_beyond_one_launch_block(engine):
    # 80 resident jobs: per-system launches cover the occupied descriptor
    # slots rounded up to 64, so slots 64..79 sit in the second block; jobs
    # on both sides equal their plain runs, and the sizes differ so the
    # chunked scan runs with several workgroups per system
    sizes = [400 + 37 * k for k in range(79)] + [20000]
    mk = lambda k: droplet(engine, sizes[k], 300 + k)
    nsteps = [6 + k % 5 for k in range(80)]
    got, p = pool_results(engine, [mk(k) for k in range(80)], nsteps, budget=1.0 * GB, max_systems=1024)
    for k in (0, 63, 64, 78, 79):
        assert_same(got[k], plain_result(mk(k), nsteps[k]), f"job {k}")


def test_full_batch_of_256_droplets(engine):
    # the batched workload at its full size: 256 droplets with N ~
    # LogNormal(ln 8000, 0.5) clipped to [2000, 40000] (~2.4M particles) in
    # one pool for 20 steps.  Size-independent checks: every job reaches
    # step 20 with finite state and all of its atoms, its net momentum stays
    # at FP32 rounding level, and the smallest, a middle and the largest job
    # equal the same systems run alone, bit for bit
    rng = np.random.default_rng(20241202)
    sizes = np.clip(np.exp(rng.normal(np.log(8000.0), 0.5, 256)), 2000, 40000).astype(int)
    assert 2.0e6 < sizes.sum() < 2.8e6
    mk = lambda k: droplet(engine, int(sizes[k]), 7000 + k)
    got, p = pool_results(engine, [mk(k) for k in range(256)], [20] * 256, budget=16 * GB, max_systems=256)
    worst = 0.0
    for k, (r, th, ts) in enumerate(got):
        assert ts == 20, k
        assert r["x"].shape == (sizes[k], 3), k
        for key in ("x", "v", "f", "e_pair", "virial"):
            assert np.isfinite(r[key]).all(), (k, key)
        # the droplets start at zero net momentum (to FP32 rounding)
        worst = max(worst, np.abs(r["v"].sum(axis=0)).max() / np.sqrt(sizes[k]))
    print("largest |net momentum| / sqrt(N) after 20 steps", worst)
    assert worst < 2e-5
    order = np.argsort(sizes)
    for k in (order[0], order[128], order[-1]):
        assert_same(got[k], plain_result(mk(k), 20), f"job {k} (N = {sizes[k]})")
    p.close()


def test_job_finished_before_a_regrow_keeps_its_last_pass(engine):
    # job a finishes on the device inside a batch of passes whose status word
    # the host has not read yet (it reads them one batch behind); job b,
    # admitted next, needs wider neighbor rows, so the pool reallocates them
    # before a's status is handled.  a's harvest pass (energy and virial of
    # its last pass) reads that pass's compact rows: a is harvested before
    # the reallocation, and its results equal a alone.
    a_sys = lambda: droplet(engine, 2000, 60)
    alone = pool_results(engine, [a_sys()], [3], poll=8)[0][0]
    p = engine.Pool(0.5 * GB, 4)
    p.tune("split", 2)
    p.tune("poll", 8)
    ja = p.submit(a_sys(), 3)
    p.advance(max_passes=8)                         # a takes 4 passes and is done on the device
    assert p.state(ja) == engine.JOB_RUNNING        # its word is not handled yet
    jb = p.submit(droplet(engine, 1500, 61, rho_est=15.0), 4)
    width = p.stat("maxnbr")
    p.run()
    assert p.stat("maxnbr") > width and p.stat("regrows") >= 1
    assert p.state(ja) == engine.JOB_DONE and p.state(jb) == engine.JOB_DONE
    got = ({k: p.gather(ja, k) for k in ("x", "v", "f", "image", "e_pair", "virial")}, p.thermo(ja),
           p.info(ja)["timestep"])
    assert_same(got, alone, "job finished before a regrow")


def test_split_change_before_a_harvest_keeps_the_last_pass(engine):
    # job a finishes at split 1 while job b runs on; the split then changes
    # to 4 before a's status is handled.  b's passes at split 4 must not
    # touch the compact-row counts a's harvest pass (energy, virial, f of its
    # last pass) reads: a's results equal a alone at split 1, bit for bit
    a_sys = lambda: droplet(engine, 2000, 62)
    alone = pool_results(engine, [a_sys()], [3], split=1, poll=8)[0][0]
    p = engine.Pool(0.5 * GB, 4)
    p.tune("split", 1)
    p.tune("poll", 8)
    # b first: its tiles come first, so its split-4 count slots would cover
    # a's tiles if the slots followed the split
    jb = p.submit(droplet(engine, 6000, 63), 40)
    ja = p.submit(a_sys(), 3)
    p.advance(max_passes=8)                         # a is done on the device, its word not handled
    assert p.state(ja) == engine.JOB_RUNNING
    p.tune("split", 4)
    p.run()
    assert p.state(ja) == engine.JOB_DONE and p.state(jb) == engine.JOB_DONE
    got = ({k: p.gather(ja, k) for k in ("x", "v", "f", "image", "e_pair", "virial")}, p.thermo(ja),
           p.info(ja)["timestep"])
    assert_same(got, alone, "job harvested after a split change")


def test_rdf_of_pool_jobs_equals_plain_runs(engine, tmp_path):
    # rdf/fast/meso on pooled jobs: each job's g(r) equals its handle's after
    # the same run, bit for bit (two fixes on one job: every 5 over all
    # pairs, every 3 over pairs with a type-2 partner and pi exact; one fix
    # sampling every step on a second job; a third job without the fix)
    def fixes(tag):
        if tag == "a":
            return (f"group b type 2\n"
                    f"fix g1 all rdf/fast/meso output {tmp_path}/a1.out nbin 25 every 5\n"
                    f"fix g2 all rdf/fast/meso output {tmp_path}/a2.out nbin 10 every 3 other b pi exact")
        if tag == "b":
            return f"fix g all rdf/fast/meso output {tmp_path}/b.out nbin 40"
        return ""
    specs = [("a", 2000, 71, 2, 23), ("b", 3000, 72, 1, 17), ("c", 2500, 73, 1, 30)]
    make = lambda t, n, seed, nt: droplet(engine, n, seed, ntypes=nt, extra=fixes(t))
    ids = {"a": ("g1", "g2"), "b": ("g",), "c": ()}
    want = {}
    for t, n, seed, nt, steps in specs:
        s = make(t, n, seed, nt)
        s.tune("split", 2)
        s.run(steps)
        want[t] = [s.fix_result(f) for f in ids[t]]
        s.close()
    got, p = pool_results(engine, [make(t, n, seed, nt) for t, n, seed, nt, _ in specs],
                          [steps for *_, steps in specs])
    for k, (t, *_rest) in enumerate(specs):
        job = k + 1
        for f, (x, g) in zip(ids[t], want[t]):
            xp, gp = p.fix_result(job, f)
            np.testing.assert_array_equal(xp, x, err_msg=f"{t}:{f}")
            np.testing.assert_array_equal(gp, g, err_msg=f"{t}:{f}")
            assert g.max() > 0, (t, f)
    with pytest.raises(RuntimeError, match="no rdf"):
        p.fix_result(3, "g")
    # a third rdf fix, a second density fix and too many windows are refused
    three = "\n".join(f"fix r{k} all rdf/fast/meso output {tmp_path}/r{k} nbin 8" for k in range(3))
    with pytest.raises(RuntimeError, match="at most 2"):
        p.submit(droplet(engine, 2000, 75, extra=three), 5)
    two = "\n".join(f"fix p{k} all density/meso output {tmp_path}/p{k} along z nbin 4 every 10 window 2" for k in range(2))
    with pytest.raises(RuntimeError, match="at most one"):
        p.submit(droplet(engine, 2000, 76, extra=two), 5)
    with pytest.raises(RuntimeError, match="exceed"):
        p.submit(droplet(engine, 2000, 77, extra=f"fix p all density/meso output {tmp_path}/q along z nbin 100 every 10 window 2"), 1000)


def test_density_windows_of_pool_jobs_equal_the_files_of_plain_runs(engine, tmp_path):
    # density/meso on pooled jobs: every window's profile equals the file a
    # handle writes for it after the same run, and the last window equals the
    # handle's current result (job a: four whole windows; job b ends inside
    # its last window, which the handle has not written)
    specs = [("a", 2000, 81, 70, "along z nbin 6 every 20 window 3", [0, 20, 40, 60]),
             ("b", 2600, 82, 27, "along x nbin 5 every 7 window 2", [0, 7, 14, 21, 28])]
    want = {}
    for t, n, seed, steps, args, centres in specs:
        d = tmp_path / t
        d.mkdir()
        s = droplet(engine, n, seed, extra=f"fix p all density/meso output {d}/prof {args}")
        s.tune("split", 2)
        s.run(steps)
        want[t] = ([np.loadtxt(d / f"prof.{m:09d}")[:, 1] for m in centres if (d / f"prof.{m:09d}").exists()],
                   s.fix_result("p"))
        s.close()
    got, p = pool_results(engine, [droplet(engine, n, seed, extra=f"fix p all density/meso output {tmp_path}/x {args}")
                                   for _, n, seed, _, args, _ in specs], [sp[3] for sp in specs])
    for k, (t, n, seed, steps, args, centres) in enumerate(specs):
        c, v = p.fix_windows(k + 1, "p")
        assert c.tolist() == centres, (t, c)
        files, (x, last) = want[t]
        assert len(files) >= len(centres) - 1
        for w, f in enumerate(files):
            np.testing.assert_allclose(v[w], f, rtol=1e-14, atol=0, err_msg=f"{t} window {centres[w]}")
            assert f.sum() > 0
        xp, lp = p.fix_result(k + 1, "p")
        np.testing.assert_array_equal(xp, x)
        if t == "b":
            np.testing.assert_array_equal(lp, last)             # the handle's current window
        np.testing.assert_array_equal(lp, v[-1])
