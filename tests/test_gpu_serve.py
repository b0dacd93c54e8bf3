"""Serving loop: each job's result equals the same simulation run alone, and
admission respects the memory budget.

Planetary jobs are compared bitwise with the batched ensemble integrating the
system by itself in one steps() call (which the ensemble tests pin to the
reference); N-body jobs bitwise with a reb_simulation driven alone through
reb_simulation_steps with the same precision."""
import time

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _systems(n):
    nb = golden("ensemble200_seed99_nbodies.npy")
    init = golden("ensemble200_seed99_init.npy")
    dt = golden("ensemble200_seed99_dt.npy")
    first = np.concatenate([[0], np.cumsum(nb)[:-1]])
    return [(init[first[s]:first[s] + nb[s]], dt[s]) for s in range(n)]


def _alone_whfast(gpu, st, dt, steps, G=1.0, safe=0):
    ens = gpu.Ensemble([st.shape[0]], st, dt, G=G, safe_mode=safe)
    ens.steps(steps)
    out = ens.state(), ens.energy()[0]
    ens.free()
    return out


def _plummer(gpu, n, seed):
    sim = gpu.Sim()
    sim.s.rand_seed = seed
    sim.add_plummer(n, 1.0, 1.0)
    sim.move_to_com()
    st = sim.state()
    sim.free()
    return st


def _alone_nbody(gpu, st, steps, precision, soft=0.01, dt=1e-3):
    sim = gpu.Sim()
    sim.s.G = 1.0
    sim.s.softening = soft
    sim.s.dt = dt
    sim.set_state(st)
    sim.set_integrator("leapfrog")
    sim.precision = precision
    sim.steps(steps)
    out = sim.state()
    sim.free()
    return out


def test_whfast_jobs_mix_equals_alone(gpu):
    systems = _systems(60)
    horizons = [40 + (37 * s) % 300 for s in range(len(systems))]
    G = [1.0 if s % 3 else 0.5 for s in range(len(systems))]
    safe = [1 if s % 7 == 0 else 0 for s in range(len(systems))]
    want = [_alone_whfast(gpu, st, dt, h, g, sm) for (st, dt), h, g, sm in zip(systems, horizons, G, safe)]
    srv = gpu.Server(tick_steps=64)
    # staggered arrivals: the second half joins while the first is running
    ids = []
    for k, ((st, dt), h) in enumerate(zip(systems, horizons)):
        ids.append(srv.submit_whfast(st, dt, G=G[k], safe_mode=safe[k], steps=h))
        if k == 30:
            time.sleep(0.02)
    for k, jid in enumerate(ids):
        assert srv.wait(jid, 60) == "done", srv.job_error(jid)
        st, (e0, e1), (ts, ta, td) = srv.result(jid)
        ref_state, ref_e = want[k]
        assert np.array_equal(st[:, :6], ref_state[:, :6]), k
        assert e1 == ref_e, k
        assert ts <= ta <= td
    s = srv.stats()
    assert s["done"] == len(systems) and s["failed"] == 0 and s["ticks"] > 1
    srv.close()


def test_whfast_job_alone_on_server(gpu):
    (st, dt), = _systems(1)
    ref_state, ref_e = _alone_whfast(gpu, st, dt, 500)
    srv = gpu.Server(tick_steps=128)      # 500 steps over four ticks
    jid = srv.submit_whfast(st, dt, steps=500)
    assert srv.wait(jid, 60) == "done"
    got, (e0, e1), _ = srv.result(jid)
    assert np.array_equal(got[:, :6], ref_state[:, :6]) and e1 == ref_e
    srv.close()


def test_nbody_jobs_mix_equals_alone(gpu):
    jobs = [(_plummer(gpu, 300, 1), "split", 30), (_plummer(gpu, 700, 2), "fp64", 20),
            (_plummer(gpu, 1500, 3), "split", 15), (_plummer(gpu, 256, 4), "exact", 25)]
    want = [_alone_nbody(gpu, st, k, p) for st, p, k in jobs]
    srv = gpu.Server(max_nbody_jobs=4)
    # planetary jobs share the device meanwhile
    extra = [srv.submit_whfast(st, dt, steps=400) for st, dt in _systems(20)]
    ids = [srv.submit_nbody(st, G=1.0, softening=0.01, dt=1e-3, steps=k, precision=p) for st, p, k in jobs]
    for k, jid in enumerate(ids):
        assert srv.wait(jid, 120) == "done", srv.job_error(jid)
        got, _, _ = srv.result(jid)
        assert np.array_equal(got[:, :6], want[k][:, :6]), k
    for jid in extra:
        assert srv.wait(jid, 60) == "done"
    srv.close()


def test_admission_by_memory_footprint(gpu):
    st = _plummer(gpu, 2000, 5)
    one = gpu.lib().reb_amd_server_nbody_bytes(2000, gpu.PRECISION["split"])
    srv = gpu.Server(memory_budget=int(1.5 * one), max_nbody_jobs=4)    # one N=2000 job at a time
    ids = [srv.submit_nbody(st, softening=0.01, steps=20, precision="split") for _ in range(3)]
    times = []
    for jid in ids:
        assert srv.wait(jid, 120) == "done", srv.job_error(jid)
        times.append(srv.result(jid)[2])
    s = srv.stats()
    assert s["peak_bytes"] <= s["budget"]
    # each job is admitted only after the previous one finished
    for a, b in zip(times, times[1:]):
        assert b[1] >= a[2]
    srv.close()


def test_job_larger_than_budget_is_rejected(gpu):
    srv = gpu.Server(memory_budget=10_000)
    with pytest.raises(gpu.RebError, match="exceeds the memory budget"):
        srv.submit_nbody(_plummer(gpu, 4096, 6), steps=1)
    srv.close()


def _wait_running(srv, jid, limit=30.0):
    t0 = time.time()
    while srv.status(jid) == "pending":
        assert time.time() - t0 < limit
        time.sleep(0.0005)


def _t_after(t0, dt, k):
    """The time reb_simulation_steps reaches after k steps: two half steps each."""
    t = t0
    for _ in range(k):
        t += dt / 2.0
        t += dt / 2.0
    return t


@pytest.mark.parametrize("safe", [0, 1])
def test_whfast_checkpoint_resume(gpu, tmp_path, safe):
    """A running planetary job written to an archive file mid-run: the file
    read back and stepped for the remaining steps (alone, or submitted again)
    gives the job's result bit for bit; a finished job's checkpoint continues
    as the simulation stepped further would."""
    (st, dt), = _systems(1)
    total = 200000
    path = str(tmp_path / "job.bin")
    srv = gpu.Server(tick_steps=500)
    jid = srv.submit_whfast(st, dt, safe_mode=safe, steps=total)
    _wait_running(srv, jid)
    srv.checkpoint(jid, path)
    srv.checkpoint(jid, path)       # a second snapshot appended later in the run
    assert srv.wait(jid, 120) == "done", srv.job_error(jid)
    done_path = str(tmp_path / "done.bin")
    srv.checkpoint(jid, done_path)
    want, _, _ = srv.result(jid)
    mid = 0
    for snap in (0, 1):
        sim = gpu.Sim.from_file(path, snap)
        k = sim.s.steps_done
        assert 0 <= k <= total
        mid += 0 < k < total
        assert sim.s.t == _t_after(0.0, dt, k) and sim.s.dt == dt
        sim.steps(total - k)
        sim.synchronize()
        assert np.array_equal(sim.state()[:, :6], want[:, :6]), (snap, k)
        # the same snapshot as a new job
        again = gpu.Sim.from_file(path, snap)
        j2 = srv.submit_simulation(again, steps=total - k)
        assert srv.wait(j2, 120) == "done", srv.job_error(j2)
        got, _, _ = srv.result(j2)
        assert np.array_equal(got[:, :6], want[:, :6]), (snap, k)
        again.free()
        if snap == 1:
            # the finished job's checkpoint continues like the stepped simulation
            fin = gpu.Sim.from_file(done_path, 0)
            assert fin.s.steps_done == total and np.array_equal(fin.state()[:, :6], want[:, :6])
            fin.steps(300)
            fin.synchronize()
            sim.steps(300)
            sim.synchronize()
            assert np.array_equal(fin.state()[:, :6], sim.state()[:, :6])
            fin.free()
        sim.free()
    assert mid >= 1, "no checkpoint landed inside the run"
    srv.close()


@pytest.mark.parametrize("precision,n,total", [("fp64", 512, 6000), ("split", 700, 6000), ("exact", 256, 3000)])
def test_nbody_checkpoint_resume(gpu, tmp_path, precision, n, total):
    """N-body jobs (batched, or alone for EXACT) checkpointed mid-run resume
    to the same result, read back alone or submitted again."""
    st = _plummer(gpu, n, 7)
    path = str(tmp_path / "nb.bin")
    srv = gpu.Server(max_nbody_jobs=4)
    jid = srv.submit_nbody(st, G=1.0, softening=0.01, dt=1e-3, steps=total, precision=precision)
    _wait_running(srv, jid)
    srv.checkpoint(jid, path)
    assert srv.wait(jid, 120) == "done", srv.job_error(jid)
    want, _, _ = srv.result(jid)
    sim = gpu.Sim.from_file(path, 0)
    k = sim.s.steps_done
    assert 0 <= k < total
    assert sim.s.t == _t_after(0.0, 1e-3, k) and sim.s.softening == 0.01
    sim.precision = precision           # an engine setting, not part of the file
    j2 = srv.submit_simulation(sim, steps=total - k)
    sim.steps(total - k)
    assert np.array_equal(sim.state()[:, :6], want[:, :6]), k
    assert srv.wait(j2, 120) == "done", srv.job_error(j2)
    got, _, _ = srv.result(j2)
    assert np.array_equal(got[:, :6], want[:, :6]), k
    sim.free()
    srv.close()


def test_checkpoint_of_pending_and_unknown_jobs(gpu, tmp_path):
    st = _plummer(gpu, 2000, 5)
    one = gpu.lib().reb_amd_server_nbody_bytes(2000, gpu.PRECISION["fp64"])
    srv = gpu.Server(memory_budget=int(1.5 * one), max_nbody_jobs=4)
    a = srv.submit_nbody(st, softening=0.01, steps=400)
    b = srv.submit_nbody(st, softening=0.01, steps=5)     # waits for a
    path = str(tmp_path / "pending.bin")
    if srv.status(b) == "pending":
        srv.checkpoint(b, path)
        sim = gpu.Sim.from_file(path, 0)
        assert sim.s.steps_done == 0 and sim.s.t == 0.0 and np.array_equal(sim.state()[:, :6], st[:, :6])
        sim.free()
    with pytest.raises(gpu.RebError, match="unknown job"):
        srv.checkpoint(12345, path)
    for jid in (a, b):
        assert srv.wait(jid, 120) == "done"
        srv.result(jid)
    srv.close()


def test_failing_jobs_do_not_affect_others(gpu):
    """Jobs whose state turns non-finite fail with a message; the planetary
    and N-body jobs running beside them give their results alone, bit for
    bit, and the server takes new work afterwards."""
    systems = _systems(12)
    want = [_alone_whfast(gpu, st, dt, 300) for st, dt in systems]
    nb = [(_plummer(gpu, 300, 11), "fp64"), (_plummer(gpu, 256, 12), "exact")]
    nb_want = [_alone_nbody(gpu, s, 40, p) for s, p in nb]
    srv = gpu.Server(tick_steps=64, max_nbody_jobs=8)
    bad = []
    bst = systems[3][0].copy()
    bst[2, 3] = np.nan                  # a planet's velocity
    bad.append(srv.submit_whfast(bst, systems[3][1], steps=300))
    ids = [srv.submit_whfast(st, dt, steps=300) for st, dt in systems]
    nb_ids = [srv.submit_nbody(s, softening=0.01, dt=1e-3, steps=40, precision=p) for s, p in nb]
    for s, p in nb:
        b = s.copy()
        b[5, 0] = np.nan
        bad.append(srv.submit_nbody(b, softening=0.01, dt=1e-3, steps=40, precision=p))
    for jid in bad:
        assert srv.wait(jid, 120) == "failed"
        assert "non-finite" in srv.job_error(jid)
        with pytest.raises(gpu.RebError, match="job failed"):
            srv.result(jid)
    for k, jid in enumerate(ids):
        assert srv.wait(jid, 120) == "done", srv.job_error(jid)
        got, _, _ = srv.result(jid)
        assert np.array_equal(got[:, :6], want[k][0][:, :6]), k
    for k, jid in enumerate(nb_ids):
        assert srv.wait(jid, 120) == "done", srv.job_error(jid)
        got, _, _ = srv.result(jid)
        assert np.array_equal(got[:, :6], nb_want[k][:, :6]), k
    st, dt = systems[0]
    again = srv.submit_whfast(st, dt, steps=300)
    assert srv.wait(again, 60) == "done"
    assert np.array_equal(srv.result(again)[0][:, :6], want[0][0][:, :6])
    s = srv.stats()
    assert s["failed"] == len(bad)
    srv.close()
