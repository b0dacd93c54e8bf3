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
