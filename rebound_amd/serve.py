"""Serving loop on the GPU (C entry points reb_amd_server_*).

A Server accepts planetary-system jobs (WHFast, democratic heliocentric) and
N-body jobs (leapfrog with direct gravity), admits them by device-memory
footprint, batches the planetary systems continuously into one device pool
and overlaps N-body jobs on their own streams.  Each job's result equals the
same simulation integrated alone by reb_simulation_steps (simulation.c:504-513).
A job's state can be written to a Simulationarchive file while it runs
(checkpoint) and a simulation read from one submitted to continue it.
"""
import ctypes

import numpy as np

from . import _abi, lib, RebError

PRECISION = {"fp64": 0, "exact": 1, "split": 2}


def _particles(state):
    state = np.asarray(state, dtype=np.float64)
    n = state.shape[0]
    parts = (_abi.Particle * n)()
    buf = np.ctypeslib.as_array(ctypes.cast(parts, ctypes.POINTER(ctypes.c_double)),
                                shape=(n * ctypes.sizeof(_abi.Particle) // 8,)).reshape(n, -1)
    buf[:, :min(11, state.shape[1])] = state[:, :11]
    return parts, n


class Server:
    def __init__(self, memory_budget=0, max_nbody_jobs=0, tick_steps=0):
        L = lib()
        self._h = L.reb_amd_server_create(int(memory_budget), int(max_nbody_jobs), int(tick_steps))
        if not self._h:
            raise RebError("reb_amd_server_create failed: " + L.reb_amd_last_error().decode())
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().reb_amd_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_id(self, jid, n):
        if jid < 0:
            raise RebError("submit failed: " + lib().reb_amd_last_error().decode())
        self._n[jid] = n
        return jid

    def submit_whfast(self, state, dt, G=1.0, safe_mode=0, steps=1):
        parts, n = _particles(state)
        return self._check_id(lib().reb_amd_server_submit_whfast(self._h, n, parts, dt, G, int(safe_mode),
                                                                 int(steps)), n)

    def submit_nbody(self, state, G=1.0, softening=0.0, dt=1e-3, steps=1, gravity="basic", precision="fp64"):
        parts, n = _particles(state)
        return self._check_id(lib().reb_amd_server_submit_nbody(self._h, n, parts, G, softening, dt, int(steps),
                                                                _abi.GRAVITY[gravity], PRECISION[precision]), n)

    def submit_simulation(self, sim, steps=1):
        """A Sim (WHFast democratic heliocentric, or order-2 leapfrog with
        direct gravity) as a job continuing from its state for `steps` steps."""
        return self._check_id(lib().reb_amd_server_submit_simulation(self._h, sim.ptr, int(steps)), sim.s.N)

    def checkpoint(self, jid, filename):
        """Appends the job's current state to an archive file; returns once written."""
        if lib().reb_amd_server_checkpoint(self._h, jid, str(filename).encode()):
            raise RebError("checkpoint of job %d: %s" % (jid, lib().reb_amd_last_error().decode()))

    def status(self, jid):
        return _abi.JOB_STATE[lib().reb_amd_server_status(self._h, jid)]

    def wait(self, jid, timeout=-1.0):
        return _abi.JOB_STATE[lib().reb_amd_server_wait(self._h, jid, float(timeout))]

    def result(self, jid):
        """(state (n x 10: x y z vx vy vz ax ay az m), (E_before, E_after),
        (t_submit, t_admit, t_done)); forgets the job."""
        n = self._n[jid]
        parts = (_abi.Particle * n)()
        e = (ctypes.c_double * 2)()
        t = (ctypes.c_double * 3)()
        if lib().reb_amd_server_result(self._h, jid, parts, n, e, t):
            raise RebError("job %d: %s" % (jid, lib().reb_amd_last_error().decode()))
        del self._n[jid]
        buf = np.ctypeslib.as_array(ctypes.cast(parts, ctypes.POINTER(ctypes.c_double)),
                                    shape=(n * ctypes.sizeof(_abi.Particle) // 8,)).reshape(n, -1)
        return buf[:, :10].copy(), (e[0], e[1]), (t[0], t[1], t[2])

    def job_error(self, jid):
        return lib().reb_amd_server_job_error(self._h, jid).decode()

    def stats(self):
        out = (ctypes.c_double * 10)()
        lib().reb_amd_server_stats(self._h, out, 10)
        keys = ("held", "pending", "running", "done", "failed", "bytes_used", "peak_bytes", "budget", "ticks",
                "pool_busy_s")
        return dict(zip(keys, list(out)))
