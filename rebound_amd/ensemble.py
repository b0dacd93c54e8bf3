"""Batched WHFast ensembles on the GPU (C entry points reb_amd_ensemble_*).

An Ensemble holds many independent planetary systems in device memory and
steps all of them in one persistent kernel launch per request (replayed from a
captured hipGraph).  Each system evolves bit-identically to a
reb_simulation with the same particles integrated by WHFast in democratic
heliocentric coordinates (integrator_whfast.c:1113-1259) through
reb_simulation_steps (simulation.c:504-513).

`planetary_systems` generates the synthetic ensemble workload with the
library's own random generator and orbit conversion (reb_random_*,
reb_particle_from_orbit), so a seed gives the same systems as a C program that
makes the same calls against the reference library.
"""
import ctypes
import math

import numpy as np

from . import _abi, lib, RebError

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "r")


def _check(rc, what, h=None):
    if rc:
        L = lib()
        msg = L.reb_amd_ensemble_error(h).decode() if h else ""
        raise RebError("%s failed: %s" % (what, msg or L.reb_amd_last_error().decode()))


def planetary_systems(n_systems, seed, kmin=2, kmax=8, G=1.0):
    """Star (m=1) plus K planets per system, K uniform in {kmin..kmax}; a_1=1,
    log-uniform period ratios in [1.3,2.5], log-uniform masses in [1e-6,1e-3],
    Rayleigh eccentricities (0.02) and inclinations (0.01), uniform angles,
    moved to the centre of mass; dt = P_1/20 with P_1 = 2 pi sqrt(1/(1+m_1)).

    Returns (n_bodies[int64], particles[(sum n_bodies) x 11 float64 in
    x y z vx vy vz ax ay az m r order], dt[float64])."""
    L = lib()
    g = L.reb_simulation_create()
    g.contents.rand_seed = seed
    uni, ray = L.reb_random_uniform, L.reb_random_rayleigh
    orbit = L.reb_particle_from_orbit
    lo_r, hi_r = math.log(1.3), math.log(2.5)
    lo_m, hi_m = math.log(1e-6), math.log(1e-3)
    two_pi = 2.0 * math.pi
    n_bodies = np.zeros(n_systems, dtype=np.int64)
    dts = np.zeros(n_systems)
    rows = []
    try:
        for s in range(n_systems):
            K = kmin + int(math.floor(uni(g, 0.0, 1.0) * float(kmax - kmin + 1)))
            K = min(K, kmax)
            r = L.reb_simulation_create()
            r.contents.G = G
            star = _abi.Particle()
            star.m = 1.0
            L.reb_simulation_add(r, star)
            a = 1.0
            for k in range(K):
                if k > 0:
                    ratio = math.exp(uni(g, lo_r, hi_r))
                    a *= math.pow(ratio, 2.0 / 3.0)
                m = math.exp(uni(g, lo_m, hi_m))
                e = ray(g, 0.02)
                inc = ray(g, 0.01)
                Om = uni(g, 0.0, two_pi)
                om = uni(g, 0.0, two_pi)
                f = uni(g, 0.0, two_pi)
                p = orbit(r.contents.G, r.contents.particles[0], m, a, e, inc, Om, om, f)
                L.reb_simulation_add(r, p)
            L.reb_simulation_move_to_com(r)
            n = r.contents.N
            arr = np.ctypeslib.as_array(ctypes.cast(r.contents.particles, ctypes.POINTER(ctypes.c_double)),
                                        shape=(n * ctypes.sizeof(_abi.Particle) // 8,))
            arr = arr.reshape(n, -1)
            rows.append(arr[:, :11].copy())
            n_bodies[s] = n
            dts[s] = 2.0 * math.pi * math.sqrt(1.0 / (1.0 + r.contents.particles[1].m)) / 20.0 if n > 1 else 1e-3
            L.reb_simulation_free(r)
    finally:
        L.reb_simulation_free(g)
    return n_bodies, np.concatenate(rows), dts


class Ensemble:
    """Independent systems stepped together on the GPU.

    n_bodies: bodies per system (1..65, central body first); state: rows of
    x y z vx vy vz ax ay az m [r] for all bodies, system after system; dt: one
    timestep per system (or a scalar); coordinates "democraticheliocentric" or
    "jacobi", the latter with optional symplectic correctors."""

    def __init__(self, n_bodies, state, dt, G=1.0, safe_mode=0, gravity="basic", coordinates="democraticheliocentric",
                 corrector=0, corrector2=0, kernel=0):
        L = lib()
        n_bodies = np.ascontiguousarray(n_bodies, dtype=np.uint64)
        state = np.asarray(state, dtype=np.float64)
        nsys = n_bodies.shape[0]
        if state.shape[0] != int(n_bodies.sum()):
            raise ValueError("state has %d rows for %d bodies" % (state.shape[0], int(n_bodies.sum())))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (nsys,)))
        parts = (_abi.Particle * state.shape[0])()
        buf = np.ctypeslib.as_array(ctypes.cast(parts, ctypes.POINTER(ctypes.c_double)),
                                    shape=(state.shape[0] * ctypes.sizeof(_abi.Particle) // 8,))
        buf = buf.reshape(state.shape[0], -1)
        buf[:, :min(11, state.shape[1])] = state[:, :11]
        self._n = state.shape[0]
        self.n_systems = nsys
        self.n_bodies = n_bodies.astype(np.int64)
        self.first = np.concatenate([[0], np.cumsum(self.n_bodies)[:-1]])
        self._h = L.reb_amd_ensemble_create(nsys, n_bodies.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), parts,
                                            dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), G, int(safe_mode),
                                            _abi.GRAVITY[gravity])
        if not self._h:
            raise RebError("reb_amd_ensemble_create failed: " + L.reb_amd_last_error().decode())
        if coordinates != "democraticheliocentric" or corrector or corrector2:
            if L.reb_amd_ensemble_set_whfast(self._h, _abi.WHFAST_COORDINATES[coordinates], int(corrector),
                                             int(corrector2)):
                msg = L.reb_amd_last_error().decode()
                self.free()
                raise RebError(msg)
        if kernel:
            if L.reb_amd_ensemble_set_whfast_kernel(self._h, int(kernel)):
                msg = L.reb_amd_last_error().decode()
                self.free()
                raise RebError(msg)

    def free(self):
        if getattr(self, "_h", None):
            lib().reb_amd_ensemble_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def steps(self, n, wait=True):
        _check(lib().reb_amd_ensemble_steps(self._h, int(n)), "reb_amd_ensemble_steps", self._h)
        if wait:
            self.wait()

    def wait(self):
        _check(lib().reb_amd_ensemble_wait(self._h), "reb_amd_ensemble_wait", self._h)

    @property
    def steps_done(self):
        return int(lib().reb_amd_ensemble_steps_done(self._h))

    def state(self):
        """(bodies x 10) x y z vx vy vz ax ay az m, creation order."""
        parts = (_abi.Particle * self._n)()
        _check(lib().reb_amd_ensemble_get_particles(self._h, parts), "reb_amd_ensemble_get_particles", self._h)
        buf = np.ctypeslib.as_array(ctypes.cast(parts, ctypes.POINTER(ctypes.c_double)),
                                    shape=(self._n * ctypes.sizeof(_abi.Particle) // 8,))
        return buf.reshape(self._n, -1)[:, :10].copy()

    def init_megno(self, var0):
        """MEGNO for every system from (bodies x >=7) initial variational
        particles x y z vx vy vz ... m (columns as state(): m is column 9 when
        given 10 columns)."""
        var0 = np.asarray(var0, dtype=np.float64)
        parts = (_abi.Particle * self._n)()
        for i, row in enumerate(var0):
            p = parts[i]
            p.x, p.y, p.z, p.vx, p.vy, p.vz = (float(v) for v in row[:6])
            p.m = float(row[9]) if var0.shape[1] >= 10 else 0.
        _check(lib().reb_amd_ensemble_init_megno(self._h, parts), "reb_amd_ensemble_init_megno", self._h)

    def megno(self):
        """(megno, lyapunov) arrays, one value per system."""
        m = np.zeros(self.n_systems)
        ly = np.zeros(self.n_systems)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().reb_amd_ensemble_megno(self._h, m.ctypes.data_as(dp), ly.ctypes.data_as(dp)),
               "reb_amd_ensemble_megno", self._h)
        return m, ly

    def var_state(self):
        """(bodies x 10) variational particles x y z vx vy vz ax ay az m."""
        parts = (_abi.Particle * self._n)()
        _check(lib().reb_amd_ensemble_get_var_particles(self._h, parts), "reb_amd_ensemble_get_var_particles",
               self._h)
        buf = np.ctypeslib.as_array(ctypes.cast(parts, ctypes.POINTER(ctypes.c_double)),
                                    shape=(self._n * ctypes.sizeof(_abi.Particle) // 8,))
        return buf.reshape(self._n, -1)[:, :10].copy()

    def system(self, s, st=None):
        st = self.state() if st is None else st
        return st[self.first[s]:self.first[s] + self.n_bodies[s]]

    def energy(self):
        out = np.zeros(self.n_systems)
        _check(lib().reb_amd_ensemble_energy(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "reb_amd_ensemble_energy", self._h)
        return out

    def flags(self):
        out = np.zeros(self.n_systems, dtype=np.uint32)
        _check(lib().reb_amd_ensemble_flags(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))),
               "reb_amd_ensemble_flags", self._h)
        return out

    def set_launch(self, grid=0, chunk=0):
        _check(lib().reb_amd_ensemble_set_launch(self._h, int(grid), int(chunk)), "reb_amd_ensemble_set_launch",
               self._h)

    def set_packing(self, tasks=0):
        """About `tasks` wave tasks (0: densest); returns the task count."""
        n = lib().reb_amd_ensemble_set_packing(self._h, int(tasks))
        if n < 0:
            raise RebError("reb_amd_ensemble_set_packing: " + lib().reb_amd_ensemble_error(self._h).decode())
        return n

    @property
    def launch(self):
        g, c, t, m = ctypes.c_int(), ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
        lib().reb_amd_ensemble_get_launch(self._h, ctypes.byref(g), ctypes.byref(c), ctypes.byref(t), ctypes.byref(m))
        return {"grid": g.value, "chunk": c.value, "tasks": t.value, "max_resident": m.value}

    def tune(self, probe_steps=200):
        best = ctypes.c_double()
        _check(lib().reb_amd_ensemble_tune(self._h, int(probe_steps), ctypes.byref(best)), "reb_amd_ensemble_tune",
               self._h)
        return best.value, self.launch
