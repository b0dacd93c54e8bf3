"""MI355X force-and-step engine behind REBOUND's C ABI.

The product is librebound_amd.so (HIP kernels for gfx950 + a C/C++ host
runtime) with the drop-in interface of include/rebound.h and the additions of
include/rebound_amd.h.  This package loads it through ctypes, the way
REBOUND's own Python package loads librebound, and offers a thin `Sim`
wrapper used by the tests and the benchmark.

There is no CPU fallback: stepping a simulation without a usable HIP device
records an error on the simulation (which `Sim` raises).
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import (BOUNDARY, COLLISION, GRAVITY, PRECISION, SIZE_MAX, STATUS, WHFAST_COORDINATES, Collision,
                   LeapfrogState, Particle, Simulation, Vec6d, WHFastState)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "librebound_amd.so")
_LIB = None

PARTICLE_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                           ("ax", "<f8"), ("ay", "<f8"), ("az", "<f8"), ("m", "<f8"), ("r", "<f8"),
                           ("name", "<u8"), ("ap", "<u8"), ("sim", "<u8")])
STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "r")


class RebError(RuntimeError):
    pass


def lib():
    """The loaded library (built by `python -m rebound_amd.build`)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("librebound_amd.so is not built: run `python -m rebound_amd.build`")
        handle = ctypes.CDLL(LIBPATH, mode=ctypes.RTLD_LOCAL)
        _abi.declare(handle)
        size = handle.reb_simulation_struct_size()
        if size != ctypes.sizeof(Simulation):
            raise RuntimeError("ABI mismatch: struct reb_simulation is %d bytes in the library, %d in the mirror"
                               % (size, ctypes.sizeof(Simulation)))
        _LIB = handle
    return _LIB


def device_count():
    return lib().reb_amd_device_count()


def device_name():
    buf = ctypes.create_string_buffer(256)
    if lib().reb_amd_device_name(buf, 256) != 0:
        return None
    return buf.value.decode()


class Sim:
    """Owns one `struct reb_simulation*` of the library."""

    def __init__(self, save_messages=True, _ptr=None):
        self.lib = lib()
        self.ptr = _ptr if _ptr is not None else self.lib.reb_simulation_create()
        self._keep = []          # ctypes callbacks must outlive the simulation
        if save_messages:
            self.s.save_messages = 1

    @classmethod
    def from_file(cls, path, snapshot=0):
        """reb_simulation_create_from_file: a snapshot of a binary file or
        simulation archive (negative snapshots count from the end)."""
        ptr = lib().reb_simulation_create_from_file(os.fsencode(path), int(snapshot))
        if not ptr:
            raise RebError("cannot read %s (%s)" % (path, lib().reb_amd_last_error().decode()))
        return cls(save_messages=True, _ptr=ptr)

    @classmethod
    def wrap(cls, address):
        """A Sim owning a simulation the library created elsewhere (an address
        or c_void_p), e.g. by reb_simulation_create_from_simulationarchive."""
        addr = address.value if isinstance(address, ctypes.c_void_p) else address
        if not addr:
            raise RebError("NULL simulation (%s)" % lib().reb_amd_last_error().decode())
        return cls(save_messages=True, _ptr=ctypes.cast(ctypes.c_void_p(addr), ctypes.POINTER(_abi.Simulation)))

    def save(self, path):
        """reb_simulation_save_to_file: a new file, or a snapshot appended to an archive."""
        self.lib.reb_simulation_save_to_file(self.ptr, os.fsencode(path))
        self.check()

    # ------------------------------------------------------------ plumbing
    @property
    def s(self):
        return self.ptr.contents

    def free(self):
        if self.ptr:
            self.lib.reb_simulation_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def messages(self):
        out = []
        if not self.s.messages:
            return out
        buf = (ctypes.c_char * 256)()
        while self.lib.reb_pop_message(self.s.messages, buf):
            raw = bytes(buf.value)
            out.append((chr(raw[0]), raw[1:].decode(errors="replace")))
        return out

    def check(self):
        """Raise on pending error messages (as REBOUND's Python layer does)."""
        msgs = self.messages()
        errs = [m for t, m in msgs if t == "e"]
        if errs:
            raise RebError("; ".join(errs))
        return msgs

    # ------------------------------------------------------------ configuration
    def set_integrator(self, name):
        st = self.lib.reb_simulation_set_integrator(self.ptr, name.encode())
        self.check()
        if name == "whfast":
            return ctypes.cast(st, ctypes.POINTER(WHFastState)).contents
        if name == "leapfrog":
            return ctypes.cast(st, ctypes.POINTER(LeapfrogState)).contents
        return st

    @property
    def precision(self):
        return self.lib.reb_amd_get_precision(self.ptr)

    @precision.setter
    def precision(self, mode):
        if isinstance(mode, str):
            mode = PRECISION[mode]
        self.lib.reb_amd_set_precision(self.ptr, mode)
        self.check()

    # ------------------------------------------------------------ particles
    def add(self, **kw):
        p = Particle()
        for k, v in kw.items():
            setattr(p, k, v)
        self.lib.reb_simulation_add(self.ptr, p)

    def add_plummer(self, n, M=1.0, R=1.0):
        self.lib.reb_simulation_add_plummer(self.ptr, n, M, R)

    def particles(self):
        """Structured numpy view of r->particles (valid until the array grows)."""
        n = self.s.N
        if n == 0:
            return np.zeros(0, dtype=PARTICLE_DTYPE)
        buf = (ctypes.c_char * (n * 112)).from_address(ctypes.addressof(self.s.particles.contents))
        return np.frombuffer(buf, dtype=PARTICLE_DTYPE, count=n)

    def state(self):
        """(N, 11) copy: x y z vx vy vz ax ay az m r."""
        p = self.particles()
        return np.stack([p[f] for f in STATE_FIELDS], axis=1).copy()

    def set_state(self, arr):
        """Add N particles from an (N, >=10) array of x y z vx vy vz ax ay az m [r]."""
        arr = np.asarray(arr, dtype=np.float64)
        for row in arr:
            p = Particle()
            for k, f in enumerate(STATE_FIELDS[: arr.shape[1]]):
                setattr(p, f, float(row[k]))
            self.lib.reb_simulation_add(self.ptr, p)
        self.check()

    def acc(self):
        p = self.particles()
        return np.stack([p["ax"], p["ay"], p["az"]], axis=1).copy()

    # ------------------------------------------------------------ stepping
    def steps(self, n):
        self.lib.reb_simulation_steps(self.ptr, n)
        self.check()

    def integrate(self, tmax):
        st = self.lib.reb_simulation_integrate(self.ptr, tmax)
        self.check()
        return st

    def update_acceleration(self):
        self.lib.reb_simulation_update_acceleration(self.ptr)
        self.check()
        return self.acc()

    def energy(self):
        e = self.lib.reb_simulation_energy(self.ptr)
        self.check()
        return e

    def synchronize(self):
        self.lib.reb_simulation_synchronize(self.ptr)
        self.check()

    def move_to_com(self):
        self.lib.reb_simulation_move_to_com(self.ptr)

    # ------------------------------------------------------------ variational sets, MEGNO
    def add_variation(self, order=1, testparticle=-1, first_a=0, first_b=0):
        if order == 1:
            i = self.lib.reb_simulation_add_variation_1st_order(self.ptr, testparticle)
        else:
            i = self.lib.reb_simulation_add_variation_2nd_order(self.ptr, testparticle, first_a, first_b)
        self.check()
        return i

    def init_megno(self, seed=None):
        if seed is None:
            self.lib.reb_simulation_init_megno(self.ptr)
        else:
            self.lib.reb_simulation_init_megno_seed(self.ptr, seed)
        self.check()

    def megno(self):
        return self.lib.reb_simulation_megno(self.ptr)

    def lyapunov(self):
        return self.lib.reb_simulation_lyapunov(self.ptr)

    def var_particles(self):
        """Structured numpy view of r->particles_var (valid until the array grows)."""
        n = self.s.N_var
        if n == 0:
            return np.zeros(0, dtype=PARTICLE_DTYPE)
        buf = (ctypes.c_char * (n * 112)).from_address(ctypes.addressof(self.s.particles_var.contents))
        return np.frombuffer(buf, dtype=PARTICLE_DTYPE, count=n)

    def var_configs(self):
        n = self.s.N_var_config
        if n == 0:
            return []
        arr = ctypes.cast(self.s.var_config, ctypes.POINTER(_abi.VarConfig))
        return [arr[i] for i in range(n)]


from .ensemble import Ensemble, planetary_systems  # noqa: E402
from .serve import Server  # noqa: E402
from .workloads import shearing_sheet  # noqa: E402
