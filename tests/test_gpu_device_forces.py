"""Device-side additional forces: the reference's additional_forces hook
(rebound.h:403, simulation.c:640-689) with the callback's work on the GPU.

A host additional_forces callback and the built-in device drag apply the same
arithmetic, a -= gamma v after gravity, so the two runs agree bit for bit."""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

GAMMA = 0.37


def _sim(gpu, state):
    sim = gpu.Sim()
    sim.s.G = 1.0
    sim.s.softening = 0.01
    sim.s.dt = 1e-3
    sim.set_state(state)
    sim.set_integrator("leapfrog")
    sim.precision = "exact"
    return sim


def test_device_drag_equals_host_callback(gpu):
    st = golden("plummer1024_seed1_state.npy")[:300]
    host = _sim(gpu, st)

    def drag(rp):
        p = host.particles()
        for a, v in (("ax", "vx"), ("ay", "vy"), ("az", "vz")):
            p[a] -= GAMMA * p[v]

    cb = gpu._abi.HEARTBEAT(drag)
    host._keep.append(cb)
    host.s.additional_forces = ctypes.cast(cb, ctypes.c_void_p)
    host.steps(40)

    dev = _sim(gpu, st)
    gamma = ctypes.c_double(GAMMA)
    dev._keep.append(gamma)
    fn = ctypes.cast(gpu.lib().reb_amd_device_force_drag, ctypes.c_void_p)
    gpu.lib().reb_amd_set_device_forces(dev.ptr, fn, ctypes.cast(ctypes.byref(gamma), ctypes.c_void_p))
    gpu.lib().reb_amd_set_resident(dev.ptr, 1)
    dev.steps(40)
    gpu.lib().reb_amd_set_resident(dev.ptr, 0)
    assert np.array_equal(dev.state()[:, :9], host.state()[:, :9])
    # and the drag did act: kinetic energy fell against the undamped run
    free = _sim(gpu, st)
    free.steps(40)
    assert not np.array_equal(free.state()[:, 3:6], dev.state()[:, 3:6])


def test_device_forces_rejected_with_whfast(gpu):
    st = golden("whfast_jac16_init.npy")[2:5]
    sim = gpu.Sim()
    sim.set_state(st)
    sim.s.dt = 0.01
    sim.set_integrator("whfast")
    gamma = ctypes.c_double(GAMMA)
    fn = ctypes.cast(gpu.lib().reb_amd_device_force_drag, ctypes.c_void_p)
    gpu.lib().reb_amd_set_device_forces(sim.ptr, fn, ctypes.cast(ctypes.byref(gamma), ctypes.c_void_p))
    with pytest.raises(gpu.RebError, match="device forces are not supported with WHFast"):
        sim.steps(1)
