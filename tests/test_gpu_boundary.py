"""Boundaries during integration against the compiled reference (OpenMP
build): OPEN removes bodies that leave the box (boundary.c:35-77), PERIODIC
wraps them and adds ghost-box images to the force (boundary.c:115-141,
gravity.c:184-232).  Leapfrog, EXACT precision: the states are bit-identical."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def boxed(gpu, boundary, gravity, root, nr, init):
    sim = gpu.Sim()
    s = sim.s
    s.boundary = gpu.BOUNDARY[boundary]
    s.gravity = gpu.GRAVITY[gravity]
    s.root_size = root
    s.N_root_x = nr
    s.N_root_y = nr
    s.N_root_z = 1
    if boundary == "periodic":
        s.N_ghost_x = 1
        s.N_ghost_y = 1
    s.softening = 0.1
    s.opening_angle2 = 0.25
    s.dt = 0.01
    sim.set_integrator("leapfrog")
    sim.set_state(init)
    sim.precision = "exact"
    return sim


def test_open_boundary_removes_bodies_bitwise(gpu):
    sim = boxed(gpu, "open", "basic", 10.0, 1, golden("open200_init.npy"))
    hist = [sim.s.N]
    for _ in range(10):
        sim.steps(10)
        hist.append(sim.s.N)
    assert hist == list(golden("open200_nhist.npy"))
    assert np.array_equal(sim.state()[:, :6], golden("open200_final.npy")[:, :6])


@pytest.mark.parametrize("gravity,steps", [("basic", 50), ("tree", 20)])
def test_periodic_leapfrog_bitwise(gpu, gravity, steps):
    sim = boxed(gpu, "periodic", gravity, 4.0, 3, golden("periodic300_init.npy"))
    sim.steps(steps)
    assert np.array_equal(sim.state()[:, :6], golden("periodic300_final_%s.npy" % gravity)[:, :6])


@pytest.mark.parametrize("k", [0, 1, 2])
def test_shear_wrap_bitwise(gpu, k):
    """The SHEAR wrap on the device (boundary.c:78-114) against the
    reference's reb_boundary_check on bodies up to 2.7 box widths outside the
    box, at three times: the integrator "none" with dt = 0 leaves t in place,
    so one step applies the wrap alone."""
    t = float(golden("shear_wrap_t.npy")[k])
    sim = gpu.Sim()
    s = sim.s
    s.boundary = gpu.BOUNDARY["shear"]
    s.OMEGA = 1.0
    s.root_size = 50.0
    s.N_root_x, s.N_root_y, s.N_root_z = 2, 2, 1
    sim.set_integrator("none")
    s.t = t
    s.dt = 0.0
    # The generator moved the bodies outside the box by writing r->particles
    # directly (reb_simulation_add refuses such bodies): add them inside,
    # then write the same positions into the host array.
    st = golden("shear_wrap_in.npy")
    inside = st.copy()
    inside[:, :3] = 0.0
    sim.set_state(inside)
    p = sim.particles()
    for k_col, f in enumerate(("x", "y", "z", "vx", "vy", "vz")):
        p[f][:] = st[:, k_col]
    sim.lib.reb_amd_mark_host_modified(sim.ptr)
    sim.steps(1)
    assert s.t == t
    assert np.array_equal(sim.state()[:, :6], golden("shear_wrap_out.npy")[k][:, :6])
    sim.free()
