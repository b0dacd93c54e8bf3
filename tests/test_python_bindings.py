"""The C functions the reference's Python package calls through ctypes.

tests/golden/python_bindings.txt lists every clibrebound.reb_* name in the
package's modules (written by tests/golden/list_python_bindings.py from the
package's source text).  Each is either exported by this library, with the
reference's signature, or listed below as outside the force-and-step path
with the reason; for those, ctypes' lookup fails naming the function, so a
script that needs one stops at that call with the function's name.  Nothing
of the package is imported or run here.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

# name -> why this engine does not provide it
NOT_PROVIDED = {}
for _n in ("reb_display_copy_data", "reb_display_init_data", "reb_display_prepare_data",
           "reb_simulation_output_screenshot", "reb_simulation_start_server", "reb_simulation_stop_server"):
    NOT_PROVIDED[_n] = "visualisation server and screenshots"
for _n in ("reb_integrator_bs_update_particles", "reb_integrator_ias15_timescale", "reb_ode_create"):
    NOT_PROVIDED[_n] = "integrators other than leapfrog, SEI and WHFast, and user ODEs"
for _n in ("reb_rotation_identity", "reb_rotation_init_angle_axis", "reb_rotation_init_from_to",
           "reb_rotation_init_orbit", "reb_rotation_init_to_new_axes", "reb_rotation_inverse", "reb_rotation_mul",
           "reb_rotation_normalize", "reb_rotation_to_orbital", "reb_simulation_irotate", "reb_particle_irotate",
           "reb_vec3d_cross", "reb_vec3d_irotate", "reb_vec3d_normalize"):
    NOT_PROVIDED[_n] = "quaternion rotations"
for _n in ("reb_particle_set_name", "reb_simulation_get_particle_by_name", "reb_simulation_remove_particle_by_name"):
    NOT_PROVIDED[_n] = "particle names"
NOT_PROVIDED["reb_simulation_add_fmt"] = "particles from format strings (the package's sim.add(string) form)"
NOT_PROVIDED["reb_particle_from_pal"] = "Pal (2009) orbital elements as input"
NOT_PROVIDED["reb_frequency_analysis"] = "frequency analysis"


def bound_names():
    rows = []
    with open(os.path.join(GOLDEN, "python_bindings.txt")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            module, name = line.split()
            rows.append((module, name))
    return rows


def exported(amd):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", amd.LIBPATH], capture_output=True, text=True, check=True)
    return {ln.split()[-1] for ln in out.stdout.splitlines() if ln.strip()}


def test_every_bound_function_is_exported_or_declared_out_of_scope(amd):
    names = {n for _, n in bound_names()}
    assert len(names) > 80
    have = exported(amd)
    missing = sorted(n for n in names if n not in have and n not in NOT_PROVIDED)
    assert not missing, "bound by the Python package, neither exported nor declared out of scope: %s" % missing
    # the out-of-scope list names only functions the package binds and this library lacks
    assert not sorted(n for n in NOT_PROVIDED if n in have)
    assert not sorted(n for n in NOT_PROVIDED if n not in names)


def test_force_and_step_calls_are_exported(amd):
    """The calls of the package's Simulation on the force-and-step path."""
    have = exported(amd)
    for n in ("reb_simulation_create", "reb_simulation_free", "reb_simulation_add", "reb_simulation_integrate",
              "reb_simulation_steps", "reb_simulation_synchronize", "reb_simulation_set_integrator",
              "reb_simulation_energy", "reb_simulation_move_to_com", "reb_simulation_com_range",
              "reb_collision_resolve_hardsphere", "reb_collision_resolve_merge", "reb_collision_resolve_halt",
              "reb_simulation_save_to_file", "reb_simulationarchive_create_from_file_with_messages",
              "reb_simulation_init_from_simulationarchive_with_messages", "reb_simulation_copy_with_messages",
              "reb_binarydata_simulation_to_stream", "reb_simulation_diff", "reb_pop_message",
              "reb_simulation_struct_size", "reb_orbit_from_particle_err", "reb_particle_from_orbit_err"):
        assert n in have, n


@pytest.mark.parametrize("name", sorted(NOT_PROVIDED))
def test_out_of_scope_call_fails_naming_the_function(amd, name):
    with pytest.raises(AttributeError) as e:
        getattr(amd.lib(), name)
    assert name in str(e.value)


class Orbit(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("d", "v", "h", "P", "n", "a", "e", "inc", "Omega", "omega", "pomega",
                                               "f", "M", "l", "theta", "T", "rhill", "pal_h", "pal_k", "pal_ix",
                                               "pal_iy", "hx", "hy", "hz", "ex", "ey", "ez")]


def test_orbit_elements_round_trip(amd):
    """reb_orbit_from_particle_err inverts reb_particle_from_orbit_err (tools.c:940-1208)
    for bound orbits, inclined and planar, prograde and retrograde; the mean
    anomaly obeys Kepler's equation; err 1 for a massless primary."""
    L = amd.lib()
    P = amd._abi.Particle
    L.reb_orbit_from_particle_err.restype = Orbit
    L.reb_orbit_from_particle_err.argtypes = [ctypes.c_double, P, P, ctypes.POINTER(ctypes.c_int)]
    L.reb_M_to_E.restype = ctypes.c_double
    L.reb_M_to_E.argtypes = [ctypes.c_double, ctypes.c_double]
    star = P()
    star.m = 1.0
    for (a, e, inc, Om, om, f) in ((1.0, 0.1, 0.3, 0.4, 1.1, 2.0), (2.5, 0.6, 2.8, 5.0, 0.2, 4.0),
                                   (0.7, 0.02, 0.0, 0.0, 0.5, 1.0), (1.3, 0.3, 1.2, 3.0, 2.0, 0.1)):
        err = ctypes.c_int(0)
        p = L.reb_particle_from_orbit_err(1.0, star, 1e-3, a, e, inc, Om, om, f, ctypes.byref(err))
        assert err.value == 0
        o = L.reb_orbit_from_particle_err(1.0, p, star, ctypes.byref(err))
        assert err.value == 0
        assert abs(o.a - a) < 1e-12 * a and abs(o.e - e) < 1e-12 and abs(o.inc - inc) < 1e-12
        if inc > 0:
            assert abs(math.remainder(o.Omega - Om, 2 * math.pi)) < 1e-11
            assert abs(math.remainder(o.omega - om, 2 * math.pi)) < 1e-10
        assert abs(math.remainder(o.f - f, 2 * math.pi)) < 1e-10
        E = L.reb_M_to_E(e, o.M)
        assert abs(math.remainder(E - e * math.sin(E) - o.M, 2 * math.pi)) < 1e-13
        assert abs(o.P - 2 * math.pi * math.sqrt(a ** 3 / (1.0 + 1e-3))) < 1e-10 * o.P
    none = P()
    err = ctypes.c_int(0)
    o = L.reb_orbit_from_particle_err(1.0, star, none, ctypes.byref(err))
    assert err.value == 1 and math.isnan(o.a)


def test_hash_and_serialized_particle_data(amd):
    L = amd.lib()
    L.reb_hash.restype = ctypes.c_uint32
    # djb2 of "" and of "AU"
    assert L.reb_hash(b"") == 5381
    assert L.reb_hash(b"AU") == ((5381 * 33 + ord("A")) * 33 + ord("U")) & 0xFFFFFFFF
    sim = amd.Sim()
    rng = np.random.default_rng(3)
    st = np.zeros((7, 11))
    st[:, :6] = rng.normal(size=(7, 6))
    st[:, 9] = rng.uniform(1, 2, 7)
    st[:, 10] = rng.uniform(0.1, 0.2, 7)
    sim.set_state(st)
    m = np.zeros(7)
    xyz = np.zeros((7, 3))
    six = np.zeros((7, 6))
    dp = ctypes.POINTER(ctypes.c_double)
    L.reb_simulation_get_serialized_particle_data(sim.ptr, m.ctypes.data_as(dp), None, xyz.ctypes.data_as(dp), None,
                                                  six.ctypes.data_as(dp))
    assert np.array_equal(m, st[:, 9]) and np.array_equal(xyz, st[:, :3]) and np.array_equal(six, st[:, :6])
    six[:, 3:] *= -1.0
    L.reb_simulation_set_serialized_particle_data(sim.ptr, None, None, None, None, six.ctypes.data_as(dp))
    assert np.array_equal(sim.state()[:, 3:6], -st[:, 3:6])
    sim.free()
