"""Host side of the drop-in API that needs no GPU: input generators (bitwise
against the reference's streams), bookkeeping, registry and error paths."""
import ctypes

import numpy as np
import pytest

from conftest import golden


def test_random_streams_bitwise(amd):
    sim = amd.Sim()
    sim.s.rand_seed = 7
    L, p = amd.lib(), sim.ptr
    u = [L.reb_random_uniform(p, -1.0, 3.0) for _ in range(256)]
    n = [L.reb_random_normal(p, 2.0) for _ in range(256)]
    w = [L.reb_random_powerlaw(p, 1.0, 4.0, -3.0) for _ in range(256)]
    r = [L.reb_random_rayleigh(p, 0.02) for _ in range(256)]
    assert np.array_equal(np.array([u, n, w, r]), golden("random_streams_seed7.npy"))


def test_plummer_bitwise(amd):
    # reb_simulation_add_plummer + move_to_com reproduce the reference's state exactly
    sim = amd.Sim()
    sim.s.rand_seed = 1
    sim.add_plummer(1024, 1.0, 1.0)
    sim.move_to_com()
    st = sim.state()
    ref = golden("plummer1024_seed1_state.npy")
    assert np.array_equal(st[:, [0, 1, 2, 3, 4, 5, 9]], ref[:, [0, 1, 2, 3, 4, 5, 9]])


def test_whfast_inputs_bitwise(amd, orc):
    # reb_particle_from_orbit with the reference's element stream
    init = golden("whfast_dh32_init.npy")
    meta = golden("whfast_dh32_meta.npy")
    sim = amd.Sim()
    sim.s.rand_seed = 2024
    L, p = amd.lib(), sim.ptr
    import math
    for s in range(4):
        K = int(meta[s, 0])
        star = amd.Particle(m=1.0)
        bodies = [star]
        a = 1.0
        for k in range(K):
            if k > 0:
                ratio = math.exp(L.reb_random_uniform(p, math.log(1.3), math.log(2.5)))
                a *= ratio ** (2.0 / 3.0)
            m = math.exp(L.reb_random_uniform(p, math.log(1e-6), math.log(1e-3)))
            e = L.reb_random_rayleigh(p, 0.02)
            inc = L.reb_random_rayleigh(p, 0.01)
            Om = L.reb_random_uniform(p, 0.0, 2 * math.pi)
            om = L.reb_random_uniform(p, 0.0, 2 * math.pi)
            f = L.reb_random_uniform(p, 0.0, 2 * math.pi)
            bodies.append(L.reb_particle_from_orbit(1.0, star, m, a, e, inc, Om, om, f))
        sub = amd.Sim()
        for b in bodies:
            L.reb_simulation_add(sub.ptr, b)
        sub.move_to_com()
        st = sub.state()
        assert np.array_equal(st[:, :6], init[s, : K + 1, :6]), s


def test_from_orbit_errors(amd):
    L = amd.lib()
    err = ctypes.c_int(0)
    p = L.reb_particle_from_orbit_err(1.0, amd.Particle(m=1.0), 0.0, 1.0, 1.0, 0, 0, 0, 0, ctypes.byref(err))
    assert err.value == 1 and np.isnan(p.x)
    L.reb_particle_from_orbit_err(1.0, amd.Particle(m=0.0), 0.0, 1.0, 0.1, 0, 0, 0, 0, ctypes.byref(err))
    assert err.value == 6


def test_add_remove_and_box(amd):
    sim = amd.Sim()
    for i in range(10):
        sim.add(x=float(i), m=1.0)
    assert sim.s.N == 10 and sim.s.N_allocated == 16
    sim.s.N_active = 6
    assert amd.lib().reb_simulation_remove_particle(sim.ptr, 2) == 0
    assert sim.s.N == 9 and sim.s.N_active == 5
    assert list(sim.state()[:4, 0]) == [0.0, 1.0, 3.0, 4.0]
    assert amd.lib().reb_simulation_remove_particle(sim.ptr, 99) == 1
    with pytest.raises(amd.RebError):
        sim.check()
    # adding outside the box of a periodic simulation is refused (test_boundary.py:57-73)
    box = amd.Sim()
    box.s.boundary = amd.BOUNDARY["periodic"]
    box.s.root_size = 2.0
    box.add(x=3.0)
    assert box.s.N == 0
    with pytest.raises(amd.RebError):
        box.check()


def test_integrator_registry(amd):
    sim = amd.Sim()
    assert sim.s.integrator.name == b"ias15"
    lf = sim.set_integrator("leapfrog")
    assert lf.order == 2
    wh = sim.set_integrator("whfast")
    assert wh.safe_mode == 1 and wh.coordinates == amd.WHFAST_COORDINATES["jacobi"]
    amd.lib().reb_simulation_set_integrator(sim.ptr, b"nonexistent")
    with pytest.raises(amd.RebError, match="Integrator not found"):
        sim.check()


def test_ghostbox_shear_matches_oracle(amd, orc):
    sim = amd.Sim()
    sim.s.boundary = amd.BOUNDARY["shear"]
    sim.s.root_size = 50.0
    sim.s.N_root_x = 2
    sim.s.N_root_y = 2
    sim.s.OMEGA = 1.0
    for t in (0.0, 0.37, 12.5):
        sim.s.t = t
        for i in (-2, -1, 0, 1, 2):
            g = amd.lib().reb_boundary_get_ghostbox(sim.ptr, i, 1, 0)
            ref = np.zeros(6)
            orc.lib().orc_ghostbox_shear(1.0, 100.0, 100.0, 50.0, t, i, 1, 0, ref)
            assert [g.x, g.y, g.z, g.vx, g.vy, g.vz] == list(ref)


def test_no_device_is_an_error_not_a_fallback(amd):
    if amd.device_count() > 0:
        pytest.skip("a device is present")
    sim = amd.Sim()
    sim.add(m=1.0)
    sim.add(x=1.0, m=1.0)
    sim.set_integrator("leapfrog")
    with pytest.raises(amd.RebError, match="No HIP device"):
        sim.steps(1)


def test_hardsphere_resolver_bitwise(amd, orc):
    """collision.c:573-665 on random pairs: the host resolver equals the oracle
    (and both use sincos exactly where the gcc-built reference does)."""
    import ctypes
    rng = np.random.default_rng(11)
    L = amd.lib()
    cor = amd._abi.COR(lambda r, v: 0.5)
    sim = amd.Sim()
    sim.s.coefficient_of_restitution = cor
    sim.s.minimum_collision_velocity = 1e-3
    sim.add(m=1.0)
    sim.add(m=1.0)
    P = sim.s.particles
    for k in range(20000):
        v = rng.normal(size=(2, 6))
        m = rng.uniform(1, 10, 2)
        rr = rng.uniform(1, 4, 2)
        st = np.zeros((2, 11))
        for i in range(2):
            P[i].x, P[i].y, P[i].z = v[i, 0], v[i, 1], v[i, 2] * 0.3
            P[i].vx, P[i].vy, P[i].vz = v[i, 3], v[i, 4], v[i, 5]
            P[i].m, P[i].r = m[i], rr[i]
            st[i] = [P[i].x, P[i].y, P[i].z, P[i].vx, P[i].vy, P[i].vz, 0, 0, 0, m[i], rr[i]]
        gb = np.array([0.5, -0.2, 0.0, 0.0, -0.75, 0.0])
        c = amd.Collision(p1=0, p2=1)
        c.gb = amd.Vec6d(*gb)
        L.reb_collision_resolve_hardsphere(sim.ptr, c)
        logn = ctypes.c_long(0)
        orc.lib().orc_hardsphere(st, 0, 1, gb, 0.5, 1e-3, ctypes.byref(logn))
        got = [P[0].vx, P[0].vy, P[0].vz, P[1].vx, P[1].vy, P[1].vz]
        assert got == list(st[:, 3:6].ravel()), k


def test_from_orbit_bitwise_vs_oracle(amd, orc):
    rng = np.random.default_rng(12)
    L = amd.lib()
    prim = amd.Particle(x=0.1, y=-0.2, z=0.05, vx=0.01, vy=0.02, vz=-0.03, m=1.3)
    pr = np.array([0.1, -0.2, 0.05, 0.01, 0.02, -0.03, 1.3])
    for k in range(20000):
        a = rng.uniform(0.5, 5)
        e = rng.uniform(0, 0.9)
        el = rng.uniform(0, 2 * np.pi, 4)
        p = L.reb_particle_from_orbit(1.0, prim, 1e-3, a, e, el[0] / 2, el[1], el[2], el[3])
        out = np.zeros(6)
        assert orc.lib().orc_particle_from_orbit(1.0, pr, 1e-3, a, e, el[0] / 2, el[1], el[2], el[3], out) == 0
        assert [p.x, p.y, p.z, p.vx, p.vy, p.vz] == list(out), k


def test_merge_resolver_bitwise(amd):
    """collision.c:674-737 against the compiled reference: merged body,
    outcome and energy_offset (with and without track_energy_offset)."""
    L = amd.lib()
    xin = golden("merge_in.npy")
    xout = golden("merge_out.npy")
    meta = golden("merge_meta.npy")
    res = golden("merge_res.npy")
    sim = amd.Sim()
    for _ in range(3):
        sim.add(m=1.0)
    P = sim.s.particles
    names = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "r")
    for c in range(xin.shape[0]):
        for i in range(3):
            for f, v in zip(names, xin[c, i]):
                setattr(P[i], f, v)
        sim.s.track_energy_offset = int(meta[c, 2])
        sim.s.N_active = int(meta[c, 3]) if meta[c, 3] >= 0 else 2 ** 64 - 1
        sim.s.energy_offset = 0.25
        sim.s.G = res[c, 2]
        out = L.reb_collision_resolve_merge(sim.ptr, amd.Collision(p1=int(meta[c, 0]), p2=int(meta[c, 1])))
        got = np.array([[getattr(P[i], f) for f in names] for i in range(3)])
        assert out == int(res[c, 0]), c
        assert np.array_equal(got, xout[c]), c
        assert sim.s.energy_offset == res[c, 1], c
