"""reb_simulation_integrate on the GPU against the compiled reference.

The fixtures come from the reference library's own reb_simulation_integrate
(simulation.c:241-502) on the same initial bodies (tests/golden/gen_golden.c,
gen_integrate): the exact-finish-time shrink of the last step and the restored
dt (simulation.c:318-341, 496-498), the exit statuses ESCAPE and ENCOUNTER of
the heartbeat checks (simulation.c:241-275), NO_PARTICLES, the user hooks
(simulation.c:520-561), the halt resolver and the default resolver
(collision.c:345-350, 667-672), collision outcomes that remove bodies with the
reference's index fix-up (collision.c:351-404; test_collisions.py:104-190),
the merge resolver on a stepping simulation (collision.c:674-737) and a host
integrator registered by name (rebound.c:86-113).  Direct gravity runs in
EXACT precision, so every comparison is bit for bit: bodies, t, dt,
dt_last_done, steps_done, the status, N, collisions_log_n and energy_offset.
"""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

META = ("t", "dt", "dt_last_done", "steps_done", "status", "N", "collisions_log_n", "energy_offset")


def make(gpu, init, integrator, dt, gravity="basic", soft=0.0):
    sim = gpu.Sim()
    sim.s.softening = soft
    sim.set_integrator(integrator)
    sim.s.gravity = gpu.GRAVITY[gravity]
    sim.s.dt = dt
    sim.set_state(golden(init))
    sim.precision = "exact"
    return sim


def replay_seed(gpu, sim, seed, per_body, n):
    """r->rand_seed as the fixture generator left it after drawing the bodies:
    `seed`, advanced by the same sequence of reb_random_* calls (per body, the
    kinds in `per_body`: u uniform, n normal), each a glibc rand_r stream on
    r->rand_seed.  The collision shuffle (collision.c:336-343) continues from
    there."""
    L = gpu.lib()
    sim.s.rand_seed = seed
    for _ in range(n):
        for kind in per_body:
            if kind == "u":
                L.reb_random_uniform(sim.ptr, 0.0, 1.0)
            else:
                L.reb_random_normal(sim.ptr, 1.0)


def check_meta(sim, status, name):
    want = golden(name)
    s = sim.s
    got = (s.t, s.dt, s.dt_last_done, s.steps_done, status, s.N, s.collisions_log_n, s.energy_offset)
    for k, (a, b) in enumerate(zip(got, want)):
        assert a == b, (name, META[k], a, b)


def check_state(sim, name):
    want = golden(name)
    got = sim.state()
    assert got.shape == want.shape, (name, got.shape, want.shape)
    assert np.array_equal(got[:, :6], want[:, :6]), name
    assert np.array_equal(got[:, 9:11], want[:, 9:11]), name


def body_view(rp):
    """(N, 14) float64 view of r->particles inside a callback (112 B per body)."""
    r = rp.contents
    if r.N == 0:
        return np.zeros((0, 14))
    buf = (ctypes.c_double * (14 * r.N)).from_address(ctypes.addressof(r.particles.contents))
    return np.frombuffer(buf, dtype=np.float64).reshape(r.N, 14)


def test_leapfrog_exact_finish_time(gpu):
    """tmax = 0.0123456 with dt = 1e-3: twelve full steps, a last step
    shrunk to the finish time and dt restored; then past tmax without the
    exact finish; then backwards (dt's sign follows the direction)."""
    sim = make(gpu, "integ_lf_init.npy", "leapfrog", 1e-3, soft=0.01)
    st = sim.integrate(0.0123456)
    check_meta(sim, st, "integ_lf_a_meta.npy")
    check_state(sim, "integ_lf_a_state.npy")
    sim.s.exact_finish_time = 0
    st = sim.integrate(0.05)
    check_meta(sim, st, "integ_lf_b_meta.npy")
    check_state(sim, "integ_lf_b_state.npy")
    sim.s.exact_finish_time = 1
    st = sim.integrate(0.0301)
    check_meta(sim, st, "integ_lf_c_meta.npy")
    check_state(sim, "integ_lf_c_state.npy")
    sim.free()


def test_whfast_exact_finish_time(gpu):
    """WHFast in Jacobi coordinates with safe mode off: the exit check
    synchronizes before it shrinks the last step."""
    sim = make(gpu, "integ_wh_init.npy", "whfast", 0.0314)
    sim.set_integrator("whfast").safe_mode = 0
    st = sim.integrate(7.77)
    check_meta(sim, st, "integ_wh_a_meta.npy")
    check_state(sim, "integ_wh_a_state.npy")
    sim.s.exact_finish_time = 0
    st = sim.integrate(20.0)
    check_meta(sim, st, "integ_wh_b_meta.npy")
    check_state(sim, "integ_wh_b_state.npy")
    sim.free()


def test_exit_max_distance_escape(gpu):
    sim = make(gpu, "integ_escape_init.npy", "leapfrog", 1e-3)
    sim.s.exit_max_distance = 1.4
    st = sim.integrate(100.0)
    assert gpu.STATUS[st] == "ESCAPE"
    check_meta(sim, st, "integ_escape_meta.npy")
    check_state(sim, "integ_escape_state.npy")
    sim.free()


def test_exit_min_distance_encounter(gpu):
    sim = make(gpu, "integ_encounter_init.npy", "leapfrog", 1e-3)
    sim.s.exit_min_distance = 0.25
    st = sim.integrate(200.0)
    assert gpu.STATUS[st] == "ENCOUNTER"
    check_meta(sim, st, "integ_encounter_meta.npy")
    check_state(sim, "integ_encounter_state.npy")
    sim.free()


def test_no_particles(gpu):
    sim = gpu.Sim()
    sim.set_integrator("leapfrog")
    st = sim.integrate(1.0)
    assert gpu.STATUS[st] == "NO_PARTICLES"
    assert sim.s.t == 0.0 and sim.s.steps_done == 0
    sim.free()


def test_hooks_heartbeat_pre_post(gpu):
    """The heartbeat sees the bodies after every step; the pre-timestep hook
    kicks body 0 and the post-timestep hook damps every velocity on the host
    copy, and the device continues from what they wrote."""
    sim = make(gpu, "integ_hooks_init.npy", "leapfrog", 1e-3, soft=0.01)
    log = []

    def hb(rp):
        r = rp.contents
        p = r.particles[5]
        log.append((r.t, float(r.N), p.x, p.vy, r.energy_offset))

    def pre(rp):
        rp.contents.particles[0].vx += 1e-7

    def post(rp):
        v = body_view(rp)
        v[:, 3:6] *= 0.9999

    cbs = [gpu._abi.HEARTBEAT(f) for f in (hb, pre, post)]
    sim._keep.extend(cbs)
    sim.s.heartbeat = cbs[0]
    sim.s.pre_timestep_modifications = ctypes.cast(cbs[1], ctypes.c_void_p)
    sim.s.post_timestep_modifications = ctypes.cast(cbs[2], ctypes.c_void_p)
    st = sim.integrate(0.0205)
    check_meta(sim, st, "integ_hooks_meta.npy")
    check_state(sim, "integ_hooks_state.npy")
    assert np.array_equal(np.array(log), golden("integ_hooks_log.npy"))
    sim.free()


@pytest.mark.parametrize("variant", [0, 1])
def test_halt(gpu, variant):
    """Variant 0: DIRECT search with no resolver set (the default halts);
    variant 1: TREE search with reb_collision_resolve_halt.  The integration
    stops with REB_STATUS_COLLISION after the step that found a collision."""
    sim = make(gpu, "integ_halt%d_init.npy" % variant, "leapfrog", 0.01)
    replay_seed(gpu, sim, 31 + variant, "uuunnn", 100)
    s = sim.s
    s.root_size = 10.0
    s.N_root_x = s.N_root_y = s.N_root_z = 1
    s.boundary = gpu.BOUNDARY["periodic"]
    s.collision = gpu.COLLISION["direct" if variant == 0 else "tree"]
    if variant == 1:
        s.collision_resolve = ctypes.cast(gpu.lib().reb_collision_resolve_halt, gpu._abi.RESOLVE)
    st = sim.integrate(10.0)
    assert gpu.STATUS[st] == "COLLISION"
    check_meta(sim, st, "integ_halt%d_meta.npy" % variant)
    check_state(sim, "integ_halt%d_state.npy" % variant)
    sim.free()


@pytest.mark.parametrize("tag", ["both_tree", "both_direct", "one_tree", "one_direct"])
def test_collision_outcomes_remove_bodies(gpu, tag):
    """test_collisions.py:104-190 with deterministic outcomes: removing both
    bodies of every collision empties the simulation after exactly five
    collisions (NO_PARTICLES); cycling the outcomes 1, 2, 3 (remove p1, p2,
    both) leaves fewer than 25 of 50 bodies.  The callback sees the same
    (p1, p2) sequence as the reference's, index fix-up included."""
    one = tag.startswith("one")
    tree = tag.endswith("tree")
    sim = make(gpu, "integ_%s_init.npy" % tag, "leapfrog", 0.001, gravity="tree" if tree else "basic")
    replay_seed(gpu, sim, 77 + (1 if tree else 0) + (2 if one else 0), "uuu", 50 if one else 10)
    s = sim.s
    s.root_size = 50000.0
    s.boundary = gpu.BOUNDARY["open"]
    s.collision = gpu.COLLISION["tree" if tree else "direct"]
    log = []

    def resolve(rp, c):
        rp.contents.collisions_log_n += 1
        out = 1 + len(log) % 3 if one else 3
        log.append((c.p1, c.p2, out))
        return out

    cb = gpu._abi.RESOLVE(resolve)
    sim._keep.append(cb)
    s.collision_resolve = cb
    if one:
        st = sim.integrate(s.dt)
        check_meta(sim, st, "integ_%s_a_meta.npy" % tag)
        st = sim.integrate(2.0 * s.dt)
        assert s.N < 25
        check_state(sim, "integ_%s_state.npy" % tag)
    else:
        st = sim.integrate(1000.0)
        assert gpu.STATUS[st] == "NO_PARTICLES" and s.collisions_log_n == 5
    check_meta(sim, st, "integ_%s_meta.npy" % tag)
    assert np.array_equal(np.array(log, dtype=np.int64).reshape(-1, 3), golden("integ_%s_log.npy" % tag))
    sim.free()


@pytest.mark.parametrize("variant", [0, 1])
def test_merge_while_stepping(gpu, variant):
    """reb_collision_resolve_merge on 300 bodies in a periodic box with one
    ring of ghost boxes, DIRECT (0) and TREE (1) searches, energy offset
    tracked: N after every step (heartbeat log), the merged bodies and the
    energy offset.  The fixtures come from the reference's OpenMP build, whose
    per-target force sums over ghost boxes are the EXACT kernel's order."""
    sim = make(gpu, "integ_merge%d_init.npy" % variant, "leapfrog", 0.01)
    replay_seed(gpu, sim, 41, "uuunnnuu", 300)
    s = sim.s
    s.root_size = 4.0
    s.N_root_x, s.N_root_y, s.N_root_z = 2, 2, 1
    s.N_ghost_x, s.N_ghost_y = 1, 1
    s.boundary = gpu.BOUNDARY["periodic"]
    s.collision = gpu.COLLISION["direct" if variant == 0 else "tree"]
    s.collision_resolve = ctypes.cast(gpu.lib().reb_collision_resolve_merge, gpu._abi.RESOLVE)
    s.track_energy_offset = 1
    log = []

    def hb(rp):
        r = rp.contents
        p = r.particles[0]
        log.append((r.t, float(r.N), p.x, p.vy, r.energy_offset))

    cb = gpu._abi.HEARTBEAT(hb)
    sim._keep.append(cb)
    s.heartbeat = cb
    st = sim.integrate(0.3)
    check_meta(sim, st, "integ_merge%d_meta.npy" % variant)
    check_state(sim, "integ_merge%d_state.npy" % variant)
    assert np.array_equal(np.array(log), golden("integ_merge%d_log.npy" % variant))
    sim.free()


def test_registered_host_integrator(gpu):
    """A symplectic Euler integrator registered by name, stepping r->particles
    on the host and calling reb_simulation_update_acceleration (device
    gravity, EXACT) inside its step."""
    L = gpu.lib()

    def step(rp, state):
        L.reb_simulation_update_acceleration(rp)
        r = rp.contents
        dt = r.dt
        v = body_view(rp)
        v[:, 3:6] += dt * v[:, 6:9]
        v[:, 0:3] += dt * v[:, 3:6]
        r.t += dt
        r.dt_last_done = dt

    cb = gpu._abi.STEP(step)
    integ = gpu._abi.Integrator()
    integ.step = ctypes.cast(cb, ctypes.c_void_p)
    name = b"test_symplectic_euler"
    L.reb_integrator_register(integ, name)
    sim = gpu.Sim()
    sim._keep.append(cb)
    sim.s.softening = 0.01
    L.reb_simulation_set_integrator(sim.ptr, name)
    sim.check()
    sim.s.gravity = gpu.GRAVITY["basic"]
    sim.s.dt = 1e-3
    sim.s.exact_finish_time = 0
    sim.set_state(golden("integ_custom_init.npy"))
    sim.precision = "exact"
    st = sim.integrate(0.01)
    check_meta(sim, st, "integ_custom_meta.npy")
    check_state(sim, "integ_custom_state.npy")
    sim.free()
