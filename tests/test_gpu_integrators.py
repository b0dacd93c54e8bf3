"""GPU parity of the integrators: leapfrog (integrator_leapfrog.c), WHFast
Jacobi and democratic heliocentric coordinates with the symplectic correctors
(integrator_whfast.c) and its Kepler solver, SEI
(integrator_sei.c) with the shear boundary (boundary.c)."""
import ctypes
import math

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def sim_from(amd, state, **fields):
    sim = amd.Sim()
    for k, v in fields.items():
        setattr(sim.s, k, v)
    sim.set_state(state)
    return sim


# ------------------------------------------------------------------ leapfrog
def test_leapfrog_energy_history_exact(gpu):
    st = golden("plummer1024_seed1_state.npy")
    sim = sim_from(gpu, st, softening=0.01, dt=1e-3)
    sim.set_integrator("leapfrog")
    sim.precision = "exact"
    E = [sim.energy()]
    for _ in range(100):
        sim.steps(10)
        E.append(sim.energy())
    assert np.array_equal(np.array(E), golden("plummer1024_seed1_lf_energy.npy"))
    fin = golden("plummer1024_seed1_lf_final.npy")
    assert np.array_equal(sim.state()[:, :9], fin[:, :9])
    assert sim.s.t == pytest.approx(1.0, abs=1e-12)


def test_leapfrog_energy_history_fp64(gpu):
    st = golden("plummer1024_seed1_state.npy")
    ref = golden("plummer1024_seed1_lf_energy.npy")
    sim = sim_from(gpu, st, softening=0.01, dt=1e-3)
    sim.set_integrator("leapfrog")
    E = [sim.energy()]
    for _ in range(100):
        sim.steps(10)
        E.append(sim.energy())
    E = np.array(E)
    # same energy-error history to 1e-10 of |E| over the whole run
    assert np.abs(E - ref).max() <= 1e-10 * abs(ref[0])


@pytest.mark.parametrize("order", [4, 6, 8])
def test_leapfrog_higher_orders(gpu, orc, order):
    # time reversibility: forward then backward returns to the start
    st = golden("plummer1024_seed1_state.npy")[:128].copy()
    sim = sim_from(gpu, st, softening=0.05, dt=1e-3)
    lf = sim.set_integrator("leapfrog")
    lf.order = order
    sim.precision = "exact"
    sim.steps(20)
    sim.s.dt = -1e-3
    sim.steps(20)
    assert np.abs(sim.state()[:, :6] - st[:, :6]).max() < 1e-12


# ------------------------------------------------------------------ Kepler
def test_kepler_vectors_bitwise(gpu):
    inp = golden("kepler_in.npy")
    out = golden("kepler_out.npy")
    s6 = np.ascontiguousarray(inp[:, :6]).copy()
    mu = np.ascontiguousarray(inp[:, 6])
    dt = np.ascontiguousarray(inp[:, 7])
    P = ctypes.POINTER(ctypes.c_double)
    rc = gpu.lib().reb_amd_kepler_solver_batch(s6.ctypes.data_as(P), mu.ctypes.data_as(P), dt.ctypes.data_as(P),
                                               len(mu))
    assert rc == 0
    bad = np.where(~np.all(s6 == out, axis=1))[0]
    assert len(bad) == 0, bad[:10]


def test_kepler_public_helper(gpu):
    p = gpu.Particle(x=1.0, vy=1.0)
    gpu.lib().reb_integrator_whfast_kepler_solver(ctypes.byref(p), 1.0, 0.5 * math.pi, None)
    assert abs(p.x) < 1e-14 and abs(p.y - 1.0) < 1e-14 and abs(p.vx + 1.0) < 1e-14


# ------------------------------------------------------------------ WHFast DH
def _systems():
    init = golden("whfast_dh32_init.npy")
    meta = golden("whfast_dh32_meta.npy")
    dts = golden("whfast_dh32_dt.npy")
    return [(init[s, : meta[s, 1]], dts[s]) for s in range(init.shape[0])]


@pytest.mark.parametrize("safe", [0, 1])
def test_whfast_dh_bitwise(gpu, safe):
    fin = golden("whfast_dh32_safe%d_final.npy" % safe)
    en = golden("whfast_dh32_safe%d_energy.npy" % safe)
    for k, (s0, dt) in enumerate(_systems()):
        sim = sim_from(gpu, s0, dt=dt)
        wh = sim.set_integrator("whfast")
        wh.coordinates = gpu.WHFAST_COORDINATES["democraticheliocentric"]
        wh.safe_mode = safe
        sim.precision = "exact"
        E = [sim.energy()]
        for _ in range(10):
            sim.steps(100)
            E.append(sim.energy())
        n = s0.shape[0]
        assert np.array_equal(sim.state()[:, :6], fin[k, :n, :6]), k
        assert np.array_equal(np.array(E), en[k]), k


def _jacobi_cases():
    init = golden("whfast_jac16_init.npy")
    nb = golden("whfast_jac16_nbodies.npy")
    dts = golden("whfast_jac16_dt.npy")
    first = np.concatenate([[0], np.cumsum(nb)[:-1]])
    return [(init[f:f + n], float(dt)) for f, n, dt in zip(first, nb, dts)]


@pytest.mark.parametrize("mode", range(8))
def test_whfast_jacobi_bitwise(gpu, mode):
    """Jacobi coordinates (the reference's default) with the symplectic
    correctors, COMPENSATED gravity and test particles, per the mode table of
    the fixtures: bit-identical bodies and energies after three
    reb_simulation_steps calls of 100 steps."""
    safe, c1, c2, comp, off, tp = (int(v) for v in golden("whfast_jac16_modes.npy")[mode])
    fin = golden("whfast_jac16_mode%d_final.npy" % mode)
    en = golden("whfast_jac16_mode%d_energy.npy" % mode)
    row = 0
    for k, (s0, dt) in enumerate(_jacobi_cases()):
        n = s0.shape[0]
        sim = sim_from(gpu, s0, dt=dt)
        wh = sim.set_integrator("whfast")
        assert wh.coordinates == gpu.WHFAST_COORDINATES["jacobi"]
        wh.safe_mode = safe
        wh.corrector = c1
        wh.corrector2 = c2
        sim.s.gravity = gpu.GRAVITY["compensated" if comp else "basic"]
        if off and n > off:
            sim.s.N_active = n - off
        sim.s.testparticle_type = tp
        sim.precision = "exact"
        E = [sim.energy()]
        for _ in range(3):
            sim.steps(100)
            E.append(sim.energy())
        assert np.array_equal(sim.state()[:, :6], fin[row:row + n, :6]), (mode, k)
        assert np.array_equal(np.array(E), en[k]), (mode, k)
        row += n


@pytest.mark.parametrize("mode", range(6))
def test_whfast_kernels_bitwise(gpu, mode):
    """The modified-kick, composition and lazy implementer's kernels in
    Jacobi coordinates (with correctors, safe mode and COMPENSATED gravity per
    the mode table): bodies with their accelerations, energies after three
    reb_simulation_steps calls of 100 steps, and r->gravity afterwards (the
    Jacobi gravity for kernels 1 and 3) as the reference leaves them."""
    kern, safe, c1, c2, comp = (int(v) for v in golden("whfast_kern16_modes.npy")[mode])
    fin = golden("whfast_kern16_mode%d_final.npy" % mode)
    en = golden("whfast_kern16_mode%d_energy.npy" % mode)
    grav = golden("whfast_kern16_mode%d_gravity.npy" % mode)
    row = 0
    for k, (s0, dt) in enumerate(_jacobi_cases()):
        n = s0.shape[0]
        sim = sim_from(gpu, s0, dt=dt)
        wh = sim.set_integrator("whfast")
        wh.kernel = kern
        wh.safe_mode = safe
        wh.corrector = c1
        wh.corrector2 = c2
        sim.s.gravity = gpu.GRAVITY["compensated" if comp else "basic"]
        sim.precision = "exact"
        E = [sim.energy()]
        for _ in range(3):
            sim.steps(100)
            E.append(sim.energy())
        assert np.array_equal(sim.state()[:, :9], fin[row:row + n, :9]), (mode, k)
        assert np.array_equal(np.array(E), en[k]), (mode, k)
        assert sim.s.gravity == grav[k]
        row += n


def test_gravity_jacobi_update_acceleration(gpu, orc):
    """reb_simulation_update_acceleration with REB_GRAVITY_JACOBI
    (gravity.c:108-163) equals the restatement bit for bit."""
    for s0, _ in _jacobi_cases()[:6]:
        sim = sim_from(gpu, s0)
        sim.s.G = 1.3
        sim.s.gravity = gpu.GRAVITY["jacobi"]
        acc = sim.update_acceleration()
        assert np.array_equal(acc, orc.gravity_jacobi(s0, 1.3))


def test_whfast_rejects_unsupported(gpu):
    s0, dt = _systems()[0]
    sim = sim_from(gpu, s0, dt=dt)
    wh = sim.set_integrator("whfast")
    wh.coordinates = gpu.WHFAST_COORDINATES["whds"]
    with pytest.raises(gpu.RebError, match="Jacobi and democratic heliocentric"):
        sim.steps(1)
    sim = sim_from(gpu, s0, dt=dt)
    wh = sim.set_integrator("whfast")
    wh.corrector = 4
    with pytest.raises(gpu.RebError, match="following orders"):
        sim.steps(1)
    sim = sim_from(gpu, s0, dt=dt)
    wh = sim.set_integrator("whfast")
    wh.coordinates = gpu.WHFAST_COORDINATES["democraticheliocentric"]
    wh.corrector = 3
    with pytest.raises(gpu.RebError, match="only compatible with Jacobi"):
        sim.steps(1)
    sim = sim_from(gpu, s0, dt=dt)
    wh = sim.set_integrator("whfast")
    wh.kernel = 4
    with pytest.raises(gpu.RebError, match="Kernel method must be 0"):
        sim.steps(1)
    sim = sim_from(gpu, s0, dt=dt)
    wh = sim.set_integrator("whfast")
    wh.coordinates = gpu.WHFAST_COORDINATES["democraticheliocentric"]
    wh.kernel = 2
    with pytest.raises(gpu.RebError, match="Non-standard kernel requires Jacobi"):
        sim.steps(1)


# ------------------------------------------------------------------ SEI
def test_sei_epicycle_closed_form(gpu):
    # no forces: H0 is exact, bodies return after one epicycle period 2 pi / Omega
    rng = np.random.default_rng(5)
    st = np.zeros((300, 11))
    st[:, 0] = rng.uniform(-5, 5, 300)
    st[:, 1] = rng.uniform(-5, 5, 300)
    st[:, 2] = rng.uniform(-1, 1, 300)
    st[:, 3:6] = rng.normal(scale=0.1, size=(300, 3))
    st[:, 4] += -1.5 * st[:, 0]
    st[:, 9] = 1.0
    sim = sim_from(gpu, st, OMEGA=1.0, dt=2 * math.pi / 200)
    sim.set_integrator("sei")
    sim.s.gravity = gpu.GRAVITY["none"]
    sim.steps(200)
    got = sim.state()
    # guiding-centre drift: y shifts by -3/2 Omega x_gc per unit time (exact solution)
    xgc = 4 * st[:, 0] + 2 * st[:, 4]
    expect_y = st[:, 1] - 1.5 * xgc * 2 * math.pi
    assert np.abs(got[:, 0] - st[:, 0]).max() < 1e-10
    assert np.abs(got[:, 1] - expect_y).max() < 1e-9
    assert np.abs(got[:, 2] - st[:, 2]).max() < 1e-10


def test_sei_operator_bitwise_vs_oracle(gpu, orc):
    rng = np.random.default_rng(6)
    st = np.zeros((500, 11))
    st[:, :3] = rng.uniform(-3, 3, (500, 3))
    st[:, 3:6] = rng.normal(size=(500, 3))
    st[:, 9] = 1e-3
    sim = sim_from(gpu, st, OMEGA=1.0, dt=1e-3 * 2 * math.pi, softening=0.1)
    sim.set_integrator("sei")
    sim.precision = "exact"
    sim.steps(1)
    ref = st.copy()
    dt = 1e-3 * 2 * math.pi
    orc.lib().orc_sei_operator(ref, 500, 1.0, 1.0, dt, 0)
    ref[:, 6:9] = orc.gravity_basic(ref, soft=0.1)
    orc.lib().orc_sei_operator(ref, 500, 1.0, 1.0, dt, 1)
    assert np.array_equal(sim.state()[:, :6], ref[:, :6])


@pytest.mark.parametrize("precision", ["exact", "fp64", "split"])
@pytest.mark.parametrize("gravity", ["basic", "compensated"])
@pytest.mark.parametrize("testparticles", [False, True])
def test_leapfrog_fused_kick_drift_bitwise(gpu, precision, gravity, testparticles):
    """With direct gravity the leapfrog's kick and closing drift run in the
    pass that finishes the accelerations; an additional_forces callback (here
    a no-op) keeps them separate.  Both give the same bits."""
    st = golden("plummer1024_seed1_state.npy")[:700]
    out = []
    for split in (False, True):
        sim = sim_from(gpu, st, softening=0.01, dt=1e-3)
        sim.set_integrator("leapfrog")
        sim.s.gravity = gpu.GRAVITY[gravity]
        sim.precision = precision
        if testparticles:
            sim.s.N_active = 600
            sim.s.testparticle_type = 1
        if split:
            cb = gpu._abi.HEARTBEAT(lambda r: None)
            sim._keep.append(cb)
            sim.s.additional_forces = ctypes.cast(cb, ctypes.c_void_p)
        sim.steps(20)
        out.append(sim.state())
        sim.free()
    assert np.array_equal(out[0][:, :9], out[1][:, :9])
