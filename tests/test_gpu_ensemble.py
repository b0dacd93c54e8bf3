"""Batched WHFast ensembles against fixtures written by the reference library.

Every system of an ensemble must evolve bit-identically to a reference
reb_simulation integrating it alone with WHFast (democratic heliocentric or
Jacobi coordinates, integrator_whfast.c:1051-1259) through
reb_simulation_steps; the fixtures hold
the reference's energies after each steps() call and its final bodies."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _fixture(tag):
    return (golden(tag + "_nbodies.npy"), golden(tag + "_init.npy"), golden(tag + "_dt.npy"),
            golden(tag + "_energy.npy"), golden(tag + "_final.npy"))


CASES = {
    # tag, safe_mode, the three steps() calls of the fixture
    "ensemble200_seed99": (0, (37, 1, 64)),
    "ensemble8_big_seed5": (1, (20, 5, 25)),
}


def _run(gpu, tag, launch=None):
    safe, calls = CASES[tag]
    nb, init, dt, en, fin = _fixture(tag)
    ens = gpu.Ensemble(nb, init, dt, G=1.0, safe_mode=safe)
    if launch:
        ens.set_launch(**launch)
    E = [ens.energy()]
    for n in calls:
        ens.steps(n)
        E.append(ens.energy())
    st = ens.state()
    ens.free()
    return np.stack(E, axis=1), st, en, fin


@pytest.mark.parametrize("tag", sorted(CASES))
def test_ensemble_bitwise(gpu, tag):
    E, st, en, fin = _run(gpu, tag)
    assert np.array_equal(E, en)
    assert np.array_equal(st[:, :6], fin[:, :6])
    assert np.array_equal(st[:, 9], fin[:, 9])


@pytest.mark.parametrize("launch", [{"grid": 3, "chunk": 0}, {"grid": 0, "chunk": 7}, {"grid": 5, "chunk": 16}])
def test_ensemble_launch_shapes_bitwise(gpu, launch):
    # fewer persistent workgroups than tasks, and requests split into several
    # graph launches, give the same bits
    E, st, en, fin = _run(gpu, "ensemble200_seed99", launch)
    assert np.array_equal(E, en)
    assert np.array_equal(st[:, :6], fin[:, :6])


@pytest.mark.parametrize("safe", [0, 1])
def test_ensemble_whfast_dh32(gpu, safe):
    # the per-simulation fixtures (steps(100) ten times), as one ensemble
    init = golden("whfast_dh32_init.npy")
    meta = golden("whfast_dh32_meta.npy")
    dts = golden("whfast_dh32_dt.npy")
    fin = golden("whfast_dh32_safe%d_final.npy" % safe)
    en = golden("whfast_dh32_safe%d_energy.npy" % safe)
    nb = meta[:, 1]
    state = np.concatenate([init[s, :nb[s]] for s in range(len(nb))])
    ens = gpu.Ensemble(nb, state, dts, safe_mode=safe)
    E = [ens.energy()]
    for _ in range(10):
        ens.steps(100)
        E.append(ens.energy())
    assert np.array_equal(np.stack(E, axis=1), en)
    st = ens.state()
    for s in range(len(nb)):
        assert np.array_equal(ens.system(s, st)[:, :6], fin[s, :nb[s], :6]), s


@pytest.mark.parametrize("mode", range(5))
def test_ensemble_whfast_jacobi(gpu, mode):
    """Jacobi coordinates with symplectic correctors: the per-simulation
    fixtures (three steps(100) calls), all sixteen systems as one ensemble."""
    safe, c1, c2, comp, off, tp = (int(v) for v in golden("whfast_jac16_modes.npy")[mode])
    assert off == 0 and tp == 0
    nb = golden("whfast_jac16_nbodies.npy")
    ens = gpu.Ensemble(nb, golden("whfast_jac16_init.npy"), golden("whfast_jac16_dt.npy"), safe_mode=safe,
                       gravity="compensated" if comp else "basic", coordinates="jacobi", corrector=c1,
                       corrector2=c2)
    E = [ens.energy()]
    for _ in range(3):
        ens.steps(100)
        E.append(ens.energy())
    assert np.array_equal(np.stack(E, axis=1), golden("whfast_jac16_mode%d_energy.npy" % mode))
    st = ens.state()
    assert np.array_equal(st[:, :6], golden("whfast_jac16_mode%d_final.npy" % mode)[:, :6])
    ens.free()


@pytest.mark.parametrize("mode", range(6))
def test_ensemble_whfast_kernels(gpu, mode):
    """The modified-kick, composition and lazy implementer's kernels: the
    per-simulation fixtures (three steps(100) calls), all sixteen systems as
    one ensemble, bodies with accelerations and energies bit for bit."""
    kern, safe, c1, c2, comp = (int(v) for v in golden("whfast_kern16_modes.npy")[mode])
    nb = golden("whfast_jac16_nbodies.npy")
    ens = gpu.Ensemble(nb, golden("whfast_jac16_init.npy"), golden("whfast_jac16_dt.npy"), safe_mode=safe,
                       gravity="compensated" if comp else "basic", coordinates="jacobi", corrector=c1,
                       corrector2=c2, kernel=kern)
    E = [ens.energy()]
    for _ in range(3):
        ens.steps(100)
        E.append(ens.energy())
    assert np.array_equal(np.stack(E, axis=1), golden("whfast_kern16_mode%d_energy.npy" % mode))
    assert np.array_equal(ens.state()[:, :9], golden("whfast_kern16_mode%d_final.npy" % mode)[:, :9])
    ens.free()


def test_ensemble_whfast_kernel_errors(gpu):
    nb, init, dt, en, fin = _fixture("ensemble200_seed99")
    with pytest.raises(gpu.RebError, match="requires Jacobi"):
        gpu.Ensemble(nb, init, dt, kernel=2)
    with pytest.raises(gpu.RebError, match="kernel method must be"):
        gpu.Ensemble(nb, init, dt, coordinates="jacobi", kernel=4)


def test_ensemble_tune_keeps_state(gpu):
    nb, init, dt, en, fin = _fixture("ensemble200_seed99")
    ens = gpu.Ensemble(nb, init, dt)
    ens.steps(37)
    before = ens.state()
    best, shape = ens.tune(probe_steps=16)
    assert best > 0 and shape["grid"] >= 1
    assert np.array_equal(ens.state(), before)
    assert ens.steps_done == 37
    ens.steps(1)
    ens.steps(64)
    assert np.array_equal(ens.energy(), en[:, 3])
    assert np.array_equal(ens.state()[:, :6], fin[:, :6])


def test_ensemble_matches_per_simulation_plugin(gpu):
    # the same systems through reb_simulation_steps one by one
    nb, init, dt, _, _ = _fixture("ensemble200_seed99")
    ens = gpu.Ensemble(nb, init, dt, gravity="compensated")
    ens.steps(50)
    st = ens.state()
    first = np.concatenate([[0], np.cumsum(nb)[:-1]])
    for s in range(0, len(nb), 23):
        sim = gpu.Sim()
        sim.set_state(init[first[s]:first[s] + nb[s]])
        sim.s.dt = dt[s]
        sim.s.gravity = gpu.GRAVITY["compensated"]
        wh = sim.set_integrator("whfast")
        wh.coordinates = gpu.WHFAST_COORDINATES["democraticheliocentric"]
        wh.safe_mode = 0
        sim.precision = "exact"
        sim.steps(50)
        assert np.array_equal(sim.state()[:, :6], st[first[s]:first[s] + nb[s], :6]), s
        sim.free()


def test_ensemble_errors(gpu):
    with pytest.raises(gpu.RebError, match="1 to 65 bodies"):
        gpu.Ensemble([66], np.zeros((66, 11)), 0.1)
    with pytest.raises(gpu.RebError, match="BASIC or REB_GRAVITY_COMPENSATED"):
        gpu.Ensemble([2], np.zeros((2, 11)), 0.1, gravity="tree")
    with pytest.raises(ValueError):
        gpu.Ensemble([3], np.zeros((2, 11)), 0.1)
    st = np.zeros((2, 11))
    st[:, 9] = 1.0
    st[1, 0] = 1.0
    with pytest.raises(gpu.RebError, match="following orders"):
        gpu.Ensemble([2], st, 0.1, coordinates="jacobi", corrector=4)
    with pytest.raises(gpu.RebError, match="only compatible with Jacobi"):
        gpu.Ensemble([2], st, 0.1, corrector=3)
