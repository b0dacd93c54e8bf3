"""GPU parity of the first-order variational equations and MEGNO with WHFast
in Jacobi coordinates (integrator_whfast.c:392-490, 1250-1258; gravity.c:533-640;
tools.c:201-375, 1313-1522): bodies, variational particles, their Jacobi
copies (p_jh_var), the MEGNO sums and the sets' lrescale, bit for bit against
the reference library at every checkpoint of the fixtures (gen_golden.c,
gen_megno).  The cases follow rebound/tests/test_megno.py (regular and chaotic
systems), add the 11th-order corrector, two sets with a mass variation in
safe_mode 0 with explicit synchronizations, and a set rescaled past 1e100."""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

VFIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m")


def _rows(arr):
    return np.stack([arr[f] for f in VFIELDS], axis=1).copy()


def _pjh_var(amd, wh):
    n = wh.N_allocated_var
    if n == 0:
        return np.zeros((0, 10))
    buf = (ctypes.c_char * (n * 112)).from_address(ctypes.addressof(wh.p_jh_var.contents))
    return _rows(np.frombuffer(buf, dtype=amd.PARTICLE_DTYPE, count=n))


def _meta(sim):
    s = sim.s
    m = [s.t, sim.megno(), sim.lyapunov(), s.megno_Ys, s.megno_Yss, s.megno_cov_Yt, s.megno_var_t, s.megno_mean_t,
         s.megno_mean_Y, s.megno_initial_t, float(s.megno_n), float(s.gravity_ignore_terms), float(s.is_synchronized),
         float(s.rand_seed)]
    return np.array(m + [vc.lrescale for vc in sim.var_configs()])


def _setup(amd, tag):
    sim = amd.Sim()
    sim.set_state(golden(tag + "_init.npy"))
    wh = sim.set_integrator("whfast")
    var0 = golden(tag + "_var0.npy")
    dt = float(golden(tag + "_k0_meta.npy")[0])   # the first checkpoint is one step from t = 0
    if tag == "megno_reg":
        sim.init_megno(seed=0)
    elif tag == "megno_chaos":
        sim.init_megno(seed=0)
        sim.move_to_com()
    elif tag == "megno_corr":
        wh.corrector = 11
        sim.init_megno(seed=7)
    else:
        if tag == "var_4b":
            wh.safe_mode = 0
            sim.add_variation()
            sim.add_variation()
        else:
            sim.add_variation()
        v = sim.var_particles()
        for k, f in enumerate(VFIELDS):
            v[f] = var0[:, k]
        sim.s.rand_seed = int(golden(tag + "_k0_meta.npy")[13])
    sim.s.dt = dt
    # what init_megno and move_to_com left (or what the case set)
    assert np.array_equal(_rows(sim.var_particles()), var0), tag
    return sim, wh


@pytest.mark.parametrize("tag", ["megno_reg", "megno_chaos", "megno_corr", "var_4b", "var_rescale"])
def test_variational_whfast_bitwise(gpu, tag):
    sim, wh = _setup(gpu, tag)
    ck = golden(tag + "_ck.npy")
    for k, n in enumerate(ck):
        if n < 0:
            sim.synchronize()
        else:
            sim.steps(int(n))
        want_state = golden("%s_k%d_state.npy" % (tag, k))
        assert np.array_equal(sim.state()[:, :10], want_state[:, :10]), (tag, k, "bodies")
        assert np.array_equal(_rows(sim.var_particles()), golden("%s_k%d_var.npy" % (tag, k))), (tag, k, "var")
        assert np.array_equal(_pjh_var(gpu, wh), golden("%s_k%d_pjhvar.npy" % (tag, k))), (tag, k, "p_jh_var")
        assert np.array_equal(_meta(sim), golden("%s_k%d_meta.npy" % (tag, k))), (tag, k, "megno/meta")
    sim.free()


def test_megno_values_as_reference_tests(gpu):
    """The regular system of test_megno.py:33-42 reaches <Y> = 2 +- 0.2 and a
    Lyapunov number of 0 +- 1e-3 (here after 3000 steps, t ~ 2400)."""
    sim, wh = _setup(gpu, "megno_reg")
    sim.steps(3000)
    assert abs(sim.megno() - 2.) < 0.2
    assert abs(sim.lyapunov()) < 1e-3
    sim.free()


def test_variational_errors(gpu):
    """integrator_whfast.c:877-910: the combinations the reference refuses."""
    sim = gpu.Sim()
    sim.set_state(golden("megno_reg_init.npy"))
    wh = sim.set_integrator("whfast")
    wh.coordinates = gpu.WHFAST_COORDINATES["democraticheliocentric"]
    sim.init_megno(seed=0)
    sim.s.dt = 0.1
    with pytest.raises(gpu.RebError, match="only compatible with Jacobi"):
        sim.steps(1)
    wh.coordinates = gpu.WHFAST_COORDINATES["jacobi"]
    wh.safe_mode = 0
    with pytest.raises(gpu.RebError, match="safe_mode=0"):
        sim.steps(1)
    sim.free()


def _bodies_at_start(gpu, tag):
    """The bodies as the case's simulation starts stepping (the chaotic case
    moved to the centre-of-mass frame after init_megno)."""
    if tag != "megno_chaos":
        return golden(tag + "_init.npy")
    sim, _ = _setup(gpu, tag)
    st = sim.state()
    sim.free()
    return st


def _ens_case(gpu, tags, copies, corrector=0):
    inits = [_bodies_at_start(gpu, t) for t in tags]
    var0s = [golden(t + "_var0.npy") for t in tags]
    dts = [float(golden(t + "_k0_meta.npy")[0]) for t in tags]
    state = np.concatenate([inits[i % len(tags)] for i in range(copies * len(tags))])
    var0 = np.concatenate([var0s[i % len(tags)] for i in range(copies * len(tags))])
    nb = [inits[i % len(tags)].shape[0] for i in range(copies * len(tags))]
    dt = [dts[i % len(tags)] for i in range(copies * len(tags))]
    ens = gpu.Ensemble(nb, state, dt, safe_mode=1, coordinates="jacobi", corrector=corrector)
    ens.init_megno(var0)
    return ens


def _ens_check(ens, tags, k_of):
    st, vs = ens.state(), ens.var_state()
    meg, ly = ens.megno()
    for s in range(ens.n_systems):
        tag = tags[s % len(tags)]
        k = k_of[tag]
        if k is None:
            continue
        lo, n = ens.first[s], ens.n_bodies[s]
        want = golden("%s_k%d_state.npy" % (tag, k))[:, :10]
        assert np.array_equal(st[lo:lo + n], want), (tag, k, s, "bodies")
        assert np.array_equal(vs[lo:lo + n], golden("%s_k%d_var.npy" % (tag, k))), (tag, k, s, "var")
        meta = golden("%s_k%d_meta.npy" % (tag, k))
        assert meg[s] == meta[1] and ly[s] == meta[2], (tag, k, s, "megno")


def test_ensemble_megno_bitwise(gpu):
    """MEGNO for many systems in one launch (reb_amd_ensemble_init_megno):
    every copy of the regular and the chaotic system of test_megno.py equals
    the reference's single simulation at each checkpoint (bodies with
    accelerations, variational particles, <Y> and the Lyapunov number), and so
    do copies of the corrector case."""
    tags = ["megno_reg", "megno_chaos"]
    ens = _ens_case(gpu, tags, 40)
    ens.tune(20)   # the probes must leave the MEGNO state as it was
    for k, n in enumerate([1, 9, 90, 900]):
        ens.steps(n)
        _ens_check(ens, tags, {"megno_reg": k, "megno_chaos": k})
    ens.steps(1000)
    _ens_check(ens, tags, {"megno_reg": None, "megno_chaos": 4})
    ens.steps(1000)
    _ens_check(ens, tags, {"megno_reg": 4, "megno_chaos": None})
    ens.free()
    ens = _ens_case(gpu, ["megno_corr"], 5, corrector=11)
    for k, n in enumerate([1, 9, 90, 400]):
        ens.steps(n)
        _ens_check(ens, ["megno_corr"], {"megno_corr": k})
    ens.free()


def test_ensemble_rescale_bitwise(gpu):
    """A set that grows past 1e100 is rescaled on the device as
    reb_simulation_rescale_var does after each step: bodies and variational
    positions and velocities equal the reference's (its variational
    accelerations are those of the MEGNO term here, which the fixture's
    simulation does not compute)."""
    ens = _ens_case(gpu, ["var_rescale"], 3)
    for k, n in enumerate([1, 1, 8, 90]):
        ens.steps(n)
        st, vs = ens.state(), ens.var_state()
        for s in range(ens.n_systems):
            lo, nb = ens.first[s], ens.n_bodies[s]
            assert np.array_equal(st[lo:lo + nb], golden("var_rescale_k%d_state.npy" % k)[:, :10]), (k, s)
            want = golden("var_rescale_k%d_var.npy" % k)
            assert np.array_equal(vs[lo:lo + nb, :6], want[:, :6]), (k, s)
            assert np.array_equal(vs[lo:lo + nb, 9], want[:, 9]), (k, s)
    ens.free()
