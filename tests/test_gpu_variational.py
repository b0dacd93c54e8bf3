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
