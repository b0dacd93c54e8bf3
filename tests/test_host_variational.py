"""Host side of the variational sets (no GPU): reb_simulation_init_megno
(tools.c:1431-1470), reb_simulation_move_to_com with variational sets
(tools.c:201-375) and reb_simulation_rescale_var (tools.c:1313-1380), against
the reference library's fixtures (gen_golden.c, gen_megno) and the
definition of the rescaling."""
import ctypes
import math

import numpy as np

from conftest import golden

VFIELDS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m")


def _rows(arr):
    return np.stack([arr[f] for f in VFIELDS], axis=1).copy()


def _sim(amd, tag):
    sim = amd.Sim()
    sim.set_state(golden(tag + "_init.npy"))
    return sim


def test_init_megno_bitwise(amd):
    sim = _sim(amd, "megno_reg")
    sim.init_megno(seed=0)
    assert np.array_equal(_rows(sim.var_particles()), golden("megno_reg_var0.npy"))
    vc = sim.var_configs()
    assert len(vc) == 1 and vc[0].order == 1 and vc[0].index == 0 and vc[0].testparticle == -1
    assert sim.s.calculate_megno == 1 and sim.megno() == 0. and sim.lyapunov() == 0.
    sim.free()


def test_move_to_com_with_variations_bitwise(amd):
    """The chaotic case of test_megno.py:58-80 calls move_to_com after
    init_megno: the first-order set is shifted as the reference does."""
    sim = _sim(amd, "megno_chaos")
    sim.init_megno(seed=0)
    sim.move_to_com()
    assert np.array_equal(_rows(sim.var_particles()), golden("megno_chaos_var0.npy"))
    sim.free()


def test_rescale_var(amd):
    sim = _sim(amd, "megno_reg")
    a = sim.add_variation()
    b = sim.add_variation()
    v = sim.var_particles()
    v["x"][a:a + 3] = [3e100, -1., 2.]
    v["vz"][a:a + 3] = [1., -7e100, 0.5]
    v["y"][b:b + 3] = [1., 2., 3.]
    before = _rows(v)
    sim.lib.reb_simulation_rescale_var(sim.ptr)
    v = sim.var_particles()
    cfg = sim.var_configs()
    assert cfg[0].lrescale == math.log(7e100)
    assert cfg[1].lrescale == 0.
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        k = VFIELDS.index(f)
        assert np.array_equal(v[f][a:a + 3], before[a:a + 3, k] / 7e100)
        assert np.array_equal(v[f][b:b + 3], before[b:b + 3, k])
    assert sim.s.did_modify_particles == 1
    sim.free()


def test_variations_survive_copy_and_file(amd, tmp_path):
    sim = _sim(amd, "megno_reg")
    sim.set_integrator("whfast")
    sim.init_megno(seed=0)
    cp = sim.lib.reb_simulation_copy(sim.ptr)
    c = amd.Sim.wrap(ctypes.cast(cp, ctypes.c_void_p).value)
    assert np.array_equal(_rows(c.var_particles()), _rows(sim.var_particles()))
    assert len(c.var_configs()) == 1
    path = str(tmp_path / "v.bin")
    sim.save(path)
    r = amd.Sim.from_file(path)
    assert np.array_equal(_rows(r.var_particles()), _rows(sim.var_particles()))
    vc = r.var_configs()
    assert len(vc) == 1 and vc[0].index == 0 and vc[0].order == 1 and r.s.calculate_megno == 1
    for s in (sim, c, r):
        s.free()
