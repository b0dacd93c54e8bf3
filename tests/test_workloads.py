"""Workload generators draw the same numbers as the reference's own calls
(fixtures written by tests/golden/gen_golden.c against the compiled reference)."""
import numpy as np

from conftest import golden


def test_shearing_sheet_bitwise(amd):
    sim = amd.shearing_sheet(2000, 50.0, 2, 99)
    assert np.array_equal(sim.state(), golden("sheet2000_init.npy"))
    assert sim.s.rand_seed == int(golden("sheet2000_seed_init.npy")[0])
    assert sim.s.coefficient_of_restitution
    assert amd.lib().reb_amd_coefficient_of_restitution_constant(sim.ptr, 3.0) == 0.5
    sim.free()
