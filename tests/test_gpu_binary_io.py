"""Continuing archived integrations on the GPU (simulationarchive.c): a
restart from a snapshot gives the bits of the uninterrupted run, as the
reference's own restart tests require (test_simulationarchive.py:118-148)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu


def test_restart_from_reference_archive(gpu):
    """The reference's archive, snapshot 4 (after 28 steps, not synchronized,
    WHFast Jacobi with the 5th-order corrector): two more steps give the
    reference's state after 30 steps."""
    sim = gpu.Sim.from_file(os.path.join(GOLDEN, "archive_whfast.bin"), 4)
    assert sim.s.is_synchronized == 0
    sim.steps(2)
    assert np.array_equal(sim.state()[:, :6], golden("archive_whfast_final_state.npy")[:, :6])
    sim.free()


def test_archive_written_while_stepping(gpu, tmp_path):
    """This library's own archive of the same run (a snapshot every 7 steps,
    written from the device state mid-integration) holds the reference's
    snapshots, and a restart from its last one reaches the reference's end."""
    sim = gpu.Sim.from_file(os.path.join(GOLDEN, "archive_whfast.bin"), 0)
    # the snapshot carries the archive's schedule; start a new one at step 0
    sim.s.simulationarchive_auto_step = 0
    path = tmp_path / "mine.bin"
    gpu.lib().reb_simulation_save_to_file_step(sim.ptr, os.fsencode(str(path)), 7)
    sim.steps(30)
    assert np.array_equal(sim.state()[:, :6], golden("archive_whfast_final_state.npy")[:, :6])
    sim.free()
    for k in range(5):
        s = gpu.Sim.from_file(path, k)
        steps, sync, n = golden("archive_whfast_snap%d_meta.npy" % k)
        assert s.s.steps_done == steps and s.s.is_synchronized == sync
        assert np.array_equal(s.state()[:, :6], golden("archive_whfast_snap%d_state.npy" % k)[:, :6]), k
        s.free()
    s = gpu.Sim.from_file(path, -1)
    s.steps(2)
    assert np.array_equal(s.state()[:, :6], golden("archive_whfast_final_state.npy")[:, :6])
    s.free()
