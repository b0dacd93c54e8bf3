"""Binary files and simulation archives (binarydata.c, simulationarchive.c).

The fixtures are files the reference library wrote: a leapfrog simulation and
a WHFast archive with a snapshot every 7 steps.  Reading and writing are host
code, so these tests run without a GPU; continuing an archived integration
on the GPU is in test_gpu_binary_io.py."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN, golden

FIELD = struct.Struct("<QQ")
BLOB = struct.Struct("<iii")


def fields(buf, pos=64):
    """(name, data) of one snapshot, from pos to its "end" field."""
    out = []
    while pos + FIELD.size <= len(buf):
        size_name, size_data = FIELD.unpack_from(buf, pos)
        pos += FIELD.size
        name = buf[pos:pos + size_name - 1].decode()
        pos += size_name
        data = buf[pos:pos + size_data]
        pos += size_data
        out.append((name, data))
        if name == "end":
            return out, pos
    raise AssertionError("no end field")


def particle_doubles(data):
    p = np.frombuffer(data, dtype=np.uint8).reshape(-1, 112)
    return np.ascontiguousarray(p[:, :88]).view(np.float64)


def test_read_reference_file(amd):
    sim = amd.Sim.from_file(os.path.join(GOLDEN, "archive_leapfrog.bin"))
    ref = golden("archive_leapfrog_state.npy")
    assert sim.s.N == 64
    assert sim.s.G == 1.0 and sim.s.softening == 0.01 and sim.s.dt == 0.01
    assert sim.s.integrator.name == b"leapfrog"
    lf = ctypes_state(amd, sim, "leapfrog")
    assert lf.order == 4
    assert np.array_equal(sim.state(), ref)
    sim.free()


def ctypes_state(amd, sim, kind):
    import ctypes
    cls = amd._abi.LeapfrogState if kind == "leapfrog" else amd._abi.WHFastState
    return ctypes.cast(sim.s.integrator.state, ctypes.POINTER(cls)).contents


def test_write_matches_reference_layout(amd, tmp_path):
    """Read then written again, the file has the reference's fields in the
    reference's order with the same data (the header's git hash and the
    particles' pointer fields aside)."""
    src = open(os.path.join(GOLDEN, "archive_leapfrog.bin"), "rb").read()
    sim = amd.Sim.from_file(os.path.join(GOLDEN, "archive_leapfrog.bin"))
    out = tmp_path / "again.bin"
    sim.save(out)
    sim.free()
    mine = out.read_bytes()
    assert mine[:36] == src[:36]             # "REBOUND Binary File. Version: 5.0.0"
    fa, ea = fields(src)
    fb, eb = fields(mine)
    assert [n for n, _ in fa] == [n for n, _ in fb]
    for (n, a), (_, b) in zip(fa, fb):
        if n == "particles":
            assert np.array_equal(particle_doubles(a), particle_doubles(b))
        elif n not in ("save_messages",):
            assert a == b, n
    assert BLOB.unpack_from(mine, eb) == (0, 0, 0)
    assert len(mine) == eb + BLOB.size


@pytest.mark.parametrize("k", range(5))
def test_read_reference_archive_snapshots(amd, k):
    sim = amd.Sim.from_file(os.path.join(GOLDEN, "archive_whfast.bin"), k)
    steps, sync, n = golden("archive_whfast_snap%d_meta.npy" % k)
    assert sim.s.steps_done == steps and sim.s.is_synchronized == sync and sim.s.N == n
    assert np.array_equal(sim.state(), golden("archive_whfast_snap%d_state.npy" % k))
    wh = ctypes_state(amd, sim, "whfast")
    # p_jh exists once the integrator has run (not in the snapshot before the first step)
    assert wh.safe_mode == 0 and wh.corrector == 5 and wh.N_allocated == (n if k else 0)
    sim.free()
    # negative snapshots count from the end
    if k == 4:
        last = amd.Sim.from_file(os.path.join(GOLDEN, "archive_whfast.bin"), -1)
        assert np.array_equal(last.state(), golden("archive_whfast_snap4_state.npy"))
        last.free()


def test_append_snapshot(amd, tmp_path):
    """A snapshot appended by this library to the reference's archive: the
    blob chain stays consistent, the old snapshots read back unchanged and the
    new one holds the appended state."""
    path = tmp_path / "arch.bin"
    path.write_bytes(open(os.path.join(GOLDEN, "archive_whfast.bin"), "rb").read())
    sim = amd.Sim.from_file(path, 2)
    sim.save(path)
    sim.free()
    buf = path.read_bytes()
    # walk the chain: header snapshot, then diffs
    _, pos = fields(buf)
    index, prev, nxt = BLOB.unpack_from(buf, pos)
    count = 1
    while nxt:
        start = pos + BLOB.size
        _, pos = fields(buf, start)
        index2, prev2, nxt2 = BLOB.unpack_from(buf, pos)
        assert prev2 == nxt and index2 == index + 1
        assert prev2 + BLOB.size == pos + BLOB.size - start
        index, prev, nxt = index2, prev2, nxt2
        count += 1
    assert count == 6 and len(buf) == pos + BLOB.size
    for k in range(5):
        s = amd.Sim.from_file(path, k)
        assert np.array_equal(s.state(), golden("archive_whfast_snap%d_state.npy" % k)), k
        s.free()
    new = amd.Sim.from_file(path, 5)
    assert np.array_equal(new.state(), golden("archive_whfast_snap2_state.npy"))
    assert new.s.steps_done == 14
    new.free()


def test_missing_file_returns_null(amd, tmp_path):
    assert not amd.lib().reb_simulation_create_from_file(os.fsencode(str(tmp_path / "none.bin")), 0)


def test_fresh_simulation_writes_reference_bytes(amd, tmp_path):
    """The same construction in this library and in the reference (seeded
    Plummer sphere, leapfrog of order 4) writes the same fields and data: the
    defaults of reb_simulation_create agree field by field."""
    import ctypes
    src = open(os.path.join(GOLDEN, "archive_leapfrog.bin"), "rb").read()
    sim = amd.Sim(save_messages=False)
    sim.s.rand_seed = 77
    sim.s.G = 1.0
    sim.s.softening = 0.01
    sim.s.dt = 0.01
    sim.add_plummer(64, 1.0, 1.0)
    sim.move_to_com()
    st = sim.set_integrator("leapfrog")
    st.order = 4
    out = tmp_path / "fresh.bin"
    sim.save(out)
    sim.free()
    fa, _ = fields(src)
    fb, _ = fields(out.read_bytes())
    assert [n for n, _ in fa] == [n for n, _ in fb]
    for (n, a), (_, b) in zip(fa, fb):
        if n == "particles":
            assert np.array_equal(particle_doubles(a), particle_doubles(b))
        else:
            assert a == b, n


def _field_offsets(buf, pos=64):
    out = {}
    while pos + FIELD.size <= len(buf):
        size_name, size_data = FIELD.unpack_from(buf, pos)
        name = bytes(buf[pos + FIELD.size:pos + FIELD.size + size_name - 1]).decode()
        out[name] = pos
        pos += FIELD.size + size_name + size_data
        if name == "end":
            break
    return out


@pytest.mark.parametrize("damage", ["truncated", "huge_data", "huge_name", "unterminated_name", "bad_name_list"])
def test_damaged_file_read_safely(amd, tmp_path, damage):
    """A truncated file, field sizes that would wrap the read position, a
    field name without its terminating NUL and a name list whose last entry
    runs past its field: the reader stops at the damage (NULL, or a
    simulation with a warning) instead of reading past the buffer."""
    import ctypes
    buf = bytearray(open(os.path.join(GOLDEN, "archive_leapfrog.bin"), "rb").read())
    off = _field_offsets(buf)
    p = off["particles"]
    if damage == "truncated":
        buf = buf[:p + FIELD.size + 40]
    elif damage == "huge_data":
        struct.pack_into("<Q", buf, p + 8, 2 ** 64 - 16)
    elif damage == "huge_name":
        struct.pack_into("<Q", buf, p, 2 ** 64 - 8)
    elif damage == "unterminated_name":
        size_name, _ = FIELD.unpack_from(buf, p)
        buf[p + FIELD.size + size_name - 1] = ord("x")
    else:
        # a name_list field holding one name with no room for its pointer
        entry = FIELD.pack(len(b"name_list\0"), 6) + b"name_list\0" + b"abcde\0"
        buf = buf[:p] + entry + buf[p:]
    path = tmp_path / ("damaged_%s.bin" % damage)
    path.write_bytes(bytes(buf))
    ptr = amd.lib().reb_simulation_create_from_file(os.fsencode(str(path)), 0)
    if ptr:
        sim = amd.Sim(save_messages=False, _ptr=ptr)
        assert sim.s.N <= 64
        assert sim.s.N == 0 or bool(sim.s.particles)
        sim.free()
    assert ctypes.sizeof(amd._abi.Simulation) > 0
