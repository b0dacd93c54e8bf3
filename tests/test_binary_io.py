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


class _Archive:
    """The archive index in the reference's layout (simulationarchive.h:30-43),
    as its Python package declares it."""

    @staticmethod
    def cls():
        import ctypes as c

        class SA(c.Structure):
            _fields_ = [("inf", c.c_void_p), ("filename", c.c_char_p), ("version", c.c_int),
                        ("major", c.c_int), ("minor", c.c_int), ("patch", c.c_int),
                        ("auto_interval", c.c_double), ("auto_walltime", c.c_double),
                        ("auto_step", c.c_uint64), ("nblobs", c.c_int64),
                        ("offset", c.POINTER(c.c_uint64)), ("t", c.POINTER(c.c_double))]
        return SA


def _archive_lib(amd):
    import ctypes as c
    L = amd.lib()
    SA = _Archive.cls()
    L.reb_simulationarchive_create_from_file.restype = c.POINTER(SA)
    L.reb_simulationarchive_create_from_file.argtypes = [c.c_char_p]
    L.reb_simulationarchive_create_from_file_with_messages.restype = c.POINTER(SA)
    L.reb_simulationarchive_create_from_file_with_messages.argtypes = [c.c_char_p, c.POINTER(c.c_int)]
    L.reb_simulationarchive_create_from_buffer_with_messages.restype = c.POINTER(SA)
    L.reb_simulationarchive_create_from_buffer_with_messages.argtypes = [c.c_char_p, c.c_size_t,
                                                                          c.POINTER(c.c_int)]
    L.reb_simulationarchive_free.argtypes = [c.POINTER(SA)]
    L.reb_simulation_create_from_simulationarchive.restype = c.c_void_p
    L.reb_simulation_create_from_simulationarchive.argtypes = [c.POINTER(SA), c.c_int64]
    L.reb_simulation_init_from_simulationarchive_with_messages.argtypes = [c.c_void_p, c.POINTER(SA), c.c_int64,
                                                                            c.POINTER(c.c_int)]
    return L, SA


def test_archive_object_index_and_snapshots(amd):
    """reb_simulationarchive_create_from_file on the reference's archive (five
    snapshots, one every 7 steps): the index holds the reference's version,
    the save interval, each snapshot's offset and time, and every snapshot read
    through the object equals the one read by reb_simulation_create_from_file."""
    import ctypes as c
    L, SA = _archive_lib(amd)
    path = os.path.join(GOLDEN, "archive_whfast.bin")
    sa = L.reb_simulationarchive_create_from_file(path.encode())
    assert sa
    a = sa.contents
    assert a.nblobs == 5 and a.version >= 5 and a.auto_step == 7
    assert (a.major, a.minor, a.patch) == (5, 0, 0)
    assert a.filename == path.encode()
    offs = [a.offset[k] for k in range(5)]
    assert offs[0] == 0 and offs == sorted(offs)
    for k in range(5):
        direct = amd.Sim.from_file(path, k)
        assert a.t[k] == direct.s.t
        ptr = L.reb_simulation_create_from_simulationarchive(sa, k)
        assert ptr
        via = amd.Sim.wrap(ptr)
        assert via.s.steps_done == direct.s.steps_done and via.s.N == direct.s.N
        assert np.array_equal(via.state(), golden("archive_whfast_snap%d_state.npy" % k))
        assert amd.lib().reb_simulation_diff(via.ptr, direct.ptr, 0) == 0
        via.free()
        direct.free()
    # negative snapshots count from the end; past the end is an error (NULL)
    last = amd.Sim.wrap(L.reb_simulation_create_from_simulationarchive(sa, -1))
    assert np.array_equal(last.state(), golden("archive_whfast_snap4_state.npy"))
    last.free()
    assert not L.reb_simulation_create_from_simulationarchive(sa, 5)
    L.reb_simulationarchive_free(sa)


def test_archive_object_with_messages_and_buffer(amd, tmp_path):
    """The entry points the reference's Python package calls: codes instead of
    messages (binarydata.h:31-45), from a file and from memory."""
    import ctypes as c
    L, SA = _archive_lib(amd)
    w = c.c_int(0)
    sa = L.reb_simulationarchive_create_from_file_with_messages(str(tmp_path / "none.bin").encode(), c.byref(w))
    assert w.value & 1                      # REB_BINARYDATA_ERROR_NOFILE
    L.reb_simulationarchive_free(sa)
    assert not L.reb_simulationarchive_create_from_file(str(tmp_path / "none.bin").encode())
    buf = open(os.path.join(GOLDEN, "archive_whfast.bin"), "rb").read()
    w = c.c_int(0)
    sa = L.reb_simulationarchive_create_from_buffer_with_messages(buf, len(buf), c.byref(w))
    assert w.value == 0 and sa.contents.nblobs == 5 and not sa.contents.filename
    sim = amd.Sim()
    w = c.c_int(0)
    L.reb_simulation_init_from_simulationarchive_with_messages(sim.ptr, sa, 3, c.byref(w))
    assert w.value == 0
    assert np.array_equal(sim.state(), golden("archive_whfast_snap3_state.npy"))
    w = c.c_int(0)
    L.reb_simulation_init_from_simulationarchive_with_messages(sim.ptr, sa, 9, c.byref(w))
    assert w.value & 32                     # REB_BINARYDATA_ERROR_OUTOFRANGE
    sim.free()
    L.reb_simulationarchive_free(sa)
    # a damaged second snapshot: the first is kept, with the corrupt-file warning
    _, end0 = fields(buf)
    cut = buf[:end0 + BLOB.size + 30]
    w = c.c_int(0)
    sa = L.reb_simulationarchive_create_from_buffer_with_messages(cut, len(cut), c.byref(w))
    assert sa.contents.nblobs == 1 and (w.value & 512)
    L.reb_simulationarchive_free(sa)


def test_diff_copy_and_stream(amd, tmp_path, capfd):
    """reb_simulation_diff (simulation.c:691-709): 0 for equal simulations, 1
    after a change, walltime ignored; the differing field printed with option
    1 and returned by reb_simulation_diff_char.  A copy through the binary
    format and the stream of a simulation read back as the same simulation."""
    import ctypes as c
    L = amd.lib()
    path = os.path.join(GOLDEN, "archive_whfast.bin")
    a = amd.Sim.from_file(path, 2)
    b = amd.Sim.from_file(path, 2)
    assert L.reb_simulation_diff(a.ptr, b.ptr, 0) == 0
    b.s.walltime = 123.0
    assert L.reb_simulation_diff(a.ptr, b.ptr, 0) == 0
    b.s.dt = 0.5
    assert L.reb_simulation_diff(a.ptr, b.ptr, 0) == 1
    capfd.readouterr()
    assert L.reb_simulation_diff(a.ptr, b.ptr, 1) == 1
    out = capfd.readouterr().out
    assert "dt:" in out and "0.5" in out and "walltime:" in out
    assert L.reb_simulation_diff(a.ptr, b.ptr, 7) == -1
    L.reb_simulation_diff_char.restype = c.c_void_p
    p = L.reb_simulation_diff_char(a.ptr, b.ptr)
    assert b"dt:" in c.string_at(p)
    L.reb_free(c.c_void_p(p))
    # stream: the simulation as a binary snapshot in memory
    bufp = c.c_void_p()
    size = c.c_size_t()
    L.reb_binarydata_simulation_to_stream(a.ptr, c.byref(bufp), c.byref(size))
    data = c.string_at(bufp, size.value)
    L.reb_free(bufp)
    f = tmp_path / "stream.bin"
    f.write_bytes(data)
    back = amd.Sim.from_file(f, 0)
    assert L.reb_simulation_diff(a.ptr, back.ptr, 0) == 0
    # copy through the binary format into a fresh simulation
    cp = amd.Sim()
    w = c.c_int(0)
    L.reb_simulation_copy_with_messages(cp.ptr, a.ptr, c.byref(w))
    assert w.value == 0
    assert L.reb_simulation_diff(a.ptr, cp.ptr, 0) == 0
    assert np.array_equal(cp.state(), a.state())
    for s in (a, b, back, cp):
        s.free()
