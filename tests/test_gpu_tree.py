"""GPU parity of the octree (tree.c), tree gravity (gravity.c:47-106) and the
collision searches with the shearing-sheet ghost boxes (collision.c), on the
small sheet written by the compiled reference (2000 bodies, 2x2 root boxes,
SEI, hard spheres with restitution 0.5)."""
import ctypes
import hashlib
import math

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def sheet(amd, collision="tree"):
    st = golden("sheet2000_init.npy")
    sim = amd.Sim()
    s = sim.s
    s.rand_seed = 99
    s.OMEGA = 1.0
    s.G = 1.546 / 400.0
    s.softening = 0.1
    s.dt = 1e-3 * 2.0 * math.pi
    s.root_size = 50.0
    s.N_root_x = 2
    s.N_root_y = 2
    s.N_root_z = 1
    s.N_ghost_x = 2
    s.N_ghost_y = 2
    s.N_ghost_z = 0
    s.boundary = amd.BOUNDARY["shear"]
    s.gravity = amd.GRAVITY["tree"]
    s.opening_angle2 = 0.25
    sim.set_integrator("sei")
    sim.set_state(st)
    # the reference drew the bodies from r->rand_seed: continue its stream from there
    s.rand_seed = int(golden("sheet2000_seed_init.npy")[0])
    return sim, st


def test_tree_structure_bitwise(gpu):
    sim, st = sheet(gpu)
    ref = golden("sheet2000_tree.npy")
    out = np.zeros((ref.shape[0] + 16, 9))
    n = gpu.lib().reb_amd_tree_dump(sim.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.shape[0])
    sim.check()
    assert n == ref.shape[0]
    assert np.array_equal(out[:n], ref)


def test_tree_gravity_bitwise(gpu):
    sim, st = sheet(gpu)
    acc = sim.update_acceleration()
    assert np.array_equal(acc, golden("sheet2000_acc_tree.npy"))


def test_tree_gravity_opening_zero_is_direct(gpu, orc):
    rng = np.random.default_rng(2)
    st = np.zeros((3000, 11))
    st[:, :3] = rng.uniform(-20, 20, (3000, 3))
    st[:, 9] = rng.uniform(0.1, 1.0, 3000)
    sim = gpu.Sim()
    sim.s.root_size = 40.0
    sim.s.gravity = gpu.GRAVITY["tree"]
    sim.s.opening_angle2 = 0.0
    sim.s.softening = 0.1
    sim.set_state(st)
    acc = sim.update_acceleration()
    ref = orc.gravity_basic(st, soft=0.1)
    assert np.abs(acc - ref).max() < 1e-12 * np.abs(ref).max()


def test_tree_errors(gpu):
    sim = gpu.Sim()
    sim.s.root_size = 10.0
    sim.s.gravity = gpu.GRAVITY["tree"]
    sim.add(x=1.0, m=1.0)
    sim.add(x=1.0, m=1.0)
    with pytest.raises(gpu.RebError, match="same coordinates"):
        sim.update_acceleration()
    sim2 = gpu.Sim()
    sim2.s.gravity = gpu.GRAVITY["tree"]
    sim2.add(x=1.0, m=1.0)
    with pytest.raises(gpu.RebError, match="root_size"):
        sim2.update_acceleration()


def _run_steps(gpu, collision, record):
    sim, st = sheet(gpu)
    sim.s.collision = gpu.COLLISION[collision]
    rec = []
    cor = gpu._abi.COR(lambda r, v: 0.5)
    if record == "constant":
        # built-in resolver with the library's constant restitution: the
        # level-parallel host resolve
        hs = ctypes.cast(gpu.lib().reb_collision_resolve_hardsphere, ctypes.c_void_p).value
        sim.s.collision_resolve = ctypes.cast(hs, gpu._abi.RESOLVE)
        gpu.lib().reb_amd_set_coefficient_of_restitution(sim.ptr, 0.5)
        return sim, rec
    if record:
        L = gpu.lib()

        def resolve(r, c):
            rec.append((c.p1, c.p2, c.gb.x, c.gb.y, c.gb.z, c.gb.vx, c.gb.vy, c.gb.vz))
            return L.reb_collision_resolve_hardsphere(r, c)

        cb = gpu._abi.RESOLVE(resolve)
        sim._keep.append(cb)
        sim.s.collision_resolve = cb
    else:
        # the built-in resolver: only the colliding bodies travel to the host
        hs = ctypes.cast(gpu.lib().reb_collision_resolve_hardsphere, ctypes.c_void_p).value
        sim.s.collision_resolve = ctypes.cast(hs, gpu._abi.RESOLVE)
    sim._keep.append(cor)
    sim.s.coefficient_of_restitution = cor
    return sim, rec


@pytest.mark.parametrize("record", [True, False, "constant"])
def test_sheet_five_steps_bitwise(gpu, record):
    sim, rec = _run_steps(gpu, "tree", record)
    counts = golden("sheet2000_collision_counts.npy")
    pairs = golden("sheet2000_collision_pairs.npy")
    gb = golden("sheet2000_collision_gb.npy")
    off = np.concatenate([[0], np.cumsum(counts)])
    for s in range(5):
        sim.steps(1)
        ref = golden("sheet2000_step%d_state.npy" % (s + 1))
        assert np.array_equal(sim.state()[:, :6], ref[:, :6]), s
        # r->collisions holds the step's list in resolve order, as the
        # reference leaves it
        nc = sim.s.N_collisions
        assert nc == counts[s]
        c = sim.s.collisions
        got = np.array([(c[k].p1, c[k].p2) for k in range(nc)], dtype=np.int64).reshape(-1, 2)
        ggb = np.array([(c[k].gb.x, c[k].gb.y, c[k].gb.z, c[k].gb.vx, c[k].gb.vy, c[k].gb.vz)
                        for k in range(nc)]).reshape(-1, 6)
        assert np.array_equal(got, pairs[off[s]:off[s + 1]]), s
        assert np.array_equal(ggb, gb[off[s]:off[s + 1]]), s
    assert sim.s.collisions_log_n > 0
    if record is True:
        got = np.array(rec)
        assert got.shape[0] == counts.sum()
        assert np.array_equal(got[:, :2].astype(np.int64), pairs)
        assert np.array_equal(got[:, 2:], gb)
    assert sim.s.rand_seed == int(golden("sheet2000_seed_after.npy")[0])
    # the pressure log is summed in the resolve order (collision.c:655-663)
    plog = golden("sheet2000_plog.npy")
    assert sim.s.collisions_plog == plog[0] and sim.s.collisions_log_n == int(plog[1])


@pytest.mark.parametrize("kind,steps", [("line", 1), ("linetree", 3)])
def test_sheet_line_searches_bitwise(gpu, kind, steps):
    """LINE (collision.c:125-196) and LINETREE (:270-331): the swept-sphere
    searches give the reference's collision lists in resolve order and its
    bodies after every step."""
    sim, rec = _run_steps(gpu, kind, True)
    counts = golden("sheet2000_%s_counts.npy" % kind)
    for s in range(steps):
        sim.steps(1)
        ref = golden("sheet2000_%s_step%d_state.npy" % (kind, s + 1))
        assert np.array_equal(sim.state()[:, :6], ref[:, :6]), s
    got = np.array(rec)
    assert got.shape[0] == counts.sum()
    assert np.array_equal(got[:, :2].astype(np.int64), golden("sheet2000_%s_pairs.npy" % kind))
    assert np.array_equal(got[:, 2:], golden("sheet2000_%s_gb.npy" % kind))


@pytest.mark.parametrize("kind", ["tree", "linetree"])
def test_search_bodies_with_many_collisions(gpu, orc, kind):
    """A body overlapping 40 approaching bodies: more collisions than the
    single-pass search keeps per body, so the search takes its second pass;
    the list in resolve order equals the oracle's search after the reference's
    shuffle."""
    from oracle import lib as olib
    rng = np.random.default_rng(3)
    n = 400
    st = np.zeros((n, 11))
    st[:, 0:2] = rng.uniform(-40, 40, (n, 2))
    st[:, 2] = rng.uniform(-2, 2, n)
    st[:, 9] = 1.0
    st[:, 10] = 0.3
    st[0, :3] = 0.0
    st[0, 10] = 6.0
    d = rng.normal(size=(40, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    st[1:41, :3] = d * rng.uniform(2.0, 5.5, (40, 1))
    st[1:41, 3:6] = -0.2 * st[1:41, :3]
    sim = gpu.Sim()
    sim.set_state(st)
    sim.s.rand_seed = 1234
    sim.s.root_size = 100.0
    sim.s.N_root_x = sim.s.N_root_y = sim.s.N_root_z = 1
    sim.s.gravity = gpu.GRAVITY["none"]
    sim.s.collision = gpu.COLLISION[kind]
    sim.s.dt = 0.05
    sim.set_integrator("leapfrog")
    rec = []

    def resolve(r, c):
        rec.append((c.p1, c.p2))
        return 0

    cb = gpu._abi.RESOLVE(resolve)
    sim._keep.append(cb)
    sim.s.collision_resolve = cb
    sim.steps(1)
    s1 = sim.state()
    g0 = np.zeros((1, 6))
    if kind == "tree":
        pairs, gbo = orc.collisions_tree(s1, 100.0, (1, 1, 1), g0)
    else:
        pairs, gbo = orc.collisions_linetree(s1, 100.0, (1, 1, 1), g0, 0.05)
    seed = ctypes.c_uint(1234)
    olib().orc_collision_shuffle(pairs, gbo, len(pairs), ctypes.byref(seed))
    assert (pairs[:, 0] == 0).sum() > 32   # beyond the per-body slots of the first pass
    assert np.array_equal(np.array(rec, dtype=np.int64), pairs)


def test_sheet_direct_search_bitwise(gpu):
    sim, rec = _run_steps(gpu, "direct", True)
    sim.steps(1)
    pairs = golden("sheet2000_direct_pairs.npy")
    got = np.array(rec)
    assert np.array_equal(got[:, :2].astype(np.int64), pairs)
    assert np.array_equal(got[:, 2:], golden("sheet2000_direct_gb.npy"))
    assert np.array_equal(sim.state()[:, :6], golden("sheet2000_direct_step1_state.npy")[:, :6])


def test_sheet_full_size_collision_pair_set(gpu, orc):
    """The full-size sheet (200000 bodies, 10 x 10 root boxes): the set of
    detected (p1, p2, ghost box) collisions equals the oracle's tree search
    (collision.c:197-269, :422-503) on the same state.  dt = 0 and no gravity,
    so the search sees the initial state."""
    sim = gpu.shearing_sheet(200000, 136.0, 10, 99)
    st = sim.state()
    sim.s.dt = 0.0
    sim.s.gravity = gpu.GRAVITY["none"]
    sim.steps(1)
    n = sim.s.N_collisions
    dt = np.dtype([("p1", "<u8"), ("p2", "<u8"), ("gb", "<f8", 6), ("ri", "<u8")])
    raw = ctypes.string_at(ctypes.cast(sim.s.collisions, ctypes.c_void_p), n * dt.itemsize)
    got = np.frombuffer(raw, dtype=dt)
    gbs = []
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            g = np.zeros(6)
            orc.lib().orc_ghostbox_shear(1.0, 1360.0, 1360.0, 136.0, 0.0, i, j, 0, g)
            gbs.append(g)
    pairs, gb = orc.collisions_tree(st, 136.0, (10, 10, 1), np.array(gbs), cap=1 << 21)
    assert n == pairs.shape[0] and n > 100000
    a = np.column_stack([got["p1"].astype(np.float64), got["p2"].astype(np.float64), got["gb"]])
    b = np.column_stack([pairs.astype(np.float64), gb])
    a = a[np.lexsort(a.T[::-1])]
    b = b[np.lexsort(b.T[::-1])]
    assert np.array_equal(a, b)
    sim.free()


def test_level_parallel_resolve_equals_sequential_at_scale(gpu):
    """A sheet of 60000 bodies (tens of thousands of collisions a step, dozens
    of levels): the built-in resolver with the constant restitution (levels
    computed on the device, resolved level-parallel on host threads) against
    the same resolver behind a restitution callback, which resolves the
    shuffled list one collision at a time as collision.c:345-404 does.  Bodies,
    pressure log and the collision list left in r->collisions agree bit for
    bit after every step."""
    a = gpu.shearing_sheet(60000, 124.0, 6, 5)
    b = gpu.shearing_sheet(60000, 124.0, 6, 5)
    cor = gpu._abi.COR(lambda r, v: 0.5)
    b._keep.append(cor)
    b.s.coefficient_of_restitution = cor
    for step in range(4):
        a.steps(1)
        b.steps(1)
        assert np.array_equal(a.state()[:, :6], b.state()[:, :6]), step
        assert a.s.collisions_plog == b.s.collisions_plog, step
        assert a.s.collisions_log_n == b.s.collisions_log_n, step
        nc = a.s.N_collisions
        assert nc == b.s.N_collisions and nc > 10000, (step, nc)
        ca = np.ctypeslib.as_array(ctypes.cast(a.s.collisions, ctypes.POINTER(ctypes.c_uint8)), (nc * 72,))
        cb = np.ctypeslib.as_array(ctypes.cast(b.s.collisions, ctypes.POINTER(ctypes.c_uint8)), (nc * 72,))
        assert np.array_equal(ca, cb), step
    a.free()
    b.free()


def test_sheet_full_size_three_steps_bitwise(gpu):
    """The benchmark's sheet at full size (200000 bodies): after each of three
    steps, the SHA-256 of every body's position and velocity in body order,
    the bounce count, the pressure log and the random seed equal the
    reference library's (serial build) on the same construction."""
    meta = golden("sheet200k_meta.npy")
    plog = golden("sheet200k_plog.npy")
    digests = golden("sheet200k_sha256.npy")
    sim = gpu.shearing_sheet(200000, 136.0, 10, 99)
    for s in range(3):
        sim.steps(1)
        st = np.ascontiguousarray(sim.state()[:, :6])
        assert hashlib.sha256(st.tobytes()).digest() == digests[s].tobytes(), s
        assert sim.s.collisions_log_n == meta[s, 1], s
        assert sim.s.rand_seed == int(np.int64(meta[s, 2]).view(np.uint64)) & 0xffffffff, s
        assert sim.s.collisions_plog == plog[s], s
    sim.free()


def test_two_simulations_resolving_from_two_threads(gpu):
    """Two sheets of 60000 bodies stepped at the same time from two host
    threads (tens of thousands of collisions a step each, through the
    level-parallel resolve and its shared pool of host threads) end bit for
    bit where the same sheets end when stepped one after the other."""
    import threading
    seeds = (5, 6)
    want = []
    for sd in seeds:
        sim = gpu.shearing_sheet(60000, 124.0, 6, sd)
        sim.steps(3)
        want.append((sim.state()[:, :6], sim.s.collisions_plog, sim.s.collisions_log_n, sim.s.rand_seed))
        sim.free()
    sims = [gpu.shearing_sheet(60000, 124.0, 6, sd) for sd in seeds]
    errors = []

    def run(sim):
        try:
            for _ in range(3):
                sim.lib.reb_simulation_steps(sim.ptr, 1)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(s,)) for s in sims]
    for t in threads:
        t.start()
    for t in threads:
        t.join(240)
        assert not t.is_alive()
    assert not errors
    for sim, (st, plog, logn, seed) in zip(sims, want):
        sim.check()
        assert np.array_equal(sim.state()[:, :6], st)
        assert (sim.s.collisions_plog, sim.s.collisions_log_n, sim.s.rand_seed) == (plog, logn, seed)
        sim.free()
