"""GPU parity of direct gravity (gravity.c:167-531) and the energy (tools.c:108-162).

EXACT is bit-identical to the reference; FP64 within 1e-14 of the summed
magnitudes sum_j |a_ij| per body; SPLIT within 1e-7 of it (FP32 r^-3), with
the 99th percentile of the plain relative error within 1e-7 as well.  At the
benchmark's full size (65536 bodies) the FP64 modes are within 1e-14 of the
exact force (terms and sum in extended precision), relative to |a_i|."""
import hashlib

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-14
TOL_SPLIT = 1e-7
# One pair's r^-3 in SPLIT carries the FP32 reciprocal square root's error
# (1 ulp of FP32, 2^-23) three times over (the cube): a force dominated by a
# single term, such as a planet's by its star, is good to 3 * 2^-23 = 3.6e-7
# of that term; many-term sums average this down to the 1e-7 above.
TOL_SPLIT_TERM = 4e-7


def make_sim(amd, state, soft, precision, G=1.0):
    sim = amd.Sim()
    sim.s.G = G
    sim.s.softening = soft
    sim.set_state(state[:, :11] if state.shape[1] >= 11 else state)
    sim.precision = precision
    return sim


def abs_sum(state, soft, rows=None):
    """sum_j |a_ij| for the target rows (all bodies as sources): the
    conditioning of each body's sum."""
    x = state[:, :3]
    m = state[:, 9]
    rows = np.arange(x.shape[0]) if rows is None else np.asarray(rows)
    out = np.zeros(len(rows))
    for b in range(0, len(rows), 256):
        t = rows[b:b + 256]
        d = x[t, None, :] - x[None, :, :]
        dist2 = (d * d).sum(-1)
        with np.errstate(divide="ignore", invalid="ignore"):
            w = m[None, :] * np.sqrt(dist2) / (dist2 + soft * soft) ** 1.5
        w[~np.isfinite(w)] = 0
        out[b:b + 256] = w.sum(1)
    return out


@pytest.mark.parametrize("which,grav", [("basic", "basic"), ("comp", "compensated")])
def test_exact_bitwise_plummer1024(gpu, which, grav):
    st = golden("plummer1024_seed1_state.npy")
    sim = make_sim(gpu, st, 0.01, "exact")
    sim.s.gravity = gpu.GRAVITY[grav]
    acc = sim.update_acceleration()
    assert np.array_equal(acc, golden("plummer1024_seed1_acc_%s.npy" % which))


@pytest.mark.parametrize("which,grav", [("basic", "basic"), ("comp", "compensated")])
def test_exact_bitwise_testparticles(gpu, which, grav):
    st = golden("plummer1024_seed1_state.npy")
    sim = make_sim(gpu, st, 0.01, "exact")
    sim.s.gravity = gpu.GRAVITY[grav]
    sim.s.N_active = 768
    sim.s.testparticle_type = 1
    acc = sim.update_acceleration()
    assert np.array_equal(acc, golden("plummer1024_seed1_acc_%s_tp1.npy" % which))


@pytest.mark.parametrize("grav", ["basic", "compensated"])
@pytest.mark.parametrize("n", [3001, 5003, 16500])
def test_exact_bitwise_larger_n(gpu, orc, grav, n):
    """The exact kernel's workgroup shapes by size (the 1024-body cases above
    take 1 target x 256 sources; 3001: 2 x 128; 5003: 4 x 64; 16500: 16 x 64),
    with ragged last tiles and workgroups, test particles and ignored terms."""
    st = orc.plummer(n, 7)[0]
    f = orc.gravity_basic if grav == "basic" else orc.gravity_compensated
    for N_active, tp, ign in ((n, 0, 0), (n - 499, 1, 1), (n * 3 // 4, 1, 2)):
        sim = make_sim(gpu, st, 0.005, "exact")
        sim.s.gravity = gpu.GRAVITY[grav]
        sim.s.N_active = N_active
        sim.s.testparticle_type = tp
        sim.s.gravity_ignore_terms = ign
        acc = sim.update_acceleration()
        ref = f(st, soft=0.005, N_active=N_active, tp_type=tp, ignore=ign)
        assert np.array_equal(acc, ref), (grav, N_active, tp, ign)


@pytest.mark.parametrize("grav", ["basic", "compensated"])
def test_exact_bitwise_row_kernel_large_n(gpu, orc, grav):
    """N=50000 takes the row-wise exact kernels; sampled target rows against
    the restatement, with test particles and ignored terms."""
    n = 50000
    st = orc.plummer(n, 9)[0]
    rng = np.random.default_rng(4)
    for N_active, tp, ign in ((n, 0, 0), (n - 777, 1, 1), (n // 2, 1, 2)):
        rows = np.unique(np.concatenate([[0, 1, 2, N_active - 1, min(N_active, n - 1), n - 1],
                                         rng.integers(0, n, 40)]))
        sim = make_sim(gpu, st, 0.005, "exact")
        sim.s.gravity = gpu.GRAVITY[grav]
        sim.s.N_active = N_active
        sim.s.testparticle_type = tp
        sim.s.gravity_ignore_terms = ign
        acc = sim.update_acceleration()
        ref = orc.gravity_rows(st, rows, soft=0.005, N_active=N_active, tp_type=tp, ignore=ign,
                               compensated=1 if grav == "compensated" else 0)
        assert np.array_equal(acc[rows], ref), (grav, N_active, tp, ign)


def test_exact_bitwise_ghostboxes(gpu):
    st = golden("periodic256_state.npy")
    sim = make_sim(gpu, st, 0.05, "exact")
    sim.s.boundary = gpu.BOUNDARY["periodic"]
    sim.s.root_size = 10.0
    sim.s.N_ghost_x = 1
    sim.s.N_ghost_y = 1
    acc = sim.update_acceleration()
    assert np.array_equal(acc, golden("periodic256_acc_basic_omp.npy"))


@pytest.mark.parametrize("grav,prec,tol", [("basic", "fp64", TOL_FP64), ("compensated", "fp64", TOL_FP64),
                                           ("basic", "split", TOL_SPLIT)])
def test_fast_modes_plummer4096(gpu, grav, prec, tol):
    st = golden("plummer4096_seed2_state.npy")
    ref = golden("plummer4096_seed2_acc_basic.npy")
    sim = make_sim(gpu, st, 0.005, prec)
    sim.s.gravity = gpu.GRAVITY[grav]
    acc = sim.update_acceleration()
    da = np.sqrt(((acc - ref) ** 2).sum(1))
    err = da / abs_sum(st, 0.005)
    rel = da / np.sqrt((ref ** 2).sum(1))
    print("%s/%s max normalised error %.3e, relative p99 %.3e max %.3e" % (
        grav, prec, err.max(), np.quantile(rel, 0.99), rel.max()))
    assert err.max() <= tol
    assert np.quantile(rel, 0.99) <= tol


@pytest.mark.parametrize("prec", ["fp64", "split"])
def test_fast_modes_testparticles_and_ignore(gpu, orc, prec):
    st = golden("plummer1024_seed1_state.npy")
    tol = TOL_FP64 if prec == "fp64" else TOL_SPLIT
    for N_active, tp, ign in ((768, 1, 0), (768, 0, 0), (1024, 0, 1), (1024, 0, 2), (500, 1, 2)):
        sim = make_sim(gpu, st, 0.01, prec)
        sim.s.N_active = N_active
        sim.s.testparticle_type = tp
        sim.s.gravity_ignore_terms = ign
        acc = sim.update_acceleration()
        ref = orc.gravity_basic(st, soft=0.01, N_active=N_active, tp_type=tp, ignore=ign)
        scale = abs_sum(st, 0.01) + 1e-300
        err = np.sqrt(((acc - ref) ** 2).sum(1)) / scale
        assert err.max() <= tol, (N_active, tp, ign, err.max())
        if ign == 2:
            assert np.all(acc[0] == 0)


def test_split_si_units(gpu, orc):
    """SPLIT in SI units: a planetary system in metres, kilograms and seconds
    (G = 6.674e-11, separations 1e9 .. 1e13 m, masses 1e20 .. 2e30 kg).  r^-3
    (down to 1e-39) and the masses (up to 2e30) lie outside FP32's normal
    range; the force must keep the per-term bound of the split arithmetic
    (the star's term dominates every planet's force)."""
    rng = np.random.default_rng(17)
    n = 600
    st = np.zeros((n, 11))
    au = 1.495978707e11
    r = au * np.exp(rng.uniform(np.log(0.05), np.log(60.0), n))
    phi = rng.uniform(0, 2 * np.pi, n)
    st[:, 0] = r * np.cos(phi)
    st[:, 1] = r * np.sin(phi)
    st[:, 2] = r * rng.normal(0, 0.02, n)
    st[:, 9] = np.exp(rng.uniform(np.log(1e20), np.log(2e27), n))
    st[0, :3] = 0.0
    st[0, 9] = 1.989e30
    G = 6.674e-11
    ref = orc.gravity_basic(st, G=G, soft=0.0)
    scale = G * abs_sum(st, 0.0)
    for prec, tol in (("split", TOL_SPLIT_TERM), ("fp64", TOL_FP64)):
        sim = make_sim(gpu, st, 0.0, prec, G=G)
        acc = sim.update_acceleration()
        sim.free()
        assert np.isfinite(acc).all()
        da = np.sqrt(((acc - ref) ** 2).sum(1))
        rel = da / np.sqrt((ref ** 2).sum(1))
        print("%s SI: normalised max %.3e relative p99 %.3e" % (prec, (da / scale).max(), np.quantile(rel, 0.99)))
        assert (da / scale).max() <= tol
        assert np.quantile(rel, 0.99) <= tol


@pytest.mark.parametrize("n", [8192, 9000, 12345])
@pytest.mark.parametrize("prec,tol", [("fp64", TOL_FP64), ("split", TOL_SPLIT)])
def test_symmetric_path(gpu, orc, n, prec, tol):
    """From 8192 bodies, with every body active, no ghost boxes and nothing
    ignored, each pair of bodies in different superblocks (1024 bodies) is
    evaluated once and added to both (Newton's third law, as the reference's
    serial loop does, gravity.c:195-214); 9000 and 12345 leave a ragged last
    superblock.  Against the reference's per-target sums: the FP64 and split
    tolerances of the one-sided kernels, momentum conserved, and the pair count
    the timed kernels report."""
    st = orc.plummer(n, 13)[0]
    ref = orc.gravity_basic(st, soft=0.005)
    sim = make_sim(gpu, st, 0.005, prec)
    gpu.lib().reb_amd_set_profiling(sim.ptr, 1)
    acc = sim.update_acceleration()
    pairs = gpu.lib().reb_amd_kernel_pairs(sim.ptr)
    sim.free()
    full, rest = divmod(n, 1024)
    assert pairs == float(n) * n - (full * 1024 ** 2 + rest ** 2)
    da = np.sqrt(((acc - ref) ** 2).sum(1))
    err = da / abs_sum(st, 0.005)
    rel = da / np.sqrt((ref ** 2).sum(1))
    print("n %d %s: normalised max %.3e relative p99 %.3e" % (n, prec, err.max(), np.quantile(rel, 0.99)))
    # split: the per-term bound for the worst body (a few close neighbours
    # dominate its sum), 1e-7 for the 99th percentile
    assert err.max() <= (tol if prec == "fp64" else TOL_SPLIT_TERM)
    assert np.quantile(rel, 0.99) <= tol
    m = st[:, 9]
    mom = np.abs((m[:, None] * acc).sum(0)).max() / (m[:, None] * np.abs(acc)).sum()
    assert mom <= (1e-13 if prec == "fp64" else 1e-6)


def test_zero_softening_self_term(gpu, orc):
    # softening 0: the self pair must be skipped, not evaluated (0 * inf)
    rng = np.random.default_rng(0)
    st = np.zeros((700, 11))
    st[:, :3] = rng.normal(size=(700, 3))
    st[:, 9] = 1.0 / 700
    for prec in ("exact", "fp64", "split"):
        sim = make_sim(gpu, st, 0.0, prec)
        acc = sim.update_acceleration()
        assert np.all(np.isfinite(acc)), prec
    ref = orc.gravity_basic(st, soft=0.0)
    sim = make_sim(gpu, st, 0.0, "exact")
    assert np.array_equal(sim.update_acceleration(), ref)


def test_energy_exact_bitwise(gpu):
    st = golden("plummer1024_seed1_state.npy")
    sim = make_sim(gpu, st, 0.01, "exact")
    assert sim.energy() == golden("plummer1024_seed1_lf_energy.npy")[0]


def test_energy_exact_bitwise_batched(gpu, orc):
    """N=3001: the pair terms (4.5M) exceed one 4M-term batch of the exact
    energy, with and without test particles."""
    st = orc.plummer(3001, 11)[0]
    for N_active, tp in ((3001, 0), (2500, 0), (2500, 1)):
        sim = make_sim(gpu, st, 0.0, "exact")
        sim.s.N_active = N_active
        sim.s.testparticle_type = tp
        assert sim.energy() == orc.energy(st, 1.0, N_active, tp), (N_active, tp)


def test_energy_parallel(gpu, orc):
    """The compensated parallel sum against an extended-precision sum (1e-14);
    the reference's own sequential sum of 8.4M terms differs from both by its
    rounding error (~1e-13 relative here)."""
    st = golden("plummer4096_seed2_state.npy")
    sim = make_sim(gpu, st, 0.005, "fp64")
    e = sim.energy()
    x = st[:, :3].astype(np.longdouble)
    m = st[:, 9].astype(np.longdouble)
    v = st[:, 3:6].astype(np.longdouble)
    kin = (np.longdouble(0.5) * m * (v * v).sum(1)).sum()
    pot = np.longdouble(0)
    for i in range(st.shape[0] - 1):
        d = x[i + 1:] - x[i]
        pot -= (m[i] * m[i + 1:] / np.sqrt((d * d).sum(1))).sum()
    exact = float(kin + pot)
    assert abs(e - exact) <= 1e-14 * abs(exact)
    assert abs(orc.energy(st) - exact) <= 1e-12 * abs(exact)


class _Hip:
    """Device memory through the HIP runtime the library itself is linked to."""

    def __init__(self):
        import ctypes
        self.c = ctypes
        self.h = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.c.c_void_p()
        assert self.h.hipMalloc(self.c.byref(p), self.c.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(self.c.c_void_p(p), arr.ctypes.data_as(self.c.c_void_p),
                                self.c.c_size_t(arr.nbytes), 1) == 0
        return p

    def get(self, p, n):
        out = np.zeros(n)
        assert self.h.hipMemcpy(out.ctypes.data_as(self.c.c_void_p), self.c.c_void_p(p),
                                self.c.c_size_t(8 * n), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)


def test_raw_entry_point(gpu, orc):
    """reb_amd_gravity_direct on device arrays owned by the caller."""
    st = golden("plummer1024_seed1_state.npy")
    hip = _Hip()
    n = st.shape[0]
    cols = [hip.put(st[:, k]) for k in (0, 1, 2, 9)]
    acc = [hip.alloc(8 * n) for _ in range(3)]
    L = gpu.lib()
    for prec, comp, ref in ((1, 0, "basic"), (1, 1, "comp"), (0, 0, "basic")):
        ws = hip.alloc(L.reb_amd_gravity_workspace_bytes(n, prec))
        rc = L.reb_amd_gravity_direct(*cols, *acc, n, n, 0, 0, 1.0, 0.01, prec, comp, ws, None)
        assert rc == 0
        assert hip.h.hipDeviceSynchronize() == 0
        got = np.stack([hip.get(a, n) for a in acc], axis=1)
        want = golden("plummer1024_seed1_acc_%s.npy" % ref)
        if prec == 1:
            assert np.array_equal(got, want)
        else:
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    hip.free()


@pytest.mark.slow
def test_fp64_large_n_sampled_rows(gpu, orc):
    """N=16384 Plummer: the fast mode against the oracle's full sum."""
    s, _ = orc.plummer(16384, 3)
    sim = make_sim(gpu, s, 0.005, "fp64")
    acc = sim.update_acceleration()
    ref = orc.gravity_basic(s, soft=0.005)
    rows = np.random.default_rng(1).choice(16384, 512, replace=False)
    err = np.sqrt(((acc[rows] - ref[rows]) ** 2).sum(1)) / abs_sum(s, 0.005, rows)
    assert err.max() <= TOL_FP64


def full_size_plummer(gpu):
    """The direct-summation benchmark's bodies: Plummer sphere of 65536
    bodies, seed 42, moved to the centre of mass (the product's generator,
    bit-identical to the reference's, see the leapfrog test below)."""
    sim = gpu.Sim()
    sim.s.rand_seed = 42
    sim.s.softening = 0.005
    sim.s.G = 1.0
    sim.add_plummer(65536, 1.0, 1.0)
    sim.move_to_com()
    return sim


def exact_rows(state, soft, rows):
    """The exact force on the target rows (every term and the sum in x87
    extended precision, 64-bit significand) and sum_j |a_ij|."""
    x = state[:, :3].astype(np.longdouble)
    m = state[:, 9].astype(np.longdouble)
    soft2 = np.longdouble(soft) * np.longdouble(soft)
    truth = np.zeros((len(rows), 3))
    cond = np.zeros(len(rows))
    for b in range(0, len(rows), 16):
        r = rows[b:b + 16]
        d = x[r, None, :] - x[None, :, :]
        r2 = (d * d).sum(2) + soft2
        w = m[None, :] / (r2 * np.sqrt(r2))
        w[np.arange(len(r)), r] = 0
        t = w[..., None] * d
        truth[b:b + 16] = (-t.sum(1)).astype(np.float64)
        cond[b:b + 16] = np.sqrt((t * t).sum(2)).sum(1).astype(np.float64)
    return truth, cond


@pytest.fixture(scope="module")
def full_size(gpu, orc):
    """Bodies, 512 sampled rows, the exact force on them and the oracle's
    (the reference's per-target sum, gravity.c:215-232)."""
    sim = full_size_plummer(gpu)
    st = sim.state()
    sim.free()
    rows = np.sort(np.random.default_rng(2024).choice(65536, 512, replace=False))
    rows[:4] = (0, 1, 2, 65535)
    rows = np.unique(rows)
    truth, cond = exact_rows(st, 0.005, rows)
    ref = orc.gravity_rows(st, rows, soft=0.005)
    return st, rows, truth, cond, ref


@pytest.mark.parametrize("grav,prec", [("basic", "fp64"), ("compensated", "fp64"), ("basic", "split")])
def test_fast_modes_full_size(gpu, full_size, grav, prec):
    """The three timed modes of the benchmark at N=65536 (softening 0.005) on
    512 sampled rows.  FP64 and COMPENSATED: plain relative error against the
    exact force <= 1e-14 on every row.  The reference's own double-precision
    sum is itself about 2e-14 from the exact force on the worst rows, so the
    comparison with the oracle allows that distance plus 1e-14.  SPLIT: error
    <= 1e-7 of sum_j |a_ij| on every row and 99th percentile of the plain
    relative error <= 1e-7."""
    st, rows, truth, cond, ref = full_size
    sim = full_size_plummer(gpu)
    sim.s.gravity = gpu.GRAVITY[grav]
    sim.precision = prec
    acc = sim.update_acceleration()[rows]
    sim.free()
    tn = np.sqrt((truth ** 2).sum(1))
    da = np.sqrt(((acc - truth) ** 2).sum(1))
    rel = da / tn
    ref_rel = np.sqrt(((ref - truth) ** 2).sum(1)) / tn
    vs_ref = np.sqrt(((acc - ref) ** 2).sum(1)) / tn
    print("%s/%s vs exact: rel max %.3e p99 %.3e normalised max %.3e | oracle vs exact: rel max %.3e" % (
        grav, prec, rel.max(), np.quantile(rel, 0.99), (da / cond).max(), ref_rel.max()))
    if prec == "fp64":
        assert rel.max() <= TOL_FP64
        assert np.all(vs_ref <= ref_rel + TOL_FP64)
    else:
        assert (da / cond).max() <= TOL_SPLIT
        assert np.quantile(rel, 0.99) <= TOL_SPLIT


def test_exact_leapfrog_full_size_two_steps_bitwise(gpu):
    """The direct-summation benchmark at full size (Plummer sphere of 65536
    bodies, seed 42, softening 0.005, leapfrog with dt 5e-4, BASIC gravity)
    in exact mode: the SHA-256 of every body's position and velocity, in body
    order, after construction and after each of two steps equals the
    reference library's (serial build) on the same construction."""
    want = golden("plummer65536_seed42_sha256.npy")
    sim = full_size_plummer(gpu)
    sim.set_integrator("leapfrog")
    sim.s.dt = 5e-4
    sim.s.gravity = gpu.GRAVITY["basic"]
    sim.precision = "exact"
    for s in range(3):
        if s:
            sim.steps(1)
        st = np.ascontiguousarray(sim.state()[:, :6])
        assert hashlib.sha256(st.tobytes()).digest() == want[s].tobytes(), s
    sim.free()
