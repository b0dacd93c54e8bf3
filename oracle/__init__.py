"""CPU oracle (test infrastructure only).

ctypes bindings of liboracle.so, the plain-C restatement of REBOUND 5.0.0's
force-and-step arithmetic (oracle.c), plus a loader for the reference library
compiled from the reference's own sources (oracle/_ref, built by
`make -C oracle ref`).  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg use this package, and only as the checker; the product
library never links or calls it.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int64, c_long, c_size_t, c_uint

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
NCOL = 11

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        U = POINTER(c_uint)
        sig = {
            "orc_rand_r": (c_int, [U]),
            "orc_random_uniform": (c_double, [U, c_double, c_double]),
            "orc_random_normal": (c_double, [U, c_double]),
            "orc_random_powerlaw": (c_double, [U, c_double, c_double, c_double]),
            "orc_random_rayleigh": (c_double, [U, c_double]),
            "orc_plummer": (None, [U, c_size_t, c_double, c_double, _dp]),
            "orc_move_to_com": (None, [_dp, c_size_t]),
            "orc_particle_from_orbit": (c_int, [c_double, _dp, c_double, c_double, c_double, c_double, c_double,
                                                c_double, c_double, _dp]),
            "orc_gravity_basic": (None, [_dp, c_size_t, c_size_t, c_int, c_int, c_double, c_double, _dp, c_int, _dp]),
            "orc_gravity_compensated": (None, [_dp, c_size_t, c_size_t, c_int, c_int, c_double, c_double, _dp]),
            "orc_gravity_rows": (None, [_dp, c_size_t, c_size_t, c_int, c_int, c_double, c_double, _dp, c_int, c_int,
                                        _ip, c_size_t, _dp]),
            "orc_energy": (c_double, [_dp, c_size_t, c_size_t, c_int, c_double]),
            "orc_leapfrog": (None, [_dp, c_size_t, c_double, c_double, c_double, c_long, POINTER(c_double), c_long,
                                    ctypes.c_void_p]),
            "orc_kepler": (c_int, [_dp, c_double, c_double]),
            "orc_whfast_dh": (None, [_dp, _dp, c_size_t, c_double, c_double, c_double, c_long, c_int, POINTER(c_int),
                                     c_int]),
            "orc_whfast_jacobi": (None, [_dp, _dp, c_size_t, c_size_t, c_int, c_double, c_double, c_int, c_double,
                                         c_long, c_int, c_int, c_int, c_int, POINTER(c_int), c_int]),
            "orc_gravity_jacobi": (None, [_dp, c_size_t, c_double]),
            "orc_sei_operator": (None, [_dp, c_size_t, c_double, c_double, c_double, c_int]),
            "orc_shear_wrap": (None, [_dp, c_size_t, c_double, c_double, c_double, c_double, c_double]),
            "orc_ghostbox_shear": (None, [c_double, c_double, c_double, c_double, c_double, c_int, c_int, c_int,
                                          _dp]),
            "orc_tree_dump": (c_long, [_dp, c_size_t, c_double, c_int, c_int, c_int, ctypes.c_void_p, c_long,
                                       ctypes.c_void_p]),
            "orc_tree_gravity": (c_int, [_dp, c_size_t, c_double, c_int, c_int, c_int, c_double, c_double, c_double,
                                         _dp, c_int, _dp, POINTER(c_long)]),
            "orc_collision_tree": (c_long, [_dp, c_size_t, c_double, c_int, c_int, c_int, _dp, c_int, _ip, _dp,
                                            c_long]),
            "orc_collision_direct": (c_long, [_dp, c_size_t, _dp, c_int, _ip, _dp, c_long]),
            "orc_collision_line": (c_long, [_dp, c_size_t, _dp, c_int, c_double, _ip, _dp, c_long]),
            "orc_collision_linetree": (c_long, [_dp, c_size_t, c_double, c_int, c_int, c_int, _dp, c_int, c_double,
                                                _ip, _dp, c_long]),
            "orc_collision_shuffle": (None, [_ip, _dp, c_long, U]),
            "orc_hardsphere": (c_double, [_dp, c_int64, c_int64, _dp, c_double, c_double, POINTER(c_long)]),
            "orc_flops": (ctypes.c_ulonglong, [c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class Seed:
    """A mutable rand_r seed (r->rand_seed)."""

    def __init__(self, value):
        self.v = c_uint(value)

    @property
    def ref(self):
        return ctypes.byref(self.v)


def plummer(n, seed, M=1.0, R=1.0, com=True):
    s = np.zeros((n, NCOL))
    sd = Seed(seed)
    lib().orc_plummer(sd.ref, n, M, R, s)
    if com:
        lib().orc_move_to_com(s, n)
    return s, sd.v.value


def gravity_basic(s, G=1.0, soft=0.0, N_active=None, tp_type=0, ignore=0, gbs=None):
    s = np.ascontiguousarray(s, dtype=np.float64)
    n = s.shape[0]
    gbs = np.zeros((1, 3)) if gbs is None else np.ascontiguousarray(gbs, dtype=np.float64)
    acc = np.zeros((n, 3))
    lib().orc_gravity_basic(s, n, n if N_active is None else N_active, tp_type, ignore, G, soft, gbs,
                            gbs.shape[0], acc)
    return acc


def gravity_compensated(s, G=1.0, soft=0.0, N_active=None, tp_type=0, ignore=0):
    s = np.ascontiguousarray(s, dtype=np.float64)
    n = s.shape[0]
    acc = np.zeros((n, 3))
    lib().orc_gravity_compensated(s, n, n if N_active is None else N_active, tp_type, ignore, G, soft, acc)
    return acc


def gravity_rows(s, rows, G=1.0, soft=0.0, N_active=None, tp_type=0, ignore=0, gbs=None, compensated=0):
    s = np.ascontiguousarray(s, dtype=np.float64)
    n = s.shape[0]
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    gbs = np.zeros((1, 3)) if gbs is None else np.ascontiguousarray(gbs, dtype=np.float64)
    acc = np.zeros((len(rows), 3))
    lib().orc_gravity_rows(s, n, n if N_active is None else N_active, tp_type, ignore, G, soft, gbs, gbs.shape[0],
                           compensated, rows, len(rows), acc)
    return acc


def energy(s, G=1.0, N_active=None, tp_type=0):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().orc_energy(s, s.shape[0], s.shape[0] if N_active is None else N_active, tp_type, G)


def leapfrog(s, G, soft, dt, steps, every=0):
    s = np.ascontiguousarray(s, dtype=np.float64).copy()
    t = c_double(0.0)
    nE = steps // every + 1 if every else 0
    e = np.zeros(max(nE, 1))
    lib().orc_leapfrog(s, s.shape[0], G, soft, dt, steps, ctypes.byref(t), every,
                       e.ctypes.data if every else None)
    return s, t.value, (e[:nE] if every else None)


def kepler(state6, mu, dt):
    st = np.ascontiguousarray(state6, dtype=np.float64).copy()
    w = lib().orc_kepler(st, mu, dt)
    return st, w


def whfast_dh(s, G, dt, steps, safe_mode, soft=0.0, ph=None, is_sync=1, did_modify=1):
    s = np.ascontiguousarray(s, dtype=np.float64).copy()
    n = s.shape[0]
    ph = np.zeros((n, 7)) if ph is None else np.ascontiguousarray(ph, dtype=np.float64).copy()
    flag = c_int(is_sync)
    lib().orc_whfast_dh(s, ph, n, G, soft, dt, steps, safe_mode, ctypes.byref(flag), did_modify)
    return s, ph, flag.value


def gravity_jacobi(s, G=1.0):
    """Accelerations of the Jacobi gravity (gravity.c:108-163), (N, 3)."""
    s = np.ascontiguousarray(s, dtype=np.float64).copy()
    lib().orc_gravity_jacobi(s, s.shape[0], G)
    return s[:, 6:9].copy()


def whfast_jacobi(s, G, dt, steps, safe_mode, soft=0.0, corrector=0, corrector2=0, N_active=None, tp_type=0,
                  compensated=0, pj=None, is_sync=1, did_modify=1, kernel=0):
    s = np.ascontiguousarray(s, dtype=np.float64).copy()
    n = s.shape[0]
    pj = np.zeros((n, 7)) if pj is None else np.ascontiguousarray(pj, dtype=np.float64).copy()
    flag = c_int(is_sync)
    lib().orc_whfast_jacobi(s, pj, n, n if N_active is None else N_active, tp_type, G, soft, compensated, dt, steps,
                            safe_mode, corrector, corrector2, kernel, ctypes.byref(flag), did_modify)
    return s, pj, flag.value


def whfast_flops_per_system_step(n_bodies, state, dts, steps, G=1.0, safe_mode=0):
    """Algorithmic FLOPs per system-step of WHFast DH on the given systems
    (counted by the restatement while it integrates them for `steps` steps,
    first step and final synchronization included)."""
    first = np.concatenate([[0], np.cumsum(n_bodies)[:-1]])
    lib().orc_flops(1)
    for s, n in enumerate(n_bodies):
        sub = np.zeros((int(n), 11))
        sub[:, :min(11, state.shape[1])] = state[first[s]:first[s] + n, :11]
        whfast_dh(sub, G, float(dts[s]), steps, safe_mode)
    return lib().orc_flops(1) / (len(n_bodies) * float(steps))


def tree_dump(s, root_size, nr):
    s = np.ascontiguousarray(s, dtype=np.float64)
    n = lib().orc_tree_dump(s, s.shape[0], root_size, nr[0], nr[1], nr[2], None, 0, None)
    if n < 0:
        raise ValueError("tree construction error")
    cells = np.zeros((n, 9))
    rs = np.zeros(nr[0] * nr[1] * nr[2] + 1, dtype=np.int64)
    lib().orc_tree_dump(s, s.shape[0], root_size, nr[0], nr[1], nr[2], cells.ctypes.data, n, rs.ctypes.data)
    return cells, rs


def tree_gravity(s, root_size, nr, G, soft, theta2, gbs):
    s = np.ascontiguousarray(s, dtype=np.float64)
    gbs = np.ascontiguousarray(gbs, dtype=np.float64).reshape(-1, 3)
    acc = np.zeros((s.shape[0], 3))
    inter = c_long(0)
    err = lib().orc_tree_gravity(s, s.shape[0], root_size, nr[0], nr[1], nr[2], G, soft, theta2, gbs, gbs.shape[0],
                                 acc, ctypes.byref(inter))
    if err:
        raise ValueError("tree construction error %d" % err)
    return acc, inter.value


def collisions_tree(s, root_size, nr, gbs6, cap=1 << 20):
    s = np.ascontiguousarray(s, dtype=np.float64)
    gbs6 = np.ascontiguousarray(gbs6, dtype=np.float64).reshape(-1, 6)
    pairs = np.zeros((cap, 2), dtype=np.int64)
    gbo = np.zeros((cap, 6))
    n = lib().orc_collision_tree(s, s.shape[0], root_size, nr[0], nr[1], nr[2], gbs6, gbs6.shape[0], pairs, gbo, cap)
    if n < 0:
        raise ValueError("tree construction error")
    return pairs[:n].copy(), gbo[:n].copy()


def collisions_direct(s, gbs6, cap=1 << 20):
    s = np.ascontiguousarray(s, dtype=np.float64)
    gbs6 = np.ascontiguousarray(gbs6, dtype=np.float64).reshape(-1, 6)
    pairs = np.zeros((cap, 2), dtype=np.int64)
    gbo = np.zeros((cap, 6))
    n = lib().orc_collision_direct(s, s.shape[0], gbs6, gbs6.shape[0], pairs, gbo, cap)
    return pairs[:n].copy(), gbo[:n].copy()


def collisions_line(s, gbs6, dt, cap=1 << 20):
    s = np.ascontiguousarray(s, dtype=np.float64)
    gbs6 = np.ascontiguousarray(gbs6, dtype=np.float64).reshape(-1, 6)
    pairs = np.zeros((cap, 2), dtype=np.int64)
    gbo = np.zeros((cap, 6))
    n = lib().orc_collision_line(s, s.shape[0], gbs6, gbs6.shape[0], dt, pairs, gbo, cap)
    return pairs[:n].copy(), gbo[:n].copy()


def collisions_linetree(s, root_size, nr, gbs6, dt, cap=1 << 20):
    s = np.ascontiguousarray(s, dtype=np.float64)
    gbs6 = np.ascontiguousarray(gbs6, dtype=np.float64).reshape(-1, 6)
    pairs = np.zeros((cap, 2), dtype=np.int64)
    gbo = np.zeros((cap, 6))
    n = lib().orc_collision_linetree(s, s.shape[0], root_size, nr[0], nr[1], nr[2], gbs6, gbs6.shape[0], dt, pairs,
                                     gbo, cap)
    if n < 0:
        raise ValueError("tree construction error")
    return pairs[:n].copy(), gbo[:n].copy()


# ------------------------------------------------------------------ the compiled reference
def ref_path(flavour="serial"):
    return os.path.join(HERE, "_ref", flavour, "librebound.so")


def ref_available(flavour="serial"):
    return os.path.exists(ref_path(flavour))


def ref_bench_path():
    return os.path.join(HERE, "_ref", "ref_bench")
