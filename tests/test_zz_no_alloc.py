"""reb_amd_gravity_direct_noalloc against reb_amd_gravity_direct.

Both entry points share their kernels and chunking; the no-allocation one
keeps the partial sums in the library's fixed device array instead of a caller
workspace, so every precision must give the same bits.  Two calls in a row, the
second with fewer bodies, check that the array's contents left by the first
call do not leak into the second.
"""
import ctypes

import numpy as np
import pytest

from conftest import golden


class _Hip:
    """Device memory through the HIP runtime the library itself is linked to."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(ctypes.c_void_p(p), arr.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_size_t(arr.nbytes), 1) == 0
        return p

    def get(self, p, n):
        out = np.zeros(n)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                ctypes.c_size_t(8 * n), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


# (precision, compensated): FP64, FP64 COMPENSATED, SPLIT, EXACT
MODES = [(0, 0), (0, 1), (2, 0), (1, 0)]


def _forces(L, hip, entry, st, n, n_active, tp, ignore, prec, comp, soft):
    cols = [hip.put(st[:n, k]) for k in (0, 1, 2, 9)]
    acc = [hip.put(np.full(n, np.nan)) for _ in range(3)]
    if entry == "workspace":
        ws = hip.alloc(max(L.reb_amd_gravity_workspace_bytes(n, prec), 256))
        rc = L.reb_amd_gravity_direct(*cols, *acc, n, n_active, tp, ignore, 1.0, soft, prec, comp, ws, None)
    else:
        rc = L.reb_amd_gravity_direct_noalloc(*cols, *acc, n, n_active, tp, ignore, 1.0, soft, prec, comp, None)
    assert rc == 0, L.reb_amd_last_error()
    assert hip.h.hipDeviceSynchronize() == 0
    return np.stack([hip.get(a, n) for a in acc], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,comp", MODES)
def test_noalloc_matches_workspace_entry_point(gpu, orc, prec, comp):
    L = gpu.lib()
    hip = _Hip()
    st = golden("plummer1024_seed1_state.npy")
    big, _ = orc.plummer(20000, 5)
    try:
        # first call: 20000 bodies fills more of the array than the second uses
        cases = [(big, 20000, 20000, 0, 0, 0.005), (st, 1024, 1024, 0, 0, 0.01),
                 (st, 1000, 700, 1, 2, 0.01)]
        for (s, n, na, tp, ig, soft) in cases:
            print("noalloc case prec=%d comp=%d n=%d n_active=%d" % (prec, comp, n, na), flush=True)
            want = _forces(L, hip, "workspace", s, n, na, tp, ig, prec, comp, soft)
            got = _forces(L, hip, "noalloc", s, n, na, tp, ig, prec, comp, soft)
            assert np.isfinite(want).all()
            assert np.array_equal(got, want), "no-allocation entry point differs (prec %d comp %d n %d)" % (
                prec, comp, n)
            hip.free()
        if prec == 1 and comp == 0:
            # EXACT also equals the reference's own sum on the committed fixture
            got = _forces(L, hip, "noalloc", st, 1024, 1024, 0, 0, prec, comp, 0.01)
            assert np.array_equal(got, golden("plummer1024_seed1_acc_basic.npy"))
    finally:
        hip.free()


@pytest.mark.gpu
def test_noalloc_rejects_oversized_input(gpu):
    """A body count whose partial sums exceed the fixed array is refused before any launch."""
    L = gpu.lib()
    cap = L.reb_amd_gravity_noalloc_max_bytes()
    assert cap == 32 << 20
    n = 4 << 20
    assert L.reb_amd_gravity_workspace_bytes(n, 0) > cap
    rc = L.reb_amd_gravity_direct_noalloc(None, None, None, None, None, None, None, n, n, 0, 0, 1.0, 0.01, 0, 0,
                                          None)
    assert rc == 4
    assert b"fixed" in L.reb_amd_last_error()
