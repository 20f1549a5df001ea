"""swarm_sort_keys_noalloc (the sort stage on device arrays, no allocation) against swarm_sort_grid.

Two calls in a row on the same GPU (the second with fewer keys and fewer cells than the first, so the scratch array
holds the first call's words), each compared bit for bit with swarm_sort_grid + swarm_read_cells of a context on a
floor with that many cells: sorted keys, permutation, cell starts and counts. Keys: random cells with inactive ones,
and one crowded cell (more agents than the counting sort's rank scan takes). Limits are refused with SWARM_EINVAL.
Device memory for the inputs and outputs comes from the HIP runtime the library itself uses.
"""
import ctypes as C

import numpy as np
import pytest

import swarm_hip as sh

pytestmark = pytest.mark.gpu

INACTIVE = 0xFFFFFFFF


class Hip:
    def __init__(self):
        sh.lib()   # the library's own HIP runtime is then the process's
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, n, dtype=np.uint32):
        out = np.zeros(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def noalloc(hip, keys, cells):
    n = keys.size
    L = sh.lib()
    L.swarm_sort_keys_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32] + [C.c_void_p] * 5
    L.swarm_noalloc_last_error.restype = C.c_char_p
    dk = hip.put(keys.astype(np.uint32))
    sk, ids, st, cn = hip.alloc(4 * n), hip.alloc(4 * n), hip.alloc(4 * cells), hip.alloc(4 * cells)
    rc = L.swarm_sort_keys_noalloc(dk, n, cells, sk, ids, st, cn, None)
    assert rc == 0, L.swarm_noalloc_last_error().decode()
    assert hip.L.hipDeviceSynchronize() == 0
    return hip.get(sk, n), hip.get(ids, n), hip.get(st, cells), hip.get(cn, cells)


def reference(scene, keys):
    sim = sh.Sim(scene, keys.size)
    sk, ids = sim.sort_grid(keys.astype(np.uint32))
    st, cn = sim.cell_table()
    sim.close()
    return sk, ids, st, cn


def make_keys(rng, n, cells, crowd):
    k = rng.integers(0, cells, n).astype(np.uint32)
    k[rng.random(n) < 0.2] = INACTIVE
    k[rng.choice(n, crowd, replace=False)] = np.uint32(cells // 2)   # one crowded cell
    return k


def test_sort_noalloc_matches_sort_grid_twice():
    hip = Hip()
    rng = np.random.default_rng(77)
    try:
        for layout, n, crowd in ((sh.LAYOUT_BOWL, 65536, 300), (sh.LAYOUT_ROOM, 4000, 40)):
            scene = sh.Scene.layout(layout)
            cells = scene.info.hash_nx * scene.info.hash_ny
            keys = make_keys(rng, n, cells, crowd)
            got = noalloc(hip, keys, cells)
            ref = reference(scene, keys)
            for g, r, what in zip(got, ref, ("sorted keys", "ids", "cell starts", "cell counts")):
                assert np.array_equal(g, r), f"{n} keys, {cells} cells: {what}"
            hip.free()
    finally:
        hip.free()


def test_sort_noalloc_refuses_sizes_beyond_its_limits():
    L = sh.lib()
    L.swarm_sort_keys_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32] + [C.c_void_p] * 5
    L.swarm_noalloc_last_error.restype = C.c_char_p
    assert L.swarm_sort_keys_noalloc(None, 65537, 100, None, None, None, None, None) == -1
    assert b"n must be" in L.swarm_noalloc_last_error()
    assert L.swarm_sort_keys_noalloc(None, 100, 262145, None, None, None, None, None) == -1
    assert b"cells must be" in L.swarm_noalloc_last_error()
