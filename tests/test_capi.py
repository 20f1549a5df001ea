"""The C ABI library: it loads, exports every function include/*.h declares, and rejects bad arguments
with the documented codes (no GPU compute here: the calls fail before touching a device)."""
import ctypes as C
import glob
import os
import re

import numpy as np

import swarm_hip as sh
from common import HERE

REPO = os.path.dirname(HERE)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(swarm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(sh.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bad_arguments_are_rejected():
    L = sh.lib()
    h = C.c_void_p()
    assert L.swarm_scene_build_layout(99, 0.008, 1.1, 1, C.byref(h)) == -1   # SWARM_EINVAL: unknown layout
    assert b"unknown layout" in L.swarm_last_error()
    assert L.swarm_step(None, 1) == -1
    assert L.swarm_model_params(-1.0, C.byref(C.c_float()), C.byref(C.c_float())) == -1
    x = np.zeros(4, np.float32)
    assert L.swarm_place_agents(sh.LAYOUT_ROOM, 10**7, 10**7, 1, 1.34, 0.26, x, x, x, x, x) == -1


def test_model_params_fold():
    tau, relax = sh.model_params(0.5)
    assert np.float32(tau) == np.float32(float.fromhex("0x1.d45d2cp-6"))     # sim_params.hpp TAU_RELAX
    assert np.float32(relax) == np.float32(float.fromhex("0x1.68de1cp-1"))   # RELAX_STEP


def test_batch_create_rejects_bad_arguments():
    L = sh.lib()
    h = C.c_void_p()
    empty = (C.c_void_p * 1)()
    sl = np.zeros(1, np.int32)
    sa = np.ones(1, np.int32)
    assert L.swarm_batch_create(None, 1, sl, sa, 1, None, C.byref(h)) == -1
    assert L.swarm_batch_create(empty, 0, sl, sa, 1, None, C.byref(h)) == -1
    assert L.swarm_batch_create(empty, 1, sl, sa, 0, None, C.byref(h)) == -1
    assert L.swarm_batch_create(empty, 1, sl, sa, 1, None, C.byref(h)) == -1   # a null scene
    assert b"null scene" in L.swarm_last_error()
    assert L.swarm_autotune(None, 3, None) == -1
    assert L.swarm_set_physics_shape(None, 128, 0) == -1
    assert L.swarm_set_projection_prune(None, 8) == -1
    assert L.swarm_set_candidate_budget(None, C.c_float(1.0)) == -1
    assert L.swarm_set_projection_lanes(None, 4) == -1


def test_metrics_entries_reject_null_contexts():
    L = sh.lib()
    out = np.zeros(68, np.uint32)
    assert L.swarm_metrics_enable(None, 1) == -1
    assert L.swarm_metrics_reset(None) == -1
    assert L.swarm_metrics_configure(None, 4, 4, C.c_float(1.0), C.c_float(0.0), C.c_float(0.0), out.ctypes.data) == -1
    assert L.swarm_metrics_grid(None, None, None, None, None, None) == -1
    assert L.swarm_metrics_record(None, *([None] * 8)) == -1
    assert L.swarm_read_metric_counters(None, out.ctypes.data) == -1
    assert L.swarm_read_metric_cells(None, out.ctypes.data) == -1
    assert L.swarm_read_metric_evac_steps(None, out.ctypes.data) == -1
    assert L.swarm_read_metric_cell_info(None, None, None) == -1
    assert L.swarm_scene_metric_grid(None, None, None, None, None, None, None) == -1


def test_venue_entries_reject_bad_arguments():
    L = C.CDLL(sh.LIB_PATH)
    L.swarm_venue_seats.argtypes = [C.c_char_p] + [C.c_void_p] * 4 + [C.c_size_t, C.c_void_p]
    L.swarm_venue_exit_map.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_void_p]
    L.swarm_venue_layout_get.argtypes = [C.c_char_p, C.c_void_p]
    L.swarm_venue_probe.argtypes = [C.c_char_p, C.c_size_t] + [C.c_void_p] * 10
    out = sh.VenueLayout()
    assert L.swarm_venue_layout_get(b"XL", C.byref(out)) == sh.EINVAL   # unknown venue
    assert L.swarm_venue_layout_get(None, C.byref(out)) == sh.EINVAL
    n = C.c_size_t()
    assert L.swarm_venue_seats(b"S", None, None, None, None, 0, C.byref(n)) == 0 and n.value == 3728
    x = np.zeros(100, np.float32)
    assert L.swarm_venue_seats(b"S", x.ctypes.data, None, None, None, 100, C.byref(n)) == sh.ESIZE   # too small
    m = C.c_int()
    e = np.zeros(4, np.int32)
    assert L.swarm_venue_exit_map(b"M", e.ctypes.data, 4, C.byref(m)) == sh.ESIZE and m.value == 36
    # points without their arrays: refused before any device work
    assert L.swarm_venue_probe(b"S", 10, None, None, None, None, None, None, None, None, None, None) == sh.EINVAL
