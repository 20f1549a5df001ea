"""ctypes binding of the oracle (liboracle.so): test infrastructure only.

The oracle is the CPU restatement of the reference's CpuCrowdSim that the GPU path is checked against.
Only the tests, the smoke check and the benchmark's CPU-baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OracleScene(C.Structure):
    _fields_ = [("xmin", C.c_float), ("ymin", C.c_float), ("xmax", C.c_float), ("ymax", C.c_float),
                ("hashSpanX", C.c_float), ("hashSpanY", C.c_float), ("hashNX", C.c_int), ("hashNY", C.c_int),
                ("percRing", C.c_int), ("projRing", C.c_int),
                ("sdfNX", C.c_int), ("sdfNY", C.c_int), ("sdfOX", C.c_float), ("sdfOY", C.c_float), ("sdfInvH", C.c_float),
                ("sdf", C.c_void_p),
                ("navRX", C.c_int), ("navRY", C.c_int), ("navOX", C.c_float), ("navOY", C.c_float),
                ("navCell", C.c_float), ("navInvCell", C.c_float), ("K", C.c_int), ("E", C.c_int),
                ("navInfo", C.c_void_p), ("navSlot", C.c_void_p), ("navDist", C.c_void_p), ("invCapS", C.c_void_p),
                ("secondsPerUnit", C.c_float), ("distQuantum", C.c_float), ("stair", C.c_void_p), ("exitRects", C.c_void_p),
                ("nWalls", C.c_int), ("wallRects", C.c_void_p), ("venue", C.c_void_p), ("venue3d", C.c_int),
                ("exitDepth", C.c_float), ("exitMaxZ", C.c_float), ("domainMax", C.c_float)]


class OracleParams(C.Structure):
    _fields_ = [("relaxStep", C.c_float), ("tau", C.c_float), ("obstacleGain", C.c_float), ("projIters", C.c_int),
                ("threads", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleParams), C.c_int]
        L.oracle_create.restype = vp
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_state.argtypes = [vp] + [_f32p] * 7
        L.oracle_set_drive.argtypes = [vp, vp, vp]
        L.oracle_set_sort_mode.argtypes = [vp, C.c_int]
        L.oracle_restore_state.argtypes = [vp] + [_f32p] * 5
        L.oracle_choice_step.argtypes = [vp]
        L.oracle_set_choice_state.argtypes = [vp, C.c_uint32, vp, vp, vp]
        L.oracle_step.argtypes = [vp, C.c_int]
        L.oracle_get_state.argtypes = [vp] + [_f32p] * 6
        L.oracle_get_grid.argtypes = [vp, _u32p, _u32p]
        L.oracle_get_cells.argtypes = [vp, _u32p, _u32p]
        L.oracle_get_assign.argtypes = [vp, _u32p]
        L.oracle_get_choice_counters.argtypes = [vp, _u32p]
        L.oracle_get_evac.argtypes = [vp, _i32p, _i32p]
        L.oracle_num_steps.argtypes = [vp]
        L.oracle_get_times.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
        L.oracle_constants.argtypes = [_f32p, C.c_int]
        L.oracle_venue_new.argtypes = [vp] * 8
        L.oracle_venue_new.restype = vp
        L.oracle_venue_free.argtypes = [vp]
        L.oracle_venue_slope.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
        L.oracle_venue_slope.restype = C.c_float
        L.oracle_incline.argtypes = [C.c_int, vp, vp, vp]
        L.oracle_set_state3.argtypes = [vp] + [_f32p] * 7
        L.oracle_get_state3.argtypes = [vp] + [_f32p] * 8
        L.oracle_get_contacts.argtypes = [vp, vp]
        L.oracle_fnv64.argtypes = [vp, C.c_uint64]
        L.oracle_fnv64.restype = C.c_uint64
        _lib = L
    return _lib


def scene_struct(info, tables):
    """An OracleScene over the scene's numpy tables (keep 'tables' alive while it is used)."""
    s = OracleScene()
    i = info
    s.xmin, s.ymin, s.xmax, s.ymax = i.xmin, i.ymin, i.xmax, i.ymax
    s.hashSpanX, s.hashSpanY, s.hashNX, s.hashNY = i.hash_span_x, i.hash_span_y, i.hash_nx, i.hash_ny
    s.percRing, s.projRing = i.perc_ring, i.proj_ring
    s.sdfNX, s.sdfNY, s.sdfOX, s.sdfOY, s.sdfInvH = i.sdf_nx, i.sdf_ny, i.sdf_ox, i.sdf_oy, i.sdf_inv_h
    s.navRX, s.navRY, s.navOX, s.navOY = i.nav_res_x, i.nav_res_y, i.nav_origin_x, i.nav_origin_y
    s.navCell, s.navInvCell, s.K, s.E = i.nav_cell, i.nav_inv_cell, i.K, i.E
    s.secondsPerUnit, s.distQuantum = i.seconds_per_unit, i.dist_quantum
    s.sdf = tables["sdf"].ctypes.data
    s.navInfo = tables["nav_info"].ctypes.data
    s.navSlot = tables["nav_slot"].ctypes.data
    s.navDist = tables["nav_dist"].ctypes.data
    s.invCapS = tables["inv_cap_s"].ctypes.data
    s.stair = tables["stair"].ctypes.data
    s.exitRects = tables["exit_rects"].ctypes.data
    s.nWalls = int(getattr(i, "n_walls", 0))
    s.wallRects = tables["walls"].ctypes.data if s.nWalls > 0 else None
    return s


def bitonic_sort(keys):
    """The reference GPU's bitonic network on keys in agent order: (sorted keys, agents); n a power of two."""
    k = np.ascontiguousarray(keys, np.uint32)
    sk, ids = np.zeros_like(k), np.zeros_like(k)
    L = lib()
    L.oracle_bitonic_sort.argtypes = [C.c_int, _u32p, _u32p, _u32p]
    if L.oracle_bitonic_sort(k.size, k, sk, ids) != 0:
        raise ValueError("the bitonic network needs a power-of-two count")
    return sk, ids


class Oracle:
    def __init__(self, info, tables, n, relax_step, tau, obstacle_gain=20.0, proj_iters=32, threads=0, venue=None):
        """venue: a seated venue plan's layout arrays (venue_layout_arrays), the incline rules' input; None elsewhere."""
        self._tables = tables
        self.h = None
        self._venue = None
        sc = scene_struct(info, tables)
        if venue is not None:
            self._venue_arrays = [np.ascontiguousarray(a) for a in venue]
            self._venue = lib().oracle_venue_new(*[a.ctypes.data for a in self._venue_arrays])
            sc.venue = self._venue
        p = OracleParams(relax_step, tau, obstacle_gain, proj_iters, threads)
        self.h = lib().oracle_create(C.byref(sc), C.byref(p), n)
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.n = n
        self.cells = info.hash_nx * info.hash_ny

    def set_state(self, a):
        f = lambda k: np.ascontiguousarray(a[k], np.float32)
        lib().oracle_set_state(self.h, f("x"), f("y"), f("vx"), f("vy"), f("w"), f("r"), f("v0"))

    def restore(self, a):
        f = lambda k: np.ascontiguousarray(a[k], np.float32)
        lib().oracle_restore_state(self.h, f("x"), f("y"), f("vx"), f("vy"), f("w"))

    def choice_step(self):
        lib().oracle_choice_step(self.h)

    def set_choice_state(self, step, assign=None, queue=None, ahead=None):
        keep = [None if a is None else np.ascontiguousarray(a, np.uint32) for a in (assign, queue, ahead)]
        ptr = [None if a is None else a.ctypes.data for a in keep]
        lib().oracle_set_choice_state(self.h, step, *ptr)

    def set_sort_mode(self, bitonic):
        """The stable counting sort (False) or the GPU's bitonic network (True; a power-of-two agent count)."""
        if lib().oracle_set_sort_mode(self.h, 1 if bitonic else 0) != 0:
            raise ValueError("the bitonic network needs a power-of-two agent count")

    def set_drive(self, dx=None, dy=None):
        if dx is None:
            lib().oracle_set_drive(self.h, None, None)
        else:
            self._dx, self._dy = np.ascontiguousarray(dx, np.float32), np.ascontiguousarray(dy, np.float32)
            lib().oracle_set_drive(self.h, self._dx.ctypes.data, self._dy.ctypes.data)

    def step(self, k=1):
        lib().oracle_step(self.h, k)

    def state(self):
        s = {k: np.zeros(self.n, np.float32) for k in ("x", "y", "vx", "vy", "speed", "w")}
        lib().oracle_get_state(self.h, s["x"], s["y"], s["vx"], s["vy"], s["speed"], s["w"])
        return s

    def grid(self):
        key, idx = np.zeros(self.n, np.uint32), np.zeros(self.n, np.uint32)
        lib().oracle_get_grid(self.h, key, idx)
        return key, idx

    def cell_table(self):
        st, cn = np.zeros(self.cells, np.uint32), np.zeros(self.cells, np.uint32)
        lib().oracle_get_cells(self.h, st, cn)
        return st, cn

    def goal(self, x, y, r, exit, wall_rule=True):
        """e_goal at points: the table's direction (choiceDirection), with the wall rule (choiceGoal) unless
        wall_rule is False: (ex, ey, applied)."""
        f = lambda a: np.ascontiguousarray(a, np.float32)
        x, y = f(x), f(y)
        r = f(np.broadcast_to(np.asarray(r, np.float32), x.shape))
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(exit, np.uint32), x.shape))
        ex, ey, ap = np.zeros(x.size, np.float32), np.zeros(x.size, np.float32), np.zeros(x.size, np.uint8)
        L = lib()
        L.oracle_goal.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4 + [C.c_int] + [C.c_void_p] * 3
        L.oracle_goal(self.h, x.size, x.ctypes.data, y.ctypes.data, r.ctypes.data, e.ctypes.data, int(bool(wall_rule)),
                      ex.ctypes.data, ey.ctypes.data, ap.ctypes.data)
        return ex, ey, ap.astype(bool)

    def choice_cell(self, x, y):
        x, y = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32)
        out = np.zeros(x.size, np.int32)
        L = lib()
        L.oracle_choice_cell.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 3
        L.oracle_choice_cell(self.h, x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
        return out

    def exit_at(self, x, y):
        x, y = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32)
        out = np.zeros(x.size, np.int32)
        L = lib()
        L.oracle_exit_at.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 3
        L.oracle_exit_at(self.h, x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
        return out

    def wall(self, x, y, reach):
        """The exact wall query at points: (direction x, direction y, clearance; 1e9 when no wall within reach)."""
        x, y = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32)
        wx, wy, wd = (np.zeros(x.size, np.float32) for _ in range(3))
        L = lib()
        L.oracle_wall.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float] + [C.c_void_p] * 3
        L.oracle_wall(self.h, x.size, x.ctypes.data, y.ctypes.data, reach, wx.ctypes.data, wy.ctypes.data, wd.ctypes.data)
        return wx, wy, wd

    def assign(self):
        a = np.zeros(self.n, np.uint32)
        lib().oracle_get_assign(self.h, a)
        return a

    def choice_counters(self):
        o = np.zeros(132, np.uint32)
        lib().oracle_get_choice_counters(self.h, o)
        return o

    def evac(self):
        st, ex = np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
        lib().oracle_get_evac(self.h, st, ex)
        return st, ex

    def times(self):
        t = np.zeros(8, np.float64)
        lib().oracle_get_times(self.h, t)
        return t

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None
        if self._venue:
            lib().oracle_venue_free(self._venue)
            self._venue = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def constants():
    out = np.zeros(64, np.float32)
    n = lib().oracle_constants(out, 64)
    return out[:n]


class AdjRef:
    """FP64 reference of the adjoint's frozen-decision model (forward-mode derivatives and finite differences)."""

    def __init__(self, info, tables, n, P, K, overlap_weight, a, tau_s, radius_scale, gain):
        L = lib()
        if not getattr(L, "_adjref_sigs", False):
            vp = C.c_void_p
            L.adjref_create.restype = vp
            L.adjref_create.argtypes = [C.POINTER(OracleScene), C.c_int, C.c_int, C.c_int, C.c_float] + [_f32p] * 7 + [C.c_float] * 3
            L.adjref_destroy.argtypes = [vp]
            L.adjref_add_step.argtypes = [vp, C.c_int, vp]
            L.adjref_set_exit.argtypes = [vp, vp]
            L.adjref_loss.restype = C.c_double
            L.adjref_loss.argtypes = [vp, vp, C.c_double, C.POINTER(C.c_double)]
            L.adjref_directional.argtypes = [vp, C.c_int, _f64p, _f64p, C.POINTER(C.c_double), C.c_int]
            L._adjref_sigs = True
        self._tables = tables
        self._scene = scene_struct(info, tables)
        f = lambda k: np.ascontiguousarray(a[k], np.float32)
        self.n = n
        self.h = L.adjref_create(C.byref(self._scene), n, P, K, overlap_weight, f("x"), f("y"), f("vx"), f("vy"), f("w"), f("r"),
                                 f("v0"), tau_s, radius_scale, gain)

    def add_step(self, step, tape_view):
        assert lib().adjref_add_step(self.h, step, C.addressof(tape_view)) == 0

    def set_exit(self, exit_recs):
        lib().adjref_set_exit(self.h, C.addressof(exit_recs))

    def loss(self, direction=None, eps=0.0):
        err = C.c_double()
        d = None if direction is None else np.ascontiguousarray(direction, np.float64)
        v = lib().adjref_loss(self.h, None if d is None else d.ctypes.data, eps, C.byref(err))
        return v, err.value

    def step_errors(self, K):
        e, a = np.zeros(K, np.float64), np.zeros(K, np.int32)
        lib().adjref_step_errors.argtypes = [C.c_void_p, _f64p, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")]
        lib().adjref_step_errors(self.h, e, a)
        return e, a

    def directional(self, dirs, anchored=False):
        """Directional derivatives (forward mode); anchored: along the tape's FP32 trajectory."""
        dirs = np.ascontiguousarray(dirs, np.float64)
        out = np.zeros(dirs.shape[0], np.float64)
        loss = C.c_double()
        lib().adjref_directional(self.h, dirs.shape[0], dirs.reshape(-1), out, C.byref(loss), 1 if anchored else 0)
        return out, loss.value

    def close(self):
        if self.h:
            lib().adjref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------- safety metrics (oracle/metrics_oracle.cpp) ----------------
COUNTER_WORDS = 68   # step, active, evacuated, activeAccum, exits[64]
CELL_WORDS = 16      # lastMass, lastSpeedMass, peakMass, occupiedSteps, stepsAbove[4], massSteps, speedMassSteps, exposure (u64 each), pad


def _metrics_sigs(L):
    if getattr(L, "_metrics_sigs", False):
        return L
    u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    L.oracle_metrics_cic.argtypes = [C.c_int, f32, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, u32, u32]
    L.oracle_metrics_speed_mass.argtypes = [C.c_int, u32, f32, u32]
    L.oracle_metrics_speed_per_unit.restype = C.c_float
    L.oracle_metrics_thresholds.argtypes = [C.c_int, f32, u8, u32]
    L.oracle_metrics_record.argtypes = ([C.c_int] + [f32] * 8 + [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, u32, C.c_int, f32, u32,
                                                                 u32, u32])
    L._metrics_sigs = True
    return L


def metrics_cic(xy, res_x, res_y, cells_per_unit, off_x, off_y):
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    n = xy.shape[0]
    cells, w = np.zeros((n, 4), np.uint32), np.zeros((n, 4), np.uint32)
    _metrics_sigs(lib()).oracle_metrics_cic(n, xy.reshape(-1), res_x, res_y, cells_per_unit, off_x, off_y, cells.reshape(-1), w.reshape(-1))
    return cells, w


def metrics_speed_mass(weight, mps):
    weight = np.ascontiguousarray(weight, np.uint32)
    mps = np.ascontiguousarray(mps, np.float32)
    out = np.zeros(weight.size, np.uint32)
    _metrics_sigs(lib()).oracle_metrics_speed_mass(weight.size, weight, mps, out)
    return out


def metrics_thresholds(walkable, included):
    walkable = np.ascontiguousarray(walkable, np.float32)
    inc = np.ascontiguousarray(included, np.uint8)
    out = np.zeros((walkable.size, 4), np.uint32)
    _metrics_sigs(lib()).oracle_metrics_thresholds(walkable.size, walkable, inc, out.reshape(-1))
    return out


class MetricsOracle:
    """The metric passes' state on the CPU: counters, per-agent evacuation steps, per-cell accumulators."""

    def __init__(self, n, res_x, res_y, cells_per_unit, off_x, off_y, thresholds, rects=None):
        self.n, self.grid = n, (res_x, res_y, float(cells_per_unit), float(off_x), float(off_y))
        self.thr = np.ascontiguousarray(thresholds, np.uint32).reshape(-1)
        self.rects = np.ascontiguousarray(np.zeros(4, np.float32) if rects is None else rects, np.float32).reshape(-1)
        self.E = 0 if rects is None else self.rects.size // 4
        self.counters = np.zeros(COUNTER_WORDS, np.uint32)
        self.evac_step = np.zeros(n, np.uint32)
        self.cells = np.zeros((res_x * res_y, CELL_WORDS), np.uint32)

    def record(self, prev, cur):
        f = lambda a: np.ascontiguousarray(a, np.float32)
        _metrics_sigs(lib()).oracle_metrics_record(self.n, f(prev["x"]), f(prev["y"]), f(prev["w"]), f(cur["x"]), f(cur["y"]), f(cur["w"]),
                                                    f(cur["vx"]), f(cur["vy"]), *self.grid, self.thr, self.E, self.rects, self.counters,
                                                    self.evac_step, self.cells.reshape(-1))


class SeatedOracle(Oracle):
    """The oracle's seated 3D step on an analytic venue: venue = its layout arrays (venue_layout_arrays), grid = (domain_min,
    domain_span, grid_res), exit = (exit_depth, exit_max_z, domain_max); nav = (info, tables) of the venue's navigation
    scene for the exit choice, or None (forced drive only)."""

    def __init__(self, venue, grid, exit, n, nav=None, threads=0, proj_iters=32):
        self._tables = None
        self.h = None
        self._venue_arrays = [np.ascontiguousarray(a) for a in venue]
        self._venue = lib().oracle_venue_new(*[a.ctypes.data for a in self._venue_arrays])
        dmin, span, res = grid
        if nav is not None:
            info, tables = nav
            self._tables = tables
            sc = scene_struct(info, tables)
        else:
            sc = OracleScene()
            self._dummy = np.zeros(4, np.float32)
            sc.sdfNX = sc.sdfNY = 2
            sc.sdf = self._dummy.ctypes.data
            sc.sdfInvH = 1.0
            sc.navRX = sc.navRY = 0
        sc.xmin = sc.ymin = dmin
        sc.xmax = sc.ymax = dmin + span
        sc.hashSpanX = sc.hashSpanY = span
        sc.hashNX = sc.hashNY = res
        sc.percRing = sc.projRing = 1
        sc.venue = self._venue
        sc.venue3d = 1
        sc.exitDepth, sc.exitMaxZ, sc.domainMax = exit
        p = OracleParams(0.0, 0.0, 20.0, proj_iters, threads)
        self.h = lib().oracle_create(C.byref(sc), C.byref(p), n)
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.n = n
        self.cells = res * res

    def set_state3(self, pos, vel):
        """pos, vel: (n, 4) as GpuBoid's position (xyz, w) and velocity (xyz, speed)."""
        pos = np.asarray(pos, np.float32)
        vel = np.asarray(vel, np.float32)
        f = lambda a: np.ascontiguousarray(a, np.float32)
        lib().oracle_set_state3(self.h, f(pos[:, 0]), f(pos[:, 1]), f(pos[:, 2]), f(vel[:, 0]), f(vel[:, 1]), f(vel[:, 2]), f(pos[:, 3]))

    def state3(self):
        """(pos, vel) as (n, 4) arrays like GpuBoid's."""
        a = [np.zeros(self.n, np.float32) for _ in range(8)]
        lib().oracle_get_state3(self.h, *a)
        x, y, z, vx, vy, vz, sp, w = a
        return np.stack([x, y, z, w], 1), np.stack([vx, vy, vz, sp], 1)

    def contacts(self):
        out = np.zeros(self.n, np.uint8)
        lib().oracle_get_contacts(self.h, out.ctypes.data)
        return out


# --- the analytic seated venue (venue_oracle.cpp) ---
VENUE_REACHES = (0.008, 0.016, 0.02)   # FEELER_REACH, PLANE_PROBE_REACH (model units) and a wide reach


def venue_layout_arrays(kat, name):
    """The layout of venue name as the oracle's inputs, from the reference's dump in venue_kat.npz."""
    g = lambda k: np.ascontiguousarray(kat["venue__%s__%s" % (name, k)])
    return (g("layout_f").astype(np.float32), g("layout_i").astype(np.int32), g("layout_gw").astype(np.float32),
            g("layout_gw_hi").astype(np.float32), g("layout_gw_row").astype(np.int32), g("layout_gw_seated").astype(np.int32),
            g("layout_door_lo").astype(np.float32), g("layout_door_hi").astype(np.float32))


_LAYOUT_F = ("stand_front", "row0", "row_depth", "clearway", "conc_front", "conc_back", "outer_face", "landing_end", "avenue_half",
             "lat0", "gangway_lo", "gangway_hi", "exit_lo", "exit_hi", "corner_inset", "seat_inset", "face_margin", "wall_above",
             "wall_clear", "first_row", "row_rise", "seat_top", "barrier_top", "concourse_z", "wall_top", "rake", "seat_width",
             "domain_half", "corner_door_lo")


def venue_layout_arrays_of(lay):
    """The oracle's layout arrays from a swarm_venue_layout (its field names), in venue_layout_arrays' order."""
    f = lambda v: np.ascontiguousarray(np.asarray(v, np.float32))
    i = lambda v: np.ascontiguousarray(np.asarray(v, np.int32))
    return (f([getattr(lay, k) for k in _LAYOUT_F]), i([lay.rows, lay.max_band_steps, lay.gangways, lay.doors, 0]),
            f(list(lay.gw_lo)), f(list(lay.gw_hi)), i(list(lay.gw_row)), i(list(lay.gw_seated)), f(list(lay.door_lo)),
            f(list(lay.door_hi)))


def fnv64(a):
    """FNV-1a 64 of an array's bytes."""
    a = np.ascontiguousarray(a)
    return int(lib().oracle_fnv64(a.ctypes.data, a.nbytes))


def incline(g):
    """seated::inclineFreeSpeed and inclineHeadwayTime (EMPIRICAL) at slopes g: (v_free, t_headway) float32."""
    g = np.ascontiguousarray(g, np.float32)
    v, t = np.zeros_like(g), np.zeros_like(g)
    lib().oracle_incline(len(g), g.ctypes.data, v.ctypes.data, t.ctypes.data)
    return v, t


def _venue_sigs(L):
    L.oracle_venue_cells.argtypes = [C.c_void_p] * 8 + [C.c_int, C.c_void_p, C.c_void_p]
    L.oracle_venue_cells.restype = C.c_int
    L.oracle_venue_eval.argtypes = [C.c_void_p] * 8 + [C.c_int] + [C.c_void_p] * 10


def venue_cells(layout):
    """The oracle's cell table: (band, index, kind, row) int32 and (l0, l1, lo, hi) float32 rows."""
    L = lib()
    _venue_sigs(L)
    ptrs = [a.ctypes.data for a in layout]
    n = L.oracle_venue_cells(*ptrs, 0, None, None)
    ci = np.zeros((n, 4), np.int32)
    cf = np.zeros((n, 4), np.float32)
    L.oracle_venue_cells(*ptrs, n, ci.ctypes.data, cf.ctypes.data)
    return ci, cf


def venue_eval(layout, p, e, reach3=VENUE_REACHES):
    """Every heightfield query at points p (n, 3) with directions e (n, 2), as the GPU probe returns them."""
    L = lib()
    _venue_sigs(L)
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(e, np.float32).reshape(-1, 2)
    r = np.ascontiguousarray(reach3, np.float32)
    n = len(p)
    o = dict(sdf_grad=np.zeros((n, 4), np.float32), floor=np.zeros(n, np.float32), cell=np.zeros((n, 3), np.int32),
             wall=np.zeros((n, 3, 3), np.float32), slope=np.zeros(n, np.float32), exit=np.zeros(n, np.int32),
             gangway=np.zeros(n, np.int32))
    L.oracle_venue_eval(*[a.ctypes.data for a in layout], n, p.ctypes.data, e.ctypes.data, r.ctypes.data, o["sdf_grad"].ctypes.data,
                        o["floor"].ctypes.data, o["cell"].ctypes.data, o["wall"].ctypes.data, o["slope"].ctypes.data,
                        o["exit"].ctypes.data, o["gangway"].ctypes.data)
    return o
