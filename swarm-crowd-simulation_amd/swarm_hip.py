"""Python binding (ctypes) of libswarm_hip.so, the MI355X per-step crowd pipeline.

The C ABI is include/swarm_hip.h; this module only marshals numpy arrays through it. It fails loudly
when the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswarm_hip.so")

LAYOUT_ROOM, LAYOUT_BOWL, LAYOUT_CORRIDOR, LAYOUT_BOWL_EIGHTH, LAYOUT_PILLAR, LAYOUT_BOWL_QUARTER = range(6)
M_PER_UNIT = 31.25
S_PER_UNIT = 17.490671
INACTIVE_CELL = 0xFFFFFFFF
CHOICE_NONE = 0xFFFFFFFF

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")


class SceneDesc(C.Structure):
    _fields_ = [("xmin", C.c_float), ("ymin", C.c_float), ("xmax", C.c_float), ("ymax", C.c_float),
                ("hash_cell", C.c_float), ("sdf_spacing", C.c_float), ("sdf_clamp", C.c_float),
                ("nav_res_x", C.c_int), ("nav_res_y", C.c_int),
                ("nav_origin_x", C.c_float), ("nav_origin_y", C.c_float), ("nav_cell", C.c_float), ("nav_inv_cell", C.c_float),
                ("nav_body", C.c_float), ("metres_per_unit", C.c_double),
                ("n_walls", C.c_int), ("n_exits", C.c_int), ("n_stairs", C.c_int),
                ("walls", C.POINTER(C.c_float)), ("exits", C.POINTER(C.c_float)), ("exit_width_m", C.POINTER(C.c_float)),
                ("stairs", C.POINTER(C.c_float)), ("stair_factor", C.POINTER(C.c_float))]


class SceneInfo(C.Structure):
    _fields_ = [("xmin", C.c_float), ("ymin", C.c_float), ("xmax", C.c_float), ("ymax", C.c_float),
                ("hash_span_x", C.c_float), ("hash_span_y", C.c_float),
                ("hash_nx", C.c_int), ("hash_ny", C.c_int), ("perc_ring", C.c_int), ("proj_ring", C.c_int),
                ("sdf_nx", C.c_int), ("sdf_ny", C.c_int),
                ("sdf_ox", C.c_float), ("sdf_oy", C.c_float), ("sdf_h", C.c_float), ("sdf_inv_h", C.c_float),
                ("sdf_min_win", C.c_int), ("nav_res_x", C.c_int), ("nav_res_y", C.c_int),
                ("nav_origin_x", C.c_float), ("nav_origin_y", C.c_float), ("nav_cell", C.c_float), ("nav_inv_cell", C.c_float),
                ("K", C.c_int), ("E", C.c_int), ("seconds_per_unit", C.c_float), ("dist_quantum", C.c_float),
                ("build_seconds", C.c_double), ("n_walls", C.c_int), ("wall_bin_entries", C.c_int), ("from_cache", C.c_int)]


class Config(C.Structure):
    _fields_ = [("tau_s", C.c_float), ("obstacle_gain", C.c_float), ("proj_iters", C.c_int), ("use_graph", C.c_int)]


_lib = None


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.swarm_last_error.restype = C.c_char_p
    L.swarm_scene_build.argtypes = [C.POINTER(SceneDesc), C.c_float, C.c_float, C.c_int, C.POINTER(vp)]
    L.swarm_scene_build_layout.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, C.POINTER(vp)]
    L.swarm_scene_info_get.argtypes = [vp, C.POINTER(SceneInfo)]
    L.swarm_scene_read.argtypes = [vp, _f32p, _f32p, _u32p, _u32p, _u16p, _f32p, _f32p, _f32p]
    L.swarm_scene_destroy.argtypes = [vp]
    L.swarm_scene_destroy.restype = None
    L.swarm_place_agents.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_uint64, C.c_double, C.c_double,
                                     _f32p, _f32p, _f32p, _f32p, _f32p]
    L.swarm_layout_slots.argtypes = [C.c_int, vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.swarm_model_params.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.swarm_create.argtypes = [vp, C.c_int, C.POINTER(Config), C.POINTER(vp)]
    L.swarm_upload.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.swarm_restore_state.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.swarm_set_drive.argtypes = [vp, vp, vp]
    L.swarm_step.argtypes = [vp, C.c_int]
    L.swarm_step_exit_choice.argtypes = [vp]
    L.swarm_set_choice_state.argtypes = [vp, C.c_uint32, vp, vp, vp]
    L.swarm_sync.argtypes = [vp]
    L.swarm_set_proj_iters.argtypes = [vp, C.c_int]
    L.swarm_read_state.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.swarm_read_grid.argtypes = [vp, _u32p, _u32p]
    L.swarm_read_prev_state.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.swarm_sort_grid.argtypes = [vp, _u32p, _u32p, _u32p]
    L.swarm_read_cells.argtypes = [vp, _u32p, _u32p]
    L.swarm_read_assign.argtypes = [vp, _u32p]
    L.swarm_read_choice_counters.argtypes = [vp, _u32p]
    L.swarm_read_evac.argtypes = [vp, _i32p, _i32p, _u32p]
    L.swarm_read_control.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    L.swarm_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.swarm_set_timing.argtypes = [vp, C.c_int]
    L.swarm_read_timings.argtypes = [vp, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
    L.swarm_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.swarm_num_cells.argtypes = [vp]
    L.swarm_batch_create.argtypes = [C.POINTER(vp), C.c_int, _i32p, _i32p, C.c_int, C.POINTER(Config), C.POINTER(vp)]
    L.swarm_num_scenarios.argtypes = [vp]
    L.swarm_read_choice_counters_all.argtypes = [vp, _u32p]
    L.swarm_read_evac_counters.argtypes = [vp, _u32p]
    L.swarm_set_physics_shape.argtypes = [vp, C.c_int, C.c_int]
    L.swarm_set_projection_prune.argtypes = [vp, C.c_int]
    L.swarm_set_candidate_budget.argtypes = [vp, C.c_float]
    L.swarm_set_projection_lanes.argtypes = [vp, C.c_int]
    L.swarm_get_projection_lanes.argtypes = [vp, C.c_void_p]
    L.swarm_get_physics_shape.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.swarm_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.swarm_destroy.argtypes = [vp]
    L.swarm_destroy.restype = None
    _lib = L
    return L


OK, EINVAL, ESIZE, ENOMEM, ESCENE, EHIP, ENONFINITE = 0, -1, -2, -3, -4, -5, -6   # swarm_hip.h status codes


class NonFiniteGradient(RuntimeError):
    """swarm_adjoint_backward's SWARM_ENONFINITE: the gradient overflowed FP32; .grad_v0 / .grad_globals hold it."""


def check(rc):
    if rc != 0:
        raise RuntimeError(f"swarm error {rc}: {lib().swarm_last_error().decode()}")


def copy_probe(nbytes, reps=10):
    """Mean ms of one launch of a plain copy kernel of nbytes (read nbytes, write nbytes)."""
    ms = C.c_double()
    lib().swarm_copy_probe.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    check(lib().swarm_copy_probe(nbytes, reps, C.byref(ms)))
    return ms.value


def model_params(tau_s=0.5):
    tau, relax = C.c_float(), C.c_float()
    check(lib().swarm_model_params(tau_s, C.byref(tau), C.byref(relax)))
    return tau.value, relax.value


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Scene:
    """Built scene tables (SDF grid, navigation/candidate table, stair field, exits)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        inf = SceneInfo()
        check(lib().swarm_scene_info_get(self.h, C.byref(inf)))
        self.info = inf

    @classmethod
    def layout(cls, layout, max_radius=0.25 / M_PER_UNIT, max_v0=2.0 * S_PER_UNIT / M_PER_UNIT, threads=0, cache_dir=None):
        """A benchmark floor's tables; cache_dir: through the checksummed scene cache there."""
        h = C.c_void_p()
        if cache_dir:
            L = lib()
            L.swarm_scene_build_layout_cached.argtypes = [C.c_int, C.c_float, C.c_float, C.c_int, C.c_char_p, C.c_void_p]
            check(L.swarm_scene_build_layout_cached(layout, max_radius, max_v0, threads, str(cache_dir).encode(), C.byref(h)))
        else:
            check(lib().swarm_scene_build_layout(layout, max_radius, max_v0, threads, C.byref(h)))
        return cls(h)

    @classmethod
    def venue(cls, name, max_radius=0.0064, max_v0=0.75, threads=0, cache_dir=None):
        """The plan of the reference's seated venue S, M or L (walls, exits past the outer face; the layout for the step's
        incline rules)."""
        h = C.c_void_p()
        L = lib()
        L.swarm_scene_build_venue.argtypes = [C.c_char_p, C.c_float, C.c_float, C.c_int, C.c_char_p, C.c_void_p]
        check(L.swarm_scene_build_venue(name.encode(), max_radius, max_v0, threads,
                                        str(cache_dir).encode() if cache_dir else None, C.byref(h)))
        return cls(h)

    @classmethod
    def seated(cls, name, threads=0, cache_dir=None):
        """The reference's own 3D step on its seated venue S, M or L: the venue's grid, its navigation field and candidate
        table on the heightfield (the reference's bit for bit), its exit test; agents carry z (Sim.upload3 / state3)."""
        h = C.c_void_p()
        L = lib()
        L.swarm_scene_build_seated.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_void_p]
        check(L.swarm_scene_build_seated(name.encode(), threads, str(cache_dir).encode() if cache_dir else None, C.byref(h)))
        return cls(h)

    @classmethod
    def rects(cls, domain, hash_cell, sdf_spacing, sdf_clamp, nav, nav_body, walls, exits, exit_width_m=None,
              stairs=None, stair_factor=None, metres_per_unit=M_PER_UNIT, max_radius=0.0064, max_v0=0.75, threads=0):
        """A rectangle world. nav = (res_x, res_y, origin_x, origin_y, cell, inv_cell)."""
        walls = np.ascontiguousarray(np.asarray(walls, np.float32).reshape(-1, 4))
        exits = np.ascontiguousarray(np.asarray(exits, np.float32).reshape(-1, 4))
        ew = np.ascontiguousarray(np.zeros(len(exits), np.float32) if exit_width_m is None else np.asarray(exit_width_m, np.float32))
        st = np.ascontiguousarray(np.zeros((0, 4), np.float32) if stairs is None else np.asarray(stairs, np.float32).reshape(-1, 4))
        sf = np.ascontiguousarray(np.zeros(len(st), np.float32) if stair_factor is None else np.asarray(stair_factor, np.float32))
        d = SceneDesc()
        d.xmin, d.ymin, d.xmax, d.ymax = domain
        d.hash_cell, d.sdf_spacing, d.sdf_clamp = hash_cell, sdf_spacing, sdf_clamp
        d.nav_res_x, d.nav_res_y, d.nav_origin_x, d.nav_origin_y, d.nav_cell, d.nav_inv_cell = nav
        d.nav_body, d.metres_per_unit = nav_body, metres_per_unit
        d.n_walls, d.n_exits, d.n_stairs = len(walls), len(exits), len(st)
        d.walls, d.exits, d.exit_width_m, d.stairs, d.stair_factor = _fp(walls), _fp(exits), _fp(ew), _fp(st), _fp(sf)
        h = C.c_void_p()
        check(lib().swarm_scene_build(C.byref(d), max_radius, max_v0, threads, C.byref(h)))
        return cls(h)

    def venue_layout(self):
        """The layout a venue scene carries (a VenueLayout up to corner_door_lo; a 3D scene also its grid and exit test),
        None for other scenes."""
        kind, out = self.venue_kind()
        return out if kind else None

    def venue_kind(self):
        """(0 no venue, 1 a venue plan, 2 a venue's own 3D step; its VenueLayout)."""
        L = lib()
        L.swarm_scene_read_venue.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        has = C.c_int()
        out = VenueLayout()
        check(L.swarm_scene_read_venue(self.h, C.byref(has), C.byref(out)))
        return has.value, out

    def tables(self):
        """Copies of every table, as numpy arrays (the oracle's inputs)."""
        i = self.info
        cells = i.nav_res_x * i.nav_res_y
        t = dict(sdf=np.zeros(i.sdf_nx * i.sdf_ny, np.float32), sdf_min=np.zeros(i.sdf_nx * i.sdf_ny, np.float32),
                 nav_info=np.zeros(cells, np.uint32), nav_slot=np.zeros(cells * i.K, np.uint32),
                 nav_dist=np.zeros(cells * i.K, np.uint16), inv_cap_s=np.zeros(max(i.E, 1), np.float32),
                 stair=np.zeros(cells, np.float32), exit_rects=np.zeros(max(4 * i.E, 4), np.float32))
        check(lib().swarm_scene_read(self.h, t["sdf"], t["sdf_min"], t["nav_info"], t["nav_slot"], t["nav_dist"],
                                     t["inv_cap_s"], t["stair"], t["exit_rects"]))
        t["walls"] = np.zeros(max(4 * i.n_walls, 4), np.float32)
        t["wall_bin_start"] = np.zeros(i.hash_nx * i.hash_ny + 1, np.uint32)
        t["wall_bin_list"] = np.zeros(max(i.wall_bin_entries, 1), np.uint32)
        L = lib()
        L.swarm_scene_read_walls.argtypes = [C.c_void_p] + [C.c_void_p] * 3
        check(L.swarm_scene_read_walls(self.h, t["walls"].ctypes.data, t["wall_bin_start"].ctypes.data, t["wall_bin_list"].ctypes.data))
        return t

    def metric_grid(self):
        """The floor's safety-metric grid: (res_x, res_y, cells_per_unit, off_x, off_y, walkable m^2 per cell)."""
        L = lib()
        L.swarm_scene_metric_grid.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        rx, ry = C.c_int(), C.c_int()
        cpu, ox, oy = C.c_float(), C.c_float(), C.c_float()
        check(L.swarm_scene_metric_grid(self.h, C.byref(rx), C.byref(ry), C.byref(cpu), C.byref(ox), C.byref(oy), None))
        walk = np.zeros(rx.value * ry.value, np.float32)
        check(L.swarm_scene_metric_grid(self.h, None, None, None, None, None, walk.ctypes.data))
        return rx.value, ry.value, cpu.value, ox.value, oy.value, walk.reshape(ry.value, rx.value)

    def close(self):
        if self.h:
            lib().swarm_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def place_agents(layout, active, total, seed, speed_mean=1.34, speed_sd=0.26):
    a = {k: np.zeros(total, np.float32) for k in ("x", "y", "r", "v0", "w")}
    check(lib().swarm_place_agents(layout, active, total, seed, speed_mean, speed_sd, a["x"], a["y"], a["r"], a["v0"], a["w"]))
    a["vx"] = np.zeros(total, np.float32)
    a["vy"] = np.zeros(total, np.float32)
    return a


def spawn_seats(seat_count, n, load=1.0, seed=1337):
    """The reference's seat spawn: the seat of each of n agents (-1: parked)."""
    out = np.zeros(n, np.int32)
    L = lib()
    L.swarm_spawn_seats.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint32, C.c_void_p]
    check(L.swarm_spawn_seats(seat_count, n, load, seed, out.ctypes.data))
    return out


VENUE_MAX_GANGWAYS = 8


class VenueLayout(C.Structure):
    """swarm_venue_layout: SeatedLayout and the venue around it."""
    G = VENUE_MAX_GANGWAYS
    _fields_ = [(n, C.c_float) for n in ("stand_front", "row0", "row_depth", "clearway")] + [("rows", C.c_int)] + \
        [(n, C.c_float) for n in ("conc_front", "conc_back", "outer_face", "landing_end", "avenue_half", "lat0", "gangway_lo",
                                  "gangway_hi", "exit_lo", "exit_hi", "corner_inset", "seat_inset", "face_margin", "wall_above",
                                  "wall_clear", "first_row", "row_rise", "seat_top", "barrier_top", "concourse_z", "wall_top",
                                  "rake")] + \
        [("max_band_steps", C.c_int), ("seat_width", C.c_float), ("domain_half", C.c_float), ("gangways", C.c_int),
         ("gw_lo", C.c_float * G), ("gw_hi", C.c_float * G), ("gw_row", C.c_int * G), ("gw_seated", C.c_int * G),
         ("doors", C.c_int), ("door_lo", C.c_float * G), ("door_hi", C.c_float * G), ("corner_door_lo", C.c_float)] + \
        [(n, C.c_float) for n in ("domain_min", "domain_span", "exit_depth", "exit_max_z")] + \
        [(n, C.c_int) for n in ("grid_res", "metric_grid_res", "exit_count", "gangway_count", "seat_total")] + \
        [(n, C.c_double) for n in ("door_width_m", "corner_door_width_m", "total_exit_width_m", "required_exit_width_m")]


def venue_layout(name):
    """The layout of the reference's seated venue S, M or L."""
    out = VenueLayout()
    L = lib()
    L.swarm_venue_layout_get.argtypes = [C.c_char_p, C.c_void_p]
    check(L.swarm_venue_layout_get(name.encode(), C.byref(out)))
    return out


def venue_seats(name):
    """Every seat: x, y, z (n,) float32 and where (n, 5) int32 = quadrant, wing, row, segment, index."""
    L = lib()
    L.swarm_venue_seats.argtypes = [C.c_char_p] + [C.c_void_p] * 4 + [C.c_size_t, C.c_void_p]
    n = C.c_size_t()
    check(L.swarm_venue_seats(name.encode(), None, None, None, None, 0, C.byref(n)))
    x, y, z = (np.zeros(n.value, np.float32) for _ in range(3))
    where = np.zeros((n.value, 5), np.int32)
    check(L.swarm_venue_seats(name.encode(), x.ctypes.data, y.ctypes.data, z.ctypes.data, where.ctypes.data, n.value, C.byref(n)))
    return x, y, z, where


PARK_COORD = 1000.0


def venue_agents(name, n, load=1.0, seed=1337):
    """The reference's seated crowd (spawnSeatedCrowd, src/crowd_sim.cpp:677-697) on venue name: n agents, round(load x
    seats) of them on seats chosen by the seat spawn, at rest, the rest parked; the reference's global radius and free
    speed (BODY_RADIUS 0.2 m, V_MAX 1.34 m/s)."""
    x, y, _, _ = venue_seats(name)
    chosen = spawn_seats(len(x), n, load, seed)
    on = chosen >= 0
    a = {k: np.zeros(n, np.float32) for k in ("x", "y", "vx", "vy", "r", "v0", "w")}
    a["x"][:] = PARK_COORD
    a["y"][:] = PARK_COORD
    a["x"][on] = x[chosen[on]]
    a["y"][on] = y[chosen[on]]
    a["w"][on] = 1.0
    a["r"][:] = np.float32(0.0064)
    a["v0"][:] = np.float32(0.75)
    return a


def venue_probe(name, p, e, reach3):
    """The heightfield queries on the GPU at points p (n, 3) with directions e (n, 2): a dict of sdf_grad (n, 4),
    floor (n,), cell (n, 3), wall (n, 3, 3) (one row per reach), slope, exit, gangway (n,)."""
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(e, np.float32).reshape(-1, 2)
    r = np.ascontiguousarray(reach3, np.float32).reshape(3)
    n = len(p)
    o = dict(sdf_grad=np.zeros((n, 4), np.float32), floor=np.zeros(n, np.float32), cell=np.zeros((n, 3), np.int32),
             wall=np.zeros((n, 3, 3), np.float32), slope=np.zeros(n, np.float32), exit=np.zeros(n, np.int32),
             gangway=np.zeros(n, np.int32))
    L = lib()
    L.swarm_venue_probe.argtypes = [C.c_char_p, C.c_size_t] + [C.c_void_p] * 10
    check(L.swarm_venue_probe(name.encode(), n, p.ctypes.data, e.ctypes.data, r.ctypes.data, o["sdf_grad"].ctypes.data,
                              o["floor"].ctypes.data, o["cell"].ctypes.data, o["wall"].ctypes.data, o["slope"].ctypes.data,
                              o["exit"].ctypes.data, o["gangway"].ctypes.data))
    return o


def venue_exit_map(name):
    """The reference's exit id of each exit of the venue's plan."""
    L = lib()
    L.swarm_venue_exit_map.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_void_p]
    n = C.c_int()
    check(L.swarm_venue_exit_map(name.encode(), None, 0, C.byref(n)))
    out = np.zeros(n.value, np.int32)
    check(L.swarm_venue_exit_map(name.encode(), out.ctypes.data, n.value, C.byref(n)))
    return out


def spawn_uniform(layout, active, total, seed=1337, speed_mean=1.34, speed_sd=0.26):
    """An open floor: uniform placement with rejection, relaxed (agent dict as place_agents, plus the relaxation's
    sweeps and last largest overlap)."""
    a = {k: np.zeros(total, np.float32) for k in ("x", "y", "r", "v0", "w")}
    sweeps, mo = C.c_int(), C.c_float()
    L = lib()
    L.swarm_spawn_uniform.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_uint32, C.c_double, C.c_double] + [C.c_void_p] * 7
    check(L.swarm_spawn_uniform(layout, active, total, seed, speed_mean, speed_sd, a["x"].ctypes.data, a["y"].ctypes.data,
                                a["r"].ctypes.data, a["v0"].ctypes.data, a["w"].ctypes.data, C.byref(sweeps), C.byref(mo)))
    a["vx"] = np.zeros(total, np.float32)
    a["vy"] = np.zeros(total, np.float32)
    return a, sweeps.value, mo.value


def layout_capacity(layout):
    n = C.c_size_t()
    check(lib().swarm_layout_slots(layout, None, None, 0, C.byref(n)))
    return n.value


class Sim:
    """The GPU step of one scene (the reference's CrowdSim surface)."""

    def __init__(self, scene, n, tau_s=0.5, obstacle_gain=20.0, proj_iters=32, use_graph=True):
        cfg = Config(tau_s, obstacle_gain, proj_iters, 1 if use_graph else 0)
        h = C.c_void_p()
        check(lib().swarm_create(scene.h, n, C.byref(cfg), C.byref(h)))
        self.h, self.n, self.scene = h, n, scene
        self.cells = lib().swarm_num_cells(self.h)

    def upload(self, a):
        f = lambda k: np.ascontiguousarray(a[k], np.float32)
        check(lib().swarm_upload(self.h, f("x"), f("y"), f("vx"), f("vy"), f("w"), f("r"), f("v0")))

    def restore(self, a):
        f = lambda k: np.ascontiguousarray(a[k], np.float32)
        check(lib().swarm_restore_state(self.h, f("x"), f("y"), f("vx"), f("vy"), f("w")))

    def upload3(self, pos, vel):
        """A seated venue's own 3D step: pos, vel (n, 4) as GpuBoid's position (x, y, z, w) and velocity (vx, vy, vz,
        speed; the speed is recomputed)."""
        pos = np.asarray(pos, np.float32)
        vel = np.asarray(vel, np.float32)
        f = lambda a: np.ascontiguousarray(a, np.float32)
        L = lib()
        L.swarm_upload3.argtypes = [C.c_void_p] + [np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")] * 7
        check(L.swarm_upload3(self.h, f(pos[:, 0]), f(pos[:, 1]), f(pos[:, 2]), f(vel[:, 0]), f(vel[:, 1]), f(vel[:, 2]), f(pos[:, 3])))

    def state3(self):
        """(pos, vel) as (n, 4) arrays like GpuBoid's: (x, y, z, w) and (vx, vy, vz, speed)."""
        a = [np.zeros(self.n, np.float32) for _ in range(8)]
        L = lib()
        L.swarm_read_state3.argtypes = [C.c_void_p] + [np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")] * 8
        check(L.swarm_read_state3(self.h, *a))
        x, y, z, vx, vy, vz, sp, w = a
        return np.stack([x, y, z, w], 1), np.stack([vx, vy, vz, sp], 1)

    def set_drive(self, dx=None, dy=None):
        if dx is None:
            check(lib().swarm_set_drive(self.h, None, None))
        else:
            self._dx, self._dy = np.ascontiguousarray(dx, np.float32), np.ascontiguousarray(dy, np.float32)
            check(lib().swarm_set_drive(self.h, self._dx.ctypes.data, self._dy.ctypes.data))

    def step(self, k=1):
        check(lib().swarm_step(self.h, k))

    def sync(self):
        check(lib().swarm_sync(self.h))

    def choice_step(self):
        check(lib().swarm_step_exit_choice(self.h))

    def set_choice_state(self, step, assign=None, queue=None, ahead=None):
        keep = [None if a is None else np.ascontiguousarray(a, np.uint32) for a in (assign, queue, ahead)]
        ptr = [None if a is None else a.ctypes.data for a in keep]
        check(lib().swarm_set_choice_state(self.h, step, *ptr))

    def nav_probe(self, x, y, r, exit):
        """The GPU's navigation lookups at points: (cell, direction (m, 2), e_goal with the wall rule (m, 2))."""
        f = lambda a: np.ascontiguousarray(a, np.float32)
        x, y = f(x), f(y)
        r = f(np.broadcast_to(np.asarray(r, np.float32), x.shape))
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(exit, np.uint32), x.shape))
        cell = np.zeros(x.size, np.int32)
        d, g = np.zeros((x.size, 2), np.float32), np.zeros((x.size, 2), np.float32)
        L = lib()
        L.swarm_nav_probe.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 7
        check(L.swarm_nav_probe(self.h, x.size, x.ctypes.data, y.ctypes.data, r.ctypes.data, e.ctypes.data, cell.ctypes.data,
                                d.ctypes.data, g.ctypes.data))
        return cell, d, g

    def set_proj_iters(self, it):
        check(lib().swarm_set_proj_iters(self.h, it))

    def state(self):
        s = {k: np.zeros(self.n, np.float32) for k in ("x", "y", "vx", "vy", "speed", "w")}
        check(lib().swarm_read_state(self.h, s["x"], s["y"], s["vx"], s["vy"], s["speed"], s["w"]))
        return s

    def prev_state(self):
        s = {k: np.zeros(self.n, np.float32) for k in ("x", "y", "vx", "vy", "speed", "w")}
        check(lib().swarm_read_prev_state(self.h, s["x"], s["y"], s["vx"], s["vy"], s["speed"], s["w"]))
        return s

    def sort_grid(self, keys):
        """The sort stage alone on the given keys (cells or INACTIVE_CELL): (sorted keys, permutation)."""
        k = np.ascontiguousarray(keys, np.uint32)
        sk, ids = np.zeros(self.n, np.uint32), np.zeros(self.n, np.uint32)
        check(lib().swarm_sort_grid(self.h, k, sk, ids))
        return sk, ids

    def grid(self):
        key, idx = np.zeros(self.n, np.uint32), np.zeros(self.n, np.uint32)
        check(lib().swarm_read_grid(self.h, key, idx))
        return key, idx

    def cell_table(self):
        st, cn = np.zeros(self.cells, np.uint32), np.zeros(self.cells, np.uint32)
        check(lib().swarm_read_cells(self.h, st, cn))
        return st, cn

    def assign(self):
        a = np.zeros(self.n, np.uint32)
        check(lib().swarm_read_assign(self.h, a))
        return a

    def choice_counters(self):
        o = np.zeros(132, np.uint32)
        check(lib().swarm_read_choice_counters(self.h, o))
        return o

    def evac(self):
        st, ex = np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
        per = np.zeros(64, np.uint32)
        check(lib().swarm_read_evac(self.h, st, ex, per))
        return st, ex, per[: self.scene.info.E + 1]

    def control(self):
        s, a, t = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().swarm_read_control(self.h, C.byref(s), C.byref(a), C.byref(t)))
        return s.value, a.value, t.value

    def set_timing(self, on, launches=False):
        """Per-pass timing; launches=True also times each later projection iteration by its dispatch's events."""
        check(lib().swarm_set_timing(self.h, (2 if launches else 1) if on else 0))

    def timings(self):
        """Mean ms per step: grid passes, exit choice, predict, projection, final, total; one projection launch; steps."""
        t = np.zeros(8, np.float64)
        check(lib().swarm_read_timings(self.h, t))
        return t

    def time_steps(self, k):
        ms = C.c_double()
        check(lib().swarm_time_steps(self.h, k, C.byref(ms)))
        return ms.value

    def set_physics_shape(self, block=0, grid=0):
        check(lib().swarm_set_physics_shape(self.h, block, grid))

    def set_projection_prune(self, every=8):
        """Candidate pruning every `every` projection iterations (0: never); results are bit-identical."""
        check(lib().swarm_set_projection_prune(self.h, every))

    def set_candidate_budget(self, moves=12.0):
        """The projection's candidate lists' displacement budget in maximal moves per step (0: the cell ring every
        iteration; default 12); bit-identical results for every budget."""
        check(lib().swarm_set_candidate_budget(self.h, moves))

    def set_projection_lanes(self, lanes=1):
        """Lanes per slot in the projection's list iterations (1, 2, 4, 8); bit-identical results."""
        check(lib().swarm_set_projection_lanes(self.h, lanes))

    def candidate_lists(self):
        """(count[n], index[48, n], contact distance[48, n]) of the projection's candidates, slot order."""
        cnt = np.zeros(self.n, np.uint32)
        idx = np.zeros((48, self.n), np.uint32)
        d = np.zeros((48, self.n), np.float32)
        lib().swarm_read_candidates.argtypes = [C.c_void_p] * 4
        check(lib().swarm_read_candidates(self.h, cnt.ctypes.data, idx.ctypes.data, d.ctypes.data))
        return cnt, idx, d

    def set_sort_mode(self, bitonic):
        """The stable counting sort (False) or the reference GPU's bitonic network (True; power-of-two n)."""
        lib().swarm_set_sort_mode.argtypes = [C.c_void_p, C.c_int]
        check(lib().swarm_set_sort_mode(self.h, 1 if bitonic else 0))

    def set_predict_lanes(self, lanes=4):
        """Lanes per slot in predict's perception ring; bit-identical results."""
        lib().swarm_set_predict_lanes.argtypes = [C.c_void_p, C.c_int]
        check(lib().swarm_set_predict_lanes(self.h, lanes))

    def set_projection_collect_lanes(self, lanes=4):
        """Lanes per slot in the projection's first iteration (ring scan + candidates); bit-identical results."""
        lib().swarm_set_projection_collect_lanes.argtypes = [C.c_void_p, C.c_int]
        check(lib().swarm_set_projection_collect_lanes(self.h, lanes))

    def lanes(self):
        """(list iterations, collecting iteration, predict) lanes per slot in force."""
        L = lib()
        L.swarm_get_lanes.argtypes = [C.c_void_p] * 4
        a, b_, c = C.c_int(), C.c_int(), C.c_int()
        check(L.swarm_get_lanes(self.h, C.byref(a), C.byref(b_), C.byref(c)))
        return a.value, b_.value, c.value

    def projection_lanes(self):
        v = C.c_int()
        check(lib().swarm_get_projection_lanes(self.h, C.byref(v)))
        return v.value

    def physics_shape(self):
        b, g = C.c_int(), C.c_int()
        check(lib().swarm_get_physics_shape(self.h, C.byref(b), C.byref(g)))
        return b.value, g.value

    def autotune(self, steps=10):
        """Keeps the fastest physics launch shape over `steps` steps from the current state (the state is restored)."""
        ms = C.c_double()
        check(lib().swarm_autotune(self.h, steps, C.byref(ms)))
        return ms.value

    def stream(self):
        p = C.c_void_p()
        check(lib().swarm_get_stream(self.h, C.byref(p)))
        return p.value

    def close(self):
        if self.h:
            lib().swarm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSim(Sim):
    """Many independent scenarios stepped together, one launch per pass: scenario s runs
    layouts[scen_layout[s]] with scen_agents[s] agents. State arrays are the scenarios' arrays concatenated in
    scenario order; grid keys are global cells and grid ids global agents."""

    def __init__(self, layouts, scen_layout, scen_agents, tau_s=0.5, obstacle_gain=20.0, proj_iters=32, use_graph=True):
        cfg = Config(tau_s, obstacle_gain, proj_iters, 1 if use_graph else 0)
        hs = (C.c_void_p * len(layouts))(*[sc.h.value if isinstance(sc.h, C.c_void_p) else sc.h for sc in layouts])
        sl = np.ascontiguousarray(scen_layout, np.int32)
        sa = np.ascontiguousarray(scen_agents, np.int32)
        h = C.c_void_p()
        check(lib().swarm_batch_create(hs, len(layouts), sl, sa, len(sl), C.byref(cfg), C.byref(h)))
        self.h, self.n = h, int(sa.sum())
        self.layouts, self.scen_layout, self.scen_agents = list(layouts), sl, sa
        self.scene = layouts[int(sl[0])]
        self.cells = lib().swarm_num_cells(self.h)
        self.scenarios = lib().swarm_num_scenarios(self.h)
        self.agent_base = np.concatenate([[0], np.cumsum(sa)[:-1]]).astype(np.int64)
        nc = [layouts[int(l)].info.hash_nx * layouts[int(l)].info.hash_ny for l in sl]
        self.scen_cells = np.asarray(nc, np.int64)
        self.cell_base = np.concatenate([[0], np.cumsum(nc)[:-1]]).astype(np.int64)

    def choice_counters_all(self):
        o = np.zeros(132 * self.scenarios, np.uint32)
        check(lib().swarm_read_choice_counters_all(self.h, o))
        return o.reshape(self.scenarios, 132)

    def evac_counters(self):
        o = np.zeros(64 * self.scenarios, np.uint32)
        check(lib().swarm_read_evac_counters(self.h, o))
        return o.reshape(self.scenarios, 64)


BATCH_LAYOUTS = (LAYOUT_ROOM, LAYOUT_CORRIDOR, LAYOUT_BOWL_EIGHTH, LAYOUT_PILLAR)


def batch_workload(scenarios=256, seed=1337, lo=1024, hi=16384, load_lo=0.3, load_hi=1.0, layouts=BATCH_LAYOUTS):
    """The ensemble workload: scenario s runs layouts[s % len(layouts)] with N_s agents (log-uniform in [lo, hi],
    the scenario's padded size), of which A_s = min(N_s, floor(load_s * capacity)) are placed (load_s ~ U[load_lo,
    load_hi], placement seed seed + s) and the rest parked. Returns (scen_layout as indices into `layouts`, N_s,
    A_s, the concatenated agent arrays)."""
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), scenarios))).astype(np.int64)
    load = rng.uniform(load_lo, load_hi, scenarios)
    scen_layout = np.arange(scenarios, dtype=np.int32) % len(layouts)
    caps = [layout_capacity(l) for l in layouts]
    active = np.array([min(int(n[s]), int(np.floor(load[s] * caps[scen_layout[s]]))) for s in range(scenarios)], np.int64)
    parts = [place_agents(layouts[scen_layout[s]], int(active[s]), int(n[s]), seed + s) for s in range(scenarios)]
    agents = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    return scen_layout, n.astype(np.int32), active, agents


# ---------------- the adjoint of K steps (include/swarm_adjoint.h) ----------------
TAPE_NBR = 48


class WallEvent(C.Structure):
    _fields_ = [(f, C.c_float) for f in ("ax", "ay", "ux", "uy", "gx", "gy", "x0", "y0", "t0")] + [("vel", C.c_uint32)]


class PredictRec(C.Structure):
    _fields_ = ([("flags", C.c_uint32), ("leader", C.c_uint32)] +
                [(f, C.c_float) for f in ("anchor_x", "anchor_y", "egoal_x", "egoal_y", "stair", "sep_x", "sep_y",
                                          "sweep_lo", "sweep_dx", "sweep_dy")] +
                [("n_events", C.c_uint32), ("plane_kind", C.c_uint32)] +
                [(f, C.c_float) for f in ("plane_nx", "plane_ny", "pc_ux", "pc_uy", "pc_gx", "pc_gy", "pc_x0", "pc_y0", "pc_t0",
                                          "goal_wall_x", "goal_wall_y")] +
                [("ev", WallEvent * 6)])


class FinalRec(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("n_events", C.c_uint32), ("sweep_lo", C.c_float), ("sweep_dx", C.c_float),
                ("sweep_dy", C.c_float), ("ev", WallEvent * 6)]


class ExitRec(C.Structure):
    _fields_ = [("anchor_x", C.c_float), ("anchor_y", C.c_float), ("base", C.c_float), ("flags", C.c_uint32)]


class TapeView(C.Structure):
    _fields_ = ([(f, C.c_void_p) for f in ("x", "y", "vx", "vy", "w", "sid", "pred", "fin", "sx", "sy", "sr", "nbr", "nbr_count",
                                            "mask", "cap_bits", "plane_bits", "overlap")])


def _adjoint_lib():
    L = lib()
    if not getattr(L, "_adjoint_sigs", False):
        vp = C.c_void_p
        L.swarm_adjoint_forward.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.POINTER(C.c_double)]
        L.swarm_adjoint_backward.argtypes = [vp, _f32p, np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
        L.swarm_adjoint_tape.argtypes = [vp, C.c_int, C.POINTER(TapeView), vp]
        L.swarm_set_model.argtypes = [vp, C.c_float, C.c_float]
        L._adjoint_sigs = True
    return L


def new_tape(n, P):
    """Host arrays for one step's tape (swarm_adjoint_tape) of a context with n agents and P iterations."""
    t = dict(x=np.zeros(n, np.float32), y=np.zeros(n, np.float32), vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32),
             w=np.zeros(n, np.float32), sid=np.zeros(n, np.uint32), pred=(PredictRec * n)(), fin=(FinalRec * n)(),
             sx=np.zeros(n, np.float32), sy=np.zeros(n, np.float32), sr=np.zeros(n, np.float32),
             nbr=np.zeros(TAPE_NBR * n, np.uint32), nbr_count=np.zeros(n, np.uint32), mask=np.zeros(P * n, np.uint64),
             cap_bits=np.zeros(n, np.uint64), plane_bits=np.zeros(n, np.uint64), overlap=np.zeros(n, np.uint64))
    view = TapeView()
    for k, a in t.items():
        setattr(view, k, C.addressof(a) if isinstance(a, C.Array) else a.ctypes.data)
    t["view"] = view
    return t


def _sim_adjoint_forward(self, K, every=20, radius_scale=1.0, overlap_weight=0.1):
    """Runs K steps from the current state keeping checkpoints every `every` steps; returns the loss."""
    loss = C.c_double()
    check(_adjoint_lib().swarm_adjoint_forward(self.h, K, every, radius_scale, overlap_weight, C.byref(loss)))
    return loss.value


def _sim_adjoint_backward(self):
    """(dL/dv0 per agent, [dL/dtau_s, dL/dradius_scale, dL/dwall_gain] per scenario) of the last adjoint_forward."""
    gv0 = np.zeros(self.n, np.float32)
    g = np.zeros(3 * getattr(self, "scenarios", 1), np.float64)
    rc = _adjoint_lib().swarm_adjoint_backward(self.h, gv0, g)
    if rc == ENONFINITE:
        e = NonFiniteGradient(f"swarm error {rc}: {lib().swarm_last_error().decode()}")
        e.grad_v0, e.grad_globals = gv0, g
        raise e
    check(rc)
    return gv0, g


def _sim_adjoint_tape(self, step, tape, exit_recs=None):
    check(_adjoint_lib().swarm_adjoint_tape(self.h, step, C.byref(tape["view"]), None if exit_recs is None else C.addressof(exit_recs)))
    return tape


def _sim_set_tape_budget(self, nbytes):
    """A forward whose whole tape fits nbytes is taped whole (no recompute); 0: always checkpoint and recompute."""
    L = _adjoint_lib()
    L.swarm_adjoint_set_tape_budget.argtypes = [C.c_void_p, C.c_size_t]
    check(L.swarm_adjoint_set_tape_budget(self.h, nbytes))


def _sim_adjoint_segments(self):
    L = _adjoint_lib()
    L.swarm_adjoint_segments.argtypes = [C.c_void_p, C.c_void_p]
    k = C.c_int()
    check(L.swarm_adjoint_segments(self.h, C.byref(k)))
    return k.value


def _sim_tape_bytes(self):
    """Device bytes the adjoint holds (tape, buffers, checkpoints)."""
    L = _adjoint_lib()
    L.swarm_adjoint_memory.argtypes = [C.c_void_p, C.c_void_p]
    b = C.c_size_t()
    check(L.swarm_adjoint_memory(self.h, C.byref(b)))
    return b.value


def _sim_set_model(self, tau_s, obstacle_gain=20.0):
    check(_adjoint_lib().swarm_set_model(self.h, tau_s, obstacle_gain))


Sim.tape_bytes = _sim_tape_bytes
Sim.set_tape_budget = _sim_set_tape_budget
Sim.adjoint_segments = _sim_adjoint_segments
Sim.adjoint_forward = _sim_adjoint_forward
Sim.adjoint_backward = _sim_adjoint_backward
Sim.adjoint_tape = _sim_adjoint_tape
Sim.set_model = _sim_set_model


def _sim_set_adjoint_terms(self, step=None, gin=None, gout=None, exit_term=True, window=0, local=False):
    """Evacuation-crossing terms of the next adjoint_backward (None: none), the exit term, the truncation window,
    the local (per-agent) estimator."""
    L = _adjoint_lib()
    if not getattr(L, "_terms_sigs", False):
        L.swarm_adjoint_set_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.swarm_read_evac_points.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L._terms_sigs = True
    if step is None:
        check(L.swarm_adjoint_set_terms(self.h, None, None, None, 1 if exit_term else 0, window, 1 if local else 0))
        return
    self._terms = (np.ascontiguousarray(step, np.int32), np.ascontiguousarray(gin, np.float32), np.ascontiguousarray(gout, np.float32))
    s, i, o = self._terms
    check(L.swarm_adjoint_set_terms(self.h, s.ctypes.data, i.ctypes.data, o.ctypes.data, 1 if exit_term else 0, window,
                                    1 if local else 0))


def _sim_evac_points(self):
    """Per agent: the step it left at, its exit, its position at that step's input and one step earlier (n x 2)."""
    L = _adjoint_lib()
    L.swarm_read_evac_points.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    st, ex = np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
    ins, bef = np.zeros((self.n, 2), np.float32), np.zeros((self.n, 2), np.float32)
    check(L.swarm_read_evac_points(self.h, st.ctypes.data, ex.ctypes.data, ins.ctypes.data, bef.ctypes.data))
    return st, ex, ins, bef


Sim.set_adjoint_terms = _sim_set_adjoint_terms
Sim.evac_points = _sim_evac_points


# ---------------- the safety metrics (swarm_hip.h) ----------------
METRIC_COUNTER_WORDS = 68   # step, active, evacuated, activeAccum, exits[64]
METRIC_CELL_WORDS = 16      # lastMass, lastSpeedMass, peakMass, occupiedSteps, stepsAbove[4], 3 x 64-bit sums, pad[2]


def _metrics_lib():
    L = lib()
    if not getattr(L, "_metrics_sigs", False):
        vp = C.c_void_p
        L.swarm_metrics_enable.argtypes = [vp, C.c_int]
        L.swarm_metrics_reset.argtypes = [vp]
        L.swarm_metrics_configure.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp]
        L.swarm_metrics_grid.argtypes = [vp, vp, vp, vp, vp, vp]
        L.swarm_metrics_record.argtypes = [vp] + [vp] * 8
        L.swarm_read_metric_counters.argtypes = [vp, vp]
        L.swarm_read_metric_cells.argtypes = [vp, vp]
        L.swarm_read_metric_evac_steps.argtypes = [vp, vp]
        L.swarm_read_metric_cell_info.argtypes = [vp, vp, vp]
        L._metrics_sigs = True
    return L


def _sim_metrics_enable(self, on=True):
    """Every step followed by the metric passes (in the captured graph); the first enable clears the grid."""
    check(_metrics_lib().swarm_metrics_enable(self.h, 1 if on else 0))


def _sim_metrics_reset(self):
    check(_metrics_lib().swarm_metrics_reset(self.h))


def _sim_metrics_configure(self, res_x, res_y, cells_per_unit, off_x, off_y, thresholds):
    thr = np.ascontiguousarray(thresholds, np.uint32).reshape(-1)
    if thr.size != 4 * res_x * res_y:
        raise ValueError("4 threshold masses per cell")
    check(_metrics_lib().swarm_metrics_configure(self.h, res_x, res_y, cells_per_unit, off_x, off_y, thr.ctypes.data))


def _sim_metrics_grid(self):
    rx, ry = C.c_int(), C.c_int()
    cpu, ox, oy = C.c_float(), C.c_float(), C.c_float()
    check(_metrics_lib().swarm_metrics_grid(self.h, C.byref(rx), C.byref(ry), C.byref(cpu), C.byref(ox), C.byref(oy)))
    return rx.value, ry.value, cpu.value, ox.value, oy.value


def _sim_metrics_record(self, prev, cur):
    """The metric passes on the caller's states (agent order): prev x, y, w; cur x, y, w, vx, vy."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (prev["x"], prev["y"], prev["w"], cur["x"], cur["y"], cur["w"], cur["vx"], cur["vy"])]
    for a in arrs:
        if a.size != self.n:
            raise ValueError("metrics_record: n values per array")
    check(_metrics_lib().swarm_metrics_record(self.h, *[a.ctypes.data for a in arrs]))


def _sim_metric_counters(self):
    out = np.zeros(METRIC_COUNTER_WORDS, np.uint32)
    check(_metrics_lib().swarm_read_metric_counters(self.h, out.ctypes.data))
    return out


def _sim_metric_cells(self):
    rx, ry = self.metrics_grid()[:2]
    out = np.zeros((rx * ry, METRIC_CELL_WORDS), np.uint32)
    check(_metrics_lib().swarm_read_metric_cells(self.h, out.ctypes.data))
    return out


def _sim_metric_evac_steps(self):
    out = np.zeros(self.n, np.uint32)
    check(_metrics_lib().swarm_read_metric_evac_steps(self.h, out.ctypes.data))
    return out


def _sim_metric_cell_info(self):
    rx, ry = self.metrics_grid()[:2]
    walk = np.zeros(rx * ry, np.float32)
    thr = np.zeros((rx * ry, 4), np.uint32)
    check(_metrics_lib().swarm_read_metric_cell_info(self.h, walk.ctypes.data, thr.ctypes.data))
    return walk, thr


Sim.metrics_enable = _sim_metrics_enable
Sim.metrics_reset = _sim_metrics_reset
Sim.metrics_configure = _sim_metrics_configure
Sim.metrics_grid = _sim_metrics_grid
Sim.metrics_record = _sim_metrics_record
Sim.metric_counters = _sim_metric_counters
Sim.metric_cells = _sim_metric_cells
Sim.metric_evac_steps = _sim_metric_evac_steps
Sim.metric_cell_info = _sim_metric_cell_info
