// extern "C" layer over HipCrowdSim and the scene builder (include/swarm_hip.h). No exception crosses it:
// std::invalid_argument -> SWARM_EINVAL, a device allocation failure -> SWARM_ENOMEM, other errors -> SWARM_EHIP.
#include "../../include/swarm_adjoint.h"
#include "../../include/swarm_hip.h"

#include <cstddef>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "hip_crowd_sim.hpp"
#include "scene.hpp"
#include "venue.hpp"
#include "venue_probe.h"

struct swarm_scene { swarm::Scene s; };
struct swarm_ctx { swarm::HipCrowdSim* sim; };

namespace {
thread_local std::string g_err;

struct SizeError : std::runtime_error {   // a caller's array is too small: SWARM_ESIZE
    using std::runtime_error::runtime_error;
};

template <class F> int guarded(F&& f) {
    try {
        f();
        return SWARM_OK;
    } catch (const SizeError& e) {
        g_err = e.what();
        return SWARM_ESIZE;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return SWARM_EINVAL;
    } catch (const swarm::NonFiniteGradient& e) {
        g_err = e.what();
        return SWARM_ENONFINITE;
    } catch (const std::bad_alloc& e) {
        g_err = e.what();
        return SWARM_ENOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return g_err.find("out of memory") != std::string::npos ? SWARM_ENOMEM : SWARM_EHIP;
    }
}

std::vector<swarm::Rect> rects(const float* p, int n) {
    std::vector<swarm::Rect> r;
    for (int i = 0; i < n; ++i) r.push_back(swarm::Rect{ p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3] });
    return r;
}
} // namespace

extern "C" {

const char* swarm_last_error(void) { return g_err.c_str(); }

int swarm_scene_build(const swarm_scene_desc* d, float max_radius, float max_v0, int threads, swarm_scene** out) {
    if (!d || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        swarm::SceneDesc sd;
        sd.xmin = d->xmin; sd.ymin = d->ymin; sd.xmax = d->xmax; sd.ymax = d->ymax;
        sd.hashCell = d->hash_cell; sd.sdfSpacing = d->sdf_spacing; sd.sdfClamp = d->sdf_clamp;
        sd.nav.resX = d->nav_res_x; sd.nav.resY = d->nav_res_y; sd.nav.originX = d->nav_origin_x; sd.nav.originY = d->nav_origin_y;
        sd.nav.cell = d->nav_cell; sd.nav.invCell = d->nav_inv_cell; sd.navBody = d->nav_body;
        sd.metresPerUnit = d->metres_per_unit;
        sd.walls = rects(d->walls, d->n_walls);
        sd.exits = rects(d->exits, d->n_exits);
        sd.stairs = rects(d->stairs, d->n_stairs);
        for (int i = 0; i < d->n_exits; ++i) sd.exitWidthM.push_back(d->exit_width_m ? d->exit_width_m[i] : 0.0f);
        for (int i = 0; i < d->n_stairs; ++i) sd.stairFactor.push_back(d->stair_factor[i]);
        if (sd.nav.resX <= 0 || sd.nav.resY <= 0 || !(sd.nav.cell > 0.0f) || !(sd.sdfSpacing > 0.0f))
            throw std::invalid_argument("swarm_scene_build: empty grid");
        auto* s = new swarm_scene{ swarm::buildScene(sd, max_radius, max_v0, threads) };
        *out = s;
    });
}

namespace {
swarm::SceneDesc descOf(const swarm_scene_desc* d) {
    swarm::SceneDesc sd;
    sd.xmin = d->xmin; sd.ymin = d->ymin; sd.xmax = d->xmax; sd.ymax = d->ymax;
    sd.hashCell = d->hash_cell; sd.sdfSpacing = d->sdf_spacing; sd.sdfClamp = d->sdf_clamp;
    sd.nav.resX = d->nav_res_x; sd.nav.resY = d->nav_res_y; sd.nav.originX = d->nav_origin_x; sd.nav.originY = d->nav_origin_y;
    sd.nav.cell = d->nav_cell; sd.nav.invCell = d->nav_inv_cell; sd.navBody = d->nav_body;
    sd.metresPerUnit = d->metres_per_unit;
    sd.walls = rects(d->walls, d->n_walls);
    sd.exits = rects(d->exits, d->n_exits);
    sd.stairs = rects(d->stairs, d->n_stairs);
    for (int i = 0; i < d->n_exits; ++i) sd.exitWidthM.push_back(d->exit_width_m ? d->exit_width_m[i] : 0.0f);
    for (int i = 0; i < d->n_stairs; ++i) sd.stairFactor.push_back(d->stair_factor[i]);
    if (sd.nav.resX <= 0 || sd.nav.resY <= 0 || !(sd.nav.cell > 0.0f) || !(sd.sdfSpacing > 0.0f))
        throw std::invalid_argument("swarm_scene_build: empty grid");
    return sd;
}
}  // namespace

int swarm_scene_build_cached(const swarm_scene_desc* d, float max_radius, float max_v0, int threads, const char* cache_dir,
                             swarm_scene** out) {
    if (!d || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        *out = new swarm_scene{ swarm::buildSceneCached(descOf(d), max_radius, max_v0, threads, cache_dir ? cache_dir : "", nullptr) };
    });
}

int swarm_scene_build_layout_cached(int layout, float max_radius, float max_v0, int threads, const char* cache_dir, swarm_scene** out) {
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        *out = new swarm_scene{ swarm::buildSceneCached(swarm::layoutDesc(layout), max_radius, max_v0, threads, cache_dir ? cache_dir : "",
                                                        nullptr) };
    });
}

int swarm_scene_build_layout(int layout, float max_radius, float max_v0, int threads, swarm_scene** out) {
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { *out = new swarm_scene{ swarm::buildScene(swarm::layoutDesc(layout), max_radius, max_v0, threads) }; });
}

int swarm_scene_info_get(const swarm_scene* sc, swarm_scene_info* i) {
    if (!sc || !i) { g_err = "null argument"; return SWARM_EINVAL; }
    const swarm::Scene& s = sc->s;
    i->xmin = s.xmin; i->ymin = s.ymin; i->xmax = s.xmax; i->ymax = s.ymax;
    i->hash_span_x = s.hashSpanX; i->hash_span_y = s.hashSpanY; i->hash_nx = s.hashNX; i->hash_ny = s.hashNY;
    i->perc_ring = s.percRing; i->proj_ring = s.projRing;
    i->sdf_nx = s.sdfNX; i->sdf_ny = s.sdfNY; i->sdf_ox = s.sdfOX; i->sdf_oy = s.sdfOY; i->sdf_h = s.sdfH; i->sdf_inv_h = s.sdfInvH;
    i->sdf_min_win = s.sdfMinWin;
    i->nav_res_x = s.nav.resX; i->nav_res_y = s.nav.resY; i->nav_origin_x = s.nav.originX; i->nav_origin_y = s.nav.originY;
    i->nav_cell = s.nav.cell; i->nav_inv_cell = s.nav.invCell;
    i->K = s.K; i->E = s.E; i->seconds_per_unit = s.secondsPerUnit; i->dist_quantum = s.distQuantum;
    i->build_seconds = s.buildSeconds;
    i->n_walls = int(s.walls.size());
    i->wall_bin_entries = int(s.wallBinList.size());
    i->from_cache = s.fromCache;
    return SWARM_OK;
}

int swarm_scene_read(const swarm_scene* sc, float* sdf, float* sdf_min, uint32_t* nav_info, uint32_t* nav_slot, uint16_t* nav_dist,
                     float* inv_cap_s, float* stair, float* exit_rects) {
    if (!sc) { g_err = "null argument"; return SWARM_EINVAL; }
    const swarm::Scene& s = sc->s;
    if (sdf) std::memcpy(sdf, s.sdf.data(), s.sdf.size() * 4);
    if (sdf_min) std::memcpy(sdf_min, s.sdfMin.data(), s.sdfMin.size() * 4);
    if (nav_info) std::memcpy(nav_info, s.navInfo.data(), s.navInfo.size() * 4);
    if (nav_slot) std::memcpy(nav_slot, s.navSlot.data(), s.navSlot.size() * 4);
    if (nav_dist) std::memcpy(nav_dist, s.navDist.data(), s.navDist.size() * 2);
    if (inv_cap_s) std::memcpy(inv_cap_s, s.invCapS.data(), s.invCapS.size() * 4);
    if (stair) std::memcpy(stair, s.stair.data(), s.stair.size() * 4);
    if (exit_rects)
        for (size_t e = 0; e < s.exits.size(); ++e) {
            exit_rects[4 * e] = s.exits[e].lox; exit_rects[4 * e + 1] = s.exits[e].loy;
            exit_rects[4 * e + 2] = s.exits[e].hix; exit_rects[4 * e + 3] = s.exits[e].hiy;
        }
    return SWARM_OK;
}

int swarm_scene_read_walls(const swarm_scene* sc, float* rects, uint32_t* bin_start, uint32_t* bin_list) {
    if (!sc) { g_err = "null argument"; return SWARM_EINVAL; }
    const swarm::Scene& s = sc->s;
    if (rects)
        for (size_t k = 0; k < s.walls.size(); ++k) {
            rects[4 * k] = s.walls[k].lox; rects[4 * k + 1] = s.walls[k].loy;
            rects[4 * k + 2] = s.walls[k].hix; rects[4 * k + 3] = s.walls[k].hiy;
        }
    if (bin_start) std::memcpy(bin_start, s.wallBinStart.data(), s.wallBinStart.size() * 4);
    if (bin_list) std::memcpy(bin_list, s.wallBinList.data(), s.wallBinList.size() * 4);
    return SWARM_OK;
}

int swarm_scene_read_venue(const swarm_scene* sc, int* has_venue, swarm_venue_layout* layout) {
    if (!sc || !has_venue) { g_err = "null argument"; return SWARM_EINVAL; }
    const swarm::Scene& s = sc->s;
    *has_venue = s.venue3d ? 2 : (s.hasVenue ? 1 : 0);
    if (layout) {
        std::memset(layout, 0, sizeof *layout);
        if (s.hasVenue) std::memcpy(layout, &s.venue, sizeof s.venue);
        if (s.venue3d) {
            layout->domain_min = s.xmin;
            layout->domain_span = s.hashSpanX;
            layout->exit_depth = s.exitDepth;
            layout->exit_max_z = s.exitMaxZ;
            layout->grid_res = s.hashNX;
            layout->exit_count = s.E;
        }
    }
    return SWARM_OK;
}

int swarm_scene_metric_grid(const swarm_scene* sc, int* res_x, int* res_y, float* cells_per_unit, float* off_x, float* off_y,
                            float* walkable_m2) {
    if (!sc) { g_err = "null argument"; return SWARM_EINVAL; }
    const swarm::Scene& s = sc->s;
    if (res_x) *res_x = s.metricResX;
    if (res_y) *res_y = s.metricResY;
    if (cells_per_unit) *cells_per_unit = s.metricCellsPerUnit;
    if (off_x) *off_x = s.metricOffX;
    if (off_y) *off_y = s.metricOffY;
    if (walkable_m2) std::memcpy(walkable_m2, s.metricWalkableM2.data(), s.metricWalkableM2.size() * 4);
    return SWARM_OK;
}

void swarm_scene_destroy(swarm_scene* s) { delete s; }

int swarm_place_agents(int layout, size_t active, size_t total, uint64_t seed, double mean, double sd, float* x, float* y,
                       float* r, float* v0, float* w) {
    if (!x || !y || !r || !v0 || !w) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::Agents a = swarm::placeAgents(layout, active, total, seed, mean, sd);
        std::memcpy(x, a.x.data(), total * 4); std::memcpy(y, a.y.data(), total * 4);
        std::memcpy(r, a.r.data(), total * 4); std::memcpy(v0, a.v0.data(), total * 4);
        std::memcpy(w, a.w.data(), total * 4);
    });
}

int swarm_spawn_seats(int seat_count, int n, double load, uint32_t seed, int32_t* chosen) {
    if (!chosen || seat_count < 0 || n < 0) { g_err = "swarm_spawn_seats: bad argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const std::vector<int> c = swarm::spawnSeats(seat_count, n, load, seed);
        for (int i = 0; i < n; ++i) chosen[i] = c[size_t(i)];
    });
}

int swarm_spawn_uniform(int layout, size_t active, size_t total, uint32_t seed, double speed_mean, double speed_sd, float* x, float* y,
                        float* r, float* v0, float* w, int* sweeps, float* max_overlap) {
    if (!x || !y || !r || !v0 || !w) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::Agents a = swarm::spawnUniform(layout, active, total, seed, speed_mean, speed_sd, sweeps, max_overlap);
        std::memcpy(x, a.x.data(), total * 4); std::memcpy(y, a.y.data(), total * 4); std::memcpy(r, a.r.data(), total * 4);
        std::memcpy(v0, a.v0.data(), total * 4); std::memcpy(w, a.w.data(), total * 4);
    });
}

int swarm_layout_slots(int layout, float* x, float* y, size_t cap, size_t* count) {
    return guarded([&] {
        std::vector<float> sx, sy;
        swarm::layoutSlots(layout, sx, sy);
        if (count) *count = sx.size();
        for (size_t i = 0; i < sx.size() && i < cap; ++i) {
            if (x) x[i] = float(double(sx[i]) / swarm::M_PER_UNIT);
            if (y) y[i] = float(double(sy[i]) / swarm::M_PER_UNIT);
        }
    });
}

int swarm_model_params(float tau_s, float* tau, float* relax) {
    if (!tau || !relax || !(tau_s > 0.0f)) { g_err = "invalid relaxation time"; return SWARM_EINVAL; }
    swarm::modelParamsFor(tau_s, *tau, *relax);
    return SWARM_OK;
}

int swarm_venue_layout_get(const char* name, swarm_venue_layout* out) {
    if (!name || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::Venue& v = swarm::venueByName(name);
        static_assert(offsetof(swarm_venue_layout, corner_door_lo) == offsetof(swarm::seated::Layout, cornerDoorLo) &&
                          sizeof(swarm::seated::Layout) == offsetof(swarm_venue_layout, corner_door_lo) + sizeof(float),
                      "swarm_venue_layout begins with the layout");
        std::memset(out, 0, sizeof *out);
        std::memcpy(out, &v.layout, sizeof v.layout);
        out->domain_min = v.domainMin; out->domain_span = v.domainSpan; out->exit_depth = v.exitDepth; out->exit_max_z = v.exitMaxZ;
        out->grid_res = v.gridRes; out->metric_grid_res = v.metricGridRes;
        out->exit_count = swarm::seated::exitCount(v.layout);
        out->gangway_count = swarm::seated::gangwayCount(v.layout);
        out->seat_total = v.seatTotal;
        out->door_width_m = v.doorWidthM; out->corner_door_width_m = v.cornerDoorWidthM;
        out->total_exit_width_m = v.totalExitWidthM; out->required_exit_width_m = v.requiredExitWidthM;
    });
}

int swarm_venue_seats(const char* name, float* x, float* y, float* z, int32_t* where, size_t cap, size_t* count) {
    if (!name) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const std::vector<swarm::VenueSeat> s = swarm::venueSeats(swarm::venueByName(name).layout);
        if (count) *count = s.size();
        if (!x && !y && !z && !where) return;
        if (s.size() > cap) throw SizeError("swarm_venue_seats: the venue has " + std::to_string(s.size()) + " seats");
        for (size_t i = 0; i < s.size(); ++i) {
            if (x) x[i] = s[i].x;
            if (y) y[i] = s[i].y;
            if (z) z[i] = s[i].z;
            if (where) {
                int32_t* w = where + 5 * i;
                w[0] = s[i].quadrant; w[1] = s[i].wing; w[2] = s[i].row; w[3] = s[i].segment; w[4] = s[i].index;
            }
        }
    });
}

int swarm_venue_probe(const char* name, size_t n, const float* p, const float* e, const float* reach3, float* sdf_grad, float* floor_h,
                      int32_t* cell, float* wall, float* slope, int32_t* exit_id, int32_t* gangway) {
    if (!name || (n > 0 && (!p || !e || !reach3))) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::Venue& v = swarm::venueByName(name);
        if (n == 0) return;
        std::vector<float> out(n * swarm::VENUE_PROBE_WORDS);
        swarm::venueProbe(v.layout, n, p, e, reach3, out.data());
        auto copy = [&](void* dst, size_t word0, size_t words) { if (dst) std::memcpy(dst, out.data() + word0 * n, words * n * 4); };
        copy(sdf_grad, 0, 4);
        copy(floor_h, 4, 1);
        copy(cell, 5, 3);
        copy(wall, 8, 9);
        copy(slope, 17, 1);
        copy(exit_id, 18, 1);
        copy(gangway, 19, 1);
    });
}

int swarm_scene_build_venue(const char* name, float max_radius, float max_v0, int threads, const char* cache_dir, swarm_scene** out) {
    if (!name || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::VenuePlan plan = swarm::venuePlan(swarm::venueByName(name));
        *out = new swarm_scene{ swarm::buildSceneCached(plan.desc, max_radius, max_v0, threads, cache_dir ? cache_dir : "", nullptr) };
    });
}

int swarm_scene_build_seated(const char* name, int threads, const char* cache_dir, swarm_scene** out) {
    if (!name || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        *out = new swarm_scene{ swarm::buildSeatedSceneCached(swarm::venueByName(name), threads, cache_dir ? cache_dir : "", nullptr) };
    });
}

int swarm_venue_exit_map(const char* name, int32_t* ref_exit, int cap, int* count) {
    if (!name) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        const swarm::VenuePlan plan = swarm::venuePlan(swarm::venueByName(name));
        if (count) *count = int(plan.refExit.size());
        if (!ref_exit) return;
        if (plan.refExit.size() > size_t(cap)) throw SizeError("swarm_venue_exit_map: the plan has " + std::to_string(plan.refExit.size()) + " exits");
        for (size_t i = 0; i < plan.refExit.size(); ++i) ref_exit[i] = plan.refExit[i];
    });
}

int swarm_create(const swarm_scene* scene, int n, const swarm_config* cfg, swarm_ctx** out) {
    if (!scene || !out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] {
        swarm::SimOptions o;
        if (cfg) {
            o.tauS = cfg->tau_s; o.obstacleGain = cfg->obstacle_gain; o.projIters = cfg->proj_iters; o.useGraph = cfg->use_graph != 0;
            if (!(o.tauS > 0.0f)) throw std::invalid_argument("swarm_create: relaxation time must be positive");
        }
        *out = new swarm_ctx{ new swarm::HipCrowdSim(scene->s, n, o) };
    });
}

#define CTX_CHECK(c) if (!(c) || !(c)->sim) { g_err = "null context"; return SWARM_EINVAL; }

int swarm_upload(swarm_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w, const float* r,
                 const float* v0) {
    CTX_CHECK(c);
    if (!x || !y || !vx || !vy || !w || !r || !v0) { g_err = "null state array"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->upload(x, y, vx, vy, w, r, v0); });
}
int swarm_restore_state(swarm_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w) {
    CTX_CHECK(c);
    if (!x || !y || !vx || !vy || !w) { g_err = "null state array"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->restoreAgentState(x, y, vx, vy, w); });
}
int swarm_set_drive(swarm_ctx* c, const float* dx, const float* dy) { CTX_CHECK(c); return guarded([&] { c->sim->setDrive(dx, dy); }); }
int swarm_step(swarm_ctx* c, int nsteps) { CTX_CHECK(c); return guarded([&] { c->sim->step(nsteps); }); }
int swarm_step_exit_choice(swarm_ctx* c) { CTX_CHECK(c); return guarded([&] { c->sim->stepExitChoice(); }); }
int swarm_set_choice_state(swarm_ctx* c, uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setChoiceState(step, assign, queue128, ahead); });
}
int swarm_sync(swarm_ctx* c) { CTX_CHECK(c); return guarded([&] { c->sim->sync(); }); }
int swarm_set_proj_iters(swarm_ctx* c, int it) { CTX_CHECK(c); return guarded([&] { c->sim->setProjectionIterations(it); }); }
int swarm_upload3(swarm_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                  const float* w) {
    CTX_CHECK(c);
    if (!x || !y || !z || !vx || !vy || !vz || !w) { g_err = "null state array"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->upload3(x, y, z, vx, vy, vz, w); });
}
int swarm_read_state3(swarm_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w) {
    CTX_CHECK(c);
    return guarded([&] {
        if (!c->sim->seated3d()) throw std::invalid_argument("swarm_read_state3: the scene is not a seated venue's own 3D step");
        c->sim->readState3(x, y, z, vx, vy, vz, speed, w);
    });
}
int swarm_read_state(swarm_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readState(x, y, vx, vy, speed, w); });
}
int swarm_read_prev_state(swarm_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readPrevState(x, y, vx, vy, speed, w); });
}
int swarm_sort_grid(swarm_ctx* c, const uint32_t* keys, uint32_t* sorted_keys, uint32_t* ids) {
    CTX_CHECK(c);
    if (!keys) { g_err = "null keys"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->sortGrid(keys, sorted_keys, ids); });
}
int swarm_nav_probe(swarm_ctx* c, int m, const float* x, const float* y, const float* r, const uint32_t* exit_id, int32_t* cell,
                    float* dir, float* goal) {
    CTX_CHECK(c);
    if (m < 0) { g_err = "negative point count"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->navProbe(m, x, y, r, exit_id, cell, dir, goal); });
}
int swarm_read_grid(swarm_ctx* c, uint32_t* key, uint32_t* id) { CTX_CHECK(c); return guarded([&] { c->sim->readGrid(key, id); }); }
int swarm_read_cells(swarm_ctx* c, uint32_t* start, uint32_t* count) { CTX_CHECK(c); return guarded([&] { c->sim->readCells(start, count); }); }
int swarm_read_assign(swarm_ctx* c, uint32_t* a) { CTX_CHECK(c); return guarded([&] { c->sim->readAssign(a); }); }
int swarm_read_choice_counters(swarm_ctx* c, uint32_t* o) { CTX_CHECK(c); return guarded([&] { c->sim->readChoiceCounters(o); }); }
int swarm_read_evac(swarm_ctx* c, int32_t* step, int32_t* exit_id, uint32_t* per_exit) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readEvac(step, exit_id, per_exit); });
}
int swarm_read_control(swarm_ctx* c, uint32_t* step, uint32_t* n_active, uint64_t* active_sum) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readControl(step, n_active, active_sum); });
}
int swarm_set_timing(swarm_ctx* c, int on) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setTimingEnabled(on != 0, on == 2); });
}
int swarm_read_timings(swarm_ctx* c, double* ms8) {
    CTX_CHECK(c);
    if (!ms8) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readTimings(ms8); });
}
int swarm_time_steps(swarm_ctx* c, int nsteps, double* ms) {
    CTX_CHECK(c);
    if (!ms) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { *ms = c->sim->timeSteps(nsteps); });
}
int swarm_copy_probe(size_t bytes, int reps, double* ms) {
    if (!ms || bytes < 16 || reps <= 0) { g_err = "invalid copy probe"; return SWARM_EINVAL; }
    return guarded([&] {
        void *a = nullptr, *b = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0, e1;
        SWARM_HIP_CHECK(hipMalloc(&a, bytes));
        SWARM_HIP_CHECK(hipMalloc(&b, bytes));
        SWARM_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        SWARM_HIP_CHECK(hipMemsetAsync(a, 0, bytes, st));
        SWARM_HIP_CHECK(hipEventCreate(&e0));
        SWARM_HIP_CHECK(hipEventCreate(&e1));
        swarm::launchCopyProbe(st, a, b, bytes);   // warm-up
        SWARM_HIP_CHECK(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) swarm::launchCopyProbe(st, a, b, bytes);
        SWARM_HIP_CHECK(hipEventRecord(e1, st));
        SWARM_HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        SWARM_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        *ms = double(t) / reps;
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
        (void)hipFree(a); (void)hipFree(b);
    });
}

int swarm_batch_create(const swarm_scene* const* layouts, int n_layouts, const int* scen_layout, const int* scen_agents, int scenarios,
                       const swarm_config* cfg, swarm_ctx** out) {
    if (!layouts || n_layouts <= 0 || !scen_layout || !scen_agents || scenarios <= 0 || !out) {
        g_err = "swarm_batch_create: null or empty argument";
        return SWARM_EINVAL;
    }
    return guarded([&] {
        swarm::SimOptions o;
        if (cfg) {
            o.tauS = cfg->tau_s; o.obstacleGain = cfg->obstacle_gain; o.projIters = cfg->proj_iters; o.useGraph = cfg->use_graph != 0;
            if (!(o.tauS > 0.0f)) throw std::invalid_argument("swarm_batch_create: relaxation time must be positive");
        }
        std::vector<const swarm::Scene*> ls;
        for (int l = 0; l < n_layouts; ++l) {
            if (!layouts[l]) throw std::invalid_argument("swarm_batch_create: null scene");
            ls.push_back(&layouts[l]->s);
        }
        const std::vector<int> sl(scen_layout, scen_layout + scenarios), sa(scen_agents, scen_agents + scenarios);
        *out = new swarm_ctx{ new swarm::HipCrowdSim(ls, sl, sa, o) };
    });
}
int swarm_num_scenarios(swarm_ctx* c) { CTX_CHECK(c); return c->sim->numScenarios(); }
int swarm_read_choice_counters_all(swarm_ctx* c, uint32_t* out) {
    CTX_CHECK(c);
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readChoiceCountersAll(out); });
}
int swarm_read_evac_counters(swarm_ctx* c, uint32_t* per_exit64) {
    CTX_CHECK(c);
    if (!per_exit64) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readEvacCounters(per_exit64); });
}
int swarm_set_physics_shape(swarm_ctx* c, int block, int grid) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setPhysicsShape(block, grid); });
}
int swarm_set_projection_prune(swarm_ctx* c, int every) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setPruneEvery(every); });
}
int swarm_set_candidate_budget(swarm_ctx* c, float moves) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setCandidateBudget(moves); });
}
int swarm_set_projection_lanes(swarm_ctx* c, int lanes) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setProjectionLanes(lanes); });
}
int swarm_read_candidates(swarm_ctx* c, uint32_t* count, uint32_t* index, float* dist) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readCandidates(count, index, dist); });
}
int swarm_set_sort_mode(swarm_ctx* c, int mode) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setSortMode(mode); });
}
int swarm_set_predict_lanes(swarm_ctx* c, int lanes) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setPredictLanes(lanes); });
}
int swarm_set_projection_collect_lanes(swarm_ctx* c, int lanes) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setCollectLanes(lanes); });
}
int swarm_get_lanes(swarm_ctx* c, int* list_lanes, int* collect_lanes, int* predict_lanes) {
    CTX_CHECK(c);
    if (list_lanes) *list_lanes = c->sim->projectionLanes();
    if (collect_lanes) *collect_lanes = c->sim->collectLanes();
    if (predict_lanes) *predict_lanes = c->sim->predictLanes();
    return SWARM_OK;
}
int swarm_get_projection_lanes(swarm_ctx* c, int* lanes) {
    CTX_CHECK(c);
    if (!lanes) { g_err = "null argument"; return SWARM_EINVAL; }
    *lanes = c->sim->projectionLanes();
    return SWARM_OK;
}
int swarm_metrics_enable(swarm_ctx* c, int on) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setMetricsEnabled(on != 0); });
}
int swarm_metrics_reset(swarm_ctx* c) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->resetMetrics(); });
}
int swarm_metrics_configure(swarm_ctx* c, int res_x, int res_y, float cells_per_unit, float off_x, float off_y, const uint32_t* thresholds4) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->configureMetrics(res_x, res_y, cells_per_unit, off_x, off_y, thresholds4); });
}
int swarm_metrics_grid(swarm_ctx* c, int* res_x, int* res_y, float* cells_per_unit, float* off_x, float* off_y) {
    CTX_CHECK(c);
    c->sim->metricGrid(res_x, res_y, cells_per_unit, off_x, off_y);
    return SWARM_OK;
}
int swarm_metrics_record(swarm_ctx* c, const float* px, const float* py, const float* pw, const float* cx, const float* cy, const float* cw,
                         const float* cvx, const float* cvy) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->recordMetrics(px, py, pw, cx, cy, cw, cvx, cvy); });
}
int swarm_read_metric_counters(swarm_ctx* c, uint32_t* out) {
    CTX_CHECK(c);
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readMetricCounters(out); });
}
int swarm_read_metric_cells(swarm_ctx* c, uint32_t* out) {
    CTX_CHECK(c);
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readMetricCells(out); });
}
int swarm_read_metric_evac_steps(swarm_ctx* c, uint32_t* out) {
    CTX_CHECK(c);
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->readMetricEvacSteps(out); });
}
int swarm_read_metric_cell_info(swarm_ctx* c, float* walkable_m2, uint32_t* thresholds4) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readMetricCellInfo(walkable_m2, thresholds4); });
}
int swarm_get_physics_shape(swarm_ctx* c, int* block, int* grid) {
    CTX_CHECK(c);
    if (block) *block = c->sim->physicsBlock();
    if (grid) *grid = c->sim->physicsGrid();
    return SWARM_OK;
}
int swarm_autotune(swarm_ctx* c, int steps, double* best_ms) {
    CTX_CHECK(c);
    return guarded([&] {
        const double ms = c->sim->autotune(steps);
        if (best_ms) *best_ms = ms;
    });
}

int swarm_adjoint_forward(swarm_ctx* c, int K, int checkpoint_every, float radius_scale, float overlap_weight, double* loss) {
    CTX_CHECK(c);
    if (!loss) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { *loss = c->sim->adjointForward(K, checkpoint_every, radius_scale, overlap_weight); });
}
int swarm_adjoint_set_tape_budget(swarm_ctx* c, size_t bytes) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setTapeBudget(bytes); });
}
int swarm_adjoint_segments(swarm_ctx* c, int* segments) {
    CTX_CHECK(c);
    if (!segments) { g_err = "null argument"; return SWARM_EINVAL; }
    *segments = c->sim->adjointSegments();
    return SWARM_OK;
}
int swarm_adjoint_memory(swarm_ctx* c, size_t* bytes) {
    CTX_CHECK(c);
    if (!bytes) { g_err = "null argument"; return SWARM_EINVAL; }
    *bytes = c->sim->adjointBytes();
    return SWARM_OK;
}
int swarm_adjoint_backward(swarm_ctx* c, float* grad_v0, double* grad_globals) {
    CTX_CHECK(c);
    if (!grad_v0 || !grad_globals) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->adjointBackward(grad_v0, grad_globals); });
}
int swarm_adjoint_tape(swarm_ctx* c, int step, const swarm_tape_view* out, swarm_exit_rec* exit) {
    CTX_CHECK(c);
    if (!out) { g_err = "null argument"; return SWARM_EINVAL; }
    return guarded([&] { c->sim->adjointTape(step, *out, exit); });
}
int swarm_adjoint_set_terms(swarm_ctx* c, const int32_t* step, const float* gin, const float* gout, int exit_term, int window, int local) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setAdjointTerms(step, gin, gout, exit_term != 0, window, local != 0); });
}
int swarm_read_evac_points(swarm_ctx* c, int32_t* step, int32_t* exit_id, float* inside, float* before) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->readEvacPoints(step, exit_id, inside, before); });
}
int swarm_set_model(swarm_ctx* c, float tau_s, float obstacle_gain) {
    CTX_CHECK(c);
    return guarded([&] { c->sim->setModel(tau_s, obstacle_gain); });
}

int swarm_get_stream(swarm_ctx* c, void** stream) {
    CTX_CHECK(c);
    if (!stream) { g_err = "null argument"; return SWARM_EINVAL; }
    *stream = c->sim->stream();
    return SWARM_OK;
}
int swarm_num_cells(swarm_ctx* c) { CTX_CHECK(c); return c->sim->numCells(); }
void swarm_destroy(swarm_ctx* c) {
    if (!c) return;
    delete c->sim;
    delete c;
}

} // extern "C"
