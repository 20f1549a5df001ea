// swarm_crowdsim_adapter.hpp - the reference-side binding of libswarm_hip: class CrowdSim's surface
// (include/swarm/crowd_sim.hpp:53-194 of the reference) over the C ABI of swarm_hip.h, for a maintainer to add to the
// reference's tree. It takes and returns the reference's own records (GpuBoid, GpuGridInfo, GpuCellData from
// include/swarm/gpu_types.hpp:17-36) and throws the reference's exception types (std::invalid_argument where
// CrowdSim does, crowd_sim.cpp:99-100,286-287,481; std::runtime_error otherwise, crowd_sim.cpp:105-111).
//
// Build it in the reference's tree (its include/ and external/glm on the include path) with this repository's
// include/ and link libswarm_hip.so. On the flat floors of this path GpuBoid::position.z and velocity.z are 0;
// velocity.w is the step's speed (physics.comp:1665). Per-agent radius and desired speed are the extra upload inputs.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include <swarm/gpu_types.hpp>

#include "swarm_hip.h"

namespace swarm {

class HipCrowdSimAdapter {
public:
    // CrowdSim(shaderDir, numBoids, venue, options) (crowd_sim.hpp:62-73): the venue is a rectangle world
    HipCrowdSimAdapter(const swarm_scene_desc& venue, int numBoids, const swarm_config& cfg, float maxRadius, float maxV0) : n_(numBoids) {
        if (numBoids <= 0) throw std::invalid_argument("HipCrowdSimAdapter: numBoids must be positive");
        check(swarm_scene_build(&venue, maxRadius, maxV0, 0, &scene_));
        const int rc = swarm_create(scene_, numBoids, &cfg, &ctx_);
        if (rc != SWARM_OK) {
            swarm_scene_destroy(scene_);
            check(rc);
        }
    }
    ~HipCrowdSimAdapter() {
        swarm_destroy(ctx_);
        swarm_scene_destroy(scene_);
    }
    HipCrowdSimAdapter(const HipCrowdSimAdapter&) = delete;
    HipCrowdSimAdapter& operator=(const HipCrowdSimAdapter&) = delete;

    // CrowdSim::upload (crowd_sim.hpp:90): AoS GpuBoid -> SoA, with each agent's radius and desired speed
    void upload(const std::vector<GpuBoid>& boids, const std::vector<float>& radius, const std::vector<float>& v0) {
        if (int(boids.size()) != n_ || int(radius.size()) != n_ || int(v0.size()) != n_)
            throw std::invalid_argument("upload: size mismatch");
        const size_t n = static_cast<size_t>(n_);
        std::vector<float> x(n), y(n), vx(n), vy(n), w(n);
        for (size_t i = 0; i < size_t(n_); ++i) {
            x[i] = boids[i].position.x;
            y[i] = boids[i].position.y;
            w[i] = boids[i].position.w;
            vx[i] = boids[i].velocity.x;
            vy[i] = boids[i].velocity.y;
        }
        check(swarm_upload(ctx_, x.data(), y.data(), vx.data(), vy.data(), w.data(), radius.data(), v0.data()));
    }
    // CrowdSim::step (crowd_sim.hpp:91): the SDF is part of the venue here
    void step() { check(swarm_step(ctx_, 1)); }
    void setProjectionIterations(int iterations) {
        if (iterations < 0) throw std::invalid_argument("setProjectionIterations: iterations must be >= 0");
        check(swarm_set_proj_iters(ctx_, iterations));
    }
    void stepExitChoice() { check(swarm_step_exit_choice(ctx_)); }

    // currentStateBuffer / previousStateBuffer readbacks (crowd_sim.hpp:132-133)
    std::vector<GpuBoid> currentState() const { return state(false); }
    std::vector<GpuBoid> previousState() const { return state(true); }
    // gridBuffer (crowd_sim.hpp:134): sorted (cell, agent) pairs
    std::vector<GpuGridInfo> grid() const {
        std::vector<uint32_t> key(static_cast<size_t>(n_)), id(static_cast<size_t>(n_));
        check(swarm_read_grid(ctx_, key.data(), id.data()));
        std::vector<GpuGridInfo> g(static_cast<size_t>(n_));
        for (size_t i = 0; i < size_t(n_); ++i) g[i] = GpuGridInfo{ key[i], id[i], 0u, 0u };
        return g;
    }
    // cellBuffer (crowd_sim.hpp:135): start and count per cell
    std::vector<GpuCellData> cells() const {
        const int nc = swarm_num_cells(ctx_);
        if (nc < 0) check(nc);
        std::vector<uint32_t> start(static_cast<size_t>(nc)), count(static_cast<size_t>(nc));
        check(swarm_read_cells(ctx_, start.data(), count.data()));
        std::vector<GpuCellData> c(static_cast<size_t>(nc));
        for (size_t i = 0; i < size_t(nc); ++i) c[i] = GpuCellData{ start[i], count[i], 0u, 0u };
        return c;
    }
    // assignBuffer (crowd_sim.hpp:109)
    std::vector<uint32_t> assignedExits() const {
        std::vector<uint32_t> a(static_cast<size_t>(n_));
        check(swarm_read_assign(ctx_, a.data()));
        return a;
    }

private:
    std::vector<GpuBoid> state(bool previous) const {
        const size_t n = static_cast<size_t>(n_);
        std::vector<float> x(n), y(n), vx(n), vy(n), s(n), w(n);
        if (previous) check(swarm_read_prev_state(ctx_, x.data(), y.data(), vx.data(), vy.data(), s.data(), w.data()));
        else check(swarm_read_state(ctx_, x.data(), y.data(), vx.data(), vy.data(), s.data(), w.data()));
        std::vector<GpuBoid> out(static_cast<size_t>(n_));
        for (size_t i = 0; i < size_t(n_); ++i) {
            out[i].position = glm::vec4(x[i], y[i], 0.0f, w[i]);
            out[i].velocity = glm::vec4(vx[i], vy[i], 0.0f, s[i]);
        }
        return out;
    }
    static void check(int rc) {
        if (rc == SWARM_OK) return;
        if (rc == SWARM_EINVAL || rc == SWARM_ESIZE) throw std::invalid_argument(swarm_last_error());
        throw std::runtime_error(swarm_last_error());
    }
    int n_;
    swarm_scene* scene_ = nullptr;
    swarm_ctx* ctx_ = nullptr;
};

}  // namespace swarm
