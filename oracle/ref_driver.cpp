// Golden-vector driver over the reference's own CPU code (src/cpu_crowd_sim.cpp, src/nav_field.cpp,
// src/exit_choice.cpp, built from /root/reference by oracle/ref_build.sh into oracle/_ref/).
// It runs the reference's CpuCrowdSim, nav::build / buildChoiceTable and nav::choiceStep on small known
// cases and writes their inputs and outputs to a binary record file; tests/golden/make_golden.py packs
// them into the committed fixtures. TEST INFRASTRUCTURE ONLY.
//
// Record format: u32 name length, name, u8 dtype ('f' float32, 'u' uint32, 'i' int32, 'b' uint8, 'd' float64),
// u32 ndim, u64 dims[ndim], raw little-endian data.
#include <swarm/cpu_crowd_sim.hpp>
#include <swarm/crowd_sim.hpp>
#include <swarm/exit_choice.hpp>
#include <swarm/metrics.hpp>
#include <swarm/nav_field.hpp>
#include <swarm/sdf.hpp>
#include <swarm/venue.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace swarm;

namespace {
FILE* g_out = nullptr;

template <class T> void put(const std::string& name, char dtype, const std::vector<T>& v, std::vector<uint64_t> dims = {}) {
    if (dims.empty()) dims = { uint64_t(v.size()) };
    const uint32_t len = uint32_t(name.size());
    std::fwrite(&len, 4, 1, g_out);
    std::fwrite(name.data(), 1, len, g_out);
    std::fwrite(&dtype, 1, 1, g_out);
    const uint32_t nd = uint32_t(dims.size());
    std::fwrite(&nd, 4, 1, g_out);
    std::fwrite(dims.data(), 8, nd, g_out);
    std::fwrite(v.data(), sizeof(T), v.size(), g_out);
}

struct Rng {   // splitmix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return float(double(next() >> 11) * (1.0 / 9007199254740992.0)); }
};

// A flat floor at z = 0 in the terrace's baked volume: the value is the height above it
std::vector<float> flatVolume() {
    std::vector<float> vol(SDF_VOXELS);
    for (int x = 0; x < SDF_RES; x++)
        for (int y = 0; y < SDF_RES; y++)
            for (int z = 0; z < SDF_RES; z++) vol[(size_t(x) * SDF_RES + y) * SDF_RES + z] = SDF_ORIGIN.z + float(z - 1) * SDF_SPACING;
    return vol;
}

// Terrace pitch scenario: n agents in a box on the fenced pitch, forced drive directions (counterflow), K steps
void terraceCase(const std::string& tag, const std::vector<float>& vol, int n, float half, float minDist, int steps, uint64_t seed,
                 int inactive, bool nanAgent, int outside, CpuCrowdSim::SortMode mode = CpuCrowdSim::SortMode::Counting) {
    Rng rng{ seed };
    std::vector<GpuBoid> b(static_cast<size_t>(n));
    std::vector<glm::vec2> drive(static_cast<size_t>(n));
    std::vector<float> px, py;
    int placed = 0;
    for (int i = 0; i < n; ++i) {
        GpuBoid& a = b[size_t(i)];
        a.velocity = glm::vec4(0.0f);
        if (i < inactive) {   // evacuated agents
            a.position = glm::vec4(PARK_COORD, PARK_COORD, PARK_COORD, 0.0f);
            drive[size_t(i)] = glm::vec2(1.0f, 0.0f);
            continue;
        }
        if (i < inactive + outside) {   // beyond the domain: parked by their first step
            a.position = glm::vec4(1.2f + 0.01f * float(i), -1.3f, BODY_RADIUS, 1.0f);
            drive[size_t(i)] = glm::vec2(0.0f, 1.0f);
            continue;
        }
        if (nanAgent && i == n - 1) {   // a NaN position: the hash resets it to the origin
            a.position = glm::vec4(std::nanf(""), 0.0f, 0.0f, 1.0f);
            drive[size_t(i)] = glm::vec2(0.0f, 1.0f);
            continue;
        }
        float x = 0, y = 0;
        for (int tries = 0; tries < 10000; ++tries) {   // a box [0.05, 0.05 + 2 half] x [-half, half], away from the origin
            x = 0.05f + 2.0f * half * rng.uniform();
            y = -half + 2.0f * half * rng.uniform();
            bool ok = true;
            for (size_t k = 0; k < px.size() && ok; ++k) ok = (px[k] - x) * (px[k] - x) + (py[k] - y) * (py[k] - y) >= minDist * minDist;
            if (ok) break;
        }
        px.push_back(x);
        py.push_back(y);
        ++placed;
        a.position = glm::vec4(x, y, BODY_RADIUS, 1.0f);
        const float ang = 6.2831853f * rng.uniform();
        const float sp = 0.5f * rng.uniform();
        a.velocity = glm::vec4(sp * std::cos(ang), sp * std::sin(ang), 0.0f, 0.0f);
        // counterflow lanes with a slight random tilt
        const float dir = (i % 2) ? 1.0f : -1.0f;
        const float tilt = 0.3f * (rng.uniform() - 0.5f);
        const glm::vec2 d = glm::normalize(glm::vec2(dir, tilt));
        drive[size_t(i)] = d;
    }
    CpuCrowdSim sim(vol, 4, mode);
    sim.setState(b);
    sim.setDriveTable(&drive);
    auto dumpState = [&](const std::vector<GpuBoid>& s, const std::string& p) {
        std::vector<float> x(static_cast<size_t>(n)), y(static_cast<size_t>(n)), z(static_cast<size_t>(n)), w(static_cast<size_t>(n)), vx(static_cast<size_t>(n)), vy(static_cast<size_t>(n)), vz(static_cast<size_t>(n)), vw(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            x[size_t(i)] = s[size_t(i)].position.x; y[size_t(i)] = s[size_t(i)].position.y; z[size_t(i)] = s[size_t(i)].position.z;
            w[size_t(i)] = s[size_t(i)].position.w; vx[size_t(i)] = s[size_t(i)].velocity.x; vy[size_t(i)] = s[size_t(i)].velocity.y;
            vz[size_t(i)] = s[size_t(i)].velocity.z; vw[size_t(i)] = s[size_t(i)].velocity.w;
        }
        put(p + "x", 'f', x); put(p + "y", 'f', y); put(p + "z", 'f', z); put(p + "w", 'f', w);
        put(p + "vx", 'f', vx); put(p + "vy", 'f', vy); put(p + "vz", 'f', vz); put(p + "speed", 'f', vw);
    };
    dumpState(b, tag + "/in/");
    std::vector<float> dx(static_cast<size_t>(n)), dy(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) { dx[size_t(i)] = drive[size_t(i)].x; dy[size_t(i)] = drive[size_t(i)].y; }
    put(tag + "/in/dx", 'f', dx);
    put(tag + "/in/dy", 'f', dy);
    for (int k = 1; k <= steps; ++k) {
        sim.step();
        const std::string p = tag + "/step" + std::to_string(k) + "/";
        dumpState(sim.state(), p);
        std::vector<uint32_t> key(static_cast<size_t>(n)), id(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) { key[size_t(i)] = sim.grid()[size_t(i)].gridID; id[size_t(i)] = sim.grid()[size_t(i)].boidID; }
        put(p + "key", 'u', key);
        put(p + "id", 'u', id);
        std::vector<uint32_t> st(sim.cells().size()); std::vector<uint32_t> cn(sim.cells().size());
        for (size_t c = 0; c < st.size(); ++c) { st[c] = sim.cells()[c].startIndex; cn[c] = sim.cells()[c].agentCount; }
        put(p + "cell_start", 'u', st);
        put(p + "cell_count", 'u', cn);
    }
}

// nav_check::roomWalls (tools/nav_check.hpp:137-153), restated for the driver's worlds
using nav::Rect;
struct Gap { int side; float from, to; };
std::vector<Rect> roomWalls(float x0, float y0, float x1, float y1, float t, const std::vector<Gap>& gaps) {
    std::vector<Rect> out;
    auto addSide = [&](int side, float a0, float a1, auto make) {
        std::vector<std::pair<float, float>> cuts;
        for (const Gap& g : gaps)
            if (g.side == side) cuts.push_back({ g.from, g.to });
        std::sort(cuts.begin(), cuts.end());
        float a = a0;
        for (const auto& c : cuts) { if (c.first > a) out.push_back(make(a, c.first)); a = c.second; }
        if (a1 > a) out.push_back(make(a, a1));
    };
    addSide(0, x0 - t, x1 + t, [&](float a, float b) { return Rect{ { a, y0 - t }, { b, y0 } }; });
    addSide(2, x0 - t, x1 + t, [&](float a, float b) { return Rect{ { a, y1 }, { b, y1 + t } }; });
    addSide(3, y0, y1, [&](float a, float b) { return Rect{ { x0 - t, a }, { x0, b } }; });
    addSide(1, y0, y1, [&](float a, float b) { return Rect{ { x1, a }, { x1 + t, b } }; });
    return out;
}
nav::Grid testGrid(float x0, float span, double cell) {
    nav::Grid g;
    g.res = int(std::lround(span / cell));
    g.origin = x0;
    g.cell = float(cell);
    g.invCell = float(1.0 / cell);
    return g;
}

void putRects(const std::string& p, const std::vector<Rect>& rs) {
    std::vector<float> v;
    for (const Rect& r : rs) { v.push_back(r.lo.x); v.push_back(r.lo.y); v.push_back(r.hi.x); v.push_back(r.hi.y); }
    put(p, 'f', v, { uint64_t(rs.size()), 4 });
}

// The reference's own lookups at probe points of a rectangle world: nav::choiceCell, nav::choiceDirection for every
// exit (src/exit_choice.cpp:512-566) and RectGeometry::exitAt (src/nav_field.cpp:155-159). Points: a seeded uniform
// sample over the grid, points beside every wall face (the NOFIT band, its 3 x 3 and 5 x 5 fallbacks), anchor
// centres and points within and just beyond GOAL_DEADZONE of them, and every exit rectangle's corners, edge midpoints
// and their neighbours one float step inside and outside.
void probeCase(const std::string& tag, const nav::RectGeometry& geo, const nav::Grid& grid, const std::vector<Rect>& walls,
               const std::vector<Rect>& exits, const nav::ChoiceTable& t) {
    std::vector<float> px, py;
    uint64_t st = 0x9E3779B97F4A7C15ull ^ uint64_t(tag.size());
    auto uni = [&]() {   // splitmix64 -> [0, 1)
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const double lo = grid.origin, span = double(grid.res) * double(grid.cell);
    for (int i = 0; i < 3000; ++i) { px.push_back(float(lo + span * uni())); py.push_back(float(lo + span * uni())); }
    const float offs[] = { 0.01f, 0.04f, 0.09f, 0.16f, 0.26f, 0.41f };
    for (const Rect& w : walls)
        for (float o : offs)
            for (int k = 0; k < 4; ++k) {
                const float a = float(uni());
                const float fx = w.lo.x + a * (w.hi.x - w.lo.x), fy = w.lo.y + a * (w.hi.y - w.lo.y);
                if (k == 0) { px.push_back(w.lo.x - o); py.push_back(fy); }
                if (k == 1) { px.push_back(w.hi.x + o); py.push_back(fy); }
                if (k == 2) { px.push_back(fx); py.push_back(w.lo.y - o); }
                if (k == 3) { px.push_back(fx); py.push_back(w.hi.y + o); }
            }
    for (int i = 0; i < 200; ++i) {   // anchors: the first candidate's anchor of a random reachable cell
        const int c = int(uni() * double(t.index.size()));
        if (c >= int(t.index.size()) || !t.reachable(size_t(c))) continue;
        const uint32_t a = t.at(size_t(c)).slot[0] & nav::CHOICE_ANCHOR_MASK;
        const glm::vec2 ac = nav::anchorCentre(t, a);
        const float d[] = { 0.0f, 3e-5f, 9.9e-5f, 1.01e-4f, 3e-4f };
        for (float dd : d) { px.push_back(ac.x + dd); py.push_back(ac.y - 0.5f * dd); }
    }
    for (const Rect& e : exits) {
        const float xs[] = { e.lo.x, 0.5f * (e.lo.x + e.hi.x), e.hi.x }, ys[] = { e.lo.y, 0.5f * (e.lo.y + e.hi.y), e.hi.y };
        for (float x : xs)
            for (float y : ys)
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        px.push_back(dx < 0 ? std::nextafter(x, -1e9f) : dx > 0 ? std::nextafter(x, 1e9f) : x);
                        py.push_back(dy < 0 ? std::nextafter(y, -1e9f) : dy > 0 ? std::nextafter(y, 1e9f) : y);
                    }
    }
    const int E = geo.exitCount();
    std::vector<int32_t> cell, at;
    std::vector<float> dx, dy;
    for (size_t i = 0; i < px.size(); ++i) {
        const glm::vec2 q(px[i], py[i]);
        cell.push_back(nav::choiceCell(t, q));
        at.push_back(geo.exitAt(q));
        for (int e = 0; e < E; ++e) {
            const glm::vec2 d = nav::choiceDirection(t, q, uint32_t(e));
            dx.push_back(d.x);
            dy.push_back(d.y);
        }
    }
    put(tag + "/probe/x", 'f', px);
    put(tag + "/probe/y", 'f', py);
    put(tag + "/probe/cell", 'i', cell);
    put(tag + "/probe/exit_at", 'i', at);
    put(tag + "/probe/dir_x", 'f', dx, { uint64_t(px.size()), uint64_t(E) });
    put(tag + "/probe/dir_y", 'f', dy, { uint64_t(px.size()), uint64_t(E) });
}

// A rectangle world's grid (cls, out, exit region), per-exit fields and candidate table
void navCase(const std::string& tag, const std::vector<Rect>& walls, const std::vector<Rect>& exits, const std::vector<float>& widths,
             const nav::Grid& grid, float body) {
    nav::RectGeometry geo(walls, exits, body);
    if (!widths.empty()) geo.setExitWidths(widths);
    const nav::Field f = nav::buildGrid(geo, grid);
    putRects(tag + "/walls", walls);
    putRects(tag + "/exits", exits);
    put(tag + "/widths", 'f', widths);
    put(tag + "/grid", 'd', std::vector<double>{ double(grid.res), double(grid.origin), double(grid.cell), double(grid.invCell), double(body) });
    put(tag + "/cls", 'b', f.cls);
    put(tag + "/out", 'b', f.out);
    std::vector<int32_t> er(f.exitRegion.begin(), f.exitRegion.end());
    put(tag + "/exit_region", 'i', er);
    for (int e = 0; e < f.exitCount; ++e) {
        const nav::ExitField ef = nav::buildExitField(f, e);
        put(tag + "/field" + std::to_string(e) + "/dist", 'f', ef.dist);
        put(tag + "/field" + std::to_string(e) + "/anchor", 'i', ef.anchor);
    }
    const nav::ChoiceTable t = nav::buildChoiceTable(geo, f);
    put(tag + "/table/index", 'u', t.index);
    std::vector<uint32_t> slots, dists;
    for (const GpuNavChoiceCell& c : t.entries) {
        for (int k = 0; k < nav::CHOICE_K; ++k) slots.push_back(c.slot[k]);
        for (int k = 0; k < nav::CHOICE_DIST_WORDS; ++k) dists.push_back(c.dist[k]);
    }
    put(tag + "/table/slots", 'u', slots, { uint64_t(t.entries.size()), uint64_t(nav::CHOICE_K) });
    put(tag + "/table/dists", 'u', dists, { uint64_t(t.entries.size()), uint64_t(nav::CHOICE_DIST_WORDS) });
    put(tag + "/table/scalars", 'f', std::vector<float>{ t.secondsPerUnit, t.distQuantum });
    put(tag + "/table/inv_cap_s", 'f', t.invCapS);
    probeCase(tag, geo, grid, walls, exits, t);
}

GpuBoid agentAt(float x, float y, bool active = true) {
    GpuBoid b{};
    b.position = glm::vec4(x, y, 0.0f, active ? 1.0f : 0.0f);
    return b;
}

void putState(const std::string& p, const std::vector<GpuBoid>& st) {
    std::vector<float> x, y, w;
    for (const GpuBoid& b : st) { x.push_back(b.position.x); y.push_back(b.position.y); w.push_back(b.position.w); }
    put(p + "x", 'f', x); put(p + "y", 'f', y); put(p + "w", 'f', w);
}
void putChoice(const std::string& p, const nav::ChoiceState& c) {
    put(p + "assign", 'u', c.assign);
    std::vector<uint32_t> cnt(c.counters.queue, c.counters.queue + 128);
    cnt.push_back(c.counters.switches); cnt.push_back(c.counters.forced); cnt.push_back(c.counters.evaluations); cnt.push_back(c.counters.eligible);
    put(p + "counters", 'u', cnt);
}

// The exit-choice known answers of test_gpu nav.exit_choice (tools/test_gpu.cpp:3968-4100) on the two-door room
// (tools/choice_check.hpp:21-36), run through the reference's CPU pass nav::choiceStep
void choiceCases() {
    const std::vector<Rect> walls = roomWalls(0.0f, 0.0f, 20.0f, 10.0f, 0.3f, { { 3, 4.4f, 5.6f }, { 1, 4.4f, 5.6f } });
    const std::vector<Rect> exits = { { { -1.0f, 4.4f }, { -0.3f, 5.6f } }, { { 20.3f, 4.4f }, { 21.0f, 5.6f } } };
    const nav::Grid grid = testGrid(-1.0f, 22.0f, 0.1);
    navCase("room2", walls, exits, { 1.2f, 1.2f }, grid, 0.0f);
    nav::RectGeometry geo(walls, exits);
    geo.setExitWidths({ 1.2f, 1.2f });
    const nav::Field f = nav::buildGrid(geo, grid);
    const nav::ChoiceTable rt = nav::buildChoiceTable(geo, f);
    const int cs = choiceSteps(DT);
    // (a) one door queued, mirrored, no queue: steps 0 and 1
    for (int variant = 0; variant < 3; ++variant) {
        const bool queueLeft = variant != 1, withQueue = variant != 2;
        std::vector<GpuBoid> st;
        for (int k = 0; k < 50 && withQueue; ++k) {
            const float x = 0.55f + 0.1f * float(k % 25);
            st.push_back(agentAt(queueLeft ? x : 20.0f - x, 4.65f + 0.2f * float(k / 25)));
        }
        const size_t first = st.size();
        for (float y : { 4.75f, 4.95f, 5.05f, 5.25f })
            for (float x : { 9.95f, 10.05f }) st.push_back(agentAt(x, y, false));
        std::vector<GpuBoid> active = st;
        for (size_t i = first; i < st.size(); ++i) active[i].position.w = 1.0f;
        nav::ChoiceState c;
        c.reset(st.size());
        const std::string p = "choice_a" + std::to_string(variant) + "/";
        putState(p + "s0/", st);
        nav::choiceStep(rt, st, c, cs);
        putChoice(p + "s0/", c);
        putState(p + "s1/", active);
        nav::choiceStep(rt, active, c, cs);
        putChoice(p + "s1/", c);
    }
    // (b) hysteresis margins: one evaluation from a set state
    for (int variant = 0; variant < 2; ++variant) {
        const uint32_t qLeft = variant ? 120u : 100u, qRight = 100u;
        const std::vector<GpuBoid> st = { agentAt(11.05f, 4.95f), agentAt(11.05f, 5.05f) };
        nav::ChoiceState c;
        c.reset(st.size());
        c.step = uint32_t(cs);
        c.assign = { 0u, 0u };
        c.counters.queue[(c.step & 1u) * nav::CHOICE_QUEUE_STRIDE + 0] = qLeft;
        c.counters.queue[(c.step & 1u) * nav::CHOICE_QUEUE_STRIDE + 1] = qRight;
        const size_t par = size_t(c.step & 1u) * nav::CHOICE_QUEUE_STRIDE * nav::CHOICE_AHEAD_BINS;
        c.ahead[par + 0 * nav::CHOICE_AHEAD_BINS] = qLeft;
        c.ahead[par + 1 * nav::CHOICE_AHEAD_BINS] = qRight;
        const std::string p = "choice_b" + std::to_string(variant) + "/";
        putState(p, st);
        put(p + "set", 'u', std::vector<uint32_t>{ c.step, qLeft, qRight });
        nav::choiceStep(rt, st, c, cs);
        putChoice(p, c);
    }
    // (ii) no oscillation: 190 agents on the centre line, all assigned left, 40 evaluations of the closed loop
    {
        std::vector<GpuBoid> st;
        for (int k = 0; k < 190; ++k) st.push_back(agentAt(0.55f + 0.1f * float(k), 5.05f));
        nav::ChoiceState c;
        c.reset(st.size());
        c.assign.assign(st.size(), 0u);
        c.counters.queue[0] = uint32_t(st.size());
        putState("choice_c/", st);
        std::vector<uint32_t> hist;
        for (int s = 0; s < 40 * cs + 1; ++s) {
            nav::choiceStep(rt, st, c, cs);
            if (s % cs == 0) hist.insert(hist.end(), c.assign.begin(), c.assign.end());
        }
        put("choice_c/assign_at_evals", 'u', hist, { uint64_t(hist.size() / st.size()), uint64_t(st.size()) });
        putChoice("choice_c/final/", c);
    }
}
// Safety metrics (src/metrics.cpp: cicDeposit, speedMass, MetricVenue::cellInfo, MetricsReference::record) on
// synthetic states over the reference's default 64 x 64 grid of 1 m cells: a dense cluster (past the crush
// threshold), agents spread over and beyond the grid (clamped to its edge), agents on exact cell and quantum
// boundaries, inactive agents and evacuations in every step
void metricsCases() {
    using namespace swarm::metrics;
    Rng rng{ 777 };
    {   // cicDeposit / speedMass known answers
        std::vector<float> xy, sp;
        std::vector<uint32_t> cells, weights, wts, sms;
        for (int k = 0; k < 2000; ++k) {
            float x = (rng.uniform() * 2.4f - 1.2f), y = (rng.uniform() * 2.4f - 1.2f);
            if (k % 10 == 0) { x = float(int(x * 31.25f)) / 31.25f; }             // integer metres (cell corners)
            if (k % 10 == 1) { y = (float(int(y * 31.25f)) + 0.5f) / 31.25f; }    // cell centres
            if (k % 10 == 2) { x = (float(int(x * 8000.0f)) + 0.5f) / 8000.0f; }  // near the 1/256 quantum's half
            xy.push_back(x); xy.push_back(y);
            const CicDeposit d = cicDeposit(glm::vec2(x, y), metrics::GRID_RES);
            for (int j = 0; j < 4; ++j) { cells.push_back(uint32_t(d.cell[j])); weights.push_back(d.weight[j]); }
        }
        for (int k = 0; k < 2000; ++k) {
            const uint32_t w = uint32_t(rng.next() % (MASS_ONE + 1));
            const float s = rng.uniform() * 3.0f;
            wts.push_back(w); sp.push_back(s); sms.push_back(speedMass(w, s));
        }
        put("metrics/cic/xy", 'f', xy, { 2000, 2 });
        put("metrics/cic/cell", 'u', cells, { 2000, 4 });
        put("metrics/cic/weight", 'u', weights, { 2000, 4 });
        put("metrics/speed/weight", 'u', wts);
        put("metrics/speed/mps", 'f', sp);
        put("metrics/speed/mass", 'u', sms);
    }
    // the cell table: walkable areas with partial and excluded cells
    MetricVenue mv;
    mv.res = metrics::GRID_RES;
    mv.walkableM2.assign(CELLS, 1.0f);
    mv.zone.assign(CELLS, 0u);
    mv.included.assign(CELLS, 1u);
    for (int c = 0; c < CELLS; ++c) {
        const uint64_t r = rng.next() % 10;
        if (r == 0) mv.walkableM2[size_t(c)] = rng.uniform();          // partial
        if (r == 1) mv.walkableM2[size_t(c)] = 0.05f * rng.uniform();  // below MIN_WALKABLE_M2
        mv.included[size_t(c)] = mv.walkableM2[size_t(c)] >= float(MIN_WALKABLE_M2) ? 1u : 0u;
    }
    const std::vector<GpuMetricCellInfo> info = mv.cellInfo();
    std::vector<float> walk(mv.walkableM2);
    std::vector<uint32_t> thr;
    for (const GpuMetricCellInfo& i : info) for (int k = 0; k < 4; ++k) thr.push_back(i.thresholdMass[k]);
    put("metrics/info/walkable_m2", 'f', walk);
    put("metrics/info/threshold_mass", 'u', thr, { uint64_t(CELLS), 4 });
    // the states
    const int n = 3000, steps = 6;
    std::vector<GpuBoid> st(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        float x, y;
        if (i < 1200) { x = 0.10f + 0.12f * rng.uniform(); y = -0.20f + 0.12f * rng.uniform(); }   // 3.75 m square: ~85 per m^2
        else if (i < 1300) { x = float(int(rng.uniform() * 60.0f) - 30) / 31.25f; y = float(int(rng.uniform() * 60.0f) - 30) / 31.25f; }
        else { x = rng.uniform() * 2.3f - 1.15f; y = rng.uniform() * 2.3f - 1.15f; }                // beyond +-1.024: clamped
        const float vx = rng.uniform() * 0.2f - 0.1f, vy = rng.uniform() * 0.2f - 0.1f;
        st[size_t(i)].position = glm::vec4(x, y, 0.0f, (i % 11 == 0) ? 0.0f : 1.0f);
        st[size_t(i)].velocity = glm::vec4(vx, vy, 0.0f, std::sqrt(vx * vx + vy * vy));
    }
    MetricsReference ref(n, info, nullptr);
    std::vector<float> states;
    auto putState = [&](const std::vector<GpuBoid>& v) {
        for (const GpuBoid& b : v) { states.push_back(b.position.x); states.push_back(b.position.y); states.push_back(b.position.w);
                                     states.push_back(b.velocity.x); states.push_back(b.velocity.y); }
    };
    putState(st);
    std::vector<uint32_t> counters, lastMass, lastSpeed;
    for (int k = 0; k < steps; ++k) {
        std::vector<GpuBoid> nx = st;
        for (int i = 0; i < n; ++i) {
            GpuBoid& b = nx[size_t(i)];
            if (b.position.w == 0.0f) continue;
            b.position.x += b.velocity.x * 0.01f;
            b.position.y += b.velocity.y * 0.01f;
            if (rng.next() % 23 == 0) b.position.w = 0.0f;   // leaves in this step
        }
        ref.record(st, nx);
        st = nx;
        putState(st);
        const GpuMetricCounters& c = ref.counters;
        counters.push_back(c.step); counters.push_back(c.active); counters.push_back(c.evacuated); counters.push_back(c.activeAccum);
        uint32_t exitSum = 0;
        for (uint32_t e : c.exits) exitSum += e;
        counters.push_back(exitSum);
        for (const GpuMetricCell& cell : ref.cells) { lastMass.push_back(cell.lastMass); lastSpeed.push_back(cell.lastSpeedMass); }
    }
    put("metrics/run/states", 'f', states, { uint64_t(steps + 1), uint64_t(n), 5 });
    put("metrics/run/counters", 'u', counters, { uint64_t(steps), 5 });
    put("metrics/run/evac_step", 'u', ref.evacStep);
    put("metrics/run/last_mass", 'u', lastMass, { uint64_t(steps), uint64_t(CELLS) });
    put("metrics/run/last_speed_mass", 'u', lastSpeed, { uint64_t(steps), uint64_t(CELLS) });
    std::vector<uint32_t> cells;
    for (const GpuMetricCell& cell : ref.cells) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(&cell);
        for (size_t j = 0; j < sizeof(GpuMetricCell) / 4; ++j) cells.push_back(w[j]);
    }
    put("metrics/run/cells", 'u', cells, { uint64_t(CELLS), uint64_t(sizeof(GpuMetricCell) / 4) });
}
// The reference's seat spawn (seatedSpawnSeats, src/crowd_sim.cpp:652-674): a partial Fisher-Yates over the seat indices
// on mt19937's raw output, the chosen seats in seat order, padded with -1, on venues S, M and L
void spawnCases() {
    const char* names[3] = { "S", "M", "L" };
    for (int v = 0; v < 3; ++v) {
        const Venue& venue = venueByName(names[v]);
        const int seats = int(seated::seats(venue.layout).size());
        int pow2 = 256;
        while (pow2 < seats) pow2 <<= 1;
        struct Case { int n; double load; unsigned seed; };
        const Case cases[] = { { seats, 1.0, 1337u }, { pow2, 1.0, 1337u }, { pow2, 0.5, 7u }, { seats, 0.3, 42u }, { pow2, 0.05, 2024u } };
        std::vector<int32_t> out, meta;
        for (const Case& c : cases) {
            const std::vector<int> ch = seatedSpawnSeats(venue, c.n, c.load, c.seed);
            meta.push_back(c.n);
            meta.push_back(int32_t(std::lround(c.load * 1000.0)));
            meta.push_back(int32_t(c.seed));
            out.insert(out.end(), ch.begin(), ch.end());
        }
        put(std::string("spawn/") + names[v] + "/seats", 'i', std::vector<int32_t>{ seats });
        put(std::string("spawn/") + names[v] + "/cases", 'i', meta);
        put(std::string("spawn/") + names[v] + "/chosen", 'i', out);
    }
}

// The analytic seated venues S, M and L (src/venue.cpp): the layout buildSeatedLayout makes (every scalar and array), the cells
// of every band, the seats, and the heightfield queries (sdfGrad, floorHeight, cellAt, wall, floorSlope, exitByLat, gangwayAt)
// at seeded probe points: uniform over the plan at the step's wall plane, above and below the surface, exactly on it, on cell
// and band boundaries (unfolded into a random octant), on the diagonals and on the axes
void venueCases() {
    const char* names[3] = { "S", "M", "L" };
    for (int v = 0; v < 3; ++v) {
        const Venue& venue = venueByName(names[v]);
        const SeatedLayout& L = venue.layout;
        const std::string p = std::string("venue/") + names[v] + "/";
        put(p + "layout_f", 'f', std::vector<float>{ L.standFront, L.row0, L.rowDepth, L.clearway, L.concFront, L.concBack, L.outerFace,
                                                     L.landingEnd, L.avenueHalf, L.lat0, L.gangwayLo, L.gangwayHi, L.exitLo, L.exitHi,
                                                     L.cornerInset, L.seatInset, L.faceMargin, L.wallAbove, L.wallClear, L.firstRow,
                                                     L.rowRise, L.seatTop, L.barrierTop, L.concourseZ, L.wallTop, L.rake, L.seatWidth,
                                                     L.domainHalf, L.cornerDoorLo });
        put(p + "layout_i", 'i', std::vector<int32_t>{ L.rows, L.maxBandSteps, L.gangways, L.doors, L.seatTotal });
        put(p + "layout_gw", 'f', std::vector<float>(L.gwLo.begin(), L.gwLo.end()));
        put(p + "layout_gw_hi", 'f', std::vector<float>(L.gwHi.begin(), L.gwHi.end()));
        put(p + "layout_gw_row", 'i', std::vector<int32_t>(L.gwRow.begin(), L.gwRow.end()));
        put(p + "layout_gw_seated", 'i', std::vector<int32_t>(L.gwSeated.begin(), L.gwSeated.end()));
        put(p + "layout_door_lo", 'f', std::vector<float>(L.doorLo.begin(), L.doorLo.end()));
        put(p + "layout_door_hi", 'f', std::vector<float>(L.doorHi.begin(), L.doorHi.end()));
        put(p + "layout_d", 'd', std::vector<double>{ L.doorWidthM, L.cornerDoorWidthM, L.totalExitWidthM, L.requiredExitWidthM });
        put(p + "venue_f", 'f', std::vector<float>{ venue.domainMin, venue.domainSpan, venue.exitDepth, venue.exitMaxZ });
        put(p + "venue_i", 'i', std::vector<int32_t>{ venue.gridRes, venue.metricGridRes, seated::exitCount(L), seated::gangwayCount(L) });
        {
            std::vector<double> w;
            for (int e = 0; e < seated::exitCount(L); ++e) w.push_back(seated::exitWidthM(L, e));
            put(p + "exit_width_m", 'd', w);
        }
        // cells: (band, c, kind, row) and (l0, l1, lo, hi)
        std::vector<int32_t> ci;
        std::vector<float> cf;
        const int nb = seated::bandCount(L);
        for (int j = 0; j < nb; ++j) {
            float lo, hi, a, b;
            int k, r;
            seated::band(L, j, lo, hi);
            for (int c = 0; seated::cell(L, j, c, a, b, k, r); ++c) {
                ci.insert(ci.end(), { j, c, k, r });
                cf.insert(cf.end(), { a, b, lo, hi });
            }
        }
        put(p + "cells_i", 'i', ci, { uint64_t(ci.size() / 4), 4 });
        put(p + "cells_f", 'f', cf, { uint64_t(cf.size() / 4), 4 });
        // seats (venue L: every 17th)
        const std::vector<seated::Seat> seats = seated::seats(L);
        const size_t stride = v == 2 ? 17 : 1;
        std::vector<float> sp;
        std::vector<int32_t> si;
        for (size_t i = 0; i < seats.size(); i += stride) {
            const seated::Seat& s = seats[i];
            sp.insert(sp.end(), { s.pos.x, s.pos.y, s.pos.z, s.lat });
            si.insert(si.end(), { int32_t(i), s.quadrant, s.wing, s.row, s.segment, s.index });
        }
        put(p + "seat_count", 'i', std::vector<int32_t>{ int32_t(seats.size()) });
        put(p + "seats_f", 'f', sp, { uint64_t(sp.size() / 4), 4 });
        put(p + "seats_i", 'i', si, { uint64_t(si.size() / 6), 6 });

        // probe points
        Rng rng{ 0x5EA7ED00ull + uint64_t(v) };
        const float H = L.landingEnd + 0.02f;
        auto U = [&](float a, float b) { return a + (b - a) * rng.uniform(); };
        auto octant = [&](float lat, float dep) {   // unfold (lat, dep) into a random octant
            const uint64_t o = rng.next();
            float x = (o & 1) ? dep : lat, y = (o & 1) ? lat : dep;
            if (o & 2) x = -x;
            if (o & 4) y = -y;
            return glm::vec2(x, y);
        };
        const int per = v == 0 ? 1024 : 512;
        std::vector<float> pts, es;
        std::vector<int32_t> kinds;
        for (int kind = 0; kind < 6; ++kind)
            for (int i = 0; i < per; ++i) {
                glm::vec2 xy;
                if (kind <= 2) xy = glm::vec2(U(-H, H), U(-H, H));
                else if (kind == 3) {   // a cell boundary: lat at l0 / l1 / inside, dep at lo / hi / just inside
                    int j = int(rng.next() % uint64_t(nb));
                    float lo, hi;
                    seated::band(L, j, lo, hi);
                    if (hi > H) hi = H;
                    int cn = 0;
                    float a, b;
                    int k, r;
                    while (seated::cell(L, j, cn, a, b, k, r)) ++cn;
                    seated::cell(L, j, int(rng.next() % uint64_t(cn)), a, b, k, r);
                    const uint64_t m = rng.next();
                    const float dep = (m % 3) == 0 ? lo : (m % 3) == 1 ? hi : U(lo, lo + std::min(0.05f, hi - lo));
                    const float bb = std::min(b, dep);
                    const float lat = ((m / 3) % 3) == 0 ? std::min(a, dep) : ((m / 3) % 3) == 1 ? bb : U(std::min(a, dep), bb);
                    xy = octant(lat, dep);
                } else if (kind == 4) {   // the diagonals
                    const float t = U(0.0f, H);
                    const uint64_t o = rng.next();
                    xy = glm::vec2((o & 1) ? -t : t, (o & 2) ? -t : t);
                } else {   // the axes
                    const float t = U(-H, H);
                    xy = (rng.next() & 1) ? glm::vec2(t, 0.0f) : glm::vec2(0.0f, t);
                }
                const float f = seated::floorHeight(L, xy);
                float z;
                const uint64_t zc = rng.next() % 4;
                if (kind == 0) z = f + BODY_RADIUS + L.wallAbove;   // the step's wall plane
                else if (kind == 1) z = f + U(-0.03f, 0.06f);
                else if (kind == 2) z = f;
                else z = zc == 0 ? f : zc == 1 ? f + BODY_RADIUS : zc == 2 ? f + BODY_RADIUS + L.wallAbove : f + U(-0.02f, 0.04f);
                const float ang = U(0.0f, 6.2831853f);
                pts.insert(pts.end(), { xy.x, xy.y, z });
                es.insert(es.end(), { std::cos(ang), std::sin(ang) });
                kinds.push_back(kind);
            }
        const size_t n = kinds.size();
        std::vector<float> grad, floorH, wallF, wallP, wallW, slope;
        std::vector<int32_t> cellK, exitL, gangway;
        for (size_t i = 0; i < n; ++i) {
            const glm::vec3 q(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            const glm::vec2 xy(q);
            const glm::vec4 g = seated::sdfGrad(L, q);
            grad.insert(grad.end(), { g.x, g.y, g.z, g.w });
            floorH.push_back(seated::floorHeight(L, xy));
            int row = 0, bandIndex = 0;
            const int kind = seated::cellAt(L, xy, &row, &bandIndex);
            cellK.insert(cellK.end(), { kind, row, bandIndex });
            const glm::vec3 wf = seated::wall(L, q, FEELER_REACH), wp = seated::wall(L, q, PLANE_PROBE_REACH), ww = seated::wall(L, q, 0.02f);
            wallF.insert(wallF.end(), { wf.x, wf.y, wf.z });
            wallP.insert(wallP.end(), { wp.x, wp.y, wp.z });
            wallW.insert(wallW.end(), { ww.x, ww.y, ww.z });
            slope.push_back(seated::floorSlope(L, xy, glm::vec2(es[2 * i], es[2 * i + 1])));
            exitL.push_back(seated::exitByLat(L, xy));
            gangway.push_back(seated::gangwayAt(L, xy));
        }
        put(p + "probe/kind", 'i', kinds);
        put(p + "probe/p", 'f', pts, { uint64_t(n), 3 });
        put(p + "probe/e", 'f', es, { uint64_t(n), 2 });
        put(p + "probe/sdf_grad", 'f', grad, { uint64_t(n), 4 });
        put(p + "probe/floor", 'f', floorH);
        put(p + "probe/cell", 'i', cellK, { uint64_t(n), 3 });
        put(p + "probe/wall_feeler", 'f', wallF, { uint64_t(n), 3 });
        put(p + "probe/wall_plane", 'f', wallP, { uint64_t(n), 3 });
        put(p + "probe/wall_wide", 'f', wallW, { uint64_t(n), 3 });
        put(p + "probe/slope", 'f', slope);
        put(p + "probe/exit", 'i', exitL);
        put(p + "probe/gangway", 'i', gangway);
    }
    {   // the default model's incline rules (seated::inclineFreeSpeed / inclineHeadwayTime at EMPIRICAL's time gaps)
        std::vector<float> g = { 0.0f, -0.0f, 1e-30f, -1e-30f, 0.1f, -0.1f, 0.25f, -0.25f, 0.5f, -0.5f, INCLINE_REF_SLOPE, -INCLINE_REF_SLOPE,
                                 0.6f, -0.6f, 2.0f, -2.0f };
        Rng rng{ 0x1C11E5ull };
        for (int i = 0; i < 1024; ++i) g.push_back(-0.7f + 1.4f * float(rng.uniform()));
        const glm::vec2 t = stairHeadwayTimes(StairFlow::Empirical);
        std::vector<float> v, h;
        for (float x : g) {
            v.push_back(seated::inclineFreeSpeed(x));
            h.push_back(seated::inclineHeadwayTime(x, t.x, t.y));
        }
        put("venue/incline/g", 'f', g);
        put("venue/incline/v_free", 'f', v);
        put("venue/incline/t_headway", 'f', h);
        put("venue/incline/consts", 'f', std::vector<float>{ V_MAX, V_STAIR_UP, V_STAIR_DOWN, T_HEADWAY, t.x, t.y, INCLINE_INV_REF_SLOPE });
    }
}

uint64_t fnv64(const void* data, size_t bytes, uint64_t h = 0xcbf29ce484222325ull) {   // FNV-1a 64
    const unsigned char* b = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) { h ^= b[i]; h *= 0x100000001b3ull; }
    return h;
}

void putBoids(const std::string& p, const std::vector<GpuBoid>& st) {
    std::vector<float> pos, vel;
    for (const GpuBoid& b : st) {
        pos.insert(pos.end(), { b.position.x, b.position.y, b.position.z, b.position.w });
        vel.insert(vel.end(), { b.velocity.x, b.velocity.y, b.velocity.z, b.velocity.w });
    }
    put(p + "pos", 'f', pos, { uint64_t(st.size()), 4 });
    put(p + "vel", 'f', vel, { uint64_t(st.size()), 4 });
}

// A seated venue's navigation field and candidate table (nav::forVenue, nav::choiceForVenue) as FNV-1a checksums of the
// cell classes, edges, exit regions, floors, the index and every entry's slots and distance words, with every 997th cell
void putNav(const std::string& p, const Venue& venue) {
    // the navigation field and the choice table the run used
    const nav::Field& f = nav::forVenue(venue);
    const nav::ChoiceTable& t = nav::choiceForVenue(venue);
    const nav::Grid& g = f.grid;
    put(p + "grid", 'd', std::vector<double>{ double(g.res), double(g.origin), double(g.cell), double(g.invCell) });
    std::vector<uint32_t> slots, dists;
    for (const GpuNavChoiceCell& c : t.entries) {
        for (int j = 0; j < nav::CHOICE_K; ++j) slots.push_back(c.slot[j]);
        for (int j = 0; j < nav::CHOICE_DIST_WORDS; ++j) dists.push_back(c.dist[j]);
    }
    const uint64_t h[] = { fnv64(f.cls.data(), f.cls.size()), fnv64(f.out.data(), f.out.size()),
                   fnv64(f.exitRegion.data(), f.exitRegion.size()), fnv64(f.floorZ.data(), f.floorZ.size() * 4),
                   fnv64(t.index.data(), t.index.size() * 4), fnv64(slots.data(), slots.size() * 4),
                   fnv64(dists.data(), dists.size() * 4) };
    std::vector<uint32_t> hw;
    for (uint64_t v : h) { hw.push_back(uint32_t(v)); hw.push_back(uint32_t(v >> 32)); }
    put(p + "fnv", 'u', hw);
    put(p + "sizes", 'i', std::vector<int32_t>{ int32_t(f.cls.size()), int32_t(t.entries.size()) });
    put(p + "scalars", 'f', std::vector<float>{ t.secondsPerUnit, t.distQuantum });
    put(p + "inv_cap_s", 'f', t.invCapS);
    std::vector<int32_t> cells;
    std::vector<uint32_t> info, sl, ds;
    std::vector<uint8_t> cls, out;
    std::vector<int32_t> er;
    std::vector<float> fz;
    for (size_t i = 0; i < f.cls.size(); i += 997) {
        cells.push_back(int32_t(i));
        cls.push_back(f.cls[i]);
        out.push_back(f.out[i]);
        er.push_back(f.exitRegion[i]);
        fz.push_back(f.floorZ[i]);
        info.push_back(t.index[i]);
        const GpuNavChoiceCell& c = t.at(i);
        for (int j = 0; j < nav::CHOICE_K; ++j) sl.push_back(c.slot[j]);
        for (int j = 0; j < nav::CHOICE_DIST_WORDS; ++j) ds.push_back(c.dist[j]);
    }
    put(p + "sample_cell", 'i', cells);
    put(p + "sample_cls", 'b', cls);
    put(p + "sample_out", 'b', out);
    put(p + "sample_exit_region", 'i', er);
    put(p + "sample_floor", 'f', fz);
    put(p + "sample_index", 'u', info);
    put(p + "sample_slots", 'u', sl, { uint64_t(cells.size()), uint64_t(nav::CHOICE_K) });
    put(p + "sample_dists", 'u', ds, { uint64_t(cells.size()), uint64_t(nav::CHOICE_DIST_WORDS) });
}

// The reference's seated 3D step (CpuCrowdSim on an analytic venue: floors, gravity, the fence, the exact walls, the
// step-up limit) on venue S, its 100 % seat spawn in 4,096 slots (3,728 seated), counting sort:
//   seated/S/drive/    a forced drive per agent (setDriveTable: seeded random directions), 40 steps
//   seated/S/default/  the default model (its navigation field, queue-aware choice, incline speed, EMPIRICAL stairs), 200 steps
// States (position, velocity) and contacts at chosen steps; the default run's choice state; its navigation field and
// choice table as FNV-1a checksums with every 997th cell written out
void seatedCases() {
    const Venue& venue = venueByName("S");
    const int n = 4096;
    int active = 0;
    const std::vector<GpuBoid> init = spawnSeatedCrowd(venue, n, 1.0, 1337u, &active);
    putBoids("seated/S/init/", init);
    {
        CpuCrowdSim sim(venue, std::vector<float>{}, 0, CpuCrowdSim::SortMode::Counting);
        std::vector<glm::vec2> drive(static_cast<size_t>(n));
        Rng rng{ 0xD21BEull };
        for (int i = 0; i < n; ++i) {
            const float a = 6.2831853f * rng.uniform();
            drive[size_t(i)] = glm::vec2(std::cos(a), std::sin(a));
        }
        std::vector<float> dv;
        for (const glm::vec2& d : drive) dv.insert(dv.end(), { d.x, d.y });
        put("seated/S/drive/dir", 'f', dv, { uint64_t(n), 2 });
        sim.setDriveTable(&drive);
        sim.setState(init);
        const int keep[] = { 1, 2, 3, 5, 10, 20, 40 };
        int k = 0;
        for (int step = 1; step <= 40; ++step) {
            sim.step();
            if (step == keep[k]) {
                const std::string p = "seated/S/drive/step" + std::to_string(step) + "/";
                putBoids(p, sim.state());
                put(p + "contacts", 'b', sim.contacts());
                ++k;
            }
        }
    }
    {
        CpuCrowdSim sim(venue, std::vector<float>{}, 0, CpuCrowdSim::SortMode::Counting);
        sim.setState(init);
        const int keep[] = { 1, 2, 3, 5, 10, 25, 50, 100, 200 };
        int k = 0;
        for (int step = 1; step <= 200; ++step) {
            sim.step();
            if (step == 1) {
                std::vector<uint32_t> key, id, start, count;
                for (const GpuGridInfo& g : sim.grid()) { key.push_back(g.gridID); id.push_back(g.boidID); }
                for (const GpuCellData& c : sim.cells()) { start.push_back(c.startIndex); count.push_back(c.agentCount); }
                put("seated/S/default/grid_key", 'u', key);
                put("seated/S/default/grid_id", 'u', id);
                put("seated/S/default/cell_start", 'u', start);
                put("seated/S/default/cell_count", 'u', count);
            }
            if (step == keep[k]) {
                const std::string p = "seated/S/default/step" + std::to_string(step) + "/";
                putBoids(p, sim.state());
                put(p + "contacts", 'b', sim.contacts());
                putChoice(p, sim.choiceState());
                ++k;
            }
        }
        putNav("seated/S/nav/", venue);
    }
    {   // venue M (two exit types beyond CHOICE_K exits: the type rule), its 100 % seat spawn in 16,384 slots: 10 steps
        const Venue& m = venueByName("M");
        const std::vector<GpuBoid> initM = spawnSeatedCrowd(m, 16384, 1.0, 1337u, &active);
        putBoids("seated/M/init/", initM);
        CpuCrowdSim sim(m, std::vector<float>{}, 0, CpuCrowdSim::SortMode::Counting);
        sim.setState(initM);
        for (int step = 1; step <= 10; ++step) sim.step();
        putBoids("seated/M/default/step10/", sim.state());
        put("seated/M/default/step10/contacts", 'b', sim.contacts());
        putChoice("seated/M/default/step10/", sim.choiceState());
        putNav("seated/M/nav/", m);
    }
}

} // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: ref_driver <out.bin>\n"); return 2; }
    g_out = std::fopen(argv[1], "wb");
    if (!g_out) return 1;
    const std::vector<float> vol = flatVolume();
    // boxes x in [0.05, 0.05 + 2 half]: every agent stays farther than OBSTACLE_RANGE from the pitch fence
    // (|x|, |y| < 0.30) and, in the first step, farther than R_PERC from the origin (the NaN agent's reset)
    terraceCase("t_sparse", vol, 512, 0.10f, 0.011f, 6, 1337, 0, false, 0);
    terraceCase("t_dense", vol, 1024, 0.085f, 0.0085f, 8, 4242, 37, true, 5);
    // the same scenes sorted by the GPU's bitonic network (SortMode::GpuBitonic: its unstable permutation)
    terraceCase("t_bsparse", vol, 512, 0.10f, 0.011f, 6, 1337, 0, false, 0, CpuCrowdSim::SortMode::GpuBitonic);
    terraceCase("t_bdense", vol, 1024, 0.085f, 0.0085f, 8, 4242, 37, true, 5, CpuCrowdSim::SortMode::GpuBitonic);
    choiceCases();
    metricsCases();
    spawnCases();
    venueCases();
    seatedCases();
    {   // the L corridor (tools/nav_check.hpp:166-177) at 0.1 m, point agents
        const std::vector<Rect> walls = { { { -0.3f, -0.3f }, { 20.3f, 0.0f } }, { { -0.3f, 0.0f }, { 0.0f, 2.3f } }, { { 0.0f, 2.0f }, { 18.0f, 2.3f } },
                                          { { 17.7f, 2.3f }, { 18.0f, 15.0f } }, { { 20.0f, 0.0f }, { 20.3f, 15.0f } } };
        navCase("lcorr", walls, { { { 17.0f, 15.025f }, { 21.0f, 16.0f } } }, {}, testGrid(-1.0f, 22.5f, 0.1), 0.0f);
    }
    {   // a room with a body radius and three exits (the table's candidate order by distance)
        const std::vector<Rect> walls = roomWalls(0.0f, 0.0f, 12.0f, 9.0f, 0.3f, { { 2, 5.4f, 6.6f }, { 0, 1.0f, 2.0f }, { 1, 6.0f, 7.5f } });
        const std::vector<Rect> exits = { { { 5.4f, 9.325f }, { 6.6f, 10.0f } }, { { 1.0f, -1.0f }, { 2.0f, -0.325f } }, { { 12.325f, 6.0f }, { 13.0f, 7.5f } } };
        navCase("room3", walls, exits, { 1.2f, 1.0f, 1.5f }, testGrid(-1.0f, 14.0f, 0.1), 0.2f);
    }
    std::fclose(g_out);
    return 0;
}
