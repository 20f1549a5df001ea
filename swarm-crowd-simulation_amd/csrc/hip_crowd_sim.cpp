// HipCrowdSim: device buffers, upload/readback and the step's launch sequence (CrowdSim::step,
// src/crowd_sim.cpp:316-412: hash, sort, cells + gather, exit choice, predict, PROJ_ITERATIONS - 1 projection
// iterations, final; then the ping-pong halves swap).
#include "hip_crowd_sim.hpp"
#include "venue.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <functional>

namespace swarm {

namespace {
// a float's shortest round-trip decimal as a double (how the reference folds constants, sim_params.hpp:36-50)
double decimalOf(float f) {
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof buf, f);
    *res.ptr = '\0';
    return std::strtod(buf, nullptr);
}
} // namespace

void modelParamsFor(float tauS, float& tau, float& relaxStep) {
    tau = float(decimalOf(tauS) / decimalOf(TIME_SCALE_S_PER_UNIT));
    relaxStep = float(std::exp(-decimalOf(DT) / decimalOf(tau)));
}

void DeviceScene::upload(const Scene& s) {
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 4));
        owned.push_back(p);
        if (bytes) SWARM_HIP_CHECK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        return p;
    };
    SceneView& v = view;
    v.xmin = s.xmin; v.ymin = s.ymin; v.xmax = s.xmax; v.ymax = s.ymax;
    v.hashSpanX = s.hashSpanX; v.hashSpanY = s.hashSpanY; v.hashNX = s.hashNX; v.hashNY = s.hashNY;
    v.percRing = s.percRing; v.projRing = s.projRing;
    v.sdfNX = s.sdfNX; v.sdfNY = s.sdfNY; v.sdfOX = s.sdfOX; v.sdfOY = s.sdfOY; v.sdfInvH = s.sdfInvH;
    v.sdf = static_cast<const float*>(put(s.sdf.data(), s.sdf.size() * 4));
    v.sdfMin = static_cast<const float*>(put(s.sdfMin.data(), s.sdfMin.size() * 4));
    v.sdfMinWin = s.sdfMinWin;
    v.navRX = s.nav.resX; v.navRY = s.nav.resY; v.navOX = s.nav.originX; v.navOY = s.nav.originY;
    v.navCell = s.nav.cell; v.navInvCell = s.nav.invCell; v.K = s.K; v.E = s.E;
    v.navInfo = static_cast<const uint32_t*>(put(s.navInfo.data(), s.navInfo.size() * 4));
    {   // one record per nav cell: class bits, K slots, K distance quanta, padded to 16 bytes
        const size_t cells = s.navInfo.size(), K = size_t(s.K);
        const size_t stride = (1 + K + (K + 1) / 2 + 3) / 4 * 4;
        if (s.navSlot.size() != cells * K || s.navDist.size() != cells * K)
            throw std::invalid_argument("DeviceScene: navigation tables of inconsistent sizes");
        std::vector<uint32_t> rec(cells * stride, 0u);
        for (size_t c = 0; c < cells; ++c) {
            uint32_t* r = rec.data() + c * stride;
            r[0] = s.navInfo[c];
            std::memcpy(r + 1, s.navSlot.data() + c * K, K * 4);
            std::memcpy(r + 1 + K, s.navDist.data() + c * K, K * 2);
        }
        v.navRec = static_cast<const uint32_t*>(put(rec.data(), rec.size() * 4));
        v.navStride = int(stride);
    }
    v.invCapS = static_cast<const float*>(put(s.invCapS.data(), s.invCapS.size() * 4));
    v.secondsPerUnit = s.secondsPerUnit; v.distQuantum = s.distQuantum;
    v.stair = static_cast<const float*>(put(s.stair.data(), s.stair.size() * 4));
    std::vector<float> ex;
    for (const Rect& r : s.exits) { ex.push_back(r.lox); ex.push_back(r.loy); ex.push_back(r.hix); ex.push_back(r.hiy); }
    v.exitRects = static_cast<const float*>(put(ex.data(), ex.size() * 4));
    std::vector<float4> wb(s.wallBinList.size());   // each bin entry's box, in the bins' (ascending box) order
    for (size_t k = 0; k < wb.size(); ++k) {
        const Rect& r = s.walls.at(s.wallBinList[k]);
        wb[k] = make_float4(r.lox, r.loy, r.hix, r.hiy);
    }
    v.wallBinStart = static_cast<const uint32_t*>(put(s.wallBinStart.data(), s.wallBinStart.size() * 4));
    v.wallBinBox = static_cast<const float4*>(put(wb.data(), wb.size() * sizeof(float4)));
    v.venue = s.hasVenue ? static_cast<const seated::Layout*>(put(&s.venue, sizeof s.venue)) : nullptr;
    v.venue3d = s.venue3d;
    v.exitDepth = s.exitDepth;
    v.exitMaxZ = s.exitMaxZ;
    v.domainMax = s.domainMax;
    v.venueCells = seated::CellTable{ nullptr, nullptr, nullptr, nullptr, nullptr };
    v.venueBlob = nullptr;
    v.venueBlobLayout = seated::CellBlob{};
    if (s.hasVenue) {   // every band's cells, generated once (the same values the generator gives), as one blob
        const CellTableData t = makeCellTable(s.venue);
        v.venueBlob = static_cast<const char*>(put(t.blob.data(), t.blob.size()));
        v.venueBlobLayout = t.blobLayout;
        v.venueCells = seated::cellTableAt(v.venueBlob, t.blobLayout);
    }
}

void DeviceScene::release() {
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
}

template <class T> T* HipCrowdSim::alloc(size_t count) {
    void* p = nullptr;
    SWARM_HIP_CHECK(hipMalloc(&p, (count > 0 ? count : 1) * sizeof(T)));
    bufs_.push_back(p);
    SWARM_HIP_CHECK(hipMemsetAsync(p, 0, (count > 0 ? count : 1) * sizeof(T), stream_));
    return static_cast<T*>(p);
}

HipCrowdSim::HipCrowdSim(const Scene& scene, int numAgents, const SimOptions& opt) : opt_(opt) {
    if (numAgents <= 0) throw std::invalid_argument("HipCrowdSim: agent count must be positive");
    init({ &scene }, { 0 }, { numAgents }, false);
}

HipCrowdSim::HipCrowdSim(const std::vector<const Scene*>& layouts, const std::vector<int>& scenLayout, const std::vector<int>& scenAgents,
                         const SimOptions& opt)
    : opt_(opt) {
    init(layouts, scenLayout, scenAgents, true);
}

void HipCrowdSim::init(const std::vector<const Scene*>& layouts, const std::vector<int>& scenLayout, const std::vector<int>& scenAgents,
                       bool batched) {
    if (opt_.projIters < 0) throw std::invalid_argument("HipCrowdSim: projection iterations must be >= 0");
    if (layouts.empty() || scenLayout.empty() || scenLayout.size() != scenAgents.size())
        throw std::invalid_argument("HipCrowdSim: a batch needs scenes and one layout and agent count per scenario");
    for (const Scene* s : layouts)
        if (!s || s->hashNX <= 0 || s->hashNY <= 0 || s->sdfNX < 2 || s->sdfNY < 2 || s->nav.resX <= 0 || s->nav.resY <= 0 || s->E < 0 ||
            s->E >= PER_EXIT_STRIDE)
            throw std::invalid_argument("HipCrowdSim: invalid scene (grids, or more than 63 exits)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("HipCrowdSim: no HIP device");
    batched_ = batched;
    for (const Scene* s : layouts)
        if (s->venue3d && batched)
            throw std::invalid_argument("HipCrowdSim: a seated venue's own 3D step runs on a plain context, not in a batch");
    venue3d_ = !batched && layouts[0]->venue3d != 0;
    S_ = int(scenLayout.size());
    const size_t nScen = static_cast<size_t>(S_);
    std::vector<uint32_t> cellBase(nScen), agentBase(nScen), lay(nScen);
    size_t nTotal = 0, cTotal = 0;
    maxExits_ = 0;
    for (int s = 0; s < S_; ++s) {
        const int l = scenLayout[size_t(s)];
        if (l < 0 || l >= int(layouts.size())) throw std::invalid_argument("HipCrowdSim: scenario layout out of range");
        if (scenAgents[size_t(s)] <= 0) throw std::invalid_argument("HipCrowdSim: a scenario needs at least one agent");
        lay[size_t(s)] = uint32_t(l);
        cellBase[size_t(s)] = uint32_t(cTotal);
        agentBase[size_t(s)] = uint32_t(nTotal);
        cTotal += size_t(layouts[size_t(l)]->hashNX) * size_t(layouts[size_t(l)]->hashNY);
        nTotal += size_t(scenAgents[size_t(s)]);
        maxExits_ = std::max(maxExits_, layouts[size_t(l)]->E);
    }
    if (nTotal > size_t(INT32_MAX) / 2 || cTotal > size_t(UINT32_MAX) / 2) throw std::invalid_argument("HipCrowdSim: batch too large");
    n_ = int(nTotal);
    C_ = int(cTotal);
    nb_ = C_ + (n_ + HASH_BLOCK - 1) / HASH_BLOCK;
    exitStride_ = batched_ ? std::max(1, maxExits_) : CHOICE_QUEUE_STRIDE;
    modelParamsFor(opt_.tauS, mp_.tau, mp_.relaxStep);
    mp_.obstacleGain = opt_.obstacleGain;
    mp_.projIters = opt_.projIters;
    SWARM_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    scenes_.resize(layouts.size());
    for (size_t l = 0; l < layouts.size(); ++l) scenes_[l].upload(*layouts[l]);
    if (!batched) {   // the safety metrics' floor grid and its threshold masses (MetricVenue::cellInfo's rule)
        const Scene& sc = *layouts[0];
        mgrid_ = MetricGrid{ sc.metricResX, sc.metricResY, sc.metricCellsPerUnit, sc.metricOffX, sc.metricOffY,
                             float(1.0 * double(WORLD_SCALE_M_PER_UNIT) / double(TIME_SCALE_S_PER_UNIT)) };
        mWalkable_ = sc.metricWalkableM2;
        static constexpr double THRESHOLD_P_M2[3] = { 1.08, 2.17, 5.0 };   // LoS E, LoS F, crush (metrics.hpp:82-94)
        mThr_.assign(mWalkable_.size() * 4, METRIC_NEVER);
        for (size_t c = 0; c < mWalkable_.size(); ++c) {
            if (!(double(mWalkable_[c]) >= 0.10)) continue;   // MIN_WALKABLE_M2 (metrics.hpp:138)
            for (int k = 0; k < 3; ++k) {
                const double m = std::ceil(THRESHOLD_P_M2[k] * double(mWalkable_[c]) * double(METRIC_MASS_ONE));
                mThr_[4 * c + size_t(k)] = uint32_t(std::min(std::max(m, 1.0), double(METRIC_NEVER)));
            }
        }
    }
    const size_t n = size_t(n_);
    const size_t S = size_t(S_);
    for (int h = 0; h < 2; ++h) {
        half_[h].x = alloc<float>(n); half_[h].y = alloc<float>(n); half_[h].vx = alloc<float>(n);
        half_[h].vy = alloc<float>(n); half_[h].spd = alloc<float>(n); half_[h].w = alloc<float>(n);
        half_[h].z = alloc<float>(n); half_[h].vz = alloc<float>(n);
    }
    seatedSlots_.pz = alloc<float>(n);
    seatedSlots_.pvz = alloc<float>(n);
    rad_ = alloc<float>(n); v0_ = alloc<float>(n);
    driveX_ = alloc<float>(n); driveY_ = alloc<float>(n);
    grid_.key = alloc<uint32_t>(n); grid_.loc = alloc<uint32_t>(n);
    grid_.bcnt = alloc<uint32_t>(size_t(nb_)); grid_.bstart = alloc<uint32_t>(size_t(nb_) + 1);
    grid_.tileSum = alloc<uint32_t>(size_t((nb_ + SCAN_TILE - 1) / SCAN_TILE)); grid_.tmp = alloc<uint32_t>(n);
    grid_.pairs = alloc<uint64_t>(n);
    sb_.sid = alloc<uint32_t>(n); sb_.skey = alloc<uint32_t>(n);
    sb_.bstart = grid_.bstart;
    sb_.sx = alloc<float>(n); sb_.sy = alloc<float>(n); sb_.sr = alloc<float>(n);
    sb_.egx = alloc<float>(n); sb_.egy = alloc<float>(n);
    sb_.pw = alloc<float>(n); sb_.pvx = alloc<float>(n); sb_.pvy = alloc<float>(n);
    sb_.planes = alloc<float4>(n);
    sb_.v0 = v0_;
    sb_.nbr = alloc<uint32_t>(size_t(NBR_CAP) * n);
    sb_.nbrD = venue3d_ ? alloc<float>(size_t(NBR_CAP) * n) : nullptr;   // the seated step's height offsets
    sb_.nbrCount = alloc<uint32_t>(n);
    sb_.listBroken = alloc<uint32_t>(S);
    sb_.nScen = S_;
    sb_.listMoves = -1.0f;
    pa_.p = alloc<float2>(n); pa_.acc = alloc<float>(n);
    pb_.p = alloc<float2>(n); pb_.acc = alloc<float>(n);
    ch_.assign = alloc<uint32_t>(n); ch_.pending = alloc<uint32_t>(n);
    ch_.queue = alloc<uint32_t>(S * QUEUE_STRIDE); ch_.counters = alloc<uint32_t>(S * COUNTER_STRIDE);
    ch_.hist = alloc<uint32_t>(S * size_t(exitStride_) * CHOICE_GAIN_BINS);
    ch_.threshold = alloc<uint32_t>(S * THRESHOLD_STRIDE);
    ch_.exitStride = exitStride_;
    ch_.ahead = alloc<uint32_t>(S * 2u * size_t(exitStride_) * CHOICE_AHEAD_BINS);
    ev_.step = alloc<int32_t>(n); ev_.exit = alloc<int32_t>(n); ev_.perExit = alloc<uint32_t>(S * PER_EXIT_STRIDE);
    ev_.inside = alloc<float2>(n); ev_.before = alloc<float2>(n);
    ctrl_ = alloc<uint32_t>(CTRL_WORDS);
    sb_.ctrl = ctrl_;
    const uint32_t sentinel = uint32_t(n_);
    SWARM_HIP_CHECK(hipMemcpyAsync(grid_.bstart + nb_, &sentinel, 4, hipMemcpyHostToDevice, stream_));
    if (batched_) {   // the per-scenario tables
        std::vector<SceneView> views;
        for (const DeviceScene& d : scenes_) views.push_back(d.view);
        std::vector<uint32_t> agentScen(n);
        for (int s = 0; s < S_; ++s)
            for (int i = 0; i < scenAgents[size_t(s)]; ++i) agentScen[size_t(agentBase[size_t(s)]) + size_t(i)] = uint32_t(s);
        SceneView* dv = alloc<SceneView>(views.size());
        uint32_t* dl = alloc<uint32_t>(S);
        uint32_t* dc = alloc<uint32_t>(S);
        uint32_t* da = alloc<uint32_t>(S);
        uint32_t* das = alloc<uint32_t>(n);
        SWARM_HIP_CHECK(hipMemcpyAsync(dv, views.data(), views.size() * sizeof(SceneView), hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(dl, lay.data(), S * 4, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(dc, cellBase.data(), S * 4, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(da, agentBase.data(), S * 4, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(das, agentScen.data(), n * 4, hipMemcpyHostToDevice, stream_));
        bt_.views = dv; bt_.scenLayout = dl; bt_.cellBase = dc; bt_.agentBase = da; bt_.agentScen = das;
        bt_.slotScen = alloc<uint32_t>(n);
    }
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
}

StepLaunch HipCrowdSim::launchDesc() const {
    StepLaunch L;
    L.n = n_;
    L.scenarios = S_;
    L.batched = batched_;
    L.s = scenes_[0].view;
    L.bt = bt_;
    L.mp = mp_;
    L.maxExits = maxExits_;
    L.cells = uint32_t(C_);
    L.physBlock = physBlock_;
    L.physGrid = physGrid_;
    // the overlap loss term reads the candidate lists after the step: they stay whole while it is on
    L.pruneEvery = lossOn_ ? 0 : pruneEvery_;
    L.projLanes = projLanes_;
    L.collectLanes = collectLanes_;
    L.predictLanes = predictLanes_;
    L.bitonic = bitonic_;
    L.listMoves = listMoves_;
    return L;
}

HipCrowdSim::~HipCrowdSim() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    dropGraph();
    for (void* p : bufs_) (void)hipFree(p);
    for (void* p : adjBufs_) (void)hipFree(p);
    for (void* p : tapeBufs_) (void)hipFree(p);
    for (void* p : mBufs_) (void)hipFree(p);
    for (char* p : ckpt_) (void)hipFree(p);
    if (finalSnap_) (void)hipFree(finalSnap_);
    for (DeviceScene& d : scenes_) d.release();
    if (stream2_) {
        (void)hipStreamSynchronize(stream2_);
        (void)hipStreamDestroy(stream2_);
    }
    for (int b = 0; b < 2; ++b) {
        if (evReplay_[b]) (void)hipEventDestroy(evReplay_[b]);
        if (evBack_[b]) (void)hipEventDestroy(evBack_[b]);
    }
    if (evMain_) (void)hipEventDestroy(evMain_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void HipCrowdSim::dropGraph() {
    for (int i = 0; i < 2; ++i) {
        if (graphExec_[i]) { (void)hipGraphExecDestroy(graphExec_[i]); graphExec_[i] = nullptr; }
        if (graph_[i]) { (void)hipGraphDestroy(graph_[i]); graph_[i] = nullptr; }
    }
    dropAdjointGraphs();
}

void HipCrowdSim::dropAdjointGraphs() {
    for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 2; ++i) {
            if (replayExec_[b][i]) { (void)hipGraphExecDestroy(replayExec_[b][i]); replayExec_[b][i] = nullptr; }
            if (replayGraph_[b][i]) { (void)hipGraphDestroy(replayGraph_[b][i]); replayGraph_[b][i] = nullptr; }
            if (backExec_[b][i]) { (void)hipGraphExecDestroy(backExec_[b][i]); backExec_[b][i] = nullptr; }
            if (backGraph_[b][i]) { (void)hipGraphDestroy(backGraph_[b][i]); backGraph_[b][i] = nullptr; }
        }
}

void HipCrowdSim::upload(const float* x, const float* y, const float* vx, const float* vy, const float* w, const float* r,
                         const float* v0) {
    const size_t b = size_t(n_) * 4;
    const size_t S = size_t(S_);
    adjK_ = 0;   // a new state: the last adjoint forward's backward no longer applies
    restoreAgentState(x, y, vx, vy, w);
    SWARM_HIP_CHECK(hipMemcpyAsync(rad_, r, b, hipMemcpyHostToDevice, stream_));
    r0_.assign(r, r + n_);
    SWARM_HIP_CHECK(hipMemcpyAsync(v0_, v0, b, hipMemcpyHostToDevice, stream_));
    // a fresh exit choice (nav::ChoiceState::reset) and evacuation record
    std::vector<uint32_t> none(static_cast<size_t>(n_), CHOICE_NONE);
    std::vector<int32_t> minus(static_cast<size_t>(n_), -1);
    std::vector<uint32_t> thr(S * THRESHOLD_STRIDE, uint32_t(CHOICE_GAIN_BINS));
    SWARM_HIP_CHECK(hipMemcpyAsync(ch_.assign, none.data(), b, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ch_.pending, 0, b, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ch_.queue, 0, S * QUEUE_STRIDE * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ch_.counters, 0, S * COUNTER_STRIDE * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ch_.hist, 0, S * size_t(exitStride_) * CHOICE_GAIN_BINS * 4, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(ch_.threshold, thr.data(), thr.size() * 4, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ch_.ahead, 0, S * 2u * size_t(exitStride_) * CHOICE_AHEAD_BINS * 4, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(ev_.step, minus.data(), b, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(ev_.exit, minus.data(), b, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ev_.perExit, 0, S * PER_EXIT_STRIDE * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(ctrl_, 0, CTRL_WORDS * 4, stream_));
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipCrowdSim::restoreAgentState(const float* x, const float* y, const float* vx, const float* vy, const float* w) {
    const size_t b = size_t(n_) * 4;
    std::vector<float> spd(static_cast<size_t>(n_));
    for (size_t i = 0; i < size_t(n_); ++i) spd[i] = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
    for (int h = 0; h < 2; ++h) {
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].x, x, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].y, y, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].vx, vx, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].vy, vy, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].w, w, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].spd, spd.data(), b, hipMemcpyHostToDevice, stream_));
    }
    parity_ = 0;
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipCrowdSim::setDrive(const float* dx, const float* dy) {
    const bool on = dx && dy;
    if (on && batched_) throw std::invalid_argument("setDrive: forced directions are a plain-step tool");
    if (on) {
        SWARM_HIP_CHECK(hipMemcpyAsync(driveX_, dx, size_t(n_) * 4, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(driveY_, dy, size_t(n_) * 4, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
    }
    if (on != drive_) dropGraph();
    drive_ = on;
}

void HipCrowdSim::setProjectionIterations(int iterations) {
    if (iterations < 0) throw std::invalid_argument("setProjectionIterations: iterations must be >= 0");
    if (iterations != mp_.projIters) {
        dropGraph();
        adjK_ = 0;   // the last adjoint forward ran another step: its backward would replay this one
    }
    mp_.projIters = iterations;
}

void HipCrowdSim::setPhysicsShape(int block, int grid) {
    if (!(block == 0 || block == 64 || block == 128 || block == 192 || block == 256) || grid < 0)
        throw std::invalid_argument("setPhysicsShape: block 0, 64, 128, 192 or 256; grid >= 0");
    if (block != physBlock_ || grid != physGrid_) dropGraph();
    physBlock_ = block;
    physGrid_ = grid;
}

void HipCrowdSim::setProjectionLanes(int lanes) {
    if (!(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8)) throw std::invalid_argument("setProjectionLanes: 1, 2, 4 or 8");
    if (lanes != projLanes_) dropGraph();
    projLanes_ = lanes;
}

void HipCrowdSim::setCollectLanes(int lanes) {
    if (!(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8)) throw std::invalid_argument("setCollectLanes: 1, 2, 4 or 8");
    if (lanes != collectLanes_) dropGraph();
    collectLanes_ = lanes;
}

void HipCrowdSim::setPredictLanes(int lanes) {
    if (!(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8)) throw std::invalid_argument("setPredictLanes: 1, 2, 4 or 8");
    if (lanes != predictLanes_) dropGraph();
    predictLanes_ = lanes;
}

void HipCrowdSim::setSortMode(int mode) {
    if (mode != 0 && mode != 1) throw std::invalid_argument("setSortMode: 0 (counting) or 1 (bitonic network)");
    if (mode == 1 && (n_ < 2 || (n_ & (n_ - 1)) != 0))
        throw std::invalid_argument("setSortMode: the bitonic network needs a power-of-two agent count");
    if ((mode == 1) != bitonic_) dropGraph();
    bitonic_ = mode == 1;
}

void HipCrowdSim::setCandidateBudget(float moves) {
    if (!(moves >= 0.0f)) throw std::invalid_argument("setCandidateBudget: moves must be >= 0 (infinity allowed)");
    if (moves != listMoves_) {
        dropGraph();
        dropAdjointGraphs();
    }
    listMoves_ = moves;
    recBudgetOff_ = false;
}

float HipCrowdSim::stepListMoves() const {
    // the overlap loss reads the candidate lists after the step: they stay whole (every pair within reach over P moves)
    if (lossOn_) return -1.0f;
    return listMoves_ >= float(mp_.projIters) ? -1.0f : listMoves_;
}

void HipCrowdSim::setPruneEvery(int every) {
    if (every < 0) throw std::invalid_argument("setPruneEvery: every must be >= 0");
    if (every != pruneEvery_) dropGraph();
    pruneEvery_ = every;
}

std::vector<std::pair<void*, size_t>> HipCrowdSim::persistentState() const {
    const size_t b = size_t(n_) * 4, S = size_t(S_);
    std::vector<std::pair<void*, size_t>> v;
    for (const AgentHalf& h : half_)
        for (float* p : { h.x, h.y, h.vx, h.vy, h.spd, h.w }) v.emplace_back(p, b);
    v.emplace_back(ch_.assign, b);
    v.emplace_back(ch_.pending, b);
    v.emplace_back(ch_.queue, S * QUEUE_STRIDE * 4);
    v.emplace_back(ch_.counters, S * COUNTER_STRIDE * 4);
    v.emplace_back(ch_.hist, S * size_t(exitStride_) * CHOICE_GAIN_BINS * 4);
    v.emplace_back(ch_.threshold, S * THRESHOLD_STRIDE * 4);
    v.emplace_back(ch_.ahead, S * 2u * size_t(exitStride_) * CHOICE_AHEAD_BINS * 4);
    v.emplace_back(ev_.step, b);
    v.emplace_back(ev_.exit, b);
    v.emplace_back(ev_.perExit, S * PER_EXIT_STRIDE * 4);
    v.emplace_back(ctrl_, CTRL_WORDS * 4);
    return v;
}

double HipCrowdSim::autotune(int steps) {
    if (steps <= 0) throw std::invalid_argument("autotune: steps must be positive");
    if (venue3d_) return 0.0;   // the seated passes have one launch shape: nothing to tune
    int dev = 0, cus = 0;
    SWARM_HIP_CHECK(hipGetDevice(&dev));
    SWARM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
    // one slot per thread at three block sizes, and persistent grids of 4 and 8 blocks per CU
    const std::pair<int, int> shapes[] = { { 64, 0 }, { 128, 0 }, { 256, 0 }, { 128, 8 * cus }, { 256, 4 * cus }, { 64, 8 * cus } };
    const auto state = persistentState();
    size_t total = 0;
    for (const auto& s : state) total += s.second;
    char* snap = nullptr;
    SWARM_HIP_CHECK(hipMalloc(&snap, total));
    auto copyAll = [&](bool save) {
        size_t off = 0;
        for (const auto& s : state) {
            char* dst = save ? snap + off : static_cast<char*>(s.first);
            const char* src = save ? static_cast<const char*>(s.first) : snap + off;
            SWARM_HIP_CHECK(hipMemcpyAsync(dst, src, s.second, hipMemcpyDeviceToDevice, stream_));
            off += s.second;
        }
    };
    const int parity0 = parity_;
    const bool timing0 = timing_;
    timing_ = false;
    double best = 0.0;
    std::pair<int, int> bestShape = { physBlock_, physGrid_ };
    auto trial = [&]() {
        step(2);   // graph build and warm-up
        const double ms = timeSteps(steps) / double(steps);
        copyAll(false);
        parity_ = parity0;
        return ms;
    };
    // One setting at a time: every option, the one in force included, timed twice in interleaved rounds, each
    // keeping its faster run, so that a drift of the clock between trials does not decide the choice
    auto pick = [&](int count, const std::function<void(int)>& apply) {
        std::vector<double> ms(size_t(count), 0.0);
        for (int round = 0; round < 2; ++round)
            for (int i = 0; i < count; ++i) {
                apply(i);
                const double t = trial();
                if (round == 0 || t < ms[size_t(i)]) ms[size_t(i)] = t;
            }
        int bi = 0;
        for (int i = 1; i < count; ++i)
            if (ms[size_t(i)] < ms[size_t(bi)]) bi = i;
        apply(bi);
        best = ms[size_t(bi)];
        return bi;
    };
    try {
        copyAll(true);
        const int shapeCount = int(sizeof(shapes) / sizeof(shapes[0]));
        bestShape = shapes[pick(shapeCount, [&](int i) { setPhysicsShape(shapes[i].first, shapes[i].second); })];
        // then, at the fastest shape, the projection's lanes per slot, predict's and the collecting iteration's
        const int listLanes[] = { 1, 2, 4, 8 }, fewLanes[] = { 1, 2, 4 };
        pick(4, [&](int i) { setProjectionLanes(listLanes[i]); });
        pick(3, [&](int i) { setPredictLanes(fewLanes[i]); });
        pick(3, [&](int i) { setCollectLanes(fewLanes[i]); });
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(snap);
        timing_ = timing0;
        throw;
    }
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
    SWARM_HIP_CHECK(hipFree(snap));
    timing_ = timing0;
    setPhysicsShape(bestShape.first, bestShape.second);
    return best;
}

void HipCrowdSim::enqueueStep(int parity) {
    const AgentHalf rd = half_[parity], wr = half_[1 - parity];
    StepBuffers b = sb_;
    b.avx = rd.vx;
    b.avy = rd.vy;
    b.listMoves = stepListMoves();
    std::vector<hipEvent_t> ev;
    if (timing_) {   // stamps: start, sorted, choice, predict, projection, final; then a pair per later iteration
        ev.resize(6 + (timeLaunches_ ? 2 * size_t(std::max(0, mp_.projIters - 2)) : 0));
        for (hipEvent_t& e : ev) SWARM_HIP_CHECK(hipEventCreate(&e));
        SWARM_HIP_CHECK(hipEventRecord(ev[0], stream_));
    }
    const StepLaunch L = launchDesc();
    launchGridPasses(stream_, L, rd, grid_, b, rad_, ctrl_);
    if (timing_) SWARM_HIP_CHECK(hipEventRecord(ev[1], stream_));
    launchChoice(stream_, L, rd, ch_, b, drive_ ? driveX_ : nullptr, drive_ ? driveY_ : nullptr, ctrl_);
    if (timing_) SWARM_HIP_CHECK(hipEventRecord(ev[2], stream_));
    if (venue3d_) launchSeatedPhysics(stream_, L, b, seatedSlots_, pa_, pb_, rd, wr, ev_, ctrl_, timing_ ? &ev[3] : nullptr);
    else
        launchPhysics(stream_, L, b, pa_, pb_, rd, wr, ev_, ctrl_, timing_ ? &ev[3] : nullptr,
                      timing_ && timeLaunches_ && mp_.projIters > 2 ? &ev[6] : nullptr);
    if (lossOn_) launchOverlap(stream_, L, b, wr, overlapWeight_, lossAcc_, nullptr);
    if (metricsOn_) {   // the step's input half rd, its output wr; in this step's slot order
        launchMetricsDeposit(stream_, n_, mgrid_, b.sid, rd.x, rd.y, rd.w, wr.x, wr.y, wr.w, wr.vx, wr.vy, scenes_[0].view.E,
                             scenes_[0].view.exitRects, mbuf_);
        launchMetricsFinalize(stream_, mgrid_, mbuf_);
    }
    if (timing_) {
        SWARM_HIP_CHECK(hipEventRecord(ev[5], stream_));
        pendingStamps_.push_back(ev);
        if (pendingStamps_.size() >= 64) collectTimings();
    }
}

void HipCrowdSim::collectTimings() {
    for (std::vector<hipEvent_t>& ev : pendingStamps_) {
        SWARM_HIP_CHECK(hipEventSynchronize(ev[5]));   // the final pass's stamp, recorded last
        float ms[5];
        for (int k = 0; k < 5; ++k) SWARM_HIP_CHECK(hipEventElapsedTime(&ms[k], ev[size_t(k)], ev[size_t(k) + 1]));
        float total = 0.0f;
        SWARM_HIP_CHECK(hipEventElapsedTime(&total, ev[0], ev[5]));
        for (int k = 0; k < 5; ++k) timeSum_[k] += ms[k];   // grid passes, choice, predict, projection, final
        timeSum_[5] += total;
        for (size_t k = 6; k + 1 < ev.size(); k += 2) {       // each later projection iteration, on its own
            float it = 0.0f;
            SWARM_HIP_CHECK(hipEventElapsedTime(&it, ev[k], ev[k + 1]));
            timeSum_[6] += it;
            timeSum_[7] += 1.0;
        }
        ++timedSteps_;
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    pendingStamps_.clear();
}

void HipCrowdSim::setTimingEnabled(bool on, bool launches) {
    if (timing_) collectTimings();
    timing_ = on;
    timeLaunches_ = on && launches;
    if (on) {
        for (double& t : timeSum_) t = 0.0;
        timedSteps_ = 0;
    }
}

int HipCrowdSim::readTimings(double* ms8) {
    collectTimings();
    const double n = timedSteps_ > 0 ? double(timedSteps_) : 1.0;
    for (int k = 0; k < 6; ++k) ms8[k] = timeSum_[k] / n;
    ms8[6] = timeSum_[7] > 0.0 ? timeSum_[6] / timeSum_[7] : 0.0;   // the mean of the later iterations' event pairs
    ms8[7] = double(timedSteps_);
    return timedSteps_;
}

void HipCrowdSim::buildGraph() {
    const int g = lossOn_ ? 1 : 0;
    SWARM_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    enqueueStep(0);
    enqueueStep(1);
    SWARM_HIP_CHECK(hipStreamEndCapture(stream_, &graph_[g]));
    SWARM_HIP_CHECK(hipGraphInstantiate(&graphExec_[g], graph_[g], nullptr, nullptr, 0));
    if (lossOn_) lossGraphOvw_ = overlapWeight_;
}

void HipCrowdSim::step(int nsteps) {
    if (nsteps < 0) throw std::invalid_argument("step: negative step count");
    int left = nsteps;
    if (opt_.useGraph && !timing_ && left >= 2) {
        if (parity_ == 1) { enqueueStep(1); parity_ = 0; --left; }
        const int g = lossOn_ ? 1 : 0;
        if (!graphExec_[g]) buildGraph();
        while (left >= 2) {
            SWARM_HIP_CHECK(hipGraphLaunch(graphExec_[g], stream_));
            left -= 2;
        }
    }
    while (left > 0) {
        enqueueStep(parity_);
        parity_ = 1 - parity_;
        --left;
    }
    SWARM_HIP_CHECK(hipGetLastError());
}

double HipCrowdSim::timeSteps(int nsteps) {
    hipEvent_t a, b;
    SWARM_HIP_CHECK(hipEventCreate(&a));
    SWARM_HIP_CHECK(hipEventCreate(&b));
    SWARM_HIP_CHECK(hipEventRecord(a, stream_));
    step(nsteps);
    SWARM_HIP_CHECK(hipEventRecord(b, stream_));
    SWARM_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    SWARM_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms);
}

void HipCrowdSim::stepExitChoice() {
    const AgentHalf rd = half_[parity_];
    StepBuffers b = sb_;
    b.avx = rd.vx;
    b.avy = rd.vy;
    const StepLaunch L = launchDesc();
    launchGridPasses(stream_, L, rd, grid_, b, rad_, ctrl_);
    launchChoice(stream_, L, rd, ch_, b, drive_ ? driveX_ : nullptr, drive_ ? driveY_ : nullptr, ctrl_);
    launchStepIncrement(stream_, ctrl_);
    SWARM_HIP_CHECK(hipGetLastError());
}

// The queue counts on the device: the half the next step reads (the step counter's parity) holds only the agents
// whose exit is not among their cell's slots, and that step's prefix pass adds each exit's queue-ahead histogram
// total. aheadTotals sums those histograms of parity p, per scenario and exit (QUEUE_STRIDE words a scenario).
std::vector<uint32_t> HipCrowdSim::aheadTotals(uint32_t p) {
    const size_t S = size_t(S_), stride = size_t(exitStride_), bins = size_t(CHOICE_AHEAD_BINS);
    std::vector<uint32_t> a(S * 2 * stride * bins);
    SWARM_HIP_CHECK(hipMemcpyAsync(a.data(), ch_.ahead, a.size() * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    std::vector<uint32_t> t(S * QUEUE_STRIDE, 0u);
    for (size_t sc = 0; sc < S; ++sc)
        for (size_t e = 0; e < std::min(stride, size_t(CHOICE_QUEUE_STRIDE)); ++e) {
            const uint32_t* h = a.data() + (sc * 2 + p) * stride * bins + e * bins;
            uint32_t sum = 0;
            for (size_t b = 0; b < bins; ++b) sum += h[b];
            t[sc * QUEUE_STRIDE + p * CHOICE_QUEUE_STRIDE + e] = sum;
        }
    return t;
}

std::vector<uint32_t> HipCrowdSim::queueCounts() {
    uint32_t step = 0;
    readControl(&step, nullptr, nullptr);
    std::vector<uint32_t> q(size_t(S_) * QUEUE_STRIDE);
    SWARM_HIP_CHECK(hipMemcpyAsync(q.data(), ch_.queue, q.size() * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    const std::vector<uint32_t> t = aheadTotals(step & 1u);
    for (size_t k = 0; k < q.size(); ++k) q[k] += t[k];
    return q;
}

void HipCrowdSim::setChoiceState(uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead) {
    std::vector<uint32_t> q = queueCounts();   // the counts as they stand (scenario 0's replaced by queue128)
    if (queue128) std::copy(queue128, queue128 + 128, q.begin());
    SWARM_HIP_CHECK(hipMemcpyAsync(ctrl_ + CTRL_STEP, &step, 4, hipMemcpyHostToDevice, stream_));
    if (assign) SWARM_HIP_CHECK(hipMemcpyAsync(ch_.assign, assign, size_t(n_) * 4, hipMemcpyHostToDevice, stream_));
    if (ahead) {   // the harness's layout: parity x 64 exits x 1024 bins, into scenario 0's exit stride
        for (int par = 0; par < 2; ++par)
            SWARM_HIP_CHECK(hipMemcpyAsync(ch_.ahead + size_t(par) * size_t(exitStride_) * CHOICE_AHEAD_BINS,
                                           ahead + size_t(par) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS,
                                           size_t(std::min(exitStride_, CHOICE_QUEUE_STRIDE)) * CHOICE_AHEAD_BINS * 4, hipMemcpyHostToDevice,
                                           stream_));
    }
    // the device form for the new parity: the next prefix pass adds the histogram totals back (modulo 2^32)
    const std::vector<uint32_t> t = aheadTotals(step & 1u);
    for (size_t k = 0; k < q.size(); ++k) q[k] -= t[k];
    SWARM_HIP_CHECK(hipMemcpyAsync(ch_.queue, q.data(), q.size() * 4, hipMemcpyHostToDevice, stream_));
    sync();
}

void HipCrowdSim::sync() { SWARM_HIP_CHECK(hipStreamSynchronize(stream_)); }

void HipCrowdSim::readState(float* x, float* y, float* vx, float* vy, float* speed, float* w) {
    const AgentHalf h = half_[parity_];
    const size_t b = size_t(n_) * 4;
    if (x) SWARM_HIP_CHECK(hipMemcpyAsync(x, h.x, b, hipMemcpyDeviceToHost, stream_));
    if (y) SWARM_HIP_CHECK(hipMemcpyAsync(y, h.y, b, hipMemcpyDeviceToHost, stream_));
    if (vx) SWARM_HIP_CHECK(hipMemcpyAsync(vx, h.vx, b, hipMemcpyDeviceToHost, stream_));
    if (vy) SWARM_HIP_CHECK(hipMemcpyAsync(vy, h.vy, b, hipMemcpyDeviceToHost, stream_));
    if (speed) SWARM_HIP_CHECK(hipMemcpyAsync(speed, h.spd, b, hipMemcpyDeviceToHost, stream_));
    if (w) SWARM_HIP_CHECK(hipMemcpyAsync(w, h.w, b, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::upload3(const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                          const float* w) {
    if (!venue3d_) throw std::invalid_argument("upload3: the scene is not a seated venue's own 3D step");
    const size_t n = size_t(n_), b = n * 4;
    std::vector<float> r(n, BODY_RADIUS), v0(n, V_MAX), spd(n);
    upload(x, y, vx, vy, w, r.data(), v0.data());
    for (size_t i = 0; i < n; ++i) spd[i] = std::sqrt((vx[i] * vx[i] + vy[i] * vy[i]) + vz[i] * vz[i]);
    for (int h = 0; h < 2; ++h) {
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].z, z, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].vz, vz, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(half_[h].spd, spd.data(), b, hipMemcpyHostToDevice, stream_));
    }
    sync();
}

void HipCrowdSim::readState3(float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w) {
    readState(x, y, vx, vy, speed, w);
    const AgentHalf h = half_[parity_];
    const size_t b = size_t(n_) * 4;
    if (z) SWARM_HIP_CHECK(hipMemcpyAsync(z, h.z, b, hipMemcpyDeviceToHost, stream_));
    if (vz) SWARM_HIP_CHECK(hipMemcpyAsync(vz, h.vz, b, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readPrevState(float* x, float* y, float* vx, float* vy, float* speed, float* w) {
    const AgentHalf h = half_[1 - parity_];
    const size_t b = size_t(n_) * 4;
    if (x) SWARM_HIP_CHECK(hipMemcpyAsync(x, h.x, b, hipMemcpyDeviceToHost, stream_));
    if (y) SWARM_HIP_CHECK(hipMemcpyAsync(y, h.y, b, hipMemcpyDeviceToHost, stream_));
    if (vx) SWARM_HIP_CHECK(hipMemcpyAsync(vx, h.vx, b, hipMemcpyDeviceToHost, stream_));
    if (vy) SWARM_HIP_CHECK(hipMemcpyAsync(vy, h.vy, b, hipMemcpyDeviceToHost, stream_));
    if (speed) SWARM_HIP_CHECK(hipMemcpyAsync(speed, h.spd, b, hipMemcpyDeviceToHost, stream_));
    if (w) SWARM_HIP_CHECK(hipMemcpyAsync(w, h.w, b, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::sortGrid(const uint32_t* keys, uint32_t* sortedKeys, uint32_t* ids) {
    if (!keys) throw std::invalid_argument("sortGrid: null keys");
    for (size_t i = 0; i < size_t(n_); ++i)
        if (keys[i] != INACTIVE_CELL && keys[i] >= uint32_t(C_)) throw std::invalid_argument("sortGrid: a key is neither a cell nor inactive");
    SWARM_HIP_CHECK(hipMemcpyAsync(grid_.key, keys, size_t(n_) * 4, hipMemcpyHostToDevice, stream_));
    StepBuffers b = sb_;
    // the counting sort ranks each agent within its cell by a scan of the cell's slots: a step's cells hold a few
    // agents (hard discs), caller keys may put any number in one; beyond FIX_RANK_MAX the stable order comes from
    // the network with whole-pair compares instead (the same order, O(n log^2 n))
    std::vector<uint32_t> cnt(size_t(C_) + 1, 0u);
    uint32_t most = 0;
    for (size_t i = 0; i < size_t(n_); ++i)
        if (keys[i] != INACTIVE_CELL) most = std::max(most, ++cnt[keys[i]]);
    if (!bitonic_ && most > FIX_RANK_MAX) {
        const uint32_t npow = keySortPow2(uint32_t(n_));
        uint64_t* pairs = grid_.pairs;
        void* tmp = nullptr;
        if (npow != uint32_t(n_)) {
            SWARM_HIP_CHECK(hipMalloc(&tmp, size_t(npow) * 8));
            pairs = static_cast<uint64_t*>(tmp);
        }
        const KeySortScratch sc{ grid_.loc, grid_.bcnt, grid_.bstart, grid_.tileSum, pairs, ctrl_ };
        launchStableKeySort(stream_, uint32_t(n_), uint32_t(C_), grid_.key, sc, nullptr, nullptr, nullptr, nullptr);
        launchSortedGather(stream_, launchDesc(), half_[parity_], pairs, b, rad_);
        const hipError_t e = hipGetLastError();
        if (tmp) {
            (void)hipStreamSynchronize(stream_);
            (void)hipFree(tmp);
        }
        SWARM_HIP_CHECK(e);
    } else {
        launchSortKeys(stream_, launchDesc(), half_[parity_], grid_, b, rad_, ctrl_);
        SWARM_HIP_CHECK(hipGetLastError());
    }
    readGrid(sortedKeys, ids);
}

void HipCrowdSim::navProbe(int m, const float* x, const float* y, const float* r, const uint32_t* ex, int32_t* cell, float* dir, float* goal) {
    if (m <= 0) return;
    if (!x || !y || !r || !ex || !cell || !dir || !goal) throw std::invalid_argument("navProbe: null argument");
    const SceneView& v = scenes_[0].view;
    for (int i = 0; i < m; ++i)
        if (v.E > 0 && ex[i] >= uint32_t(v.E) && ex[i] != CHOICE_NONE) throw std::invalid_argument("navProbe: no such exit");
    const size_t b = size_t(m) * 4;
    char* d = nullptr;
    SWARM_HIP_CHECK(hipMalloc(&d, 9 * b));
    try {
        float* dx = reinterpret_cast<float*>(d);
        float* dy = dx + m;
        float* dr = dy + m;
        uint32_t* de = reinterpret_cast<uint32_t*>(dr + m);
        int32_t* dc = reinterpret_cast<int32_t*>(de + m);
        float* ddir = reinterpret_cast<float*>(dc + m);
        float* dgoal = ddir + 2 * size_t(m);
        SWARM_HIP_CHECK(hipMemcpyAsync(dx, x, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(dy, y, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(dr, r, b, hipMemcpyHostToDevice, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(de, ex, b, hipMemcpyHostToDevice, stream_));
        launchNavProbe(stream_, v, m, dx, dy, dr, de, dc, ddir, dgoal);
        SWARM_HIP_CHECK(hipGetLastError());
        SWARM_HIP_CHECK(hipMemcpyAsync(cell, dc, b, hipMemcpyDeviceToHost, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(dir, ddir, 2 * b, hipMemcpyDeviceToHost, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(goal, dgoal, 2 * b, hipMemcpyDeviceToHost, stream_));
        sync();
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(d);
        throw;
    }
    SWARM_HIP_CHECK(hipFree(d));
}

void HipCrowdSim::readCandidates(uint32_t* count, uint32_t* index, float* dist) {
    const size_t n = size_t(n_);
    if (count) SWARM_HIP_CHECK(hipMemcpyAsync(count, sb_.nbrCount, n * 4, hipMemcpyDeviceToHost, stream_));
    if (index) SWARM_HIP_CHECK(hipMemcpyAsync(index, sb_.nbr, size_t(NBR_CAP) * n * 4, hipMemcpyDeviceToHost, stream_));
    if (dist && venue3d_) SWARM_HIP_CHECK(hipMemcpyAsync(dist, sb_.nbrD, size_t(NBR_CAP) * n * 4, hipMemcpyDeviceToHost, stream_));
    // the plain step's lists hold indices only: a pair's contact distance is the radii's sum (slot order), as the
    // collecting scan forms it
    std::vector<uint32_t> cnt, idx;
    std::vector<float> rad;
    if (dist && !venue3d_) {
        cnt.resize(n); idx.resize(size_t(NBR_CAP) * n); rad.resize(n);
        SWARM_HIP_CHECK(hipMemcpyAsync(cnt.data(), sb_.nbrCount, n * 4, hipMemcpyDeviceToHost, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(idx.data(), sb_.nbr, size_t(NBR_CAP) * n * 4, hipMemcpyDeviceToHost, stream_));
        SWARM_HIP_CHECK(hipMemcpyAsync(rad.data(), sb_.sr, n * 4, hipMemcpyDeviceToHost, stream_));
    }
    sync();
    if (dist && !venue3d_) {
        std::fill(dist, dist + size_t(NBR_CAP) * n, 0.0f);
        for (size_t s = 0; s < n; ++s) {
            const uint32_t c = cnt[s] == NBR_OVERFLOW ? 0u : std::min(cnt[s], uint32_t(NBR_CAP));
            for (uint32_t j = 0; j < c; ++j) {
                const uint32_t k = idx[size_t(j) * n + s];
                if (k < n) dist[size_t(j) * n + s] = rad[s] + rad[k];
            }
        }
    }
}

void HipCrowdSim::readGrid(uint32_t* key, uint32_t* id) {
    const size_t b = size_t(n_) * 4;
    if (key) SWARM_HIP_CHECK(hipMemcpyAsync(key, sb_.skey, b, hipMemcpyDeviceToHost, stream_));
    if (id) SWARM_HIP_CHECK(hipMemcpyAsync(id, sb_.sid, b, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readCells(uint32_t* start, uint32_t* count) {
    // the reference's cell table (start 0 for an empty cell) from the bucket starts: cell c holds [bstart[c], bstart[c + 1])
    std::vector<uint32_t> bs(static_cast<size_t>(C_) + 1);
    SWARM_HIP_CHECK(hipMemcpyAsync(bs.data(), grid_.bstart, bs.size() * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    for (size_t c = 0; c < size_t(C_); ++c) {
        const uint32_t cnt = bs[c + 1] - bs[c];
        if (start) start[c] = cnt ? bs[c] : 0u;
        if (count) count[c] = cnt;
    }
}

void HipCrowdSim::readAssign(uint32_t* assign) {
    SWARM_HIP_CHECK(hipMemcpyAsync(assign, ch_.assign, size_t(n_) * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readChoiceCounters(uint32_t* out) {
    const std::vector<uint32_t> q = queueCounts();
    std::copy(q.begin(), q.begin() + 128, out);
    SWARM_HIP_CHECK(hipMemcpyAsync(out + 128, ch_.counters, 4 * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readChoiceCountersAll(uint32_t* out) {
    const std::vector<uint32_t> q = queueCounts();
    std::vector<uint32_t> c(size_t(S_) * COUNTER_STRIDE);
    SWARM_HIP_CHECK(hipMemcpyAsync(c.data(), ch_.counters, c.size() * 4, hipMemcpyDeviceToHost, stream_));
    sync();
    for (size_t s = 0; s < size_t(S_); ++s) {
        std::memcpy(out + s * 132, q.data() + s * QUEUE_STRIDE, 128 * 4);
        std::memcpy(out + s * 132 + 128, c.data() + s * COUNTER_STRIDE, 4 * 4);
    }
}

void HipCrowdSim::readEvac(int32_t* step, int32_t* exitId, uint32_t* perExit) {
    const size_t b = size_t(n_) * 4;
    if (step) SWARM_HIP_CHECK(hipMemcpyAsync(step, ev_.step, b, hipMemcpyDeviceToHost, stream_));
    if (exitId) SWARM_HIP_CHECK(hipMemcpyAsync(exitId, ev_.exit, b, hipMemcpyDeviceToHost, stream_));
    if (perExit) SWARM_HIP_CHECK(hipMemcpyAsync(perExit, ev_.perExit, size_t(scenes_[0].view.E + 1) * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readEvacCounters(uint32_t* perExit64) {
    SWARM_HIP_CHECK(hipMemcpyAsync(perExit64, ev_.perExit, size_t(S_) * PER_EXIT_STRIDE * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readControl(uint32_t* step, uint32_t* nActive, uint64_t* activeSum) {
    uint32_t c[CTRL_WORDS];
    SWARM_HIP_CHECK(hipMemcpyAsync(c, ctrl_, sizeof c, hipMemcpyDeviceToHost, stream_));
    sync();
    if (step) *step = c[CTRL_STEP];
    if (nActive) *nActive = c[CTRL_NACTIVE];
    if (activeSum) *activeSum = uint64_t(c[CTRL_ACTIVE_SUM_LO]) | (uint64_t(c[CTRL_ACTIVE_SUM_LO + 1]) << 32);
}

} // namespace swarm

// ---------------------------------------------------------------------------------------------------------------
// The adjoint of K steps. The forward runs the plain step (graph replays) with the loss terms and keeps the
// persistent state every 'every' steps; the backward replays each segment from its checkpoint with the taped
// kernels (the same arithmetic: the same bits) into a segment's tape and runs the adjoint kernels over it backwards.
namespace swarm {

void HipCrowdSim::setModel(float tauS, float obstacleGain) {
    if (!(tauS > 0.0f)) throw std::invalid_argument("setModel: the relaxation time must be positive");
    opt_.tauS = tauS;
    opt_.obstacleGain = obstacleGain;
    float tau = 0.0f, relax = 0.0f;
    modelParamsFor(tauS, tau, relax);
    if (tau != mp_.tau || relax != mp_.relaxStep || obstacleGain != mp_.obstacleGain) adjK_ = 0;   // another model than the last forward's
    mp_.tau = tau;
    mp_.relaxStep = relax;
    mp_.obstacleGain = obstacleGain;
    dropGraph();
}

void HipCrowdSim::snapshotTo(char* dst) {
    size_t off = 0;
    for (const auto& s : persistentState()) {
        SWARM_HIP_CHECK(hipMemcpyAsync(dst + off, s.first, s.second, hipMemcpyDeviceToDevice, stream_));
        off += s.second;
    }
}

void HipCrowdSim::restoreFrom(const char* src, hipStream_t st) {
    size_t off = 0;
    for (const auto& s : persistentState()) {
        SWARM_HIP_CHECK(hipMemcpyAsync(s.first, src + off, s.second, hipMemcpyDeviceToDevice, st));
        off += s.second;
    }
}

void HipCrowdSim::allocAdjoint(int every, bool dual) {
    auto dalloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 16));
        adjBufs_.push_back(p);
        adjBytes_ += bytes;
        return p;
    };
    const size_t n = size_t(n_);
    if (!lossAcc_) {
        lossAcc_ = static_cast<double*>(dalloc(8 * size_t(S_)));
        exitRec_ = static_cast<swarm_exit_rec*>(dalloc(n * sizeof(swarm_exit_rec)));
        finX_ = static_cast<float*>(dalloc(n * 4));
        finY_ = static_cast<float*>(dalloc(n * 4));
        float** agent[] = { &adj_.gx, &adj_.gy, &adj_.gvx, &adj_.gvy, &adj_.hx, &adj_.hy, &adj_.hvx, &adj_.hvy, &adj_.gv0, &adj_.gR };
        for (float** f : agent) *f = static_cast<float*>(dalloc(n * 4));
        float2** slot2[] = { &adj_.gq, &adj_.gq2, &adj_.w[0], &adj_.w[1], &adj_.gself[0], &adj_.gself[1], &adj_.gpv, &adj_.gsep, &adj_.gp, &adj_.gpo };
        for (float2** f : slot2) *f = static_cast<float2*>(dalloc(n * 8));
        float** slot1[] = { &adj_.gpc, &adj_.gr, &adj_.gs };
        for (float** f : slot1) *f = static_cast<float*>(dalloc(n * 4));
        adj_.flags = static_cast<uint32_t*>(dalloc(n * 4));
        adj_.lead = static_cast<uint32_t*>(dalloc(n * 4));
        adj_.gtau = static_cast<float*>(dalloc(4 * size_t(S_)));
        adj_.ggain = static_cast<float*>(dalloc(4 * size_t(S_)));
        tape_[0].acc0 = static_cast<float*>(dalloc(n * 4));
        tape_[0].acc1 = static_cast<float*>(dalloc(n * 4));
        tape_[0].err = static_cast<uint32_t*>(dalloc(16));
        stateBase_ = static_cast<int*>(dalloc(16));
        adjG0_ = adj_.gx;
    }
    // the mask and position arrays hold P iterations per step: a tape made for fewer iterations is made again
    const int Pnow = mp_.projIters;
    const bool haveDual = tape_[1].inx != nullptr;
    if (every <= adjCap_ && Pnow <= tapeP_ && (haveDual || !dual)) return;
    every = std::max(every, adjCap_);
    dropAdjointGraphs();   // they hold the old tape's addresses
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
    if (stream2_) SWARM_HIP_CHECK(hipStreamSynchronize(stream2_));
    for (void* p : tapeBufs_) (void)hipFree(p);   // the smaller tape
    tapeBufs_.clear();
    adjBytes_ -= tapeBytes_;
    tapeBytes_ = 0;
    auto talloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 16));
        tapeBufs_.push_back(p);
        tapeBytes_ += bytes;
        adjBytes_ += bytes;
        return p;
    };
    // a segment's tape: 'every' steps (two of them for the pipelined backward)
    const size_t C = size_t(every), P = size_t(Pnow), nbs = size_t(nb_) + 1;
    const uint32_t sentinel = uint32_t(n_);
    for (int tb = 0; tb < 2; ++tb) {
        Tape& T = tape_[tb];
        float* acc0 = tape_[0].acc0;
        float* acc1 = tape_[0].acc1;
        uint32_t* err = tape_[0].err;
        T = Tape{};
        T.acc0 = acc0; T.acc1 = acc1; T.err = err;   // the replay's scratch and error word are shared
        if (tb == 1 && !dual) continue;
        float** f1[] = { &T.inx, &T.iny, &T.invx, &T.invy, &T.inw, &T.outx, &T.outy, &T.sx, &T.sy,
                         &T.sr, &T.egx, &T.egy, &T.pw, &T.pvx, &T.pvy };
        for (float** f : f1) *f = static_cast<float*>(talloc(C * n * 4));
        T.sid = static_cast<uint32_t*>(talloc(C * n * 4));
        T.skey = static_cast<uint32_t*>(talloc(C * n * 4));
        T.bstart = static_cast<uint32_t*>(talloc(C * nbs * 4));
        T.planes = static_cast<float4*>(talloc(C * n * 16));
        T.nbr = static_cast<uint32_t*>(talloc(C * size_t(NBR_CAP) * n * 4));
        T.nbrCount = static_cast<uint32_t*>(talloc(C * n * 4));
        T.pred = static_cast<swarm_predict_rec*>(talloc(C * n * sizeof(swarm_predict_rec)));
        T.fin = static_cast<swarm_final_rec*>(talloc(C * n * sizeof(swarm_final_rec)));
        T.mask = static_cast<uint64_t*>(talloc(C * P * n * 8));
        T.capBits = static_cast<uint64_t*>(talloc(C * n * 8));
        T.planeBits = static_cast<uint64_t*>(talloc(C * n * 8));
        T.overlap = static_cast<uint64_t*>(talloc(C * n * 8));
        T.q = static_cast<float2*>(talloc(C * (P + 1) * n * 8));
        for (size_t k = 0; k < C; ++k)
            SWARM_HIP_CHECK(hipMemcpyAsync(T.bstart + k * nbs + size_t(nb_), &sentinel, 4, hipMemcpyHostToDevice, stream_));
    }
    SWARM_HIP_CHECK(hipStreamSynchronize(stream_));
    tapedSeg_[0] = tapedSeg_[1] = -1;
    adjCap_ = every;
    tapeP_ = Pnow;
}

size_t HipCrowdSim::tapeBytesPerStep() const {   // allocAdjoint's tape, per step
    const size_t n = size_t(n_), P = size_t(mp_.projIters), nbs = size_t(nb_) + 1;
    return n * (15 * 4 + 2 * 4 + 16 + size_t(NBR_CAP) * 4 + 4 + sizeof(swarm_predict_rec) + sizeof(swarm_final_rec) + P * 8 + 3 * 8 +
                (P + 1) * 8) + nbs * 4;
}

// A taped step whose candidate lists broke their budget (the tape's contact masks index the lists) makes the call run
// again, from its start, with whole lists: rare, and the results are the same either way.
double HipCrowdSim::adjointForward(int K, int every, float radiusScale, float overlapWeight) {
    try {
        return adjointForwardRun(K, every, radiusScale, overlapWeight);
    } catch (const TapeBudgetBroken&) {
        restoreFrom(ckpt_[0], stream_);   // the run's start state
        parity_ = ckptParity_[0];
        recBudgetOff_ = true;
        dropAdjointGraphs();
        return adjointForwardRun(K, every, radiusScale, overlapWeight);
    }
}

double HipCrowdSim::adjointForwardRun(int K, int every, float radiusScale, float overlapWeight) {
    if (K <= 0 || every <= 0) throw std::invalid_argument("adjointForward: K and the checkpoint interval must be positive");
    if (mp_.projIters < 2 || mp_.projIters > SWARM_TAPE_MAX_ITERS)
        throw std::invalid_argument("adjointForward: the adjoint needs 2 to 32 projection iterations");
    if (r0_.size() != size_t(n_)) throw std::invalid_argument("adjointForward: upload a state first");
    if (!(radiusScale > 0.0f)) throw std::invalid_argument("adjointForward: the radius scale must be positive");
    if (scenes_[0].view.venue)
        throw std::invalid_argument("adjointForward: a venue plan's incline rules have no adjoint (a rectangle world's step only)");
    // the segment length: the whole run when its tape fits the budget (no recompute), else the checkpoint interval;
    // with several segments, a second segment tape when the two fit ADJ_DUAL_LIMIT (the pipelined backward)
    const int segLen = size_t(K) * tapeBytesPerStep() <= tapeBudget_ ? K : std::min(every, K);
    const int segs = (K + segLen - 1) / segLen;
    const bool dual = segs > 1 && 2 * size_t(segLen) * tapeBytesPerStep() <= ADJ_DUAL_LIMIT;
    allocAdjoint(segLen, dual);
    // this run's radii
    std::vector<float> r(r0_.size());
    for (size_t i = 0; i < r.size(); ++i) r[i] = radiusScale * r0_[i];
    SWARM_HIP_CHECK(hipMemcpyAsync(rad_, r.data(), r.size() * 4, hipMemcpyHostToDevice, stream_));
    if (segLen != adjEvery_ || overlapWeight != overlapWeight_) dropAdjointGraphs();
    adjK_ = K;
    adjEvery_ = segLen;
    adjSegs_ = segs;
    dual_ = dual;
    adjScale_ = radiusScale;
    overlapWeight_ = overlapWeight;
    tapedSeg_[0] = tapedSeg_[1] = -1;
    snapBytes_ = 0;
    for (const auto& s : persistentState()) snapBytes_ += s.second;
    if (!finalSnap_) {
        SWARM_HIP_CHECK(hipMalloc(&finalSnap_, snapBytes_));
        adjBytes_ += snapBytes_;
    }
    while (int(ckpt_.size()) < adjSegs_) {
        char* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, snapBytes_));
        ckpt_.push_back(p);
        adjBytes_ += snapBytes_;
    }
    ckptParity_.assign(size_t(adjSegs_), 0);
    SWARM_HIP_CHECK(hipMemsetAsync(lossAcc_, 0, 8 * size_t(S_), stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(tape_[0].err, 0, 4, stream_));
    lossOn_ = true;   // graph_[1]: the step with the loss terms, for this overlap weight
    if (graphExec_[1] && lossGraphOvw_ != overlapWeight_) {
        (void)hipGraphExecDestroy(graphExec_[1]); graphExec_[1] = nullptr;
        (void)hipGraphDestroy(graph_[1]); graph_[1] = nullptr;
    }
    for (int seg = 0; seg < adjSegs_; ++seg) {
        snapshotTo(ckpt_[size_t(seg)]);
        ckptParity_[size_t(seg)] = parity_;
        // the last segment runs taped: the backward starts there and needs no replay of it (a run of one segment,
        // every >= K, is taped whole and never replayed)
        if (seg == adjSegs_ - 1) recordSegment(seg, tapeBuf(seg), stream_);
        else step(std::min(adjEvery_, K - seg * adjEvery_));
    }
    lossOn_ = false;
    dropGraph();
    launchExitLoss(stream_, launchDesc(), half_[parity_], ch_.assign, exitRec_, lossAcc_);
    SWARM_HIP_CHECK(hipGetLastError());
    SWARM_HIP_CHECK(hipMemcpyAsync(finX_, half_[parity_].x, size_t(n_) * 4, hipMemcpyDeviceToDevice, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(finY_, half_[parity_].y, size_t(n_) * 4, hipMemcpyDeviceToDevice, stream_));
    snapshotTo(finalSnap_);
    finalParity_ = parity_;
    double loss = 0.0;
    SWARM_HIP_CHECK(hipMemcpyAsync(&loss, lossAcc_, 8, hipMemcpyDeviceToHost, stream_));
    checkTapeErrors();   // synchronises
    return loss;
}

// the forward's final state again, after a segment's replay from its checkpoint moved the state back
void HipCrowdSim::restoreFinal() {
    restoreFrom(finalSnap_, stream_);
    parity_ = finalParity_;
}

// the replays' error word (synchronises the context's stream)
void HipCrowdSim::checkTapeErrors() {
    uint32_t err = 0;
    SWARM_HIP_CHECK(hipMemcpyAsync(&err, tape_[0].err, 4, hipMemcpyDeviceToHost, stream_));
    sync();
    if (err & 1u) throw std::runtime_error("adjoint: a slot has more than 48 contact candidates (unsupported by the adjoint)");
    if (err & 2u) throw std::runtime_error("adjoint: more wall pushes in one resolveStatic than the tape holds");
    if (err & 4u) throw TapeBudgetBroken();
}

// the taped forward's candidate budget (StepBuffers::listMoves): the context's, unless a taped step broke it once
float HipCrowdSim::recListMoves() const {
    return !recBudgetOff_ && listMoves_ > 0.0f && listMoves_ < float(mp_.projIters) ? listMoves_ : -1.0f;
}

// one taped step into tape buffer 'buf' (its step k), on stream st
void HipCrowdSim::tapedStep(int buf, int k, hipStream_t st) {
    const Tape& T = tape_[buf];
    const size_t n = size_t(n_), K = size_t(k), P = size_t(mp_.projIters), nbs = size_t(nb_) + 1;
    const AgentHalf rd = half_[parity_], wr = half_[1 - parity_];
    {   // the step's input state into the tape, one launch
        TapeCopy c{};
        const float* src[5] = { rd.x, rd.y, rd.vx, rd.vy, rd.w };
        float* dst[5] = { T.inx + K * n, T.iny + K * n, T.invx + K * n, T.invy + K * n, T.inw + K * n };
        for (int a = 0; a < 5; ++a) { c.src[a] = src[a]; c.dst[a] = dst[a]; }
        c.count = 5;
        launchTapeCopy(st, n_, c);
    }
    StepBuffers sb = sb_;
    sb.listMoves = recListMoves();
    sb.sid = T.sid + K * n; sb.skey = T.skey + K * n; sb.bstart = T.bstart + K * nbs;
    sb.sx = T.sx + K * n; sb.sy = T.sy + K * n; sb.sr = T.sr + K * n;
    sb.egx = T.egx + K * n; sb.egy = T.egy + K * n;
    sb.pw = T.pw + K * n; sb.pvx = T.pvx + K * n; sb.pvy = T.pvy + K * n;
    sb.planes = T.planes + K * n;
    sb.nbr = T.nbr + K * size_t(NBR_CAP) * n; sb.nbrD = nullptr; sb.nbrCount = T.nbrCount + K * n;
    sb.avx = rd.vx; sb.avy = rd.vy;
    GridBuffers gb = grid_;
    gb.bstart = T.bstart + K * nbs;
    const StepLaunch L = launchDesc();
    launchGridPasses(st, L, rd, gb, sb, rad_, ctrl_);
    launchChoice(st, L, rd, ch_, sb, drive_ ? driveX_ : nullptr, drive_ ? driveY_ : nullptr, ctrl_);
    TapeStep tp{ T.pred + K * n, T.fin + K * n, T.mask + K * P * n, T.capBits + K * n, T.planeBits + K * n,
                 T.overlap + K * n, ch_.assign, (drive_ || (!batched_ && L.s.E == 0)) ? 1 : 0, T.err };
    launchPhysicsRecord(st, L, sb, T.q + K * (P + 1) * n, T.acc0, T.acc1, rd, wr, ev_, ctrl_, tp);
    launchOverlap(st, L, sb, wr, overlapWeight_, lossAcc_ + 0, &tp);
    {   // and its output positions
        TapeCopy c{};
        c.src[0] = wr.x; c.dst[0] = T.outx + K * n;
        c.src[1] = wr.y; c.dst[1] = T.outy + K * n;
        c.count = 2;
        launchTapeCopy(st, n_, c);
    }
    SWARM_HIP_CHECK(hipGetLastError());
    parity_ = 1 - parity_;
}

// segment 'seg' replayed from its checkpoint into its tape buffer, on the context's stream (synchronous)
void HipCrowdSim::tapeSegment(int seg) {
    if (seg < 0 || seg >= adjSegs_) throw std::invalid_argument("adjoint: no such segment");
    restoreFrom(ckpt_[size_t(seg)], stream_);
    parity_ = ckptParity_[size_t(seg)];
    recordSegment(seg, tapeBuf(seg), stream_);
}

// segment 'seg''s taped steps from the current state (its checkpoint, or the forward's own state at its start) into
// tape buffer 'buf', enqueued on stream st; the replays' errors are read by checkTapeErrors
void HipCrowdSim::recordSegment(int seg, int buf, hipStream_t st) {
    const int s0 = seg * adjEvery_, s1 = std::min(adjK_, s0 + adjEvery_);
    const int len = s1 - s0;
    // a full segment of at most ADJ_GRAPH_MAX_STEPS steps: one captured replay per tape buffer and starting parity
    // (longer graphs launch slower than the same kernels launched eagerly)
    if (opt_.useGraph && len == adjEvery_ && len <= ADJ_GRAPH_MAX_STEPS) {
        const int p = parity_;
        if (!replayExec_[buf][p]) {
            SWARM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < len; ++k) tapedStep(buf, k, st);
            SWARM_HIP_CHECK(hipStreamEndCapture(st, &replayGraph_[buf][p]));
            SWARM_HIP_CHECK(hipGraphInstantiate(&replayExec_[buf][p], replayGraph_[buf][p], nullptr, nullptr, 0));
        }
        SWARM_HIP_CHECK(hipGraphLaunch(replayExec_[buf][p], st));
        parity_ = (len & 1) ? 1 - p : p;
    } else {
        for (int k = 0; k < len; ++k) tapedStep(buf, k, st);
    }
    tapedSeg_[buf] = seg;
}

AdjStep HipCrowdSim::adjStepView(int buf, int k) const {
    const Tape& T = tape_[buf];
    const size_t n = size_t(n_), K = size_t(k), P = size_t(mp_.projIters), nbs = size_t(nb_) + 1;
    AdjStep t{};
    t.n = n_; t.P = mp_.projIters;
    t.sid = T.sid + K * n; t.skey = T.skey + K * n; t.bstart = T.bstart + K * nbs;
    t.sx = T.sx + K * n; t.sy = T.sy + K * n; t.sr = T.sr + K * n;
    t.invx = T.invx + K * n; t.invy = T.invy + K * n;
    t.outx = T.outx + K * n; t.outy = T.outy + K * n;
    t.pred = T.pred + K * n; t.fin = T.fin + K * n;
    t.nbr = T.nbr + K * size_t(NBR_CAP) * n; t.cnt = T.nbrCount + K * n;
    t.mask = T.mask + K * P * n; t.capBits = T.capBits + K * n; t.planeBits = T.planeBits + K * n;
    t.overlap = T.overlap + K * n;
    t.q = T.q + K * (P + 1) * n;
    t.planes = T.planes + K * n;
    t.v0 = v0_;
    t.local = local_ ? 1 : 0;
    return t;
}

void HipCrowdSim::adjointBackward(float* gradV0, double* gradGlobals) {
    try {
        adjointBackwardRun(gradV0, gradGlobals);
    } catch (const TapeBudgetBroken&) {   // a segment taped again broke its budget: every segment taped again without
        recBudgetOff_ = true;
        dropAdjointGraphs();
        tapedSeg_[0] = tapedSeg_[1] = -1;
        adjointBackwardRun(gradV0, gradGlobals);
    }
}

void HipCrowdSim::adjointBackwardRun(float* gradV0, double* gradGlobals) {
    if (adjK_ <= 0) throw std::invalid_argument("adjointBackward: run adjointForward first");
    const size_t n = size_t(n_), S = size_t(S_);
    // the exit term's adjoint from the forward's final positions
    if (exitTermOn_) {
        AgentHalf fin = half_[parity_];
        fin.x = finX_;
        fin.y = finY_;
        launchAdjointSeed(stream_, n_, fin, exitRec_, adj_);
    } else {
        launchAdjointZero(stream_, n_, adj_);
    }
    SWARM_HIP_CHECK(hipMemsetAsync(adj_.gv0, 0, n * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(adj_.gR, 0, n * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(adj_.gtau, 0, S * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(adj_.ggain, 0, S * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(tape_[0].err, 0, 4, stream_));
    // Pipelined (two tape buffers): segment s runs backwards on the context's stream while segment s - 1 replays from
    // its checkpoint into the other buffer on stream2_. A replay writes the live state and its own buffer only; the
    // backward reads its buffer, the adjoint's buffers and constants. Events order the buffer reuse: a replay waits
    // for the backward that last read its buffer, a backward for the replay that filled it.
    const bool pipe = dual_ && adjSegs_ > 1;
    bool replayPending[2] = { false, false }, backDone[2] = { false, false };
    if (pipe) {
        if (!stream2_) SWARM_HIP_CHECK(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
        for (int b = 0; b < 2; ++b) {
            if (!evReplay_[b]) SWARM_HIP_CHECK(hipEventCreateWithFlags(&evReplay_[b], hipEventDisableTiming));
            if (!evBack_[b]) SWARM_HIP_CHECK(hipEventCreateWithFlags(&evBack_[b], hipEventDisableTiming));
        }
        if (!evMain_) SWARM_HIP_CHECK(hipEventCreateWithFlags(&evMain_, hipEventDisableTiming));
        SWARM_HIP_CHECK(hipEventRecord(evMain_, stream_));
        SWARM_HIP_CHECK(hipStreamWaitEvent(stream2_, evMain_, 0));
    }
    auto enqueueReplay = [&](int seg) {   // on stream2_
        const int b = tapeBuf(seg);
        if (tapedSeg_[b] == seg) return;
        if (backDone[b]) SWARM_HIP_CHECK(hipStreamWaitEvent(stream2_, evBack_[b], 0));
        restoreFrom(ckpt_[size_t(seg)], stream2_);
        parity_ = ckptParity_[size_t(seg)];
        recordSegment(seg, b, stream2_);
        SWARM_HIP_CHECK(hipEventRecord(evReplay_[b], stream2_));
        replayPending[b] = true;
    };
    for (int seg = adjSegs_ - 1; seg >= 0; --seg) {
        const int buf = tapeBuf(seg);
        if (pipe) {
            enqueueReplay(seg);   // (already there: the forward's last segment, or enqueued one segment earlier)
            if (replayPending[buf]) {
                SWARM_HIP_CHECK(hipStreamWaitEvent(stream_, evReplay_[buf], 0));
                replayPending[buf] = false;
            }
            if (seg > 0) enqueueReplay(seg - 1);   // runs while this segment goes backwards
        } else if (tapedSeg_[buf] != seg) {
            tapeSegment(seg);
        }
        const int len = std::min(adjK_, (seg + 1) * adjEvery_) - seg * adjEvery_;
        // a full segment whose clearing pattern repeats from segment to segment: one captured backward per tape
        // buffer and ping-pong half; the segment's first state comes from the device word stateBase_
        if (opt_.useGraph && len == adjEvery_ && len <= ADJ_GRAPH_MAX_STEPS && (window_ == 0 || adjEvery_ % window_ == 0)) {
            launchAdjointSetBase(stream_, stateBase_, seg * adjEvery_);
            const int h = adj_.gx == adjG0_ ? 0 : 1;
            const AdjBufs before = adj_;
            if (!backExec_[buf][h]) {
                SWARM_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                backwardSegment(-1, len, buf);
                SWARM_HIP_CHECK(hipStreamEndCapture(stream_, &backGraph_[buf][h]));
                SWARM_HIP_CHECK(hipGraphInstantiate(&backExec_[buf][h], backGraph_[buf][h], nullptr, nullptr, 0));
                adj_ = before;
            }
            SWARM_HIP_CHECK(hipGraphLaunch(backExec_[buf][h], stream_));
            if (len & 1) {
                std::swap(adj_.gx, adj_.hx); std::swap(adj_.gy, adj_.hy);
                std::swap(adj_.gvx, adj_.hvx); std::swap(adj_.gvy, adj_.hvy);
            }
        } else {
            backwardSegment(seg, len, buf);
        }
        SWARM_HIP_CHECK(hipGetLastError());
        if (pipe) {
            SWARM_HIP_CHECK(hipEventRecord(evBack_[buf], stream_));
            backDone[buf] = true;
        }
    }
    if (adjSegs_ > 1) restoreFinal();   // a segment's replay restored an earlier state (the last replay ran before this)
    std::vector<float> gR(n), gtau(S), ggain(S);
    SWARM_HIP_CHECK(hipMemcpyAsync(gradV0, adj_.gv0, n * 4, hipMemcpyDeviceToHost, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(gR.data(), adj_.gR, n * 4, hipMemcpyDeviceToHost, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(gtau.data(), adj_.gtau, S * 4, hipMemcpyDeviceToHost, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(ggain.data(), adj_.ggain, S * 4, hipMemcpyDeviceToHost, stream_));
    checkTapeErrors();   // synchronises
    // per scenario: dL/dtau_s (tau = tau_s / T), dL/dscale = sum r0 dL/dr, dL/dgain
    std::vector<double> gs(S, 0.0);
    std::vector<uint32_t> base(S + 1, 0);
    if (batched_) {
        std::vector<uint32_t> ab(S);
        SWARM_HIP_CHECK(hipMemcpy(ab.data(), bt_.agentBase, S * 4, hipMemcpyDeviceToHost));
        for (size_t s = 0; s < S; ++s) base[s] = ab[s];
    }
    base[S] = uint32_t(n);
    for (size_t s = 0; s < S; ++s)
        for (uint32_t i = base[s]; i < base[s + 1]; ++i) gs[s] += double(r0_[i]) * double(gR[i]);
    for (size_t s = 0; s < S; ++s) {
        gradGlobals[3 * s + 0] = double(gtau[s]) / double(TIME_SCALE_S_PER_UNIT);
        gradGlobals[3 * s + 1] = gs[s];
        gradGlobals[3 * s + 2] = double(ggain[s]);
    }
    bool finite = true;
    for (size_t i = 0; i < n && finite; ++i) finite = std::isfinite(gradV0[i]);
    for (size_t i = 0; i < 3 * S && finite; ++i) finite = std::isfinite(gradGlobals[i]);
    if (!finite)
        throw NonFiniteGradient("adjointBackward: the gradient overflowed FP32 (inf or NaN); the pathwise derivative of this "
                                "horizon exceeds the float range (shorten K or use the truncation window)");
}

// the backward of one segment's 'len' steps from tape buffer 'buf'; seg < 0: the segment's first state is read from
// stateBase_ (the captured form), and the clearing pattern is that of a segment starting at a multiple of the window
void HipCrowdSim::backwardSegment(int seg, int len, int buf) {
    const StepLaunch L = launchDesc();
    for (int k = len - 1; k >= 0; --k) {
        const int rel = seg < 0 ? k : seg * adjEvery_ + k;   // the step (relative to the segment when captured)
        if (seedsOn_) {
            if (seg < 0) launchAdjointEventSeedRel(stream_, n_, stateBase_, k + 1, seedStep_, seedIn_, seedOut_, adj_);
            else launchAdjointEventSeed(stream_, n_, rel + 1, seedStep_, seedIn_, seedOut_, adj_);
        }
        launchAdjointStep(stream_, L, adjStepView(buf, k), adj_, overlapWeight_);
        std::swap(adj_.gx, adj_.hx); std::swap(adj_.gy, adj_.hy);
        std::swap(adj_.gvx, adj_.hvx); std::swap(adj_.gvy, adj_.hvy);
        if (window_ > 0 && rel % window_ == 0) launchAdjointZero(stream_, n_, adj_);
    }
}

void HipCrowdSim::adjointTape(int step, const swarm_tape_view& out, swarm_exit_rec* exit) {
    try {
        adjointTapeRun(step, out, exit);
    } catch (const TapeBudgetBroken&) {
        recBudgetOff_ = true;
        dropAdjointGraphs();
        tapedSeg_[0] = tapedSeg_[1] = -1;
        adjointTapeRun(step, out, exit);
    }
}

void HipCrowdSim::adjointTapeRun(int step, const swarm_tape_view& out, swarm_exit_rec* exit) {
    if (adjK_ <= 0 || step < 0 || step >= adjK_) throw std::invalid_argument("adjointTape: no such step of the last forward");
    const int seg = step / adjEvery_;
    const int buf = tapeBuf(seg);
    if (tapedSeg_[buf] != seg) tapeSegment(seg);
    const Tape& T = tape_[buf];
    const size_t n = size_t(n_), K = size_t(step - seg * adjEvery_), P = size_t(mp_.projIters);
    auto get = [&](void* dst, const void* src, size_t bytes) {
        if (dst) SWARM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
    };
    get(out.x, T.inx + K * n, n * 4); get(out.y, T.iny + K * n, n * 4);
    get(out.vx, T.invx + K * n, n * 4); get(out.vy, T.invy + K * n, n * 4); get(out.w, T.inw + K * n, n * 4);
    get(out.sid, T.sid + K * n, n * 4);
    get(out.pred, T.pred + K * n, n * sizeof(swarm_predict_rec));
    get(out.fin, T.fin + K * n, n * sizeof(swarm_final_rec));
    get(out.sx, T.sx + K * n, n * 4); get(out.sy, T.sy + K * n, n * 4); get(out.sr, T.sr + K * n, n * 4);
    get(out.nbr, T.nbr + K * size_t(NBR_CAP) * n, size_t(NBR_CAP) * n * 4);
    get(out.nbr_count, T.nbrCount + K * n, n * 4);
    get(out.mask, T.mask + K * P * n, P * n * 8);
    get(out.cap_bits, T.capBits + K * n, n * 8);
    get(out.plane_bits, T.planeBits + K * n, n * 8);
    get(out.overlap, T.overlap + K * n, n * 8);
    if (exit && step == adjK_ - 1) get(exit, exitRec_, n * sizeof(swarm_exit_rec));
    if (adjSegs_ > 1) restoreFinal();
    checkTapeErrors();   // synchronises
}

} // namespace swarm

namespace swarm {

void HipCrowdSim::setAdjointTerms(const int32_t* step, const float* gin, const float* gout, bool exitTerm, int window, bool local) {
    if (window < 0) throw std::invalid_argument("setAdjointTerms: the window must be >= 0");
    if (window != window_ || local != local_ || (step && gin && gout) != seedsOn_) dropAdjointGraphs();
    exitTermOn_ = exitTerm;
    window_ = window;
    local_ = local;
    seedsOn_ = step && gin && gout;
    if (!seedsOn_) return;
    const size_t n = size_t(n_);
    if (!seedStep_) {
        void* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, n * 4)); adjBufs_.push_back(p); seedStep_ = static_cast<int32_t*>(p);
        SWARM_HIP_CHECK(hipMalloc(&p, n * 8)); adjBufs_.push_back(p); seedIn_ = static_cast<float2*>(p);
        SWARM_HIP_CHECK(hipMalloc(&p, n * 8)); adjBufs_.push_back(p); seedOut_ = static_cast<float2*>(p);
    }
    SWARM_HIP_CHECK(hipMemcpyAsync(seedStep_, step, n * 4, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(seedIn_, gin, n * 8, hipMemcpyHostToDevice, stream_));
    SWARM_HIP_CHECK(hipMemcpyAsync(seedOut_, gout, n * 8, hipMemcpyHostToDevice, stream_));
    sync();
}

void HipCrowdSim::readEvacPoints(int32_t* step, int32_t* exitId, float* inside, float* before) {
    const size_t n = size_t(n_);
    if (step) SWARM_HIP_CHECK(hipMemcpyAsync(step, ev_.step, n * 4, hipMemcpyDeviceToHost, stream_));
    if (exitId) SWARM_HIP_CHECK(hipMemcpyAsync(exitId, ev_.exit, n * 4, hipMemcpyDeviceToHost, stream_));
    if (inside) SWARM_HIP_CHECK(hipMemcpyAsync(inside, ev_.inside, n * 8, hipMemcpyDeviceToHost, stream_));
    if (before) SWARM_HIP_CHECK(hipMemcpyAsync(before, ev_.before, n * 8, hipMemcpyDeviceToHost, stream_));
    sync();
}

// ---- the safety metrics ----
void HipCrowdSim::allocMetrics() {
    for (void* p : mBufs_) (void)hipFree(p);
    mBufs_.clear();
    auto dalloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        SWARM_HIP_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 16));
        mBufs_.push_back(p);
        return p;
    };
    const size_t cells = size_t(mgrid_.resX) * size_t(mgrid_.resY);
    mbuf_.counters = static_cast<uint32_t*>(dalloc(METRIC_COUNTER_WORDS * 4));
    mbuf_.evacStep = static_cast<uint32_t*>(dalloc(size_t(n_) * 4));
    mbuf_.mass = static_cast<uint32_t*>(dalloc(cells * 8));
    mbuf_.cells = static_cast<uint32_t*>(dalloc(cells * METRIC_CELL_WORDS * 4));
    uint32_t* thr = static_cast<uint32_t*>(dalloc(cells * 16));
    SWARM_HIP_CHECK(hipMemcpyAsync(thr, mThr_.data(), cells * 16, hipMemcpyHostToDevice, stream_));
    mbuf_.thr = thr;
    mScratch_ = static_cast<float*>(dalloc(size_t(n_) * 8 * 4));
    resetMetrics();
}

void HipCrowdSim::setMetricsEnabled(bool on) {
    if (on && batched_) throw std::invalid_argument("setMetricsEnabled: the safety metrics run on plain contexts");
    if (on && !mbuf_.counters) {
        if (mgrid_.resX <= 0 || mgrid_.resY <= 0) throw std::invalid_argument("setMetricsEnabled: the scene has no metric grid");
        allocMetrics();
    }
    if (on != metricsOn_) dropGraph();
    metricsOn_ = on;
}

void HipCrowdSim::resetMetrics() {
    if (!mbuf_.counters) return;
    const size_t cells = size_t(mgrid_.resX) * size_t(mgrid_.resY);
    SWARM_HIP_CHECK(hipMemsetAsync(mbuf_.counters, 0, METRIC_COUNTER_WORDS * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(mbuf_.evacStep, 0, size_t(n_) * 4, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(mbuf_.mass, 0, cells * 8, stream_));
    SWARM_HIP_CHECK(hipMemsetAsync(mbuf_.cells, 0, cells * METRIC_CELL_WORDS * 4, stream_));
}

void HipCrowdSim::configureMetrics(int resX, int resY, float cellsPerUnit, float offX, float offY, const uint32_t* thresholds4) {
    if (batched_) throw std::invalid_argument("configureMetrics: the safety metrics run on plain contexts");
    if (resX <= 0 || resY <= 0 || size_t(resX) * size_t(resY) > (size_t(1) << 26) || !(cellsPerUnit > 0.0f) || !thresholds4)
        throw std::invalid_argument("configureMetrics: a grid of 1 .. 2^26 cells, positive cells per unit and thresholds");
    mgrid_.resX = resX; mgrid_.resY = resY;
    mgrid_.cellsPerUnit = cellsPerUnit; mgrid_.offX = offX; mgrid_.offY = offY;
    const size_t cells = size_t(resX) * size_t(resY);
    mThr_.assign(thresholds4, thresholds4 + cells * 4);
    mWalkable_.assign(cells, 0.0f);
    allocMetrics();
    dropGraph();
}

void HipCrowdSim::metricGrid(int* resX, int* resY, float* cellsPerUnit, float* offX, float* offY) const {
    if (resX) *resX = mgrid_.resX;
    if (resY) *resY = mgrid_.resY;
    if (cellsPerUnit) *cellsPerUnit = mgrid_.cellsPerUnit;
    if (offX) *offX = mgrid_.offX;
    if (offY) *offY = mgrid_.offY;
}

void HipCrowdSim::recordMetrics(const float* px, const float* py, const float* pw, const float* cx, const float* cy, const float* cw,
                                const float* cvx, const float* cvy) {
    if (batched_) throw std::invalid_argument("recordMetrics: the safety metrics run on plain contexts");
    if (!mbuf_.counters) allocMetrics();
    const size_t n = size_t(n_);
    const float* src[8] = { px, py, pw, cx, cy, cw, cvx, cvy };
    for (int k = 0; k < 8; ++k) {
        if (!src[k]) throw std::invalid_argument("recordMetrics: null state array");
        SWARM_HIP_CHECK(hipMemcpyAsync(mScratch_ + size_t(k) * n, src[k], n * 4, hipMemcpyHostToDevice, stream_));
    }
    const float* d = mScratch_;
    launchMetricsDeposit(stream_, n_, mgrid_, nullptr, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n, d + 7 * n,
                         scenes_[0].view.E, scenes_[0].view.exitRects, mbuf_);
    launchMetricsFinalize(stream_, mgrid_, mbuf_);
    SWARM_HIP_CHECK(hipGetLastError());
    sync();
}

void HipCrowdSim::readMetricCounters(uint32_t* out) {
    if (!mbuf_.counters) throw std::invalid_argument("readMetricCounters: the metrics are not enabled");
    SWARM_HIP_CHECK(hipMemcpyAsync(out, mbuf_.counters, METRIC_COUNTER_WORDS * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readMetricCells(uint32_t* out) {
    if (!mbuf_.counters) throw std::invalid_argument("readMetricCells: the metrics are not enabled");
    const size_t cells = size_t(mgrid_.resX) * size_t(mgrid_.resY);
    SWARM_HIP_CHECK(hipMemcpyAsync(out, mbuf_.cells, cells * METRIC_CELL_WORDS * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readMetricEvacSteps(uint32_t* out) {
    if (!mbuf_.counters) throw std::invalid_argument("readMetricEvacSteps: the metrics are not enabled");
    SWARM_HIP_CHECK(hipMemcpyAsync(out, mbuf_.evacStep, size_t(n_) * 4, hipMemcpyDeviceToHost, stream_));
    sync();
}

void HipCrowdSim::readMetricCellInfo(float* walkableM2, uint32_t* thresholds4) {
    if (walkableM2) std::copy(mWalkable_.begin(), mWalkable_.end(), walkableM2);
    if (thresholds4) std::copy(mThr_.begin(), mThr_.end(), thresholds4);
}

} // namespace swarm
