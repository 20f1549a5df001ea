// The checksummed cache of built scenes: the reference's navigation-field and choice-table caches
// (src/nav_field.cpp:468-688, src/exit_choice.cpp:335-445) for this build's rectangle worlds.
//
// What is kept as the reference keeps it: a 64-bit key that hashes (FNV-1a) a text of everything the tables depend on
// (format version, the floor's description with its float bits, the radius and speed bounds that size the rings),
// the key in the file name, a file of magic + payload + the payload's FNV-1a checksum, written to a temporary file
// and renamed (a concurrent reader never sees a partial file), and a load that refuses a wrong magic, version, key,
// size or checksum, after which the caller builds the scene again and rewrites the file.
// What differs: one file holds every table of the scene (SDF and its window minimum, the candidate table, the stair
// field, exits, walls and their index, the metric grid), since this build makes them in one pass.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "scene.hpp"
#include "venue.hpp"

namespace swarm {

namespace {

constexpr uint32_t SCENE_CACHE_VERSION = 3;
constexpr char MAGIC[8] = { 'S', 'W', 'S', 'C', 'E', 'N', 'E', '1' };

struct Fnv1a {   // FNV-1a, 64 bits (Fowler, Noll, Vo)
    uint64_t h = 0xcbf29ce484222325ull;
    void bytes(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 0x100000001b3ull;
        }
    }
};

std::string bits(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    char s[16];
    std::snprintf(s, sizeof s, "%08x", b);
    return s;
}

void rects(std::ostringstream& s, const char* tag, const std::vector<Rect>& rs) {
    s << tag << rs.size() << "|";
    for (const Rect& r : rs) s << bits(r.lox) << "," << bits(r.loy) << "," << bits(r.hix) << "," << bits(r.hiy) << "|";
}

class Writer {
public:
    void raw(const void* p, size_t n) { body_.write(static_cast<const char*>(p), std::streamsize(n)); }
    template <class T> void pod(const T& v) { raw(&v, sizeof v); }
    template <class T> void vec(const std::vector<T>& v) {
        const uint64_t n = v.size();
        pod(n);
        raw(v.data(), v.size() * sizeof(T));
    }
    std::string payload() const { return body_.str(); }

private:
    std::ostringstream body_{ std::ios::binary };
};

class Reader {
public:
    bool open(const std::string& file) {
        std::ifstream in(file, std::ios::binary);
        if (!in) return false;
        data_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        if (data_.size() < 16 || std::memcmp(data_.data(), MAGIC, 8) != 0) return false;
        end_ = data_.size() - 8;   // the checksum follows the payload
        pos_ = 8;
        uint64_t sum;
        std::memcpy(&sum, data_.data() + end_, 8);
        Fnv1a h;
        h.bytes(data_.data() + 8, end_ - 8);
        return h.h == sum;
    }
    bool raw(void* p, size_t n) {
        if (n > end_ - pos_) return false;
        std::memcpy(p, data_.data() + pos_, n);
        pos_ += n;
        return true;
    }
    template <class T> bool pod(T& v) { return raw(&v, sizeof v); }
    template <class T> bool vec(std::vector<T>& v, uint64_t limit) {
        uint64_t n = 0;
        if (!pod(n) || n > limit || n * sizeof(T) > end_ - pos_) return false;
        v.resize(size_t(n));
        return raw(v.data(), size_t(n) * sizeof(T));
    }
    bool done() const { return pos_ == end_; }

private:
    std::vector<char> data_;
    size_t pos_ = 0, end_ = 0;
};

void write(Writer& w, const Scene& s) {
    for (float v : { s.xmin, s.ymin, s.xmax, s.ymax, s.hashSpanX, s.hashSpanY }) w.pod(v);
    for (int v : { s.hashNX, s.hashNY, s.percRing, s.projRing, s.sdfNX, s.sdfNY, s.sdfMinWin }) w.pod(v);
    for (float v : { s.sdfOX, s.sdfOY, s.sdfH, s.sdfInvH }) w.pod(v);
    w.vec(s.sdf);
    w.vec(s.sdfMin);
    w.pod(s.nav);
    w.pod(s.K);
    w.pod(s.E);
    w.vec(s.navInfo);
    w.vec(s.navSlot);
    w.vec(s.navDist);
    w.vec(s.invCapS);
    w.pod(s.secondsPerUnit);
    w.pod(s.distQuantum);
    w.vec(s.stair);
    w.vec(s.exits);
    w.vec(s.walls);
    w.vec(s.wallBinStart);
    w.vec(s.wallBinList);
    w.pod(s.maxRadius);
    w.pod(s.maxV0);
    w.pod(s.metricResX);
    w.pod(s.metricResY);
    w.pod(s.metricCellsPerUnit);
    w.pod(s.metricOffX);
    w.pod(s.metricOffY);
    w.vec(s.metricWalkableM2);
    w.pod(s.hasVenue);
    w.pod(s.venue);
    w.pod(s.venue3d);
    w.pod(s.exitDepth);
    w.pod(s.exitMaxZ);
    w.pod(s.domainMax);
}

bool read(Reader& r, Scene& s) {
    constexpr uint64_t LIM = uint64_t(1) << 34;   // no table of a scene holds more entries
    bool ok = true;
    for (float* v : { &s.xmin, &s.ymin, &s.xmax, &s.ymax, &s.hashSpanX, &s.hashSpanY }) ok = ok && r.pod(*v);
    for (int* v : { &s.hashNX, &s.hashNY, &s.percRing, &s.projRing, &s.sdfNX, &s.sdfNY, &s.sdfMinWin }) ok = ok && r.pod(*v);
    for (float* v : { &s.sdfOX, &s.sdfOY, &s.sdfH, &s.sdfInvH }) ok = ok && r.pod(*v);
    ok = ok && r.vec(s.sdf, LIM) && r.vec(s.sdfMin, LIM) && r.pod(s.nav) && r.pod(s.K) && r.pod(s.E) && r.vec(s.navInfo, LIM) &&
         r.vec(s.navSlot, LIM) && r.vec(s.navDist, LIM) && r.vec(s.invCapS, LIM) && r.pod(s.secondsPerUnit) && r.pod(s.distQuantum) &&
         r.vec(s.stair, LIM) && r.vec(s.exits, LIM) && r.vec(s.walls, LIM) && r.vec(s.wallBinStart, LIM) && r.vec(s.wallBinList, LIM) &&
         r.pod(s.maxRadius) && r.pod(s.maxV0) && r.pod(s.metricResX) && r.pod(s.metricResY) && r.pod(s.metricCellsPerUnit) &&
         r.pod(s.metricOffX) && r.pod(s.metricOffY) && r.vec(s.metricWalkableM2, LIM) && r.pod(s.hasVenue) && r.pod(s.venue) && r.pod(s.venue3d) && r.pod(s.exitDepth) &&
         r.pod(s.exitMaxZ) && r.pod(s.domainMax);
    if (!ok || !r.done()) return false;
    // the sizes every reader of the tables relies on
    const size_t sdfN = size_t(s.sdfNX) * size_t(s.sdfNY), navN = s.nav.cells(), hashN = size_t(s.hashNX) * size_t(s.hashNY);
    return s.sdfNX > 1 && s.sdfNY > 1 && s.hashNX > 0 && s.hashNY > 0 && s.K > 0 && s.E >= 0 && s.sdf.size() == sdfN &&
           s.sdfMin.size() == sdfN && s.navInfo.size() == navN && s.stair.size() == navN && s.navSlot.size() == navN * size_t(s.K) &&
           s.navDist.size() == navN * size_t(s.K) && s.invCapS.size() == size_t(s.E) && s.exits.size() == (s.venue3d ? size_t(0) : size_t(s.E)) &&
           s.wallBinStart.size() == hashN + 1 && s.wallBinStart.back() == s.wallBinList.size() &&
           s.metricWalkableM2.size() == size_t(s.metricResX) * size_t(s.metricResY);
}

}  // namespace

uint64_t sceneCacheKey(const SceneDesc& d, float maxRadius, float maxV0) {
    std::ostringstream s;
    s << "swarm-scene|" << SCENE_CACHE_VERSION << "|";
    for (float v : { d.xmin, d.ymin, d.xmax, d.ymax, d.hashCell, d.sdfSpacing, d.sdfClamp, d.navBody, maxRadius, maxV0 }) s << bits(v) << "|";
    s << d.nav.resX << "|" << d.nav.resY << "|" << bits(d.nav.originX) << "|" << bits(d.nav.originY) << "|" << bits(d.nav.cell) << "|"
      << bits(d.nav.invCell) << "|";
    char b[40];
    std::snprintf(b, sizeof b, "%.17g|", d.metresPerUnit);
    s << b;
    rects(s, "W", d.walls);
    rects(s, "E", d.exits);
    rects(s, "S", d.stairs);
    for (float v : d.stairFactor) s << bits(v) << "|";
    s << "w|";
    for (float v : d.exitWidthM) s << bits(v) << "|";
    s << "v|" << (d.hasVenue ? 1 : 0) << "|";
    if (d.hasVenue) {   // the layout's bytes, in hex
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&d.venue);
        char h[3];
        for (size_t i = 0; i < sizeof d.venue; ++i) {
            std::snprintf(h, sizeof h, "%02x", b[i]);
            s << h;
        }
    }
    const std::string text = s.str();
    Fnv1a h;
    h.bytes(text.data(), text.size());
    return h.h;
}

std::string sceneCacheFile(const std::string& dir, uint64_t key) {
    char b[64];
    std::snprintf(b, sizeof b, "scene_%016llx.bin", (unsigned long long)key);
    return (std::filesystem::path(dir) / b).string();
}

bool saveScene(const Scene& s, uint64_t key, const std::string& file) {
    std::error_code ec;
    std::filesystem::create_directories(std::filesystem::path(file).parent_path(), ec);
    Writer w;
    w.pod(SCENE_CACHE_VERSION);
    w.pod(key);
    write(w, s);
    const std::string payload = w.payload();
    Fnv1a h;
    h.bytes(payload.data(), payload.size());
    const std::string tmp = file + ".tmp" + std::to_string(std::chrono::steady_clock::now().time_since_epoch().count());
    {
        std::ofstream o(tmp, std::ios::binary | std::ios::trunc);
        if (!o) return false;
        o.write(MAGIC, 8);
        o.write(payload.data(), std::streamsize(payload.size()));
        o.write(reinterpret_cast<const char*>(&h.h), 8);
        if (!o) return false;
    }
    std::filesystem::rename(tmp, file, ec);
    if (ec) {
        std::filesystem::remove(tmp, ec);
        return false;
    }
    return true;
}

bool loadScene(Scene& s, uint64_t key, const std::string& file) {
    Reader r;
    if (!r.open(file)) return false;
    uint32_t version = 0;
    uint64_t k = 0;
    if (!r.pod(version) || !r.pod(k) || version != SCENE_CACHE_VERSION || k != key) return false;
    Scene out;
    if (!read(r, out)) return false;
    s = std::move(out);
    return true;
}

Scene buildSeatedSceneCached(const Venue& v, int threads, const std::string& dir, int* fromCache) {
    if (fromCache) *fromCache = 0;
    if (dir.empty()) return buildSeatedScene(v, threads);
    const auto t0 = std::chrono::steady_clock::now();
    std::ostringstream t;   // everything the tables depend on: the layout's bytes, the domain, the grid
    t << "swarm-seated-scene|" << SCENE_CACHE_VERSION << "|" << v.gridRes << "|" << bits(v.domainMin) << "|" << bits(v.domainSpan) << "|"
      << bits(v.exitDepth) << "|" << bits(v.exitMaxZ) << "|";
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&v.layout);
    char h[3];
    for (size_t i = 0; i < sizeof v.layout; ++i) {
        std::snprintf(h, sizeof h, "%02x", b[i]);
        t << h;
    }
    for (int e = 0; e < seated::exitCount(v.layout); ++e) {
        char w[40];
        std::snprintf(w, sizeof w, "|%.17g", venueExitWidthM(v, e));
        t << w;
    }
    const std::string text = t.str();
    Fnv1a f;
    f.bytes(text.data(), text.size());
    const uint64_t key = f.h;
    const std::string file = sceneCacheFile(dir, key);
    Scene s;
    if (loadScene(s, key, file)) {
        s.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s.fromCache = 1;
        if (fromCache) *fromCache = 1;
        return s;
    }
    s = buildSeatedScene(v, threads);
    (void)saveScene(s, key, file);
    return s;
}

Scene buildSceneCached(const SceneDesc& d, float maxRadius, float maxV0, int threads, const std::string& dir, int* fromCache) {
    if (fromCache) *fromCache = 0;
    if (dir.empty()) return buildScene(d, maxRadius, maxV0, threads);
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t key = sceneCacheKey(d, maxRadius, maxV0);
    const std::string file = sceneCacheFile(dir, key);
    Scene s;
    if (loadScene(s, key, file)) {
        s.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s.fromCache = 1;
        if (fromCache) *fromCache = 1;
        return s;
    }
    s = buildScene(d, maxRadius, maxV0, threads);
    (void)saveScene(s, key, file);   // a cache that cannot be written is no error: the next run builds again
    return s;
}

}  // namespace swarm
