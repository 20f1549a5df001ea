// Host check of the seated venues' cell table (the cell source of the GPU step) against the generator (the cell source
// of the host queries): every query of the heightfield through seated::CellTable -- the bucketed lateral search and
// the windowed wall and distance scans -- must return the generator path's bits at every point. Points: uniform over
// the domain at several heights, on and next to every cell and band boundary (ties of the nearest-cell searches), and
// non-finite ones; the boundary points also through a table whose buckets all start at their band's last cell (the
// search's walk back); and the band lookup's row estimate (a product with the reciprocal) against the reference's
// division at every row front. Prints "ok <points>" or the first difference. Built and run by test_venue_table.py (no GPU needed).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "venue.hpp"

using namespace swarm;
using seated::F2;
using seated::F3;
using seated::F4;

static long g_checked = 0;

template <class T> static bool same(const T& a, const T& b) { return std::memcmp(&a, &b, sizeof(T)) == 0; }

static bool fail(const char* venue, const char* what, F3 p, const char* extra = "") {
    std::printf("DIFF venue %s: %s at (%a, %a, %a) %s\n", venue, what, p.x, p.y, p.z, extra);
    return false;
}

static bool checkPoint(const char* nm, const seated::Layout& L, const seated::CellTable& tb, F3 p, F2 step, F2 e) {
    const seated::GenCells gen{ L };
    const F2 xy{ p.x, p.y };
    ++g_checked;
    if (!same(seated::cellKeyC(L, tb, xy), seated::cellKeyC(L, gen, xy))) return fail(nm, "cellKey", p);
    if (!same(seated::floorHeightC(L, tb, xy), seated::floorHeightC(L, gen, xy))) return fail(nm, "floorHeight", p);
    int r0 = 0, b0 = 0, r1 = 0, b1 = 0;
    if (seated::kindAtC(L, tb, xy, &r0, &b0) != seated::kindAtC(L, gen, xy, &r1, &b1) || r0 != r1 || b0 != b1) return fail(nm, "kindAt", p);
    if (!same(seated::sdfGradC(L, tb, p), seated::sdfGradC(L, gen, p))) return fail(nm, "sdfGrad", p);
    const float reaches[4] = { 0x1.0624dep-7f, seated::STEP_UP_SAMPLE, 0x1.0624dep-6f, 0.5f };
    for (float r : reaches)
        if (!same(seated::wallC(L, tb, p, r), seated::wallC(L, gen, p, r))) return fail(nm, "wall", p, std::to_string(r).c_str());
    if (!same(seated::floorSlopeC(L, tb, xy, e), seated::floorSlopeC(L, gen, xy, e))) return fail(nm, "floorSlope", p);
    if (seated::gangwayAtC(L, tb, xy) != seated::gangwayAtC(L, gen, xy)) return fail(nm, "gangwayAt", p);
    const F2 q{ p.x + step.x, p.y + step.y };
    const float f0 = seated::floorHeightC(L, gen, xy);
    F2 n0{}, n1{};
    const float a0 = seated::stepUpScanC(L, tb, xy, q, f0, n0), a1 = seated::stepUpScanC(L, gen, xy, q, f0, n1);
    if (!same(a0, a1) || !same(n0, n1)) return fail(nm, "stepUpScan", p);
    return true;
}

// bandOf as the reference evaluates it (src/venue.cpp: the row from the correctly rounded quotient, then corrected)
static int bandOfDivided(const seated::Layout& L, float dep) {
    const int nb = seated::bandCount(L);
    if (dep < L.standFront) return 0;
    if (dep < L.row0) return 1;
    if (dep < L.concFront) {
        int k = seated::cli(int(std::floor((dep - L.row0) / L.rowDepth)), 0, L.rows - 1);
        if (dep < seated::rowFront(L, k)) --k;
        else if (k + 1 < L.rows && dep >= seated::rowFront(L, k + 1)) ++k;
        return 2 + 2 * k + (dep >= seated::rowFront(L, k) + L.clearway ? 1 : 0);
    }
    if (dep < L.concBack) return nb - 4;
    if (dep < L.outerFace) return nb - 3;
    if (dep < L.landingEnd) return nb - 2;
    return nb - 1;
}

// bandOf against the divided form at every row front and clearway edge, 64 floats either side, and at random depths
static bool checkBands(const char* nm, const seated::Layout& L, std::mt19937& rng) {
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    auto one = [&](float dep) {
        ++g_checked;
        if (seated::bandOf(L, dep) != bandOfDivided(L, dep)) {
            std::printf("DIFF venue %s: bandOf at depth %a: %d, divided %d\n", nm, dep, seated::bandOf(L, dep), bandOfDivided(L, dep));
            return false;
        }
        return true;
    };
    for (int k = 0; k <= L.rows; ++k) {
        const float edges[2] = { seated::rowFront(L, k), seated::rowFront(L, k) + L.clearway };
        for (float e : edges) {
            float lo = e, hi = e;
            for (int i = 0; i < 64; ++i) {
                if (!one(lo) || !one(hi)) return false;
                lo = std::nextafter(lo, -1.0f);
                hi = std::nextafter(hi, 2e9f);
            }
        }
    }
    for (int i = 0; i < 1000000; ++i)
        if (!one(u(rng) * L.domainHalf)) return false;
    return true;
}

int main(int argc, char** argv) {
    const int perVenue = argc > 1 ? std::atoi(argv[1]) : 200000;
    for (const char* nm : { "S", "M", "L" }) {
        const Venue& v = venueByName(nm);
        const seated::Layout& L = v.layout;
        const CellTableData data = makeCellTable(L);
        const seated::CellTable tb = data.view();
        // the same table with every bucket at its band's last cell: every lateral search then walks back from the end
        CellTableData worst = data;
        for (size_t j = 0; j + 1 < worst.start.size(); ++j)
            for (int b = 0; b < seated::CELL_BUCKETS; ++b) worst.bucket[j * seated::CELL_BUCKETS + size_t(b)] = worst.start[j + 1] - 1;
        const seated::CellTable tbWorst = worst.view();
        std::mt19937 rng(12345u);
        if (!checkBands(nm, L, rng)) return 1;
        std::uniform_real_distribution<float> u(-1.0f, 1.0f), u01(0.0f, 1.0f);
        const float H = L.domainHalf;
        const float zs[5] = { 0.0f, L.firstRow, L.concourseZ, L.wallTop, L.barrierTop };
        auto randStep = [&] { return F2{ u(rng) * 0.004f, u(rng) * 0.004f }; };
        auto randDir = [&] {
            const float a = u(rng) * 3.14159265f;
            return F2{ std::cos(a), std::sin(a) };
        };
        // uniform over the domain, at the venue's typical heights plus a body
        for (int k = 0; k < perVenue; ++k) {
            const float z = zs[k % 5] + u01(rng) * 0.03f;
            if (!checkPoint(nm, L, tb, F3{ u(rng) * H, u(rng) * H, z }, randStep(), randDir())) return 1;
        }
        // on and next to every cell boundary and band edge, in random octants
        const int nb = seated::bandCount(L);
        for (int j = 0; j < nb; ++j) {
            float lo, hi;
            seated::band(L, j, lo, hi);
            if (hi > 1e8f) hi = lo + 0.05f;
            seated::Cell c{};
            for (int i = 0; seated::cellAt(L, j, i, c); ++i) {
                const float lats[6] = { c.l0, c.l1, std::nextafter(c.l0, -1.0f), std::nextafter(c.l1, 2e9f), 0.5f * (c.l0 + c.l1),
                                        std::nextafter(c.l1, -1.0f) };
                const float deps[5] = { lo, hi, std::nextafter(hi, -1.0f), 0.5f * (lo + hi), std::nextafter(lo, 2e9f) };
                for (float lat : lats)
                    for (float dep : deps) {
                        if (!(lat >= 0.0f) || lat > dep || dep > H) continue;
                        const float z = zs[(i + j) % 5] + 0.01f;
                        const int oct = int(rng() & 7u);
                        float x = (oct & 1) ? lat : dep, y = (oct & 1) ? dep : lat;
                        if (oct & 2) x = -x;
                        if (oct & 4) y = -y;
                        const F2 st = randStep(), e = randDir();
                        if (!checkPoint(nm, L, tb, F3{ x, y, z }, st, e) || !checkPoint(nm, L, tbWorst, F3{ x, y, z }, st, e)) return 1;
                    }
            }
        }
        // on and next to the table's bucket edges (the lateral search's start), in every band
        for (int j = 0; j < nb; ++j) {
            float lo, hi;
            seated::band(L, j, lo, hi);
            const float dep = hi > 1e8f ? lo + 0.01f : 0.5f * (lo + hi);
            const float invW = data.invW[size_t(j)];
            for (int b = 1; b < seated::CELL_BUCKETS && invW > 0.0f; ++b) {
                const float edge = float(b) / invW;
                const float lats[3] = { edge, std::nextafter(edge, -1.0f), std::nextafter(edge, 2e9f) };
                for (float lat : lats) {
                    if (!(lat >= 0.0f) || lat > dep || dep > H) continue;
                    if (!checkPoint(nm, L, tb, F3{ lat, dep, L.firstRow }, randStep(), randDir())) return 1;
                    if (!checkPoint(nm, L, tb, F3{ -dep, -lat, L.concourseZ }, randStep(), randDir())) return 1;
                }
            }
        }
        // non-finite points
        const float bad[3] = { NAN, INFINITY, -INFINITY };
        for (float b : bad)
            for (int k = 0; k < 3; ++k) {
                F3 p{ 0.1f, 0.2f, 0.0f };
                (k == 0 ? p.x : (k == 1 ? p.y : p.z)) = b;
                if (!checkPoint(nm, L, tb, p, F2{ 0.001f, 0.0f }, F2{ 1.0f, 0.0f })) return 1;
            }
    }
    std::printf("ok %ld\n", g_checked);
    return 0;
}
