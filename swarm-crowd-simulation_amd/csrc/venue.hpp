// The reference's analytic seated venues S, M and L on the host: the layout generator, the seats, and the venue's plan
// as a rectangle world for the per-step pipeline.
#pragma once

#include <string>
#include <vector>

#include "scene.hpp"
#include "seated_venue.h"

namespace swarm {

// SeatedVenueParams (include/swarm/venue.hpp:60-99 of the reference): metres unless stated
struct VenueParams {
    std::string name;
    double walkwayWidthM, barrierThicknessM, avenueWidthM, sideWallThicknessM;
    int rows;
    double rowDepthM, clearwayM, seatWidthM, gangwayWidthM, seatBlockInsetM;
    int maxSeatsBetweenGangways, maxSeatsDeadEnd;
    double concourseWidthM, outerWallThicknessM, landingDepthM, rowRiseM, seatBlockTopM, barrierTopM, wallTopAboveConcourseM;
    int exitsPerQuadrant;
    double exitRatePerMMin, egressTimeMin, minExitWidthM;
    bool doorPerGangwayHead;   // false: one door per wing (venue S)
    int gridRes;               // neighbour grid cells per axis (0: the terrace's 40)
    int metricGridRes;         // 1 m metric cells per axis (0: 64)
};

// A venue: its parameters, the neighbour-grid domain and its layout, with the exit sizing rule's figures (metres)
struct Venue {
    VenueParams params;
    float domainMin, domainSpan;
    int gridRes, metricGridRes;
    float exitDepth, exitMaxZ;
    seated::Layout layout;
    double doorWidthM, cornerDoorWidthM, totalExitWidthM, requiredExitWidthM;
    int seatTotal;
};

// "S", "M" or "L" (either case); std::invalid_argument otherwise
const Venue& venueByName(const std::string& name);
// The layout of a parameter set in a domain of half-width domainHalf, with the generator's design rules checked
// (std::runtime_error naming the rule on a violation)
Venue makeVenue(const VenueParams& p);

// A seat: the spawn position (clearway centre, the tread's rest height) and where it is
struct VenueSeat {
    float x, y, z, lat;
    int quadrant, wing, row, segment, index;
};
// Every seat, in the order quadrant, wing, row, segment, index (seated::seats, src/venue.cpp:1054-1083)
std::vector<VenueSeat> venueSeats(const seated::Layout& L);
int seatCountIn(float a, float b, float seatWidth);

// The exit clear width (metres) of reference exit e (seated::exitWidthM, src/venue.cpp:867-874)
double venueExitWidthM(const Venue& v, int e);

// A layout's cell table (seated::CellTable) in host memory: every band's cells as the generator gives them and the
// lateral buckets of the table's search, as arrays (view()) and packed into one blob (blob, blobLayout; for the device);
// std::logic_error when a band is not sorted (a cell with l0 > l1, or starting before the previous cell's l1), since the
// table's queries search sorted bands, or when it has 65,536 cells or more (the buckets hold 16-bit indices)
struct CellTableData {
    std::vector<int> start;
    std::vector<float> l1;
    std::vector<int4> cell;
    std::vector<uint16_t> bucket;
    std::vector<float> invW;
    std::vector<char> blob;
    seated::CellBlob blobLayout{};
    seated::CellTable view() const { return seated::CellTable{ start.data(), l1.data(), cell.data(), bucket.data(), invW.data() }; }
};
CellTableData makeCellTable(const seated::Layout& L);

// The venue's plan for the per-step pipeline: walls are the footprints of the seat blocks, the barrier, the side and
// outer walls (each stands higher than a body above every floor beside it) and the fenced pitch; agents leave when
// max(|x|, |y|) passes the outer wall's outer face (the domain's edge). Exit rectangles lie beyond that face, one per
// avenue and door; the corner door, L-shaped there, is two rectangles, the second appended after the reference's exits.
// refExit[e] is the reference's exit id of plan exit e. The plan carries the layout: the step applies the reference's
// default model's incline rules on the gangway inclines (the free speed and the headway time gap by the floor's slope
// along e_goal), no stair bands.
struct VenuePlan {
    SceneDesc desc;
    std::vector<int> refExit;
};
VenuePlan venuePlan(const Venue& v);

// The venue for its own 3D step: the reference's neighbour grid (gridRes cells over the domain), its navigation field on
// the heightfield (0.05 m cells, SeatedGeometry) with the candidate table and the exit-type rule, its exit test (past the
// outer face below exitMaxZ, or out of the domain); the heightfield itself is the scene (no distance grid, walls or exit
// rectangles). Every table is the reference's bit for bit.
NavGrid venueNavGrid(const Venue& v);
Scene buildSeatedScene(const Venue& v, int threads = 0);
// The same through the checksummed scene cache in dir (empty: no cache)
Scene buildSeatedSceneCached(const Venue& v, int threads, const std::string& dir, int* fromCache);

}  // namespace swarm
