// The batched evaluation of the analytic seated venue on the GPU (venue_probe.hip)
#pragma once

#include <cstddef>

#include "seated_venue.h"

namespace swarm {

// Words per point of venueProbe's output
constexpr size_t VENUE_PROBE_WORDS = 20;

// Evaluates every query at n points p (n x 3) with directions e (n x 2) on the current device, synchronously. out holds
// 20 n words, as consecutive arrays: sdfGrad n x 4, floorHeight n, (kind, row, band) n x 3 (int32), wall at reach[0..2]
// n x 9, floorSlope n, exitByLat n (int32), gangwayAt n (int32).
void venueProbe(const seated::Layout& L, size_t n, const float* p, const float* e, const float reach[3], float* out);

}  // namespace swarm
