// The oracle's analytic seated venue (venue_oracle.cpp) as the oracle's seated step reads it: TEST INFRASTRUCTURE ONLY.
// A venue is an opaque handle from oracle_venue_new. Each function restates the reference's seated:: function named.
#pragma once

namespace vo {
void sdfGrad(const void* v, float x, float y, float z, float out[4]);             // seated::sdfGrad (gradient, distance)
void wall(const void* v, float x, float y, float z, float reach, float out[3]);   // seated::wall (away xy, clearance)
float floorHeight(const void* v, float x, float y);                                // seated::floorHeight
int cellAt(const void* v, float x, float y, int* row);                            // seated::cellAt (kind; row)
void cellKey(const void* v, float x, float y, float out[2]);                      // seated::cellKey (band, l0)
float floorSlope(const void* v, float x, float y, float ex, float ey);            // seated::floorSlope
int exitByLat(const void* v, float x, float y);                                   // seated::exitByLat
// seated::stepUpScan (src/venue.cpp:807-835): the fraction of the move p0 -> p1 allowed from start floor f0; n the face's normal
float stepUpScan(const void* v, float p0x, float p0y, float p1x, float p1y, float f0, float n[2]);
float rake(const void* v);
float wallAbove(const void* v);
float wallClear(const void* v);
float faceMargin(const void* v);
}  // namespace vo
