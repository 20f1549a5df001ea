// The analytic seated venue evaluated on the GPU at a batch of points: every query of the heightfield (the signed distance
// and its gradient, the floor height, the cell under the point, the nearest wall at three reaches, the incline slope along
// a direction, the exit and the gangway). One thread per point; the layout travels as a kernel argument (it lives in
// scalar registers), so a query reads nothing from memory but its point.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "seated_venue.h"
#include "venue_probe.h"

namespace swarm {

namespace {

constexpr int PROBE_BLOCK = 256;

__global__ __launch_bounds__(PROBE_BLOCK) void k_venue_probe(seated::Layout L, uint32_t n, const float* __restrict__ p,
                                                              const float* __restrict__ e, float r0, float r1, float r2,
                                                              float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const seated::F3 q{ p[3 * i], p[3 * i + 1], p[3 * i + 2] };
    const seated::F2 xy{ q.x, q.y };
    float* grad = out;                   // n x 4
    float* floorH = out + 4 * size_t(n);  // n
    int* cell = reinterpret_cast<int*>(out + 5 * size_t(n));   // n x 3
    float* wall = out + 8 * size_t(n);    // n x 9
    float* slope = out + 17 * size_t(n);  // n
    int* exitId = reinterpret_cast<int*>(out + 18 * size_t(n));
    int* gangway = reinterpret_cast<int*>(out + 19 * size_t(n));
    const seated::F4 g = seated::sdfGrad(L, q);
    grad[4 * i] = g.x; grad[4 * i + 1] = g.y; grad[4 * i + 2] = g.z; grad[4 * i + 3] = g.w;
    floorH[i] = seated::floorHeight(L, xy);
    int row = 0, band = 0;
    const int kind = seated::kindAt(L, xy, &row, &band);
    cell[3 * i] = kind; cell[3 * i + 1] = row; cell[3 * i + 2] = band;
    const float reach[3] = { r0, r1, r2 };
    for (int k = 0; k < 3; ++k) {
        const seated::F3 w = seated::wall(L, q, reach[k]);
        wall[9 * i + 3 * k] = w.x; wall[9 * i + 3 * k + 1] = w.y; wall[9 * i + 3 * k + 2] = w.z;
    }
    slope[i] = seated::floorSlope(L, xy, seated::F2{ e[2 * i], e[2 * i + 1] });
    exitId[i] = seated::exitByLat(L, xy);
    gangway[i] = seated::gangwayAt(L, xy);
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("venue probe: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

void venueProbe(const seated::Layout& L, size_t n, const float* p, const float* e, const float reach[3], float* out) {
    if (n == 0) return;
    if (n > (size_t(1) << 26)) throw std::invalid_argument("venue probe: at most 2^26 points per call");
    float* dp = nullptr;
    float* de = nullptr;
    float* dout = nullptr;
    auto release = [&] {
        if (dp) (void)hipFree(dp);
        if (de) (void)hipFree(de);
        if (dout) (void)hipFree(dout);
    };
    try {
        check(hipMalloc(&dp, n * 3 * sizeof(float)), "hipMalloc");
        check(hipMalloc(&de, n * 2 * sizeof(float)), "hipMalloc");
        check(hipMalloc(&dout, n * VENUE_PROBE_WORDS * sizeof(float)), "hipMalloc");
        check(hipMemcpy(dp, p, n * 3 * sizeof(float), hipMemcpyHostToDevice), "copy in");
        check(hipMemcpy(de, e, n * 2 * sizeof(float), hipMemcpyHostToDevice), "copy in");
        const uint32_t blocks = uint32_t((n + PROBE_BLOCK - 1) / PROBE_BLOCK);
        k_venue_probe<<<blocks, PROBE_BLOCK, 0, 0>>>(L, uint32_t(n), dp, de, reach[0], reach[1], reach[2], dout);
        check(hipGetLastError(), "launch");
        check(hipMemcpy(out, dout, n * VENUE_PROBE_WORDS * sizeof(float), hipMemcpyDeviceToHost), "copy out");
    } catch (...) {
        release();
        throw;
    }
    release();
}

}  // namespace swarm
