// swarm_sort_keys_noalloc: the sort stage of swarm_sort_grid (CrowdSim::sortGrid) on device arrays of the caller,
// allocating nothing. Its working memory is one file-scope device array, 'scratch' below: each GPU the library runs
// on holds its own copy from the first call on that GPU until the process ends. For callers that cannot allocate
// device memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// The sort is the stable one of the counting sort (ascending (key, index), inactive keys last), computed on the
// bitonic network with whole 64-bit compares, and the cell table comes from the counting pass: the same results as
// swarm_sort_grid followed by swarm_read_cells, bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/swarm_hip.h"
#include "crowd_kernels.h"

namespace {

using namespace swarm;

constexpr uint32_t MAX_KEYS = SWARM_NOALLOC_MAX_KEYS;
constexpr uint32_t MAX_CELLS = SWARM_NOALLOC_MAX_CELLS;
constexpr size_t ALIGN_WORDS = 64;   // 256 bytes
constexpr size_t up(size_t w) { return (w + ALIGN_WORDS - 1) / ALIGN_WORDS * ALIGN_WORDS; }
// the largest segments: loc n, bcnt and bstart C + n / HASH_BLOCK (+ 1), tileSum, ctrl, pairs (2 words each) n
constexpr size_t NB_MAX = size_t(MAX_CELLS) + (size_t(MAX_KEYS) + HASH_BLOCK - 1) / HASH_BLOCK;
constexpr size_t TILES_MAX = (NB_MAX + SCAN_TILE - 1) / SCAN_TILE;
constexpr size_t OFF_LOC = 0;
constexpr size_t OFF_BCNT = OFF_LOC + up(MAX_KEYS);
constexpr size_t OFF_BSTART = OFF_BCNT + up(NB_MAX);
constexpr size_t OFF_TILES = OFF_BSTART + up(NB_MAX + 1);
constexpr size_t OFF_CTRL = OFF_TILES + up(TILES_MAX);
constexpr size_t OFF_PAIRS = OFF_CTRL + up(CTRL_WORDS);
constexpr size_t SCRATCH_WORDS = OFF_PAIRS + 2 * size_t(MAX_KEYS);
static_assert(SCRATCH_WORDS * 4 == SWARM_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

std::mutex g_mu;
constexpr int MAX_DEVICES = 64;
uint32_t* g_ptr[MAX_DEVICES] = {};   // per device: the address of its copy of 'scratch'
thread_local std::string g_msg;

}  // namespace

// The working memory (device side). Host code never passes this name to a kernel or a copy: it stands for the array
// only in the device's address space; host code uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) uint32_t scratch[SCRATCH_WORDS];

extern "C" const char* swarm_noalloc_last_error(void) { return g_msg.c_str(); }

extern "C" int swarm_sort_keys_noalloc(const uint32_t* keys, uint32_t n, uint32_t cells, uint32_t* sorted_keys, uint32_t* ids,
                                       uint32_t* cell_start, uint32_t* cell_count, void* stream) {
    if (n == 0 || n > MAX_KEYS) {
        g_msg = "swarm_sort_keys_noalloc: n must be 1 .. " + std::to_string(MAX_KEYS) + " (got " + std::to_string(n) + ")";
        return SWARM_EINVAL;
    }
    if (cells == 0 || cells > MAX_CELLS) {
        g_msg = "swarm_sort_keys_noalloc: cells must be 1 .. " + std::to_string(MAX_CELLS) + " (got " + std::to_string(cells) + ")";
        return SWARM_EINVAL;
    }
    if (!keys) {
        g_msg = "swarm_sort_keys_noalloc: null keys";
        return SWARM_EINVAL;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        g_msg = "swarm_sort_keys_noalloc: no current HIP device";
        return SWARM_EHIP;
    }
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_ptr[dev]) {   // once per device, with that device current: its copy's address
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_msg = std::string("swarm_sort_keys_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
                return SWARM_EHIP;
            }
            g_ptr[dev] = static_cast<uint32_t*>(p);
        }
        base = g_ptr[dev];
    }
    const KeySortScratch s{ base + OFF_LOC, base + OFF_BCNT, base + OFF_BSTART, base + OFF_TILES,
                            reinterpret_cast<uint64_t*>(base + OFF_PAIRS), base + OFF_CTRL };
    const hipStream_t st = static_cast<hipStream_t>(stream);
    launchStableKeySort(st, n, cells, keys, s, sorted_keys, ids, cell_start, cell_count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_msg = std::string("swarm_sort_keys_noalloc: ") + hipGetErrorString(e);
        return SWARM_EHIP;
    }
    return SWARM_OK;
}
