// sf_grid.hip -- the uniform-grid searcher on caller device arrays.
//
// CudaGNBaseSearcher::BuildGNSearcher (seepage_flows_cuda/src/kiri_pbs_cuda/
// searcher/cuda_neighbor_searcher.cu:36-51) hashes a particle set's positions
// to cell keys (Pos2GridHash, helper_thrust.cuh:56-130), sorts the set stably
// by key (SortData, :64-116) and counts cellStart (CountingInCell_CUDA +
// exclusive_scan).  sf_grid_build does the same for positions the caller
// holds in device memory (packed float3, the reference's layout) and returns
// the sorted keys, the stable permutation (perm[new] = old) and cellStart
// (numCells + 1 entries), leaving the caller's arrays unpermuted.
//
// Work arrays: the LSD radix sort ping-pongs between one pair of key/value
// arrays of its own and the caller's output pair; which pair the first pass
// writes is chosen from the pass count so that the last pass lands in the
// caller's outputs.  The hashed keys start in whichever pair the first pass
// does not write.
#include <mutex>

#include "sf_internal.h"

namespace sf {
namespace {

__global__ __launch_bounds__(kBlock) void k_hash3(SFConst k, const float* __restrict__ pos3, int n,
                                                  uint32_t* __restrict__ keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int cx, cy, cz;
    cell_of(k, pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], cx, cy, cz);
    keys[i] = cell_key(k, cx, cy, cz);
}

// cellStart[c] = number of keys below c (lower bound in the sorted keys), c in [0, C]
__global__ __launch_bounds__(kBlock) void k_cell_start_search(const uint32_t* __restrict__ sk, int n, uint32_t C,
                                                              uint32_t* __restrict__ cs) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c > C) return;
    uint32_t lo = 0, hi = (uint32_t)n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sk[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    cs[c] = lo;
}

}  // namespace

// Work-array words for n particles: two arrays of n plus the radix histogram.
size_t grid_work_words(size_t n) {
    const size_t tiles = (n + kRsTile - 1) / kRsTile;
    return 2 * n + 256 * (tiles > 0 ? tiles : 1);
}

int grid_validate(const float* pos3, size_t n, float cell_size, sf_int3 g, uint32_t* sorted_keys, uint32_t* perm,
                  uint32_t* cell_start) {
    if (!(cell_size > 0.f) || !std::isfinite(cell_size)) return set_error(SF_E_INVALID, "cell size must be > 0");
    if (g.x < 1 || g.y < 1 || g.z < 1) return set_error(SF_E_INVALID, "grid size must be >= 1 per axis");
    if ((uint64_t)g.x * (uint64_t)g.y * (uint64_t)g.z >= (1ull << 31))
        return set_error(SF_E_INVALID, "more than 2^31 grid cells");
    if (n >= ((size_t)1 << 27)) return set_error(SF_E_CAPACITY, "more than 2^27 particles");
    if (!cell_start) return set_error(SF_E_INVALID, "null cell_start");
    if (n > 0 && (!pos3 || !sorted_keys || !perm)) return set_error(SF_E_INVALID, "null device array");
    return SF_OK;
}

// The build on a work area of grid_work_words(n) words.
int grid_build_with(uint32_t* work, const float* pos3, size_t n, sf_float3 lo, float cell_size, sf_int3 g,
                    uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start, hipStream_t s) {
    SFConst k{};
    k.lox = lo.x;
    k.loy = lo.y;
    k.loz = lo.z;
    k.h = cell_size;
    k.gx = g.x;
    k.gy = g.y;
    k.gz = g.z;
    k.C = (uint32_t)g.x * (uint32_t)g.y * (uint32_t)g.z;
    const int ni = (int)n;
    if (ni > 0) {
        const RadixPlan plan = make_radix_plan(k.C);
        uint32_t* wk = work;
        uint32_t* wv = work + n;
        uint32_t* hist = work + 2 * n;
        // odd pass count: first pass into the caller's pair; even: into the work pair
        const bool odd = plan.passes & 1;
        SortBuffers b = odd ? SortBuffers{sorted_keys, perm, wk, wv, hist} : SortBuffers{wk, wv, sorted_keys, perm, hist};
        uint32_t* keys0 = odd ? wk : sorted_keys;
        k_hash3<<<(ni + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, pos3, ni, keys0);
        const uint32_t *sk = nullptr, *pm = nullptr;
        int rc = radix_sort(keys0, ni, plan, b, &sk, &pm, s);
        if (rc) return rc;
        if (sk != sorted_keys || pm != perm) return set_error(SF_E_STATE, "radix sort did not end in the outputs");
    }
    k_cell_start_search<<<(k.C + 1 + kBlock - 1) / kBlock, kBlock, 0, s>>>(sorted_keys, ni, k.C, cell_start);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

}  // namespace sf

using namespace sf;

extern "C" {

int sf_grid_build(const float* pos3, size_t n, sf_float3 lowest, float cell_size, sf_int3 grid_size,
                  uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start, void* stream) {
    int rc;
    if ((rc = grid_validate(pos3, n, cell_size, grid_size, sorted_keys, perm, cell_start))) return rc;
    hipStream_t s = (hipStream_t)stream;
    DevBuf<uint32_t> work;
    if ((rc = work.alloc(grid_work_words(n)))) return rc;
    if ((rc = grid_build_with(work.p, pos3, n, lowest, cell_size, grid_size, sorted_keys, perm, cell_start, s)))
        return rc;
    SF_HIP(hipStreamSynchronize(s));  // the work array is released on return
    return SF_OK;
}

}  // extern "C"

// ---- the same build without allocation -------------------------------------
namespace sf {
namespace {
// Largest particle count of sf_grid_build_noalloc and its work area:
// 2 * 2^20 words of key / value ping-pong plus the radix histogram of 2^20 / kRsTile tiles
// (8.25 MiB per GPU, held until the process ends).
constexpr size_t kNoAllocMaxParticles = (size_t)1 << 20;
constexpr size_t kScratchWords = 2 * kNoAllocMaxParticles + 256 * (kNoAllocMaxParticles / kRsTile);
static_assert(kNoAllocMaxParticles % kRsTile == 0, "whole radix tiles");

__device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[kMaxDevices] = {};  // the array's address on each device

// The array's device address on the current device, looked up once per device
// (each GPU holds its own copy at its own address; host code never passes the
// array's name itself to a kernel).
int scratch_on_current_device(uint32_t** out) {
    int dev = -1;
    SF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(SF_E_INVALID, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[dev] = (uint32_t*)p;
    }
    *out = g_scratch_dev[dev];
    return SF_OK;
}
}  // namespace
}  // namespace sf

extern "C" {

int sf_grid_build_noalloc(const float* pos3, size_t n, sf_float3 lowest, float cell_size, sf_int3 grid_size,
                          uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start, void* stream) {
    int rc;
    if ((rc = grid_validate(pos3, n, cell_size, grid_size, sorted_keys, perm, cell_start))) return rc;
    if (n > kNoAllocMaxParticles)
        return set_error(SF_E_CAPACITY, "sf_grid_build_noalloc takes at most 1048576 particles (use sf_grid_build)");
    uint32_t* work = nullptr;
    if ((rc = scratch_on_current_device(&work))) return rc;
    static_assert(kScratchWords >= 2 * kNoAllocMaxParticles + 256 * ((kNoAllocMaxParticles + kRsTile - 1) / kRsTile),
                  "work area of grid_build_with");
    return grid_build_with(work, pos3, n, lowest, cell_size, grid_size, sorted_keys, perm, cell_start,
                           (hipStream_t)stream);
}

}  // extern "C"
