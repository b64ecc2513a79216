// sf_searcher.hip -- uniform-grid neighbour searcher for gfx950.
//
// Replaces CudaGNBaseSearcher::BuildGNSearcher / CudaGNSearcher::SortData
// (cuda_neighbor_searcher.cu:36-116): FP32 cell hash, a stable LSD radix sort
// of 32-bit cell keys (wave64 ballot matching for in-wave ranks, LDS digit
// counters per wave, a scanned per-block digit histogram), cellStart from the
// sorted keys, and one gather of the persistent particle state.
#include "sf_internal.h"

namespace sf {

// ---------------------------------------------------------------------------
// hash: keys[i] = cell of pm[i]  (Pos2GridHash, helper_thrust.cuh:65-89)
template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_hash(SFConst k, const float4* __restrict__ pm,
                                                 const uint16_t* __restrict__ scene, uint32_t* __restrict__ keys,
                                                 int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = pm[i];
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    uint32_t key = cell_key(k, cx, cy, cz);
    if constexpr (kBatch) key += (uint32_t)scene[i] * k.C;
    keys[i] = key;
}
void launch_hash(const SFConst& k, const float4* pm, const uint16_t* scene, uint32_t* keys, int n, hipStream_t s) {
    if (n <= 0) return;
    if (scene)
        k_hash<true><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, pm, scene, keys, n);
    else
        k_hash<false><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, pm, nullptr, keys, n);
}
void launch_hash_boundary(const SFConst& k, const float4* bp, uint32_t* keys, int nb, hipStream_t s) {
    launch_hash(k, bp, nullptr, keys, nb, s);
}

// ---------------------------------------------------------------------------
// radix sort
RadixPlan make_radix_plan(uint32_t num_cells) {
    RadixPlan p;
    int K = 0;
    if (num_cells > 1) {
        uint32_t m = num_cells - 1;
        while (m) {
            ++K;
            m >>= 1;
        }
    }
    if (K < 1) K = 1;
    p.passes = (K + 7) / 8;
    int sh = 0;
    for (int i = 0; i < p.passes; ++i) {
        int rem_passes = p.passes - i, rem_bits = K - sh;
        int b = (rem_bits + rem_passes - 1) / rem_passes;
        p.shift[i] = sh;
        p.bits[i] = b;
        sh += b;
    }
    return p;
}

// lanes of the wave whose digit equals this lane's digit (among valid lanes)
SF_DEV uint64_t match_digit(uint32_t d, int bits, bool valid) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// per-block digit counts -> hist[d * nblocks + block]
__global__ __launch_bounds__(kRsThreads) void k_rs_upsweep(const uint32_t* __restrict__ keys, int n, int shift,
                                                           int bits, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const int t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    const int base = blockIdx.x * kRsTile;
    const uint64_t lt = __lanemask_lt();
#pragma unroll 4
    for (int r = 0; r < kRsItems; ++r) {
        const int idx = base + r * kRsThreads + t;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & mask : 0u;
        const uint64_t peers = match_digit(d, bits, valid);
        if (valid && (peers & lt) == 0) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    if (t <= (int)mask) hist[(size_t)t * gridDim.x + blockIdx.x] = cnt[t];
}

// exclusive scan of hist[0, total) in place by one workgroup of 1024
// threads; *total_out (if given) receives the sum
SF_DEV void scan_range(uint32_t* __restrict__ hist, int total, uint32_t* __restrict__ total_out) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < total; base += 1024 * 4) {
        uint32_t v[4];
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + t * 4 + q;
            v[q] = idx < total ? hist[idx] : 0u;
            s += v[q];
        }
        // inclusive wave scan of s
        uint32_t x = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        if (w == 0) {
            uint32_t ws = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                uint32_t y = __shfl_up(ws, o, 64);
                if (lane >= o) ws += y;
            }
            if (lane < 16) wsum[lane] = ws;  // inclusive over waves
        }
        __syncthreads();
        const uint32_t carry = carry_s;
        uint32_t run = carry + (w > 0 ? wsum[w - 1] : 0u) + x - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + t * 4 + q;
            if (idx < total) hist[idx] = run;
            run += v[q];
        }
        __syncthreads();
        if (t == 1023) carry_s = carry + wsum[15];
        __syncthreads();
    }
    if (total_out && t == 0) *total_out = carry_s;
    __syncthreads();
}
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ hist, int total) {
    scan_range(hist, total, nullptr);
}
// two independent scans, [0, n1) and [n1, n1 + n2), in one launch
__global__ __launch_bounds__(1024) void k_rs_scan2(uint32_t* __restrict__ part, int n1, int n2,
                                                  uint32_t* __restrict__ total_out) {
    scan_range(part, n1, total_out);
    scan_range(part + n1, n2, nullptr);
}

// stable scatter: each wave ranks a contiguous chunk of 64*kRsItems keys
__global__ __launch_bounds__(kRsThreads) void k_rs_downsweep(const uint32_t* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                             uint32_t* __restrict__ vout, int n, int shift, int bits,
                                                             const uint32_t* __restrict__ hist) {
    __shared__ uint32_t wcnt[4][256];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int q = 0; q < 4; ++q) wcnt[q][t] = 0;
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    const uint64_t lt = __lanemask_lt();
    const int base = blockIdx.x * kRsTile + w * (64 * kRsItems);
    uint32_t key[kRsItems], val[kRsItems], rank[kRsItems];
    volatile uint32_t* wc = wcnt[w];
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const int idx = base + r * 64 + lane;
        const bool valid = idx < n;
        key[r] = valid ? kin[idx] : 0xffffffffu;
        val[r] = valid ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = valid ? (key[r] >> shift) & mask : 0u;
        const uint64_t peers = match_digit(d, bits, valid);
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        const uint32_t c = wc[d];
        rank[r] = c + before;
        if (valid && before == 0) wc[d] = c + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        const uint32_t c0 = wcnt[0][t], c1 = wcnt[1][t], c2 = wcnt[2][t];
        const uint32_t g = t <= (int)mask ? hist[(size_t)t * gridDim.x + blockIdx.x] : 0u;
        __syncthreads();
        wcnt[0][t] = g;
        wcnt[1][t] = g + c0;
        wcnt[2][t] = g + c0 + c1;
        wcnt[3][t] = g + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const int idx = base + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t pos = wcnt[w][d] + rank[r];
            if (pos < (uint32_t)n) {  // always true for a consistent histogram; keeps a bug in bounds
                kout[pos] = key[r];
                vout[pos] = val[r];
            }
        }
    }
}

int radix_sort(const uint32_t* keys, int n, const RadixPlan& plan, const SortBuffers& b, const uint32_t** sorted_keys,
               const uint32_t** perm, hipStream_t s) {
    *sorted_keys = b.kA;
    *perm = b.vA;
    if (n <= 0) return SF_OK;
    const int nblocks = (n + kRsTile - 1) / kRsTile;
    const uint32_t* kin = keys;
    const uint32_t* vin = nullptr;  // identity on the first pass
    uint32_t* ko = b.kA;
    uint32_t* vo = b.vA;
    for (int p = 0; p < plan.passes; ++p) {
        k_rs_upsweep<<<nblocks, kRsThreads, 0, s>>>(kin, n, plan.shift[p], plan.bits[p], b.hist);
        const int total = nblocks << plan.bits[p];
        k_rs_scan<<<1, 1024, 0, s>>>(b.hist, total);
        k_rs_downsweep<<<nblocks, kRsThreads, 0, s>>>(kin, vin, ko, vo, n, plan.shift[p], plan.bits[p], b.hist);
        *sorted_keys = ko;
        *perm = vo;
        kin = ko;
        vin = vo;
        ko = (ko == b.kA) ? b.kB : b.kA;
        vo = (vo == b.vA) ? b.vB : b.vA;
    }
    SF_HIP(hipGetLastError());
    return SF_OK;
}

// ---------------------------------------------------------------------------
// cellStart[c] = number of particles with key < c, c in [0, C]
// (CountingInCell_CUDA + exclusive_scan, cuda_neighbor_searcher.cu:45-47)
__global__ __launch_bounds__(kBlock) void k_cell_start(const uint32_t* __restrict__ sk, int n, uint32_t C,
                                                       const uint32_t* __restrict__ wpre, uint32_t* __restrict__ cs,
                                                       uint2* __restrict__ ct, uint32_t* __restrict__ zero_c,
                                                       PhasePos pp, IncStart inc) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    // pph[t] = pm[plist[t]]: coalesced stores, loads that walk each phase's
    // slots in increasing order
    if (pp.pph && c < (uint32_t)n) pp.pph[c] = ldo(pp.pm, ldo(pp.plist, c));
    if (c > C) return;
    if (zero_c && c < C) zero_c[c] = 0u;  // the next incremental sort's newcomer counters
    uint32_t lo = 0;
    if (inc.stay_pre) {
        // particles with a new key below c: the stayers (key unchanged) of the
        // old cells below c and the newcomers of the cells below c
        lo = inc.stay_pre[cs[c]] + inc.in_start[c];
    } else {
        uint32_t hi = (uint32_t)n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sk[mid] < c) lo = mid + 1;
            else hi = mid;
        }
    }
    cs[c] = lo;
    if (ct) ct[c] = make_uint2(lo, wpre ? wpre[lo] : 0u);
}
void launch_cell_start(const uint32_t* sorted_keys, int n, uint32_t C, const uint32_t* wpre, uint32_t* cell_start,
                       uint2* ct, uint32_t* zero_c, PhasePos pp, IncStart inc, hipStream_t s) {
    uint32_t total = C + 1;
    if (pp.pph && (uint32_t)n > total) total = (uint32_t)n;
    k_cell_start<<<(total + kBlock - 1) / kBlock, kBlock, 0, s>>>(sorted_keys, n, C, wpre, cell_start, ct, zero_c,
                                                                  pp, inc);
}

// ---------------------------------------------------------------------------
// reorder of the persistent state (the zipped payload of sort_by_key,
// cuda_neighbor_searcher.cu:90-111)
// Each block gathers one water-prefix tile (kWpTile slots) and stores the
// tile's water count in wpart[block] for the prefix that follows.
constexpr int kWpItems = 8;
constexpr int kWpTile = kBlock * kWpItems;
__global__ __launch_bounds__(kBlock) void k_gather(GatherArgs a, int n, uint32_t* __restrict__ wpart) {
    __shared__ uint32_t ws[kBlock / 64];
    const int t = threadIdx.x;
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kWpItems; ++q) {
        const int i = blockIdx.x * kWpTile + q * kBlock + t;
        if (i >= n) break;
        uint32_t p = a.perm[i];
        if (p >= (uint32_t)n) p = (uint32_t)i;  // a permutation of [0, n) never takes this branch
        const float4 pm = a.pm_in[p];
        a.pm_out[i] = pm;
        a.vr_out[i] = a.vr_in[p];
        a.cs_out[i] = a.cs_in[p];
        if (a.multi) {
            a.cda_out[i] = a.cda_in[p];
            a.amc_out[i] = a.amc_in[p];
        }
        const uint32_t water = pm.w > 0.f ? 1u : 0u;
        a.wflag[i] = (uint8_t)water;
        c += water;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((t & 63) == 0) ws[t >> 6] = c;
    __syncthreads();
    if (t == 0) {
        uint32_t sum = 0;
        for (int q = 0; q < kBlock / 64; ++q) sum += ws[q];
        wpart[blockIdx.x] = sum;
    }
}

__global__ __launch_bounds__(kBlock) void k_boundary_gather(const uint32_t* __restrict__ perm,
                                                            const float4* __restrict__ in, float4* __restrict__ out,
                                                            int nb) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nb) return;
    uint32_t p = perm[i];
    if (p >= (uint32_t)nb) p = (uint32_t)i;
    out[i] = in[p];
}
void launch_boundary_gather(const uint32_t* perm, const float4* bp_in, float4* bp_out, int nb, hipStream_t s) {
    if (nb <= 0) return;
    k_boundary_gather<<<(nb + kBlock - 1) / kBlock, kBlock, 0, s>>>(perm, bp_in, bp_out, nb);
}

// ComputeBoundaryVolume_CUDA (cuda_sph_system_gpu.cuh:18-66): bpv.w enters as
// +-1 (label sign) and leaves as +-V_b.
__global__ __launch_bounds__(kBlock) void k_boundary_volume(SFConst k, float4* __restrict__ bpv,
                                                            const uint32_t* __restrict__ bcs, int nb) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nb) return;
    const float4 pi = bpv[i];
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    float sum = 0.f;
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t e = bcs[c1];
        for (uint32_t j = bcs[c0]; j < e; ++j) {
            if ((int)j == i) continue;
            const float4 pj = bpv[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            sum += poly6_r2(k, dx * dx + dy * dy + dz * dz);
        }
    });
    const float v = 1.f / fmaxf(sum, kEps);
    // every thread reads only .xyz of other slots, so writing .w is race-free
    bpv[i].w = pi.w < 0.f ? -v : v;
}
void launch_boundary_volume(const SFConst& k, float4* bpv, const uint32_t* bcs, int nb, hipStream_t s) {
    if (nb <= 0) return;
    k_boundary_volume<<<(nb + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, bpv, bcs, nb);
}

// Cells whose 27-cell neighbourhood holds boundary particles: the passes'
// boundary loops skip the other cells' 9 empty column ranges.
__global__ __launch_bounds__(kBlock) void k_boundary_near(SFConst k, const uint32_t* __restrict__ bcs,
                                                          uint8_t* __restrict__ near) {
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= k.C) return;
    const int cz = (int)(c % (uint32_t)k.gz), cy = (int)((c / (uint32_t)k.gz) % (uint32_t)k.gy);
    const int cx = (int)(c / ((uint32_t)k.gz * (uint32_t)k.gy));
    bool any = false;
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) { any = any || bcs[c1] > bcs[c0]; });
    near[c] = any ? 1 : 0;
}
void launch_boundary_near(const SFConst& k, const uint32_t* bcs, uint8_t* near, hipStream_t s) {
    if (k.C == 0) return;
    k_boundary_near<<<(k.C + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, bcs, near);
}

// ---------------------------------------------------------------------------
// wpre[i] = number of water particles among sorted slots [0, i), i in [0, n]:
// lets a neighbour loop see how many water / sand particles a contiguous
// cell range holds (wpre[e] - wpre[s]) and skip ranges without the phase it
// needs.  Two kernels around a tiny scan of per-block counts.
int wprefix_blocks(int n) { return (n + kWpTile - 1) / kWpTile; }

// Exclusive prefix of a per-slot value v(i), i in [0, n): block partial sums,
// one-workgroup scan of the partials, then per-slot prefixes; slot n gets the
// total.  Used for the water prefix and the incremental sort.
template <class V>
SF_DEV void tile_sum(V v, int n, int tile, uint32_t* __restrict__ part_out) {
    __shared__ uint32_t ws[kBlock / 64];
    const int t = threadIdx.x;
    const int base = tile * kWpTile;
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kWpItems; ++q) {
        const int i = base + q * kBlock + t;
        c += i < n ? v(i) : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((t & 63) == 0) ws[t >> 6] = c;
    __syncthreads();
    if (t == 0) {
        uint32_t sum = 0;
        for (int q = 0; q < kBlock / 64; ++q) sum += ws[q];
        *part_out = sum;
    }
}
// tile's slot prefixes from its scanned partial (thread t owns kWpItems
// consecutive slots); slot n (one past the end) receives the total.
// emit(i, prefix, value) is called for every slot i < n.
// Up to kDirectTiles tiles, each prefix block sums the tile counts it needs
// itself instead of waiting for a one-block scan launch: the sum of cnt[q]
// over q in [lo, b) (before) and over [lo, hi) (all), by the whole block.
constexpr int kDirectTiles = 4096;
int g_direct_tiles = kDirectTiles;  // sf_set_prefix_scan_threshold
SF_DEV void block_sums(const uint32_t* __restrict__ cnt, uint32_t lo, uint32_t b, uint32_t hi, uint32_t& before,
                       uint32_t& all) {
    __shared__ uint32_t red[2][kBlock / 64];
    uint32_t a = 0, p = 0;
    for (uint32_t q = lo + threadIdx.x; q < hi; q += kBlock) {
        const uint32_t c = cnt[q];
        a += c;
        p += q < b ? c : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        p += __shfl_down(p, o, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = p;
    }
    __syncthreads();
    all = 0;
    before = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        all += red[0][w];
        before += red[1][w];
    }
}

template <class V, class Emit>
SF_DEV void tile_prefix(V v, int n, int tile, uint32_t part, uint32_t* __restrict__ out, Emit&& emit) {
    __shared__ uint32_t wsum[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int base = tile * kWpTile + t * kWpItems;
    uint32_t f[kWpItems];
    uint32_t s = 0;
    if (base + kWpItems <= n) {
        v.load8(base, f);  // whole run in range: vector loads
    } else {
#pragma unroll
        for (int q = 0; q < kWpItems; ++q) f[q] = base + q < n ? v(base + q) : 0u;
    }
#pragma unroll
    for (int q = 0; q < kWpItems; ++q) s += f[q];
    uint32_t x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t run = part + x - s;
    for (int q = 0; q < w; ++q) run += wsum[q];
#pragma unroll
    for (int q = 0; q < kWpItems; ++q) {
        const int i = base + q;
        if (i <= n) out[i] = run;
        if (i < n) emit(i, run, f[q]);
        run += f[q];
    }
}
struct NoEmit {
    SF_DEV void operator()(int, uint32_t, uint32_t) const {}
};

// Exclusive prefixes of two per-slot values in three launches: v1 over
// [0, n1) into out1[0, n1], v2 over [0, n2) into out2[0, n2].  Blocks
// [0, nb1) work on the first, [nb1, nb1 + nb2) on the second.
template <class V1, class V2>
__global__ __launch_bounds__(kBlock) void k_vsum2(V1 v1, int n1, int nb1, V2 v2, int n2, uint32_t* __restrict__ part) {
    const int b = blockIdx.x;
    if (b < nb1) tile_sum(v1, n1, b, part + b);
    else tile_sum(v2, n2, b - nb1, part + b);
}
template <class V1, class V2>
// scanned = 0: part holds the raw block sums, and each block sums those
// before it within its own prefix (the block holding slot n thereby ends on
// the total, as the scan would give it).
__global__ __launch_bounds__(kBlock) void k_vprefix2(V1 v1, int n1, int nb1, uint32_t* __restrict__ out1, V2 v2,
                                                     int n2, uint32_t* __restrict__ out2,
                                                     const uint32_t* __restrict__ part, int scanned) {
    const int b = blockIdx.x;
    uint32_t pb = 0;
    if (scanned) {
        pb = part[b];
    } else {
        uint32_t all;
        if (b < nb1) block_sums(part, 0u, (uint32_t)b, (uint32_t)nb1, pb, all);
        else block_sums(part, (uint32_t)nb1, (uint32_t)b, gridDim.x, pb, all);
    }
    if (b < nb1) tile_prefix(v1, n1, b, pb, out1, NoEmit{});
    else tile_prefix(v2, n2, b - nb1, pb, out2, NoEmit{});
}
int vprefix2_partials(int n1, int n2) { return wprefix_blocks(n1 + 1) + wprefix_blocks(n2 + 1); }
template <class V1, class V2>
void launch_vprefix2(V1 v1, int n1, uint32_t* out1, V2 v2, int n2, uint32_t* out2, uint32_t* part, hipStream_t s) {
    // blocks cover slots [0, n]; the block holding slot n counts nothing and
    // receives the total from the prefix of the blocks before it
    const int nb1 = wprefix_blocks(n1 + 1), nb2 = wprefix_blocks(n2 + 1);
    k_vsum2<<<nb1 + nb2, kBlock, 0, s>>>(v1, n1, nb1, v2, n2, part);
    const int scanned = nb1 + nb2 > g_direct_tiles ? 1 : 0;
    if (scanned) k_rs_scan2<<<1, 1024, 0, s>>>(part, nb1, nb2, nullptr);
    k_vprefix2<<<nb1 + nb2, kBlock, 0, s>>>(v1, n1, nb1, out1, v2, n2, out2, part, scanned);
}
// Per-slot values of the prefixes; load8(i, f) reads slots i .. i + 7 (i a
// multiple of 8, all below n) with vector loads.
SF_DEV void load8_u32(const uint32_t* a, int i, uint4& x, uint4& y) {
    x = *reinterpret_cast<const uint4*>(a + i);
    y = *reinterpret_cast<const uint4*>(a + i + 4);
}
struct StayVal {
    const uint32_t* key;
    const uint32_t* old;
    SF_DEV uint32_t operator()(int i) const { return key[i] == old[i] ? 1u : 0u; }
    SF_DEV void load8(int i, uint32_t* f) const {
        uint4 k0, k1, o0, o1;
        load8_u32(key, i, k0, k1);
        load8_u32(old, i, o0, o1);
        f[0] = k0.x == o0.x; f[1] = k0.y == o0.y; f[2] = k0.z == o0.z; f[3] = k0.w == o0.w;
        f[4] = k1.x == o1.x; f[5] = k1.y == o1.y; f[6] = k1.z == o1.z; f[7] = k1.w == o1.w;
    }
};
struct ArrayVal {
    const uint32_t* a;
    SF_DEV uint32_t operator()(int i) const { return a[i]; }
    SF_DEV void load8(int i, uint32_t* f) const {
        uint4 x, y;
        load8_u32(a, i, x, y);
        f[0] = x.x; f[1] = x.y; f[2] = x.z; f[3] = x.w;
        f[4] = y.x; f[5] = y.y; f[6] = y.z; f[7] = y.w;
    }
};
struct FlagVal {
    const uint8_t* f;
    SF_DEV uint32_t operator()(int i) const { return f[i]; }
    SF_DEV void load8(int i, uint32_t* o) const {
        const uint2 w = *reinterpret_cast<const uint2*>(f + i);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[q] = (w.x >> (8 * q)) & 0xffu;
            o[4 + q] = (w.y >> (8 * q)) & 0xffu;
        }
    }
};

// Water prefix of the gathered slots (k_gather left the tile counts in
// wpart) and the slots grouped by phase (water first, each group in slot
// order) so that a pass over one phase can give every XCD an equal
// contiguous share of it.
// With scanned = 0, wpart holds the raw per-tile counts and each block sums
// the counts before its tile, and all of them, itself (a few thousand tiles
// at most), instead of waiting for a one-block scan launch.
__global__ __launch_bounds__(kBlock) void k_wprefix_plist(const uint8_t* __restrict__ wflag, int n,
                                                          const uint32_t* __restrict__ wpart, int scanned,
                                                          uint32_t* __restrict__ wpre, uint32_t* __restrict__ plist) {
    uint32_t nwater, part;
    if (scanned) {
        nwater = wpart[gridDim.x];
        part = wpart[blockIdx.x];
    } else {
        block_sums(wpart, 0u, blockIdx.x, gridDim.x, part, nwater);
    }
    tile_prefix(FlagVal{wflag}, n, (int)blockIdx.x, part, wpre, [&](int i, uint32_t pre, uint32_t water) {
        plist[water ? pre : nwater + (uint32_t)i - pre] = (uint32_t)i;
    });
}
int set_direct_tiles(int tiles) {
    const int prev = g_direct_tiles;
    g_direct_tiles = tiles < 0 ? kDirectTiles : tiles;
    return prev;
}
void launch_gather_wprefix(const GatherArgs& a, int n, uint32_t* wpart, uint32_t* wpre, uint32_t* plist,
                           hipStream_t s) {
    const int nb = wprefix_blocks(n + 1);
    k_gather<<<nb, kBlock, 0, s>>>(a, n, wpart);
    const int scanned = nb > g_direct_tiles ? 1 : 0;
    if (scanned) k_rs_scan2<<<1, 1024, 0, s>>>(wpart, nb, 0, wpart + nb);
    k_wprefix_plist<<<nb, kBlock, 0, s>>>(a.wflag, n, wpart, scanned, wpre, plist);
}

// ---------------------------------------------------------------------------
// Incremental stable sort.  The particles are in the previous step's sorted
// order and most keep their cell, so the stable sort by the new keys is
// rebuilt without radix passes: within each new cell the particles are
// ordered by current index, so a particle's output slot is
//   newStart(c) + #(stayers of c before it) + #(newcomers of c before it),
// where stayers of c are the particles of c's old block whose key is still c
// (a prefix over "key unchanged" flags counts them) and newcomers are the
// few particles whose key changed to c (collected per cell, in the order the
// atomics leave; a newcomer's rank is the count of smaller indices in its
// cell's list, which that order does not change).
// newStart(c) = stayers in cells < c + newcomers in cells < c.
// The result is the stable sort's, element for element.
__global__ __launch_bounds__(kBlock) void k_inc_flag(const uint32_t* __restrict__ key, const uint32_t* __restrict__ old,
                                                     int n, uint32_t* __restrict__ in_count,
                                                     uint32_t* __restrict__ slot) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    if (c != old[i]) slot[i] = atomicAdd(in_count + c, 1u);
}
__global__ __launch_bounds__(kBlock) void k_inc_fill(const uint32_t* __restrict__ key, const uint32_t* __restrict__ old,
                                                     int n, const uint32_t* __restrict__ in_start,
                                                     const uint32_t* __restrict__ slot, uint32_t* __restrict__ in_list) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    if (c != old[i]) in_list[in_start[c] + slot[i]] = (uint32_t)i;
}
__global__ __launch_bounds__(kBlock) void k_inc_place(const uint32_t* __restrict__ key, const uint32_t* __restrict__ old,
                                                      int n, const uint32_t* __restrict__ old_start,
                                                      const uint32_t* __restrict__ stay_pre,
                                                      const uint32_t* __restrict__ in_start,
                                                      const uint32_t* __restrict__ in_list, uint32_t* __restrict__ perm,
                                                      uint32_t* __restrict__ sorted_keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    const uint32_t oa = old_start[c], ob = old_start[c + 1];
    const uint32_t sa = stay_pre[oa];
    const uint32_t s = in_start[c], e = in_start[c + 1];
    uint32_t rank_in = 0;
    for (uint32_t q = s; q < e; ++q) rank_in += in_list[q] < (uint32_t)i ? 1u : 0u;
    uint32_t rank_stay;
    if (c == old[i]) {
        rank_stay = stay_pre[i] - sa;
    } else {
        const uint32_t u = (uint32_t)i;
        rank_stay = u < oa ? 0u : (u >= ob ? stay_pre[ob] - sa : stay_pre[u] - sa);
    }
    const uint32_t pos = sa + s + rank_stay + rank_in;
    perm[pos] = (uint32_t)i;
    sorted_keys[pos] = c;
}

void launch_incremental_sort(const uint32_t* key, const uint32_t* old_key, const uint32_t* old_start, int n,
                             uint32_t C, const IncSortBuffers& b, uint32_t* perm, uint32_t* sorted_keys,
                             bool flags_done, hipStream_t s) {
    if (n <= 0) return;
    const int gn = (n + kBlock - 1) / kBlock;
    // in_count is zero here: the cell-start kernel after every sort clears it
    if (!flags_done) k_inc_flag<<<gn, kBlock, 0, s>>>(key, old_key, n, b.in_count, b.slot);
    launch_vprefix2(StayVal{key, old_key}, n, b.stay_pre, ArrayVal{b.in_count}, (int)C, b.in_start, b.part, s);
    k_inc_fill<<<gn, kBlock, 0, s>>>(key, old_key, n, b.in_start, b.slot, b.in_list);
    k_inc_place<<<gn, kBlock, 0, s>>>(key, old_key, n, old_start, b.stay_pre, b.in_start, b.in_list, perm,
                                       sorted_keys);
}

// in-range neighbour counts: r^2 = dx*dx + dy*dy + dz*dz in FP32, left to
// right without contraction, compared with h^2
__global__ __launch_bounds__(kBlock) void k_nbr_count(SFConst k, const float4* __restrict__ pm,
                                                      const uint16_t* __restrict__ scene,
                                                      const uint32_t* __restrict__ cs, uint32_t* __restrict__ cnt,
                                                      uint32_t* __restrict__ cand, int n) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 pi = pm[i];
    const uint32_t cb = scene ? (uint32_t)scene[i] * k.C : 0u;
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    uint32_t c = 0, cc = 0;
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t s0 = cs[cb + c0], e = cs[cb + c1];
        cc += e - s0;
        for (uint32_t j = s0; j < e; ++j) {
#pragma clang fp contract(off)
            // written out here (not via helpers) so the pragma governs every
            // operation and nothing is fused into an FMA
            const float4 pj = pm[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
            const float r2 = (xx + yy) + zz;
            c += r2 < k.h2 ? 1u : 0u;
        }
    });
    cnt[i] = c;
    cand[i] = cc;
}
void launch_nbr_count(const SFConst& k, const float4* pm, const uint16_t* scene, const uint32_t* cs, uint32_t* cnt,
                      uint32_t* cand, int n, hipStream_t s) {
    if (n <= 0) return;
    k_nbr_count<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, pm, scene, cs, cnt, cand, n);
}

}  // namespace sf
