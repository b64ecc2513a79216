// sf_solver.hip -- the per-step solver passes of the coupled SPH-DEM seepage
// step on gfx950 (one thread per particle over the cell-sorted arrays; each
// thread walks its 27-cell neighbourhood as 9 contiguous z-ranges).
//
// Pass structure (dependencies of cuda_wcsph_sf_solver.cpp:44-177):
//   P1 density + EOS            <- ComputeSFDensity_CUDA, ComputeSFPressure(ByTait)_CUDA
//   P2 avg flow / voidage / sat <- ComputeSFAvgFlow_CUDA
//   P3 sand voidage | water adhesion
//                               <- ComputeSFSandVoidage_CUDA, Compute(Multi)SFWaterAdhesionForces_CUDA
//   P4 sand momentum + colour   <- Compute(Multi)SFSandLinearMomentum_CUDA, ComputeSFWetSandColor_CUDA
//   P5 water momentum           <- ComputeSFWaterLinearMomentum_CUDA
//   CFL (optional), advect      <- ComputeSubTimeStepsByCFL, CudaSFParticles::Advect,
//                                  SFWaterBoundaryConstrain_CUDA
// ExtraForces (acc = g) is folded into P4/P5, which write the final acc.
#include <type_traits>

#include "sf_internal.h"
#include "sf_nlist.h"

namespace sf {

// Tait EOS B((rho/rho0)^7 - 1) (cuda_wcsph_sf_solver_gpu.cuh:34) evaluated as
// B * u * (1 + x + ... + x^6) with u = (rho - rho0) / rho0, x = rho / rho0,
// which avoids the cancellation of x^7 - 1 near rest density.
SF_DEV float eos(const SFConst& k, float rho) {
    float p;
    if (k.tait) {
        const float x = rho / k.rho0;
        const float u = (rho - k.rho0) / k.rho0;
        const float s = 1.f + x * (1.f + x * (1.f + x * (1.f + x * (1.f + x * (1.f + x)))));
        p = k.stiff * u * s;
    } else {
        p = k.stiff * (rho - k.rho0);  // ComputeSFPressure_CUDA (cuda_sph_sf_solver_gpu.cuh:497)
    }
    if (p < 0.f) p *= k.neg_scale;
    return p;
}

// First cell key of slot i's scene (batched handles), 0 otherwise.
template <bool B>
SF_DEV uint32_t scene_base(const SFConst& k, const Fields& f, int i) {
    if constexpr (B) return (uint32_t)ldo(f.scene, (uint32_t)i) * k.C;
    else return 0u;
}

// Every S-th slot j of [s, e) whose test(j, pm[j]) holds goes to the list,
// in slot order.  Four candidates per round: four independent loads in
// flight per lane (the loop is otherwise bound by one dependent global load
// per candidate) from one base address with immediate offsets, and one
// list-full check per round (a list at CAP or more is flushed first; a round
// then leaves at most CAP + 3 entries, which the list has room for).
// The pairs reach apply in the same order whenever the flushes fall.
template <uint32_t S, int NB, int CAP, class Test, class Fetch, class Apply>
SF_DEV void filter_push(NList<NB, CAP>& L, const float4* __restrict__ pm, uint32_t s, uint32_t e, Test&& test,
                        Fetch&& fetch, Apply&& apply) {
    static_assert(CAP >= 4, "a round pushes up to four candidates");
    uint32_t j = s;
    for (; j + 3u * S < e; j += 4u * S) {
        // slot offsets stay below 2^31 bytes, so base + 16 q S does not wrap
        const float4* q = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(pm) + j * 16u);
        const float4 a = q[0], b = q[S], c = q[2u * S], d = q[3u * S];
        const bool ka = test(j, a), kb = test(j + S, b), kc = test(j + 2u * S, c), kd = test(j + 3u * S, d);
        if (__ballot(L.cnt >= CAP)) nl_flush2(L, fetch, apply);
        L.slot[L.cnt * NB] = j;
        L.cnt += ka ? 1 : 0;
        L.slot[L.cnt * NB] = j + S;
        L.cnt += kb ? 1 : 0;
        L.slot[L.cnt * NB] = j + 2u * S;
        L.cnt += kc ? 1 : 0;
        L.slot[L.cnt * NB] = j + 3u * S;
        L.cnt += kd ? 1 : 0;
    }
    for (; j < e; j += S) {
        const float4 a = ldo(pm, j);
        nl_push2(L, test(j, a), j, fetch, apply);
    }
}

// Phase-index ranges of one phase over the 9 neighbour columns, in the order
// of for_columns: a column's slot range [s, e) holds the water candidates
// pph[wpre[s], wpre[e]) and the sand candidates pph[nw + s - wpre[s],
// nw + e - wpre[e]), each in slot order (the cell table holds both ends).
template <class F>
SF_DEV void for_phase_ranges(const SFConst& k, const Fields& f, uint32_t cb, int cx, int cy, int cz, bool water,
                             uint32_t nw, F&& fn) {
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint2 a = ldo(f.ct, cb + c0), b = ldo(f.ct, cb + c1);
        const uint32_t t0 = water ? a.y : nw + (a.x - a.y), t1 = water ? b.y : nw + (b.x - b.y);
        if (t0 < t1) fn(t0, t1);
    });
}

SF_DEV float sq3(float d) { return d * d * d; }

// ---------------------------------------------------------------------------
// Wave-staged neighbour columns.  The 64 lanes of a wave hold consecutive
// slots of the cell-sorted order, so they sit in a few consecutive cells of
// one grid column (x, y), rarely two or three, and their 27-cell
// neighbourhoods overlap almost entirely: lane by lane, the eight or so lanes
// of one cell would load exactly the same candidates.  Here the lanes that
// share a column (a "segment") take, for each of the 9 neighbour columns, the
// union of their z-ranges -- one contiguous slot range, from the first lane's
// range start to the last lane's range end -- load it once with coalesced
// loads into the wave's LDS stage, and each lane then walks its own window
// of the union from LDS.  Each wave owns its stage; the lanes of one wave
// order their LDS accesses with wave-scope fences only.
constexpr int kStage = 128;  // candidates staged per chunk and wave

SF_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <class T>
SF_DEV T lane_value(T v, int lane) {
    return __builtin_amdgcn_readlane(v, lane);
}

// For the lanes with `active` set: fn(j, staged_j) over every slot j of the
// lane's 9 column ranges (cells z-1..z+1), column by column.  prep(pm[j]) is what gets staged (once per wave, not per
// lane).  need(nw, ns), given the water / sand counts of a column's union
// range, may skip the whole column.  Must be called by all lanes of the wave
// (uniform control flow).
// With L > 1 lanes per particle (adjacent lanes, same cell), lane sub of a
// particle takes the slots j = first + sub + L m of each of its ranges
// [first, last), in increasing j, however the ranges fall into stage chunks
// (the same split as a per-lane walk of the ranges).
// With Water set, the walk covers only the water candidates: phase indices
// [wpre[s], wpre[e]) of each column range, staged from pph.  range(a, b) sees
// each of the lane's own 9 cell-table pairs (range start, range end).
template <int L = 1, bool Water = false, class Prep, class Need, class Fn, class Range>
SF_DEV void staged_walk(const SFConst& k, const Fields& f, float4* __restrict__ stage, bool active, uint32_t cb,
                        int cx, int cy, int cz, Prep&& prep, Need&& need, Fn&& fn, Range&& range) {
    static_assert((L & (L - 1)) == 0, "lanes per particle: a power of two");
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t sub = (uint32_t)lane & (uint32_t)(L - 1);
    const uint32_t col = cb + (uint32_t)((cx * k.gy + cy) * k.gz);
    const int zlo = cz > 0 ? cz - 1 : 0;
    const int zhi = cz + 1 < k.gz ? cz + 1 : k.gz - 1;
    uint64_t todo = __ballot(active);
    while (todo) {
        const int lead = __builtin_ctzll(todo);
        const uint32_t lcol = lane_value(col, lead);
        const bool mine = active && col == lcol;
        const uint64_t seg = __ballot(mine);
        todo &= ~seg;
        const int last = 63 - __builtin_clzll(seg);
        const int sx = lane_value(cx, lead), sy = lane_value(cy, lead);
        const uint32_t scb = lane_value(cb, lead);
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = sx + dx;
            if (x < 0 || x >= k.gx) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int y = sy + dy;
                if (y < 0 || y >= k.gy) continue;
                const uint32_t base = scb + (uint32_t)((x * k.gy + y) * k.gz);
                uint2 a = make_uint2(0u, 0u), b = make_uint2(0u, 0u);
                if (mine) {
                    a = ldo(f.ct, base + (uint32_t)zlo);
                    b = ldo(f.ct, base + (uint32_t)zhi + 1u);
                    range(a, b);
                }
                const uint32_t ax = Water ? a.y : a.x, bx = Water ? b.y : b.x;
                const uint32_t us = lane_value(ax, lead), ue = lane_value(bx, last);
                if (us >= ue) continue;
                const uint32_t nw = lane_value(b.y, last) - lane_value(a.y, lead);
                if (!need(nw, (lane_value(b.x, last) - lane_value(a.x, lead)) - nw)) continue;
                const float4* __restrict__ src = Water ? f.pph : f.pm;
                for (uint32_t u = us; u < ue; u += (uint32_t)kStage) {
                    const uint32_t m = ue - u < (uint32_t)kStage ? ue - u : (uint32_t)kStage;
                    for (uint32_t q = (uint32_t)lane; q < m; q += 64u) stage[q] = prep(ldo(src, u + q));
                    wave_sync();
                    const uint32_t wlo = ax > u ? ax : u, whi = bx < u + m ? bx : u + m;
                    constexpr uint32_t S = (uint32_t)L;
                    uint32_t j = wlo + ((ax + sub - wlo) & (S - 1u));  // j = ax + sub (mod L)
                    for (; j + 3u * S < whi; j += 4u * S) {
                        const float4 q0 = stage[j - u], q1 = stage[j + S - u], q2 = stage[j + 2u * S - u],
                                     q3 = stage[j + 3u * S - u];
                        fn(j, q0);
                        fn(j + S, q1);
                        fn(j + 2u * S, q2);
                        fn(j + 3u * S, q3);
                    }
                    for (; j < whi; j += S) fn(j, stage[j - u]);
                    wave_sync();
                }
            }
        }
    }
}

// The filtered walk of the water momentum pass: every candidate j of the
// particle's 9 column ranges with test(pm[j]) goes to the lane's list, lane
// sub (of S per particle) taking j = s + sub + S m of each range [s, e) in
// increasing j, exactly as filter_push over the ranges.  The filter is bound
// by the texture address unit, which costs the same per wave load
// instruction whatever it carries, so where the wave's ranges overlap it
// loads them once: when every particle of the wave sits in one grid column
// (most waves of an ordered region), the lanes walk the same 9 columns, and
// a column whose union range (first lane's start to last lane's end) fits
// KS candidates is staged in LDS with coalesced loads (one per lane up to
// 64) and filtered from there.  Other columns, and waves spread over several columns, filter from
// memory per lane.  Either way a lane's list gets the same entries in the
// same order, so which way a column goes does not change the sums.  Must be
// called by all lanes of the wave.
template <uint32_t S, uint32_t KS, int NB, int CAP, class Test, class Fetch, class Apply>
SF_DEV void column_filter(const SFConst& k, const Fields& f, float4* __restrict__ stage, bool active, uint32_t cb,
                          int cx, int cy, int cz, NList<NB, CAP>& L, Test&& test, Fetch&& fetch, Apply&& apply) {
    static_assert(CAP >= 4, "a round pushes up to four candidates");
    static_assert(KS <= 128, "at most two staged candidates per lane");
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t sub = (uint32_t)lane % S;
    const uint64_t act = __ballot(active);
    if (!act) return;
    const int lead = __builtin_ctzll(act), last = 63 - __builtin_clzll(act);
    const uint32_t col = cb + (uint32_t)((cx * k.gy + cy) * k.gz);
    const bool one = !__ballot(active && col != lane_value(col, lead));  // wave-uniform
    if (!one && !active) return;
    // lanes without a particle follow the lead's columns (with empty ranges)
    const int wx = active ? cx : lane_value(cx, lead), wy = active ? cy : lane_value(cy, lead);
    const int wz = active ? cz : lane_value(cz, lead);
    for_columns(k, wx, wy, wz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t rs = active ? ldo(f.ct, cb + c0).x : 0u, re = active ? ldo(f.ct, cb + c1).x : 0u;
        if (one) {
            const uint32_t us = lane_value(rs, lead), ue = lane_value(re, last);
            if (us >= ue) return;
            if (ue - us <= KS) {
                for (uint32_t q = (uint32_t)lane; q < ue - us; q += 64u) stage[q] = ldo(f.pm, us + q);
                wave_sync();
                uint32_t j = rs + sub;
                for (; j + 3u * S < re; j += 4u * S) {
                    const float4 a = stage[j - us], b = stage[j + S - us], c = stage[j + 2u * S - us],
                                 d = stage[j + 3u * S - us];
                    const bool ka = test(a), kb = test(b), kc = test(c), kd = test(d);
                    if (__ballot(L.cnt >= CAP)) nl_flush2(L, fetch, apply);
                    L.slot[L.cnt * NB] = j;
                    L.cnt += ka ? 1 : 0;
                    L.slot[L.cnt * NB] = j + S;
                    L.cnt += kb ? 1 : 0;
                    L.slot[L.cnt * NB] = j + 2u * S;
                    L.cnt += kc ? 1 : 0;
                    L.slot[L.cnt * NB] = j + 3u * S;
                    L.cnt += kd ? 1 : 0;
                }
                for (; j < re; j += S) nl_push2(L, test(stage[j - us]), j, fetch, apply);
                wave_sync();
                return;
            }
        }
        if (active)
            filter_push<S>(L, f.pm, rs + sub, re, [&](uint32_t, const float4& pj) { return test(pj); }, fetch, apply);
    });
}

// Lanes per particle of the interface passes (P2, P3), whose work sits at the
// water-sand interface (PassCfg::lanes): several lanes per particle let the
// few thousand waves of a small handle hide load latency.
template <class L>
void with_lanes(int lanes, L&& l) {
    switch (lanes) {
    case 4: l(std::integral_constant<int, 4>{}); break;
    case 2: l(std::integral_constant<int, 2>{}); break;
    default: l(std::integral_constant<int, 1>{}); break;
    }
}
// Sums x over the LP lanes of a particle (adjacent lanes, fixed butterfly).
template <int LP>
SF_DEV float lane_sum(float x) {
#pragma unroll
    for (int o = 1; o < LP; o <<= 1) x += __shfl_xor(x, o);
    return x;
}

// Host-side launch dispatch over (batched?, workgroup size, list capacity);
// the choices are kNbChoices x kCapChoices.
template <int V>
using IC = std::integral_constant<int, V>;
template <class L>
void dispatch(const Fields& f, PassCfg c, L&& l) {
    auto with_cap = [&](auto B, auto NB) {
        if (c.cap == 8) l(B, NB, IC<8>{});
        else l(B, NB, IC<16>{});
    };
    auto with_nb = [&](auto B) {
        if (c.nb == 64) with_cap(B, IC<64>{});
        else if (c.nb == 128) with_cap(B, IC<128>{});
        else with_cap(B, IC<256>{});
    };
    if (f.scene) with_nb(std::true_type{});
    else with_nb(std::false_type{});
}

// ---------------------------------------------------------------------------
// Interface lists.  The avg-flow and voidage/adhesion passes (P2, P3) need
// each particle's neighbours of the other phase only.  Most particles have
// none in their 9 column ranges (the cell table gives each range's water and
// sand counts).  For those the empty sums' results are written directly
// (write_no_interface); the rest are listed, water and sand apart
// (list_interface), and P2 and P3 run over the lists.  In a step the density
// pass does both, since it already loads every water particle's 9 cell-table
// pairs; k_interface does the same for a lone avg-flow run.  The counters
// start at zero: the water momentum pass (after P2 and P3 have read them)
// clears them for the next step, and a lone density or avg-flow run clears
// them first (launch_zero_counts).  The list order the atomics leave does
// not matter: each listed particle's sums are its own lanes' work in
// candidate order.

// What P2 then P3 leave for particle i with empty sums; rpi = (rho, p).
// Later passes read the sand's voidage / saturation (drag, capillary bridges)
// and the water's wt or records; the rest is output only and written here when the
// fields are read directly (full_empty) or taped (den4), else on request.
SF_DEV void write_empty_outputs(const Fields& f, int i, bool sand) {
    f.flow[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    f.adh[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!sand) {
        f.es[i] = make_float2(1.f, 0.f);
        f.xf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}
// Water particle i's (rho~, p / rho~^2) from (rho, p) and its voidage eps:
// wt, or the neighbour records with pi = (position, mass) (Fields::rec0).
SF_DEV void write_water_terms(const Fields& f, int i, const float4& pi, float2 rpi, float eps) {
    const float rt = rpi.x / fmaxf(kEps, eps);
    const float pt = rpi.y / fmaxf(kEps, rt * rt);
    if (!f.rec || f.den4) f.wt[i] = make_float2(rt, pt);
    if (f.rec) {
        const float4 v = ldo(f.vr, (uint32_t)i);
        f.rec0[i] = make_float4(pi.x, pi.y, pi.z, rt);
        f.rec1[i] = make_float4(v.x, v.y, v.z, pt);
    }
}
SF_DEV void write_no_interface(const Fields& f, int i, bool sand, float2 rpi, const float4& pi) {
    if (sand) {
        f.es[i] = make_float2(0.f, 0.f);
    } else {
        write_water_terms(f, i, pi, rpi, 1.f);
    }
    if (f.full_empty || f.den4) write_empty_outputs(f, i, sand);
    if (f.den4) {
        reinterpret_cast<float*>(f.den4)[4 * i] = 0.f;
        reinterpret_cast<float*>(f.den4)[4 * i + 1] = 0.f;
    }
}
// Appends phase index t of every listed lane to its phase's list (water at
// ilist[0, nw), sand at ilist[nw, n)): one atomic per wave and phase, each
// listed lane taking its rank among the wave's.  A list never grows past its
// phase's size (that would take a missed clear of the counters).  Must be
// called by all lanes of the wave.
SF_DEV void list_interface(const Fields& f, bool listed, bool sand, uint32_t t, uint32_t nw, uint32_t n) {
    const int lane = (int)(threadIdx.x & 63u);
    for (int ph = 0; ph < 2; ++ph) {
        const bool mine = listed && sand == (ph == 1);
        const uint64_t m = __ballot(mine);
        if (m == 0) continue;
        const int leader = __builtin_ctzll(m);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(f.icount + ph, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, leader, 64);
        const uint32_t q = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (mine && q < (ph ? n - nw : nw)) f.ilist[(ph ? nw : 0u) + q] = t;
    }
}

// ---------------------------------------------------------------------------
// P1: water density (self counted twice, :466 + loop) and pressure; sand gets
// rho_sand and p = 0.  Only water candidates are walked (sand adds nothing);
// branch-free: max(h^2 - r^2, 0) vanishes outside the support.  The kDenL
// lanes of a particle split both the candidates and the boundary particles.
// The boundary particles of the 27 cells around (cx, cy, cz) for lane sub
// of L lanes per particle: fn(b) for j = s + sub + L m of each column's
// range [s, e), columns in for_columns order and j increasing (the order of
// a plain walk).  The nine ranges, then each range's first candidate of the
// lane, are loaded together: a lane near a wall has about one candidate per
// column, so a column-by-column walk waited on three dependent loads per
// column.
template <int L, class Fn>
SF_DEV void boundary_walk(const SFConst& k, const Fields& f, int cx, int cy, int cz, uint32_t sub, Fn&& fn) {
    const int zlo = cz > 0 ? cz - 1 : 0;
    const int zhi = cz + 1 < k.gz ? cz + 1 : k.gz - 1;
    uint32_t s[9], e[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int x = cx + q / 3 - 1, y = cy + q % 3 - 1;
        s[q] = 0u;
        e[q] = 0u;
        if (x >= 0 && x < k.gx && y >= 0 && y < k.gy) {
            const uint32_t base = (uint32_t)((x * k.gy + y) * k.gz);
            s[q] = ldo(f.bcell, base + (uint32_t)zlo) + sub;
            e[q] = ldo(f.bcell, base + (uint32_t)zhi + 1u);
        }
    }
    float4 b0[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) b0[q] = s[q] < e[q] ? ldo(f.bpv, s[q]) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        if (s[q] < e[q]) fn(b0[q]);
        for (uint32_t j = s[q] + (uint32_t)L; j < e[q]; j += (uint32_t)L) fn(ldo(f.bpv, j));
    }
}

constexpr int kDenL = 2;  // lanes per water particle in the density pass
template <bool B, int NB>
__global__ __launch_bounds__(NB) void k_density(SFConst k, Fields f, int n) {
    __shared__ float4 stage_lds[(NB / 64) * kStage];
    float4* stage = stage_lds + (threadIdx.x >> 6) * kStage;
    constexpr uint32_t PB = NB / kDenL;  // water particles per block
    const uint32_t nw = *f.nwater, ns = (uint32_t)n - nw;
    const PhaseBlock pb = phase_block((nw + PB - 1) / PB, (ns + NB - 1) / NB);
    if (pb.phase < 0) return;
    if (pb.phase == 1) {  // sand: rho_sand, p = 0; listed if a column range holds water
        const uint32_t t = pb.blk * NB + threadIdx.x;
        const bool valid = t < ns;
        int i = 0;
        uint32_t other = 0;
        if (valid) {
            i = (int)ldo(f.plist, nw + t);
            // every grain has rho_sand and p = 0; the passes use the constant,
            // so the field is written here only when it is read directly or
            // taped, else on request (launch_fill_empty)
            if (f.full_empty || f.den4) f.rp[i] = make_float2(k.rho_s, 0.f);
            const float4 p = ldo(f.pph, nw + t);
            if (f.rec) f.rec0[i] = p;
            const uint32_t cb = scene_base<B>(k, f, i);
            int cx, cy, cz;
            cell_of(k, p.x, p.y, p.z, cx, cy, cz);
            for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
                other += ldo(f.ct, cb + c1).y - ldo(f.ct, cb + c0).y;
            });
        }
        const bool listed = valid && other != 0u;
        list_interface(f, listed, true, nw + t, nw, (uint32_t)n);
        if (valid) f.iface[t] = listed ? 1 : 0;
        if (valid) {  // every grain's velocity and saturation by phase index (P2 sets the listed ones' .w)
            const float4 v = ldo(f.vr, (uint32_t)i);
            f.vq[nw + t] = make_float4(v.x, v.y, v.z, 0.f);
        }
        if (valid && !listed) write_no_interface(f, i, true, make_float2(k.rho_s, 0.f), make_float4(0.f, 0.f, 0.f, 0.f));
        return;
    }
    const uint32_t t = pb.blk * PB + threadIdx.x / kDenL;
    const bool lead = (threadIdx.x % kDenL) == 0;
    const bool water = t < nw;
    const int i = water ? (int)ldo(f.plist, t) : 0;
    const float4 pi = water ? ldo(f.pm, (uint32_t)i) : make_float4(0.f, 0.f, 0.f, -1.f);
    const uint32_t cb = water ? scene_base<B>(k, f, i) : 0u;
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    float sf = 0.f, sb = 0.f;
    uint32_t other = 0;  // sand in the particle's 9 column ranges
    auto add = [&](uint32_t, const float4& pj) {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        // h^2 - r^2 as one fused chain (three FMAs, one rounding each)
        const float d = fmaxf(fmaf(-dz, dz, fmaf(-dy, dy, fmaf(-dx, dx, k.h2))), 0.f);
        sf += pj.w * sq3(d);
    };
    auto count = [&](uint2 a, uint2 b) { other += (b.x - a.x) - (b.y - a.y); };
    // A wave whose particles all sit in one grid column stages the water
    // candidates of its 9 columns once; a wave over several columns would
    // stage once per column it holds, one after another, so its lanes walk
    // their own ranges instead.  Both give each lane the same candidates in
    // the same order (staged_walk), hence the same sums.
    const uint64_t act = __ballot(water);
    const uint32_t col = cb + (uint32_t)((cx * k.gy + cy) * k.gz);
    const bool one = !__ballot(water && col != lane_value(col, act ? __builtin_ctzll(act) : 0));
    if (one) {
        staged_walk<kDenL, true>(
            k, f, stage, water, cb, cx, cy, cz, [](float4 p) { return p; }, [](uint32_t, uint32_t) { return true; },
            add, count);
    } else if (water) {
        const uint32_t sub = threadIdx.x % kDenL;
        for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
            const uint2 a = ldo(f.ct, cb + c0), b = ldo(f.ct, cb + c1);
            count(a, b);
            uint32_t j = a.y + sub;
            for (; j + 3u * kDenL < b.y; j += 4u * kDenL) {
                const float4 q0 = ldo(f.pph, j), q1 = ldo(f.pph, j + kDenL), q2 = ldo(f.pph, j + 2u * kDenL),
                             q3 = ldo(f.pph, j + 3u * kDenL);
                add(j, q0);
                add(j + kDenL, q1);
                add(j + 2u * kDenL, q2);
                add(j + 3u * kDenL, q3);
            }
            for (; j < b.y; j += kDenL) add(j, ldo(f.pph, j));
        });
    }
    if (water && ldo(f.bnear, cell_key(k, cx, cy, cz))) {  // boundary particles within the 27 cells
        const uint32_t sub = threadIdx.x % kDenL;
        auto badd = [&](const float4& b) {
            const float dx = pi.x - b.x, dy = pi.y - b.y, dz = pi.z - b.z;
            const float d = fmaxf(k.h2 - (dx * dx + dy * dy + dz * dz), 0.f);
            sb += fabsf(b.w) * sq3(d);
        };
        // column by column (this pass keeps its registers low), four
        // candidates' loads in flight at a time, summed in increasing j
        for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
            const uint32_t be = ldo(f.bcell, c1);
            uint32_t j = ldo(f.bcell, c0) + sub;
            for (; j + 3u * kDenL < be; j += 4u * kDenL) {
                const float4 b0 = ldo(f.bpv, j), b1 = ldo(f.bpv, j + kDenL), b2 = ldo(f.bpv, j + 2u * kDenL),
                             b3 = ldo(f.bpv, j + 3u * kDenL);
                badd(b0);
                badd(b1);
                badd(b2);
                badd(b3);
            }
            for (; j < be; j += kDenL) badd(ldo(f.bpv, j));
        });
    }
#pragma unroll
    for (int o = 1; o < kDenL; o <<= 1) {
        sf += __shfl_xor(sf, o);
        sb += __shfl_xor(sb, o);
    }
    const bool listed = water && lead && other != 0u;
    list_interface(f, listed, false, t, nw, (uint32_t)n);
    if (!water || !lead) return;
    f.wl[t] = listed ? 1 : 0;
    const float rho = pi.w * k.w0 + k.w_coef * (sf + k.rho0 * sb);
    const float2 rpi = make_float2(rho, eos(k, rho));
    f.rp[i] = rpi;
    if (!listed) {
        write_no_interface(f, i, false, rpi, pi);
    } else {
        f.vq[t] = ldo(f.vr, (uint32_t)i);
        f.wq[t] = make_float4(rpi.x, rpi.y, 0.f, 0.f);  // voidage: avg flow
    }
}
void launch_density(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s) {
    if (n <= 0) return;
    dispatch(f, c, [&](auto B, auto NB, auto) {
        k_density<decltype(B)::value, decltype(NB)::value><<<phase_grid(n * kDenL, NB), (int)NB, 0, s>>>(k, f, n);
    });
}

// ---------------------------------------------------------------------------
// k_interface: the interface lists and the empty sums' results, as the
// density pass leaves them, for a lone avg-flow run.
__global__ void k_zero_counts(uint32_t* __restrict__ c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}
template <bool B>
__global__ __launch_bounds__(kBlock) void k_interface(SFConst k, Fields f, int n) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = t < (uint32_t)n;
    const uint32_t nw = *f.nwater;
    const bool sand_i = t >= nw;
    int i = 0;
    uint32_t other = 0;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) {
        i = (int)ldo(f.plist, t);
        pi = ldo(f.pph, t);
        if (f.rec && sand_i) f.rec0[i] = pi;
        const uint32_t cb = scene_base<B>(k, f, i);
        int cx, cy, cz;
        cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
        for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
            const uint2 a = ldo(f.ct, cb + c0), b = ldo(f.ct, cb + c1);
            const uint32_t w = b.y - a.y;
            other += sand_i ? w : (b.x - a.x) - w;
        });
    }
    const bool listed = valid && other != 0u;
    list_interface(f, listed, sand_i, t, nw, (uint32_t)n);
    if (valid && !sand_i) f.wl[t] = listed ? 1 : 0;
    if (valid && sand_i) f.iface[t - nw] = listed ? 1 : 0;
    if (valid && !listed) write_no_interface(f, i, sand_i, ldo(f.rp, (uint32_t)i), pi);
}
__global__ __launch_bounds__(kBlock) void k_fill_empty(SFConst k, Fields f, int n) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint32_t)n) return;
    const uint32_t nw = *f.nwater;
    const bool sand = t >= nw;
    const int i = (int)ldo(f.plist, t);
    if (sand) f.rp[i] = make_float2(k.rho_s, 0.f);
    if (sand ? ldo(f.iface, t - nw) : ldo(f.wl, t)) return;
    write_empty_outputs(f, i, sand);
}
void launch_fill_empty(const SFConst& k, const Fields& f, int n, hipStream_t s) {
    if (n > 0) k_fill_empty<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, f, n);
}
void launch_zero_counts(const Fields& f, hipStream_t s) { k_zero_counts<<<1, 64, 0, s>>>(f.icount); }
void launch_interface(const SFConst& k, const Fields& f, int n, hipStream_t s) {
    if (f.scene) k_interface<true><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, f, n);
    else k_interface<false><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, f, n);
}

template <int NB, int LP>
struct PhaseSlot {
    int i;         // the particle's slot
    uint32_t t;    // its phase index (plist[t] = i)
    uint32_t sub;  // its lane among LP
    bool sand;
};
// The listed particle of this thread (false: none).
template <int NB, int LP>
SF_DEV bool phase_slot(const Fields& f, uint32_t nw, PhaseSlot<NB, LP>& ps) {
    constexpr uint32_t PB = NB / LP;  // particles per block
    const uint32_t cw = f.icount[0], cs = f.icount[1];
    const PhaseBlock pb = phase_block((cw + PB - 1) / PB, (cs + PB - 1) / PB);
    if (pb.phase < 0) return false;
    ps.sand = pb.phase == 1;
    const uint32_t q = pb.blk * PB + threadIdx.x / LP;
    ps.sub = threadIdx.x % LP;
    if (q >= (ps.sand ? cs : cw)) return false;
    ps.t = ldo(f.ilist, (ps.sand ? nw : 0u) + q);
    ps.i = (int)ldo(f.plist, ps.t);
    return true;
}
// enough blocks for every particle listed
inline int phase_slot_grid(int n, int lanes, int nb) { return phase_grid(n * lanes, nb); }

// ---------------------------------------------------------------------------
// P2 (ComputeSFAvgFlow_CUDA): water i -> voidage 1 - sum_sand V W and the
// average sand velocity; sand i -> average water velocity and saturation.
// (Sand voidage from this pass is overwritten by P3 before anything reads it,
// so a sand particle only needs its water neighbours, a water particle only
// its sand neighbours.)
// Water particles also store wt = (rho / max(eps, voidage), p / max(eps, that^2)),
// the per-particle factors the water momentum pass needs from every neighbour.
// Candidates are filtered four at a time into the lane's list; the flush
// fetches each kept neighbour's fields (by phase index u: every neighbour
// within h is listed, see Fields::vq) ahead of its use.
template <bool B, int NB, int CAP, int LP>
__global__ __launch_bounds__(NB) void k_avg_flow(SFConst k, Fields f, int n) {
    SF_NLIST(L);
    const uint32_t nw = *f.nwater;
    PhaseSlot<NB, LP> ps;
    if (!phase_slot(f, nw, ps)) return;
    const int i = ps.i;
    const uint32_t sub = ps.sub;
    const bool sand_i = ps.sand;
    const float4 pi = ldo(f.pm, (uint32_t)i);
    const uint32_t cb = scene_base<B>(k, f, i);
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    float sandv = 0.f, vsum = 0.f;
    v3 av = mk3(0.f, 0.f, 0.f);
    const float inv_rho_s = k.rho_s != 0.f ? 1.f / k.rho_s : 0.f;
    auto in_range = [&](uint32_t, const float4& pj) {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        return r2 < k.h2;
    };
    struct Nbr {
        float4 pj, vj;
        float rho;
    };
    // sand i: water neighbours (rho_j needed); water i: sand neighbours
    // neighbours within h are listed: their fields by phase index
    auto fetch = [&](uint32_t u) { return Nbr{ldo(f.pph, u), ldo(f.vq, u), sand_i ? ldo(f.wq, u).x : 0.f}; };
    auto pair = [&](uint32_t, const Nbr& q) {
        const float4 pj = q.pj;
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float w = k.w_coef * sq3(k.h2 - r2);
        const float vw = (sand_i ? pj.w * frcp(q.rho) : -pj.w * inv_rho_s) * w;
        if (!sand_i) sandv += vw;
        av += xyz(q.vj) * vw;
        vsum += vw;
    };
    for_phase_ranges(k, f, cb, cx, cy, cz, sand_i, nw, [&](uint32_t t0, uint32_t t1) {
        filter_push<LP>(L, f.pph, t0 + sub, t1, in_range, fetch, pair);
    });
    nl_flush2(L, fetch, pair);
    sandv = lane_sum<LP>(sandv);
    vsum = lane_sum<LP>(vsum);
    av = mk3(lane_sum<LP>(av.x), lane_sum<LP>(av.y), lane_sum<LP>(av.z));
    if (sub != 0) return;
    const v3 fl = vsum != 0.f ? av * (1.f / vsum) : mk3(0.f, 0.f, 0.f);
    f.flow[i] = f4(fl, 0.f);
    if (f.den4) reinterpret_cast<float*>(f.den4)[4 * i] = vsum;
    const float eps_i = 1.f - sandv;
    f.es[i] = make_float2(eps_i, sand_i ? vsum : 0.f);
    if (sand_i) {  // saturation, for P3 and (with the velocity) P4
        reinterpret_cast<float*>(f.wq)[4 * ps.t + 3] = vsum;
        reinterpret_cast<float*>(f.vq)[4 * ps.t + 3] = vsum;
    }
    if (!sand_i) {
        write_water_terms(f, i, pi, ldo(f.rp, (uint32_t)i), eps_i);
        reinterpret_cast<float*>(f.wq)[4 * ps.t + 2] = eps_i;
    }
}
void launch_avg_flow(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s) {
    if (n <= 0) return;
    dispatch(f, c, [&](auto B, auto NB, auto CAP) {
        with_lanes(c.lanes, [&](auto LP) {
            k_avg_flow<decltype(B)::value, decltype(NB)::value, decltype(CAP)::value, decltype(LP)::value>
                <<<phase_slot_grid(n, LP, NB), (int)NB, 0, s>>>(k, f, n);
        });
    });
}

// ---------------------------------------------------------------------------
// P3: sand i -> voidage from water neighbours; water i -> adhesion force.
// Same walk as P2.
template <bool B, int NB, int CAP, int LP>
__global__ __launch_bounds__(NB) void k_voidage_adhesion(SFConst k, Fields f, int n) {
    SF_NLIST(L);
    const uint32_t nw = *f.nwater;
    PhaseSlot<NB, LP> ps;
    if (!phase_slot(f, nw, ps)) return;
    const int i = ps.i;
    const uint32_t sub = ps.sub;
    const float4 pi = ldo(f.pm, (uint32_t)i);
    const uint32_t cb = scene_base<B>(k, f, i);
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    auto in_range = [&](uint32_t, const float4& pj) {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        return r2 < k.h2;
    };
    struct Nbr {
        float4 pj;
        float2 x;  // sand i: (rho_j, voidage_j) of a water j; water i: saturation of a sand j in y
    };
    // neighbours within h are listed: water (rho, voidage) or sand saturation
    // (in y) by phase index
    auto fetch = [&](uint32_t u) {
        const float4 w = ldo(f.wq, u);
        return Nbr{ldo(f.pph, u), ps.sand ? make_float2(w.x, w.z) : make_float2(0.f, w.w)};
    };
    if (ps.sand) {
        float num = 0.f, den = 0.f;
        auto pair = [&](uint32_t, const Nbr& q) {
            const float4 pj = q.pj;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float vw = (pj.w * frcp(q.x.x)) * (k.w_coef * sq3(k.h2 - r2));
            num += q.x.y * vw;
            den += vw;
        };
        for_phase_ranges(k, f, cb, cx, cy, cz, true, nw, [&](uint32_t t0, uint32_t t1) {
            filter_push<LP>(L, f.pph, t0 + sub, t1, in_range, fetch, pair);
        });
        nl_flush2(L, fetch, pair);
        num = lane_sum<LP>(num);
        den = lane_sum<LP>(den);
        if (sub != 0) return;
        f.es[i].x = den != 0.f ? num / den : 0.f;
        f.adh[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (f.den4) reinterpret_cast<float*>(f.den4)[4 * i + 1] = den;
        return;
    }
    const float mi = pi.w;
    const float inv_rho_s = k.rho_s != 0.f ? 1.f / k.rho_s : 0.f;
    float vs = 0.f;
    v3 F = mk3(0.f, 0.f, 0.f);
    auto pair = [&](uint32_t, const Nbr& q) {
        const float4 pj = q.pj;
        const v3 dp = mk3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
        const float r2 = dot(dp, dp);
        const float mj = -pj.w;
        vs += (mj * inv_rho_s) * (k.w_coef * sq3(k.h2 - r2));
        const float l = fsqrt(r2);
        const float A = adhesion_A(k, l);
        if (A != 0.f) {
            const float G = bezier(k.gadh, q.x.y);
            F += dp * (G * mi * mj * A * frcp(fmaxf(kEps, l)));
        }
    };
    for_phase_ranges(k, f, cb, cx, cy, cz, false, nw, [&](uint32_t t0, uint32_t t1) {
        filter_push<LP>(L, f.pph, t0 + sub, t1, in_range, fetch, pair);
    });
    nl_flush2(L, fetch, pair);
    vs = lane_sum<LP>(vs);
    F = mk3(lane_sum<LP>(F.x), lane_sum<LP>(F.y), lane_sum<LP>(F.z));
    if (sub != 0) return;
    f.adh[i] = f4(F, 0.f);
    const float4 avg = vs != 0.f ? f4(F * (1.f / vs), 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    f.xf[i] = avg;
    // what a grain's buoyancy sums need of this particle (sand momentum
    // pass): m / rho~ and m p / max(eps, rho_sand rho~), both zero unless the
    // voidage lies in (0, 1]
    const float4 w = ldo(f.wq, ps.t);  // (rho, p, voidage)
    float a = 0.f, g = 0.f;
    if (w.z > 0.f && w.z <= 1.f) {
        const float dj = w.x * frcp(fmaxf(kEps, w.z));
        a = mi * frcp(dj);
        g = mi * w.y * frcp(fmaxf(kEps, k.rho_s * dj));
    }
    f.xq[ps.t] = make_float4(avg.x, avg.y, avg.z, g);
    f.wbq[ps.t] = make_float4(pi.x, pi.y, pi.z, a);
    if (f.den4) reinterpret_cast<float*>(f.den4)[4 * i + 1] = vs;
}
void launch_voidage_adhesion(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s) {
    if (n <= 0) return;
    dispatch(f, c, [&](auto B, auto NB, auto CAP) {
        with_lanes(c.lanes, [&](auto LP) {
            k_voidage_adhesion<decltype(B)::value, decltype(NB)::value, decltype(CAP)::value, decltype(LP)::value>
                <<<phase_slot_grid(n, LP, NB), (int)NB, 0, s>>>(k, f, n);
        });
    });
}

// ---------------------------------------------------------------------------
// P4: sand linear momentum (DEM + capillary + buoyancy + adhesion + boundary
// DEM + walls + drag), acc = g + 2 (f + df) / m, then the wet-sand colour.
//
// Sand-sand forces vanish beyond r_c = 2R + s_rupture (DEM needs overlap,
// the bridge H < s_rupture), and r_c < h / 2, so along each axis a grain only
// reaches its own cell and the one neighbour on its nearer side: the contact
// search walks 2 x 2 x 2 cells instead of 27 (same non-zero terms).

constexpr int kSmL = 2;  // lanes per sand particle (as in the water momentum pass)
template <bool B, int NB, int CAP>
__global__ __launch_bounds__(NB) void k_sand_momentum(SFConst k, Fields f, int n) {
    SF_NLIST(L);
    constexpr uint32_t PB = NB / kSmL;  // sand particles per block
    const uint32_t nw = *f.nwater, ns = (uint32_t)n - nw;
    const PhaseBlock pb = phase_block(0u, (ns + PB - 1) / PB);
    const uint32_t t = pb.blk * PB + threadIdx.x / kSmL;
    const uint32_t sub = threadIdx.x % kSmL;  // this lane's share of the particle's candidates
    if (pb.phase < 0 || t >= ns) return;
    const int i = (int)ldo(f.plist, nw + t);
    const uint32_t self = nw + t;  // this particle's phase index
    const float4 pi = ldo(f.pm, (uint32_t)i);
    if (pi.w >= 0.f) return;
    const uint32_t cb = scene_base<B>(k, f, i);
    if constexpr (B) {  // the scene's friction, capillary curve and drag coefficient
        const SceneConst sc = f.sct[ldo(f.scene, (uint32_t)i)];
        k.tan2 = sc.tan2;
        k.tanphi = sc.tanphi;
        k.cd = sc.cd;
        k.gcap = sc.gcap;
    }
    const float mi = -pi.w;
    const v3 xi = xyz(pi);
    const v3 vi = xyz(ldo(f.vr, (uint32_t)i));
    const float rho_i = k.rho_s;  // the density pass's value for every grain
    const float2 esi = ldo(f.es, (uint32_t)i);
    const float sat_i = esi.y;
    const float twoR = 2.f * k.R;
    const float cap_r = twoR + k.cap_srup;
    const float cap_r2 = cap_r * cap_r;
    const float vol_i = mi / fmaxf(kEps, rho_i);
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    v3 fsum = mk3(0.f, 0.f, 0.f);
    float vw = 0.f;

    // sand-sand: DEM contact or liquid bridge
    struct SandNbr {
        float4 pj, vj;  // vj.w: saturation
    };
    // candidates are phase indices u (pph[u] = pm[plist[u]], vq[u] = the
    // velocity and saturation of slot plist[u])
    auto fetch_sand = [&](uint32_t u) { return SandNbr{ldo(f.pph, u), ldo(f.vq, u)}; };
    auto contact = [&](uint32_t, const SandNbr& q) {
        const float4 pj = q.pj;
        const v3 dji = mk3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        const float dist = fsqrt(dot(dji, dji));
        if (dist < twoR) {
            fsum += dem_contact(dji, dist, xyz(q.vj) - vi, twoR, k.kn, k.ks, k.tan2);
        } else {
            const float H = dist - twoR;
            if (H > 0.f && H < k.cap_srup) fsum += capillary(k, dji, dist, H, (sat_i + q.vj.w) * 0.5f);
        }
    };
    {
        const float rcm = cap_r + 1e-3f * k.h;  // conservative against FP32 cell edges
        int xa, xb, ya, yb, za, zb;
        near_span(pi.x, k.lox, k.h, cx, k.gx, rcm, xa, xb);
        near_span(pi.y, k.loy, k.h, cy, k.gy, rcm, ya, yb);
        near_span(pi.z, k.loz, k.h, cz, k.gz, rcm, za, zb);
        for (int x = xa; x <= xb; ++x)
            for (int y = ya; y <= yb; ++y) {
                const uint32_t base = cb + (uint32_t)((x * k.gy + y) * k.gz);
                const uint2 a = ldo(f.ct, base + (uint32_t)za), b = ldo(f.ct, base + (uint32_t)zb + 1u);
                // the sand candidates of the range, as phase indices
                filter_push<kSmL>(L, f.pph, nw + (a.x - a.y) + sub, nw + (b.x - b.y), [&](uint32_t u, const float4& pj) {
                    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    return u != self && r2 < cap_r2;
                }, fetch_sand, contact);
            }
        nl_flush2(L, fetch_sand, contact);
    }
    // water neighbours: buoyancy (ComputeSFSandBuoyancyForces) + adhesion
    // (ComputeSFSandAdhesionTerm)
    struct WaterNbr {
        float4 pa;  // position, m_j / rho~_j
        float4 xg;  // avgAdhesion, m_j p_j / max(eps, rho_sand rho~_j)
    };
    // a water neighbour within h has this grain in its columns, so it is
    // listed and the voidage/adhesion pass wrote its record at its phase
    // index u (the two factors zero where rho~_j is undefined: voidage
    // outside (0, 1])
    auto fetch_water = [&](uint32_t u) { return WaterNbr{ldo(f.wbq, u), ldo(f.xq, u)}; };
    auto water = [&](uint32_t, const WaterNbr& q) {
        const float4 pj = q.pa;
        const v3 dji = mk3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        const float r2 = dot(dji, dji);
        const float w = k.w_coef * sq3(k.h2 - r2);
        vw += pj.w * w;
        const float gs = spiky_s(k, fsqrt(r2));
        // gradW(x_i - x_j) = gs * (x_i - x_j) = -gs * dji
        fsum += dji * (mi * q.xg.w * gs);
        fsum += xyz(q.xg) * (w * vol_i);
    };
    // a grain with no water in its 9 column ranges (the density pass flagged
    // the others) has nothing to walk
    if (ldo(f.iface, t)) {
        for_phase_ranges(k, f, cb, cx, cy, cz, true, nw, [&](uint32_t t0, uint32_t t1) {
            filter_push<kSmL>(L, f.pph, t0 + sub, t1, [&](uint32_t, const float4& pj) {
                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                return r2 < k.h2;
            }, fetch_water, water);
        });
    }
    nl_flush2(L, fetch_water, water);
    if (k.blabel1 && ldo(f.bnear, cell_key(k, cx, cy, cz))) {  // ComputeSFUniRadiusDemBoundaryForces: label 1 only
        for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
            const uint32_t be = ldo(f.bcell, c1);
            for (uint32_t j = ldo(f.bcell, c0) + sub; j < be; j += kSmL) {
                const float4 b = ldo(f.bpv, j);
                if (b.w >= 0.f) continue;
                fsum += dem_force(xi - xyz(b), -vi, twoR, k.kn, k.ks, k.tan2);
            }
        });
    }
#pragma unroll
    for (int o = 1; o < kSmL; o <<= 1) {
        fsum.x += __shfl_xor(fsum.x, o);
        fsum.y += __shfl_xor(fsum.y, o);
        fsum.z += __shfl_xor(fsum.z, o);
        vw += __shfl_xor(vw, o);
    }
    if (sub != 0) return;
    fsum += world_walls(k, xi, vi);
    v3 df = mk3(0.f, 0.f, 0.f);
    const float eps_i = esi.x;
    if (eps_i < 1.f && eps_i > 0.f) {
        const float cd = k.multi ? ldo(f.cda, (uint32_t)i).x : k.cd;
        const float ome = 1.f - eps_i;
        const float dc = cd * (eps_i * eps_i * eps_i) / (ome * ome);
        if (dc != 0.f) {
            const v3 fl = xyz(ldo(f.flow, (uint32_t)i));
            df = (fl - vi) * (k.grav_abs * k.rho0 * (eps_i * eps_i) / dc * mi / rho_i);
        }
    }
    const float4 drag = vw != 0.f ? f4(df * (1.f / vw), 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    f.xf[i] = drag;
    if (f.rec) f.rec1[i] = drag;
    if (f.den4) reinterpret_cast<float*>(f.den4)[4 * i + 2] = vw;
    const v3 a = mk3(k.gvx, k.gvy, k.gvz) + (fsum + df) * (2.f / mi);
    f.acc[i] = f4(a, 0.f);
    // ComputeSFWetSandColor_CUDA (:952-977)
    float4 c = ldo((const float4*)f.cs, (uint32_t)i);
    if (c.w < sat_i) {
        c.x = k.dryr + (k.wetr - k.dryr) * sat_i;
        c.y = k.dryg + (k.wetg - k.dryg) * sat_i;
        c.z = k.dryb + (k.wetb - k.dryb) * sat_i;
        c.w = sat_i;
        f.cs[i] = c;
    }
}
void launch_sand_momentum(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s) {
    if (n <= 0) return;
    dispatch(f, c, [&](auto B, auto NB, auto CAP) {
        k_sand_momentum<decltype(B)::value, decltype(NB)::value, decltype(CAP)::value>
            <<<phase_grid(n * kSmL, NB), (int)NB, 0, s>>>(k, f, n);
    });
}

// ---------------------------------------------------------------------------
// P5: water linear momentum (pressure + artificial viscosity + seepage +
// boundary pressure/viscosity), acc = g + a - F_adh / m.  kWmL adjacent lanes
// share a particle: each filters every kWmL-th candidate of the ranges into
// its own list (column_filter, which stages the candidates in LDS where the
// wave's ranges overlap), and the partial sums are combined with shuffles.
constexpr int kWmL = 4;
constexpr uint32_t kWmStage = 64;  // longest union range staged per wave
// A neighbour's data in the water momentum pass: r0 = (position, w) with
// w > 0 for water, w = -mass for sand.  WmRecords reads the two records
// (Fields::rec0; every water particle has mass f.mw); WmSlots reads the
// slot arrays pm (w = mass), wt and vr, and a grain's xf on its own.
struct WmRecords {
    float4 r0, r1;
    SF_DEV static WmRecords load(const Fields& f, uint32_t j) { return {ldo(f.rec0, j), ldo(f.rec1, j)}; }
    SF_DEV float water_mass(const Fields& f) const { return f.mw; }
    SF_DEV float2 wt() const { return make_float2(r0.w, r1.w); }
    SF_DEV v3 vel() const { return xyz(r1); }
    SF_DEV v3 drag(const Fields&, uint32_t) const { return xyz(r1); }
};
struct WmSlots {
    float4 r0;
    float2 w;
    float4 v;
    SF_DEV static WmSlots load(const Fields& f, uint32_t j) { return {ldo(f.pm, j), ldo(f.wt, j), ldo(f.vr, j)}; }
    SF_DEV float water_mass(const Fields&) const { return r0.w; }
    SF_DEV float2 wt() const { return w; }
    SF_DEV v3 vel() const { return xyz(v); }
    SF_DEV v3 drag(const Fields& f, uint32_t j) const { return xyz(ldo(f.xf, j)); }
};
template <bool B, class Nbr, int NB, int CAP>
__global__ __launch_bounds__(NB) void k_water_momentum(SFConst k, Fields f, int n) {
    SF_NLIST(L);
    __shared__ float4 stage_lds[(NB / 64) * kWmStage];
    float4* stage = stage_lds + (threadIdx.x >> 6) * kWmStage;
    // P2 and P3 have read the interface lists: the next density pass lists
    // into counters that start at zero
    if (blockIdx.x == 0 && threadIdx.x < 2) f.icount[threadIdx.x] = 0u;
    constexpr uint32_t PB = NB / kWmL;  // water particles per block
    const uint32_t nw = *f.nwater;
    const PhaseBlock pb = phase_block((nw + PB - 1) / PB, 0u);
    if (pb.phase < 0) return;  // the whole block
    const uint32_t t = pb.blk * PB + threadIdx.x / kWmL;
    const uint32_t sub = threadIdx.x % kWmL;  // this lane's share of the particle's candidates
    // lanes past the last particle take part in the staged walk without a range
    const int i = t < nw ? (int)ldo(f.plist, t) : 0;
    const Nbr self = Nbr::load(f, (uint32_t)i);
    const float4 pi = self.r0;  // position; w > 0 for water
    const bool active = t < nw && pi.w >= 0.f;
    const uint32_t cb = scene_base<B>(k, f, i);
    const float mi = self.water_mass(f);
    const v3 vi = self.vel();
    const float2 wti = self.wt();  // (rho~_i, p_i / rho~_i^2)
    const float p_i = ldo(f.rp, (uint32_t)i).y;
    const float rt_i = wti.x;
    const float inv_rt_i = frcp(rt_i);
    const float bpress = -k.rho0 * (p_i / fmaxf(kEps, k.rho0 * k.rho0));
    const float bvisc = -k.bnu / (2.f * k.rho0);
    int cx, cy, cz;
    cell_of(k, pi.x, pi.y, pi.z, cx, cy, cz);
    v3 a = mk3(0.f, 0.f, 0.f);
    auto fetch = [&](uint32_t j) { return Nbr::load(f, j); };
    auto pair = [&](uint32_t j, const Nbr& q) {
        const float4 pj = q.r0;
        const v3 d = mk3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
        const float r2 = dot(d, d);
        const float rs = frsq(r2);
        const float r = r2 * rs;
        const float hr = k.h - r;
        const float gs = r >= kEps ? k.g_coef * hr * hr * rs : 0.f;  // spiky gradient / r
        if (pj.w > 0.f) {  // water j: ComputeSFWaterPressureTerm + ArtificialViscosity
            const float2 wtj = q.wt();
            const float mj = q.water_mass(f);
            float s = -mj * (wti.y + wtj.y);
            const v3 dv = vi - q.vel();
            const float dd = dot(dv, d);
            const float pij = -k.nu * frcp(rt_i + wtj.x) * (dd * frcp(r2 + kEps));
            s += dd < 0.f ? -mj * pij : 0.f;
            a += d * (s * gs);
        } else {  // sand j: ComputeSFWaterSeepageTerm
            const float mj = -pj.w;
            const float rho_j = k.rho_s;  // the density pass gives every grain rho_sand
            const float w = k.w_coef * sq3(k.h2 - r2);
            a += d * (-mj * p_i * frcp(fmaxf(kEps, rt_i * rho_j)) * gs);
            a += q.drag(f, j) * (-w * inv_rt_i);
        }
    };
    // the particle's own slot passes too: d = 0 makes its term exactly zero
    // (the spiky factor is 0 below r = eps), and acc adds g afterwards
    auto in_range = [&](const float4& pj) {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        return r2 < k.h2;
    };
    column_filter<kWmL, kWmStage>(k, f, stage, active, cb, cx, cy, cz, L, in_range, fetch, pair);
    nl_flush2(L, fetch, pair);
    if (!active) return;
    if (ldo(f.bnear, cell_key(k, cx, cy, cz))) {  // boundary particles within the 27 cells
        boundary_walk<kWmL>(k, f, cx, cy, cz, sub, [&](const float4& b) {
            const v3 d = mk3(pi.x - b.x, pi.y - b.y, pi.z - b.z);
            const float r2 = dot(d, d);
            if (r2 >= k.h2) return;
            const float rs = frsq(r2);
            const float r = r2 * rs;
            const float hr = k.h - r;
            const float gs = r >= kEps ? k.g_coef * hr * hr * rs : 0.f;
            const float vb = fabsf(b.w);
            float s = vb * bpress;  // ComputeBoundaryPressure (densityi = rho0, :944)
            const float dd = dot(vi, d);
            if (dd < 0.f) s += -vb * k.rho0 * (bvisc * (dd * frcp(r2 + kEps)));  // ComputeBoundaryViscosity
            a += d * (s * gs);
        });
    }
#pragma unroll
    for (int o = 1; o < kWmL; o <<= 1) {
        a.x += __shfl_xor(a.x, o);
        a.y += __shfl_xor(a.y, o);
        a.z += __shfl_xor(a.z, o);
    }
    if (sub != 0) return;
    // adhesion: P3 wrote it for the listed water; the rest has none
    const v3 ad = ldo(f.wl, t) ? xyz(ldo(f.adh, (uint32_t)i)) : mk3(0.f, 0.f, 0.f);
    const v3 acc = mk3(k.gvx, k.gvy, k.gvz) + a - ad * (1.f / mi);
    f.acc[i] = f4(acc, 0.f);
}
void launch_water_momentum(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s) {
    if (n <= 0) return;
    dispatch(f, c, [&](auto B, auto NB, auto CAP) {
        constexpr bool b = decltype(B)::value;
        constexpr int nb = decltype(NB)::value, cap = decltype(CAP)::value;
        if (f.rec) k_water_momentum<b, WmRecords, nb, cap><<<phase_grid(n * kWmL, NB), nb, 0, s>>>(k, f, n);
        else k_water_momentum<b, WmSlots, nb, cap><<<phase_grid(n * kWmL, NB), nb, 0, s>>>(k, f, n);
    });
}

// ---------------------------------------------------------------------------
// CFL sub-step count (ComputeSubTimeStepsByCFL, cuda_wcsph_sf_solver.cu:35-58),
// computed on the device so the step needs no host round trip.
constexpr int kCflBlocks = 1024;
int cfl_partials(int) { return kCflBlocks; }

__global__ __launch_bounds__(kBlock) void k_cfl_partial(const float4* __restrict__ acc, int n, float* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ float sm[kBlock / 64];
    float m = 0.f;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float4 a = acc[i];
        const float xx = a.x * a.x, yy = a.y * a.y, zz = a.z * a.z;
        m = fmaxf(m, (xx + yy) + zz);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = sm[0];
        for (int q = 1; q < kBlock / 64; ++q) r = fmaxf(r, sm[q]);
        part[blockIdx.x] = r;
    }
}
__global__ __launch_bounds__(kBlock) void k_cfl_final(const float* __restrict__ part, int np, CflOut* out, float h,
                                                      float render_interval, float m, float c, float scale,
                                                      float params_dt, FrameDev* frame) {
#pragma clang fp contract(off)
    __shared__ float sm[kBlock / 64];
    float r = 0.f;
    for (int q = threadIdx.x; q < np; q += kBlock) r = fmaxf(r, part[q]);
    for (int o = 32; o > 0; o >>= 1) r = fmaxf(r, __shfl_down(r, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mr = sm[0];
        for (int q = 1; q < kBlock / 64; ++q) mr = fmaxf(mr, sm[q]);
        const float amax = sqrtf(mr);
        const float fmax_ = amax * m;
        const float dts = 0.4f * h / c;
        const float dtf = 0.25f * sqrtf(h * m / fmax_);
        const float inner = dtf < dts ? dtf : dts;
        const float lim = scale * inner;
        const float dt = params_dt < lim ? params_dt : lim;
        FrameDev fr;
        if (frame) fr = *frame;
        const float ri = (frame && (fr.flags & kFrameAdaptive)) ? fr.remaining : render_interval;
        const float q = ri / dt;
        uint32_t k = q >= 4294967295.f ? 0xffffffffu : (q >= 1.f ? (uint32_t)q : 1u);
        CflOut o;
        o.dt = dt;
        o.nsub = k;
        o.amax = amax;
        o.ri = ri;
        *out = o;
        if (frame) {
            // the host loop's updates, in its float arithmetic (main.cpp:246,
            // cuda_sf_system.cu:95)
            const float step = ri / (float)k;
            if (fr.flags & kFrameElapsed) fr.elapsed = fr.elapsed + step;
            if (fr.flags & kFrameAdaptive) fr.remaining = ri - step;
            fr.count += 1;
            *frame = fr;
        }
    }
}
void launch_cfl(const SFConst& k, const float4* acc, int n, float* partials, CflOut* out, float render_interval,
                float sph_mass, float speed_of_sound, float scale, float params_dt, FrameDev* frame, hipStream_t s) {
    int nb = (n + kBlock - 1) / kBlock;
    if (nb > kCflBlocks) nb = kCflBlocks;
    if (nb < 1) nb = 1;
    k_cfl_partial<<<nb, kBlock, 0, s>>>(acc, n, partials);
    k_cfl_final<<<1, kBlock, 0, s>>>(partials, nb, out, k.h, render_interval, sph_mass, speed_of_sound, scale,
                                      params_dt, frame);
}

// ---------------------------------------------------------------------------
// advect: sand damping (AccDampingForSand), symplectic Euler, then the water
// box clamp and label-1 boundary push (SFWaterBoundaryConstrain_CUDA).
SF_DEV float sgnf(float v) { return (float)((0.f < v) - (v < 0.f)); }
__global__ __launch_bounds__(kBlock) void k_advect(SFConst k, AdvectArgs a, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float dt = a.dt_dev ? a.dt_dev->ri / (float)a.dt_dev->nsub : a.dt;
    const float4 p = a.pm_in[i];
    const float4 v = a.vr_in[i];
    float4 ac = a.acc[i];
    if (p.w < 0.f) {
        const float qx = ac.x * (v.x + 0.5f * dt * ac.x);
        const float qy = ac.y * (v.y + 0.5f * dt * ac.y);
        const float qz = ac.z * (v.z + 0.5f * dt * ac.z);
        ac.x *= 1.f - k.damping * sgnf(qx);
        ac.y *= 1.f - k.damping * sgnf(qy);
        ac.z *= 1.f - k.damping * sgnf(qz);
        a.acc[i] = ac;  // the reference stores the damped acceleration
    }
    v3 nv = mk3(v.x + dt * ac.x, v.y + dt * ac.y, v.z + dt * ac.z);
    v3 np = mk3(p.x + dt * nv.x, p.y + dt * nv.y, p.z + dt * nv.z);
    if (p.w > 0.f) {
        const v3 pre = np;
        const float r2 = 2.f * k.rw;
        if (np.x > k.hix - r2) { np.x = k.hix - r2; nv.x = fminf(nv.x, 0.f); }
        if (np.x < k.lox + r2) { np.x = k.lox + r2; nv.x = fmaxf(nv.x, 0.f); }
        if (np.y > k.hiy - r2) { np.y = k.hiy - r2; nv.y = fminf(nv.y, 0.f); }
        if (np.y < k.loy + r2) { np.y = k.loy + r2; nv.y = fmaxf(nv.y, 0.f); }
        if (np.z > k.hiz - r2) { np.z = k.hiz - r2; nv.z = fminf(nv.z, 0.f); }
        if (np.z < k.loz + r2) { np.z = k.loz + r2; nv.z = fmaxf(nv.z, 0.f); }
        if (k.blabel1) {
            int cx, cy, cz;
            cell_of(k, pre.x, pre.y, pre.z, cx, cy, cz);
            for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
                const uint32_t be = a.bcell[c1];
                for (uint32_t j = a.bcell[c0]; j < be; ++j) {
                    const float4 b = a.bpv[j];
                    if (b.w >= 0.f) continue;
                    const v3 dp = pre - xyz(b);
                    const float dl = sqrtf(dot(dp, dp));
                    const float ov = dl - r2;
                    if (ov < 0.f) {
                        nv = nv * 0.f;
                        np += dp * (1.f / dl) * (-ov);
                    }
                }
            });
        }
    }
    a.pm_out[i] = f4(np, p.w);
    a.vr_out[i] = f4(nv, v.w);
    if (a.keys_out) {  // the next build's hash and change flag (k_hash, k_inc_flag)
        int cx, cy, cz;
        cell_of(k, np.x, np.y, np.z, cx, cy, cz);
        uint32_t key = cell_key(k, cx, cy, cz);
        if (a.scene) key += (uint32_t)a.scene[i] * k.C;
        a.keys_out[i] = key;
        if (key != a.old_keys[i]) a.slot[i] = atomicAdd(a.in_count + key, 1u);
    }
}
void launch_advect(const SFConst& k, const AdvectArgs& a, int n, hipStream_t s) {
    if (n <= 0) return;
    k_advect<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(k, a, n);
}

}  // namespace sf
