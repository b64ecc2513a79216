// sf_adjoint.h -- reverse (adjoint) passes of one fixed-dt step and the
// helpers of the differentiable window (sf_diff.cpp).
#pragma once
#include "sf_internal.h"

namespace sf {

// Forward values of one recorded step, all in that step's cell-sorted order
// (the order the passes ran in).
struct TapeStep {
    const float4* pm;    // positions (+mass / -mass) after the gather
    const float4* vr;    // velocities
    const float2* rp;    // density, pressure
    const float2* es;    // voidage (water: P2, sand: P3), saturation
    const float2* wt;    // water: rho~, p / rho~^2
    const float4* flow;  // average flow velocity
    const float4* adh;   // water adhesion force
    const float4* xf;    // avgAdhesion (water) / avgDrag (sand)
    const float4* acc;   // acceleration before the sand damping of advect
    const float4* den4;  // (P2 weight sum, P3 denominator, P4 vw, -)
    const uint2* ct;     // cell table of the step
    const float4* bpv;   // boundary particles (static)
    const uint32_t* bcell;
    const uint8_t* bnear;  // cells with boundary particles in their 27-cell neighbourhood
};

// Adjoint accumulators, same order as the step's TapeStep.
struct AdjFields {
    float4* gx;     // d loss / d position   (in: after advect; out: before)
    float4* gv;     // d loss / d velocity
    float4* gacc;   // d loss / d acceleration (before damping)
    float2* grp;    // (density, pressure)
    float2* ges;    // (voidage, saturation)
    float2* gwt;    // (rho~, p / rho~^2)
    float4* gflow;  // average flow velocity (sand)
    float4* gadh;   // adhesion force (water)
    float4* gxf;    // avgAdhesion (water) / avgDrag (sand)
    float4* o4;     // per-pass adjoints of raw sums: P4 (f, vw), P2 sand (av, vsum)
    float2* o2;     // P2 water (sandv, -)
    float4* gth;    // per-slot parameter adjoints (cd, c0, cmc, tan(friction angle)), summed over steps
    float4* o4b;    // P3 water (F, vs): apart from o4, which P4's pairs still read while P3's are formed
    float2* o2b;    // P3 sand (num, den)
};

// The reverse of one step (advect, P5, P4, P3, P2, P1) into a; sat_seed (may
// be null) adds d loss / d saturation of this step.
void launch_adj_step(const SFConst& k, const TapeStep& t, const AdjFields& a, const float* sat_seed, float dt, int n,
                     hipStream_t s);
// Lanes per particle the reverse pair passes use for n particles (1, 2, 4 or 8).
int adj_lanes(int n);
// Undo the step's gather: prev[perm[k]] = cur[k] for position and velocity adjoints.
void launch_adj_scatter(const uint32_t* perm, const float4* gx, const float4* gv, float4* gx_prev, float4* gv_prev,
                        int n, hipStream_t s);
void launch_iota(uint32_t* out, int n, hipStream_t s);
void launch_gather_u32(const uint32_t* perm, const uint32_t* in, uint32_t* out, int n, hipStream_t s);
// dst[id[k] * comps + c] = component c (first `comps`, from `first`) of src[k]
void launch_scatter_by_id(const uint32_t* id, const float* src, int src_stride, int first, int comps, float* dst, int n,
                          hipStream_t s);
// dst[k].xyz (comps 3) or dst[k].x (comps 1) = src[id[k] * comps + c]; src null -> 0
void launch_gather_by_id(const uint32_t* id, const float* src, int comps, float4* dst, int n, hipStream_t s);
void launch_gather_by_id1(const uint32_t* id, const float* src, float* dst, int n, hipStream_t s);
// vr[k].xyz = vel3[id[k]]
void launch_set_vel_by_id(const uint32_t* id, const float* vel3, float4* vr, int n, hipStream_t s);


// Up to four device-to-device copies in one launch (the taped step's
// copies, which as separate copies cost a launch each).  Sizes are in bytes,
// multiples of 4; pointers 16-byte aligned.
struct CopyJob {
    const void* src;
    void* dst;
    size_t bytes;
};
struct CopyJobs {
    CopyJob j[4];
    int count;
};
void launch_copy_jobs(const CopyJobs& c, hipStream_t s);

}  // namespace sf
