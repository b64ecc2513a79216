// sf_views.hip -- the particle state in the reference's array layout.
//
// CudaSFParticles hands out one device array per field
// (seepage_flows_cuda/include/kiri_pbs_cuda/particle/cuda_sf_particles.cuh:170-190):
// packed float3 (12-byte elements), float, size_t labels, float2.  This
// library keeps the state in 16-byte float4 groups instead (position with a
// signed mass, velocity with radius, colour with maxSaturation, ...).  The
// kernels here copy between the two: out after a step for every view a caller
// holds, in before the next step for the views of persistent fields.  Pure
// copies, so a value that round-trips unchanged keeps its bits.
#include "sf_internal.h"

namespace sf {

namespace {
__device__ inline void put3(float* d, int i, float x, float y, float z) {
    d[3 * i] = x;
    d[3 * i + 1] = y;
    d[3 * i + 2] = z;
}

__global__ void __launch_bounds__(256) k_views_out(ViewArgs a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pm = a.pm[i];
    const bool sand = pm.w < 0.f;
    if (a.pos) put3(a.pos, i, pm.x, pm.y, pm.z);
    if (a.label) a.label[i] = sand ? 1u : 0u;
    if (a.mass) a.mass[i] = fabsf(pm.w);
    if (a.vel || a.radius) {
        const float4 v = a.vr[i];
        if (a.vel) put3(a.vel, i, v.x, v.y, v.z);
        if (a.radius) a.radius[i] = v.w;
    }
    if (a.accv) {
        const float4 v = a.acc[i];
        put3(a.accv, i, v.x, v.y, v.z);
    }
    if (a.col || a.maxsat) {
        const float4 v = a.cs[i];
        if (a.col) put3(a.col, i, v.x, v.y, v.z);
        if (a.maxsat) a.maxsat[i] = v.w;
    }
    if (a.density || a.pressure) {
        const float2 v = a.rp[i];
        if (a.density) a.density[i] = v.x;
        if (a.pressure) a.pressure[i] = v.y;
    }
    if (a.voidage || a.saturation) {
        const float2 v = a.es[i];
        if (a.voidage) a.voidage[i] = v.x;
        if (a.saturation) a.saturation[i] = v.y;
    }
    if (a.avgflow) {
        const float4 v = a.flow[i];
        put3(a.avgflow, i, v.x, v.y, v.z);
    }
    if (a.avgdrag || a.avgadh) {
        // one buffer: avgDrag on sand slots, avgAdhesion on water slots; the
        // other phase's entry of each view is 0
        const float4 v = a.xf[i];
        if (a.avgdrag) put3(a.avgdrag, i, sand ? v.x : 0.f, sand ? v.y : 0.f, sand ? v.z : 0.f);
        if (a.avgadh) put3(a.avgadh, i, sand ? 0.f : v.x, sand ? 0.f : v.y, sand ? 0.f : v.z);
    }
    if (a.adhesion) {
        const float4 v = a.adh[i];
        put3(a.adhesion, i, sand ? 0.f : v.x, sand ? 0.f : v.y, sand ? 0.f : v.z);
    }
    if (a.cda3) {
        const float4 v = a.cda[i];
        put3(a.cda3, i, v.x, v.y, v.z);
    }
    if (a.amc2) a.amc2[i] = a.amc[i];
}

__global__ void __launch_bounds__(256) k_views_in(ViewArgs a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (a.pos || a.mass || a.label) {
        float4 pm = a.pm[i];
        if (a.pos) {
            pm.x = a.pos[3 * i];
            pm.y = a.pos[3 * i + 1];
            pm.z = a.pos[3 * i + 2];
        }
        const bool sand = a.label ? a.label[i] != 0u : pm.w < 0.f;
        const float m = a.mass ? a.mass[i] : fabsf(pm.w);
        pm.w = sand ? -m : m;
        a.pm[i] = pm;
    }
    if (a.vel || a.radius) {
        float4 v = a.vr[i];
        if (a.vel) {
            v.x = a.vel[3 * i];
            v.y = a.vel[3 * i + 1];
            v.z = a.vel[3 * i + 2];
        }
        if (a.radius) v.w = a.radius[i];
        a.vr[i] = v;
    }
    if (a.col || a.maxsat) {
        float4 v = a.cs[i];
        if (a.col) {
            v.x = a.col[3 * i];
            v.y = a.col[3 * i + 1];
            v.z = a.col[3 * i + 2];
        }
        if (a.maxsat) v.w = a.maxsat[i];
        a.cs[i] = v;
    }
    if (a.cda3) {
        float4 v = a.cda[i];
        v.x = a.cda3[3 * i];
        v.y = a.cda3[3 * i + 1];
        v.z = a.cda3[3 * i + 2];
        a.cda[i] = v;
    }
    if (a.amc2) a.amc[i] = a.amc2[i];
}

__global__ void __launch_bounds__(256) k_widen_u32(const uint32_t* in, uint64_t* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}
}  // namespace

void launch_views_out(const ViewArgs& a, int n, hipStream_t s) {
    if (n > 0) k_views_out<<<(n + 255) / 256, 256, 0, s>>>(a, n);
}
void launch_views_in(const ViewArgs& a, int n, hipStream_t s) {
    if (n > 0) k_views_in<<<(n + 255) / 256, 256, 0, s>>>(a, n);
}
void launch_widen_u32(const uint32_t* in, uint64_t* out, int n, hipStream_t s) {
    if (n > 0) k_widen_u32<<<(n + 255) / 256, 256, 0, s>>>(in, out, n);
}

}  // namespace sf
