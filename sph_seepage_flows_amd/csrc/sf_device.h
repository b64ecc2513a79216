// sf_device.h -- device-side math for the gfx950 seepage kernels.
//
// Storage layout (all arrays in the particles' current, cell-sorted order):
//   pm  float4  (x, y, z, +mass for water / -mass for sand)
//   vr  float4  (vx, vy, vz, radius)
//   cs  float4  (r, g, b, maxSaturation)
//   rp  float2  (density, pressure)
//   es  float2  (voidage, saturation)
//   flow/adh/xf/acc float4 (xyz used); xf holds avgAdhesion on water slots
//   and avgDrag on sand slots (each is only ever read from the other phase).
//   bpv float4  (x, y, z, +volume for boundary label 0 / -volume for label 1)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SF_DEV __device__ __forceinline__

namespace sf {

constexpr float kEps = 1e-6f;  // KIRI_EPSILON (kiri_pbs_pch.cuh:52)
constexpr int kBlock = 256;    // 4 wave64s per workgroup

struct v3 {
    float x, y, z;
};
SF_DEV v3 mk3(float x, float y, float z) { return {x, y, z}; }
SF_DEV v3 xyz(float4 a) { return {a.x, a.y, a.z}; }
SF_DEV v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
SF_DEV v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SF_DEV v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
SF_DEV v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
SF_DEV v3 operator*(float s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
SF_DEV v3& operator+=(v3& a, v3 b) {
    a.x += b.x; a.y += b.y; a.z += b.z;
    return a;
}
SF_DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SF_DEV float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Loads through a 32-bit byte offset from a uniform base, so the compiler can
// use the SGPR-base + VGPR-offset addressing form (no 64-bit VALU address
// math per candidate).  Particle indices stay below 2^27 (checked at create).
template <class T>
SF_DEV T ldo(const T* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (j * (uint32_t)sizeof(T)));
}

// Fast reciprocal / square root (1 ulp hardware ops) for the physics; the
// grid hash keeps IEEE division.
SF_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
SF_DEV float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
SF_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Per-scene constants, precomputed on the host in FP32 the way the reference's
// functor constructors compute them (cuda_sph_kernel.cuh:79-118,189-203,
// cuda_sf_utils.cuh:36-57, cuda_dem_solver_common_gpu.cuh:63-65).
struct Bez {
    float py0, py1, px1, py2, b, c, d, e;
    float ipx1, ie;  // 1 / px1, 1 / e
};
struct SFConst {
    // grid (CudaBoundaryParams)
    float lox, loy, loz, h;
    float hix, hiy, hiz, h2;
    int gx, gy, gz;
    uint32_t C;  // gx*gy*gz
    // SPH kernels
    float w_coef, w0, g_coef, adh_coef, ih;  // ih = 1 / h
    // water
    float rho0, stiff, neg_scale, nu, bnu, rw;
    // sand
    float rho_s, R, iR, kn, ks, tan2, tanphi, damping, young, poisson, cd;
    // capillary bridge
    float cap_vb, cap_srup, cap_vb_piR, cap_cos, cap_sin, cap_pi;
    Bez gcap, gadh;
    // gravity, colours
    float gvx, gvy, gvz, grav_abs;
    float dryr, dryg, dryb, wetr, wetg, wetb;
    int tait;      // WCSPH Tait EOS (1) or linear EOS (0)
    int multi;     // MULTI_SF: per-particle cd, fixed adhesion curve
    int blabel1;   // number of boundary particles with label 1 (0 = skip those loops)
};

// Batched scenes: every scene shares the box, grid and the constants above
// except these, which it takes from its own row of a per-scene table (the
// sand-momentum pass is the only reader).  Scene s owns the fixed slot range
// [off[s], off[s+1]) and the cell keys [s*C, (s+1)*C).
struct SceneConst {
    float tan2, tanphi, cd, pad;
    Bez gcap;
    float pad2[2];
};

// ---- grid hash (helper_thrust.cuh:56-63), FP32 and exactly as written -----
SF_DEV int cell_coord(float q, int g) {
    if (!(q >= 0.f)) return 0;
    if (q >= (float)g) return g - 1;
    return (int)q;
}
SF_DEV void cell_of(const SFConst& k, float px, float py, float pz, int& cx, int& cy, int& cz) {
#pragma clang fp contract(off)
    float rx = px - k.lox, ry = py - k.loy, rz = pz - k.loz;
    cx = cell_coord(rx / k.h, k.gx);
    cy = cell_coord(ry / k.h, k.gy);
    cz = cell_coord(rz / k.h, k.gz);
}
SF_DEV uint32_t cell_key(const SFConst& k, int cx, int cy, int cz) {
    return (uint32_t)((cx * k.gy + cy) * k.gz + cz);
}

// The 27-cell neighbourhood as 9 (dx, dy) columns; in each column the cells
// z-1, z, z+1 have consecutive keys, so their particles form one contiguous
// range [cs[first], cs[last + 1]) of the sorted arrays.  Out-of-grid cells are
// skipped, as the reference's sentinel hash does (helper_thrust.cuh:123-128).
template <class F>
SF_DEV void for_columns(const SFConst& k, int cx, int cy, int cz, F&& f) {
    const int zlo = cz > 0 ? cz - 1 : 0;
    const int zhi = cz + 1 < k.gz ? cz + 1 : k.gz - 1;
    for (int dx = -1; dx <= 1; ++dx) {
        const int x = cx + dx;
        if (x < 0 || x >= k.gx) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= k.gy) continue;
            const uint32_t base = (uint32_t)((x * k.gy + y) * k.gz);
            f(base + (uint32_t)zlo, base + (uint32_t)zhi + 1u);
        }
    }
}

// Cells along one axis that can hold a partner within rc < h / 2 of
// coordinate p in cell c: the own cell and the neighbour on the nearer side.
SF_DEV void near_span(float p, float lo, float h, int c, int g, float rc, int& a, int& b) {
    const float cl = lo + (float)c * h;
    a = (c > 0 && p - cl < rc) ? c - 1 : c;
    b = (c < g - 1 && cl + h - p < rc) ? c + 1 : c;
}

// ---- SPH kernels on r^2 (Poly6 cuda_sph_kernel.cuh:79-98, Spiky :100-118) ----
SF_DEV float poly6_r2(const SFConst& k, float r2) {
    if (r2 >= k.h2) return 0.f;
    float d = k.h2 - r2;
    return k.w_coef * d * d * d;
}
// returns the scalar s with gradW(r) = s * r
SF_DEV float spiky_s(const SFConst& k, float r) {
    if (r >= k.h || r < kEps) return 0.f;
    float d = k.h - r;
    return k.g_coef * d * d * frcp(r);
}
// AdhesionAkinci13 (:189-203).  The base -4r^2/h + 6r - 2h equals
// (2/h)(2r - h)(h - r); on h/2 < r <= h both differences are exact in FP32
// (Sterbenz), so the factored form has no cancellation where the base goes
// to zero (the expanded form loses most digits there, and its quarter power
// amplifies that).  The quarter power is two hardware square roots.
SF_DEV float adhesion_base(const SFConst& k, float r) { return 2.f * (2.f * r - k.h) * (k.h - r) * k.ih; }
SF_DEV float adhesion_A(const SFConst& k, float r) {
    if (2.f * r > k.h && r <= k.h) return k.adh_coef * fsqrt(fsqrt(fmaxf(adhesion_base(k, r), 0.f)));
    return 0.f;
}
// QuadraticBezierCoeff (cuda_sf_utils.cuh:36-82)
SF_DEV float bezier(const Bez& g, float sr) {
    if (sr < 0.f) return g.py0;
    if (sr >= 1.f) return g.py2;
    if (sr <= g.px1) {
        float t = sr * g.ipx1, omt = 1.f - t;
        return omt * omt * g.py0 + 2.f * t * omt * g.py1 + t * t * g.py1;
    }
    float t = (g.b + fsqrt(g.c + g.d * (g.px1 - sr))) * g.ie;
    float omt = 1.f - t;
    return omt * omt * g.py1 + 2.f * t * omt * g.py1 + t * t * g.py2;
}

// ComputeDemForces (cuda_dem_solver_common_gpu.cuh:20-51) for a pair known to
// overlap (dist < rij); dist = |dij| > 0.
SF_DEV v3 dem_contact(v3 dij, float dist, v3 vij, float rij, float kn, float ks, float tan2) {
    float pen = rij - dist;
    v3 n = dij * frcp(dist);
    float de = dot(vij, n);
    v3 vt = vij - de * n;
    v3 fn = (kn * pen) * n;
    v3 fs = (-ks) * vt;
    float maxfs = dot(fn, fn) * tan2;
    float fs2 = dot(fs, fs);
    if (fs2 > maxfs) fs = fs * (fsqrt(maxfs) * frsq(fs2));
    return -fn - fs;
}
SF_DEV v3 dem_force(v3 dij, v3 vij, float rij, float kn, float ks, float tan2) {
    float dist = fsqrt(dot(dij, dij));
    if (rij - dist > 0.f) return dem_contact(dij, dist, vij, rij, kn, ks, tan2);
    return mk3(0.f, 0.f, 0.f);
}
// ComputeDemCapillaryForces (:53-90) inside the window 0 < H < s_rupture
SF_DEV v3 capillary(const SFConst& k, v3 dij, float dist, float H, float sr) {
    v3 N = dij * frcp(dist);
    float cc = bezier(k.gcap, sr);
    float R = k.R;
    float d = -H + fsqrt(H * H + k.cap_vb_piR);
    float phi = fsqrt(2.f * H * k.iR * (-1.f + fsqrt(1.f + k.cap_vb_piR * frcp(H * H))));
    float neck = -2.f * k.cap_pi * cc * R * k.cap_cos * frcp(1.f + H * frcp(2.f * d));
    float surf = -2.f * k.cap_pi * cc * R * phi * k.cap_sin;
    return -N * (neck + surf);
}
// ComputeDemWorldBoundaryForces (:132-199); z-low trigger 2*rij as written (:191)
SF_DEV v3 world_walls(const SFConst& k, v3 p, v3 v) {
    const float rij = k.rw + k.R;
    v3 f = mk3(0.f, 0.f, 0.f);
    v3 mv = -v;
    if (p.x > k.hix - rij) f += dem_force(mk3(1.f, 0.f, 0.f) * fabsf(p.x - k.hix), mv, rij, k.kn, k.ks, k.tan2);
    if (p.x < k.lox + rij) f += dem_force(mk3(-1.f, 0.f, 0.f) * fabsf(p.x - k.lox), mv, rij, k.kn, k.ks, k.tan2);
    if (p.y > k.hiy - rij) f += dem_force(mk3(0.f, 1.f, 0.f) * fabsf(p.y - k.hiy), mv, rij, k.kn, k.ks, k.tan2);
    if (p.y < k.loy + rij) f += dem_force(mk3(0.f, -1.f, 0.f) * fabsf(p.y - k.loy), mv, rij, k.kn, k.ks, k.tan2);
    if (p.z > k.hiz - rij) f += dem_force(mk3(0.f, 0.f, 1.f) * fabsf(p.z - k.hiz), mv, rij, k.kn, k.ks, k.tan2);
    if (p.z < k.loz + 2.f * rij) f += dem_force(mk3(0.f, 0.f, -1.f) * fabsf(p.z - k.loz), mv, rij, k.kn, k.ks, k.tan2);
    return f;
}

// Block mapping of a pass over two phases of a slot list: na blocks of phase
// 0, then nb blocks of phase 1.  Each XCD (hardware blocks b with b % 8 = x)
// first runs its contiguous 1/8 share of the phase-0 blocks, then its share
// of the phase-1 blocks; surplus hardware blocks get phase -1 and exit.  The
// grid needs at least na + nb + 16 blocks (phase_grid).
struct PhaseBlock {
    int phase;
    uint32_t blk;  // block within the phase
};
SF_DEV PhaseBlock phase_block(uint32_t na, uint32_t nb) {
    const uint32_t b = blockIdx.x, x = b & 7u, l = b >> 3;
    const uint32_t qa = na >> 3, ra = na & 7u, qb = nb >> 3, rb = nb & 7u;
    const uint32_t ca = qa + (x < ra ? 1u : 0u), cbn = qb + (x < rb ? 1u : 0u);
    if (l < ca) return {0, x * qa + (x < ra ? x : ra) + l};
    if (l - ca < cbn) return {1, x * qb + (x < rb ? x : rb) + (l - ca)};
    return {-1, 0u};
}
inline int phase_grid(int n, int nb) { return (n + nb - 1) / nb + 16; }

}  // namespace sf
