// sf_adjoint.hip -- reverse passes of one fixed-dt seepage step on gfx950.
//
// The forward step (sf_solver.hip) maps the cell-sorted state (x, v) through
// P1 density/EOS -> P2 avg flow/voidage/saturation -> P3 sand voidage | water
// adhesion -> P4 sand momentum -> P5 water momentum -> advect.  Its reverse
// runs the same passes backwards over the recorded values of the step:
//   A0 advect, A5 water momentum, A4 sand momentum (local drag/walls, then
//   pairs), A3 voidage/adhesion, A2 avg flow (+ the rho~ factors), A1 density,
// then undoes the step's gather.  Every pair pass is written in gather form:
// thread k adds to its own adjoints the terms of pairs where k is the centre
// (i-side, weighted by k's output adjoint) and of pairs where a neighbour j is
// the centre (j-side, weighted by j's output adjoint); no atomics, so the
// result does not depend on scheduling.  Pair Jacobians come from the generic
// pair terms of sf_dual.h evaluated with dual-number inputs.
//
// Scope: SF mode (one global drag coefficient), no label-1 collider boundary
// particles (the caller checks both).  Piecewise terms (Coulomb cap, clamps,
// damping sign, viscosity switch) are differentiated on the branch the forward
// took.
#include "sf_adjoint.h"
#include "sf_dual.h"

namespace sf {

namespace {

// Reverse passes run one wave per workgroup: small windows (tens of thousands
// of particles) still spread over every CU, and the pair kernels are latency
// bound (dependent loads of neighbour data), not LDS or occupancy bound.
constexpr int kAB = 64;

template <int K>
SF_DEV void contract3(const G3<Dn<K>>& o, float3 ob, float* g) {
#pragma unroll
    for (int q = 0; q < K; ++q) g[q] = ob.x * o.x.d[q] + ob.y * o.y.d[q] + ob.z * o.z.d[q];
}
SF_DEV float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
SF_DEV void add3(float4& a, float x, float y, float z) {
    a.x += x;
    a.y += y;
    a.z += z;
}
SF_DEV void add3(float4& a, v3 b) { add3(a, b.x, b.y, b.z); }

template <class F>
SF_DEV void neighbours(const SFConst& k, const TapeStep& t, float4 p, F&& fn) {
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t e = t.ct[c1].x;
        for (uint32_t j = t.ct[c0].x; j < e; ++j) fn(j);
    });
}
template <class F>
SF_DEV void boundary_neighbours(const SFConst& k, const TapeStep& t, float4 p, F&& fn) {
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t e = t.bcell[c1];
        for (uint32_t j = t.bcell[c0]; j < e; ++j) fn(j);
    });
}

SF_DEV float sgn1(float v) { return (float)((0.f < v) - (v < 0.f)); }

// ---- A0: advect (symplectic Euler, sand damping, water box clamp) ---------
__global__ __launch_bounds__(kAB) void k_adj_advect(SFConst k, TapeStep t, AdjFields a, const float* sat_seed,
                                                       float dt, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 p = t.pm[i], v = t.vr[i], ac = t.acc[i];
    const float4 gx = a.gx[i], gv = a.gv[i];
    float gnp[3] = {gx.x, gx.y, gx.z}, gnv[3] = {gv.x, gv.y, gv.z};
    float dmp[3] = {1.f, 1.f, 1.f};
    if (p.w < 0.f) {
        const float qx = ac.x * (v.x + 0.5f * dt * ac.x);
        const float qy = ac.y * (v.y + 0.5f * dt * ac.y);
        const float qz = ac.z * (v.z + 0.5f * dt * ac.z);
        dmp[0] = 1.f - k.damping * sgn1(qx);
        dmp[1] = 1.f - k.damping * sgn1(qy);
        dmp[2] = 1.f - k.damping * sgn1(qz);
    } else {
        // the clamps of SFWaterBoundaryConstrain on the updated state
        const float r2 = 2.f * k.rw;
        const float pc[3] = {p.x, p.y, p.z}, vc[3] = {v.x, v.y, v.z}, ak[3] = {ac.x, ac.y, ac.z};
        const float hi[3] = {k.hix, k.hiy, k.hiz}, lo[3] = {k.lox, k.loy, k.loz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float nv = vc[c] + dt * ak[c];
            float np = pc[c] + dt * nv;
            float pass = 1.f;  // d tv / d nv through the clamps
            bool fixed = false;
            if (np > hi[c] - r2) {
                np = hi[c] - r2;
                pass *= nv < 0.f ? 1.f : 0.f;
                nv = fminf(nv, 0.f);
                fixed = true;
            }
            if (np < lo[c] + r2) {
                np = lo[c] + r2;
                pass *= nv > 0.f ? 1.f : 0.f;
                nv = fmaxf(nv, 0.f);
                fixed = true;
            }
            if (fixed) gnp[c] = 0.f;
            gnv[c] *= pass;
        }
    }
    float4 ox = make_float4(0.f, 0.f, 0.f, 0.f), ov = ox, oa = ox;
    float* pox = &ox.x;
    float* pov = &ov.x;
    float* poa = &oa.x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float gtot = gnv[c] + dt * gnp[c];  // nv feeds tv and np = x + dt nv
        pox[c] = gnp[c];
        pov[c] = gtot;
        poa[c] = dt * gtot * dmp[c];
    }
    a.gx[i] = ox;
    a.gv[i] = ov;
    a.gacc[i] = oa;
    a.grp[i] = make_float2(0.f, 0.f);
    a.ges[i] = make_float2(0.f, sat_seed ? sat_seed[i] : 0.f);
    a.gwt[i] = make_float2(0.f, 0.f);
    a.gflow[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    a.gadh[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    a.gxf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---- A5: water momentum ----------------------------------------------------
__global__ __launch_bounds__(kAB) void k_adj_water(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const float4 vk = t.vr[i];
    const bool water = pk.w > 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = a.gx[i], gv = a.gv[i];
    float gp = 0.f, grt = 0.f, gpw = 0.f;
    float4 gdr = make_float4(0.f, 0.f, 0.f, 0.f);
    const float3 ak = f3(a.gacc[i]);
    float2 wtk = water ? t.wt[i] : make_float2(0.f, 0.f);
    const float p_k = t.rp[i].y;
    const float4 xfk = t.xf[i];
    using D8 = Dn<8>;
    using D6 = Dn<6>;
    using D5 = Dn<5>;
    neighbours(k, t, pk, [&](uint32_t j) {
        if ((int)j == i) return;
        const float4 pj = t.pm[j];
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        if (dx * dx + dy * dy + dz * dz >= k.h2) return;
        const float4 vj = t.vr[j];
        if (water) {  // i-side: k is the centre
            if (pj.w > 0.f) {
                const float2 wtj = t.wt[j];
                G3<D8> d = {dvar<8>(dx, 0), dvar<8>(dy, 1), dvar<8>(dz, 2)};
                G3<D8> dv = {dvar<8>(vk.x - vj.x, 3), dvar<8>(vk.y - vj.y, 4), dvar<8>(vk.z - vj.z, 5)};
                G3<D8> o = t_water_water(k, d, dv, dvar<8>(wtk.y, 6), D8(wtj.y), dvar<8>(wtk.x, 7), D8(wtj.x), pj.w);
                float g[8];
                contract3(o, ak, g);
                add3(gx, g[0], g[1], g[2]);
                add3(gv, g[3], g[4], g[5]);
                gpw += g[6];
                grt += g[7];
            } else {
                const float4 drj = t.xf[j];
                G3<D5> d = {dvar<5>(dx, 0), dvar<5>(dy, 1), dvar<5>(dz, 2)};
                G3<D5> o = t_water_sand(k, d, dvar<5>(p_k, 3), dvar<5>(wtk.x, 4), lift<D5>(drj.x, drj.y, drj.z),
                                        -pj.w, t.rp[j].x);
                float g[5];
                contract3(o, ak, g);
                add3(gx, g[0], g[1], g[2]);
                gp += g[3];
                grt += g[4];
            }
        }
        if (pj.w > 0.f) {  // j-side: water j is the centre, k its neighbour
            const float3 aj = f3(a.gacc[j]);
            const float2 wtj = t.wt[j];
            const float4 ppj = t.pm[j];
            if (water) {
                G3<D8> d = {dvar<8>(-dx, 0), dvar<8>(-dy, 1), dvar<8>(-dz, 2)};
                G3<D8> dv = {dvar<8>(vj.x - vk.x, 3), dvar<8>(vj.y - vk.y, 4), dvar<8>(vj.z - vk.z, 5)};
                G3<D8> o = t_water_water(k, d, dv, D8(wtj.y), dvar<8>(wtk.y, 6), D8(wtj.x), dvar<8>(wtk.x, 7), mk);
                float g[8];
                contract3(o, aj, g);
                add3(gx, -g[0], -g[1], -g[2]);
                add3(gv, -g[3], -g[4], -g[5]);
                gpw += g[6];
                grt += g[7];
            } else {
                G3<D6> d = {dvar<6>(-dx, 0), dvar<6>(-dy, 1), dvar<6>(-dz, 2)};
                G3<D6> dr = {dvar<6>(xfk.x, 3), dvar<6>(xfk.y, 4), dvar<6>(xfk.z, 5)};
                G3<D6> o = t_water_sand(k, d, D6(t.rp[j].y), D6(t.wt[j].x), dr, mk, t.rp[i].x);
                float g[6];
                contract3(o, aj, g);
                add3(gx, -g[0], -g[1], -g[2]);
                add3(gdr, g[3], g[4], g[5]);
                (void)ppj;
            }
        }
    });
    if (water) {
        boundary_neighbours(k, t, pk, [&](uint32_t b) {
            const float4 bp = t.bpv[b];
            const float dx = pk.x - bp.x, dy = pk.y - bp.y, dz = pk.z - bp.z;
            if (dx * dx + dy * dy + dz * dz >= k.h2) return;
            using D7 = Dn<7>;
            G3<D7> d = {dvar<7>(dx, 0), dvar<7>(dy, 1), dvar<7>(dz, 2)};
            G3<D7> vv = {dvar<7>(vk.x, 3), dvar<7>(vk.y, 4), dvar<7>(vk.z, 5)};
            G3<D7> o = t_water_boundary(k, d, vv, dvar<7>(p_k, 6), fabsf(bp.w));
            float g[7];
            contract3(o, ak, g);
            add3(gx, g[0], g[1], g[2]);
            add3(gv, g[3], g[4], g[5]);
            gp += g[6];
        });
        a.gadh[i] = make_float4(-ak.x / mk, -ak.y / mk, -ak.z / mk, 0.f);
        a.gwt[i] = make_float2(grt, gpw);
        a.grp[i].y += gp;
    } else {
        a.gxf[i] = gdr;  // d loss / d avgDrag of sand k
    }
    a.gx[i] = gx;
    a.gv[i] = gv;
}

// ---- A4 (local): sand i's force and vw adjoints, drag, walls ------------------
__global__ __launch_bounds__(kAB) void k_adj_sand_local(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 p = t.pm[i];
    if (p.w >= 0.f) {
        a.o4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float m = -p.w;
    const float4 v = t.vr[i];
    const float3 ab = f3(a.gacc[i]);
    const float3 fb = make_float3(2.f * ab.x / m, 2.f * ab.y / m, 2.f * ab.z / m);
    const float vw = t.den4[i].z;
    const float3 xfb = f3(a.gxf[i]);  // adjoint of avgDrag = df / vw
    float3 dfb = fb;
    float vwb = 0.f;
    const float eps = t.es[i].x;
    const float rho = t.rp[i].x;
    const float4 fl = t.flow[i];
    float4 gx = a.gx[i], gv = a.gv[i], gth = a.gth[i];
    if (eps < 1.f && eps > 0.f) {
        // c(eps, cd) = g rho0 eps^2 / (cd eps^3 / (1 - eps)^2) * m / rho
        using D2 = Dn<2>;
        const D2 e = dvar<2>(eps, 0), cd = dvar<2>(k.cd, 1);
        const D2 ome = 1.f - e;
        const D2 dc = cd * (e * e * e) / (ome * ome);
        if (dc.v != 0.f) {
            const D2 c = k.grav_abs * k.rho0 * (e * e) / dc * m / rho;
            const float3 df = make_float3((fl.x - v.x) * c.v, (fl.y - v.y) * c.v, (fl.z - v.z) * c.v);
            if (vw != 0.f) {
                const float ivw = 1.f / vw;
                dfb.x += xfb.x * ivw;
                dfb.y += xfb.y * ivw;
                dfb.z += xfb.z * ivw;
                vwb = -(xfb.x * df.x + xfb.y * df.y + xfb.z * df.z) * ivw * ivw;
            }
            const float cb = dfb.x * (fl.x - v.x) + dfb.y * (fl.y - v.y) + dfb.z * (fl.z - v.z);
            add3(a.gflow[i], dfb.x * c.v, dfb.y * c.v, dfb.z * c.v);
            add3(gv, -dfb.x * c.v, -dfb.y * c.v, -dfb.z * c.v);
            a.ges[i].x += cb * c.d[0];
            gth.x += cb * c.d[1];
        }
    }
    {  // world walls
        using D7 = Dn<7>;
        G3<D7> pp = {dvar<7>(p.x, 0), dvar<7>(p.y, 1), dvar<7>(p.z, 2)};
        G3<D7> vv = {dvar<7>(v.x, 3), dvar<7>(v.y, 4), dvar<7>(v.z, 5)};
        G3<D7> o = t_walls(k, pp, vv, dvar<7>(k.tanphi, 6));
        float g[7];
        contract3(o, fb, g);
        add3(gx, g[0], g[1], g[2]);
        add3(gv, g[3], g[4], g[5]);
        gth.w += g[6];
    }
    a.gx[i] = gx;
    a.gv[i] = gv;
    a.gth[i] = gth;
    a.o4[i] = make_float4(fb.x, fb.y, fb.z, vwb);
}

// ---- A4 (pairs): sand-sand contacts / bridges, sand-water buoyancy + adhesion term
__global__ __launch_bounds__(kAB) void k_adj_sand_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const float4 vk = t.vr[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    const float twoR = 2.f * k.R;
    float4 gx = a.gx[i], gv = a.gv[i], gth = a.gth[i];
    float gsat = 0.f, grho = 0.f, geps = 0.f, gp = 0.f;
    float4 gadh = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ok = a.o4[i];
    const float3 fk = make_float3(ok.x, ok.y, ok.z);
    const float sat_k = t.es[i].y;
    const float rho_k = t.rp[i].x;
    neighbours(k, t, pk, [&](uint32_t j) {
        if ((int)j == i) return;
        const float4 pj = t.pm[j];
        const float dx = pj.x - pk.x, dy = pj.y - pk.y, dz = pj.z - pk.z;  // x_j - x_k
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (sand && pj.w < 0.f) {
            const float dist = sqrtf(r2);
            const float4 vj = t.vr[j];
            const float4 oj = a.o4[j];
            const float3 fj = make_float3(oj.x, oj.y, oj.z);
            if (dist < twoR) {
                using D7 = Dn<7>;
                {  // i-side: f_k += dem(x_j - x_k, v_j - v_k)
                    G3<D7> d = {dvar<7>(dx, 0), dvar<7>(dy, 1), dvar<7>(dz, 2)};
                    G3<D7> vv = {dvar<7>(vj.x - vk.x, 3), dvar<7>(vj.y - vk.y, 4), dvar<7>(vj.z - vk.z, 5)};
                    G3<D7> o = t_dem(k, d, vv, twoR, dvar<7>(k.tanphi, 6));
                    float g[7];
                    contract3(o, fk, g);
                    add3(gx, -g[0], -g[1], -g[2]);
                    add3(gv, -g[3], -g[4], -g[5]);
                    gth.w += g[6];
                }
                {  // j-side: f_j += dem(x_k - x_j, v_k - v_j)
                    G3<D7> d = {dvar<7>(-dx, 0), dvar<7>(-dy, 1), dvar<7>(-dz, 2)};
                    G3<D7> vv = {dvar<7>(vk.x - vj.x, 3), dvar<7>(vk.y - vj.y, 4), dvar<7>(vk.z - vj.z, 5)};
                    G3<D7> o = t_dem(k, d, vv, twoR, D7(k.tanphi));
                    float g[7];
                    contract3(o, fj, g);
                    add3(gx, g[0], g[1], g[2]);
                    add3(gv, g[3], g[4], g[5]);
                }
            } else {
                const float H = dist - twoR;
                if (H > 0.f && H < k.cap_srup) {
                    using D6 = Dn<6>;
                    const float sr = (sat_k + t.es[j].y) * 0.5f;
                    {  // i-side
                        G3<D6> d = {dvar<6>(dx, 0), dvar<6>(dy, 1), dvar<6>(dz, 2)};
                        G3<D6> o = t_capillary(k, d, dvar<6>(sr, 3), dvar<6>(k.gcap.py0, 4), dvar<6>(k.gcap.py1, 5));
                        float g[6];
                        contract3(o, fk, g);
                        add3(gx, -g[0], -g[1], -g[2]);
                        gsat += 0.5f * g[3];
                        gth.y += g[4];
                        gth.z += g[5];
                    }
                    {  // j-side
                        G3<D6> d = {dvar<6>(-dx, 0), dvar<6>(-dy, 1), dvar<6>(-dz, 2)};
                        G3<D6> o = t_capillary(k, d, dvar<6>(sr, 3), D6(k.gcap.py0), D6(k.gcap.py1));
                        float g[6];
                        contract3(o, fj, g);
                        add3(gx, g[0], g[1], g[2]);
                        gsat += 0.5f * g[3];
                    }
                }
            }
            return;
        }
        if (r2 >= k.h2) return;
        if (sand && pj.w > 0.f) {  // i-side: water j acts on sand k
            using D3 = Dn<3>;
            const float2 rpj = t.rp[j];
            const float4 xfj = t.xf[j];
            G3<D3> d = {dvar<3>(dx, 0), dvar<3>(dy, 1), dvar<3>(dz, 2)};
            G3<D3> f;
            D3 vw;
            t_sand_water(k, d, D3(rpj.x), D3(t.es[j].x), D3(rpj.y), lift<D3>(xfj.x, xfj.y, xfj.z), mk, pj.w, rho_k,
                         mk / fmaxf(kEps, rho_k), f, vw);
            float g[3];
            contract3(f, fk, g);
#pragma unroll
            for (int q = 0; q < 3; ++q) g[q] += ok.w * vw.d[q];
            add3(gx, -g[0], -g[1], -g[2]);
        } else if (!sand && pj.w < 0.f) {  // j-side: water k acts on sand j
            using D9 = Dn<9>;
            const float4 oj = a.o4[j];
            const float mi = -pj.w;
            const float rho_i = t.rp[j].x;
            const float4 xfk = t.xf[i];
            const float2 esk = t.es[i];
            const float2 rpk = t.rp[i];
            G3<D9> d = {dvar<9>(-dx, 0), dvar<9>(-dy, 1), dvar<9>(-dz, 2)};  // x_k - x_j
            G3<D9> ad = {dvar<9>(xfk.x, 6), dvar<9>(xfk.y, 7), dvar<9>(xfk.z, 8)};
            G3<D9> f;
            D9 vw;
            t_sand_water(k, d, dvar<9>(rpk.x, 3), dvar<9>(esk.x, 4), dvar<9>(rpk.y, 5), ad, mi, mk, rho_i,
                         mi / fmaxf(kEps, rho_i), f, vw);
            float g[9];
            contract3(f, make_float3(oj.x, oj.y, oj.z), g);
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] += oj.w * vw.d[q];
            add3(gx, g[0], g[1], g[2]);
            grho += g[3];
            geps += g[4];
            gp += g[5];
            add3(gadh, g[6], g[7], g[8]);
        }
    });
    a.gx[i] = gx;
    a.gv[i] = gv;
    a.gth[i] = gth;
    if (sand) {
        a.ges[i].y += gsat;
    } else {
        a.grp[i].x += grho;
        a.grp[i].y += gp;
        a.ges[i].x += geps;
        a.gxf[i] = gadh;  // d loss / d avgAdhesion of water k
    }
}

// ---- A3 (local) -------------------------------------------------------------
__global__ __launch_bounds__(kAB) void k_adj_voidage_local(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 p = t.pm[i];
    const float den = t.den4[i].y;
    if (p.w < 0.f) {  // eps = num / den
        const float eb = a.ges[i].x;
        const float eps = t.es[i].x;
        a.o2[i] = den != 0.f ? make_float2(eb / den, -eb * eps / den) : make_float2(0.f, 0.f);
        a.o4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {  // adh = F, avgAdh = F / vs
        const float3 F = f3(t.adh[i]);
        const float3 ab = f3(a.gadh[i]);
        const float3 xb = f3(a.gxf[i]);
        float4 o = make_float4(ab.x, ab.y, ab.z, 0.f);
        if (den != 0.f) {
            const float iv = 1.f / den;
            o.x += xb.x * iv;
            o.y += xb.y * iv;
            o.z += xb.z * iv;
            o.w = -(xb.x * F.x + xb.y * F.y + xb.z * F.z) * iv * iv;
        }
        a.o4[i] = o;
        a.o2[i] = make_float2(0.f, 0.f);
    }
}

// ---- A3 (pairs) ---------------------------------------------------------------
__global__ __launch_bounds__(kAB) void k_adj_voidage_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = a.gx[i];
    float gsat = 0.f, geps = 0.f, grho = 0.f;
    const float2 o2k = a.o2[i];
    const float4 o4k = a.o4[i];
    const float rho_k = t.rp[i].x;
    const float eps_k = t.es[i].x;
    neighbours(k, t, pk, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        if ((pj.w < 0.f) == sand) return;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;  // x_k - x_j
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        const v3 d = mk3(dx, dy, dz);
        using D4 = Dn<4>;
        if (sand) {  // j is water
            const float Vj = pj.w / t.rp[j].x;
            add3(gx, dpoly6(k, d, r2) * ((o2k.x * t.es[j].x + o2k.y) * Vj));  // i-side: sand voidage sums
            // j-side: F_j and vs_j of water j with sand k as neighbour; dp = x_j - x_k
            const float4 oj = a.o4[j];
            G3<D4> dp = {dvar<4>(-dx, 0), dvar<4>(-dy, 1), dvar<4>(-dz, 2)};
            G3<D4> F;
            D4 vs;
            t_adhesion(k, dp, dvar<4>(t.es[i].y, 3), pj.w, mk, F, vs);
            float g[4];
            contract3(F, make_float3(oj.x, oj.y, oj.z), g);
#pragma unroll
            for (int q = 0; q < 4; ++q) g[q] += oj.w * vs.d[q];
            add3(gx, -g[0], -g[1], -g[2]);
            gsat += g[3];
        } else {  // k water, j sand
            {  // i-side: F_k and vs_k
                G3<D4> dp = {dvar<4>(dx, 0), dvar<4>(dy, 1), dvar<4>(dz, 2)};
                G3<D4> F;
                D4 vs;
                t_adhesion(k, dp, D4(t.es[j].y), mk, -pj.w, F, vs);
                float g[4];
                contract3(F, make_float3(o4k.x, o4k.y, o4k.z), g);
#pragma unroll
                for (int q = 0; q < 3; ++q) g[q] += o4k.w * vs.d[q];
                add3(gx, g[0], g[1], g[2]);
            }
            // j-side: sand j's voidage num += eps_k V_k W, den += V_k W
            const float2 oj = a.o2[j];
            const float Vk = mk / rho_k;
            const float w = k.w_coef * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
            const float c = oj.x * eps_k + oj.y;
            geps += oj.x * Vk * w;
            grho += c * (-mk / (rho_k * rho_k)) * w;
            add3(gx, dpoly6(k, d, r2) * (c * Vk));
        }
    });
    a.gx[i] = gx;
    if (sand) {
        a.ges[i].y += gsat;
    } else {
        a.ges[i].x += geps;
        a.grp[i].x += grho;
    }
}

// ---- A2 (local): rho~ factors of water, average-flow normalisation of sand ----
__global__ __launch_bounds__(kAB) void k_adj_flow_local(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 p = t.pm[i];
    if (p.w > 0.f) {
        const float2 rp = t.rp[i];
        const float eps = t.es[i].x;
        const float rt = t.wt[i].x;
        float2 gwt = a.gwt[i];
        float2 grp = a.grp[i];
        float geps = a.ges[i].x;
        // pw = p / max(eps, rt^2)
        const float q = rt * rt;
        grp.y += gwt.y / fmaxf(kEps, q);
        if (q > kEps) gwt.x += -gwt.y * rp.y * 2.f * rt / (q * q);
        // rt = rho / max(eps, voidage)
        grp.x += gwt.x / fmaxf(kEps, eps);
        if (eps > kEps) geps += -gwt.x * rp.x / (eps * eps);
        a.grp[i] = grp;
        a.ges[i].x = geps;
        a.o2[i] = make_float2(-geps, 0.f);  // voidage = 1 - sandv
        a.o4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        const float vsum = t.den4[i].x;
        const float3 fb = f3(a.gflow[i]);
        const float sb = a.ges[i].y;
        float4 o = make_float4(0.f, 0.f, 0.f, sb);
        if (vsum != 0.f) {
            const float iv = 1.f / vsum;
            const float3 fl = f3(t.flow[i]);
            o.x = fb.x * iv;
            o.y = fb.y * iv;
            o.z = fb.z * iv;
            o.w = sb - (fb.x * fl.x + fb.y * fl.y + fb.z * fl.z) * iv;
        }
        a.o4[i] = o;
        a.o2[i] = make_float2(0.f, 0.f);
    }
}

// ---- A2 (pairs) ---------------------------------------------------------------
__global__ __launch_bounds__(kAB) void k_adj_flow_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = a.gx[i], gv = a.gv[i];
    float grho = 0.f;
    const float2 o2k = a.o2[i];
    const float4 o4k = a.o4[i];
    const float rho_k = t.rp[i].x;
    const float4 vk = t.vr[i];
    const float inv_rho_s = k.rho_s != 0.f ? 1.f / k.rho_s : 0.f;
    neighbours(k, t, pk, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        if ((pj.w < 0.f) == sand) return;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;  // x_k - x_j
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        const v3 dw = dpoly6(k, mk3(dx, dy, dz), r2);
        const float w = k.w_coef * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
        if (!sand) {  // k water, j sand
            add3(gx, dw * (o2k.x * (-pj.w * inv_rho_s)));  // i-side: sandv_k
            const float4 oj = a.o4[j];  // j-side: sand j's vsum / av with water k
            const float Vk = mk / rho_k;
            const float g = oj.x * vk.x + oj.y * vk.y + oj.z * vk.z + oj.w;
            add3(gx, dw * (g * Vk));
            grho += g * (-mk / (rho_k * rho_k)) * w;
            add3(gv, oj.x * Vk * w, oj.y * Vk * w, oj.z * Vk * w);
        } else {  // k sand, j water
            const float4 vj = t.vr[j];
            const float Vj = pj.w / t.rp[j].x;
            const float g = o4k.x * vj.x + o4k.y * vj.y + o4k.z * vj.z + o4k.w;
            add3(gx, dw * (g * Vj));  // i-side: vsum_k, av_k
            add3(gx, dw * (a.o2[j].x * (mk * inv_rho_s)));  // j-side: sandv_j
        }
    });
    a.gx[i] = gx;
    a.gv[i] = gv;
    if (!sand) a.grp[i].x += grho;
}

// ---- A1: density and EOS ------------------------------------------------------
__global__ __launch_bounds__(kAB) void k_adj_density_local(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    if (t.pm[i].w < 0.f) {
        a.grp[i] = make_float2(0.f, 0.f);
        return;
    }
    const float2 rp = t.rp[i];
    float dp;  // d p / d rho
    if (k.tait) {
        const float x = rp.x / k.rho0;
        const float x2 = x * x, x3 = x2 * x;
        dp = 7.f * k.stiff / k.rho0 * (x3 * x3);
    } else {
        dp = k.stiff;
    }
    if (rp.y < 0.f || (rp.y == 0.f && rp.x < k.rho0)) dp *= k.neg_scale;
    const float2 g = a.grp[i];
    a.grp[i] = make_float2(g.x + g.y * dp, 0.f);
}
__global__ __launch_bounds__(kAB) void k_adj_density_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    if (pk.w < 0.f) return;
    const float rb = a.grp[i].x;
    float4 gx = a.gx[i];
    neighbours(k, t, pk, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        if (pj.w < 0.f || (int)j == i) return;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        // i-side (k's density over water j) + j-side (j's density over k)
        add3(gx, dpoly6(k, mk3(dx, dy, dz), r2) * (rb * pj.w + a.grp[j].x * pk.w));
    });
    boundary_neighbours(k, t, pk, [&](uint32_t b) {
        const float4 bp = t.bpv[b];
        const float dx = pk.x - bp.x, dy = pk.y - bp.y, dz = pk.z - bp.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        add3(gx, dpoly6(k, mk3(dx, dy, dz), r2) * (rb * k.rho0 * fabsf(bp.w)));
    });
    a.gx[i] = gx;
}

__global__ __launch_bounds__(kAB) void k_adj_scatter(const uint32_t* __restrict__ perm, const float4* __restrict__ gx,
                                                        const float4* __restrict__ gv, float4* __restrict__ gxp,
                                                        float4* __restrict__ gvp, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = perm[i];
    gxp[o] = gx[i];
    gvp[o] = gv[i];
}
__global__ __launch_bounds__(kAB) void k_iota(uint32_t* out, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
}
__global__ __launch_bounds__(kAB) void k_gather_u32(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}
__global__ __launch_bounds__(kAB) void k_scatter_by_id(const uint32_t* __restrict__ id, const float* __restrict__ src,
                                                          int stride, int first, int comps, float* __restrict__ dst,
                                                          int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = id[i];
    for (int c = 0; c < comps; ++c) dst[(size_t)o * comps + c] = src[(size_t)i * stride + first + c];
}
__global__ __launch_bounds__(kAB) void k_gather_by_id(const uint32_t* __restrict__ id, const float* __restrict__ src,
                                                         int comps, float4* __restrict__ dst, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (src) {
        const size_t o = (size_t)id[i] * comps;
        r.x = src[o];
        if (comps > 1) {
            r.y = src[o + 1];
            r.z = src[o + 2];
        }
    }
    dst[i] = r;
}
__global__ __launch_bounds__(kAB) void k_gather_by_id1(const uint32_t* __restrict__ id, const float* __restrict__ src,
                                                          float* __restrict__ dst, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i < n) dst[i] = src ? src[id[i]] : 0.f;
}
__global__ __launch_bounds__(kAB) void k_set_vel_by_id(const uint32_t* __restrict__ id, const float* __restrict__ vel3,
                                                          float4* __restrict__ vr, int n) {
    const int i = blockIdx.x * kAB + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)id[i] * 3;
    float4 v = vr[i];
    v.x = vel3[o];
    v.y = vel3[o + 1];
    v.z = vel3[o + 2];
    vr[i] = v;
}

inline int blocks(int n) { return (n + kAB - 1) / kAB; }

}  // namespace

void launch_adj_step(const SFConst& k, const TapeStep& t, const AdjFields& a, const float* sat_seed, float dt, int n,
                     hipStream_t s) {
    if (n <= 0) return;
    const int g = blocks(n);
    k_adj_advect<<<g, kAB, 0, s>>>(k, t, a, sat_seed, dt, n);
    k_adj_water<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_sand_local<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_sand_pairs<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_voidage_local<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_voidage_pairs<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_flow_local<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_flow_pairs<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_density_local<<<g, kAB, 0, s>>>(k, t, a, n);
    k_adj_density_pairs<<<g, kAB, 0, s>>>(k, t, a, n);
}
void launch_adj_scatter(const uint32_t* perm, const float4* gx, const float4* gv, float4* gx_prev, float4* gv_prev,
                        int n, hipStream_t s) {
    if (n > 0) k_adj_scatter<<<blocks(n), kAB, 0, s>>>(perm, gx, gv, gx_prev, gv_prev, n);
}
void launch_iota(uint32_t* out, int n, hipStream_t s) {
    if (n > 0) k_iota<<<blocks(n), kAB, 0, s>>>(out, n);
}
void launch_gather_u32(const uint32_t* perm, const uint32_t* in, uint32_t* out, int n, hipStream_t s) {
    if (n > 0) k_gather_u32<<<blocks(n), kAB, 0, s>>>(perm, in, out, n);
}
void launch_scatter_by_id(const uint32_t* id, const float* src, int src_stride, int first, int comps, float* dst, int n,
                          hipStream_t s) {
    if (n > 0) k_scatter_by_id<<<blocks(n), kAB, 0, s>>>(id, src, src_stride, first, comps, dst, n);
}
void launch_gather_by_id(const uint32_t* id, const float* src, int comps, float4* dst, int n, hipStream_t s) {
    if (n > 0) k_gather_by_id<<<blocks(n), kAB, 0, s>>>(id, src, comps, dst, n);
}
void launch_gather_by_id1(const uint32_t* id, const float* src, float* dst, int n, hipStream_t s) {
    if (n > 0) k_gather_by_id1<<<blocks(n), kAB, 0, s>>>(id, src, dst, n);
}
void launch_set_vel_by_id(const uint32_t* id, const float* vel3, float4* vr, int n, hipStream_t s) {
    if (n > 0) k_set_vel_by_id<<<blocks(n), kAB, 0, s>>>(id, vel3, vr, n);
}

}  // namespace sf
