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
#include "sf_nlist.h"

namespace sf {

namespace {

// Reverse passes run one wave per workgroup: small windows (tens of thousands
// of particles) still spread over every CU, and the pair kernels are latency
// bound (dependent loads of neighbour data), not LDS or occupancy bound.
constexpr int kAB = 64;

// The pair kernels filter their candidates into per-lane LDS lists
// (sf_nlist.h) and run the pair bodies over the lists, so a wave executes a
// body only for kept pairs instead of whenever some lane keeps one.  A lane
// still takes its candidates in slot order, so its sums keep their order.
constexpr int kAdjCap = 16;
#define SF_ADJ_LIST(name)                  \
    constexpr int NB = kAB, CAP = kAdjCap; \
    SF_NLIST(name)
// The loads most pair bodies start with, issued a pair ahead by the flush.
struct AdjPre {
    float4 pj, vj, wj;  // position, velocity and one per-kernel adjoint field of j
};

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

// Lanes per particle: L consecutive lanes share one particle; lane `sub`
// takes every L-th slot (from the sub-th) of each neighbour range, and the
// partial sums are combined by a fixed butterfly (deterministic for a given
// L).  Small windows get L > 1 so that their few thousand waves hide load
// latency.
template <int L>
SF_DEV float lsum(float x) {
    if constexpr (L >= 2) x += __shfl_xor(x, 1);
    if constexpr (L >= 4) x += __shfl_xor(x, 2);
    if constexpr (L >= 8) x += __shfl_xor(x, 4);
    return x;
}
template <int L>
SF_DEV float4 lsum(float4 a) {
    return make_float4(lsum<L>(a.x), lsum<L>(a.y), lsum<L>(a.z), lsum<L>(a.w));
}
SF_DEV float4 operator+(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

// Phases a neighbour loop needs: ranges without any are skipped through the
// tape's cell table (cellStart, water count before it).
constexpr int kNeedWater = 1, kNeedSand = 2, kNeedBoth = 3;

template <int L, class F>
SF_DEV void neighbours(const SFConst& k, const TapeStep& t, float4 p, int sub, int need, F&& fn) {
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint2 a = t.ct[c0], b = t.ct[c1];
        const uint32_t nw = b.y - a.y, ns = (b.x - a.x) - nw;
        if (!(((need & kNeedWater) && nw) || ((need & kNeedSand) && ns))) return;
        for (uint32_t j = a.x + (uint32_t)sub; j < b.x; j += (uint32_t)L) fn(j);
    });
}
// Partners of a sand particle: water within h and sand within reach of a
// sand-sand force (DEM overlap or liquid bridge, r < rc = 2R + s_rupture <= h).
// A column holding water is walked whole; a sand-only column only if it lies
// in the near span (along each axis the own cell and the neighbour on the
// nearer side, the forward pass's contact search), and only over its near
// cells.
template <int L, class F>
SF_DEV void sand_partners(const SFConst& k, const TapeStep& t, float4 p, int sub, float rc, F&& fn) {
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    int xa, xb, ya, yb, za, zb;
    near_span(p.x, k.lox, k.h, cx, k.gx, rc, xa, xb);
    near_span(p.y, k.loy, k.h, cy, k.gy, rc, ya, yb);
    near_span(p.z, k.loz, k.h, cz, k.gz, rc, za, zb);
    const int zlo = cz > 0 ? cz - 1 : 0;
    const int zhi = cz + 1 < k.gz ? cz + 1 : k.gz - 1;
    for (int x = cx - 1; x <= cx + 1; ++x) {
        if (x < 0 || x >= k.gx) continue;
        for (int y = cy - 1; y <= cy + 1; ++y) {
            if (y < 0 || y >= k.gy) continue;
            const uint32_t base = (uint32_t)((x * k.gy + y) * k.gz);
            const uint2 a = t.ct[base + (uint32_t)zlo], b = t.ct[base + (uint32_t)zhi + 1u];
            uint32_t s0, e0;
            if (b.y != a.y) {  // water in the column
                s0 = a.x;
                e0 = b.x;
            } else if (x >= xa && x <= xb && y >= ya && y <= yb) {
                s0 = t.ct[base + (uint32_t)za].x;
                e0 = t.ct[base + (uint32_t)zb + 1u].x;
            } else {
                continue;
            }
            for (uint32_t j = s0 + (uint32_t)sub; j < e0; j += (uint32_t)L) fn(j);
        }
    }
}
template <int L, class F>
SF_DEV void boundary_neighbours(const SFConst& k, const TapeStep& t, float4 p, int sub, F&& fn) {
    int cx, cy, cz;
    cell_of(k, p.x, p.y, p.z, cx, cy, cz);
    if (!t.bnear[cell_key(k, cx, cy, cz)]) return;  // no boundary particle in the 27 cells
    for_columns(k, cx, cy, cz, [&](uint32_t c0, uint32_t c1) {
        const uint32_t e = t.bcell[c1];
        for (uint32_t j = t.bcell[c0] + (uint32_t)sub; j < e; j += (uint32_t)L) fn(j);
    });
}

SF_DEV float sgn1(float v) { return (float)((0.f < v) - (v < 0.f)); }

// The per-particle steps between two pair passes run as the tail of the
// earlier pair kernel, in the thread that finishes particle i: they read and
// write only particle i's adjoints, and no pair kernel reads, for a
// neighbour, a field that the tail of the same kernel writes (P3's adjoints
// go to o4b / o2b while P4's pairs still read o4).
SF_DEV void adj_sand_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i);
SF_DEV void adj_voidage_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i);
SF_DEV void adj_flow_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i);
SF_DEV void adj_density_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i);

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
template <int L>
__global__ __launch_bounds__(kAB) void k_adj_water(SFConst k, TapeStep t, AdjFields a, int n) {
    const int tid = blockIdx.x * kAB + threadIdx.x;
    const int i = tid / L, sub = tid % L;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const float4 vk = t.vr[i];
    const bool water = pk.w > 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f), gv = gx, gdr = gx;
    float gp = 0.f, grt = 0.f, gpw = 0.f;
    const float3 ak = f3(a.gacc[i]);
    float2 wtk = water ? t.wt[i] : make_float2(0.f, 0.f);
    const float p_k = t.rp[i].y;
    const float4 xfk = t.xf[i];
    using D8 = Dn<8>;
    SF_ADJ_LIST(Lst);
    auto fetch = [&](uint32_t j) { return AdjPre{t.pm[j], t.vr[j], a.gacc[j]}; };
    auto body = [&](uint32_t j, const AdjPre& q) {
        const float4 pj = q.pj;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;  // x_k - x_j
        if (dx * dx + dy * dy + dz * dz >= k.h2) return;
        if (!water && pj.w < 0.f) return;
        if (water && pj.w > 0.f) {
            // water-water: the pair term is antisymmetric up to the mass ratio,
            // T(j <- k) = -(m_k / m_j) T(k <- j), so one Jacobian of T(k <- j)
            // serves both sides, contracted with a_k - (m_k / m_j) a_j.
            const float4 vj = q.vj;
            const float2 wtj = t.wt[j];
            const float3 aj = f3(q.wj);
            const float c = mk / pj.w;
            const float3 comb = make_float3(ak.x - c * aj.x, ak.y - c * aj.y, ak.z - c * aj.z);
            G3<D8> d = {dvar<8>(dx, 0), dvar<8>(dy, 1), dvar<8>(dz, 2)};
            G3<D8> dv = {dvar<8>(vk.x - vj.x, 3), dvar<8>(vk.y - vj.y, 4), dvar<8>(vk.z - vj.z, 5)};
            G3<D8> o = t_water_water(k, d, dv, dvar<8>(wtk.y, 6), D8(wtj.y), dvar<8>(wtk.x, 7), D8(wtj.x), pj.w);
            float g[8];
            contract3(o, comb, g);
            add3(gx, g[0], g[1], g[2]);
            add3(gv, g[3], g[4], g[5]);
            gpw += g[6];
            grt += g[7];
        } else if (water) {  // i-side: sand j's seepage terms on water k
            const float4 drj = t.xf[j];
            v3 gd = mk3(0.f, 0.f, 0.f), gdr_unused = gd;
            vjp_water_sand(k, mk3(dx, dy, dz), p_k, wtk.x, xyz(drj), -pj.w, t.rp[j].x, mk3(ak.x, ak.y, ak.z), gd, gp,
                           grt, gdr_unused);
            add3(gx, gd);
        } else {  // j-side: water j's seepage terms with sand k (k's avgDrag)
            const float3 aj = f3(q.wj);
            v3 gd = mk3(0.f, 0.f, 0.f), gdr3 = gd;
            float gp_unused = 0.f, grt_unused = 0.f;
            vjp_water_sand(k, mk3(-dx, -dy, -dz), t.rp[j].y, t.wt[j].x, xyz(xfk), mk, t.rp[i].x,
                           mk3(aj.x, aj.y, aj.z), gd, gp_unused, grt_unused, gdr3);
            add3(gx, -gd.x, -gd.y, -gd.z);
            add3(gdr, gdr3);
        }
    };
    neighbours<L>(k, t, pk, sub, water ? kNeedBoth : kNeedWater, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        const bool keep = (int)j != i && dx * dx + dy * dy + dz * dz < k.h2 && (water || pj.w > 0.f);
        nl_push2(Lst, keep, j, fetch, body);
    });
    nl_flush2(Lst, fetch, body);
    if (water) {
        boundary_neighbours<L>(k, t, pk, sub, [&](uint32_t b) {
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
    }
    gx = lsum<L>(gx);
    gv = lsum<L>(gv);
    gdr = lsum<L>(gdr);
    gp = lsum<L>(gp);
    grt = lsum<L>(grt);
    gpw = lsum<L>(gpw);
    if (sub != 0) return;
    if (water) {
        a.gadh[i] = make_float4(-ak.x / mk, -ak.y / mk, -ak.z / mk, 0.f);
        a.gwt[i] = make_float2(grt, gpw);
        a.grp[i].y += gp;
    } else {
        a.gxf[i] = gdr;  // d loss / d avgDrag of sand k
    }
    a.gx[i] = a.gx[i] + gx;
    a.gv[i] = a.gv[i] + gv;
    adj_sand_local(k, t, a, i);
}

// ---- A4 (local, the tail of A5): sand i's force and vw adjoints, drag, walls ----
SF_DEV void adj_sand_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i) {
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
template <int L>
__global__ __launch_bounds__(kAB) void k_adj_sand_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int tid = blockIdx.x * kAB + threadIdx.x;
    const int i = tid / L, sub = tid % L;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const float4 vk = t.vr[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    const float twoR = 2.f * k.R;
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f), gv = gx, gth = gx, gadh = gx;
    float gsat = 0.f, grho = 0.f, geps = 0.f, gp = 0.f;
    const float4 ok = a.o4[i];
    const float3 fk = make_float3(ok.x, ok.y, ok.z);
    const v3 fk3 = mk3(ok.x, ok.y, ok.z);
    const float sat_k = t.es[i].y;
    const float rho_k = t.rp[i].x;
    const float cap_r = twoR + k.cap_srup;
    SF_ADJ_LIST(Lst);
    auto fetch = [&](uint32_t j) { return AdjPre{t.pm[j], t.vr[j], a.o4[j]}; };
    auto pair = [&](uint32_t j, const AdjPre& q) {
        const float4 pj = q.pj;
        const float dx = pj.x - pk.x, dy = pj.y - pk.y, dz = pj.z - pk.z;  // x_j - x_k
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (sand && pj.w < 0.f) {
            // sand-sand: DEM and bridge terms are antisymmetric, T(j <- k) = -T(k <- j),
            // so one Jacobian of T(k <- j) serves both sides (contracted with f_j - f_k)
            if (r2 >= cap_r * cap_r) return;
            const float dist = sqrtf(r2);
            if (dist >= cap_r) return;
            const float4 vj = q.vj;
            const float4 oj = q.wj;
            const float3 dfb = make_float3(oj.x - fk.x, oj.y - fk.y, oj.z - fk.z);
            if (dist < twoR) {
                using D7 = Dn<7>;
                G3<D7> d = {dvar<7>(dx, 0), dvar<7>(dy, 1), dvar<7>(dz, 2)};
                G3<D7> vv = {dvar<7>(vj.x - vk.x, 3), dvar<7>(vj.y - vk.y, 4), dvar<7>(vj.z - vk.z, 5)};
                G3<D7> o = t_dem(k, d, vv, twoR, dvar<7>(k.tanphi, 6));
                float g[7];
                contract3(o, dfb, g);
                add3(gx, g[0], g[1], g[2]);
                add3(gv, g[3], g[4], g[5]);
                gth.w += fk.x * o.x.d[6] + fk.y * o.y.d[6] + fk.z * o.z.d[6];
            } else {
                const float H = dist - twoR;
                if (H > 0.f && H < k.cap_srup) {
                    using D6 = Dn<6>;
                    const float sr = (sat_k + t.es[j].y) * 0.5f;
                    G3<D6> d = {dvar<6>(dx, 0), dvar<6>(dy, 1), dvar<6>(dz, 2)};
                    G3<D6> o = t_capillary(k, d, dvar<6>(sr, 3), dvar<6>(k.gcap.py0, 4), dvar<6>(k.gcap.py1, 5));
                    float g[6], h[6];
                    contract3(o, dfb, g);
                    contract3(o, fk, h);
                    add3(gx, g[0], g[1], g[2]);
                    gsat += -0.5f * g[3];
                    gth.y += h[4];
                    gth.z += h[5];
                }
            }
            return;
        }
        if (r2 >= k.h2) return;
        if (sand && pj.w > 0.f) {  // i-side: water j acts on sand k
            const float2 rpj = t.rp[j];
            v3 gd = mk3(0.f, 0.f, 0.f), ga_unused = gd;
            float gr_unused = 0.f, ge_unused = 0.f, gp_unused = 0.f;
            vjp_sand_water(k, mk3(dx, dy, dz), rpj.x, t.es[j].x, rpj.y, xyz(t.xf[j]), mk, pj.w, rho_k,
                           mk / fmaxf(kEps, rho_k), fk3, ok.w, gd, gr_unused, ge_unused, gp_unused, ga_unused);
            add3(gx, -gd.x, -gd.y, -gd.z);
        } else if (!sand && pj.w < 0.f) {  // j-side: water k acts on sand j
            const float4 oj = q.wj;
            const float mi = -pj.w;
            const float rho_i = t.rp[j].x;
            const float2 rpk = t.rp[i];
            v3 gd = mk3(0.f, 0.f, 0.f), ga = gd;
            vjp_sand_water(k, mk3(-dx, -dy, -dz), rpk.x, t.es[i].x, rpk.y, xyz(t.xf[i]), mi, mk, rho_i,
                           mi / fmaxf(kEps, rho_i), mk3(oj.x, oj.y, oj.z), oj.w, gd, grho, geps, gp, ga);
            add3(gx, gd);
            add3(gadh, ga);
        }
    };
    const float cap_r2 = cap_r * cap_r;
    auto filter = [&](uint32_t j) {
        const float4 pj = t.pm[j];
        const float dx = pj.x - pk.x, dy = pj.y - pk.y, dz = pj.z - pk.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool sj = pj.w < 0.f;
        const bool keep = (int)j != i && ((sand && sj) ? r2 < cap_r2 : (r2 < k.h2 && sand != sj));
        nl_push2(Lst, keep, j, fetch, pair);
    };
    if (sand) {
        sand_partners<L>(k, t, pk, sub, cap_r + 1e-3f * k.h, filter);
    } else {
        neighbours<L>(k, t, pk, sub, kNeedSand, filter);
    }
    nl_flush2(Lst, fetch, pair);
    gx = lsum<L>(gx);
    gv = lsum<L>(gv);
    gth = lsum<L>(gth);
    gadh = lsum<L>(gadh);
    gsat = lsum<L>(gsat);
    grho = lsum<L>(grho);
    geps = lsum<L>(geps);
    gp = lsum<L>(gp);
    if (sub != 0) return;
    a.gx[i] = a.gx[i] + gx;
    a.gv[i] = a.gv[i] + gv;
    a.gth[i] = a.gth[i] + gth;
    if (sand) {
        a.ges[i].y += gsat;
    } else {
        a.grp[i].x += grho;
        a.grp[i].y += gp;
        a.ges[i].x += geps;
        a.gxf[i] = gadh;  // d loss / d avgAdhesion of water k
    }
    adj_voidage_local(k, t, a, i);
}

// ---- A3 (local, the tail of A4's pairs) --------------------------------------
SF_DEV void adj_voidage_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i) {
    const float4 p = t.pm[i];
    const float den = t.den4[i].y;
    if (p.w < 0.f) {  // eps = num / den
        const float eb = a.ges[i].x;
        const float eps = t.es[i].x;
        a.o2b[i] = den != 0.f ? make_float2(eb / den, -eb * eps / den) : make_float2(0.f, 0.f);
        a.o4b[i] = make_float4(0.f, 0.f, 0.f, 0.f);
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
        a.o4b[i] = o;
        a.o2b[i] = make_float2(0.f, 0.f);
    }
}

// ---- A3 (pairs) ---------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(kAB) void k_adj_voidage_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int tid = blockIdx.x * kAB + threadIdx.x;
    const int i = tid / L, sub = tid % L;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f);
    float gsat = 0.f, geps = 0.f, grho = 0.f;
    const float2 o2k = a.o2b[i];
    const float4 o4k = a.o4b[i];
    const float rho_k = t.rp[i].x;
    const float eps_k = t.es[i].x;
    const float sat_k = t.es[i].y;
    SF_ADJ_LIST(Lst);
    // (rho_j, eps_j, sat_j) in vj; sand k: o4 of water j, water k: o2 of sand j in wj
    auto fetch = [&](uint32_t j) {
        const float2 rp = t.rp[j], es = t.es[j];
        const float2 o2 = sand ? make_float2(0.f, 0.f) : a.o2b[j];
        return AdjPre{t.pm[j], make_float4(rp.x, es.x, es.y, 0.f),
                      sand ? a.o4b[j] : make_float4(o2.x, o2.y, 0.f, 0.f)};
    };
    auto body = [&](uint32_t j, const AdjPre& q) {
        const float4 pj = q.pj;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;  // x_k - x_j
        const float r2 = dx * dx + dy * dy + dz * dz;
        const v3 d = mk3(dx, dy, dz);
        if (sand) {  // j is water
            const float Vj = pj.w / q.vj.x;
            add3(gx, dpoly6(k, d, r2) * ((o2k.x * q.vj.y + o2k.y) * Vj));  // i-side: sand voidage sums
            // j-side: F_j and vs_j of water j with sand k as neighbour; dp = x_j - x_k
            const float4 oj = q.wj;
            v3 gd = mk3(0.f, 0.f, 0.f);
            vjp_adhesion(k, mk3(-dx, -dy, -dz), sat_k, pj.w, mk, mk3(oj.x, oj.y, oj.z), oj.w, gd, gsat);
            add3(gx, -gd.x, -gd.y, -gd.z);
        } else {  // k water, j sand
            {  // i-side: F_k and vs_k
                v3 gd = mk3(0.f, 0.f, 0.f);
                float gsat_unused = 0.f;
                vjp_adhesion(k, d, q.vj.z, mk, -pj.w, mk3(o4k.x, o4k.y, o4k.z), o4k.w, gd, gsat_unused);
                add3(gx, gd);
            }
            // j-side: sand j's voidage num += eps_k V_k W, den += V_k W
            const float2 oj = make_float2(q.wj.x, q.wj.y);
            const float Vk = mk / rho_k;
            const float w = k.w_coef * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
            const float c = oj.x * eps_k + oj.y;
            geps += oj.x * Vk * w;
            grho += c * (-mk / (rho_k * rho_k)) * w;
            add3(gx, dpoly6(k, d, r2) * (c * Vk));
        }
    };
    neighbours<L>(k, t, pk, sub, sand ? kNeedWater : kNeedSand, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        const bool keep = (pj.w < 0.f) != sand && dx * dx + dy * dy + dz * dz < k.h2;
        nl_push2(Lst, keep, j, fetch, body);
    });
    nl_flush2(Lst, fetch, body);
    gx = lsum<L>(gx);
    gsat = lsum<L>(gsat);
    geps = lsum<L>(geps);
    grho = lsum<L>(grho);
    if (sub != 0) return;
    a.gx[i] = a.gx[i] + gx;
    if (sand) {
        a.ges[i].y += gsat;
    } else {
        a.ges[i].x += geps;
        a.grp[i].x += grho;
    }
    adj_flow_local(k, t, a, i);
}

// ---- A2 (local, the tail of A3's pairs): rho~ factors of water, sand average flow
SF_DEV void adj_flow_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i) {
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
template <int L>
__global__ __launch_bounds__(kAB) void k_adj_flow_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int tid = blockIdx.x * kAB + threadIdx.x;
    const int i = tid / L, sub = tid % L;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    const bool sand = pk.w < 0.f;
    const float mk = fabsf(pk.w);
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f), gv = gx;
    float grho = 0.f;
    const float2 o2k = a.o2[i];
    const float4 o4k = a.o4[i];
    const float rho_k = t.rp[i].x;
    const float4 vk = t.vr[i];
    const float inv_rho_s = k.rho_s != 0.f ? 1.f / k.rho_s : 0.f;
    SF_ADJ_LIST(Lst);
    auto fetch = [&](uint32_t j) { return AdjPre{t.pm[j], t.vr[j], a.o4[j]}; };
    auto body = [&](uint32_t j, const AdjPre& q) {
        const float4 pj = q.pj;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;  // x_k - x_j
        const float r2 = dx * dx + dy * dy + dz * dz;
        const v3 dw = dpoly6(k, mk3(dx, dy, dz), r2);
        const float w = k.w_coef * (k.h2 - r2) * (k.h2 - r2) * (k.h2 - r2);
        if (!sand) {  // k water, j sand
            add3(gx, dw * (o2k.x * (-pj.w * inv_rho_s)));  // i-side: sandv_k
            const float4 oj = q.wj;  // j-side: sand j's vsum / av with water k
            const float Vk = mk / rho_k;
            const float g = oj.x * vk.x + oj.y * vk.y + oj.z * vk.z + oj.w;
            add3(gx, dw * (g * Vk));
            grho += g * (-mk / (rho_k * rho_k)) * w;
            add3(gv, oj.x * Vk * w, oj.y * Vk * w, oj.z * Vk * w);
        } else {  // k sand, j water
            const float4 vj = q.vj;
            const float Vj = pj.w / t.rp[j].x;
            const float g = o4k.x * vj.x + o4k.y * vj.y + o4k.z * vj.z + o4k.w;
            add3(gx, dw * (g * Vj));  // i-side: vsum_k, av_k
            add3(gx, dw * (a.o2[j].x * (mk * inv_rho_s)));  // j-side: sandv_j
        }
    };
    neighbours<L>(k, t, pk, sub, sand ? kNeedWater : kNeedSand, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        const bool keep = (pj.w < 0.f) != sand && dx * dx + dy * dy + dz * dz < k.h2;
        nl_push2(Lst, keep, j, fetch, body);
    });
    nl_flush2(Lst, fetch, body);
    gx = lsum<L>(gx);
    gv = lsum<L>(gv);
    grho = lsum<L>(grho);
    if (sub != 0) return;
    a.gx[i] = a.gx[i] + gx;
    a.gv[i] = a.gv[i] + gv;
    if (!sand) a.grp[i].x += grho;
    adj_density_local(k, t, a, i);
}

// ---- A1: density and EOS (the EOS step is the tail of A2's pairs) --------------
SF_DEV void adj_density_local(const SFConst& k, const TapeStep& t, const AdjFields& a, int i) {
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
template <int L>
__global__ __launch_bounds__(kAB) void k_adj_density_pairs(SFConst k, TapeStep t, AdjFields a, int n) {
    const int tid = blockIdx.x * kAB + threadIdx.x;
    const int i = tid / L, sub = tid % L;
    if (i >= n) return;
    const float4 pk = t.pm[i];
    if (pk.w < 0.f) return;
    const float rb = a.grp[i].x;
    float4 gx = make_float4(0.f, 0.f, 0.f, 0.f);
    neighbours<L>(k, t, pk, sub, kNeedWater, [&](uint32_t j) {
        const float4 pj = t.pm[j];
        if (pj.w < 0.f || (int)j == i) return;
        const float dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        // i-side (k's density over water j) + j-side (j's density over k)
        add3(gx, dpoly6(k, mk3(dx, dy, dz), r2) * (rb * pj.w + a.grp[j].x * pk.w));
    });
    boundary_neighbours<L>(k, t, pk, sub, [&](uint32_t b) {
        const float4 bp = t.bpv[b];
        const float dx = pk.x - bp.x, dy = pk.y - bp.y, dz = pk.z - bp.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= k.h2) return;
        add3(gx, dpoly6(k, mk3(dx, dy, dz), r2) * (rb * k.rho0 * fabsf(bp.w)));
    });
    gx = lsum<L>(gx);
    if (sub != 0) return;
    a.gx[i] = a.gx[i] + gx;
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

namespace {
template <int L>
void launch_pairs_step(const SFConst& k, const TapeStep& t, const AdjFields& a, const float* sat_seed, float dt, int n,
                       hipStream_t s) {
    const int g = blocks(n), gl = blocks(n * L);
    // each pair kernel ends with the next per-particle step (see adj_sand_local)
    k_adj_advect<<<g, kAB, 0, s>>>(k, t, a, sat_seed, dt, n);
    k_adj_water<L><<<gl, kAB, 0, s>>>(k, t, a, n);
    k_adj_sand_pairs<L><<<gl, kAB, 0, s>>>(k, t, a, n);
    k_adj_voidage_pairs<L><<<gl, kAB, 0, s>>>(k, t, a, n);
    k_adj_flow_pairs<L><<<gl, kAB, 0, s>>>(k, t, a, n);
    k_adj_density_pairs<L><<<gl, kAB, 0, s>>>(k, t, a, n);
}
}  // namespace

int adj_lanes(int n) {
    // about 8 waves per SIMD (256 CUs x 4 SIMDs) or more
    const long target = 8L * 1024 * 64;
    return (long)n * 4 <= target ? 8 : (long)n * 2 <= target ? 4 : (long)n <= target ? 2 : 1;
}

void launch_adj_step(const SFConst& k, const TapeStep& t, const AdjFields& a, const float* sat_seed, float dt, int n,
                     hipStream_t s) {
    if (n <= 0) return;
    switch (adj_lanes(n)) {
    case 8: launch_pairs_step<8>(k, t, a, sat_seed, dt, n, s); break;
    case 4: launch_pairs_step<4>(k, t, a, sat_seed, dt, n, s); break;
    case 2: launch_pairs_step<2>(k, t, a, sat_seed, dt, n, s); break;
    default: launch_pairs_step<1>(k, t, a, sat_seed, dt, n, s); break;
    }
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


__global__ __launch_bounds__(256) void k_copy_jobs(CopyJobs c) {
    const uint32_t stride = gridDim.x * 256u, t0 = blockIdx.x * 256u + threadIdx.x;
    for (int q = 0; q < c.count; ++q) {
        const uint32_t n16 = (uint32_t)(c.j[q].bytes / 16u), n4 = (uint32_t)(c.j[q].bytes / 4u);
        const uint4* s16 = static_cast<const uint4*>(c.j[q].src);
        uint4* d16 = static_cast<uint4*>(c.j[q].dst);
        for (uint32_t w = t0; w < n16; w += stride) d16[w] = s16[w];
        const uint32_t* s4 = static_cast<const uint32_t*>(c.j[q].src);
        uint32_t* d4 = static_cast<uint32_t*>(c.j[q].dst);
        for (uint32_t w = 4u * n16 + t0; w < n4; w += stride) d4[w] = s4[w];
    }
}
void launch_copy_jobs(const CopyJobs& c, hipStream_t s) {
    size_t most = 0;
    for (int q = 0; q < c.count; ++q) most = c.j[q].bytes > most ? c.j[q].bytes : most;
    if (most == 0) return;
    size_t g = (most / 16u + 255u) / 256u;
    if (g < 1) g = 1;
    if (g > 1024) g = 1024;
    k_copy_jobs<<<(int)g, 256, 0, s>>>(c);
}

}  // namespace sf
