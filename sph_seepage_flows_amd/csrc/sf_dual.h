// sf_dual.h -- forward-mode dual numbers and the pair terms of the seepage
// step written once over a generic scalar S (float, or Dn<K> carrying K
// tangent components).  The reverse passes evaluate a pair term with its
// inputs seeded as tangents and contract the result with the output adjoint,
// which gives the vector-Jacobian product of that term exactly (to FP32) for
// the branch the forward took.
//
// Each term restates the expression the forward kernels evaluate
// (sf_solver.hip, sf_device.h), which follow the reference's functions:
//   water pressure + viscosity   ComputeSFWaterPressureTerm / ArtificialViscosity (cuda_sph_sf_solver_gpu.cuh:343-406)
//   seepage                      ComputeSFWaterSeepageTerm (:275-312)
//   boundary pressure/viscosity  cuda_sph_solver_common_gpu.cuh:89-139
//   DEM contact                  ComputeDemForces (cuda_dem_solver_common_gpu.cuh:20-51)
//   liquid bridge                ComputeDemCapillaryForces (:53-90) with QuadraticBezierCoeff (cuda_sf_utils.cuh:36-82)
//   world walls                  ComputeDemWorldBoundaryForces (:132-199)
//   buoyancy + adhesion term     ComputeSFSandBuoyancyForces / ComputeSFSandAdhesionTerm (cuda_sph_sf_solver_gpu.cuh:240-341)
//   water adhesion               ComputeSFWaterAdhesionForces (:175-207), AdhesionAkinci13 (cuda_sph_kernel.cuh:189-203)
#pragma once
#include "sf_device.h"

namespace sf {

template <int K>
struct Dn {
    float v;
    float d[K];
    SF_DEV Dn() : v(0.f) {
#pragma unroll
        for (int q = 0; q < K; ++q) d[q] = 0.f;
    }
    SF_DEV Dn(float x) : v(x) {
#pragma unroll
        for (int q = 0; q < K; ++q) d[q] = 0.f;
    }
};
template <int K>
SF_DEV Dn<K> dvar(float x, int idx) {  // idx: a compile-time constant at every call site
    Dn<K> r(x);
    r.d[idx] = 1.f;
    return r;
}

SF_DEV float val(float x) { return x; }
template <int K>
SF_DEV float val(const Dn<K>& x) { return x.v; }

#define SF_DN_LOOP(expr)      \
    _Pragma("unroll") for (int q = 0; q < K; ++q) expr

template <int K> SF_DEV Dn<K> operator+(const Dn<K>& a, const Dn<K>& b) { Dn<K> r; r.v = a.v + b.v; SF_DN_LOOP(r.d[q] = a.d[q] + b.d[q]); return r; }
template <int K> SF_DEV Dn<K> operator-(const Dn<K>& a, const Dn<K>& b) { Dn<K> r; r.v = a.v - b.v; SF_DN_LOOP(r.d[q] = a.d[q] - b.d[q]); return r; }
template <int K> SF_DEV Dn<K> operator-(const Dn<K>& a) { Dn<K> r; r.v = -a.v; SF_DN_LOOP(r.d[q] = -a.d[q]); return r; }
template <int K> SF_DEV Dn<K> operator*(const Dn<K>& a, const Dn<K>& b) { Dn<K> r; r.v = a.v * b.v; SF_DN_LOOP(r.d[q] = a.d[q] * b.v + a.v * b.d[q]); return r; }
template <int K> SF_DEV Dn<K> operator/(const Dn<K>& a, const Dn<K>& b) {
    Dn<K> r;
    const float ib = 1.f / b.v;
    r.v = a.v * ib;
    SF_DN_LOOP(r.d[q] = (a.d[q] - r.v * b.d[q]) * ib);
    return r;
}
template <int K> SF_DEV Dn<K> operator+(const Dn<K>& a, float b) { Dn<K> r = a; r.v += b; return r; }
template <int K> SF_DEV Dn<K> operator+(float a, const Dn<K>& b) { return b + a; }
template <int K> SF_DEV Dn<K> operator-(const Dn<K>& a, float b) { Dn<K> r = a; r.v -= b; return r; }
template <int K> SF_DEV Dn<K> operator-(float a, const Dn<K>& b) { Dn<K> r; r.v = a - b.v; SF_DN_LOOP(r.d[q] = -b.d[q]); return r; }
template <int K> SF_DEV Dn<K> operator*(const Dn<K>& a, float b) { Dn<K> r; r.v = a.v * b; SF_DN_LOOP(r.d[q] = a.d[q] * b); return r; }
template <int K> SF_DEV Dn<K> operator*(float a, const Dn<K>& b) { return b * a; }
template <int K> SF_DEV Dn<K> operator/(const Dn<K>& a, float b) { return a * (1.f / b); }
template <int K> SF_DEV Dn<K> operator/(float a, const Dn<K>& b) {
    Dn<K> r;
    const float ib = 1.f / b.v;
    r.v = a * ib;
    SF_DN_LOOP(r.d[q] = -r.v * b.d[q] * ib);
    return r;
}

SF_DEV float gsqrt(float x) { return sqrtf(x); }
template <int K> SF_DEV Dn<K> gsqrt(const Dn<K>& a) {
    Dn<K> r;
    r.v = sqrtf(a.v);
    const float g = r.v > 0.f ? 0.5f / r.v : 0.f;
    SF_DN_LOOP(r.d[q] = a.d[q] * g);
    return r;
}
SF_DEV float gpow(float x, float p) { return powf(x, p); }
template <int K> SF_DEV Dn<K> gpow(const Dn<K>& a, float p) {
    Dn<K> r;
    r.v = powf(a.v, p);
    const float g = a.v > 0.f ? p * powf(a.v, p - 1.f) : 0.f;
    SF_DN_LOOP(r.d[q] = a.d[q] * g);
    return r;
}
SF_DEV float gabs(float x) { return fabsf(x); }
template <int K> SF_DEV Dn<K> gabs(const Dn<K>& a) { return a.v < 0.f ? -a : a; }
// fmaxf(x, c) / fminf(x, c) against a constant
SF_DEV float gmax(float x, float c) { return fmaxf(x, c); }
template <int K> SF_DEV Dn<K> gmax(const Dn<K>& a, float c) { return a.v >= c ? a : Dn<K>(c); }

template <class S>
struct G3 {
    S x, y, z;
};
template <class S> SF_DEV G3<S> g3(S x, S y, S z) { return {x, y, z}; }
template <class S> SF_DEV G3<S> operator+(const G3<S>& a, const G3<S>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> SF_DEV G3<S> operator-(const G3<S>& a, const G3<S>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> SF_DEV G3<S> operator-(const G3<S>& a) { return {-a.x, -a.y, -a.z}; }
template <class S, class T> SF_DEV G3<S> operator*(const G3<S>& a, const T& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> SF_DEV S dot(const G3<S>& a, const G3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> SF_DEV G3<S> operator/(const G3<S>& a, const S& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class S> SF_DEV G3<S> zero3() { return {S(0.f), S(0.f), S(0.f)}; }
template <class S> SF_DEV G3<S> lift(float x, float y, float z) { return {S(x), S(y), S(z)}; }

// ---- SPH kernels (Poly6 on r^2, Spiky gradient scalar gs with gradW = gs * r) ----
template <class S> SF_DEV S g_poly6_r2(const SFConst& k, const S& r2) {
    if (val(r2) >= k.h2) return S(0.f);
    S d = k.h2 - r2;
    return k.w_coef * d * d * d;
}
template <class S> SF_DEV S g_spiky_s(const SFConst& k, const S& r) {
    if (val(r) >= k.h || val(r) < kEps) return S(0.f);
    S d = k.h - r;
    return k.g_coef * d * d / r;
}
// d/dx_i of W(|x_i - x_j|) for d = x_i - x_j inside the support
SF_DEV v3 dpoly6(const SFConst& k, v3 d, float r2) {
    const float q = k.h2 - r2;
    return d * (-6.f * k.w_coef * q * q);
}

// QuadraticBezierCoeff with differentiable end values py0, py1
template <class S> SF_DEV S g_bezier(const Bez& g, const S& py0, const S& py1, const S& sr) {
    if (val(sr) < 0.f) return py0;
    if (val(sr) >= 1.f) return S(g.py2);
    if (val(sr) <= g.px1) {
        S t = sr / g.px1, omt = 1.f - t;
        return omt * omt * py0 + 2.f * t * omt * py1 + t * t * py1;
    }
    S t = (g.b + gsqrt(g.c + g.d * (g.px1 - sr))) / g.e;
    S omt = 1.f - t;
    return omt * omt * py1 + 2.f * t * omt * py1 + t * t * g.py2;
}

// ---- water i <- water j (pressure + artificial viscosity), d = x_i - x_j, dv = v_i - v_j
template <class S>
SF_DEV G3<S> t_water_water(const SFConst& k, const G3<S>& d, const G3<S>& dv, const S& pw_i, const S& pw_j,
                           const S& rt_i, const S& rt_j, float m_j) {
    S r2 = dot(d, d);
    S gs = g_spiky_s(k, gsqrt(r2));
    S s = -m_j * (pw_i + pw_j);
    S dd = dot(dv, d);
    if (val(dd) < 0.f) {
        S pij = -k.nu / (rt_i + rt_j) * (dd / (r2 + kEps));
        s = s - m_j * pij;
    }
    return d * (s * gs);
}
// ---- water i <- sand j (seepage): pressure buoyancy + drag back-reaction
template <class S>
SF_DEV G3<S> t_water_sand(const SFConst& k, const G3<S>& d, const S& p_i, const S& rt_i, const G3<S>& drag_j,
                          float m_j, float rho_j) {
    S r2 = dot(d, d);
    S gs = g_spiky_s(k, gsqrt(r2));
    S w = g_poly6_r2(k, r2);
    return d * (-m_j * p_i / gmax(rt_i * rho_j, kEps) * gs) + drag_j * (-w / rt_i);
}
// ---- water i <- boundary particle of volume vb, d = x_i - x_b
template <class S>
SF_DEV G3<S> t_water_boundary(const SFConst& k, const G3<S>& d, const G3<S>& v_i, const S& p_i, float vb) {
    S r2 = dot(d, d);
    S gs = g_spiky_s(k, gsqrt(r2));
    S s = vb * (-k.rho0 * (p_i / fmaxf(kEps, k.rho0 * k.rho0)));
    S dd = dot(v_i, d);
    const float bvisc = -k.bnu / (2.f * k.rho0);
    if (val(dd) < 0.f) s = s + (-vb * k.rho0) * (bvisc * (dd / (r2 + kEps)));
    return d * (s * gs);
}
// ---- sand i <- sand j DEM contact, d = x_j - x_i, vji = v_j - v_i (caller checks |d| < 2R)
template <class S>
SF_DEV G3<S> t_dem(const SFConst& k, const G3<S>& d, const G3<S>& vji, float rij, const S& tanphi) {
    S dist = gsqrt(dot(d, d));
    S pen = rij - dist;
    G3<S> n = d / dist;
    S de = dot(vji, n);
    G3<S> vt = vji - n * de;
    G3<S> fn = n * (k.kn * pen);
    G3<S> fs = vt * (-k.ks);
    S maxfs = dot(fn, fn) * (tanphi * tanphi);
    S fs2 = dot(fs, fs);
    if (val(fs2) > val(maxfs)) fs = fs * (gsqrt(maxfs) / gsqrt(fs2));
    return -fn - fs;
}
template <class S>
SF_DEV G3<S> t_dem_any(const SFConst& k, const G3<S>& d, const G3<S>& vji, float rij, const S& tanphi) {
    if (rij - sqrtf(val(dot(d, d))) > 0.f) return t_dem(k, d, vji, rij, tanphi);
    return zero3<S>();
}
// ---- sand i <- sand j liquid bridge (caller checks 0 < H < s_rupture)
template <class S>
SF_DEV G3<S> t_capillary(const SFConst& k, const G3<S>& d, const S& sr, const S& c0, const S& cmc) {
    S dist = gsqrt(dot(d, d));
    S H = dist - 2.f * k.R;
    G3<S> N = d / dist;
    S cc = g_bezier(k.gcap, c0, cmc, sr);
    S dd = -H + gsqrt(H * H + k.cap_vb_piR);
    S phi = gsqrt(2.f * H / k.R * (-1.f + gsqrt(1.f + k.cap_vb_piR / (H * H))));
    S neck = -2.f * k.cap_pi * cc * k.R * k.cap_cos / (1.f + H / (2.f * dd));
    S surf = -2.f * k.cap_pi * cc * k.R * phi * k.cap_sin;
    return -N * (neck + surf);
}
// ---- sand i: world walls
template <class S>
SF_DEV G3<S> t_walls(const SFConst& k, const G3<S>& p, const G3<S>& v, const S& tanphi) {
    const float rij = k.rw + k.R;
    G3<S> f = zero3<S>();
    G3<S> mv = -v;
    const S z(0.f);
    if (val(p.x) > k.hix - rij) f = f + t_dem_any(k, g3(gabs(p.x - k.hix), z, z), mv, rij, tanphi);
    if (val(p.x) < k.lox + rij) f = f + t_dem_any(k, g3(-gabs(p.x - k.lox), z, z), mv, rij, tanphi);
    if (val(p.y) > k.hiy - rij) f = f + t_dem_any(k, g3(z, gabs(p.y - k.hiy), z), mv, rij, tanphi);
    if (val(p.y) < k.loy + rij) f = f + t_dem_any(k, g3(z, -gabs(p.y - k.loy), z), mv, rij, tanphi);
    if (val(p.z) > k.hiz - rij) f = f + t_dem_any(k, g3(z, z, gabs(p.z - k.hiz)), mv, rij, tanphi);
    if (val(p.z) < k.loz + 2.f * rij) f = f + t_dem_any(k, g3(z, z, -gabs(p.z - k.loz)), mv, rij, tanphi);
    return f;
}
// ---- sand i <- water j: buoyancy (0 < eps_j <= 1) and the adhesion term; also vw.
// d = x_j - x_i.  out: force (xyz) and vw contribution.
template <class S>
SF_DEV void t_sand_water(const SFConst& k, const G3<S>& d, const S& rho_j, const S& eps_j, const S& p_j,
                         const G3<S>& adh_j, float m_i, float m_j, float rho_i, float vol_i, G3<S>& f, S& vw) {
    S r2 = dot(d, d);
    S w = g_poly6_r2(k, r2);
    f = zero3<S>();
    vw = S(0.f);
    if (val(eps_j) > 0.f && val(eps_j) <= 1.f) {
        S dj = rho_j / gmax(eps_j, kEps);
        vw = m_j / dj * w;
        S gs = g_spiky_s(k, gsqrt(r2));
        f = d * (m_i * m_j * p_j / gmax(rho_i * dj, kEps) * gs);
    }
    f = f + adh_j * (w * vol_i);
}
// ---- water i <- sand j adhesion force F and the sand volume weight vs; dp = x_i - x_j
template <class S>
SF_DEV void t_adhesion(const SFConst& k, const G3<S>& dp, const S& sat_j, float m_i, float m_j, G3<S>& F, S& vs) {
    S l2 = dot(dp, dp);
    vs = (m_j / k.rho_s) * g_poly6_r2(k, l2);
    S l = gsqrt(l2);
    F = zero3<S>();
    if (2.f * val(l) > k.h && val(l) <= k.h) {
        S base = gmax(2.f * (2.f * l - k.h) * (k.h - l) * k.ih, 0.f);
        S A = k.adh_coef * gpow(base, 0.25f);
        if (val(A) != 0.f) {
            const S G = g_bezier(k.gadh, S(k.gadh.py0), S(k.gadh.py1), sat_j);
            F = dp * (G * m_i * m_j * A / gmax(l, kEps));
        }
    }
}

// ---- hand-derived vector-Jacobian products of the two hottest mixed-phase terms

// d gs / d r for the Spiky scalar gs(r) = g (h - r)^2 / r (0 outside [eps, h))
SF_DEV float spiky_ds(const SFConst& k, float r) {
    if (r >= k.h || r < kEps) return 0.f;
    const float q = k.h - r, ir = 1.f / r;
    return k.g_coef * (-2.f * q * ir - q * q * ir * ir);
}

// t_sand_water (sand i <- water j, d = x_j - x_i) contracted with (fb, vwb):
// adds the adjoints of d, rho_j, eps_j, p_j and adh_j.
SF_DEV void vjp_sand_water(const SFConst& k, v3 d, float rho_j, float eps_j, float p_j, v3 adh_j, float m_i, float m_j,
                           float rho_i, float vol_i, v3 fb, float vwb, v3& gd, float& grho, float& geps, float& gp,
                           v3& gadh) {
    const float r2 = dot(d, d);
    if (r2 >= k.h2) return;
    const float q2 = k.h2 - r2;
    const float w = k.w_coef * q2 * q2 * q2;
    const v3 dw = d * (-6.f * k.w_coef * q2 * q2);  // d w / d d
    const float fd = dot(fb, d);
    if (eps_j > 0.f && eps_j <= 1.f) {
        const float ep = fmaxf(eps_j, kEps);
        const float dj = rho_j / ep;
        const float X = rho_i * dj, Xc = fmaxf(X, kEps);
        const float B = m_i * m_j * p_j / Xc;
        const float r = sqrtf(r2);
        const float gs = spiky_s(k, r), gsd = spiky_ds(k, r);
        gd += fb * (B * gs);
        if (r > 0.f) gd += d * (fd * B * gsd / r);
        const float Bb = fd * gs;
        gp += Bb * m_i * m_j / Xc;
        float djb = X > kEps ? -Bb * B / X * rho_i : 0.f;
        const float vw = m_j / dj * w;
        gd += dw * (vwb * m_j / dj);
        djb += -vwb * vw / dj;
        grho += djb / ep;
        if (eps_j > kEps) geps += djb * (-rho_j / (ep * ep));
    }
    gadh += fb * (w * vol_i);
    gd += dw * (dot(fb, adh_j) * vol_i);
}

// t_water_sand (water i <- sand j, d = x_i - x_j) contracted with ab: adds
// the adjoints of d, p_i, rt_i and drag_j.
SF_DEV void vjp_water_sand(const SFConst& k, v3 d, float p_i, float rt_i, v3 drag_j, float m_j, float rho_j, v3 ab,
                           v3& gd, float& gp, float& grt, v3& gdrag) {
    const float r2 = dot(d, d);
    const float r = sqrtf(r2);
    const float gs = spiky_s(k, r), gsd = spiky_ds(k, r);
    const float Y = rt_i * rho_j, Yc = fmaxf(Y, kEps);
    const float C = -m_j / Yc;
    const float q = dot(ab, d);
    const float ad = dot(ab, drag_j);
    float w = 0.f;
    v3 dw = mk3(0.f, 0.f, 0.f);
    if (r2 < k.h2) {
        const float q2 = k.h2 - r2;
        w = k.w_coef * q2 * q2 * q2;
        dw = d * (-6.f * k.w_coef * q2 * q2);
    }
    gd += ab * (C * p_i * gs);
    if (r > 0.f) gd += d * (q * C * p_i * gsd / r);
    gd += dw * (-ad / rt_i);
    gp += q * C * gs;
    if (Y > kEps) grt += q * p_i * gs * (m_j / (rt_i * rt_i * rho_j));
    grt += ad * w / (rt_i * rt_i);
    gdrag += ab * (-w / rt_i);
}

// t_adhesion (water i <- sand j, dp = x_i - x_j) contracted with (Fb, vsb):
// adds the adjoints of dp and sat_j.
SF_DEV void vjp_adhesion(const SFConst& k, v3 dp, float sat_j, float m_i, float m_j, v3 Fb, float vsb, v3& gdp,
                         float& gsat) {
    const float l2 = dot(dp, dp);
    if (l2 >= k.h2) return;
    const float q2 = k.h2 - l2;
    gdp += dp * (vsb * (m_j / k.rho_s) * (-6.f * k.w_coef * q2 * q2));
    const float l = sqrtf(l2);
    if (!(2.f * l > k.h && l <= k.h)) return;
    const float base = adhesion_base(k, l);
    if (!(base > 0.f)) return;
    const float A = k.adh_coef * powf(base, 0.25f);
    const float dA = k.adh_coef * 0.25f * powf(base, -0.75f) * (-8.f * l / k.h + 6.f);
    const Dn<1> G = g_bezier(k.gadh, Dn<1>(k.gadh.py0), Dn<1>(k.gadh.py1), dvar<1>(sat_j, 0));
    const float L = fmaxf(l, kEps), mm = m_i * m_j;
    const float phi = G.v * mm * A / L;
    const float dphi = G.v * mm * (dA / L - (l > kEps ? A / (L * L) : 0.f));  // d phi / d l
    const float q = dot(Fb, dp);
    gdp += Fb * phi;
    if (l > 0.f) gdp += dp * (q * dphi / l);
    gsat += q * G.d[0] * mm * A / L;
}

}  // namespace sf
