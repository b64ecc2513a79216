/*
 * sf_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's coupled SPH-DEM seepage step, templated
 * on the arithmetic type (double = the FP64 oracle that the GPU path is checked
 * against; float = the FP32 CPU baseline timed beside the GPU).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker.  The product (libsf_seepage.so) never
 * links, loads or calls it.
 *
 * Every function follows the reference as text (paths relative to the
 * reference repository, "K/" = seepage_flows_cuda/include/kiri_pbs_cuda/):
 *   grid hash ............ K/thrust_helper/helper_thrust.cuh:56-130
 *   searcher ............. seepage_flows_cuda/src/kiri_pbs_cuda/searcher/cuda_neighbor_searcher.cu:19-139,
 *                          K/searcher/cuda_neighbor_searcher_gpu.cuh:18-25
 *   W, gradW, A .......... K/kernel/cuda_sph_kernel.cuh:79-118,189-203
 *   Bezier G ............. K/solver/seepageflow/cuda_sf_utils.cuh:36-82
 *   DEM / capillary / walls K/solver/dem/cuda_dem_solver_common_gpu.cuh:20-90,132-199
 *   boundary SPH terms ... K/solver/sph/cuda_sph_solver_common_gpu.cuh:70-139
 *   solver kernels ....... K/solver/seepageflow/cuda_sph_sf_solver_gpu.cuh:21-1074
 *   Tait EOS ............. K/solver/seepageflow/cuda_wcsph_sf_solver_gpu.cuh:19-44
 *   advect + damping ..... K/particle/cuda_sf_particles.cuh:19-40, src/.../particle/cuda_sf_particles.cu:18-48
 *   boundary volume ...... K/system/cuda_sph_system_gpu.cuh:18-66
 *   world boundary ....... src/.../emitter/cuda_boundary_emitter.cpp:12-80
 *   pass order / CFL ..... src/.../solver/seepage_flow/cuda_wcsph_sf_solver.cpp:34-179,
 *                          cuda_wcsph_sf_solver.cu:35-58, cuda_sph_sf_solver.cpp:14-152,
 *                          cuda_sph_sf_solver.cu:355-402
 *
 * Deterministic: OpenMP splits only the per-particle loops; each particle's
 * sums run in the reference's loop order (cell m = 0..26, then each helper's
 * j-loop), so results do not depend on the thread count.
 */
#include "../include/sf_seepage.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

/* unqualified math calls: std:: for float / double, the Dual overloads below otherwise */
using std::cos;
using std::fabs;
using std::pow;
using std::sin;
using std::sqrt;

/* Forward-mode dual number: value + N tangent components (FP64 host tangent). */
template <int N> struct Dual {
    double v;
    double d[N];
    Dual(double x = 0.0) : v(x) { for (int q = 0; q < N; ++q) d[q] = 0.0; }
    template <class I, class = typename std::enable_if<std::is_integral<I>::value>::type>
    Dual(I x) : Dual((double)x) {}
    explicit operator float() const { return (float)v; }
    explicit operator double() const { return v; }
    friend Dual operator+(const Dual& a, const Dual& b) { Dual r(a.v + b.v); for (int q = 0; q < N; ++q) r.d[q] = a.d[q] + b.d[q]; return r; }
    friend Dual operator-(const Dual& a, const Dual& b) { Dual r(a.v - b.v); for (int q = 0; q < N; ++q) r.d[q] = a.d[q] - b.d[q]; return r; }
    friend Dual operator-(const Dual& a) { Dual r(-a.v); for (int q = 0; q < N; ++q) r.d[q] = -a.d[q]; return r; }
    friend Dual operator*(const Dual& a, const Dual& b) { Dual r(a.v * b.v); for (int q = 0; q < N; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q]; return r; }
    friend Dual operator/(const Dual& a, const Dual& b) {
        Dual r(a.v / b.v);
        for (int q = 0; q < N; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
        return r;
    }
    Dual& operator+=(const Dual& b) { return *this = *this + b; }
    Dual& operator-=(const Dual& b) { return *this = *this - b; }
    Dual& operator*=(const Dual& b) { return *this = *this * b; }
    friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
    friend Dual sqrt(const Dual& a) {
        Dual r(std::sqrt(a.v));
        const double g = r.v > 0.0 ? 0.5 / r.v : 0.0;
        for (int q = 0; q < N; ++q) r.d[q] = a.d[q] * g;
        return r;
    }
    friend Dual pow(const Dual& a, const Dual& b) { /* a^b, b constant or a > 0 */
        Dual r(std::pow(a.v, b.v));
        const double ga = a.v != 0.0 ? b.v * std::pow(a.v, b.v - 1.0) : (b.v == 1.0 ? 1.0 : 0.0);
        const double gb = a.v > 0.0 ? r.v * std::log(a.v) : 0.0;
        for (int q = 0; q < N; ++q) r.d[q] = ga * a.d[q] + gb * b.d[q];
        return r;
    }
    friend Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
    friend Dual cos(const Dual& a) { Dual r(std::cos(a.v)); for (int q = 0; q < N; ++q) r.d[q] = -std::sin(a.v) * a.d[q]; return r; }
    friend Dual sin(const Dual& a) { Dual r(std::sin(a.v)); for (int q = 0; q < N; ++q) r.d[q] = std::cos(a.v) * a.d[q]; return r; }
};

constexpr float kEps = 1e-6f;                               /* KIRI_EPSILON kiri_pbs_pch.cuh:52 */
constexpr double kPiD = 3.141592653589793238462643383279502884197; /* KIRI_PI :54 */

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return {x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class T> inline V3<T>& operator*=(V3<T>& a, T s) { a.x *= s; a.y *= s; a.z *= s; return a; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> inline T lengthSquared(V3<T> a) { return dot(a, a); }     /* cuda_math_utils.cuh:74 */
template <class T> inline T sgn(T v) { return T((T(0) < v) - (v < T(0))); } /* cuda_math_utils.cuh:76-82 */

/* ---- grid hash: always FP32, helper_thrust.cuh:56-63 ----------------------
 * (int)(q) truncates toward zero, then clamps to [0, g-1].  For NaN and
 * beyond-int-range quotients (undefined in C++) this restatement returns 0
 * for NaN / negatives and g-1 for large values. */
inline int cell_coord(float q, int g) {
    if (!(q >= 0.f)) return 0;
    if (q >= (float)g) return g - 1;
    return (int)q;
}
struct Grid {
    float lox, loy, loz, h;
    int gx, gy, gz;
};
inline void cell_xyz(const Grid& G, float px, float py, float pz, int& cx, int& cy, int& cz) {
    float rx = px - G.lox, ry = py - G.loy, rz = pz - G.loz;
    cx = cell_coord(rx / G.h, G.gx);
    cy = cell_coord(ry / G.h, G.gy);
    cz = cell_coord(rz / G.h, G.gz);
}
inline uint32_t cell_key(const Grid& G, int cx, int cy, int cz) {
    return (uint32_t)((cx * G.gy + cy) * G.gz + cz); /* helper_thrust.cuh:87,126 */
}

/* ---- SPH kernels, cuda_sph_kernel.cuh ------------------------------------ */
template <class T> struct Poly6 { /* :79-98 */
    T h, coef, h2;
    explicit Poly6(T r) : h(r) {
        h2 = h * h;
        T ih = T(1) / h, ih3 = ih * ih * ih, ih9 = ih3 * ih3 * ih3;
        coef = T(315) * ih9 / (T(64) * T(kPiD));
    }
    T operator()(T r) const {
        T r2 = r * r;
        if (r2 >= h2) return T(0);
        T d = h2 - r2;
        return coef * d * d * d;
    }
};
template <class T> struct SpikyGrad { /* :100-118 */
    T h, coef;
    explicit SpikyGrad(T r) : h(r) {
        T h2 = h * h, h6 = h2 * h2 * h2;
        coef = T(-45) / (T(kPiD) * h6);
    }
    V3<T> operator()(V3<T> r) const {
        T rl = length(r);
        if (rl >= h || rl < T(kEps)) return mk<T>(0, 0, 0);
        T d = h - rl;
        return (coef * d * d / rl) * r;
    }
};
template <class T> struct Adhesion { /* AdhesionAkinci13 :189-203 */
    T coef, h;
    explicit Adhesion(T r) : h(r) { coef = T(0.007) / pow(h, T(3.25)); }
    T operator()(T r) const {
        if (T(2) * r > h && r <= h) return coef * pow(T(-4) * r * r / h + T(6) * r - T(2) * h, T(0.25));
        return T(0);
    }
};
template <class T> struct Bezier { /* QuadraticBezierCoeff cuda_sf_utils.cuh:36-82 */
    T a, b, c, d, e, py0, py1, px1, py2;
    Bezier(T py0_, T py1_, T px1_, T py2_) : py0(py0_), py1(py1_), px1(px1_), py2(py2_) {
        a = (px1 + T(1)) / T(4);
        b = T(-2) * a;
        c = b * b;
        d = T(-4) * (T(1) + b - px1);
        e = T(2) * (T(1) + b - px1);
    }
    T rx2t(T sr) const { return (b + sqrt(c + d * (px1 - sr))) / e; }
    T operator()(T sr) const {
        if (sr < T(0)) return py0;
        if (sr >= T(1)) return py2;
        if (sr <= px1) {
            T t = sr / px1, omt = T(1) - t;
            return omt * omt * py0 + T(2) * t * omt * py1 + t * t * py1; /* note: t^2 * py1 (:76) */
        }
        T t = rx2t(sr), omt = T(1) - t;
        return omt * omt * py1 + T(2) * t * omt * py1 + t * t * py2;
    }
};

/* ---- DEM, cuda_dem_solver_common_gpu.cuh ---------------------------------- */
template <class T>
V3<T> dem_force(V3<T> dij, V3<T> vij, T rij, T kn, T ks, T tanphi) { /* :20-51 */
    V3<T> f = mk<T>(0, 0, 0);
    T dist = length(dij);
    T pen = rij - dist;
    if (pen > T(0)) {
        V3<T> n = dij / dist;
        T de = dot(vij, n);
        V3<T> vt = vij - de * n;
        V3<T> fn = (kn * pen) * n;
        V3<T> fs = (-ks) * vt;
        T maxfs = lengthSquared(fn) * pow(tanphi, T(2));
        if (lengthSquared(fs) > maxfs) {
            T ratio = sqrt(maxfs) / length(fs);
            fs *= ratio;
        }
        f = -fn - fs;
    }
    return f;
}
template <class T> struct CapConst { /* rupture / bridge constants, :63-65 */
    T theta, vb, srup;
    explicit CapConst(T R) {
        theta = T(30) / T(180) * T(kPiD);
        vb = T(4) / T(3) * T(kPiD) * pow(R, T(3)) * T(0.01) * T(0.01);
        srup = (T(1) + T(0.5) * theta) * (pow(vb, T(1) / T(3)) + T(0.1) * pow(vb, T(2) / T(3)));
    }
};
template <class T>
V3<T> capillary_force(V3<T> dij, T R, T sr, const Bezier<T>& G, const CapConst<T>& K) { /* :53-90 */
    V3<T> f = mk<T>(0, 0, 0);
    T dist = length(dij);
    T H = dist - (R + R);
    if (H < K.srup && H > T(0)) {
        V3<T> N = dij / dist;
        T cc = G(sr);
        T d = -H + sqrt(H * H + K.vb / (T(kPiD) * R));
        T phi = sqrt(T(2) * H / R * (T(-1) + sqrt(T(1) + K.vb / (T(kPiD) * R * H * H))));
        T neck = T(-2) * T(kPiD) * cc * R * cos(K.theta) / (T(1) + H / (T(2) * d));
        T surf = T(-2) * T(kPiD) * cc * R * phi * sin(K.theta);
        f = -N * (neck + surf);
    }
    return f;
}
template <class T>
void world_boundary_force(V3<T>& f, V3<T> p, V3<T> v, T R, T Rb, T young, T poisson, T tanphi,
                          V3<T> lo, V3<T> hi) { /* :132-199 */
    T rij = Rb + R, kn = young * R, ks = kn * poisson;
    V3<T> mv = -v;
    if (p.x > hi.x - rij) f += dem_force(mk<T>(1, 0, 0) * fabs(p.x - hi.x), mv, rij, kn, ks, tanphi);
    if (p.x < lo.x + rij) f += dem_force(mk<T>(-1, 0, 0) * fabs(p.x - lo.x), mv, rij, kn, ks, tanphi);
    if (p.y > hi.y - rij) f += dem_force(mk<T>(0, 1, 0) * fabs(p.y - hi.y), mv, rij, kn, ks, tanphi);
    if (p.y < lo.y + rij) f += dem_force(mk<T>(0, -1, 0) * fabs(p.y - lo.y), mv, rij, kn, ks, tanphi);
    if (p.z > hi.z - rij) f += dem_force(mk<T>(0, 0, 1) * fabs(p.z - hi.z), mv, rij, kn, ks, tanphi);
    if (p.z < lo.z + T(2) * rij) /* 2*rij trigger as written at :191 */
        f += dem_force(mk<T>(0, 0, -1) * fabs(p.z - lo.z), mv, rij, kn, ks, tanphi);
}

/* ---- the oracle system ----------------------------------------------------- */
struct Base {
    virtual ~Base() {}
};

template <class T> struct Sys : Base {
    sf_params P;
    sf_boundary_params B;
    sf_solver_options O;
    Grid G;
    size_t C = 0; /* gx*gy*gz (cellStart has C+1 entries) */
    size_t n = 0, nb = 0, nsub = 1;
    bool multi = false;

    std::vector<V3<T>> pos, vel, acc, col, avgflow, avgdrag, adh, avgadh, cda;
    std::vector<T> mass, radius, density, pressure, voidage, saturation, maxsat, amc0, amc1;
    std::vector<uint64_t> label;
    std::vector<uint32_t> keys, perm, sorted_keys, cellStart;
    std::vector<uint32_t> ids; /* upload slot of each particle, follows every sort */
    /* differentiated parameters (sf_cd, sf_c0, sf_cmc, dem_tan_friction_angle):
     * read from P unless a tangent run seeded them */
    bool th_set = false;
    T th[4];
    T p_cd() const { return th_set ? th[0] : T(P.sf_cd); }
    T p_c0() const { return th_set ? th[1] : T(P.sf_c0); }
    T p_cmc() const { return th_set ? th[2] : T(P.sf_cmc); }
    T p_tan() const { return th_set ? th[3] : T(P.dem_tan_friction_angle); }
    std::vector<V3<T>> bpos;
    std::vector<uint64_t> blabel;
    std::vector<T> bvol;
    std::vector<uint32_t> bcellStart;

    void cellof(V3<T> p, int& cx, int& cy, int& cz) const {
        cell_xyz(G, (float)p.x, (float)p.y, (float)p.z, cx, cy, cz);
    }
    /* neighbour cell m (order m/9-1, (m%9)/3-1, m%3-1) or -1 when outside */
    long ncell(int cx, int cy, int cz, int m) const {
        int x = cx + m / 9 - 1, y = cy + (m % 9) / 3 - 1, z = cz + m % 3 - 1;
        if (x >= 0 && x < G.gx && y >= 0 && y < G.gy && z >= 0 && z < G.gz) return (long)cell_key(G, x, y, z);
        return -1; /* sentinel gx*gy*gz, helper_thrust.cuh:123-128 */
    }

    /* Inflow emitter (host code of the reference, FP32 as written; this file
     * is compiled without fused multiply-add): CudaEmitter samples and axes
     * (cuda_emitter.cpp:12-97, cuda_emitter.cuh:33-55), the emission schedule
     * of CudaSFSystem::OnUpdateSolver (cuda_sf_system.cu:86-115) and
     * AddSphParticles (cuda_sf_particles.cu:50-68).  cap = max particle count. */
    struct Emitter {
        bool enable = false, run = false;
        float pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, col[3] = {0, 0, 0}, a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
        std::vector<float> s1, s2;
        size_t counter = 0, emitted = 0, cap = 0;
        float elapsed = 0.f;
    } em;
    void emit_water() {
        const size_t k = em.s1.size();
        if (n + k >= em.cap) { /* Size() + p.size() < MaxSize() failed: emitter off */
            em.enable = false;
            return;
        }
        for (size_t q = 0; q < k; ++q) {
            float p[3];
            for (int c = 0; c < 3; ++c) p[c] = em.pos[c] + em.s1[q] * em.a1[c] + em.s2[q] * em.a2[c];
            pos.push_back(mk<T>(p[0], p[1], p[2]));
            vel.push_back(mk<T>(em.vel[0], em.vel[1], em.vel[2]));
            col.push_back(mk<T>(em.col[0], em.col[1], em.col[2]));
            for (auto* a : {&acc, &avgflow, &avgdrag, &adh, &avgadh, &cda}) a->push_back(mk<T>(0, 0, 0));
            mass.push_back(T(P.sph_mass));
            radius.push_back(T(P.sph_particle_radius));
            for (auto* a : {&density, &pressure, &voidage, &saturation, &maxsat, &amc0, &amc1}) a->push_back(T(0));
            label.push_back(0);
            ids.push_back((uint32_t)ids.size());
            keys.push_back(0);
            perm.push_back(0);
            sorted_keys.push_back(0);
        }
        n += k;
        em.emitted += k;
    }
    void emit_after_update(float render_interval) {
        if (!em.enable || !em.run || em.s1.empty()) return;
        const float speed = std::sqrt(em.vel[0] * em.vel[0] + em.vel[1] * em.vel[1] + em.vel[2] * em.vel[2]);
        const float two_r = 2.f * P.sph_particle_radius;
        if (P.solver_type == SF_WCSPH_SOLVER) {
            em.elapsed += render_interval / static_cast<float>(nsub);
            if ((speed * em.elapsed) / two_r >= static_cast<float>(em.counter)) {
                emit_water();
                em.counter++;
            }
        } else {
            size_t step_num = (size_t)(two_r / (speed * P.dt));
            if (step_num == 0) step_num = 1; /* the reference divides by zero here */
            if (em.counter++ % step_num == 0) emit_water();
        }
    }

    void resize(size_t k) {
        n = k;
        for (auto* a : {&pos, &vel, &acc, &col, &avgflow, &avgdrag, &adh, &avgadh, &cda}) a->assign(k, mk<T>(0, 0, 0));
        for (auto* a : {&mass, &radius, &density, &pressure, &voidage, &saturation, &maxsat, &amc0, &amc1}) a->assign(k, T(0));
        label.assign(k, 0);
        ids.resize(k);
        std::iota(ids.begin(), ids.end(), 0u);
        keys.assign(k, 0);
        perm.assign(k, 0);
        sorted_keys.assign(k, 0);
    }

    /* stable counting of keys -> cellStart (CountingInCell_CUDA + exclusive_scan) */
    static void cell_start_from_keys(const std::vector<uint32_t>& k, size_t C, std::vector<uint32_t>& cs) {
        cs.assign(C + 1, 0);
        for (uint32_t key : k) cs[key] += 1;
        uint32_t run = 0;
        for (size_t c = 0; c <= C; ++c) {
            uint32_t v = cs[c];
            cs[c] = run;
            run += v;
        }
    }
    static std::vector<uint32_t> stable_order(const std::vector<uint32_t>& k) {
        std::vector<uint32_t> idx(k.size());
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
        return idx;
    }
    template <class A> static void permute(A& a, const std::vector<uint32_t>& p) {
        A t(a.size());
        for (size_t i = 0; i < p.size(); ++i) t[i] = a[p[i]];
        a.swap(t);
    }

    /* BuildGNSearcher + SortData (cuda_neighbor_searcher.cu:36-116) */
    void build_neighbors() {
        for (size_t i = 0; i < n; ++i) {
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            keys[i] = cell_key(G, cx, cy, cz);
        }
        perm = stable_order(keys);
        for (size_t i = 0; i < n; ++i) sorted_keys[i] = keys[perm[i]];
        permute(label, perm);
        permute(ids, perm);
        permute(pos, perm);
        permute(vel, perm);
        permute(col, perm);
        permute(radius, perm);
        permute(mass, perm);
        permute(maxsat, perm);
        if (multi) {
            permute(cda, perm);
            permute(amc0, perm);
            permute(amc1, perm);
        }
        cell_start_from_keys(keys, C, cellStart);
    }

    void upload_boundary(const float* p3, const uint64_t* lab, size_t k) {
        nb = k;
        std::vector<uint32_t> bk(k);
        bpos.resize(k);
        blabel.resize(k);
        for (size_t i = 0; i < k; ++i) {
            bpos[i] = mk<T>(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]);
            blabel[i] = lab ? lab[i] : 0;
            int cx, cy, cz;
            cell_xyz(G, p3[3 * i], p3[3 * i + 1], p3[3 * i + 2], cx, cy, cz);
            bk[i] = cell_key(G, cx, cy, cz);
        }
        auto bp = stable_order(bk);
        permute(bpos, bp);
        permute(blabel, bp);
        cell_start_from_keys(bk, C, bcellStart);
        /* ComputeBoundaryVolume_CUDA, cuda_sph_system_gpu.cuh:18-66 (kernel radius = cell size) */
        Poly6<T> W((T)G.h);
        bvol.assign(k, T(0));
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)k; ++i) {
            int cx, cy, cz;
            cellof(bpos[i], cx, cy, cz);
            T acc_ = T(0);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j)
                    if ((long)j != i) acc_ += W(length(bpos[i] - bpos[j]));
            }
            bvol[i] = T(1) / std::max(acc_, T(kEps));
        }
    }

    /* ---- solver passes (cuda_wcsph_sf_solver.cpp:34-179) ---- */
    void extra_forces() { /* cuda_sph_sf_solver.cu:390-402 */
        V3<T> g = mk<T>(P.gravity.x, P.gravity.y, P.gravity.z);
        for (size_t i = 0; i < n; ++i) acc[i] = g;
    }

    void compute_density() { /* ComputeSFDensity_CUDA :437-480 */
        Poly6<T> W((T)B.kernel_radius);
        T rho0 = P.sph_density, rho1 = P.dem_density;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] == 1) {
                density[i] = rho1;
                continue;
            }
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            T d = mass[i] * W(T(0));
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j) /* ComputeSFFluidDensity :408-430 */
                    if (label[i] == label[j]) d += mass[j] * W(length(pos[i] - pos[j]));
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j) /* ComputeBoundaryDensity */
                    d += rho0 * bvol[j] * W(length(pos[i] - bpos[j]));
            }
            density[i] = d;
        }
    }

    void compute_pressure() {
        T rho0 = P.sph_density, B_ = P.sph_stiff;
        bool tait = P.solver_type == SF_WCSPH_SOLVER;
        T neg = tait ? T(O.negative_scale) : T(1); /* ComputeSFPressure_CUDA passes 1 (cuda_sph_sf_solver.cu:63) */
        for (size_t i = 0; i < n; ++i) {
            if (label[i] == 0) {
                T p = tait ? B_ * (pow(density[i] / rho0, T(7)) - T(1)) /* wcsph_gpu.cuh:34 */
                           : B_ * (density[i] - rho0);                       /* sph_sf_gpu.cuh:497 */
                if (p < T(0)) p *= neg;
                pressure[i] = p;
            } else if (label[i] == 1) {
                pressure[i] = T(0);
            }
        }
    }

    void compute_avg_flow() { /* ComputeSFAvgFlow_CUDA :510-572 */
        Poly6<T> W((T)B.kernel_radius);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            T vs = 0, vw = 0, sandv = 0;
            V3<T> avs = mk<T>(0, 0, 0), avw = mk<T>(0, 0, 0);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j) /* ComputeSFSandVolume :111-135 */
                    if (label[j] == 1 && density[j] != T(0)) {
                        T vj = mass[j] / density[j];
                        sandv += vj * W(length(pos[i] - pos[j]));
                    }
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j) { /* ComputeSFAvgFlow :137-173 */
                    T vj = mass[j] / density[j];
                    T wij = W(length(pos[i] - pos[j]));
                    if (label[i] == 1 && label[j] == 0) {
                        avs += vel[j] * vj * wij;
                        vs += vj * wij;
                    } else if (label[i] == 0 && label[j] == 1) {
                        avw += vel[j] * vj * wij;
                        vw += vj * wij;
                    }
                }
            }
            if (label[i] == 0 && vw != T(0)) avgflow[i] = avw / vw;
            else if (label[i] == 1 && vs != T(0)) avgflow[i] = avs / vs;
            else avgflow[i] = mk<T>(0, 0, 0);
            voidage[i] = T(1) - sandv;
            saturation[i] = vs;
        }
    }

    void compute_sand_voidage() { /* ComputeSFSandVoidage_CUDA :680-721 */
        Poly6<T> W((T)B.kernel_radius);
        std::vector<T> out(voidage);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] != 1) continue;
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            T num = 0, den = 0;
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j)
                    if (label[i] == 1 && label[j] == 0) {
                        T vj = mass[j] / density[j];
                        T wij = W(length(pos[i] - pos[j]));
                        num += voidage[j] * vj * wij;
                        den += vj * wij;
                    }
            }
            out[i] = den != T(0) ? num / den : T(0);
        }
        voidage.swap(out);
    }

    void compute_water_adhesion() { /* Compute(Multi)SFWaterAdhesionForces_CUDA :574-678 */
        Poly6<T> W((T)B.kernel_radius);
        Adhesion<T> A((T)B.kernel_radius);
        Bezier<T> G = multi ? Bezier<T>(T(0), T(1.5), T(0.5), T(0.8)) /* :667 */
                            : Bezier<T>(P.sf_a0, P.sf_amc, P.sf_amc_p, P.sf_asat); /* cuda_sph_sf_solver.cu:119 */
        for (size_t i = 0; i < n; ++i) { /* Advect resets both to 0 each step (cuda_sph_sf_solver.cu:384-385) */
            adh[i] = mk<T>(0, 0, 0);
            avgadh[i] = mk<T>(0, 0, 0);
        }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] != 0) continue;
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            T vs = 0;
            V3<T> F = mk<T>(0, 0, 0);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j) /* ComputeSFWaterAdhesionForces :175-207 */
                    if (label[i] == 0 && label[j] == 1) {
                        V3<T> dp = pos[i] - pos[j];
                        T l = length(dp);
                        T wij = W(l), aij = A(l);
                        F += G(saturation[j]) * mass[i] * mass[j] * aij * dp / std::max(T(kEps), l);
                        vs += mass[j] / density[j] * wij;
                    }
            }
            adh[i] = F;
            if (vs != T(0)) avgadh[i] = F / vs;
        }
    }

    void compute_sand_momentum() { /* Compute(Multi)SFSandLinearMomentum_CUDA :723-897 */
        Poly6<T> W((T)B.kernel_radius);
        SpikyGrad<T> dW((T)B.kernel_radius);
        T R = P.dem_particle_radius, Rw = P.sph_particle_radius;
        T young = P.dem_young, poisson = P.dem_poisson, tanphi = p_tan();
        T kn = young * R, ks = kn * poisson, rij = T(2) * R;
        Bezier<T> G(p_c0(), p_cmc(), T(P.sf_cmc_p), T(P.sf_csat)); /* cuda_sph_sf_solver.cu:249,308 */
        CapConst<T> K(R);
        T grav = fabs((T)P.gravity.y), rho0 = P.sph_density;
        V3<T> lo = mk<T>(B.lowest_point.x, B.lowest_point.y, B.lowest_point.z);
        V3<T> hi = mk<T>(B.highest_point.x, B.highest_point.y, B.highest_point.z);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] != 1) continue;
            T vw = 0;
            V3<T> f = mk<T>(0, 0, 0), df = mk<T>(0, 0, 0);
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                uint32_t s = cellStart[c], e = cellStart[c + 1];
                for (uint32_t j = s; j < e; ++j) /* ComputeSFUniRadiusDemForces :21-50 */
                    if ((long)j != i && label[i] == label[j])
                        f += dem_force(pos[j] - pos[i], vel[j] - vel[i], rij, kn, ks, tanphi);
                for (uint32_t j = s; j < e; ++j) /* ComputeSFUniRadiusDemCapillaryForces :83-109 */
                    if ((long)j != i && label[i] == label[j])
                        f += capillary_force(pos[j] - pos[i], R, (saturation[i] + saturation[j]) / T(2), G, K);
                for (uint32_t j = s; j < e; ++j) /* ComputeSFSandBuoyancyForces :240-273 */
                    if ((long)j != i && label[i] == 1 && label[j] == 0 && voidage[j] > T(0) && voidage[j] <= T(1)) {
                        T di = density[i];
                        T dj = density[j] / std::max(T(kEps), voidage[j]);
                        V3<T> dij = pos[i] - pos[j];
                        T wij = W(length(dij));
                        V3<T> nw = dW(dij);
                        vw += mass[j] / dj * wij;
                        f += (-mass[i] * mass[j] * pressure[j] / std::max(T(kEps), di * dj)) * nw;
                    }
                for (uint32_t j = s; j < e; ++j) /* ComputeSFSandAdhesionTerm :314-341 */
                    if (label[i] == 1 && label[j] == 0) {
                        V3<T> dij = pos[i] - pos[j];
                        T wij = W(length(dij));
                        T v = mass[i] / std::max(T(kEps), density[i]);
                        f += wij * v * avgadh[j];
                    }
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j) /* ComputeSFUniRadiusDemBoundaryForces :52-81 */
                    if (blabel[j] == 1)
                        f += dem_force(pos[i] - bpos[j], -vel[i], rij, kn, ks, tanphi);
            }
            world_boundary_force(f, pos[i], vel[i], R, Rw, young, poisson, tanphi, lo, hi);
            T eps = voidage[i];
            if (eps < T(1) && eps > T(0)) {
                T cd = multi ? cda[i].x : p_cd();
                T dc = cd * pow(eps, T(3)) / pow(T(1) - eps, T(2));
                if (dc != T(0))
                    df = (grav * rho0 * pow(eps, T(2)) / dc) * (avgflow[i] - vel[i]) * mass[i] / density[i];
            }
            avgdrag[i] = vw != T(0) ? df / vw : mk<T>(0, 0, 0);
            acc[i] += T(2) * (f + df) / mass[i];
        }
    }

    void compute_water_momentum() { /* ComputeSFWaterLinearMomentum_CUDA :899-950 */
        Poly6<T> W((T)B.kernel_radius);
        SpikyGrad<T> dW((T)B.kernel_radius);
        T rho0 = P.sph_density, nu = P.sph_nu, bnu = P.sph_bnu;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] != 0) continue;
            V3<T> a = mk<T>(0, 0, 0);
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                uint32_t s = cellStart[c], e = cellStart[c + 1];
                for (uint32_t j = s; j < e; ++j) /* ComputeSFWaterPressureTerm :343-369 */
                    if ((long)j != i && label[i] == label[j]) {
                        T di = density[i] / std::max(T(kEps), voidage[i]);
                        T dj = density[j] / std::max(T(kEps), voidage[j]);
                        a += (-mass[j] * (pressure[i] / std::max(T(kEps), di * di) + pressure[j] / std::max(T(kEps), dj * dj))) *
                             dW(pos[i] - pos[j]);
                    }
                for (uint32_t j = s; j < e; ++j) /* ComputeSFWaterArtificialViscosity :371-406 */
                    if ((long)j != i && label[i] == label[j]) {
                        V3<T> dp = pos[i] - pos[j];
                        V3<T> dv = vel[i] - vel[j];
                        T dd = dot(dv, dp);
                        if (dd < T(0)) {
                            T di = density[i] / std::max(T(kEps), voidage[i]);
                            T dj = density[j] / std::max(T(kEps), voidage[j]);
                            T pij = -nu / (di + dj) * (dd / (lengthSquared(dp) + T(kEps)));
                            a += (-mass[j] * pij) * dW(dp);
                        }
                    }
                for (uint32_t j = s; j < e; ++j) /* ComputeSFWaterSeepageTerm :275-312 */
                    if ((long)j != i && label[i] == 0 && label[j] == 1) {
                        T di = density[i] / std::max(T(kEps), voidage[i]);
                        T dj = density[j];
                        V3<T> dij = pos[i] - pos[j];
                        T wij = W(length(dij));
                        V3<T> nw = dW(dij);
                        V3<T> buoy = (-mass[j] * pressure[i] / std::max(T(kEps), di * dj)) * nw;
                        V3<T> drag = (-wij / di) * avgdrag[j];
                        a += buoy + drag;
                    }
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j) /* ComputeBoundaryPressure :89-108 */
                    a += (-rho0 * bvol[j] * (pressure[i] / std::max(T(kEps), rho0 * rho0))) * dW(pos[i] - bpos[j]);
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j) { /* ComputeBoundaryViscosity :110-139 */
                    V3<T> dp = pos[i] - bpos[j];
                    T dd = dot(vel[i], dp);
                    if (dd < T(0)) {
                        T pij = -bnu / (T(2) * rho0) * (dd / (lengthSquared(dp) + T(kEps)));
                        a += (-bvol[j] * rho0 * pij) * dW(dp);
                    }
                }
            }
            acc[i] += a - adh[i] / mass[i];
        }
    }

    void compute_wet_sand_color() { /* ComputeSFWetSandColor_CUDA :952-977 */
        V3<T> dry = mk<T>(P.sf_dry_sand_color.x, P.sf_dry_sand_color.y, P.sf_dry_sand_color.z);
        V3<T> wet = mk<T>(P.sf_wet_sand_color.x, P.sf_wet_sand_color.y, P.sf_wet_sand_color.z);
        for (size_t i = 0; i < n; ++i) {
            if (label[i] != 1) continue;
            if (maxsat[i] < saturation[i]) {
                col[i] = dry + (wet - dry) * saturation[i];
                maxsat[i] = saturation[i];
            }
        }
    }

    /* ComputeSubTimeStepsByCFL cuda_wcsph_sf_solver.cu:35-58 (FP32 host arithmetic as written) */
    size_t cfl_sub_steps(float render_interval) {
        float amax = 0.f;
        for (size_t i = 0; i < n; ++i) {
            float l = sqrt((float)acc[i].x * (float)acc[i].x + (float)acc[i].y * (float)acc[i].y +
                                (float)acc[i].z * (float)acc[i].z);
            amax = std::max(amax, l);
        }
        float h = B.kernel_radius, m = P.sph_mass;
        float fmax_ = amax * m;
        float dts = 0.4f * h / O.speed_of_sound;
        float dtf = 0.25f * sqrt(h * m / fmax_);
        float dt = std::min(O.time_step_limit_scale * std::min(dts, dtf), P.dt);
        size_t k = (size_t)(render_interval / dt);
        return k < 1 ? 1 : k; /* the reference divides by 0 here; the restatement keeps >= 1 */
    }

    void advect(T dt) { /* CudaSphSFSolver::Advect cuda_sph_sf_solver.cu:355-388 */
        T damping = P.dem_damping;
        for (size_t i = 0; i < n; ++i) { /* AccDampingForSand cuda_sf_particles.cuh:19-40 */
            if (label[i] == 1) {
                V3<T> a = acc[i];
                V3<T> q = a * (vel[i] + T(0.5) * dt * a);
                V3<T> s = mk<T>(sgn(q.x), sgn(q.y), sgn(q.z));
                acc[i] = a * (mk<T>(1, 1, 1) - damping * s);
            }
        }
        for (size_t i = 0; i < n; ++i) vel[i] = vel[i] + dt * acc[i];
        for (size_t i = 0; i < n; ++i) pos[i] = pos[i] + dt * vel[i];
        /* SFWaterBoundaryConstrain_CUDA :979-1074 */
        T r = P.sph_particle_radius;
        V3<T> lo = mk<T>(B.lowest_point.x, B.lowest_point.y, B.lowest_point.z);
        V3<T> hi = mk<T>(B.highest_point.x, B.highest_point.y, B.highest_point.z);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            if (label[i] != 0) continue;
            V3<T> tp = pos[i], tv = vel[i];
            if (tp.x > hi.x - T(2) * r) { tp.x = hi.x - T(2) * r; tv.x = std::min(tv.x, T(0)); }
            if (tp.x < lo.x + T(2) * r) { tp.x = lo.x + T(2) * r; tv.x = std::max(tv.x, T(0)); }
            if (tp.y > hi.y - T(2) * r) { tp.y = hi.y - T(2) * r; tv.y = std::min(tv.y, T(0)); }
            if (tp.y < lo.y + T(2) * r) { tp.y = lo.y + T(2) * r; tv.y = std::max(tv.y, T(0)); }
            if (tp.z > hi.z - T(2) * r) { tp.z = hi.z - T(2) * r; tv.z = std::min(tv.z, T(0)); }
            if (tp.z < lo.z + T(2) * r) { tp.z = lo.z + T(2) * r; tv.z = std::max(tv.z, T(0)); }
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = bcellStart[c]; j < bcellStart[c + 1]; ++j)
                    if (blabel[j] == 1) {
                        T d = length(pos[i] - bpos[j]);
                        T ov = d - T(2) * r;
                        if (ov < T(0)) {
                            V3<T> nn = (pos[i] - bpos[j]) / d;
                            tv *= T(0);
                            tp += nn * -ov;
                        }
                    }
            }
            pos[i] = tp; /* each thread writes only its own slot after reading pos[i] */
            vel[i] = tv;
        }
    }

    void solve() {
        extra_forces();
        compute_density();
        compute_pressure();
        compute_avg_flow();
        compute_sand_voidage();
        compute_water_adhesion();
        compute_sand_momentum();
        compute_water_momentum();
        compute_wet_sand_color();
    }
    /* one sub-step with a caller-fixed dt */
    void step(T dt) {
        build_neighbors();
        solve();
        advect(dt);
    }
    /* CudaBaseSystem::UpdateSystem -> OnUpdateSolver */
    void update_system(float render_interval) {
        build_neighbors();
        if (P.solver_type == SF_WCSPH_SOLVER) {
            solve();
            nsub = cfl_sub_steps(render_interval);
            advect((T)(render_interval / (float)nsub));
        } else {
            size_t k = (size_t)(render_interval / P.dt); /* cuda_sph_sf_solver.cpp:23 */
            nsub = k < 1 ? 1 : k;
            solve();
            advect((T)P.dt);
        }
        emit_after_update(render_interval);
    }

    /* in-range neighbour count: r^2 < h^2 with r^2 = dx*dx + dy*dy + dz*dz in
     * FP32, evaluated left to right without fused multiply-add */
    void neighbor_counts(std::vector<uint32_t>& cnt, std::vector<uint32_t>& cand) const {
        cnt.assign(n, 0);
        cand.assign(n, 0);
        float h = B.kernel_radius, h2 = h * h;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            int cx, cy, cz;
            cellof(pos[i], cx, cy, cz);
            uint32_t k = 0, kc = 0;
            float px = (float)pos[i].x, py = (float)pos[i].y, pz = (float)pos[i].z;
            for (int m = 0; m < 27; ++m) {
                long c = ncell(cx, cy, cz, m);
                if (c < 0) continue;
                for (uint32_t j = cellStart[c]; j < cellStart[c + 1]; ++j) {
                    volatile float dx = px - (float)pos[j].x, dy = py - (float)pos[j].y, dz = pz - (float)pos[j].z;
                    volatile float xx = dx * dx, yy = dy * dy, zz = dz * dz;
                    volatile float s = xx + yy;
                    volatile float r2 = s + zz;
                    k += (r2 < h2) ? 1u : 0u;
                    kc += 1u;
                }
            }
            cnt[i] = k;
            cand[i] = kc;
        }
    }
};

} // namespace orc

/* ============================ C interface ================================= */
using namespace orc;

namespace {
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
struct Handle {
    int prec; /* 32 or 64 */
    Base* s;
};
template <class T> Sys<T>* S(Handle* h) { return static_cast<Sys<T>*>(h->s); }
} // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_threads(int k) {
#ifdef _OPENMP
    if (k > 0) omp_set_num_threads(k);
    return omp_get_max_threads();
#else
    (void)k;
    return 1;
#endif
}

/* ---- scalar restatements used by the pinning tests (double) ---- */
double orc_poly6(double h, double r) { return Poly6<double>(h)(r); }
void orc_spiky_grad(double h, const double* r3, double* out3) {
    V3<double> g = SpikyGrad<double>(h)(mk(r3[0], r3[1], r3[2]));
    out3[0] = g.x; out3[1] = g.y; out3[2] = g.z;
}
double orc_adhesion(double h, double r) { return Adhesion<double>(h)(r); }
double orc_bezier(double py0, double py1, double px1, double py2, double sr) { return Bezier<double>(py0, py1, px1, py2)(sr); }
void orc_dem_force(const double* d3, const double* v3, double rij, double kn, double ks, double tanphi, double* out3) {
    V3<double> f = dem_force(mk(d3[0], d3[1], d3[2]), mk(v3[0], v3[1], v3[2]), rij, kn, ks, tanphi);
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
void orc_capillary_force(const double* d3, double R, double sr, const double* bez4, double* out3) {
    Bezier<double> G(bez4[0], bez4[1], bez4[2], bez4[3]);
    CapConst<double> K(R);
    V3<double> f = capillary_force(mk(d3[0], d3[1], d3[2]), R, sr, G, K);
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
double orc_capillary_rupture(double R) { return CapConst<double>(R).srup; }
void orc_world_boundary_force(const double* p3, const double* v3, double R, double Rb, double young, double poisson,
                              double tanphi, const double* lo3, const double* hi3, double* out3) {
    V3<double> f = mk(0.0, 0.0, 0.0);
    world_boundary_force(f, mk(p3[0], p3[1], p3[2]), mk(v3[0], v3[1], v3[2]), R, Rb, young, poisson, tanphi,
                         mk(lo3[0], lo3[1], lo3[2]), mk(hi3[0], hi3[1], hi3[2]));
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
/* FP32 hash of packed float positions */
void orc_hash(const float* pos3, size_t n, const float* lo3, float h, const int* g3, uint32_t* keys) {
    Grid G{lo3[0], lo3[1], lo3[2], h, g3[0], g3[1], g3[2]};
    for (size_t i = 0; i < n; ++i) {
        int cx, cy, cz;
        cell_xyz(G, pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], cx, cy, cz);
        keys[i] = cell_key(G, cx, cy, cz);
    }
}
/* BuildWorldBoundary cuda_boundary_emitter.cpp:12-80: loops `for (size_t i = 0; i <= sides.x; ++i)`
 * compare a size_t with a float; pos = lowest + i * spacing in FP32. */
size_t orc_world_boundary(const float* lo3, const float* hi3, float r, float* pos3, uint64_t* lab) {
    float sp = r * 2.f;
    float sx = (hi3[0] - lo3[0]) / sp, sy = (hi3[1] - lo3[1]) / sp, sz = (hi3[2] - lo3[2]) / sp;
    size_t k = 0;
    auto put = [&](float x, float y, float z) {
        if (pos3) { pos3[3 * k] = x; pos3[3 * k + 1] = y; pos3[3 * k + 2] = z; lab[k] = 0; }
        ++k;
    };
    for (size_t i = 0; (float)i <= sx; ++i) for (size_t j = 0; (float)j <= sz; ++j) put(lo3[0] + (float)i * sp, lo3[1], lo3[2] + (float)j * sp);
    for (size_t i = 0; (float)i <= sx; ++i) for (size_t j = 0; (float)j <= sz; ++j) put(lo3[0] + (float)i * sp, hi3[1], lo3[2] + (float)j * sp);
    for (size_t i = 0; (float)i <= sx; ++i) for (size_t j = 0; (float)j <= sy; ++j) put(lo3[0] + (float)i * sp, lo3[1] + (float)j * sp, lo3[2]);
    for (size_t i = 0; (float)i <= sx; ++i) for (size_t j = 0; (float)j <= sy; ++j) put(lo3[0] + (float)i * sp, lo3[1] + (float)j * sp, hi3[2]);
    for (size_t i = 0; (float)i <= sy; ++i) for (size_t j = 0; (float)j <= sz; ++j) put(lo3[0], lo3[1] + (float)i * sp, lo3[2] + (float)j * sp);
    for (size_t i = 0; (float)i <= sy; ++i) for (size_t j = 0; (float)j <= sz; ++j) put(hi3[0], lo3[1] + (float)i * sp, lo3[2] + (float)j * sp);
    return k;
}

/* ---- system handle ---- */
void* orc_create(int precision, const sf_params* p, const sf_boundary_params* b, const sf_solver_options* o) {
    if (!p || !b) { fail(-1, "null params"); return nullptr; }
    Handle* h = new Handle();
    h->prec = precision == 32 ? 32 : 64;
    auto init = [&](auto* s) {
        s->P = *p;
        s->B = *b;
        s->O = o ? *o : sf_solver_options{0.f, 3.f, 100.f};
        s->G = Grid{b->lowest_point.x, b->lowest_point.y, b->lowest_point.z, b->kernel_radius,
                    b->grid_size.x, b->grid_size.y, b->grid_size.z};
        s->C = (size_t)b->grid_size.x * b->grid_size.y * b->grid_size.z;
        s->multi = p->sf_type == SF_TYPE_MULTI_SF;
        s->cell_start_from_keys({}, s->C, s->cellStart);
        s->cell_start_from_keys({}, s->C, s->bcellStart);
        h->s = s;
    };
    if (h->prec == 32) init(new Sys<float>()); else init(new Sys<double>());
    return h;
}
void orc_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return;
    delete h->s;
    delete h;
}
void orc_set_params(void* hv, const sf_params* p) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) { S<float>(h)->P = *p; S<float>(h)->multi = p->sf_type == SF_TYPE_MULTI_SF; }
    else { S<double>(h)->P = *p; S<double>(h)->multi = p->sf_type == SF_TYPE_MULTI_SF; }
}

extern "C++" {
template <class T>
static void upload_T(Sys<T>* s, size_t n, const float* pos3, const float* vel3, const float* col3, const uint64_t* label,
                     const float* mass, const float* radius, const float* cda3, const float* amc2) {
    s->resize(n);
    for (size_t i = 0; i < n; ++i) {
        s->pos[i] = mk<T>(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]);
        if (vel3) s->vel[i] = mk<T>(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]);
        if (col3) s->col[i] = mk<T>(col3[3 * i], col3[3 * i + 1], col3[3 * i + 2]);
        s->label[i] = label[i];
        s->mass[i] = mass[i];
        s->radius[i] = radius ? radius[i] : T(0);
        if (cda3) s->cda[i] = mk<T>(cda3[3 * i], cda3[3 * i + 1], cda3[3 * i + 2]);
        if (amc2) { s->amc0[i] = amc2[2 * i]; s->amc1[i] = amc2[2 * i + 1]; }
    }
}
}
void orc_upload_particles(void* hv, size_t n, const float* pos3, const float* vel3, const float* col3, const uint64_t* label,
                          const float* mass, const float* radius, const float* cda3, const float* amc2) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) upload_T(S<float>(h), n, pos3, vel3, col3, label, mass, radius, cda3, amc2);
    else upload_T(S<double>(h), n, pos3, vel3, col3, label, mass, radius, cda3, amc2);
}
/* state in double precision (positions etc. as the oracle holds them) */
extern "C++" {
template <class T>
static void upload_state_T(Sys<T>* s, size_t n, const double* pos3, const double* vel3) {
    for (size_t i = 0; i < n && i < s->n; ++i) {
        s->pos[i] = mk<T>((T)pos3[3 * i], (T)pos3[3 * i + 1], (T)pos3[3 * i + 2]);
        s->vel[i] = mk<T>((T)vel3[3 * i], (T)vel3[3 * i + 1], (T)vel3[3 * i + 2]);
    }
}
}
void orc_upload_boundary(void* hv, size_t nb, const float* pos3, const uint64_t* label) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) S<float>(h)->upload_boundary(pos3, label, nb);
    else S<double>(h)->upload_boundary(pos3, label, nb);
}
void orc_build_neighbors(void* hv) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) S<float>(h)->build_neighbors(); else S<double>(h)->build_neighbors();
}
void orc_step(void* hv, float dt, int nsteps) {
    Handle* h = (Handle*)hv;
    for (int k = 0; k < nsteps; ++k) {
        if (h->prec == 32) S<float>(h)->step(dt); else S<double>(h)->step((double)dt);
    }
}
void orc_update_system(void* hv, float render_interval) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) S<float>(h)->update_system(render_interval); else S<double>(h)->update_system(render_interval);
}
extern "C++" {
static void norm3(const float v[3], float out[3]) {
    const float inv = 1.f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int c = 0; c < 3; ++c) out[c] = v[c] * inv;
}
static void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}
template <class T>
static void set_emitter_T(Sys<T>* s, int enable, int run, const float* pos3, const float* vel3, const float* col3,
                          int type, float radius, float width, float height, size_t cap) {
    auto& e = s->em;
    const size_t counter = e.counter, emitted = e.emitted;
    const float elapsed = e.elapsed;
    e = typename Sys<T>::Emitter();
    e.counter = counter; /* system state (CudaSFSystem members), kept across reconfiguration */
    e.emitted = emitted;
    e.elapsed = elapsed;
    e.enable = enable != 0;
    e.run = run != 0;
    e.cap = cap;
    for (int c = 0; c < 3; ++c) {
        e.pos[c] = pos3[c];
        e.vel[c] = vel3[c];
        e.col[c] = col3[c];
    }
    float axis[3], t[3];
    norm3(e.vel, axis);
    const float ex[3] = {1.f, 0.f, 0.f}, ey[3] = {0.f, 1.f, 0.f};
    if (std::fabs(axis[0]) == 1.f && std::fabs(axis[1]) == 0.f && std::fabs(axis[2]) == 0.f) cross3(axis, ey, t);
    else cross3(axis, ex, t);
    norm3(t, e.a1);
    cross3(axis, e.a1, t);
    norm3(t, e.a2);
    if (!e.enable) return;
    const float offset = s->P.sph_particle_radius * 2.f;
    const float a = type == 2 ? width : radius, b = type == 2 ? height : radius;
    for (float i = -a; i < a; i += offset)
        for (float j = -b; j < b; j += offset) {
            if (type == 1 && !(std::sqrt(i * i + j * j) <= a)) continue; /* circle: length(p) <= radius */
            e.s1.push_back(i);
            e.s2.push_back(j);
        }
}
}
/* type 0 square, 1 circle, 2 rectangle; cap = the handle's max particle count */
void orc_set_emitter(void* hv, int enable, int run, const float* pos3, const float* vel3, const float* col3, int type,
                     float radius, float width, float height, size_t cap) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) set_emitter_T(S<float>(h), enable, run, pos3, vel3, col3, type, radius, width, height, cap);
    else set_emitter_T(S<double>(h), enable, run, pos3, vel3, col3, type, radius, width, height, cap);
}
void orc_emitter_state(void* hv, int* enabled, size_t* points, size_t* emitted) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) {
        auto& e = S<float>(h)->em;
        *enabled = e.enable; *points = e.s1.size(); *emitted = e.emitted;
    } else {
        auto& e = S<double>(h)->em;
        *enabled = e.enable; *points = e.s1.size(); *emitted = e.emitted;
    }
}
size_t orc_num_sub_steps(void* hv) {
    Handle* h = (Handle*)hv;
    return h->prec == 32 ? S<float>(h)->nsub : S<double>(h)->nsub;
}
size_t orc_size(void* hv) {
    Handle* h = (Handle*)hv;
    return h->prec == 32 ? S<float>(h)->n : S<double>(h)->n;
}

/* field download: float-valued fields as double[k*n], integer fields as uint32 / uint64 */
extern "C++" {
template <class T>
static int get_T(Sys<T>* s, int field, void* dst) {
    auto v3 = [&](const std::vector<V3<T>>& a) {
        double* d = (double*)dst;
        for (size_t i = 0; i < a.size(); ++i) { d[3 * i] = a[i].x; d[3 * i + 1] = a[i].y; d[3 * i + 2] = a[i].z; }
        return 0;
    };
    auto v1 = [&](const std::vector<T>& a) {
        double* d = (double*)dst;
        for (size_t i = 0; i < a.size(); ++i) d[i] = a[i];
        return 0;
    };
    auto u32 = [&](const std::vector<uint32_t>& a) { std::memcpy(dst, a.data(), a.size() * 4); return 0; };
    switch (field) {
    case SF_FIELD_POS: return v3(s->pos);
    case SF_FIELD_VEL: return v3(s->vel);
    case SF_FIELD_ACC: return v3(s->acc);
    case SF_FIELD_COL: return v3(s->col);
    case SF_FIELD_LABEL: std::memcpy(dst, s->label.data(), s->n * 8); return 0;
    case SF_FIELD_MASS: return v1(s->mass);
    case SF_FIELD_RADIUS: return v1(s->radius);
    case SF_FIELD_DENSITY: return v1(s->density);
    case SF_FIELD_PRESSURE: return v1(s->pressure);
    case SF_FIELD_VOIDAGE: return v1(s->voidage);
    case SF_FIELD_SATURATION: return v1(s->saturation);
    case SF_FIELD_MAX_SATURATION: return v1(s->maxsat);
    case SF_FIELD_AVG_FLOW_VEL: return v3(s->avgflow);
    case SF_FIELD_AVG_DRAG_FORCE: {
        std::vector<V3<T>> t(s->avgdrag);
        for (size_t i = 0; i < s->n; ++i) if (s->label[i] != 1) t[i] = mk<T>(0, 0, 0);
        return v3(t);
    }
    case SF_FIELD_ADHESION_FORCE: return v3(s->adh);
    case SF_FIELD_AVG_ADHESION_FORCE: return v3(s->avgadh);
    case SF_FIELD_CDA0ASAT: return v3(s->cda);
    case SF_FIELD_PARTICLE2CELL: return u32(s->keys);
    case SF_FIELD_SORT_PERM: return u32(s->perm);
    case SF_FIELD_SORTED_KEYS: return u32(s->sorted_keys);
    case SF_FIELD_CELL_START: return u32(s->cellStart);
    case SF_FIELD_NBR_COUNT: { std::vector<uint32_t> a, b; s->neighbor_counts(a, b); return u32(a); }
    case SF_FIELD_NBR_CANDIDATES: { std::vector<uint32_t> a, b; s->neighbor_counts(a, b); return u32(b); }
    case SF_FIELD_B_POS: return v3(s->bpos);
    case SF_FIELD_B_LABEL: std::memcpy(dst, s->blabel.data(), s->nb * 8); return 0;
    case SF_FIELD_B_VOLUME: return v1(s->bvol);
    case SF_FIELD_B_CELL_START: return u32(s->bcellStart);
    default: return fail(-6, "unknown field");
    }
}
}
int orc_get(void* hv, int field, void* dst) {
    Handle* h = (Handle*)hv;
    return h->prec == 32 ? get_T(S<float>(h), field, dst) : get_T(S<double>(h), field, dst);
}
/* upload slot of each particle in the current order */
void orc_get_ids(void* hv, uint32_t* out) {
    Handle* h = (Handle*)hv;
    const auto& v = h->prec == 32 ? S<float>(h)->ids : S<double>(h)->ids;
    std::memcpy(out, v.data(), v.size() * 4);
}
/* initial maxSaturation history (current order) */
void orc_set_maxsat(void* hv, const float* ms) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) { auto* s = S<float>(h); for (size_t i = 0; i < s->n; ++i) s->maxsat[i] = ms[i]; }
    else { auto* s = S<double>(h); for (size_t i = 0; i < s->n; ++i) s->maxsat[i] = ms[i]; }
}
/* overwrite positions / velocities (current order) from double arrays */
void orc_set_state(void* hv, const double* pos3, const double* vel3) {
    Handle* h = (Handle*)hv;
    if (h->prec == 32) upload_state_T(S<float>(h), S<float>(h)->n, pos3, vel3);
    else upload_state_T(S<double>(h), S<double>(h)->n, pos3, vel3);
}

/* FP64 forward-mode tangent of a fixed-dt window (up to 4 directions at once).
 * Direction q seeds d(initial velocity) = vdot3[q], d(initial position) =
 * xdot3[q] (n x 3, upload order; either may be NULL) and d(sf_cd, sf_c0,
 * sf_cmc, dem_tan_friction_angle) = pdot4[q].
 * Outputs in upload order: final positions, the last step's saturation and
 * their tangents (posdot3 / satdot: ndir blocks). */
int orc_tangent_window(const sf_params* p, const sf_boundary_params* b, const sf_solver_options* o, size_t n,
                       const float* pos3, const float* vel3, const uint64_t* label, const float* mass,
                       const float* radius, const float* maxsat, size_t nb, const float* bpos3, const uint64_t* blabel,
                       float dt, int steps, int ndir, const double* vdot3, const double* xdot3, const double* pdot4,
                       double* pos_out3,
                       double* sat_out, double* posdot3, double* satdot) {
    constexpr int N = 4;
    using D = Dual<N>;
    if (ndir < 0 || ndir > N) return fail(-1, "ndir must be in [0, 4]");
    if (p->sf_type == SF_TYPE_MULTI_SF) return fail(-1, "tangent runs need SF mode");
    Sys<D> s;
    s.P = *p;
    s.B = *b;
    s.O = o ? *o : sf_solver_options{0.f, 3.f, 100.f};
    s.G = Grid{b->lowest_point.x, b->lowest_point.y, b->lowest_point.z, b->kernel_radius, b->grid_size.x,
               b->grid_size.y, b->grid_size.z};
    s.C = (size_t)b->grid_size.x * b->grid_size.y * b->grid_size.z;
    s.multi = false;
    s.cell_start_from_keys({}, s.C, s.cellStart);
    s.cell_start_from_keys({}, s.C, s.bcellStart);
    upload_T(&s, n, pos3, vel3, nullptr, label, mass, radius, nullptr, nullptr);
    if (maxsat)
        for (size_t i = 0; i < n; ++i) s.maxsat[i] = maxsat[i];
    for (int q = 0; q < ndir; ++q)
        for (size_t i = 0; i < n; ++i) {
            if (vdot3) {
                s.vel[i].x.d[q] = vdot3[(q * n + i) * 3];
                s.vel[i].y.d[q] = vdot3[(q * n + i) * 3 + 1];
                s.vel[i].z.d[q] = vdot3[(q * n + i) * 3 + 2];
            }
            if (xdot3) { /* the grid hash reads the value only */
                s.pos[i].x.d[q] = xdot3[(q * n + i) * 3];
                s.pos[i].y.d[q] = xdot3[(q * n + i) * 3 + 1];
                s.pos[i].z.d[q] = xdot3[(q * n + i) * 3 + 2];
            }
        }
    const float pv[4] = {p->sf_cd, p->sf_c0, p->sf_cmc, p->dem_tan_friction_angle};
    s.th_set = true;
    for (int c = 0; c < 4; ++c) {
        s.th[c] = D(pv[c]);
        for (int q = 0; q < ndir; ++q) s.th[c].d[q] = pdot4 ? pdot4[4 * q + c] : 0.0;
    }
    if (nb) s.upload_boundary(bpos3, blabel, nb);
    for (int k = 0; k < steps; ++k) s.step(D((double)dt));
    for (size_t k = 0; k < n; ++k) {
        const size_t id = s.ids[k];
        const V3<D>& x = s.pos[k];
        pos_out3[3 * id] = x.x.v;
        pos_out3[3 * id + 1] = x.y.v;
        pos_out3[3 * id + 2] = x.z.v;
        sat_out[id] = s.saturation[k].v;
        for (int q = 0; q < ndir; ++q) {
            posdot3[(q * n + id) * 3] = x.x.d[q];
            posdot3[(q * n + id) * 3 + 1] = x.y.d[q];
            posdot3[(q * n + id) * 3 + 2] = x.z.d[q];
            satdot[q * n + id] = s.saturation[k].d[q];
        }
    }
    return 0;
}

} /* extern "C" */
