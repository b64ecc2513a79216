// sf_emitter.cpp -- the inflow emitter of the seepage system, host side:
// sample points of CudaEmitter (cuda_emitter.cpp:12-97, axes from the ctor
// cuda_emitter.cuh:33-55), the emission schedule of CudaSFSystem::
// OnUpdateSolver (cuda_sf_system.cu:86-115) and AddSphParticles
// (cuda_sf_particles.cu:50-68).  The float arithmetic is the reference's host
// arithmetic: written order, no fused multiply-add, normalize(v) =
// v * (1 / sqrtf(dot(v, v))) (helper_math.h:79-82, 1338-1342).
#include <cmath>
#include <vector>

#include "sf_system.h"

namespace sf {
namespace {

struct F3 {
    float x, y, z;
};
F3 f3(const sf_float3& a) { return {a.x, a.y, a.z}; }
sf_float3 sf3(const F3& a) { return {a.x, a.y, a.z}; }
float dot3(F3 a, F3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
F3 cross3(F3 a, F3 b) {
#pragma clang fp contract(off)
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
F3 normalize3(F3 v) {
#pragma clang fp contract(off)
    const float inv = 1.f / std::sqrt(dot3(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

// The reference's float loops: for (i = -a; i < a; i += 2r) for (j = -b; j < b; j += 2r)
void grid_samples(float r, float a, float b, bool circle, std::vector<float>& out) {
#pragma clang fp contract(off)
    out.clear();
    const float offset = r * 2.f;
    for (float i = -a; i < a; i += offset)
        for (float j = -b; j < b; j += offset) {
            if (circle && !(std::sqrt(i * i + j * j) <= a)) continue;
            out.push_back(i);
            out.push_back(j);
        }
}

// AddSphParticles: water of the emitter's velocity and colour, appended after
// the current particles (the next neighbour build sorts them in).
int append_water(sf_system* h, const std::vector<F3>& p) {
    const size_t k = p.size();
    std::vector<float4> pm(k), vr(k), cs(k), cda(k);
    std::vector<float2> amc(k);
    const EmitterState& e = h->emitter;
    for (size_t q = 0; q < k; ++q) {
        pm[q] = make_float4(p[q].x, p[q].y, p[q].z, h->P.sph_mass);
        vr[q] = make_float4(e.vel.x, e.vel.y, e.vel.z, h->P.sph_particle_radius);
        cs[q] = make_float4(e.col.x, e.col.y, e.col.z, 0.f);
        cda[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        amc[q] = make_float2(0.f, 0.f);
    }
    const size_t n = h->n;
    hipStream_t s = h->stream;
    SF_HIP(hipMemcpyAsync(h->pm[h->cur].p + n, pm.data(), k * sizeof(float4), hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(h->vr[h->cur].p + n, vr.data(), k * sizeof(float4), hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(h->cs[h->cur_cs].p + n, cs.data(), k * sizeof(float4), hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(h->cda[h->cur_aux].p + n, cda.data(), k * sizeof(float4), hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(h->amc[h->cur_aux].p + n, amc.data(), k * sizeof(float2), hipMemcpyHostToDevice, s));
    SF_HIP(hipStreamSynchronize(s));
    h->n = n + k;
    if (k && h->water_mass_state == 0) {
        h->water_mass_state = 1;
        h->water_mass = h->P.sph_mass;
    } else if (k && h->water_mass != h->P.sph_mass) {
        h->water_mass_state = 2;
    }
    h->built = false;  // the new slots have no previous keys: full sort next build
    h->drop_graph();
    diff_release(h->diff);  // a recorded window no longer matches the particles
    h->diff = nullptr;
    return SF_OK;
}

// CudaEmitter::Emit: emitPosition + s1 * axis1 + s2 * axis2, then the
// capacity check of OnUpdateSolver (Size() + count < MaxSize()).
int emit(sf_system* h) {
#pragma clang fp contract(off)
    EmitterState& e = h->emitter;
    const size_t k = e.samples.size() / 2;
    if (h->n + k >= h->cap) {
        e.enable = false;
        return SF_OK;
    }
    std::vector<F3> p(k);
    const F3 o = f3(e.pos), a1 = f3(e.axis1), a2 = f3(e.axis2);
    for (size_t q = 0; q < k; ++q) {
        const float s1 = e.samples[2 * q], s2 = e.samples[2 * q + 1];
        p[q] = {o.x + s1 * a1.x + s2 * a2.x, o.y + s1 * a1.y + s2 * a2.y, o.z + s1 * a1.z + s2 * a2.z};
    }
    e.emitted += k;
    return append_water(h, p);
}

}  // namespace

bool emitter_wcsph_clock(const sf_system* h) {
    const EmitterState& e = h->emitter;
    return e.enable && e.run && !e.samples.empty() && h->P.solver_type == SF_WCSPH_SOLVER;
}

// The WCSPH schedule's test, on the clock as it stands (cuda_sf_system.cu:95-104).
int emitter_wcsph_check(sf_system* h) {
#pragma clang fp contract(off)
    EmitterState& e = h->emitter;
    const float speed = std::sqrt(dot3(f3(e.vel), f3(e.vel)));
    const float two_r = 2.f * h->P.sph_particle_radius;
    if ((speed * e.elapsed) / two_r >= static_cast<float>(e.counter)) {
        const int rc = emit(h);
        if (rc) return rc;
        ++e.counter;
    }
    return SF_OK;
}

// Sub-steps after which the WCSPH test is certainly still false when each
// advances the clock by at most step_max: the float test
// fl(fl(speed * t) / 2r) >= counter is below counter while
// speed * t / 2r * (1 + u)^2 < counter, and t grows by at most
// (t + step_max)(1 + u) per sub-step (u = 2^-24, the float rounding bound).
size_t emitter_quiet_steps(const sf_system* h, double step_max, size_t cap) {
    const EmitterState& e = h->emitter;
    const double speed = (double)std::sqrt(dot3(f3(e.vel), f3(e.vel)));
    const double two_r = (double)(2.f * h->P.sph_particle_radius);
    const double u = 1.0 / 16777216.0;
    const double counter = (double)e.counter;
    if (speed * (double)e.elapsed / two_r * (1 + u) * (1 + u) >= counter) return 0;
    double t = (double)e.elapsed;
    size_t j = 0;
    while (j < cap) {
        t = (t + step_max) * (1 + u);
        if (speed * t / two_r * (1 + u) * (1 + u) >= counter) break;
        ++j;
    }
    return j;
}

int emit_after_update(sf_system* h, float render_interval) {
#pragma clang fp contract(off)
    EmitterState& e = h->emitter;
    if (!e.enable || !e.run || e.samples.empty()) return SF_OK;
    const float speed = std::sqrt(dot3(f3(e.vel), f3(e.vel)));
    const float two_r = 2.f * h->P.sph_particle_radius;
    if (h->P.solver_type == SF_WCSPH_SOLVER) {
        e.elapsed += render_interval / static_cast<float>(h->nsub);
        return emitter_wcsph_check(h);
    } else {
        // size_t(2 r / (|v| dt)); a period below one step emits every call
        // (the reference would divide by zero)
        const float q = two_r / (speed * h->P.dt);
        // (a quotient beyond size_t, from a near-zero speed, never comes round)
        size_t period = q < 1.8e19f ? (size_t)q : SIZE_MAX;
        if (period == 0) period = 1;
        if (e.counter++ % period == 0) return emit(h);
    }
    return SF_OK;
}

}  // namespace sf

using namespace sf;

extern "C" {

int sf_set_emitter(sf_handle h, const sf_emitter_params* p) {
#pragma clang fp contract(off)
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!p) return set_error(SF_E_INVALID, "null emitter parameters");
    if (h->nscenes) return set_error(SF_E_STATE, "batched handles have no emitter");
    const float v[12] = {p->emit_pos.x, p->emit_pos.y, p->emit_pos.z, p->emit_vel.x, p->emit_vel.y, p->emit_vel.z,
                         p->emit_col.x, p->emit_col.y, p->emit_col.z, p->emit_radius, p->emit_width, p->emit_height};
    for (float x : v)
        if (!std::isfinite(x)) return set_error(SF_E_INVALID, "emitter parameters must be finite");
    if (p->emit_type < SF_EMITTER_SQUARE || p->emit_type > SF_EMITTER_RECTANGLE)
        return set_error(SF_E_INVALID, "emit_type must be SF_EMITTER_SQUARE, _CIRCLE or _RECTANGLE");
    const F3 vel = f3(p->emit_vel);
    if (dot3(vel, vel) == 0.f) return set_error(SF_E_INVALID, "emit_vel must not be zero");
    const float r = h->P.sph_particle_radius;
    if (!(r > 0.f)) return set_error(SF_E_INVALID, "sph_particle_radius must be > 0");
    EmitterState e;
    // the emission counter, elapsed time and count belong to the system
    // (CudaSFSystem members, zero from its construction), not to the emitter
    e.counter = h->emitter.counter;
    e.elapsed = h->emitter.elapsed;
    e.emitted = h->emitter.emitted;
    e.enable = p->enable != 0;
    e.run = p->run != 0;
    e.pos = p->emit_pos;
    e.vel = p->emit_vel;
    e.col = p->emit_col;
    const F3 axis = normalize3(vel);
    const F3 a1 = (std::fabs(axis.x) == 1.f && std::fabs(axis.y) == 0.f && std::fabs(axis.z) == 0.f)
                      ? normalize3(cross3(axis, F3{0.f, 1.f, 0.f}))
                      : normalize3(cross3(axis, F3{1.f, 0.f, 0.f}));
    e.axis1 = sf3(a1);
    e.axis2 = sf3(normalize3(cross3(axis, a1)));
    if (e.enable) {  // the CudaSFSystem ctor builds the samples only for an enabled emitter
        // bounded loops: at most 4096 steps per axis, each step advancing the
        // float counter (at huge extents i += 2r could round back to i)
        const float ea = p->emit_type == SF_EMITTER_RECTANGLE ? p->emit_width : p->emit_radius;
        const float eb = p->emit_type == SF_EMITTER_RECTANGLE ? p->emit_height : p->emit_radius;
        for (float ext : {ea, eb})
            if (std::fabs(ext) / r > 4096.f || -std::fabs(ext) + 2.f * r == -std::fabs(ext))
                return set_error(SF_E_INVALID, "emitter extent must stay within 4096 particle spacings");
        switch (p->emit_type) {
        case SF_EMITTER_SQUARE: grid_samples(r, p->emit_radius, p->emit_radius, false, e.samples); break;
        case SF_EMITTER_CIRCLE: grid_samples(r, p->emit_radius, p->emit_radius, true, e.samples); break;
        default: grid_samples(r, p->emit_width, p->emit_height, false, e.samples); break;
        }
    }
    h->emitter = std::move(e);
    return SF_OK;
}

int sf_get_emitter(sf_handle h, int* enabled, size_t* num_points, size_t* emitted) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (enabled) *enabled = h->emitter.enable ? 1 : 0;
    if (num_points) *num_points = h->emitter.samples.size() / 2;
    if (emitted) *emitted = h->emitter.emitted;
    return SF_OK;
}

}  // extern "C"
