// sf_api.cpp -- the C-ABI entry points that own device memory and order the
// passes of one step on the handle's HIP stream.
//
// Per step (sf_step):  hash -> radix sort -> cellStart -> gather (state is
// physically reordered, as CudaGNSearcher::SortData does) -> P1..P5 -> advect.
// sf_update_system adds the device-side CFL sub-step count of the WCSPH solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sf_system.h"

namespace sf {

int hip_error(hipError_t e, const char* what) {
    return set_error(SF_E_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
int validate(const sf_params* p, const sf_boundary_params* b);
int make_dirs(const std::string& path);

// Constants exactly as the reference's functor constructors form them in FP32.
static SFConst make_const(const sf_params& P, const sf_boundary_params& B, const sf_solver_options& O, int blabel1) {
    const float PI = 3.141592653589793238462643383279502884197f;  // KIRI_PI
    SFConst k{};
    k.lox = B.lowest_point.x;
    k.loy = B.lowest_point.y;
    k.loz = B.lowest_point.z;
    k.hix = B.highest_point.x;
    k.hiy = B.highest_point.y;
    k.hiz = B.highest_point.z;
    k.h = B.kernel_radius;
    k.gx = B.grid_size.x;
    k.gy = B.grid_size.y;
    k.gz = B.grid_size.z;
    k.C = (uint32_t)((size_t)k.gx * k.gy * k.gz);
    const float h = k.h;
    k.h2 = h * h;
    {  // Poly6Kernel (cuda_sph_kernel.cuh:82-89)
        float ih = 1.f / h, ih3 = ih * ih * ih, ih9 = ih3 * ih3 * ih3;
        k.w_coef = 315.f * ih9 / (64.f * PI);
        k.w0 = k.w_coef * k.h2 * k.h2 * k.h2;
    }
    {  // SpikyKernelGrad (:103-108)
        float h2 = h * h, h6 = h2 * h2 * h2;
        k.g_coef = -45.f / (PI * h6);
    }
    k.adh_coef = 0.007f / powf(h, 3.25f);  // AdhesionAkinci13 (:192-195)
    k.ih = 1.f / h;
    k.rho0 = P.sph_density;
    k.stiff = P.sph_stiff;
    k.tait = P.solver_type == SF_WCSPH_SOLVER ? 1 : 0;
    k.neg_scale = k.tait ? O.negative_scale : 1.f;
    k.nu = P.sph_nu;
    k.bnu = P.sph_bnu;
    k.rw = P.sph_particle_radius;
    k.rho_s = P.dem_density;
    k.R = P.dem_particle_radius;
    k.iR = 1.f / k.R;
    k.young = P.dem_young;
    k.poisson = P.dem_poisson;
    k.kn = P.dem_young * k.R;
    k.ks = k.kn * P.dem_poisson;
    k.tanphi = P.dem_tan_friction_angle;
    k.tan2 = powf(k.tanphi, 2.f);
    k.damping = P.dem_damping;
    k.cd = P.sf_cd;
    {  // liquid bridge constants (cuda_dem_solver_common_gpu.cuh:63-65)
        float theta = 30.f / 180.f * PI;
        float vb = 4.f / 3.f * PI * powf(k.R, 3.f) * 0.01f * 0.01f;
        k.cap_vb = vb;
        k.cap_srup = (1.f + 0.5f * theta) * (powf(vb, 1.f / 3.f) + 0.1f * powf(vb, 2.f / 3.f));
        k.cap_vb_piR = vb / (PI * k.R);
        k.cap_cos = cosf(theta);
        k.cap_sin = sinf(theta);
        k.cap_pi = PI;
    }
    auto bez = [](float py0, float py1, float px1, float py2) {  // QuadraticBezierCoeff ctor (cuda_sf_utils.cuh:42-57)
        Bez g;
        g.py0 = py0;
        g.py1 = py1;
        g.px1 = px1;
        g.py2 = py2;
        float a = (px1 + 1.f) / 4.f;
        g.b = -2.f * a;
        g.c = g.b * g.b;
        g.d = -4.f * (1.f + g.b - px1);
        g.e = 2.f * (1.f + g.b - px1);
        g.ipx1 = 1.f / px1;
        g.ie = 1.f / g.e;
        return g;
    };
    k.gcap = bez(P.sf_c0, P.sf_cmc, P.sf_cmc_p, P.sf_csat);  // cuda_sph_sf_solver.cu:249,308
    k.multi = P.sf_type == SF_TYPE_MULTI_SF ? 1 : 0;
    k.gadh = k.multi ? bez(0.f, 1.5f, 0.5f, 0.8f)                             // cuda_sph_sf_solver_gpu.cuh:667
                     : bez(P.sf_a0, P.sf_amc, P.sf_amc_p, P.sf_asat);         // cuda_sph_sf_solver.cu:119
    k.gvx = P.gravity.x;
    k.gvy = P.gravity.y;
    k.gvz = P.gravity.z;
    k.grav_abs = fabsf(P.gravity.y);  // abs(params.gravity.y) (cuda_wcsph_sf_solver.cpp:106)
    k.dryr = P.sf_dry_sand_color.x;
    k.dryg = P.sf_dry_sand_color.y;
    k.dryb = P.sf_dry_sand_color.z;
    k.wetr = P.sf_wet_sand_color.x;
    k.wetg = P.sf_wet_sand_color.y;
    k.wetb = P.sf_wet_sand_color.z;
    k.blabel1 = blabel1;
    return k;
}

}  // namespace sf

using namespace sf;


namespace sf {

int ensure_sort_capacity(sf_system* h, size_t m) {
    if (m <= h->sort_cap && h->kA.p) return SF_OK;
    size_t c = std::max<size_t>(m, 1);
    int rc;
    if ((rc = h->keys.alloc(c)) || (rc = h->kA.alloc(c)) || (rc = h->kB.alloc(c)) || (rc = h->vA.alloc(c)) ||
        (rc = h->vB.alloc(c)))
        return rc;
    size_t nblocks = (c + kRsTile - 1) / kRsTile;
    if ((rc = h->hist.alloc(256 * nblocks))) return rc;
    if ((rc = h->wpre.alloc(c + 1)) || (rc = h->wpart.alloc((size_t)wprefix_blocks((int)c + 1) + 1)) ||
        (rc = h->plist.alloc(c)) || (rc = h->wflag.alloc(c)) || (rc = h->pph.alloc(c)) || (rc = h->ilist.alloc(c)) ||
        (rc = h->icount.alloc(2)) || (rc = h->iface.alloc(c)) || (rc = h->wl.alloc(c)) || (rc = h->vq.alloc(c)) || (rc = h->wq.alloc(c)) ||
        (rc = h->xq.alloc(c)) || (rc = h->wbq.alloc(c)) || (rc = h->rec0.alloc(c)) || (rc = h->rec1.alloc(c)))
        return rc;
    // the interface counters start at zero (each step's water momentum pass
    // clears them for the next)
    SF_HIP(hipMemsetAsync(h->icount.p, 0, 2 * sizeof(uint32_t), h->stream));
    if ((rc = h->skeys[0].alloc(c)) || (rc = h->skeys[1].alloc(c)) || (rc = h->sperm.alloc(c)) ||
        (rc = h->slot.alloc(c)) || (rc = h->in_list.alloc(c)) || (rc = h->stay_pre.alloc(c + 1)))
        return rc;
    h->sort_cap = c;
    return ensure_cell_capacity(h);
}

// arrays of the incremental sort sized by the cell count (C + 1) and the
// prefix partials for max(n, C) + 1 slots
int ensure_cell_capacity(sf_system* h) {
    const size_t C1 = (size_t)h->C + 1;
    int rc;
    if (h->in_count.n < C1) {
        if ((rc = h->in_count.alloc(C1)) || (rc = h->in_start.alloc(C1))) return rc;
        // zero before the first incremental sort; afterwards every cell-start pass clears it
        SF_HIP(hipMemsetAsync(h->in_count.p, 0, C1 * sizeof(uint32_t), h->stream));
    }
    const size_t np = (size_t)vprefix2_partials((int)h->sort_cap, (int)h->C);
    if (h->part2.n < np && (rc = h->part2.alloc(np))) return rc;
    return SF_OK;
}

template <class T>
int zero(sf_system* h, DevBuf<T>& b) {
    SF_HIP(hipMemsetAsync(b.p, 0, b.n * sizeof(T), h->stream));
    return SF_OK;
}

// ev (optional): events recorded before each pass of the step
#define SF_MARK(ev, k) \
    if ((ev) && (ev)[k]) SF_HIP(hipEventRecord((ev)[k], s))

int build_neighbors(sf_system* h, hipEvent_t* ev) {
    const int n = (int)h->n;
    hipStream_t s = h->stream;
    const bool inc = h->sort_inc && h->built && h->sorted_keys;
    const bool fused = inc && h->keys_fresh;  // the last advect hashed and flagged
    h->keys_fresh = false;
    SF_MARK(ev, SF_PASS_HASH);
    if (!fused) launch_hash(h->K, h->pm[h->cur].p, h->nscenes ? h->scene.p : nullptr, h->keys.p, n, s);
    SortBuffers sb{h->kA.p, h->vA.p, h->kB.p, h->vB.p, h->hist.p};
    SF_MARK(ev, SF_PASS_SORT);
    if (inc) {
        // the particles are in the last sort's order: rebuild the stable sort
        // from the previous keys and cellStart
        uint32_t* out = h->sorted_keys == h->skeys[0].p ? h->skeys[1].p : h->skeys[0].p;
        IncSortBuffers ib{h->in_count.p, h->in_start.p, h->slot.p, h->in_list.p, h->stay_pre.p, h->part2.p};
        launch_incremental_sort(h->keys.p, h->sorted_keys, h->cell.p, n, h->C, ib, h->sperm.p, out, fused, s);
        h->sorted_keys = out;
        h->perm = h->sperm.p;
    } else {
        int rc = radix_sort(h->keys.p, n, h->plan, sb, &h->sorted_keys, &h->perm, s);
        if (rc) return rc;
    }
    GatherArgs g;
    g.perm = h->perm;
    g.pm_in = h->pm[h->cur].p;
    g.vr_in = h->vr[h->cur].p;
    g.cs_in = h->cs[h->cur_cs].p;
    g.cda_in = h->cda[h->cur_aux].p;
    g.amc_in = h->amc[h->cur_aux].p;
    g.pm_out = h->pm[h->cur ^ 1].p;
    g.vr_out = h->vr[h->cur ^ 1].p;
    g.cs_out = h->cs[h->cur_cs ^ 1].p;
    g.cda_out = h->cda[h->cur_aux ^ 1].p;
    g.amc_out = h->amc[h->cur_aux ^ 1].p;
    g.wflag = h->wflag.p;
    g.multi = h->K.multi;
    SF_MARK(ev, SF_PASS_GATHER);
    launch_gather_wprefix(g, n, h->wpart.p, h->wpre.p, h->plist.p, s);
    SF_MARK(ev, SF_PASS_CELL_START);
    launch_cell_start(h->sorted_keys, n, h->C, h->wpre.p, h->cell.p, h->ct.p, h->in_count.p,
                      PhasePos{h->plist.p, g.pm_out, h->pph.p},
                      // the incremental sort wrote stay_pre and in_start only for n > 0
                      inc && n > 0 ? IncStart{h->stay_pre.p, h->in_start.p} : IncStart{nullptr, nullptr}, s);
    h->cur ^= 1;
    h->cur_cs ^= 1;
    if (h->K.multi) h->cur_aux ^= 1;
    h->built = true;
    h->build_pending = true;
    SF_HIP(hipGetLastError());
    return SF_OK;
}

int solve(sf_system* h, hipEvent_t* ev) { return solve_with(h, h->fields(), ev); }

int solve_with(sf_system* h, const Fields& f, hipEvent_t* ev) {
    const int n = (int)h->n;
    hipStream_t s = h->stream;
    h->empty_pending = !(f.full_empty || f.den4);
    h->n_built = h->n;
    SF_MARK(ev, SF_PASS_DENSITY);
    launch_density(h->K, f, n, h->cfg[0], s);
    SF_MARK(ev, SF_PASS_AVG_FLOW);
    launch_avg_flow(h->K, f, n, h->cfg[1], s);
    SF_MARK(ev, SF_PASS_VOIDAGE_ADHESION);
    launch_voidage_adhesion(h->K, f, n, h->cfg[2], s);
    SF_MARK(ev, SF_PASS_SAND_MOMENTUM);
    launch_sand_momentum(h->K, f, n, h->cfg[3], s);
    SF_MARK(ev, SF_PASS_WATER_MOMENTUM);
    launch_water_momentum(h->K, f, n, h->cfg[4], s);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

int advect(sf_system* h, float dt, const CflOut* dt_dev, hipEvent_t* ev) {
    hipStream_t s = h->stream;
    SF_MARK(ev, SF_PASS_ADVECT);
    AdvectArgs a;
    a.pm_in = h->pm[h->cur].p;
    a.vr_in = h->vr[h->cur].p;
    a.acc = h->acc.p;
    a.pm_out = h->pm[h->cur ^ 1].p;
    a.vr_out = h->vr[h->cur ^ 1].p;
    a.bpv = h->bpv.p;
    a.bcell = h->bcell.p;
    a.dt = dt;
    a.dt_dev = dt_dev;
    // the next build's hash and change flags, when it will be incremental
    const bool fuse = h->sort_inc && h->built && h->sorted_keys && !h->pm_shared;
    a.keys_out = fuse ? h->keys.p : nullptr;
    a.old_keys = h->sorted_keys;
    a.in_count = h->in_count.p;
    a.slot = h->slot.p;
    a.scene = h->nscenes ? h->scene.p : nullptr;
    launch_advect(h->K, a, (int)h->n, s);
    h->keys_fresh = fuse;
    h->build_pending = false;
    h->cur ^= 1;
    SF_MARK(ev, SF_PASS_COUNT);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

int one_step(sf_system* h, float dt, hipEvent_t* ev) {
    int rc;
    if ((rc = build_neighbors(h, ev))) return rc;
    if ((rc = solve(h, ev))) return rc;
    return advect(h, dt, nullptr, ev);
}

// CudaWCSphSFSolver::UpdateSolver after the neighbour build
// (cuda_wcsph_sf_solver.cpp:34-179): the passes, the CFL sub-step count for
// render_interval (or the frame's remaining time, see FrameDev) and the
// advect with dt = render_interval / nsub, all without a host round trip.
int wcsph_substep_tail(sf_system* h, float render_interval, FrameDev* frame) {
    int rc;
    if ((rc = solve(h))) return rc;
    launch_cfl(h->K, h->acc.p, (int)h->n, h->cfl_part.p, h->cfl_out.p, render_interval, h->P.sph_mass,
               h->O.speed_of_sound, h->O.time_step_limit_scale, h->P.dt, frame, h->stream);
    return advect(h, 0.f, h->cfl_out.p);
}

// size_t(q) of the reference's nSub = renderInterval / params.dt, at least 1;
// a quotient beyond 2^62 (which no frame could run) saturates there rather
// than converting out of range
size_t sub_step_count(float q) {
    if (!(q >= 1.f)) return 1;
    return q < 4.6e18f ? (size_t)q : ((size_t)1 << 62);
}

// CudaWCSphSFSolver::UpdateSolver / CudaSphSFSolver::UpdateSolver on the
// current neighbour build: the passes, the sub-step count and the advect;
// records ev1 after the advect and leaves h->nsub set (reads the WCSPH count
// back from the device).
int update_solver(sf_system* h, float render_interval) {
    int rc;
    if (h->P.solver_type == SF_WCSPH_SOLVER) {
        if ((rc = wcsph_substep_tail(h, render_interval, nullptr))) return rc;
        CflOut o;
        SF_HIP(hipMemcpyAsync(&o, h->cfl_out.p, sizeof(CflOut), hipMemcpyDeviceToHost, h->stream));
        SF_HIP(hipEventRecord(h->ev1, h->stream));
        SF_HIP(hipEventSynchronize(h->ev1));
        h->nsub = o.nsub;
    } else {
        // CudaSphSFSolver::UpdateSolver: nSub = renderInterval / params.dt, advect with params.dt
        h->nsub = sub_step_count(render_interval / h->P.dt);
        if ((rc = solve(h))) return rc;
        if ((rc = advect(h, h->P.dt, nullptr))) return rc;
        SF_HIP(hipEventRecord(h->ev1, h->stream));
        SF_HIP(hipEventSynchronize(h->ev1));
    }
    return SF_OK;
}

int check_handle(sf_system* h) {
    if (!h) return set_error(SF_E_INVALID, "null handle");
    return SF_OK;
}
int set_device(sf_system* h) {
    SF_HIP(hipSetDevice(h->device));
    return SF_OK;
}

// bit of each field that has a reference-layout view
constexpr uint32_t kViewable = (1u << 18) - 1u;  // SF_FIELD_POS .. SF_FIELD_AMCAMCP
constexpr uint32_t kPersistentViews = (1u << SF_FIELD_POS) | (1u << SF_FIELD_VEL) | (1u << SF_FIELD_COL) |
                                      (1u << SF_FIELD_LABEL) | (1u << SF_FIELD_MASS) | (1u << SF_FIELD_RADIUS) |
                                      (1u << SF_FIELD_MAX_SATURATION) | (1u << SF_FIELD_CDA0ASAT) |
                                      (1u << SF_FIELD_AMCAMCP);
constexpr uint32_t kIndexViews =
    (1u << SF_FIELD_SORTED_KEYS) | (1u << SF_FIELD_CELL_START) | (1u << SF_FIELD_B_CELL_START);

static ViewArgs view_args(sf_system* h, uint32_t bits) {
    auto v = [&](int f) -> void* { return (bits >> f) & 1u ? (void*)h->views[f].p : nullptr; };
    ViewArgs a;
    a.pm = h->pm[h->cur].p;
    a.vr = h->vr[h->cur].p;
    a.cs = h->cs[h->cur_cs].p;
    a.cda = h->cda[h->cur_aux].p;
    a.amc = h->amc[h->cur_aux].p;
    a.acc = h->acc.p;
    a.rp = h->rp.p;
    a.es = h->es.p;
    a.flow = h->flow.p;
    a.xf = h->xf.p;
    a.adh = h->adh.p;
    a.pos = (float*)v(SF_FIELD_POS);
    a.vel = (float*)v(SF_FIELD_VEL);
    a.accv = (float*)v(SF_FIELD_ACC);
    a.col = (float*)v(SF_FIELD_COL);
    a.label = (uint64_t*)v(SF_FIELD_LABEL);
    a.mass = (float*)v(SF_FIELD_MASS);
    a.radius = (float*)v(SF_FIELD_RADIUS);
    a.density = (float*)v(SF_FIELD_DENSITY);
    a.pressure = (float*)v(SF_FIELD_PRESSURE);
    a.voidage = (float*)v(SF_FIELD_VOIDAGE);
    a.saturation = (float*)v(SF_FIELD_SATURATION);
    a.maxsat = (float*)v(SF_FIELD_MAX_SATURATION);
    a.avgflow = (float*)v(SF_FIELD_AVG_FLOW_VEL);
    a.avgdrag = (float*)v(SF_FIELD_AVG_DRAG_FORCE);
    a.adhesion = (float*)v(SF_FIELD_ADHESION_FORCE);
    a.avgadh = (float*)v(SF_FIELD_AVG_ADHESION_FORCE);
    a.cda3 = (float*)v(SF_FIELD_CDA0ASAT);
    a.amc2 = (float2*)v(SF_FIELD_AMCAMCP);
    return a;
}

int drop_fresh_keys(sf_system* h) {
    if (!h->keys_fresh) return SF_OK;
    // the last advect hashed the next build's keys and took the newcomers'
    // ranks in in_count; the next build hashes and ranks again from zero
    SF_HIP(hipMemsetAsync(h->in_count.p, 0, h->in_count.n * sizeof(uint32_t), h->stream));
    h->keys_fresh = false;
    return SF_OK;
}

int fill_empty(sf_system* h) {
    if (!h->empty_pending) return SF_OK;
    Fields f = h->fields();
    f.nwater = h->wpre.p + h->n_built;
    launch_fill_empty(h->K, f, (int)h->n_built, h->stream);
    SF_HIP(hipGetLastError());
    h->empty_pending = false;
    return SF_OK;
}

int views_in(sf_system* h) {
    const uint32_t bits = h->view_bits & kPersistentViews;
    if (!bits || !h->uploaded) return SF_OK;
    launch_views_in(view_args(h, bits), (int)h->n, h->stream);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

int views_out(sf_system* h) { return views_out_bits(h, h->view_bits); }

int views_out_bits(sf_system* h, uint32_t bits) {
    if (!bits || !h->uploaded) return SF_OK;
    constexpr uint32_t kEmptySumViews = (1u << SF_FIELD_VOIDAGE) | (1u << SF_FIELD_SATURATION) |
                                        (1u << SF_FIELD_AVG_FLOW_VEL) | (1u << SF_FIELD_ADHESION_FORCE) |
                                        (1u << SF_FIELD_AVG_ADHESION_FORCE) | (1u << SF_FIELD_DENSITY) |
                                        (1u << SF_FIELD_PRESSURE);
    if (bits & kEmptySumViews) {
        int rc = fill_empty(h);
        if (rc) return rc;
    }
    if (bits & kViewable) launch_views_out(view_args(h, bits & kViewable), (int)h->n, h->stream);
    if (((bits >> SF_FIELD_SORTED_KEYS) & 1u) && h->built && h->sorted_keys)
        launch_widen_u32(h->sorted_keys, (uint64_t*)h->views[SF_FIELD_SORTED_KEYS].p, (int)h->n, h->stream);
    if ((bits >> SF_FIELD_CELL_START) & 1u)
        launch_widen_u32(h->cell.p, (uint64_t*)h->views[SF_FIELD_CELL_START].p, (int)h->C + 1, h->stream);
    if ((bits >> SF_FIELD_B_CELL_START) & 1u)
        launch_widen_u32(h->bcell.p, (uint64_t*)h->views[SF_FIELD_B_CELL_START].p, (int)h->local_cells() + 1,
                         h->stream);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

// host-side unpack of a device float4 array into k floats per element
int fetch4(sf_system* h, const float4* src, size_t n, std::vector<float4>& out) {
    out.resize(n);
    if (n == 0) return SF_OK;
    SF_HIP(hipMemcpyAsync(out.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}
template <class T>
int fetch(sf_system* h, const T* src, size_t n, std::vector<T>& out) {
    out.resize(n);
    if (n == 0) return SF_OK;
    SF_HIP(hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

// One neighbour pass with an explicit launch shape (autotuner).
void launch_pass(sf_system* h, int q, PassCfg c) {
    const Fields f = h->fields();
    const int n = (int)h->n;
    switch (q) {
    case 0:  // the density pass lists the interface particles
        launch_zero_counts(f, h->stream);
        launch_density(h->K, f, n, c, h->stream);
        break;
    case 1:  // in a step the density pass builds the lists
        launch_zero_counts(f, h->stream);
        launch_interface(h->K, f, n, h->stream);
        launch_avg_flow(h->K, f, n, c, h->stream);
        break;
    case 2: launch_voidage_adhesion(h->K, f, n, c, h->stream); break;
    case 3: launch_sand_momentum(h->K, f, n, c, h->stream); break;
    default: launch_water_momentum(h->K, f, n, c, h->stream); break;
    }
}

// sf_params of scene s: the handle's parameters with the scene's overrides.
sf_params scene_params(const sf_params& P, const sf_scene_params& q) {
    sf_params r = P;
    r.dem_tan_friction_angle = q.dem_tan_friction_angle;
    r.sf_c0 = q.sf_c0;
    r.sf_cmc = q.sf_cmc;
    r.sf_cmc_p = q.sf_cmc_p;
    r.sf_csat = q.sf_csat;
    r.sf_cd = q.sf_cd;
    return r;
}

// Per-scene constants from make_const of each scene's own sf_params, so a
// scene computes exactly what a single-scene handle with those parameters does.
int upload_scene_table(sf_system* h) {
    if (!h->nscenes) return SF_OK;
    std::vector<SceneConst> t(h->nscenes);
    for (size_t q = 0; q < h->nscenes; ++q) {
        const sf_params ps = scene_params(h->P, h->sp[q]);
        int rc = validate(&ps, &h->B);
        if (rc) return set_error(SF_E_INVALID, "scene " + std::to_string(q) + ": " + sf_last_error());
        const SFConst k = make_const(ps, h->B, h->O, h->blabel1);
        t[q].tan2 = k.tan2;
        t[q].tanphi = k.tanphi;
        t[q].cd = k.cd;
        t[q].pad = 0.f;
        t[q].gcap = k.gcap;
    }
    SF_HIP(hipMemcpyAsync(h->sct.p, t.data(), t.size() * sizeof(SceneConst), hipMemcpyHostToDevice, h->stream));
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

}  // namespace sf

extern "C" {

int sf_create(const sf_params* p, const sf_boundary_params* b, const sf_solver_options* opts, size_t max_num,
              int device, sf_handle* out) {
    if (!out) return set_error(SF_E_INVALID, "null output handle");
    *out = nullptr;
    int rc = validate(p, b);
    if (rc) return rc;
    if (max_num > ((size_t)1 << 27)) return set_error(SF_E_CAPACITY, "max_num above 2^27 particles");
    sf_system* h = new sf_system();
    h->P = *p;
    h->B = *b;
    h->O = opts ? *opts : sf_solver_options{0.f, 3.f, 100.f};
    if (device < 0) {
        hipError_t e = hipGetDevice(&h->device);
        if (e != hipSuccess) {
            delete h;
            return hip_error(e, "hipGetDevice");
        }
    } else {
        h->device = device;
    }
    auto bail = [&](int code) {
        sf_destroy(h);
        return code;
    };
    if ((rc = set_device(h))) return bail(rc);
    {
        hipError_t e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return bail(hip_error(e, "hipStreamCreate"));
        h->stream = h->own_stream;
        if ((e = hipEventCreate(&h->ev0)) != hipSuccess || (e = hipEventCreate(&h->ev1)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming)) != hipSuccess)
            return bail(hip_error(e, "hipEventCreate"));
    }
    h->cap = max_num;
    h->C = (uint32_t)((size_t)b->grid_size.x * b->grid_size.y * b->grid_size.z);
    h->plan = make_radix_plan(h->C);
    h->K = make_const(h->P, h->B, h->O, 0);
    const size_t c = std::max<size_t>(max_num, 1);
    for (int q = 0; q < 2; ++q) {
        if ((rc = h->pm[q].alloc(c)) || (rc = h->vr[q].alloc(c)) || (rc = h->cs[q].alloc(c)) ||
            (rc = h->cda[q].alloc(c)) || (rc = h->amc[q].alloc(c)))
            return bail(rc);
    }
    if ((rc = h->rp.alloc(c)) || (rc = h->es.alloc(c)) || (rc = h->wt.alloc(c)) || (rc = h->flow.alloc(c)) || (rc = h->adh.alloc(c)) ||
        (rc = h->xf.alloc(c)) || (rc = h->acc.alloc(c)))
        return bail(rc);
    if ((rc = h->cell.alloc((size_t)h->C + 1)) || (rc = h->bcell.alloc((size_t)h->local_cells() + 1)) ||
        (rc = h->bnear.alloc((size_t)h->local_cells())) || (rc = h->ct.alloc((size_t)h->C + 1)))
        return bail(rc);
    if ((rc = h->cfl_part.alloc(cfl_partials(0))) || (rc = h->cfl_out.alloc(1)) ||
        (rc = h->frame_dev.alloc(1)))
        return bail(rc);
    if ((rc = ensure_sort_capacity(h, c))) return bail(rc);
    if ((rc = h->bpv.alloc(1))) return bail(rc);
    // zero everything like CudaArray::Clear (cuda_array.cuh:51-54)
    for (int q = 0; q < 2; ++q) {
        if ((rc = zero(h, h->pm[q])) || (rc = zero(h, h->vr[q])) || (rc = zero(h, h->cs[q])) ||
            (rc = zero(h, h->cda[q])) || (rc = zero(h, h->amc[q])))
            return bail(rc);
    }
    if ((rc = zero(h, h->rp)) || (rc = zero(h, h->es)) || (rc = zero(h, h->flow)) || (rc = zero(h, h->adh)) ||
        (rc = zero(h, h->xf)) || (rc = zero(h, h->acc)) || (rc = zero(h, h->cell)) || (rc = zero(h, h->bcell)) ||
        (rc = zero(h, h->bnear)) ||
        (rc = zero(h, h->ct)) || (rc = zero(h, h->wt)) || (rc = zero(h, h->wpre)) || (rc = zero(h, h->plist)) ||
        (rc = zero(h, h->keys)) || (rc = zero(h, h->bpv)))
        return bail(rc);
    {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return bail(hip_error(e, "hipStreamSynchronize"));
    }
    *out = h;
    return SF_OK;
}

int sf_destroy(sf_handle h) {
    if (!h) return SF_OK;
    (void)hipSetDevice(h->device);
    if (h->own_stream) (void)hipStreamSynchronize(h->own_stream);
    h->drop_graph();
    diff_release(h->diff);
    h->diff = nullptr;
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->ev_order) (void)hipEventDestroy(h->ev_order);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete h;
    return SF_OK;
}

int sf_set_params(sf_handle h, const sf_params* p) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if ((rc = validate(p, &h->B))) return rc;
    if (p->sf_type != h->P.sf_type && h->uploaded)
        return set_error(SF_E_STATE, "sf_type cannot change after particles are uploaded");
    const sf_params old = h->P;
    h->P = *p;
    h->K = make_const(h->P, h->B, h->O, h->blabel1);
    if ((rc = upload_scene_table(h))) {
        h->P = old;
        h->K = make_const(h->P, h->B, h->O, h->blabel1);
        return rc;
    }
    h->drop_graph();
    return SF_OK;
}

int sf_set_stream(sf_handle h, void* stream) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    hipStream_t next = stream ? (hipStream_t)stream : h->own_stream;
    if (next != h->stream) {
        // the handle's work still queued on the previous stream runs before
        // anything the new stream gets from it
        SF_HIP(hipEventRecord(h->ev_order, h->stream));
        SF_HIP(hipStreamWaitEvent(next, h->ev_order, 0));
        h->stream = next;
    }
    return SF_OK;
}

int sf_synchronize(sf_handle h) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_upload_particles(sf_handle h, size_t n, const float* pos3, const float* vel3, const float* col3,
                        const uint64_t* label, const float* mass, const float* radius, const float* cda0asat3,
                        const float* amcamcp2) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (n > h->cap) return set_error(SF_E_CAPACITY, "particle count exceeds the handle's max_num");
    if (h->nscenes && n != h->cap)
        return set_error(SF_E_INVALID, "a batched handle takes exactly the sum of its scene counts");
    if (n > 0 && (!pos3 || !label || !mass)) return set_error(SF_E_INVALID, "pos3, label and mass are required");
    std::vector<float4> pm(n), vr(n), cs(n), cda(n);
    std::vector<float2> amc(n);
    int wstate = 0;
    float wmass = 0.f;
    for (size_t i = 0; i < n; ++i) {
        if (label[i] > 1) return set_error(SF_E_INVALID, "labels must be 0 (water) or 1 (sand)");
        if (!(mass[i] > 0.f) || !std::isfinite(mass[i])) return set_error(SF_E_INVALID, "masses must be finite and > 0");
        if (label[i] == 0 && wstate == 0) {
            wstate = 1;
            wmass = mass[i];
        } else if (label[i] == 0 && mass[i] != wmass) {
            wstate = 2;
        }
        const float m = label[i] == 1 ? -mass[i] : mass[i];
        pm[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], m);
        vr[i] = vel3 ? make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], radius ? radius[i] : 0.f)
                     : make_float4(0.f, 0.f, 0.f, radius ? radius[i] : 0.f);
        cs[i] = col3 ? make_float4(col3[3 * i], col3[3 * i + 1], col3[3 * i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
        cda[i] = cda0asat3 ? make_float4(cda0asat3[3 * i], cda0asat3[3 * i + 1], cda0asat3[3 * i + 2], 0.f)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
        amc[i] = amcamcp2 ? make_float2(amcamcp2[2 * i], amcamcp2[2 * i + 1]) : make_float2(0.f, 0.f);
    }
    h->cur = h->cur_cs = h->cur_aux = 0;
    if (n) {
        SF_HIP(hipMemcpyAsync(h->pm[0].p, pm.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        SF_HIP(hipMemcpyAsync(h->vr[0].p, vr.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        SF_HIP(hipMemcpyAsync(h->cs[0].p, cs.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        SF_HIP(hipMemcpyAsync(h->cda[0].p, cda.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        SF_HIP(hipMemcpyAsync(h->amc[0].p, amc.data(), n * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    }
    SF_HIP(hipStreamSynchronize(h->stream));
    h->n = n;
    h->uploaded = true;
    h->built = false;
    h->build_pending = false;
    h->water_mass_state = wstate;
    h->water_mass = wmass;
    h->drop_graph();
    diff_release(h->diff);  // a recorded window no longer matches the particles
    h->diff = nullptr;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_upload_boundary(sf_handle h, size_t nb, const float* pos3, const uint64_t* label) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (nb > 0 && !pos3) return set_error(SF_E_INVALID, "null boundary positions");
    if (nb > ((size_t)1 << 27)) return set_error(SF_E_CAPACITY, "more than 2^27 boundary particles");
    std::vector<float4> bp(nb);
    int l1 = 0;
    for (size_t i = 0; i < nb; ++i) {
        const uint64_t l = label ? label[i] : 0;
        if (l > 1) return set_error(SF_E_INVALID, "boundary labels must be 0 or 1");
        l1 += l == 1;
        bp[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], l == 1 ? -1.f : 1.f);
    }
    if ((rc = ensure_sort_capacity(h, std::max(h->cap, nb)))) return rc;
    if ((rc = h->bpv.alloc(std::max<size_t>(nb, 1)))) return rc;
    DevBuf<float4> tmp;
    if ((rc = tmp.alloc(std::max<size_t>(nb, 1)))) return rc;
    if (nb) SF_HIP(hipMemcpyAsync(tmp.p, bp.data(), nb * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    // CudaGNBoundarySearcher::SortData (cuda_neighbor_searcher.cu:125-139): stable sort by cell
    launch_hash_boundary(h->K, tmp.p, h->keys.p, (int)nb, h->stream);
    SortBuffers sb{h->kA.p, h->vA.p, h->kB.p, h->vB.p, h->hist.p};
    const uint32_t *sk = nullptr, *bperm = nullptr;
    if ((rc = radix_sort(h->keys.p, (int)nb, h->plan, sb, &sk, &bperm, h->stream))) return rc;
    launch_cell_start(sk, (int)nb, h->local_cells(), nullptr, h->bcell.p, nullptr, nullptr, PhasePos{},
                      IncStart{nullptr, nullptr}, h->stream);
    launch_boundary_gather(bperm, tmp.p, h->bpv.p, (int)nb, h->stream);
    launch_boundary_volume(h->K, h->bpv.p, h->bcell.p, (int)nb, h->stream);
    launch_boundary_near(h->K, h->bcell.p, h->bnear.p, h->stream);
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(h->stream));
    h->nb = nb;
    h->blabel1 = l1;
    h->K = make_const(h->P, h->B, h->O, h->blabel1);
    h->built = false;  // the particle searcher buffers were reused
    h->build_pending = false;
    h->drop_graph();
    if ((rc = upload_scene_table(h))) return rc;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_build_neighbors(sf_handle h) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if ((rc = views_in(h)) || (rc = build_neighbors(h)) || (rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_step(sf_handle h, float dt) { return sf_step_n(h, dt, 1); }

int sf_step_n(sf_handle h, float dt, int n_steps) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(dt) || !(dt > 0.f)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    if (n_steps < 0) return set_error(SF_E_INVALID, "n_steps must be >= 0");
    if ((rc = views_in(h))) return rc;
    for (int q = 0; q < n_steps; ++q)
        if ((rc = one_step(h, dt))) return rc;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_step_profiled(sf_handle h, float dt, int n_steps, float* pass_ms, float* total_ms) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(dt) || !(dt > 0.f)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    if (n_steps < 1 || n_steps > 100000) return set_error(SF_E_INVALID, "n_steps must be in [1, 100000]");
    const size_t per = SF_PASS_COUNT + 1;
    const size_t need = pass_ms ? per * (size_t)n_steps : 0;
    while (h->ev_pool.size() < need) {
        hipEvent_t e;
        SF_HIP(hipEventCreate(&e));
        h->ev_pool.push_back(e);
    }
    hipStream_t s = h->stream;
    if ((rc = views_in(h))) return rc;
    SF_HIP(hipEventRecord(h->ev0, s));
    for (int q = 0; q < n_steps; ++q)
        if ((rc = one_step(h, dt, pass_ms ? h->ev_pool.data() + per * q : nullptr))) return rc;
    SF_HIP(hipEventRecord(h->ev1, s));
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(s));
    if (total_ms) SF_HIP(hipEventElapsedTime(total_ms, h->ev0, h->ev1));
    if (pass_ms) {
        for (int k = 0; k < SF_PASS_COUNT; ++k) pass_ms[k] = 0.f;
        for (int q = 0; q < n_steps; ++q) {
            hipEvent_t* e = h->ev_pool.data() + per * q;
            for (int k = 0; k < SF_PASS_COUNT; ++k) {
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, e[k], e[k + 1]));
                pass_ms[k] += ms;
            }
        }
        for (int k = 0; k < SF_PASS_COUNT; ++k) pass_ms[k] /= (float)n_steps;
    }
    return SF_OK;
}

int sf_step_timed(sf_handle h, float dt, int n_steps, int pass, float* pass_ms, float* total_ms) {
    return sf_step_timed_every(h, dt, n_steps, pass, 1, pass_ms, total_ms);
}

int sf_step_timed_every(sf_handle h, float dt, int n_steps, int pass, int every, float* pass_ms, float* total_ms) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(dt) || !(dt > 0.f)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    if (n_steps < 1 || n_steps > 100000) return set_error(SF_E_INVALID, "n_steps must be in [1, 100000]");
    if (pass < 0 || pass >= SF_PASS_COUNT) return set_error(SF_E_INVALID, "pass must be one of the step's passes");
    if (every < 1 || every > n_steps) return set_error(SF_E_INVALID, "every must be in [1, n_steps]");
    const int sampled = (n_steps + every - 1) / every;  // steps 0, every, 2 every, ...
    const size_t need = 2 * (size_t)sampled;
    while (h->ev_pool.size() < need) {
        hipEvent_t e;
        SF_HIP(hipEventCreate(&e));
        h->ev_pool.push_back(e);
    }
    hipStream_t s = h->stream;
    std::vector<hipEvent_t> marks(SF_PASS_COUNT + 1, nullptr);
    if ((rc = views_in(h))) return rc;
    SF_HIP(hipEventRecord(h->ev0, s));
    for (int q = 0; q < n_steps; ++q) {
        const bool mark = q % every == 0;
        marks[pass] = mark ? h->ev_pool[2 * (q / every)] : nullptr;
        marks[pass + 1] = mark ? h->ev_pool[2 * (q / every) + 1] : nullptr;
        if ((rc = one_step(h, dt, marks.data()))) return rc;
    }
    SF_HIP(hipEventRecord(h->ev1, s));
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(s));
    if (total_ms) SF_HIP(hipEventElapsedTime(total_ms, h->ev0, h->ev1));
    if (pass_ms) {
        float sum = 0.f;
        for (int q = 0; q < sampled; ++q) {
            float ms = 0.f;
            SF_HIP(hipEventElapsedTime(&ms, h->ev_pool[2 * q], h->ev_pool[2 * q + 1]));
            sum += ms;
        }
        *pass_ms = sum / (float)sampled;
    }
    return SF_OK;
}

int sf_update_system(sf_handle h, float render_interval, float* ms_out) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(render_interval) || !(render_interval > 0.f))
        return set_error(SF_E_INVALID, "render interval must be finite and > 0");
    if (h->nscenes) return set_error(SF_E_STATE, "batched handles step with a fixed dt (sf_step, sf_step_graph)");
    if ((rc = views_in(h))) return rc;
    SF_HIP(hipEventRecord(h->ev0, h->stream));
    if ((rc = build_neighbors(h)) || (rc = update_solver(h, render_interval))) return rc;
    // CudaSFSystem::OnUpdateSolver emits after the solver step (cuda_sf_system.cu:86-115)
    if ((rc = emit_after_update(h, render_interval))) return rc;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    if (ms_out) {
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *ms_out = ms;
    }
    return SF_OK;
}

int sf_update_solver(sf_handle h, float render_interval, float* ms_out) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(render_interval) || !(render_interval > 0.f))
        return set_error(SF_E_INVALID, "render interval must be finite and > 0");
    if (h->nscenes) return set_error(SF_E_STATE, "batched handles step with a fixed dt (sf_step, sf_step_graph)");
    if (!h->build_pending)
        return set_error(SF_E_STATE, "sf_update_solver needs a neighbour build (sf_build_neighbors) since the last step");
    if ((rc = views_in(h))) return rc;
    SF_HIP(hipEventRecord(h->ev0, h->stream));
    if ((rc = update_solver(h, render_interval))) return rc;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    if (ms_out) {
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *ms_out = ms;
    }
    return SF_OK;
}

}  // extern "C"

namespace sf {

// Largest chunk of sub-steps enqueued between two host reads.
constexpr size_t kFrameChunk = 4096;

// k_cfl_final's sub-step count for remaining time r and CFL dt d.
static uint32_t cfl_count(float r, float d) {
    const float q = r / d;
    return q >= 4294967295.f ? 0xffffffffu : (q >= 1.f ? (uint32_t)q : 1u);
}

// Sub-steps the adaptive loop certainly still runs from remaining time r.
// The CFL dt never exceeds dt_max, so every count is at least
// cfl_count(r, dt_max), and r - r / k grows with r and with k: the loop run
// with dt_max ends no later than the real one.  This replays it in the same
// float arithmetic.
static size_t certain_sub_steps(float r, float dt_max, size_t cap) {
#pragma clang fp contract(off)
    size_t k = 0;
    while (r > 1e-6f && k < cap) {
        r = r - r / (float)cfl_count(r, dt_max);
        ++k;
    }
    return k;
}

// One WCSPH frame (main.cpp:235-259): the host enqueues as many sub-steps as
// the loop certainly runs (or, with a running emitter, as many as certainly
// pass without an emission), then reads the device loop state once.
static int frame_wcsph(sf_system* h, float ri, bool adaptive, size_t* steps) {
#pragma clang fp contract(off)
    int rc;
    // k_cfl_final: dt = min(params dt, scale * min(dtf, 0.4 h / c)) <= dt_max
    const float dts = 0.4f * h->K.h / h->O.speed_of_sound;
    const float lim = h->O.time_step_limit_scale * dts;
    const float dt_max = h->P.dt < lim ? h->P.dt : lim;
    // the fixed loop runs the count the previous call left (main.cpp:251-257)
    const size_t todo = h->nsub;
    bool clock = emitter_wcsph_clock(h);
    FrameDev fd{ri, h->emitter.elapsed, 0u, (adaptive ? kFrameAdaptive : 0u) | (clock ? kFrameElapsed : 0u)};
    SF_HIP(hipMemcpyAsync(h->frame_dev.p, &fd, sizeof(fd), hipMemcpyHostToDevice, h->stream));
    float r = ri;
    size_t done = 0;
    for (;;) {
        size_t len;
        if (adaptive) {
            if (!(r > 1e-6f)) break;
            len = certain_sub_steps(r, dt_max, kFrameChunk);
        } else {
            if (done >= todo) break;
            len = std::min(todo - done, kFrameChunk);
        }
        if (clock) {
            // a sub-step advances the clock by r / k <= 2 dt_max (k = floor(r / dt) >= 1),
            // or, with a fixed interval, by at most ri / cfl_count(ri, dt_max)
            const double step_max = adaptive ? 2.0 * (double)dt_max : (double)(ri / (float)cfl_count(ri, dt_max));
            len = std::min(len, emitter_quiet_steps(h, step_max * (1.0 + 1.0 / 4194304.0), len) + 1);
        }
        for (size_t k = 0; k < len; ++k) {
            if ((rc = build_neighbors(h))) return rc;
            if ((rc = wcsph_substep_tail(h, ri, h->frame_dev.p))) return rc;
        }
        CflOut o;
        SF_HIP(hipMemcpyAsync(&fd, h->frame_dev.p, sizeof(fd), hipMemcpyDeviceToHost, h->stream));
        SF_HIP(hipMemcpyAsync(&o, h->cfl_out.p, sizeof(o), hipMemcpyDeviceToHost, h->stream));
        SF_HIP(hipStreamSynchronize(h->stream));
        done += len;
        h->nsub = o.nsub;
        r = fd.remaining;
        if (clock) {
            h->emitter.elapsed = fd.elapsed;
            if ((rc = emitter_wcsph_check(h))) return rc;
            clock = emitter_wcsph_clock(h);  // off once the emitter met the capacity
        }
    }
    *steps = done;
    return SF_OK;
}

// One SPH-solver frame: the sub-step count is ri / params.dt on the host
// (CudaSphSFSolver::UpdateSolver, cuda_sph_sf_solver.cpp:19-30), so nothing
// is read back.
static int frame_sph(sf_system* h, float ri, bool adaptive, size_t* steps) {
#pragma clang fp contract(off)
    int rc;
    const size_t todo = h->nsub;
    float r = ri;
    size_t done = 0;
    for (;;) {
        float cur;
        if (adaptive) {
            if (!(r > 1e-6f)) break;
            cur = r;
        } else {
            if (done >= todo) break;
            cur = ri;
        }
        h->nsub = sub_step_count(cur / h->P.dt);
        if ((rc = build_neighbors(h)) || (rc = solve(h)) || (rc = advect(h, h->P.dt, nullptr))) return rc;
        if ((rc = emit_after_update(h, cur))) return rc;
        if (adaptive) r = r - r / static_cast<float>(h->nsub);
        ++done;
    }
    *steps = done;
    return SF_OK;
}

}  // namespace sf

extern "C" {

int sf_update_frame(sf_handle h, float render_interval, int adaptive, size_t* sub_steps, float* ms_out) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(render_interval) || !(render_interval > 0.f))
        return set_error(SF_E_INVALID, "render interval must be finite and > 0");
    if (h->nscenes) return set_error(SF_E_STATE, "batched handles step with a fixed dt (sf_step, sf_step_graph)");
    if ((rc = views_in(h))) return rc;
    SF_HIP(hipEventRecord(h->ev0, h->stream));
    size_t steps = 0;
    rc = h->P.solver_type == SF_WCSPH_SOLVER ? frame_wcsph(h, render_interval, adaptive != 0, &steps)
                                             : frame_sph(h, render_interval, adaptive != 0, &steps);
    if (rc) return rc;
    SF_HIP(hipEventRecord(h->ev1, h->stream));
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    if (sub_steps) *sub_steps = steps;
    if (ms_out) {
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *ms_out = ms;
    }
    return SF_OK;
}

int sf_num_sub_steps(sf_handle h, size_t* out) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!out) return set_error(SF_E_INVALID, "null output");
    *out = h->nsub;
    return SF_OK;
}

int sf_size(sf_handle h, size_t* n, size_t* nb, size_t* num_cells) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (n) *n = h->n;
    if (nb) *nb = h->nb;
    if (num_cells) *num_cells = h->C;
    return SF_OK;
}

int sf_download(sf_handle h, int field, void* dst, size_t dst_bytes) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    const size_t n = h->n, nb = h->nb;
    auto need = [&](size_t bytes) -> int {
        if (!dst && bytes) return set_error(SF_E_INVALID, "null destination");
        if (dst_bytes < bytes) return set_error(SF_E_INVALID, "destination buffer too small");
        return SF_OK;
    };
    std::vector<float4> v4;
    std::vector<float2> v2;
    std::vector<uint32_t> u;
    auto put3 = [&](const float4* src, size_t cnt, int mask_label /* -1 none, 0 water only, 1 sand only */) -> int {
        int r;
        if ((r = need(cnt * 12))) return r;
        if ((r = fetch4(h, src, cnt, v4))) return r;
        std::vector<float4> pmh;
        if (mask_label >= 0 && (r = fetch4(h, h->pm[h->cur].p, cnt, pmh))) return r;
        float* d = (float*)dst;
        for (size_t i = 0; i < cnt; ++i) {
            bool keep = mask_label < 0 || (mask_label == 1) == (pmh[i].w < 0.f);
            d[3 * i] = keep ? v4[i].x : 0.f;
            d[3 * i + 1] = keep ? v4[i].y : 0.f;
            d[3 * i + 2] = keep ? v4[i].z : 0.f;
        }
        return SF_OK;
    };
    auto put1w = [&](const float4* src, size_t cnt, bool absval) -> int {
        int r;
        if ((r = need(cnt * 4))) return r;
        if ((r = fetch4(h, src, cnt, v4))) return r;
        float* d = (float*)dst;
        for (size_t i = 0; i < cnt; ++i) d[i] = absval ? fabsf(v4[i].w) : v4[i].w;
        return SF_OK;
    };
    auto put2 = [&](const float2* src, size_t cnt, int comp) -> int {
        int r;
        if ((r = need(cnt * 4))) return r;
        if ((r = fetch(h, src, cnt, v2))) return r;
        float* d = (float*)dst;
        for (size_t i = 0; i < cnt; ++i) d[i] = comp ? v2[i].y : v2[i].x;
        return SF_OK;
    };
    auto putu = [&](const uint32_t* src, size_t cnt) -> int {
        int r;
        if ((r = need(cnt * 4))) return r;
        if (cnt == 0) return SF_OK;
        if (!src) return set_error(SF_E_STATE, "searcher has not run yet");
        SF_HIP(hipMemcpyAsync(dst, src, cnt * 4, hipMemcpyDeviceToHost, h->stream));
        SF_HIP(hipStreamSynchronize(h->stream));
        return SF_OK;
    };
    auto labels = [&](const float4* src, size_t cnt) -> int {
        int r;
        if ((r = need(cnt * 8))) return r;
        if ((r = fetch4(h, src, cnt, v4))) return r;
        uint64_t* d = (uint64_t*)dst;
        for (size_t i = 0; i < cnt; ++i) d[i] = v4[i].w < 0.f ? 1u : 0u;
        return SF_OK;
    };
    switch (field) {
    case SF_FIELD_VOIDAGE: case SF_FIELD_SATURATION: case SF_FIELD_AVG_FLOW_VEL: case SF_FIELD_ADHESION_FORCE:
    case SF_FIELD_AVG_ADHESION_FORCE: case SF_FIELD_DENSITY: case SF_FIELD_PRESSURE:
        if ((rc = fill_empty(h))) return rc;
        break;
    default: break;
    }
    switch (field) {
    case SF_FIELD_POS: return put3(h->pm[h->cur].p, n, -1);
    case SF_FIELD_VEL: return put3(h->vr[h->cur].p, n, -1);
    case SF_FIELD_ACC: return put3(h->acc.p, n, -1);
    case SF_FIELD_COL: return put3(h->cs[h->cur_cs].p, n, -1);
    case SF_FIELD_LABEL: return labels(h->pm[h->cur].p, n);
    case SF_FIELD_MASS: return put1w(h->pm[h->cur].p, n, true);
    case SF_FIELD_RADIUS: return put1w(h->vr[h->cur].p, n, false);
    case SF_FIELD_DENSITY: return put2(h->rp.p, n, 0);
    case SF_FIELD_PRESSURE: return put2(h->rp.p, n, 1);
    case SF_FIELD_VOIDAGE: return put2(h->es.p, n, 0);
    case SF_FIELD_SATURATION: return put2(h->es.p, n, 1);
    case SF_FIELD_MAX_SATURATION: return put1w(h->cs[h->cur_cs].p, n, false);
    case SF_FIELD_AVG_FLOW_VEL: return put3(h->flow.p, n, -1);
    case SF_FIELD_AVG_DRAG_FORCE: return put3(h->xf.p, n, 1);
    case SF_FIELD_ADHESION_FORCE: return put3(h->adh.p, n, 0);
    case SF_FIELD_AVG_ADHESION_FORCE: return put3(h->xf.p, n, 0);
    case SF_FIELD_CDA0ASAT: return put3(h->cda[h->cur_aux].p, n, -1);
    case SF_FIELD_AMCAMCP: {
        if ((rc = need(n * 8))) return rc;
        if ((rc = fetch(h, h->amc[h->cur_aux].p, n, v2))) return rc;
        if (n) std::memcpy(dst, v2.data(), n * 8);
        return SF_OK;
    }
    case SF_FIELD_PARTICLE2CELL: return putu(h->built ? h->keys.p : nullptr, n);
    case SF_FIELD_SORT_PERM: return putu(h->built ? h->perm : nullptr, n);
    case SF_FIELD_SORTED_KEYS: return putu(h->built ? h->sorted_keys : nullptr, n);
    case SF_FIELD_CELL_START: return putu(h->cell.p, (size_t)h->C + 1);
    case SF_FIELD_NBR_COUNT:
    case SF_FIELD_NBR_CANDIDATES: {
        if ((rc = need(n * 4))) return rc;
        if (!h->built) return set_error(SF_E_STATE, "searcher has not run yet");
        if (h->nbr_cnt.n < std::max<size_t>(n, 1)) {
            if ((rc = h->nbr_cnt.alloc(std::max<size_t>(n, 1))) || (rc = h->nbr_cand.alloc(std::max<size_t>(n, 1))))
                return rc;
        }
        // counts refer to the positions the last neighbour build sorted: the
        // pre-advect state is gone after a step, so count on the current
        // positions with the current cellStart only directly after
        // sf_build_neighbors (callers use it that way).
        launch_nbr_count(h->K, h->pm[h->cur].p, h->nscenes ? h->scene.p : nullptr, h->cell.p, h->nbr_cnt.p,
                         h->nbr_cand.p, (int)n, h->stream);
        SF_HIP(hipGetLastError());
        return putu(field == SF_FIELD_NBR_COUNT ? h->nbr_cnt.p : h->nbr_cand.p, n);
    }
    case SF_FIELD_B_POS: return put3(h->bpv.p, nb, -1);
    case SF_FIELD_B_LABEL: return labels(h->bpv.p, nb);
    case SF_FIELD_B_VOLUME: return put1w(h->bpv.p, nb, true);
    case SF_FIELD_B_CELL_START: return putu(h->bcell.p, (size_t)h->local_cells() + 1);
    default: return set_error(SF_E_NOTFOUND, "unknown field id " + std::to_string(field));
    }
}

int sf_set_field(sf_handle h, int field, const void* src, size_t src_bytes) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    const size_t n = h->n;
    if (n == 0) return SF_OK;
    if (!src) return set_error(SF_E_INVALID, "null source");
    const float* s = (const float*)src;
    auto patch4 = [&](float4* dev, int comps, int first) -> int {  // overwrite comps floats from component `first`
        if (src_bytes < n * comps * sizeof(float)) return set_error(SF_E_INVALID, "source buffer too small");
        std::vector<float4> v;
        int r;
        if ((r = fetch4(h, dev, n, v))) return r;
        for (size_t i = 0; i < n; ++i) {
            float* e = &v[i].x;
            for (int c = 0; c < comps; ++c) {
                if (!std::isfinite(s[comps * i + c])) return set_error(SF_E_INVALID, "non-finite value");
                e[first + c] = s[comps * i + c];
            }
        }
        SF_HIP(hipMemcpyAsync(dev, v.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
        if ((r = views_out(h))) return r;
        SF_HIP(hipStreamSynchronize(h->stream));
        return SF_OK;
    };
    switch (field) {
    case SF_FIELD_VEL: return patch4(h->vr[h->cur].p, 3, 0);
    case SF_FIELD_COL: return patch4(h->cs[h->cur_cs].p, 3, 0);
    case SF_FIELD_MAX_SATURATION: return patch4(h->cs[h->cur_cs].p, 1, 3);
    case SF_FIELD_CDA0ASAT: return patch4(h->cda[h->cur_aux].p, 3, 0);
    case SF_FIELD_AMCAMCP: {
        if (src_bytes < n * 8) return set_error(SF_E_INVALID, "source buffer too small");
        SF_HIP(hipMemcpyAsync(h->amc[h->cur_aux].p, src, n * 8, hipMemcpyHostToDevice, h->stream));
        if ((rc = views_out(h))) return rc;
        SF_HIP(hipStreamSynchronize(h->stream));
        return SF_OK;
    }
    default: return set_error(SF_E_INVALID, "field cannot be set (vel, col, max_saturation, cda0asat, amcamcp)");
    }
}

int sf_device_buffer(sf_handle h, int field, void** dptr, size_t* stride) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!dptr || !stride) return set_error(SF_E_INVALID, "null output");
    void* p = nullptr;
    size_t s = 16;
    switch (field) {
    case SF_FIELD_POS: case SF_FIELD_MASS: case SF_FIELD_LABEL:
        // the caller may now write positions: every later build hashes them
        h->pm_shared = true;
        if ((rc = set_device(h)) || (rc = drop_fresh_keys(h))) return rc;
        h->drop_graph();
        p = h->pm[h->cur].p;
        break;
    case SF_FIELD_VEL: case SF_FIELD_RADIUS: p = h->vr[h->cur].p; break;
    case SF_FIELD_ACC: p = h->acc.p; break;
    case SF_FIELD_COL: case SF_FIELD_MAX_SATURATION: p = h->cs[h->cur_cs].p; break;
    case SF_FIELD_DENSITY: case SF_FIELD_PRESSURE: p = h->rp.p; s = 8; break;
    case SF_FIELD_VOIDAGE: case SF_FIELD_SATURATION: p = h->es.p; s = 8; break;
    case SF_FIELD_AVG_FLOW_VEL: p = h->flow.p; break;
    case SF_FIELD_AVG_DRAG_FORCE: case SF_FIELD_AVG_ADHESION_FORCE: p = h->xf.p; break;
    case SF_FIELD_ADHESION_FORCE: p = h->adh.p; break;
    case SF_FIELD_CDA0ASAT: p = h->cda[h->cur_aux].p; break;
    case SF_FIELD_AMCAMCP: p = h->amc[h->cur_aux].p; s = 8; break;
    case SF_FIELD_PARTICLE2CELL: p = h->keys.p; s = 4; break;
    case SF_FIELD_SORT_PERM: p = (void*)h->perm; s = 4; break;
    case SF_FIELD_SORTED_KEYS: p = (void*)h->sorted_keys; s = 4; break;
    case SF_FIELD_CELL_START: p = h->cell.p; s = 4; break;
    case SF_FIELD_B_POS: case SF_FIELD_B_VOLUME: case SF_FIELD_B_LABEL: p = h->bpv.p; break;
    case SF_FIELD_B_CELL_START: p = h->bcell.p; s = 4; break;
    default: return set_error(SF_E_NOTFOUND, "field has no device buffer");
    }
    switch (field) {
    case SF_FIELD_VOIDAGE: case SF_FIELD_SATURATION: case SF_FIELD_AVG_FLOW_VEL: case SF_FIELD_DENSITY:
    case SF_FIELD_PRESSURE: case SF_FIELD_AVG_DRAG_FORCE: case SF_FIELD_AVG_ADHESION_FORCE:
    case SF_FIELD_ADHESION_FORCE: {
        // the caller reads these buffers directly: every later step writes
        // the empty sums' constants too
        h->full_empty = true;
        h->drop_graph();
        if ((rc = set_device(h)) || (rc = fill_empty(h))) return rc;
        SF_HIP(hipStreamSynchronize(h->stream));
        break;
    }
    default: break;
    }
    *dptr = p;
    *stride = s;
    return SF_OK;
}

int sf_particle_view(sf_handle h, int field, void** dptr) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!dptr) return set_error(SF_E_INVALID, "null output");
    if (field < 0 || field >= SF_FIELD_COUNT_ || !(((kViewable | kIndexViews) >> field) & 1u))
        return set_error(SF_E_NOTFOUND, "field has no reference-layout view");
    if (!((h->view_bits >> field) & 1u)) {
        size_t bytes;
        switch (field) {
        case SF_FIELD_LABEL: case SF_FIELD_SORTED_KEYS: bytes = 8 * h->cap; break;
        case SF_FIELD_AMCAMCP: bytes = 8 * h->cap; break;
        case SF_FIELD_MASS: case SF_FIELD_RADIUS: case SF_FIELD_DENSITY: case SF_FIELD_PRESSURE:
        case SF_FIELD_VOIDAGE: case SF_FIELD_SATURATION: case SF_FIELD_MAX_SATURATION: bytes = 4 * h->cap; break;
        case SF_FIELD_CELL_START: bytes = 8 * ((size_t)h->C + 1); break;
        case SF_FIELD_B_CELL_START: bytes = 8 * ((size_t)h->local_cells() + 1); break;
        default: bytes = 12 * h->cap; break;  // float3
        }
        if ((rc = h->views[field].alloc(bytes))) return rc;
        SF_HIP(hipMemsetAsync(h->views[field].p, 0, h->views[field].n, h->stream));
        h->view_bits |= 1u << field;
        if (field == SF_FIELD_POS || field == SF_FIELD_MASS || field == SF_FIELD_LABEL) {
            // the caller may now write positions: every later build hashes them
            h->pm_shared = true;
            if ((rc = drop_fresh_keys(h))) return rc;
        }
        h->drop_graph();
        // fill this view only: the others may hold writes for the next step
        if ((rc = views_out_bits(h, 1u << field))) return rc;
        SF_HIP(hipStreamSynchronize(h->stream));
    }
    *dptr = h->views[field].p;
    return SF_OK;
}

int sf_export_bgeo_device(const char* folder, const char* name, const float* pos3, const float* col3,
                          const float* radius, const uint64_t* label, size_t n) {
    if (!folder || !name) return set_error(SF_E_INVALID, "null folder or name");
    if (n > 0 && (!pos3 || !col3 || !radius || !label)) return set_error(SF_E_INVALID, "null device array");
    std::vector<float> pos(3 * n), col(3 * n), rad(n);
    std::vector<uint64_t> lab(n);
    if (n) {
        SF_HIP(hipMemcpy(pos.data(), pos3, 12 * n, hipMemcpyDeviceToHost));
        SF_HIP(hipMemcpy(col.data(), col3, 12 * n, hipMemcpyDeviceToHost));
        SF_HIP(hipMemcpy(rad.data(), radius, 4 * n, hipMemcpyDeviceToHost));
        SF_HIP(hipMemcpy(lab.data(), label, 8 * n, hipMemcpyDeviceToHost));
    }
    int rc;
    if ((rc = make_dirs(folder))) return rc;
    std::string path = std::string(folder) + "/" + name + ".bgeo";
    return sf_write_bgeo(path.c_str(), n, pos.data(), col.data(), rad.data(), lab.data());
}

int sf_export_bgeo(sf_handle h, const char* folder, const char* name) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!folder || !name) return set_error(SF_E_INVALID, "null folder or name");
    const size_t n = h->n;
    std::vector<float> pos(3 * n), col(3 * n), rad(n);
    std::vector<uint64_t> lab(n);
    if ((rc = sf_download(h, SF_FIELD_POS, pos.data(), pos.size() * 4))) return rc;
    if ((rc = sf_download(h, SF_FIELD_COL, col.data(), col.size() * 4))) return rc;
    if ((rc = sf_download(h, SF_FIELD_RADIUS, rad.data(), rad.size() * 4))) return rc;
    if ((rc = sf_download(h, SF_FIELD_LABEL, lab.data(), lab.size() * 8))) return rc;
    if ((rc = make_dirs(folder))) return rc;
    std::string path = std::string(folder) + "/" + name + ".bgeo";
    return sf_write_bgeo(path.c_str(), n, pos.data(), col.data(), rad.data(), lab.data());
}

int sf_batch_create(const sf_params* p, const sf_boundary_params* b, const sf_solver_options* opts, size_t n_scenes,
                    const size_t* counts, const sf_scene_params* scenes, int device, sf_handle* out) {
    if (!out) return set_error(SF_E_INVALID, "null output handle");
    *out = nullptr;
    if (n_scenes == 0 || n_scenes > 65535) return set_error(SF_E_INVALID, "n_scenes must be in [1, 65535]");
    if (!counts || !scenes) return set_error(SF_E_INVALID, "null counts or scene parameters");
    int rc = validate(p, b);
    if (rc) return rc;
    std::vector<size_t> off(n_scenes + 1, 0);
    for (size_t q = 0; q < n_scenes; ++q) off[q + 1] = off[q] + counts[q];
    const uint64_t cells = (uint64_t)b->grid_size.x * b->grid_size.y * b->grid_size.z;
    if (cells * n_scenes >= ((uint64_t)1 << 31)) return set_error(SF_E_CAPACITY, "scenes x cells above 2^31");
    sf_handle h = nullptr;
    if ((rc = sf_create(p, b, opts, off.back(), device, &h))) return rc;
    h->nscenes = n_scenes;
    h->soff = off;
    h->sp.assign(scenes, scenes + n_scenes);
    h->C = (uint32_t)(cells * n_scenes);
    h->plan = make_radix_plan(h->C);
    std::vector<uint16_t> sid(std::max<size_t>(off.back(), 1), 0);
    for (size_t q = 0; q < n_scenes; ++q)
        for (size_t i = off[q]; i < off[q + 1]; ++i) sid[i] = (uint16_t)q;
    auto bail = [&](int code) {
        sf_destroy(h);
        return code;
    };
    if ((rc = h->cell.alloc((size_t)h->C + 1)) || (rc = h->ct.alloc((size_t)h->C + 1)) ||
        (rc = h->scene.alloc(sid.size())) || (rc = h->sct.alloc(n_scenes)) || (rc = ensure_cell_capacity(h)))
        return bail(rc);
    {
        hipError_t e;
        if ((e = hipMemsetAsync(h->cell.p, 0, h->cell.n * sizeof(uint32_t), h->stream)) != hipSuccess ||
            (e = hipMemsetAsync(h->ct.p, 0, h->ct.n * sizeof(uint2), h->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(h->scene.p, sid.data(), sid.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                                h->stream)) != hipSuccess)
            return bail(hip_error(e, "batch buffers"));
    }
    if ((rc = upload_scene_table(h))) return bail(rc);
    *out = h;
    return SF_OK;
}

int sf_batch_set_scene_params(sf_handle h, const sf_scene_params* scenes) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->nscenes) return set_error(SF_E_STATE, "not a batched handle");
    if (!scenes) return set_error(SF_E_INVALID, "null scene parameters");
    std::vector<sf_scene_params> old = h->sp;
    h->sp.assign(scenes, scenes + h->nscenes);
    if ((rc = upload_scene_table(h))) {
        h->sp = old;
        (void)upload_scene_table(h);
        return rc;
    }
    return SF_OK;
}

int sf_batch_scene_offsets(sf_handle h, size_t* n_scenes, size_t* offsets, size_t capacity) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (n_scenes) *n_scenes = h->nscenes;
    if (!offsets) return SF_OK;
    if (!h->nscenes) return set_error(SF_E_STATE, "not a batched handle");
    if (capacity < h->nscenes + 1) return set_error(SF_E_INVALID, "offsets needs n_scenes + 1 entries");
    std::copy(h->soff.begin(), h->soff.end(), offsets);
    return SF_OK;
}

int sf_step_graph(sf_handle h, float dt, int n_steps) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (!std::isfinite(dt) || !(dt > 0.f)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    if (n_steps < 0) return set_error(SF_E_INVALID, "n_steps must be >= 0");
    if ((rc = views_in(h))) return rc;
    while (n_steps >= 2 && !h->keys_cycle()) {  // at most two plain steps first: see keys_cycle
        if ((rc = one_step(h, dt))) return rc;
        --n_steps;
    }
    if (n_steps >= 2 && (h->graph_dirty || !h->gexec || h->graph_dt != dt || h->graph_parity != h->parity() ||
                         h->graph_keys != h->sorted_keys || h->graph_fresh != h->keys_fresh)) {
        h->drop_graph();
        const int parity = h->parity();
        const uint32_t* keys = h->sorted_keys;
        const bool fresh = h->keys_fresh;
        SF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        rc = one_step(h, dt);
        if (!rc) rc = one_step(h, dt);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : hip_error(e, "hipStreamEndCapture");
        }
        h->graph = g;
        SF_HIP(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
        h->graph_dt = dt;
        h->graph_parity = parity;
        h->graph_keys = keys;
        h->graph_fresh = fresh;
        h->graph_dirty = false;
        if (h->parity() != parity || h->sorted_keys != keys || h->keys_fresh != fresh)
            return set_error(SF_E_STATE, "two steps did not restore the buffer parity");
    }
    for (int q = 0; q + 1 < n_steps; q += 2) SF_HIP(hipGraphLaunch(h->gexec, h->stream));
    if (n_steps >= 2) {  // the replayed solves, as solve_with records them
        h->empty_pending = !h->full_empty;
        h->n_built = h->n;
    }
    if (n_steps & 1)
        if ((rc = one_step(h, dt))) return rc;
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_set_launch_config(sf_handle h, int pass, int nb, int cap) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (pass < SF_PASS_DENSITY || pass > SF_PASS_WATER_MOMENTUM)
        return set_error(SF_E_INVALID, "pass must be one of the neighbour passes");
    if (std::find(std::begin(kNbChoices), std::end(kNbChoices), nb) == std::end(kNbChoices) ||
        std::find(std::begin(kCapChoices), std::end(kCapChoices), cap) == std::end(kCapChoices))
        return set_error(SF_E_INVALID, "nb must be 64, 128 or 256 and cap 8 or 16");
    PassCfg& c = h->cfg[pass - SF_PASS_DENSITY];
    c.nb = nb;
    c.cap = cap;
    h->drop_graph();
    return SF_OK;
}

int sf_set_interface_lanes(sf_handle h, int lanes) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (lanes != 1 && lanes != 2 && lanes != 4) return set_error(SF_E_INVALID, "lanes must be 1, 2 or 4");
    h->cfg[SF_PASS_AVG_FLOW - SF_PASS_DENSITY].lanes = lanes;
    h->cfg[SF_PASS_VOIDAGE_ADHESION - SF_PASS_DENSITY].lanes = lanes;
    h->drop_graph();
    return SF_OK;
}

int sf_set_prefix_scan_threshold(int tiles) { return set_direct_tiles(tiles); }

int sf_set_sort_strategy(sf_handle h, int strategy) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (strategy != SF_SORT_RADIX && strategy != SF_SORT_INCREMENTAL)
        return set_error(SF_E_INVALID, "strategy must be SF_SORT_RADIX or SF_SORT_INCREMENTAL");
    h->sort_inc = strategy == SF_SORT_INCREMENTAL;
    h->drop_graph();
    return SF_OK;
}

int sf_get_launch_config(sf_handle h, int pass, int* nb, int* cap) {
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (pass < SF_PASS_DENSITY || pass > SF_PASS_WATER_MOMENTUM)
        return set_error(SF_E_INVALID, "pass must be one of the neighbour passes");
    if (nb) *nb = h->cfg[pass - SF_PASS_DENSITY].nb;
    if (cap) *cap = h->cfg[pass - SF_PASS_DENSITY].cap;
    return SF_OK;
}

int sf_autotune(sf_handle h, int reps, float* best_ms, float* default_ms) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded) return set_error(SF_E_STATE, "no particles uploaded");
    if (reps < 1 || reps > 1000) return set_error(SF_E_INVALID, "reps must be in [1, 1000]");
    // A neighbour build reorders the particles exactly as the next step's
    // first pass would; every neighbour pass is then idempotent (it rewrites
    // the same outputs from unchanged inputs), so each is timed in place, in
    // dependency order, without disturbing the trajectory.
    if ((rc = build_neighbors(h))) return rc;
    if ((rc = solve(h))) return rc;
    for (int q = 0; q < 5; ++q) {
        const bool lists = q >= 3;
        float best = 0.f, dflt = 0.f;
        PassCfg pick{256, 16, h->cfg[q].lanes};
        for (int nb : kNbChoices)
            for (int cap : kCapChoices) {
                if (!lists && cap != 16) continue;
                const PassCfg c{nb, cap, h->cfg[q].lanes};
                launch_pass(h, q, c);  // warm-up
                SF_HIP(hipEventRecord(h->ev0, h->stream));
                for (int r = 0; r < reps; ++r) launch_pass(h, q, c);
                SF_HIP(hipEventRecord(h->ev1, h->stream));
                SF_HIP(hipEventSynchronize(h->ev1));
                SF_HIP(hipGetLastError());
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
                ms /= (float)reps;
                if (nb == 256 && cap == 16) dflt = ms;
                if (best == 0.f || ms < best) {
                    best = ms;
                    pick = c;
                }
            }
        h->cfg[q] = pick;
        if (best_ms) best_ms[q] = best;
        if (default_ms) default_ms[q] = dflt;
    }
    h->drop_graph();
    return SF_OK;
}

}  // extern "C"
