// sf_system.h -- the handle behind sf_handle and the step drivers shared by
// the C-ABI translation units (sf_api.cpp, sf_diff.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sf_internal.h"

namespace sf {
struct DiffState;  // differentiable-window tape (sf_diff.cpp)
void diff_release(DiffState* d);
}  // namespace sf

using namespace sf;

// Inflow emitter (host side, sf_emitter.cpp).
struct EmitterState {
    bool enable = false, run = false;
    sf_float3 pos{}, vel{}, col{}, axis1{}, axis2{};
    std::vector<float> samples;  // (s1, s2) pairs in the emitter plane
    size_t counter = 0, emitted = 0;
    float elapsed = 0.f;
};

struct sf_system {
    sf_params P{};
    sf_boundary_params B{};
    sf_solver_options O{};
    SFConst K{};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    size_t cap = 0, n = 0, nb = 0;
    uint32_t C = 0;
    RadixPlan plan;
    int blabel1 = 0;
    // persistent state, ping-pong
    DevBuf<float4> pm[2], vr[2], cs[2], cda[2];
    DevBuf<float2> amc[2];
    int cur = 0;      // pm / vr
    int cur_cs = 0;   // cs
    int cur_aux = 0;  // cda / amc
    // searcher
    DevBuf<uint32_t> keys, kA, kB, vA, vB, hist, cell, bcell, wpre, wpart, plist;
    DevBuf<float4> pph;  // positions in phase order (Fields::pph)
    DevBuf<uint32_t> ilist, icount;  // interface particles (Fields::ilist)
    DevBuf<uint8_t> iface;            // sand with water in its columns (Fields::iface)
    DevBuf<uint8_t> wl;               // listed water (Fields::wl)
    // the density pass writes every empty-sum field (sf_device_buffer handed
    // one of those buffers out); otherwise empty_pending marks a solve whose
    // unlisted particles still need them before a read (n_built particles)
    bool full_empty = false;
    bool empty_pending = false;
    size_t n_built = 0;
    DevBuf<uint8_t> bnear;            // cells with boundary particles near (Fields::bnear)
    DevBuf<float4> vq, wq, xq, wbq;   // listed particles' fields by phase index (Fields::wq)
    DevBuf<float4> rec0, rec1;        // water momentum neighbour records (Fields::rec0)
    DevBuf<uint8_t> wflag;
    DevBuf<uint2> ct;
    // incremental sort (sf_set_sort_strategy): sorted-key ping-pong, permutation, work arrays
    DevBuf<uint32_t> skeys[2], sperm, slot, in_list, stay_pre, in_count, in_start, part2;
    bool sort_inc = true;
    size_t sort_cap = 0;
    const uint32_t* sorted_keys = nullptr;
    // keys holds the hash of the current positions and the incremental
    // sort's change flags are taken (the last advect wrote them); cleared by
    // anything else that changes positions
    bool keys_fresh = false;
    bool pm_shared = false;  // the position buffer was handed out (sf_device_buffer)
    // water masses of the particles: 0 no water, 1 all equal to water_mass,
    // 2 not all equal (set by the upload and the emitter)
    int water_mass_state = 0;
    float water_mass = 0.f;
    // the neighbour records are used while the water masses are equal and
    // known on the host (a caller holding the position / mass buffer may
    // change them)
    bool records_on() const { return water_mass_state < 2 && !pm_shared; }
    const uint32_t* perm = nullptr;
    // transient fields
    DevBuf<float2> rp, es, wt;
    DevBuf<float4> flow, adh, xf, acc;
    // boundary
    DevBuf<float4> bpv;
    // CFL
    DevBuf<float> cfl_part;
    DevBuf<CflOut> cfl_out;
    DevBuf<FrameDev> frame_dev;  // sf_update_frame's loop state
    size_t nsub = 1;
    // batched scenes (nscenes == 0: one scene)
    size_t nscenes = 0;
    std::vector<size_t> soff;
    std::vector<sf_scene_params> sp;
    DevBuf<uint16_t> scene;
    DevBuf<SceneConst> sct;
    // launch shapes of the neighbour passes (SF_PASS_DENSITY ..)
    PassCfg cfg[5];
    // graph of two fixed-dt steps
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float graph_dt = 0.f;
    int graph_parity = -1;
    const uint32_t* graph_keys = nullptr;  // sorted_keys when the graph was captured
    bool graph_fresh = false;              // keys_fresh when the graph was captured
    bool graph_dirty = true;
    // differentiable window (sf_diff_*), null until first used
    DiffState* diff = nullptr;
    EmitterState emitter;
    // reference-layout views handed out (sf_particle_view), by field id;
    // view_bits has bit f set for each
    DevBuf<uint8_t> views[SF_FIELD_COUNT_];
    uint32_t view_bits = 0;
    // the last neighbour build has not been followed by an advect yet
    // (sf_update_solver runs on that build)
    bool build_pending = false;
    // diagnostics
    DevBuf<uint32_t> nbr_cnt, nbr_cand;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_order = nullptr;  // orders a stream switch (sf_set_stream)
    std::vector<hipEvent_t> ev_pool;  // sf_step_profiled
    bool uploaded = false;
    bool built = false;

    Fields fields() {
        Fields f;
        f.pm = pm[cur].p;
        f.vr = vr[cur].p;
        f.cs = cs[cur_cs].p;
        f.cda = cda[cur_aux].p;
        f.ct = ct.p;
        f.bpv = bpv.p;
        f.bcell = bcell.p;
        f.bnear = bnear.p;
        f.rp = rp.p;
        f.es = es.p;
        f.wt = wt.p;
        f.flow = flow.p;
        f.adh = adh.p;
        f.xf = xf.p;
        f.acc = acc.p;
        f.den4 = nullptr;
        f.scene = nscenes ? scene.p : nullptr;
        f.sct = nscenes ? sct.p : nullptr;
        f.plist = plist.p;
        f.nwater = wpre.p + n;
        f.pph = pph.p;
        f.ilist = ilist.p;
        f.icount = icount.p;
        f.iface = iface.p;
        f.wl = wl.p;
        f.full_empty = full_empty ? 1 : 0;
        f.vq = vq.p;
        f.wq = wq.p;
        f.xq = xq.p;
        f.wbq = wbq.p;
        f.rec0 = rec0.p;
        f.rec1 = rec1.p;
        f.mw = water_mass;
        f.rec = records_on() ? 1 : 0;
        return f;
    }
    uint32_t local_cells() const { return K.C; }
    int parity() const { return cur | (cur_cs << 1) | (cur_aux << 2); }
    // Two steps leave every buffer and flag where they found them, so that a
    // captured pair of steps can be replayed: with the incremental sort this
    // holds once the previous sorted keys sit in its two buffers (after a
    // first build and one incremental sort; before, the first sort is a radix
    // sort whose keys sit elsewhere) and the last advect wrote the next keys;
    // the radix sort always ends in the same buffer.
    bool keys_cycle() const {
        return !sort_inc || (built && keys_fresh && (sorted_keys == skeys[0].p || sorted_keys == skeys[1].p));
    }
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_dirty = true;
    }
};

namespace sf {
int ensure_sort_capacity(sf_system* h, size_t m);
int ensure_cell_capacity(sf_system* h);
int build_neighbors(sf_system* h, hipEvent_t* ev = nullptr);
int solve(sf_system* h, hipEvent_t* ev = nullptr);
int solve_with(sf_system* h, const Fields& f, hipEvent_t* ev = nullptr);
int advect(sf_system* h, float dt, const CflOut* dt_dev, hipEvent_t* ev = nullptr);
int wcsph_substep_tail(sf_system* h, float render_interval, FrameDev* frame);
int update_solver(sf_system* h, float render_interval);
int one_step(sf_system* h, float dt, hipEvent_t* ev = nullptr);
int check_handle(sf_system* h);
// Inflow emitter after a sf_update_system step (sf_emitter.cpp);
// render_interval as passed to that call.
int emit_after_update(sf_system* h, float render_interval);
// the WCSPH emission clock runs (emitter on, running, built, WCSPH solver)
bool emitter_wcsph_clock(const sf_system* h);
// the WCSPH emission test on the clock as it stands (emits when due)
int emitter_wcsph_check(sf_system* h);
// sub-steps, each advancing the clock by at most step_max, after which the
// WCSPH emission test is certainly still false (at most cap)
size_t emitter_quiet_steps(const sf_system* h, double step_max, size_t cap);
int set_device(sf_system* h);
// Views of the reference layout: views_in takes what callers wrote into the
// persistent fields' views into the state (start of a stepping call),
// views_out refreshes every view from the state (end of a call that changed it).
int views_in(sf_system* h);
int views_out(sf_system* h);
// Before anything outside a step may move particles: forgets the next
// build's keys and newcomer ranks that the last advect took (the build then
// hashes and ranks them again).
int drop_fresh_keys(sf_system* h);
int views_out_bits(sf_system* h, uint32_t bits);  // the views in `bits` only
// Writes the empty sums' fields of the unlisted particles of the last solve
// when they are still pending (before the host reads those fields).
int fill_empty(sf_system* h);
int fetch4(sf_system* h, const float4* src, size_t n, std::vector<float4>& out);
}  // namespace sf
