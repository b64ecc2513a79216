// sf_diff.cpp -- differentiable window of fixed-dt steps (no counterpart in
// the reference): a forward run that keeps a checkpoint of the state every K
// steps, and a backward run that, segment by segment from the last, recomputes
// the K steps from their checkpoint into a tape (sorted state, permutation,
// cell table, every pass's outputs) and runs the reverse passes of
// sf_adjoint.hip over it.  Memory: one state per checkpoint plus K taped
// steps (about 180 bytes per particle each).
//
// Particle identity: every field crossing this interface is in "id order",
// the slot order of the state the window started from; ids follow each
// step's sort permutation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sf_adjoint.h"
#include "sf_system.h"

namespace sf {

struct StateCopy {
    DevBuf<float4> pm, vr, cs;
    DevBuf<uint32_t> id;
};
struct StepTapeBufs {
    DevBuf<float4> pm, vr, flow, adh, xf, acc, den4;
    DevBuf<float2> rp, es, wt;
    DevBuf<uint2> ct;
    DevBuf<uint32_t> perm;
};
struct DiffState {
    size_t n = 0;
    int steps = 0, every = 0;
    float dt = 0.f;
    bool have_forward = false;
    bool last_taped = false;  // the forward already taped the last segment
    DevBuf<uint32_t> id, id_tmp;
    std::vector<StateCopy> ck;  // state before steps 0, K, 2K, ...
    StateCopy fin;              // state after the window
    DevBuf<float2> fin_rp, fin_es;  // density/pressure, voidage/saturation of the window's last step
    std::vector<StepTapeBufs> tape;
    DevBuf<float4> gx[2], gv[2], gacc, gflow, gadh, gxf, o4, o4b, gth, tmp4;
    DevBuf<float2> grp, ges, gwt, o2, o2b;
    DevBuf<float> sat_seed, tmpf;
};
void diff_release(DiffState* d) { delete d; }

namespace {

int copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) SF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    return SF_OK;
}

int save_state(sf_system* h, StateCopy& c) {
    const size_t n = h->n;
    int rc;
    if (c.pm.n < n && ((rc = c.pm.alloc(n)) || (rc = c.vr.alloc(n)) || (rc = c.cs.alloc(n)) || (rc = c.id.alloc(n))))
        return rc;
    if ((rc = copy(c.pm.p, h->pm[h->cur].p, n * sizeof(float4), h->stream)) ||
        (rc = copy(c.vr.p, h->vr[h->cur].p, n * sizeof(float4), h->stream)) ||
        (rc = copy(c.cs.p, h->cs[h->cur_cs].p, n * sizeof(float4), h->stream)) ||
        (rc = copy(c.id.p, h->diff->id.p, n * sizeof(uint32_t), h->stream)))
        return rc;
    return SF_OK;
}
int load_state(sf_system* h, const StateCopy& c) {
    const size_t n = h->n;
    int rc;
    if ((rc = copy(h->pm[h->cur].p, c.pm.p, n * sizeof(float4), h->stream)) ||
        (rc = copy(h->vr[h->cur].p, c.vr.p, n * sizeof(float4), h->stream)) ||
        (rc = copy(h->cs[h->cur_cs].p, c.cs.p, n * sizeof(float4), h->stream)) ||
        (rc = copy(h->diff->id.p, c.id.p, n * sizeof(uint32_t), h->stream)))
        return rc;
    h->built = false;
    return SF_OK;
}

// one step; ids follow the sort; with `tape`, every value the reverse needs is kept
int diff_step(sf_system* h, float dt, StepTapeBufs* tape) {
    DiffState* d = h->diff;
    const int n = (int)h->n;
    hipStream_t s = h->stream;
    int rc;
    if ((rc = build_neighbors(h))) return rc;
    launch_gather_u32(h->perm, d->id.p, d->id_tmp.p, n, s);
    std::swap(d->id.p, d->id_tmp.p);
    std::swap(d->id.n, d->id_tmp.n);
    if (!tape) {
        if ((rc = solve(h))) return rc;
        return advect(h, dt, nullptr);
    }
    StepTapeBufs& T = *tape;
    CopyJobs before{{{h->pm[h->cur].p, T.pm.p, n * sizeof(float4)},
                     {h->vr[h->cur].p, T.vr.p, n * sizeof(float4)},
                     {h->perm, T.perm.p, n * sizeof(uint32_t)},
                     {h->ct.p, T.ct.p, ((size_t)h->C + 1) * sizeof(uint2)}},
                    4};
    launch_copy_jobs(before, s);
    Fields f = h->fields();
    f.rp = T.rp.p;
    f.es = T.es.p;
    f.wt = T.wt.p;
    f.flow = T.flow.p;
    f.adh = T.adh.p;
    f.xf = T.xf.p;
    f.den4 = T.den4.p;
    if ((rc = solve_with(h, f))) return rc;
    // the handle's own transient fields keep the last step's values for downloads
    CopyJobs after{{{h->acc.p, T.acc.p, n * sizeof(float4)},
                    {T.rp.p, h->rp.p, n * sizeof(float2)},
                    {T.es.p, h->es.p, n * sizeof(float2)},
                    {nullptr, nullptr, 0}},
                   3};
    launch_copy_jobs(after, s);
    SF_HIP(hipGetLastError());
    return advect(h, dt, nullptr);
}

int check_diff_handle(sf_system* h) {
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!h->uploaded || h->n == 0) return set_error(SF_E_STATE, "no particles uploaded");
    if (h->nscenes) return set_error(SF_E_STATE, "the differentiable window runs on single-scene handles");
    if (h->K.multi) return set_error(SF_E_STATE, "the differentiable window needs SF mode (one global drag coefficient)");
    if (h->blabel1) return set_error(SF_E_STATE, "the differentiable window does not support label-1 collider particles");
    return SF_OK;
}

int alloc_diff(sf_system* h) {
    if (!h->diff) h->diff = new DiffState();
    DiffState* d = h->diff;
    const size_t n = h->n;
    if (d->n == n) return SF_OK;
    int rc;
    if ((rc = d->id.alloc(n)) || (rc = d->id_tmp.alloc(n))) return rc;
    for (int q = 0; q < 2; ++q)
        if ((rc = d->gx[q].alloc(n)) || (rc = d->gv[q].alloc(n))) return rc;
    if ((rc = d->gacc.alloc(n)) || (rc = d->gflow.alloc(n)) || (rc = d->gadh.alloc(n)) || (rc = d->gxf.alloc(n)) ||
        (rc = d->o4.alloc(n)) || (rc = d->o4b.alloc(n)) || (rc = d->o2b.alloc(n)) || (rc = d->gth.alloc(n)) || (rc = d->tmp4.alloc(n)) || (rc = d->grp.alloc(n)) ||
        (rc = d->ges.alloc(n)) || (rc = d->gwt.alloc(n)) || (rc = d->o2.alloc(n)) || (rc = d->sat_seed.alloc(n)) ||
        (rc = d->tmpf.alloc(3 * n)) || (rc = d->fin_rp.alloc(n)) || (rc = d->fin_es.alloc(n)))
        return rc;
    d->ck.clear();
    d->tape.clear();
    d->fin = StateCopy();
    d->n = n;
    d->have_forward = false;
    return SF_OK;
}

int alloc_tape(sf_system* h, int slots) {
    DiffState* d = h->diff;
    const size_t n = h->n;
    int rc;
    while ((int)d->tape.size() < slots) {
        d->tape.emplace_back();
        StepTapeBufs& T = d->tape.back();
        if ((rc = T.pm.alloc(n)) || (rc = T.vr.alloc(n)) || (rc = T.flow.alloc(n)) || (rc = T.adh.alloc(n)) ||
            (rc = T.xf.alloc(n)) || (rc = T.acc.alloc(n)) || (rc = T.den4.alloc(n)) || (rc = T.rp.alloc(n)) ||
            (rc = T.es.alloc(n)) || (rc = T.wt.alloc(n)) || (rc = T.ct.alloc((size_t)h->C + 1)) ||
            (rc = T.perm.alloc(n)))
            return rc;
    }
    return SF_OK;
}

// host or device source -> device pointer (staged in `stage` for host data)
int to_device(sf_system* h, const float* src, size_t count, int on_device, float* stage, const float** out) {
    if (!src) {
        *out = nullptr;
        return SF_OK;
    }
    if (on_device) {
        *out = src;
        return SF_OK;
    }
    SF_HIP(hipMemcpyAsync(stage, src, count * sizeof(float), hipMemcpyHostToDevice, h->stream));
    *out = stage;
    return SF_OK;
}

}  // namespace
}  // namespace sf

extern "C" {

int sf_diff_forward(sf_handle h, float dt, int n_steps, int checkpoint_every) {
    int rc;
    if ((rc = check_diff_handle(h))) return rc;
    if (!std::isfinite(dt) || !(dt > 0.f)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    if (n_steps < 1 || checkpoint_every < 1) return set_error(SF_E_INVALID, "n_steps and checkpoint_every must be >= 1");
    if ((rc = alloc_diff(h))) return rc;
    DiffState* d = h->diff;
    const int nck = (n_steps + checkpoint_every - 1) / checkpoint_every;
    if ((int)d->ck.size() < nck) d->ck.resize(nck);
    d->steps = n_steps;
    d->every = checkpoint_every;
    d->dt = dt;
    // the last segment is taped right away, so the backward starts without recomputing it
    if ((rc = alloc_tape(h, checkpoint_every))) return rc;
    const int last0 = (nck - 1) * checkpoint_every;
    d->last_taped = false;
    if ((rc = views_in(h))) return rc;
    launch_iota(d->id.p, (int)h->n, h->stream);
    for (int q = 0; q < n_steps; ++q) {
        if (q % checkpoint_every == 0 && (rc = save_state(h, d->ck[q / checkpoint_every]))) return rc;
        if ((rc = diff_step(h, dt, q >= last0 ? &d->tape[q - last0] : nullptr))) return rc;
    }
    d->last_taped = true;
    if ((rc = save_state(h, d->fin)) || (rc = copy(d->fin_rp.p, h->rp.p, h->n * sizeof(float2), h->stream)) ||
        (rc = copy(d->fin_es.p, h->es.p, h->n * sizeof(float2), h->stream)) || (rc = views_out(h)))
        return rc;
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(h->stream));
    d->have_forward = true;
    return SF_OK;
}

int sf_diff_reset(sf_handle h, const float* vel3, int on_device) {
    int rc;
    if ((rc = check_diff_handle(h))) return rc;
    DiffState* d = h->diff;
    if (!d || !d->have_forward) return set_error(SF_E_STATE, "no differentiable window has run yet");
    if ((rc = load_state(h, d->ck[0]))) return rc;
    if (vel3) {
        const float* dv = nullptr;
        if ((rc = to_device(h, vel3, 3 * h->n, on_device, d->tmpf.p, &dv))) return rc;
        launch_set_vel_by_id(d->ck[0].id.p, dv, h->vr[h->cur].p, (int)h->n, h->stream);
    }
    if ((rc = views_out(h))) return rc;
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_diff_get(sf_handle h, int field, float* dst, int on_device) {
    int rc;
    if ((rc = check_diff_handle(h))) return rc;
    DiffState* d = h->diff;
    if (!d || !d->have_forward) return set_error(SF_E_STATE, "no differentiable window has run yet");
    if (!dst) return set_error(SF_E_INVALID, "null destination");
    const int n = (int)h->n;
    float* out = on_device ? dst : d->tmpf.p;
    size_t count = 0;
    switch (field) {
    case SF_FIELD_POS:
        launch_scatter_by_id(d->id.p, (const float*)h->pm[h->cur].p, 4, 0, 3, out, n, h->stream);
        count = 3 * (size_t)n;
        break;
    case SF_FIELD_VEL:
        launch_scatter_by_id(d->id.p, (const float*)h->vr[h->cur].p, 4, 0, 3, out, n, h->stream);
        count = 3 * (size_t)n;
        break;
    case SF_FIELD_SATURATION:  // the window's last step
        launch_scatter_by_id(d->id.p, (const float*)h->es.p, 2, 1, 1, out, n, h->stream);
        count = (size_t)n;
        break;
    default: return set_error(SF_E_INVALID, "sf_diff_get takes SF_FIELD_POS, SF_FIELD_VEL or SF_FIELD_SATURATION");
    }
    SF_HIP(hipGetLastError());
    if (!on_device) SF_HIP(hipMemcpyAsync(dst, out, count * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    SF_HIP(hipStreamSynchronize(h->stream));
    return SF_OK;
}

int sf_diff_backward(sf_handle h, const float* pos_bar3, const float* vel_bar3, const float* sat_bar,
                     float* param_bar4, float* pos0_bar3, float* vel0_bar3, int on_device) {
    int rc;
    if ((rc = check_diff_handle(h))) return rc;
    DiffState* d = h->diff;
    if (!d || !d->have_forward) return set_error(SF_E_STATE, "sf_diff_forward has not run");
    const int n = (int)h->n;
    hipStream_t s = h->stream;
    if ((rc = alloc_tape(h, d->every))) return rc;
    // seeds, gathered from id order into the final slot order
    {
        DevBuf<float> stage;
        if (!on_device && (rc = stage.alloc(3 * (size_t)n))) return rc;
        const float* src = nullptr;
        if ((rc = to_device(h, pos_bar3, 3 * (size_t)n, on_device, stage.p, &src))) return rc;
        launch_gather_by_id(d->fin.id.p, src, 3, d->gx[0].p, n, s);
        if ((rc = to_device(h, vel_bar3, 3 * (size_t)n, on_device, stage.p, &src))) return rc;
        launch_gather_by_id(d->fin.id.p, src, 3, d->gv[0].p, n, s);
        if ((rc = to_device(h, sat_bar, (size_t)n, on_device, stage.p, &src))) return rc;
        launch_gather_by_id1(d->fin.id.p, src, d->sat_seed.p, n, s);
        SF_HIP(hipStreamSynchronize(s));  // `stage` is released at the end of this block
    }
    SF_HIP(hipMemsetAsync(d->gth.p, 0, (size_t)n * sizeof(float4), s));
    int cur = 0;
    const int nck = (d->steps + d->every - 1) / d->every;
    for (int c = nck - 1; c >= 0; --c) {
        const int s0 = c * d->every, s1 = std::min(d->steps, s0 + d->every);
        if (c < nck - 1 || !d->last_taped) {
            if ((rc = load_state(h, d->ck[c]))) return rc;
            for (int q = s0; q < s1; ++q)
                if ((rc = diff_step(h, d->dt, &d->tape[q - s0]))) return rc;
        }
        for (int q = s1 - 1; q >= s0; --q) {
            const StepTapeBufs& T = d->tape[q - s0];
            TapeStep t{T.pm.p, T.vr.p, T.rp.p, T.es.p, T.wt.p, T.flow.p, T.adh.p, T.xf.p, T.acc.p, T.den4.p,
                       T.ct.p, h->bpv.p, h->bcell.p, h->bnear.p};
            AdjFields a{d->gx[cur].p, d->gv[cur].p, d->gacc.p, d->grp.p, d->ges.p, d->gwt.p, d->gflow.p,
                        d->gadh.p, d->gxf.p, d->o4.p, d->o2.p, d->gth.p, d->o4b.p, d->o2b.p};
            launch_adj_step(h->K, t, a, q == d->steps - 1 ? d->sat_seed.p : nullptr, d->dt, n, s);
            launch_adj_scatter(T.perm.p, d->gx[cur].p, d->gv[cur].p, d->gx[cur ^ 1].p, d->gv[cur ^ 1].p, n, s);
            cur ^= 1;
        }
        SF_HIP(hipGetLastError());
    }
    // gx / gv now hold the adjoints of the window's initial state, in id order
    if (pos0_bar3 || vel0_bar3) {
        float* out = d->tmpf.p;
        for (int which = 0; which < 2; ++which) {
            float* dst = which ? vel0_bar3 : pos0_bar3;
            if (!dst) continue;
            const float4* src = which ? d->gv[cur].p : d->gx[cur].p;
            launch_iota(d->id_tmp.p, n, s);
            launch_scatter_by_id(d->id_tmp.p, (const float*)src, 4, 0, 3, on_device ? dst : out, n, s);
            if (!on_device)
                SF_HIP(hipMemcpyAsync(dst, out, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
            SF_HIP(hipStreamSynchronize(s));
        }
    }
    if (param_bar4) {
        std::vector<float4> g(n);
        SF_HIP(hipMemcpyAsync(g.data(), d->gth.p, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
        double acc[4] = {0, 0, 0, 0};
        for (const float4& e : g) {
            acc[0] += e.x;
            acc[1] += e.y;
            acc[2] += e.z;
            acc[3] += e.w;
        }
        for (int q = 0; q < 4; ++q) param_bar4[q] = (float)acc[q];
    }
    d->last_taped = false;  // the tape slots now hold the first segment
    // leave the handle at the end of the window, as after sf_diff_forward
    if ((rc = load_state(h, d->fin)) || (rc = copy(h->rp.p, d->fin_rp.p, (size_t)n * sizeof(float2), s)) ||
        (rc = copy(h->es.p, d->fin_es.p, (size_t)n * sizeof(float2), s)) || (rc = views_out(h)))
        return rc;
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(s));
    return SF_OK;
}

}  // extern "C"
