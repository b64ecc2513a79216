// sf_internal.h -- host-side structures shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sf_seepage.h"
#include "sf_device.h"

namespace sf {

// Thread-local error text behind sf_last_error().
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);
#define SF_HIP(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return ::sf::hip_error(e_, #call);  \
    } while (0)

// Owning device buffer (the role of CudaArray, cuda_array.cuh:19-61).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return hip_error(e, "hipMalloc");
        }
        n = count;
        return SF_OK;
    }
};

// LSD radix sort plan over keys in [0, C): balanced digit widths <= 8 bits.
struct RadixPlan {
    int passes = 1;
    int shift[4] = {0, 0, 0, 0};
    int bits[4] = {1, 0, 0, 0};
};
RadixPlan make_radix_plan(uint32_t num_cells);

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;

struct SortBuffers {
    uint32_t* kA;
    uint32_t* vA;
    uint32_t* kB;
    uint32_t* vB;
    uint32_t* hist;  // 256 * ceil(n / kRsTile) entries
};
// Stable sort of keys[0, n): on return *sorted_keys / *perm point into the
// ping-pong buffers (perm[new] = old index).
int radix_sort(const uint32_t* keys, int n, const RadixPlan& plan, const SortBuffers& b, const uint32_t** sorted_keys,
               const uint32_t** perm, hipStream_t s);

// ---- searcher launchers (sf_searcher.hip) ----
// scene (batched handles): per-slot scene id, keys offset by scene * C; null for one scene
void launch_hash(const SFConst& k, const float4* pm, const uint16_t* scene, uint32_t* keys, int n, hipStream_t s);
// Positions in phase order, written by the cell-start launch when pph is set:
// pph[t] = pm[plist[t]] for t < n.
struct PhasePos {
    const uint32_t* plist;
    const float4* pm;
    float4* pph;
};
// After an incremental sort: cellStart[c] is then the stayers before c's old
// start plus the newcomers of the cells before c.
struct IncStart {
    const uint32_t* stay_pre;  // n + 1
    const uint32_t* in_start;  // C + 1
};
// cellStart[c] for c in [0, C] plus the cell table ct[c] = (cellStart[c], wpre[cellStart[c]]);
// zero_c (may be null) is cleared over [0, C) (the incremental sort's newcomer counters).
// With inc.stay_pre set, cell_start holds the previous sort's cellStart on
// entry and each entry is updated in place from it; otherwise each cell's
// start is found by binary search over sorted_keys.
void launch_cell_start(const uint32_t* sorted_keys, int n, uint32_t C, const uint32_t* wpre, uint32_t* cell_start,
                       uint2* ct, uint32_t* zero_c, PhasePos pp, IncStart inc, hipStream_t s);
struct GatherArgs {
    const uint32_t* perm;
    const float4 *pm_in, *vr_in, *cs_in, *cda_in;
    const float2* amc_in;
    float4 *pm_out, *vr_out, *cs_out, *cda_out;
    float2* amc_out;
    uint8_t* wflag;  // out: 1 for water slots (read by the water prefix)
    int multi;
};
// Reorders the persistent state by perm, then wpre[0..n] = water count among
// slots [0, i) and plist = water slots then sand slots, each in slot order.
// wpart holds wprefix_blocks(n + 1) + 1 entries.
// Largest tile count up to which the prefix launches sum the earlier tiles
// themselves (above it a one-block scan runs in between); < 0 restores the
// default.  Returns the previous value.  Process-wide.
int set_direct_tiles(int tiles);
void launch_gather_wprefix(const GatherArgs& a, int n, uint32_t* wpart, uint32_t* wpre, uint32_t* plist,
                           hipStream_t s);
void launch_boundary_gather(const uint32_t* perm, const float4* bp_in, float4* bp_out, int nb, hipStream_t s);
void launch_boundary_volume(const SFConst& k, float4* bpv, const uint32_t* bcs, int nb, hipStream_t s);
// near[c] for every local cell c: whether any of its 9 neighbour column
// ranges of the boundary cellStart bcs is non-empty.
void launch_boundary_near(const SFConst& k, const uint32_t* bcs, uint8_t* near, hipStream_t s);
void launch_nbr_count(const SFConst& k, const float4* pm, const uint16_t* scene, const uint32_t* cs, uint32_t* cnt,
                      uint32_t* cand, int n, hipStream_t s);
void launch_hash_boundary(const SFConst& k, const float4* bp, uint32_t* keys, int nb, hipStream_t s);
// Incremental stable sort (the particles are in the previous sort's order):
// key / old_key per slot, old_start = the previous cellStart (C + 1).
struct IncSortBuffers {
    uint32_t* in_count;  // C + 1, zero on entry (launch_cell_start clears it)
    uint32_t* in_start;  // C + 1
    uint32_t* slot;      // n
    uint32_t* in_list;   // n
    uint32_t* stay_pre;  // n + 1
    uint32_t* part;      // vprefix2_partials(n, C)
};
// flags_done: the newcomer ranks (slot) and counts (in_count) were already
// taken, by the advect that wrote key
void launch_incremental_sort(const uint32_t* key, const uint32_t* old_key, const uint32_t* old_start, int n,
                             uint32_t C, const IncSortBuffers& b, uint32_t* perm, uint32_t* sorted_keys,
                             bool flags_done, hipStream_t s);
// partial-sum counts of the prefix kernels
int wprefix_blocks(int n);
int vprefix2_partials(int n1, int n2);

// ---- solver launchers (sf_solver.hip) ----
struct Fields {
    const float4* pm;
    const float4* vr;
    float4* cs;
    const float4* cda;
    const uint2* ct;  // (cellStart[c], water prefix at cellStart[c]), c in [0, C]
    const float4* bpv;
    const uint32_t* bcell;
    // bnear[c] = 1 when a boundary particle lies in cell c's 27-cell
    // neighbourhood (local cell keys; set at boundary upload)
    const uint8_t* bnear;
    float2* rp;
    float2* es;
    float2* wt;  // water: (rho / max(eps, voidage), p / max(eps, (rho / max(eps, voidage))^2))
    float4* flow;
    float4* adh;
    float4* xf;
    float4* acc;
    // differentiable window only (null otherwise): per-particle sums the
    // reverse passes need: (P2 weight sum, P3 sand-voidage / water vs
    // denominator, P4 vw, unused)
    float4* den4;
    // batched handles only (null otherwise): per-slot scene id, per-scene constants
    const uint16_t* scene;
    const SceneConst* sct;
    // slots by phase: plist[0, nw) the water slots, plist[nw, n) the sand
    // slots, each in slot order; *nwater = nw (device)
    const uint32_t* plist;
    const uint32_t* nwater;
    // pph[t] = pm[plist[t]]: positions and masses in phase order, so the
    // water candidates of a column range [s, e) are pph[wpre[s], wpre[e]) and
    // the sand ones pph[nw + s - wpre[s], nw + e - wpre[e])
    const float4* pph;
    // interface particles of the avg-flow / voidage-adhesion passes (phase
    // indices t): water at ilist[0, icount[0]), sand at ilist[nw, nw + icount[1])
    uint32_t* ilist;
    uint32_t* icount;
    // iface[ts] = 1 when sand particle ts (index among the sand, plist[nw + ts])
    // has water in its 9 column ranges (written by the density pass, read by
    // the sand momentum pass to skip an empty water walk)
    uint8_t* iface;
    // wl[t] = 1 when water particle t (phase index) is listed (density pass).
    // Unlisted particles' avg flow, adhesion, water voidage / saturation and
    // avgAdhesion are the empty sums' constants; the density pass writes them
    // only when full_empty is set (or den4, the differentiable window),
    // otherwise they are filled in on request (launch_fill_empty).  The
    // fields a later pass reads (sand voidage / saturation, water wt) are
    // always written.
    uint8_t* wl;
    int full_empty;
    // Fields of the listed particles (those with the other phase in their 9
    // column ranges) by phase index t, written only for those:
    //   vq[t] = the velocity (density pass); for every grain, not only the
    //   listed ones, with its saturation in vq[t].w (0 from the density
    //   pass, the avg-flow pass's value for the listed);
    //   water: wq[t] = (rho, p) (density pass), voidage in wq[t].z (avg
    //   flow), xq[t] = avgAdhesion (voidage/adhesion pass);
    //   sand: saturation in wq[t].w (avg flow).
    // Every other-phase neighbour within h of a particle is listed, so the
    // interface and sand momentum passes read these without the slot
    // indirection plist[t].
    float4* vq;
    float4* wq;
    float4* xq;
    // wbq[t] = (position, m / rho~) of listed water t, and xq[t].w = its
    // m p / max(eps, rho_sand rho~) (voidage/adhesion pass; sand momentum's
    // buoyancy pairs)
    float4* wbq;
    // Neighbour records of the water momentum pass, by slot, written when
    // rec is set (every water particle has the same mass mw): water
    // rec0 = (position, rho~), rec1 = (velocity, p / rho~^2); sand
    // rec0 = (position, -mass), rec1 = avgDrag.  A neighbour then costs two
    // 16-byte loads instead of the three of pm, wt and vr (or pm and xf).
    // With rec set, wt is written only for the differentiable window (den4).
    float4* rec0;
    float4* rec1;
    float mw;
    int rec;
};
// Launch shape of a neighbour pass: workgroup size and the per-lane LDS
// candidate-list capacity (the latter only for the passes that use one).
struct PassCfg {
    int nb = 256;
    int cap = 16;
    int lanes = 4;  // lanes per particle (interface passes P2, P3)
};
constexpr int kNbChoices[3] = {64, 128, 256};
constexpr int kCapChoices[2] = {8, 16};
// The density pass also lists the interface particles of the avg-flow and
// voidage/adhesion passes (and writes the empty sums' results of the rest)
// into counters that start at zero; the water momentum pass clears them for
// the next step.  A caller that runs the density pass alone clears them
// first (launch_zero_counts); one that runs the avg-flow pass alone clears
// them and builds the lists with launch_interface.
void launch_density(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s);
void launch_avg_flow(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s);
void launch_zero_counts(const Fields& f, hipStream_t s);
void launch_interface(const SFConst& k, const Fields& f, int n, hipStream_t s);
// The empty sums' constants of every unlisted particle's avg flow, adhesion,
// water voidage / saturation and avgAdhesion, and every grain's (rho_sand, 0)
// density / pressure (what full_empty writes), over the n particles of the
// last neighbour build.
void launch_fill_empty(const SFConst& k, const Fields& f, int n, hipStream_t s);
void launch_voidage_adhesion(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s);
void launch_sand_momentum(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s);
void launch_water_momentum(const SFConst& k, const Fields& f, int n, PassCfg c, hipStream_t s);
// CFL reduction: partial maxima of |acc| -> finalize writes CflOut
struct CflOut {
    float dt;
    uint32_t nsub;
    float amax;
    float ri;  // the render interval this sub-step used (dt of the advect = ri / nsub)
};
// Device-side state of a frame loop (main.cpp:235-247 run without a host
// round trip per sub-step).  With kFrameAdaptive the finalize takes the
// render interval from `remaining` and then applies the loop's update
// remaining -= remaining / nsub; with kFrameElapsed it also advances the
// emitter clock elapsed += ri / nsub.  `count` counts the sub-steps.
struct FrameDev {
    float remaining;
    float elapsed;
    uint32_t count;
    uint32_t flags;
};
constexpr uint32_t kFrameAdaptive = 1u;
constexpr uint32_t kFrameElapsed = 2u;
int cfl_partials(int n);
// frame (may be null): see FrameDev; when null the render interval is the argument
void launch_cfl(const SFConst& k, const float4* acc, int n, float* partials, CflOut* out, float render_interval,
                float sph_mass, float speed_of_sound, float scale, float params_dt, FrameDev* frame, hipStream_t s);
struct AdvectArgs {
    const float4* pm_in;
    const float4* vr_in;
    float4* acc;
    float4* pm_out;
    float4* vr_out;
    const float4* bpv;
    const uint32_t* bcell;
    float dt;
    const CflOut* dt_dev;  // when non-null, dt = ri / nsub from the device
    // The next neighbour build's hash and change flags, written here from the
    // new positions when keys_out is non-null (incremental sort): keys_out[i]
    // = cell key of the new position (offset by scene * C for batched
    // handles), and where it differs from old_keys[i] (this step's sorted key
    // of slot i) slot[i] = the newcomer rank taken from in_count.
    uint32_t* keys_out;
    const uint32_t* old_keys;
    uint32_t* in_count;
    uint32_t* slot;
    const uint16_t* scene;
};
void launch_advect(const SFConst& k, const AdvectArgs& a, int n, hipStream_t s);

}  // namespace sf

namespace sf {
// ---- reference-layout views (sf_views.hip) ----
// The library's storage (current buffers) and the caller-visible arrays in
// the layout of CudaSFParticles::Get*Ptr(): packed float3, float, uint64
// labels, float2.  A view pointer is null when that view was not handed out.
struct ViewArgs {
    float4* pm;  // xyz position, w = +mass (water) / -mass (sand)
    float4* vr;  // xyz velocity, w radius
    float4* cs;  // xyz colour, w maxSaturation
    float4* cda;
    float2* amc;
    const float4* acc;
    const float2* rp;    // density, pressure
    const float2* es;    // voidage, saturation
    const float4* flow;  // avg flow velocity
    const float4* xf;    // avgDrag (sand slots) / avgAdhesion (water slots)
    const float4* adh;   // adhesion (water slots)
    float* pos;
    float* vel;
    float* accv;
    float* col;
    uint64_t* label;
    float* mass;
    float* radius;
    float* density;
    float* pressure;
    float* voidage;
    float* saturation;
    float* maxsat;
    float* avgflow;
    float* avgdrag;
    float* adhesion;
    float* avgadh;
    float* cda3;
    float2* amc2;
};
// storage -> every view that is set
void launch_views_out(const ViewArgs& a, int n, hipStream_t s);
// persistent views that are set (pos, vel, col, label, mass, radius,
// maxsat, cda3, amc2) -> storage
void launch_views_in(const ViewArgs& a, int n, hipStream_t s);
// out[i] = in[i] widened to 64 bits (cell starts as the reference's size_t)
void launch_widen_u32(const uint32_t* in, uint64_t* out, int n, hipStream_t s);
}  // namespace sf
