// sf_nlist.h -- per-lane candidate lists in LDS for the neighbour passes
// (sf_solver.hip) and their reverse passes (sf_adjoint.hip).
#pragma once
#include "sf_device.h"

namespace sf {

// Per-thread candidate lists in LDS.  A neighbour loop first filters the
// candidates of its cell ranges (cheap: one load, r^2, compare) and appends
// the survivors to the lane's list; the expensive pair bodies then run over
// the compacted lists, so a wave no longer executes every body for every
// candidate that some lane happens to keep.  The flush is decided wave-wide
// before appending (a single append, or a round of four in filter_push), so
// every list stays within CAP + 3.  NB is the workgroup size.
template <int NB, int CAP>
struct NList {
    uint32_t* slot;  // this lane's column of the LDS list: slot[k * NB]
    int cnt;
};
// CAP entries per lane trigger a flush; three more slots take the overflow
// of a four-candidate round (filter_push checks once per round).
#define SF_NLIST(name)                                   \
    __shared__ uint32_t name##_lds[(CAP + 3) * NB];      \
    NList<NB, CAP> name{name##_lds + threadIdx.x, 0}

// The pair body is split in two: fetch(j) issues every load a pair needs,
// with no branch on loaded data, and apply(j, data) does the arithmetic.  The flush is software-pipelined: the next pair's loads are
// issued before the current pair's arithmetic, so a flush no longer waits on
// one pair's dependent load chain at a time.  apply has a single call site,
// so every pair gets the same instruction sequence however the flushes fall.
template <int NB, int CAP, class Fetch, class Apply>
SF_DEV void nl_flush2(NList<NB, CAP>& L, Fetch&& fetch, Apply&& apply) {
    if (L.cnt > 0) {
        uint32_t j = L.slot[0];
        auto cur = fetch(j);
        for (int q = 0; q < L.cnt; ++q) {
            const uint32_t jn = L.slot[(q + 1 < L.cnt ? q + 1 : q) * NB];
            const auto nxt = fetch(jn);
            apply(j, cur);
            j = jn;
            cur = nxt;
        }
    }
    L.cnt = 0;
}
template <int NB, int CAP, class Fetch, class Apply>
SF_DEV void nl_push2(NList<NB, CAP>& L, bool keep, uint32_t j, Fetch&& fetch, Apply&& apply) {
    if (__ballot(L.cnt >= CAP)) nl_flush2(L, fetch, apply);
    L.slot[L.cnt * NB] = j;
    L.cnt += keep ? 1 : 0;
}

}  // namespace sf
