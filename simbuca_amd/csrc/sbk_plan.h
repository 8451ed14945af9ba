// sbk_plan.h -- host-only planning of the Coulomb work list (no HIP calls).
//
// A Coulomb evaluation of clouds c = 0..C-1 (n_c ions at offset off_c in
// the engine's arrays) is split into work items: one target block of a
// cloud against one slice of that cloud's sources.  Each slice writes its
// own partial-sum plane CP[slice], so the partial buffer must hold at least
// max_slices planes; check_plan() states that and every other property the
// kernels rely on, and the engine calls it on every list it builds.
#ifndef SBK_PLAN_H
#define SBK_PLAN_H

#include <string>
#include <vector>

#include "sbk_internal.h"

namespace sbk {

struct PlanCloud {
    int n = 0, off = 0;
    bool on = true;    // false: the cloud gets no items (Coulomb switched off)
};

struct Plan {
    std::vector<CoulombItem> items;
    std::vector<int> slices;   // per cloud (1 for clouds without items)
    int max_slices = 1;
    std::vector<int> xrun;     // 9 entries: XCD x takes items [xrun[x], xrun[x+1]) = x * grid / 8 ...
    int grid = 0;              // work-groups to launch = items, 8 runs padded with empty items
};

// Item-cost cap: unless exact, a cloud whose items would each cost more than
// 1/kItemShare of one resident work-group's share of the evaluation (its
// target block times its slice, against sum n^2 / capacity) gets its slices
// halved while each keeps at least kMinCapChunk sources (16 sub-tiles: 4 for
// each of 4 source streams).  Clouds of unequal sizes give items of unequal
// cost, and the longest set the tail of the dispatch.
constexpr long long kItemShare = 4;
constexpr int kMinCapChunk = 16 * kSubTile;

// targets_per_block: targets of one work-group (kBlock / groups * targets per
// thread).  slices_req > 0 fixes the slice count; 0 splits automatically so
// that the items make about eight rounds of `capacity` resident work-groups.
// quantum: a slice's length is a multiple of it (kSubTile * groups).
// exact: one slice per cloud (the FP64-exact sum is serial over sources).
// split_blk > 0 (only for kernels that run a block's filled target sets):
// with the automatic split, clouds whose items would be large against one
// work-group's share take blocks of split_blk targets.
Plan plan_coulomb(const std::vector<PlanCloud> &clouds, int targets_per_block, int quantum,
                  int slices_req, long long capacity, bool exact, int split_blk = 0);

// Every item targets and sources one cloud inside its range, its slice is
// below the cloud's slice count and that count at most max_slices; each
// cloud's items cover every (target, source) pair exactly once; the list is
// 8 equally long XCD runs, in which empty items (all zero) only pad.  Returns
// an empty string when all hold, else what failed.
std::string check_plan(const std::vector<PlanCloud> &clouds, const Plan &p);

}  // namespace sbk

#endif
