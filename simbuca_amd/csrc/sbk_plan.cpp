// sbk_plan.cpp -- the Coulomb work list of an engine's clouds (host only).
#include "sbk_plan.h"

#include <algorithm>
#include <cstdio>

namespace sbk {

Plan plan_coulomb(const std::vector<PlanCloud> &clouds, int blk, int quantum, int sreq,
                  long long capacity, bool exact, int split_blk)
{
    Plan L;
    L.slices.assign(clouds.size(), 1);
    // automatic split: about eight rounds of resident work-groups; short
    // items keep the last, partially filled round (the tail) short
    const long long target = 8 * capacity;
    // target blocks, pairs, and the items the list would hold with slices of
    // 512 sources
    long long total_blocks = 0, pairs = 0, items512 = 0;
    for (const PlanCloud &c : clouds)
        if (c.n > 0 && c.on) {
            total_blocks += (c.n + blk - 1) / blk;
            pairs += (long long)c.n * c.n;
            items512 += (long long)((c.n + blk - 1) / blk) * std::max(1, c.n / 512);
        }
    // slice count and length of a cloud of n ions when the list holds
    // total_blocks target blocks
    auto split = [&](int n, long long total_blocks, int &S, int &chunk) {
        S = 1;
        if (!exact) {
            if (sreq > 0)
                S = sreq;
            else {
                long long want = total_blocks > 0 ? (target + total_blocks / 2) / total_blocks : 1;
                // at least 512 sources per slice (partial sums stay a few % of the traffic),
                // relaxed to 64 when items of 512 sources would not fill one round of
                // the device (single clouds up to about 5,000 ions)
                long long cap = std::max(1, n / (items512 < capacity ? 64 : 512));
                S = (int)std::max(1LL, std::min(want, cap));
            }
        }
        S = std::max(1, std::min(S, n));
        // slice length: a multiple of the sources one LDS tile gives all wave
        // groups (each group takes its own kSubTile of every tile), so that
        // only a cloud's last slice ends in a partial tile
        chunk = (n + S - 1) / S;
        if (!exact) chunk = ((chunk + quantum - 1) / quantum) * quantum;
        S = (n + chunk - 1) / chunk;
    };
    // Target blocks per cloud.  The slice cap keeps the items of mid-size
    // clouds large (a few hundred sources against a whole block); when the
    // kernel runs a block's filled target sets only (split_blk > 0), a cloud
    // whose items would each cost more than a fifth of one resident
    // work-group's share of the evaluation takes blocks of split_blk
    // targets instead, which gives the last rounds smaller items without
    // shortening any source stream.  A single large cloud never qualifies:
    // its items are an eighth of the share.
    std::vector<int> cblk(clouds.size(), blk);
    if (!exact && sreq == 0 && split_blk > 0 && split_blk < blk && capacity > 0) {
        for (size_t ci = 0; ci < clouds.size(); ++ci) {
            const PlanCloud &c = clouds[ci];
            if (c.n <= 0 || !c.on) continue;
            int S, chunk;
            split(c.n, total_blocks, S, chunk);
            const long long item = (long long)std::min(blk, c.n) * std::min(chunk, c.n);
            if (5 * item * capacity > pairs) cblk[ci] = split_blk;
        }
        total_blocks = 0;
        for (size_t ci = 0; ci < clouds.size(); ++ci)
            if (clouds[ci].n > 0 && clouds[ci].on)
                total_blocks += (clouds[ci].n + cblk[ci] - 1) / cblk[ci];
    }
    for (size_t ci = 0; ci < clouds.size(); ++ci) {
        const PlanCloud &c = clouds[ci];
        if (c.n <= 0 || !c.on) continue;
        const int cb = cblk[ci];
        const int ib = (c.n + cb - 1) / cb;
        int S, chunk;
        split(c.n, total_blocks, S, chunk);
        // the item-cost cap (sbk_plan.h): below kMinCapChunk sources a
        // slice's per-item set-up costs more than the tail it saves
        if (!exact && capacity > 0 && pairs > 0) {
            const long long tb = std::min(cb, c.n);
            while (chunk >= 2 * kMinCapChunk && kItemShare * tb * chunk * capacity > pairs) {
                chunk = ((chunk / 2 + quantum - 1) / quantum) * quantum;
                S = (c.n + chunk - 1) / chunk;
            }
        }
        // slice boundaries (relative to the cloud's first ion)
        std::vector<int> cut;
        for (int s = 0; s <= S; ++s) cut.push_back(std::min(c.n, s * chunk));
        L.slices[ci] = S;
        L.max_slices = std::max(L.max_slices, S);
        for (int s = 0; s < S; ++s) {
            const int j0 = c.off + cut[s], j1 = c.off + cut[s + 1];
            for (int b = 0; b < ib; ++b) {
                CoulombItem w{};
                w.cloud = (int)ci;
                w.i0 = c.off + b * cb;
                w.i1 = c.off + std::min(c.n, (b + 1) * cb);
                w.j0 = j0;
                w.j1 = j1;
                w.slice = s;
                L.items.push_back(w);
            }
        }
    }
    // Work-group b runs on XCD b % 8 and takes the b / 8-th item of that
    // XCD's contiguous run (xcd_item in the kernels); an XCD starts its
    // work-groups in order.  The list is cut into eight contiguous runs of
    // about equal work, which keeps each cloud's items, and so its sources,
    // in one or two XCDs' L2; each run is ordered costliest first, so that
    // short items fill the XCD's last, partly empty round.  An item costs its
    // sources times the sets of 64 targets it fills: every lane loops over
    // all the sources, and the streamed kernel runs a cloud's last block with
    // only the sets it fills.  Runs of unequal counts are fine: the grid is
    // 8 times the longest, and the work-groups past a run's end leave.
    const int nb = (int)L.items.size();
    auto cost = [](const CoulombItem &u) {
        return (long long)((u.i1 - u.i0 + 63) / 64) * (u.j1 - u.j0);
    };
    long long total = 0;
    for (const CoulombItem &u : L.items) total += cost(u);
    L.xrun.assign(9, nb);
    L.xrun[0] = 0;
    long long acc = 0;
    for (int k = 0, x = 1; k < nb && x < 8; ++k) {
        acc += cost(L.items[k]);
        while (x < 8 && acc * 8 >= total * x) L.xrun[x++] = k + 1;
    }
    int longest = 0;
    for (int x = 0; x < 8; ++x) {
        std::stable_sort(L.items.begin() + L.xrun[x], L.items.begin() + L.xrun[x + 1],
                         [&cost](const CoulombItem &u, const CoulombItem &v) {
                             return cost(u) > cost(v);
                         });
        longest = std::max(longest, L.xrun[x + 1] - L.xrun[x]);
    }
    // Lay the runs out at x * longest and pad each to `longest` with empty
    // items, so that the kernels' equal-count mapping (8 runs of `longest`)
    // gives work-group b the (b / 8)-th item of XCD b % 8's run.
    if (nb > 0) {
        std::vector<CoulombItem> laid((size_t)8 * longest);   // value-initialized: empty
        for (int x = 0; x < 8; ++x)
            std::copy(L.items.begin() + L.xrun[x], L.items.begin() + L.xrun[x + 1],
                      laid.begin() + (size_t)x * longest);
        for (int x = 0; x < 9; ++x) L.xrun[x] = std::min(x, 8) * longest;
        L.items.swap(laid);
    }
    L.grid = 8 * longest;
    return L;
}

std::string check_plan(const std::vector<PlanCloud> &clouds, const Plan &p)
{
    char buf[256];
    if (p.slices.size() != clouds.size()) return "one slice count per cloud is needed";
    // the XCD runs: 8 equally long runs, x at x * grid / 8, that make up the list
    const int nb = (int)p.items.size();
    if (p.grid != nb || nb % 8 != 0 || p.xrun.size() != 9) return "the list is not 8 equal XCD runs";
    for (int x = 0; x <= 8; ++x)
        if (p.xrun[x] != x * (nb / 8)) return "the XCD runs do not start at multiples of grid / 8";
    for (size_t ci = 0; ci < clouds.size(); ++ci)
        if (p.slices[ci] < 1 || p.slices[ci] > p.max_slices) {
            std::snprintf(buf, sizeof(buf), "cloud %zu has %d slices, max_slices %d", ci,
                          p.slices[ci], p.max_slices);
            return buf;
        }
    // per cloud: pairs covered, summed over items, must be n^2 with no overlap
    std::vector<long long> covered(clouds.size(), 0);
    for (size_t k = 0; k < p.items.size(); ++k) {
        const CoulombItem &w = p.items[k];
        if (w.i0 == w.i1 && w.i0 == 0 && w.j0 == 0 && w.j1 == 0 && w.cloud == 0 && w.slice == 0)
            continue;   // an empty item padding an XCD's run
        if (w.cloud < 0 || w.cloud >= (int)clouds.size()) return "item names no cloud";
        const PlanCloud &c = clouds[w.cloud];
        const int lo = c.off, hi = c.off + c.n;
        if (!c.on || w.i0 < lo || w.i1 > hi || w.i0 >= w.i1 || w.j0 < lo || w.j1 > hi ||
            w.j0 >= w.j1) {
            std::snprintf(buf, sizeof(buf), "item %zu [%d,%d)x[%d,%d) leaves cloud %d [%d,%d)", k,
                          w.i0, w.i1, w.j0, w.j1, w.cloud, lo, hi);
            return buf;
        }
        if (w.slice < 0 || w.slice >= p.slices[w.cloud]) {
            std::snprintf(buf, sizeof(buf), "item %zu writes slice %d of cloud %d with %d slices",
                          k, w.slice, w.cloud, p.slices[w.cloud]);
            return buf;
        }
        covered[w.cloud] += (long long)(w.i1 - w.i0) * (w.j1 - w.j0);
    }
    // overlap: items of one slice partition the targets, and the slices of
    // one target block partition the sources
    for (size_t ci = 0; ci < clouds.size(); ++ci) {
        const PlanCloud &c = clouds[ci];
        const long long want = (c.on && c.n > 0) ? (long long)c.n * c.n : 0;
        if (covered[ci] != want) {
            std::snprintf(buf, sizeof(buf), "cloud %zu: items cover %lld pairs of %lld", ci,
                          covered[ci], want);
            return buf;
        }
    }
    std::vector<std::vector<int>> tgt(clouds.size()), src(clouds.size());
    for (size_t ci = 0; ci < clouds.size(); ++ci) {
        tgt[ci].assign((size_t)std::max(0, clouds[ci].n), 0);
        src[ci].assign((size_t)std::max(0, clouds[ci].n), 0);
    }
    for (const CoulombItem &w : p.items) {
        const int off = clouds[w.cloud].off;
        if (w.slice == 0)
            for (int i = w.i0; i < w.i1; ++i) tgt[w.cloud][i - off]++;
        if (w.i0 == off)
            for (int j = w.j0; j < w.j1; ++j) src[w.cloud][j - off]++;
    }
    for (size_t ci = 0; ci < clouds.size(); ++ci) {
        if (!clouds[ci].on) continue;
        for (int i = 0; i < clouds[ci].n; ++i)
            if (tgt[ci][i] != 1 || src[ci][i] != 1) {
                std::snprintf(buf, sizeof(buf), "cloud %zu ion %d: %d target blocks, %d slices", ci,
                              i, tgt[ci][i], src[ci][i]);
                return buf;
            }
    }
    return std::string();
}

}  // namespace sbk
