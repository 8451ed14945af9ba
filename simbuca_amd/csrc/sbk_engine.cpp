// sbk_engine.cpp -- host side of the device-resident Dormand-Prince engine
// (export set sbk_* of include/simbuca_hip.h).
//
// The engine replaces the reference's per-step host loop
// (bin/ion_fly.cpp:83-101 -> step(), bin/ode.cpp:459-494 ->
// Dormand_Prince_5, bin/ode.cpp:61-218 -> force(), bin/force.cpp:34-175).
// Ion state lives in HBM from sbk_add_cloud to sbk_destroy; one attempt is
// 7 x (Coulomb kernel + stage kernel) + controller + commit, all enqueued
// on the engine's stream (optionally replayed from one hipGraph), and the
// host only reads the per-cloud scalars back between batches of attempts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/simbuca_hip.h"
#include "sbk_internal.h"
#include "sbk_plan.h"

namespace {

// bin/globals.h:6-10
const double kPi = 3.14159265358979323;
const double kElCharge = 1.60217733e-19;
const double kKe = 8.987551787E9;

thread_local std::string g_err;

}  // namespace

int sbk::record_error(int code, const char *msg)
{
    g_err = msg;
    return code;
}

namespace {

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(SBK_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                               \
    } while (0)

struct HostCloud {
    int n = 0, off = 0;
    sbk_fields f{};
    double lifetime = 0.0;
    std::vector<double> pos, vel, mass;   // in device (slot) order after finalize
    std::vector<int> charge;
    std::vector<int> perm;                // slot -> caller's index (empty: identity)
    std::vector<double> swift;            // SWIFT samples, copied from the fields
    std::vector<double> emap;             // field map (E_r, E_z per node), copied from the fields
    int xcross_on = 0;                    // record crossings of the plane x = xcross
    double xcross = 0.0;
};

// A crossing taken off the device list, with the ion's caller index.
struct HostCrossing {
    int cloud, ion;
    long long step;
    double t, p[3], v[3];
};

struct ItemList {
    sbk::CoulombItem *d_items = nullptr;
    int *d_slices = nullptr;
    int nitems = 0;
    int max_slices = 1;
    std::vector<int> slices;
};

}  // namespace

struct sbk_engine {
    sbk_config cfg{};
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    bool finalized = false;
    std::vector<HostCloud> clouds;
    int ntot = 0;
    long long stride = 0;

    // device memory
    double *d_f64 = nullptr;   // all FP64 SoA arrays
    float *d_qf = nullptr;
    sbk::IonBlock *d_blocks = nullptr;
    double *d_errb = nullptr;
    std::vector<int> blk0, nblk;
    double *d_cp = nullptr;
    int cp_slices = 0;
    sbk::CloudParams *d_params = nullptr;
    double *d_swift = nullptr;             // every cloud's SWIFT samples
    size_t swift_cap = 0;
    std::vector<size_t> swift_off;
    sbk::CloudState *d_state = nullptr;
    sbk::CloudState *h_state = nullptr;   // pinned mirror
    double *d_kc = nullptr, *d_qd = nullptr;
    bool unit_charges = false;   // every ion has charge +1
    bool spatial = false;        // ions kept in a spatial order (mixed precisions)
    int *d_src = nullptr;        // gather indices of a re-sort
    long long attempts_at_order = 0;
    long long reorders = 0, reorders_at_reset = 0;
    sbk::EngineDev dev{};

    // crossing list, allocated once a cloud records crossings
    sbk::XCrossing *d_xc = nullptr;
    int *d_xc_count = nullptr;
    int xc_cap = 0;
    std::vector<HostCrossing> xc_host;   // entries drained from the device

    // launch shape (tpt 0: automatic)
    int tpt = 0;
    int groups = 4;
    int slices_req = 0;
    int local_req = -1;          // -1: automatic
    int kernel_req = -1;         // SBK_KERNEL_*; -1: automatic
    std::map<std::tuple<int, int, int, int, int, int, int, long long>, ItemList> lists;
    // clouds the attempt work list covers: all (empty mask, id 0), or those
    // still active in an advance once the others hold a tenth of its pairs
    std::vector<char> list_mask;
    long long list_mask_id = 0, list_masks = 0;
    double list_mask_pairs = 0.0;   // sum of n^2 over the covered Coulomb clouds

    // persistent attempts: grid-barrier counters and the timeout flag
    unsigned *d_bar = nullptr;
    int *d_timeout = nullptr;
    bool pers_launched = false;   // a persistent launch ran since the states were last read
    long long pers_count = 0;     // persistent launches since the last reset_stats
    int coop = 0;                 // the device supports cooperative launches

    // graph of one attempt
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    // timing
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
    double coulomb_ms = 0.0;
    long long coulomb_launches = 0;
    std::vector<long long> attempts_at_reset, steps_at_reset, evals_at_reset;
};

namespace {

sbk::CloudParams make_params(const sbk_fields &f, int off, int n)
{
    sbk::CloudParams p{};
    p.trap = f.trap;
    p.tof = f.tof;
    p.coulomb = f.coulomb;
    p.ramp = f.ramp;
    p.kappa = f.kappa;
    p.Vdc = f.Vdc;
    p.Vrf = f.Vrf;
    p.r02 = f.r0 * f.r0;                       // bin/force.cpp:42-43
    p.z02 = f.z0 * f.z0;
    p.omega_rf = f.freq_rf * 2 * kPi;          // bin/force.cpp:44
    p.E_kick = f.E_kick;
    p.ramp_t0 = f.ramp_t0;
    p.ramp_t1 = f.ramp_t1;
    p.newVrf = f.newVrf;
    p.newVdc = f.newVdc;
    p.cpen = (f.V0 != 0.0) ? f.V0 / (2.0 * f.d0 * f.d0) : 0.0;
    p.Bz = f.Bz;
    p.Ed = f.Ed;
    p.fd = f.fd;
    p.phid = f.phid;
    p.Uq_rq2 = (f.Uq != 0.0) ? f.Uq / (f.rq * f.rq) : 0.0;
    p.fq = f.fq;
    p.phiq = f.phiq;
    p.delta = f.delta;
    p.fd_rate = f.fd_rate;
    p.fq_rate = f.fq_rate;
    p.Uo_ro4 = (f.Uo != 0.0) ? f.Uo / (f.ro * f.ro * f.ro * f.ro) : 0.0;
    p.fo = f.fo;
    p.phio = f.phio;
    p.rw_order = f.rw_order;
    p.Uw_rwl = (f.Uw != 0.0) ? ((f.rw_order == 1) ? f.Uw / f.rw : f.Uw / (f.rw * f.rw)) : 0.0;
    p.fw = f.fw;
    p.phiw = f.phiw;
    p.Eax = f.Eax;
    p.fax = f.fax;
    p.phiax = f.phiax;
    p.B2 = f.B2;
    p.inv_K = (f.K0 != 0.0 && f.p_mbar != 0.0) ? (f.p_mbar / 1013.25) * (273.15 / f.T_gas) / f.K0
                                                : 0.0;
    p.swift = nullptr;   // set by cloud_params
    p.swift_n = f.swift_n;
    p.swift_kind = f.swift_kind;
    p.swift_t0 = f.swift_t0;
    p.swift_dt = f.swift_dt;
    p.swift_r = f.swift_r;
    p.c4k = (f.V0 != 0.0 && f.C4 != 0.0) ? f.V0 * f.C4 / (2.0 * f.d0 * f.d0 * f.d0 * f.d0) : 0.0;
    p.c6k = (f.V0 != 0.0 && f.C6 != 0.0)
                ? f.V0 * f.C6 / (2.0 * f.d0 * f.d0 * f.d0 * f.d0 * f.d0 * f.d0) : 0.0;
    p.emap = nullptr;    // set by cloud_params
    p.emap_nr = f.emap_nr;
    p.emap_nz = f.emap_nz;
    p.emap_dr = f.emap_dr;
    p.emap_z0 = f.emap_z0;
    p.emap_dz = f.emap_dz;
    p.emap_scale = f.emap_scale;
    p.off = off;
    p.n = n;
    p.slices = 1;
    return p;
}

int kp(int precision)
{
    return precision;  // SBK_* and sbk::KP_* share values
}

// Spatial order of one cloud: recursive bisection along the longest axis of
// the bounding box, where the first part holds a multiple of the largest
// power of two A <= 1024 with 2A <= len ions.  Every aligned run of
// 64 * 2^k <= 1024 ions (the Coulomb kernel's target blocks and source
// sub-tiles, which start at the cloud's first ion) is then one subtree, so
// the ions of a block are close together.  A leaf (at most 64 ions) is
// sorted by position, so the order depends only on the positions, not on
// the order they came in: re-sorting a sorted cloud changes nothing.
// pos is AoS [n][3].  A NaN coordinate sorts as +infinity, so the
// comparisons stay a strict weak order whatever the positions hold.
inline double order_key(double v) { return std::isnan(v) ? INFINITY : v; }

void bisect(const double *pos, int *idx, int len)
{
    if (len <= 64) {
        std::sort(idx, idx + len, [pos](int a, int b) {
            const double *p = pos + 3 * (size_t)a, *q = pos + 3 * (size_t)b;
            for (int k = 0; k < 3; ++k) {
                const double u = order_key(p[k]), v = order_key(q[k]);
                if (u != v) return u < v;
            }
            return a < b;
        });
        return;
    }
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < len; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = pos[3 * (size_t)idx[i] + k];
            mn[k] = std::min(mn[k], v);
            mx[k] = std::max(mx[k], v);
        }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
    int A = 1024;
    while (A > 64 && 2 * A > len) A /= 2;
    const int m = std::max(A, (int)std::lround(len / (2.0 * A)) * A);   // A <= m < len
    std::nth_element(idx, idx + m, idx + len, [pos, ax](int a, int b) {
        return order_key(pos[3 * (size_t)a + ax]) < order_key(pos[3 * (size_t)b + ax]);
    });
    bisect(pos, idx, m);
    bisect(pos, idx + m, len - m);
}

std::vector<int> spatial_order(const double *pos, int n)
{
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    if (n > 0) bisect(pos, ord.data(), n);
    return ord;
}

template <typename T>
void apply_order(std::vector<T> &v, const std::vector<int> &ord, int width)
{
    std::vector<T> out(v.size());
    for (size_t i = 0; i < ord.size(); ++i)
        for (int k = 0; k < width; ++k) out[width * i + k] = v[width * (size_t)ord[i] + k];
    v.swap(out);
}

// caller's index of the ion in slot i of a cloud
inline int caller_index(const HostCloud &c, int i)
{
    return c.perm.empty() ? i : c.perm[i];
}

// The attempt graph's kernels hold the work list, partial-sum buffer and
// ion arrays by address: the stream is drained first, so no replay still in
// flight reads memory that the caller frees or reallocates next.
void destroy_graph(sbk_engine *e)
{
    if ((e->graph_exec || e->graph) && e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->graph_exec = nullptr;
    e->graph = nullptr;
}

void free_lists(sbk_engine *e)
{
    destroy_graph(e);   // first: the graph reads the lists
    if (e->stream) (void)hipStreamSynchronize(e->stream);   // eager launches read them too
    for (auto &kv : e->lists) {
        (void)hipFree(kv.second.d_items);
        (void)hipFree(kv.second.d_slices);
    }
    e->lists.clear();
}

// Sum of n^2 over the clouds with a Coulomb term, all of them or those whose
// state is active.
double coulomb_pairs(const sbk_engine *e, bool active_only)
{
    double p = 0.0;
    for (size_t i = 0; i < e->clouds.size(); ++i)
        if (e->clouds[i].f.coulomb && (!active_only || e->h_state[i].active))
            p += (double)e->clouds[i].n * e->clouds[i].n;
    return p;
}

// Sets the clouds the attempt work list covers (an empty mask: all).  The
// attempt graph and the lists of the previous mask go; the stream must be
// idle of attempts (the callers have just read the states).
void set_list_mask(sbk_engine *e, std::vector<char> mask)
{
    if (mask == e->list_mask) return;
    destroy_graph(e);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto it = e->lists.begin(); it != e->lists.end();) {
        if (std::get<7>(it->first) != 0) {
            (void)hipFree(it->second.d_items);
            (void)hipFree(it->second.d_slices);
            it = e->lists.erase(it);
        } else {
            ++it;
        }
    }
    e->list_mask = std::move(mask);
    e->list_mask_id = e->list_mask.empty() ? 0 : ++e->list_masks;
    e->list_mask_pairs = e->list_mask.empty() ? coulomb_pairs(e, false) : 0.0;
}

// In a run of attempts, clouds that reached their end time stay in the work
// list and their items leave at once, but the list's XCD runs were cut for
// all clouds: once the active clouds hold less than nine tenths of the
// list's pairs, the list is planned again for them alone, so their items are
// spread over all XCDs again (the ensembles' large clouds need the most
// attempts; measured on the 256-cloud ensemble: nine tenths 1% faster than
// three quarters, one half 1% slower).  Deterministic: it depends on the
// cloud states only.
void narrow_list(sbk_engine *e)
{
    const double act = coulomb_pairs(e, true);
    if (act >= 0.9 * e->list_mask_pairs) return;
    std::vector<char> mask(e->clouds.size(), 0);
    for (size_t i = 0; i < e->clouds.size(); ++i) mask[i] = e->h_state[i].active ? 1 : 0;
    set_list_mask(e, std::move(mask));
    e->list_mask_pairs = act;
}

// Whether the mixed Coulomb kernel uses FP32 offsets for far sources.  It
// pays off once a target block is a small part of its cloud and the work
// items are long.  Fitted to autotuner runs on MI355X (each shape timed with
// the read-back of its partial sums, automatic slices): on for single clouds
// of 10,000-14,000 ions (5-16% faster than off) and ensembles whose
// pair-weighted mean cloud size is 2,800-12,000 ions (13-30%); off for single
// clouds of 500-8,000 ions (off 2-24% faster: most sources are near a small
// cloud's blocks, and the items, split finely to fill the device, are short)
// and for ensembles of 64-512-ion clouds (off 19% faster).  Hence both the
// total pairs and the pair-weighted mean cloud size must be large.
int local_coords(const sbk_engine *e, int precision)
{
    if (!e->spatial || (precision != SBK_MIXED && precision != SBK_MIXED_KAHAN)) return 0;
    if (e->local_req >= 0) return e->local_req;
    double n2 = 0.0, n3 = 0.0;
    for (auto &c : e->clouds) {
        n2 += (double)c.n * c.n;
        n3 += (double)c.n * c.n * c.n;
    }
    return n2 >= 8e7 && n3 / n2 >= 2048.0 ? 1 : 0;
}

// Targets per thread.  0 (the default) picks by size, fitted to MI355X
// autotuner runs: 4 once a cloud fills the device with long work items
// (pair-weighted mean size >= 32,768: 65,536 ions 3% faster than 2, 32,768
// ions 1%; at 20,000 ions 2 is 5% faster than 4); 2 with local coordinates
// (single clouds of 10,000-20,000 ions and ensembles of 2,800-12,000-ion
// clouds: the far path's box of 128 targets) and for many pairs in total;
// 1 for small clouds, where more, shorter work-groups win.
int targets_per_thread(const sbk_engine *e, int precision)
{
    if (e->tpt > 0) return e->tpt;
    double n2 = 0.0, n3 = 0.0;
    for (auto &c : e->clouds) {
        n2 += (double)c.n * c.n;
        n3 += (double)c.n * c.n * c.n;
    }
    if (n2 > 0.0 && n3 / n2 >= 32768.0) return 4;
    return (n2 >= 2.5e8 || local_coords(e, precision)) ? 2 : 1;
}

// Coulomb kernel of a launch: the streamed kernel exists for the mixed
// precisions with 4 source groups and 1, 2 or 4 targets per thread.
int coulomb_kernel(const sbk_engine *e, int precision)
{
    if (precision != SBK_MIXED && precision != SBK_MIXED_KAHAN) return sbk::KK_TILED;
    if (e->groups != 4) return sbk::KK_TILED;
    const int tpt = targets_per_thread(e, precision);
    if (tpt != 1 && tpt != 2 && tpt != 4) return sbk::KK_TILED;
    // automatic: the streamed kernel, measured on MI355X as fast or faster than the
    // tiled one on every benchmark workload (65,536 ions -0.4%, 262,144 ions -0.3%)
    return e->kernel_req == SBK_KERNEL_TILED ? sbk::KK_TILED : sbk::KK_STREAM;
}

// Builds (or returns the cached) Coulomb work list for a precision and
// launch shape.  `only_coulomb` keeps the clouds whose Coulomb switch is on.
int get_list(sbk_engine *e, int precision, bool only_coulomb, ItemList **out)
{
    const bool exact = precision == SBK_FP64_EXACT;
    const int tpt = exact ? 1 : targets_per_thread(e, precision);
    const int q = exact ? 1 : e->groups;
    const int sreq = exact ? 1 : e->slices_req;
    const int kern = coulomb_kernel(e, precision);
    const long long mask_id = only_coulomb ? e->list_mask_id : 0;
    auto key = std::make_tuple(precision, tpt, q, sreq, only_coulomb ? 1 : 0,
                               local_coords(e, precision), kern, mask_id);
    auto it = e->lists.find(key);
    if (it != e->lists.end()) {
        *out = &it->second;
        return 0;
    }
    const int blk = sbk::kBlock / q * tpt;   // targets per work-group
    const long long capacity =
        (long long)sbk::coulomb_blocks_per_cu(precision, tpt, e->unit_charges ? 1 : 0, q,
                                              local_coords(e, precision), kern) *
        e->cus;
    std::vector<sbk::PlanCloud> pc(e->clouds.size());
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) {
        pc[ci].n = e->clouds[ci].n;
        pc[ci].off = e->clouds[ci].off;
        pc[ci].on = !only_coulomb ||
                    (e->clouds[ci].f.coulomb && (mask_id == 0 || e->list_mask[ci]));
    }
    // the streamed kernel runs a block's filled sets of 64 targets only, so
    // mid-size clouds may take half-size blocks (sbk_plan.h)
    const int split_blk = (kern == sbk::KK_STREAM && tpt == 4) ? blk / 2 : 0;
    sbk::Plan plan = sbk::plan_coulomb(pc, blk, sbk::kSubTile * q, sreq, capacity, exact, split_blk);
    // a list narrowed to the clouds still active keeps the requested slice
    // count, chosen for all clouds; when its items would not fill one round
    // of resident work-groups, the automatic split serves it better
    if (mask_id != 0 && !exact && sreq > 0 && (long long)plan.items.size() < capacity)
        plan = sbk::plan_coulomb(pc, blk, sbk::kSubTile * q, 0, capacity, exact, split_blk);
    const std::string bad = sbk::check_plan(pc, plan);
    if (!bad.empty()) return fail(SBK_ESTATE, "Coulomb work list is inconsistent: %s", bad.c_str());
    ItemList L;
    std::vector<sbk::CoulombItem> &items = plan.items;
    L.slices = plan.slices;
    L.max_slices = plan.max_slices;
    L.nitems = (int)items.size();
    if (L.nitems > 0) {
        HIP_TRY(hipMalloc(&L.d_items, sizeof(sbk::CoulombItem) * items.size()));
        HIP_TRY(hipMemcpy(L.d_items, items.data(), sizeof(sbk::CoulombItem) * items.size(),
                          hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&L.d_slices, sizeof(int) * std::max<size_t>(1, L.slices.size())));
    HIP_TRY(hipMemcpy(L.d_slices, L.slices.data(), sizeof(int) * L.slices.size(),
                      hipMemcpyHostToDevice));
    if (L.max_slices > e->cp_slices) {
        destroy_graph(e);
        HIP_TRY(hipStreamSynchronize(e->stream));   // launches in flight may still write d_cp
        if (e->d_cp) HIP_TRY(hipFree(e->d_cp));
        e->d_cp = nullptr;
        HIP_TRY(hipMalloc(&e->d_cp, sizeof(double) * 3 * e->stride * L.max_slices));
        e->cp_slices = L.max_slices;
        e->dev.CP = e->d_cp;
    }
    auto res = e->lists.emplace(key, L);
    *out = &res.first->second;
    return 0;
}

// The partial-sum buffer holds cp_slices planes; a list whose items write
// more planes must never be launched (get_list grows the buffer first).
int check_partials(const sbk_engine *e, const ItemList *L)
{
    if (L->max_slices > e->cp_slices || (L->nitems > 0 && !e->d_cp))
        return fail(SBK_ESTATE, "partial-sum buffer holds %d slices, the work list writes %d",
                    e->cp_slices, L->max_slices);
    return 0;
}

sbk::CloudParams cloud_params(const sbk_engine *e, size_t ci)
{
    const HostCloud &c = e->clouds[ci];
    sbk::CloudParams p = make_params(c.f, c.off, c.n);
    p.blk0 = e->blk0[ci];
    p.nblk = e->nblk[ci];
    // Reusing an accepted attempt's stage-7 derivative as the next stage 1 is
    // exact when stage 7 sees the new step's state and time: y7 is bit-equal
    // to y_{n+1}, and with stage times t + c7 h = t + h = t_{n+1}.  With the
    // reference's frozen time it is exact only if no term depends on time; the
    // trap ramp is evaluated at the step's start time, so it never is.
    const bool timed = (c.f.trap && c.f.Vrf != 0.0) || c.f.Ed != 0.0 || c.f.Uq != 0.0 ||
                       c.f.Uo != 0.0 || c.f.Uw != 0.0 || c.f.Eax != 0.0 || c.f.swift_n > 0;
    if (c.f.swift_n > 0 && e->d_swift) p.swift = e->d_swift + e->swift_off[ci];
    if (c.f.emap_nr > 0 && e->d_swift) p.emap = e->d_swift + e->swift_off[ci] + c.swift.size();
    p.fsal = !c.f.ramp && (e->cfg.time_mode == SBK_TIME_STAGE || !timed);
    p.xcross_on = c.xcross_on;
    p.xcross = c.xcross;
    p.ion0 = c.off;
    for (size_t i = 0; i < c.perm.size(); ++i)
        if (c.perm[i] == 0) p.ion0 = c.off + (int)i;
    return p;
}

// Entries of the crossing list: one per ion of the engine, at least 65,536
// and at most 2^24 (72 B each).
int crossing_capacity(long long ntot)
{
    return (int)std::min<long long>(1LL << 24, std::max<long long>(65536, ntot));
}

// Allocates the crossing list once a cloud records crossings.  The attempt
// graph holds the device descriptor by value, so it is rebuilt.  The stream
// must be idle.
int ensure_crossings(sbk_engine *e)
{
    if (e->d_xc) return 0;
    bool any = false;
    for (auto &c : e->clouds) any = any || c.xcross_on;
    if (!any) return 0;
    e->xc_cap = crossing_capacity(e->ntot);
    HIP_TRY(hipMalloc(&e->d_xc, sizeof(sbk::XCrossing) * (size_t)e->xc_cap));
    HIP_TRY(hipMalloc(&e->d_xc_count, sizeof(int)));
    HIP_TRY(hipMemsetAsync(e->d_xc_count, 0, sizeof(int), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->dev.xc = e->d_xc;
    e->dev.xc_count = e->d_xc_count;
    e->dev.xc_cap = e->xc_cap;
    destroy_graph(e);
    return 0;
}

// Moves the device list's entries to the host, with the caller indices of
// the current ion order, and empties it.  The stream must be idle.
int drain_crossings(sbk_engine *e)
{
    if (!e->d_xc) return 0;
    int count = 0;
    HIP_TRY(hipMemcpyAsync(&count, e->d_xc_count, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const int kept = std::min(count, e->xc_cap);
    if (kept > 0) {
        std::vector<sbk::XCrossing> v(kept);
        HIP_TRY(hipMemcpyAsync(v.data(), e->d_xc, sizeof(sbk::XCrossing) * (size_t)kept,
                               hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (const sbk::XCrossing &x : v) {
            const HostCloud &c = e->clouds[x.cloud];
            HostCrossing h;
            h.cloud = x.cloud;
            h.ion = caller_index(c, x.slot - c.off);
            h.step = x.step;
            h.t = x.t;
            for (int k = 0; k < 3; ++k) {
                h.p[k] = x.p[k];
                h.v[k] = x.v[k];
            }
            e->xc_host.push_back(h);
        }
    }
    HIP_TRY(hipMemsetAsync(e->d_xc_count, 0, sizeof(int), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (count > e->xc_cap)
        return fail(SBK_ESTATE, "the crossing list overflowed (%d crossings, room for %d): take them "
                                "more often", count, e->xc_cap);
    return 0;
}

// Concatenates every cloud's SWIFT samples, each followed by its field map,
// into one device buffer (grown as needed; the stream must be idle, since
// the old buffer may be freed).
int upload_swift(sbk_engine *e)
{
    std::vector<double> all;
    e->swift_off.assign(e->clouds.size(), 0);
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) {
        e->swift_off[ci] = all.size();
        all.insert(all.end(), e->clouds[ci].swift.begin(), e->clouds[ci].swift.end());
        all.insert(all.end(), e->clouds[ci].emap.begin(), e->clouds[ci].emap.end());
    }
    if (all.empty()) return 0;
    if (all.size() > e->swift_cap) {
        if (e->d_swift) HIP_TRY(hipFree(e->d_swift));
        e->d_swift = nullptr;
        HIP_TRY(hipMalloc(&e->d_swift, sizeof(double) * all.size()));
        e->swift_cap = all.size();
    }
    HIP_TRY(hipMemcpy(e->d_swift, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice));
    return 0;
}

int upload_params(sbk_engine *e)
{
    std::vector<sbk::CloudParams> ps;
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) ps.push_back(cloud_params(e, ci));
    HIP_TRY(hipMemcpyAsync(e->d_params, ps.data(), sizeof(sbk::CloudParams) * ps.size(),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));   // ps is pageable and goes out of scope
    return 0;
}

// per-ion constants of one cloud: qm = charge*e/mass (bin/force.cpp:71),
// kc = coulomb_scale*ke*e*charge*e/mass (bin/force.cpp:153)
void ion_constants(const HostCloud &c, double *qm, double *kc, double *qd, float *qf)
{
    for (int i = 0; i < c.n; ++i) {
        qm[i] = c.charge[i] * kElCharge / c.mass[i];
        kc[i] = c.f.coulomb_scale * kKe * kElCharge * c.charge[i] * kElCharge / c.mass[i];
        qd[i] = (double)c.charge[i];
        qf[i] = (float)c.charge[i];
    }
}

// index of the FP64 SoA arrays inside d_f64
enum {
    A_P0 = 0, A_V0 = 3, A_PS = 6, A_KV = 9 /* KV2..KV7: 18 */, A_KA = 27 /* KA1..KA7: 21 */,
    A_DER = 48 /* 3 */, A_QM = 51, A_KC = 52, A_QD = 53, A_COUNT = 54
};

int finalize(sbk_engine *e)
{
    if (e->finalized) return 0;
    if (e->clouds.empty()) return fail(SBK_ESTATE, "engine has no clouds");
    e->ntot = 0;
    for (auto &c : e->clouds) {
        c.off = e->ntot;
        e->ntot += c.n;
    }
    e->stride = std::max<long long>(64, ((long long)e->ntot + 63) / 64 * 64);
    const long long S = e->stride;
    e->spatial = e->cfg.spatial_order &&
                 (e->cfg.precision == SBK_MIXED || e->cfg.precision == SBK_MIXED_KAHAN);
    if (e->spatial) {
        for (auto &c : e->clouds) {
            const std::vector<int> ord = spatial_order(c.pos.data(), c.n);
            apply_order(c.pos, ord, 3);
            apply_order(c.vel, ord, 3);
            apply_order(c.mass, ord, 1);
            apply_order(c.charge, ord, 1);
            c.perm = ord;
        }
        HIP_TRY(hipMalloc(&e->d_src, sizeof(int) * S));
    }
    HIP_TRY(hipMalloc(&e->d_f64, sizeof(double) * S * A_COUNT));
    HIP_TRY(hipMemsetAsync(e->d_f64, 0, sizeof(double) * S * A_COUNT, e->stream));
    HIP_TRY(hipMalloc(&e->d_qf, sizeof(float) * S));
    // ion-block table: at most 256 ions of one cloud per block
    std::vector<sbk::IonBlock> blocks;
    e->blk0.assign(e->clouds.size(), 0);
    e->nblk.assign(e->clouds.size(), 0);
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) {
        const HostCloud &c = e->clouds[ci];
        e->blk0[ci] = (int)blocks.size();
        for (int b = 0; b < c.n; b += 256)
            blocks.push_back(sbk::IonBlock{(int)ci, c.off + b, c.off + std::min(c.n, b + 256), 0});
        e->nblk[ci] = (int)blocks.size() - e->blk0[ci];
    }
    HIP_TRY(hipMalloc(&e->d_blocks, sizeof(sbk::IonBlock) * std::max<size_t>(1, blocks.size())));
    HIP_TRY(hipMalloc(&e->d_errb, sizeof(double) * std::max<size_t>(1, blocks.size())));
    if (!blocks.empty())
        HIP_TRY(hipMemcpyAsync(e->d_blocks, blocks.data(), sizeof(sbk::IonBlock) * blocks.size(),
                               hipMemcpyHostToDevice, e->stream));
    e->dev.nblocks = (int)blocks.size();
    HIP_TRY(hipMalloc(&e->d_params, sizeof(sbk::CloudParams) * e->clouds.size()));
    HIP_TRY(hipMalloc(&e->d_state, sizeof(sbk::CloudState) * e->clouds.size()));
    HIP_TRY(hipMalloc(&e->d_bar, sizeof(unsigned) * 2));
    HIP_TRY(hipMalloc(&e->d_timeout, sizeof(int)));
    HIP_TRY(hipMemsetAsync(e->d_bar, 0, sizeof(unsigned) * 2, e->stream));
    HIP_TRY(hipMemsetAsync(e->d_timeout, 0, sizeof(int), e->stream));
    HIP_TRY(hipHostMalloc(&e->h_state, sizeof(sbk::CloudState) * e->clouds.size(),
                          hipHostMallocDefault));

    // host staging in SoA, then one copy per array
    std::vector<double> soa(S * 6, 0.0), cst(S * 3, 0.0);
    std::vector<float> qf(S, 0.f);
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) {
        HostCloud &c = e->clouds[ci];
        for (int i = 0; i < c.n; ++i) {
            for (int k = 0; k < 3; ++k) {
                soa[k * S + c.off + i] = c.pos[3 * i + k];
                soa[(3 + k) * S + c.off + i] = c.vel[3 * i + k];
            }
        }
        ion_constants(c, &cst[c.off], &cst[S + c.off], &cst[2 * S + c.off], &qf[c.off]);
    }
    double *base = e->d_f64;
    HIP_TRY(hipMemcpyAsync(base + A_P0 * S, soa.data(), sizeof(double) * S * 6,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(base + A_QM * S, cst.data(), sizeof(double) * S * 3,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->d_qf, qf.data(), sizeof(float) * S, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));   // pageable staging goes out of scope
    if (int rc = upload_swift(e)) return rc;
    std::vector<sbk::CloudState> st(e->clouds.size());
    for (size_t ci = 0; ci < e->clouds.size(); ++ci) {
        sbk::CloudState s{};
        s.t = e->clouds[ci].lifetime;
        s.h = e->cfg.h0;
        s.hnext = e->cfg.h0;       // _ode_vars::Init_ode, bin/ode.cpp:539-546
        s.errold = 1.0e-4;
        s.t_end = s.t;
        s.t_prev = s.t;
        s.budget = -1;
        st[ci] = s;
    }
    HIP_TRY(hipMemcpyAsync(e->d_state, st.data(), sizeof(sbk::CloudState) * st.size(),
                           hipMemcpyHostToDevice, e->stream));
    int rc = upload_params(e);
    if (rc) return rc;

    sbk::EngineDev &d = e->dev;
    for (int k = 0; k < 3; ++k) {
        d.P0[k] = base + (A_P0 + k) * S;
        d.V0[k] = base + (A_V0 + k) * S;
        d.PS[k] = base + (A_PS + k) * S;
        d.KV[0][k] = nullptr;
        d.KA[0][k] = nullptr;
        d.KV[1][k] = d.V0[k];
        for (int s = 2; s <= 7; ++s) d.KV[s][k] = base + (A_KV + (s - 2) * 3 + k) * S;
        for (int s = 1; s <= 7; ++s) d.KA[s][k] = base + (A_KA + (s - 1) * 3 + k) * S;
    }
    d.ERRB = e->d_errb;
    d.derivs_out = base + A_DER * S;
    d.stride = S;
    d.qm = base + A_QM * S;
    e->d_kc = base + A_KC * S;
    e->d_qd = base + A_QD * S;
    d.blocks = e->d_blocks;
    d.params = e->d_params;
    d.state = e->d_state;
    d.ntot = e->ntot;
    d.nclouds = (int)e->clouds.size();
    d.time_mode = e->cfg.time_mode;
    d.adaptive = e->cfg.adaptive;
    d.rtol = e->cfg.rtol;
    d.atol = e->cfg.atol;
    d.beta = e->cfg.beta;
    if (int rc2 = ensure_crossings(e)) return rc2;
    d.CP = nullptr;
    d.slices = nullptr;
    e->unit_charges = true;
    for (auto &c : e->clouds)
        for (int q : c.charge) e->unit_charges = e->unit_charges && q == 1;
    e->attempts_at_reset.assign(e->clouds.size(), 0);
    e->steps_at_reset.assign(e->clouds.size(), 0);
    e->evals_at_reset.assign(e->clouds.size(), 0);
    e->finalized = true;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

sbk::CoulombArgs coulomb_args(sbk_engine *e, const ItemList &L, double *const *pos, bool gated)
{
    sbk::CoulombArgs a{};
    a.tx = pos[0];
    a.ty = pos[1];
    a.tz = pos[2];
    a.sx = pos[0];
    a.sy = pos[1];
    a.sz = pos[2];
    a.sqd = e->d_qd;
    a.sqf = e->d_qf;
    a.kc = e->d_kc;
    a.out = e->d_cp;
    a.stride = e->stride;
    a.items = L.d_items;
    a.nitems = L.nitems;
    a.self_skip = 1;
    a.unit_weights = e->unit_charges ? 1 : 0;
    a.state = gated ? e->d_state : nullptr;
    a.eps_f = 1e-23f;   // (float) softening of bin/cunbody.cpp:11; keeps the self pair finite
    a.eps_d = 0.0;      // FP64-exact: no softening (bin/force.cpp:163)
    return a;
}

int launch_coulomb_timed(sbk_engine *e, int precision, const sbk::CoulombArgs &a)
{
    std::pair<hipEvent_t, hipEvent_t> *evp = nullptr;
    if (e->cfg.timing) {
        if (e->ev_used == e->ev.size()) {
            hipEvent_t a0, a1;
            HIP_TRY(hipEventCreate(&a0));
            HIP_TRY(hipEventCreate(&a1));
            e->ev.emplace_back(a0, a1);
        }
        evp = &e->ev[e->ev_used++];
        HIP_TRY(hipEventRecord(evp->first, e->stream));
    }
    sbk::CoulombArgs b = a;
    if (precision == SBK_FP64 || precision == SBK_FP64_RSQ)
        b.eps_d = 1e-40;   // the guard of bin/force.cpp:136
    b.local_coords = local_coords(e, precision);
    if (sbk::launch_coulomb(kp(precision), targets_per_thread(e, precision), e->groups,
                            coulomb_kernel(e, precision), b, e->stream))
        return fail(SBK_EHIP, "Coulomb kernel launch failed: %s",
                    hipGetErrorString(hipGetLastError()));
    if (evp) HIP_TRY(hipEventRecord(evp->second, e->stream));
    return 0;
}

int harvest_events(sbk_engine *e)
{
    for (size_t i = 0; i < e->ev_used; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e->ev[i].first, e->ev[i].second));
        e->coulomb_ms += ms;
        e->coulomb_launches++;
    }
    e->ev_used = 0;
    return 0;
}

// Enqueues one attempt for all active clouds.
int enqueue_attempt(sbk_engine *e)
{
    ItemList *L = nullptr;
    bool any_coulomb = false;
    for (auto &c : e->clouds) any_coulomb |= (c.f.coulomb && c.n > 0);
    int rc = get_list(e, e->cfg.precision, true, &L);
    if (rc) return rc;
    if ((rc = check_partials(e, L))) return rc;
    sbk::EngineDev d = e->dev;
    d.slices = L->d_slices;
    for (int s = 1; s <= 7; ++s) {
        if (any_coulomb && L->nitems > 0) {
            sbk::CoulombArgs a = coulomb_args(e, *L, s == 1 ? d.P0 : d.PS, true);
            a.skip_valid_k1 = s == 1;
            rc = launch_coulomb_timed(e, e->cfg.precision, a);
            if (rc) return rc;
        }
        if (sbk::launch_stage(s, d, e->stream)) return fail(SBK_EHIP, "stage kernel launch failed");
    }
    if (sbk::launch_controller(d, e->stream)) return fail(SBK_EHIP, "controller launch failed");
    if (sbk::launch_commit(d, e->stream)) return fail(SBK_EHIP, "commit launch failed");
    return 0;
}

// Grid of the persistent attempt kernel for this engine (0: none): enough
// work-groups for the largest phase, at most what is resident at once.
int persistent_grid(sbk_engine *e, const ItemList &L)
{
    const int precision = e->cfg.precision;
    const int tpt = precision == SBK_FP64_EXACT ? 1 : targets_per_thread(e, precision);
    if (precision != SBK_FP64_EXACT && e->groups != 4) return 0;
    const int per_cu = sbk::persistent_blocks_per_cu(kp(precision), tpt, e->unit_charges ? 1 : 0,
                                                     local_coords(e, precision));
    if (per_cu <= 0) return 0;
    const long long cap = (long long)per_cu * e->cus;
    const long long want = std::max<long long>({(long long)L.nitems, (long long)e->dev.nblocks,
                                                (long long)e->clouds.size(), 1LL});
    return (int)std::min(cap, want);
}

// Whether the next attempts run as one persistent launch (persistent = 1;
// a shape without such a kernel is an error, not a silent fallback).
int use_persistent(sbk_engine *e, bool *use)
{
    *use = false;
    if (e->cfg.persistent != 1 || e->cfg.timing) return 0;
    const int precision = e->cfg.precision;
    const int tpt = precision == SBK_FP64_EXACT ? 1 : targets_per_thread(e, precision);
    bool ok = e->coop != 0;
    if (precision != SBK_FP64_EXACT && (e->groups != 4 || (tpt != 1 && tpt != 2 && tpt != 4)))
        ok = false;
    if (ok)
        ok = sbk::persistent_blocks_per_cu(kp(precision), tpt, e->unit_charges ? 1 : 0,
                                           local_coords(e, precision)) > 0;
    if (!ok)
        return fail(SBK_EINVAL,
                    "persistent = 1: no persistent attempt kernel for this shape (needs "
                    "cooperative launch, 4 source groups and 1, 2 or 4 targets per thread; "
                    "have %d groups, %d targets per thread)", e->groups, tpt);
    *use = ok;
    return 0;
}

// `count` attempts as one cooperative launch (see attempts_kernel).
int run_persistent(sbk_engine *e, int count)
{
    ItemList *L = nullptr;
    int rc = get_list(e, e->cfg.precision, true, &L);
    if (rc) return rc;
    if ((rc = check_partials(e, L))) return rc;
    bool any_coulomb = false;
    for (auto &c : e->clouds) any_coulomb |= (c.f.coulomb && c.n > 0);
    const int precision = e->cfg.precision;
    sbk::PersistArgs p{};
    p.d = e->dev;
    p.d.slices = L->d_slices;
    p.a = coulomb_args(e, *L, e->dev.P0, true);
    p.a.local_coords = local_coords(e, precision);
    if (precision == SBK_FP64 || precision == SBK_FP64_RSQ) p.a.eps_d = 1e-40;   // bin/force.cpp:136
    p.bar = e->d_bar;
    p.timeout = e->d_timeout;
    p.spin_limit = e->cfg.persistent_spin > 0 ? (unsigned)e->cfg.persistent_spin : 1u << 22;
    p.attempts = count;
    p.any_coulomb = any_coulomb && L->nitems > 0;
    const int grid = persistent_grid(e, *L);
    if (grid <= 0) return fail(SBK_ESTATE, "no persistent attempt kernel for this launch shape");
    HIP_TRY(hipMemsetAsync(e->d_bar, 0, sizeof(unsigned) * 2, e->stream));
    const int tpt = precision == SBK_FP64_EXACT ? 1 : targets_per_thread(e, precision);
    const hipError_t err =
        (hipError_t)sbk::launch_attempts_persistent(kp(precision), tpt, p, grid, e->stream);
    if (err != hipSuccess) {
        (void)hipGetLastError();   // reported here
        // a refused launch enqueued nothing; cooperative queues are a limited
        // resource of the device, which another process may hold
        return fail(SBK_EHIP, "persistent attempt kernel launch failed (grid %d): %s", grid,
                    hipGetErrorString(err));
    }
    e->pers_launched = true;
    e->pers_count++;
    return 0;
}

int run_attempts(sbk_engine *e, int count)
{
    if (count <= 0) return 0;
    bool pers = false;
    if (int rc = use_persistent(e, &pers)) return rc;
    if (pers) return run_persistent(e, count);
    if (e->cfg.use_graph && !e->cfg.timing) {
        if (!e->graph_exec) {
            // build the list before capture: allocation is not allowed while capturing
            ItemList *L = nullptr;
            int rc = get_list(e, e->cfg.precision, true, &L);
            if (rc) return rc;
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_attempt(e);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ce != hipSuccess) return fail(SBK_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
            e->graph = g;
            HIP_TRY(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < count; ++i) HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
        return 0;
    }
    for (int i = 0; i < count; ++i) {
        int rc = enqueue_attempt(e);
        if (rc) return rc;
    }
    return 0;
}

int read_states(sbk_engine *e)
{
    HIP_TRY(hipMemcpyAsync(e->h_state, e->d_state, sizeof(sbk::CloudState) * e->clouds.size(),
                           hipMemcpyDeviceToHost, e->stream));
    int timeout = 0;
    if (e->pers_launched)
        HIP_TRY(hipMemcpyAsync(&timeout, e->d_timeout, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->pers_launched = false;
    if (timeout) {
        (void)hipMemsetAsync(e->d_timeout, 0, sizeof(int), e->stream);
        (void)hipStreamSynchronize(e->stream);
        return fail(SBK_EHIP, "a grid barrier of the persistent attempt kernel timed out; the "
                              "state is undefined (use persistent = -1)");
    }
    return harvest_events(e);
}

// Re-sorts every cloud into a spatial order of its current positions and
// permutes the per-ion arrays that live from one attempt to the next
// (positions, velocities, the stage-1 derivative and the ion constants).
// The stage arrays are dead between attempts and serve as the gather buffer.
int reorder(sbk_engine *e)
{
    if (!e->spatial) return 0;
    // recorded crossings name slots of the order that is about to change
    if (int rc = drain_crossings(e)) return rc;
    const long long S = e->stride;
    const int n = e->ntot;
    std::vector<double> soa(3 * (size_t)S);
    for (int k = 0; k < 3; ++k)
        HIP_TRY(hipMemcpyAsync(soa.data() + k * S, e->dev.P0[k], sizeof(double) * n,
                               hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<int> src(n);
    // each cloud's order depends on its own positions only, and it writes its
    // own range of src: the clouds of an ensemble are shared among host
    // threads in contiguous runs of about equal bisection work (n log n)
    auto sort_clouds = [&](size_t c0, size_t c1) {
        std::vector<double> aos;
        for (size_t ci = c0; ci < c1; ++ci) {
            HostCloud &c = e->clouds[ci];
            aos.resize(3 * (size_t)c.n);
            for (int i = 0; i < c.n; ++i)
                for (int k = 0; k < 3; ++k) aos[3 * (size_t)i + k] = soa[k * S + c.off + i];
            const std::vector<int> ord = spatial_order(aos.data(), c.n);
            std::vector<int> perm(c.n);
            for (int i = 0; i < c.n; ++i) {
                src[c.off + i] = c.off + ord[i];
                perm[i] = caller_index(c, ord[i]);
            }
            apply_order(c.mass, ord, 1);
            apply_order(c.charge, ord, 1);
            c.perm.swap(perm);
        }
    };
    const size_t nc = e->clouds.size();
    size_t nt = std::min<size_t>({(size_t)std::max(1u, std::thread::hardware_concurrency()),
                                  (size_t)16, nc});
    if (n < 16384) nt = 1;   // thread start-up would cost more than the sort
    if (nt <= 1) {
        sort_clouds(0, nc);
    } else {
        std::vector<double> work(nc + 1, 0.0);
        for (size_t ci = 0; ci < nc; ++ci) {
            const double m = (double)std::max(1, e->clouds[ci].n);
            work[ci + 1] = work[ci] + m * std::log2(m + 1.0);
        }
        std::vector<std::thread> pool;
        size_t c0 = 0;
        for (size_t t = 1; t <= nt && c0 < nc; ++t) {
            size_t c1 = c0;
            while (c1 < nc && (t == nt || work[c1 + 1] * nt <= work[nc] * t)) ++c1;
            if (c1 == c0) c1 = c0 + 1;
            pool.emplace_back(sort_clouds, c0, c1);
            c0 = c1;
        }
        for (auto &th : pool) th.join();
    }
    HIP_TRY(hipMemcpyAsync(e->d_src, src.data(), sizeof(int) * n, hipMemcpyHostToDevice, e->stream));
    double *base = e->d_f64;
    double *tmp = base + A_PS * S;
    double *arrays[12] = {e->dev.P0[0],    e->dev.P0[1],    e->dev.P0[2],    e->dev.V0[0],
                          e->dev.V0[1],    e->dev.V0[2],    e->dev.KA[1][0], e->dev.KA[1][1],
                          e->dev.KA[1][2], base + A_QM * S, base + A_KC * S, base + A_QD * S};
    for (double *a : arrays) {
        if (sbk::launch_gather(e->d_src, a, tmp, n, e->stream))
            return fail(SBK_EHIP, "gather kernel launch failed");
        HIP_TRY(hipMemcpyAsync(a, tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, e->stream));
    }
    float *tmpf = reinterpret_cast<float *>(tmp);
    if (sbk::launch_gather_f32(e->d_src, e->d_qf, tmpf, n, e->stream))
        return fail(SBK_EHIP, "gather kernel launch failed");
    HIP_TRY(hipMemcpyAsync(e->d_qf, tmpf, sizeof(float) * n, hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));   // src goes out of scope
    e->reorders++;
    return upload_params(e);   // the slot of each cloud's ion 0 moved
}

// Attempts run so far as the re-sort interval counts them: one per launch of
// the attempt sequence, which is the most attempts any cloud has made (a sum
// over the clouds would re-sort an ensemble of C clouds C times as often).
long long attempts_run(const sbk_engine *e)
{
    long long att = 0;
    for (size_t i = 0; i < e->clouds.size(); ++i) att = std::max(att, e->h_state[i].attempts);
    return att;
}

// Re-sorts when reorder_interval attempts have run since the last sort; the
// per-cloud states must be fresh (read_states) and the stream idle.
int maybe_reorder(sbk_engine *e)
{
    if (!e->spatial || e->cfg.reorder_interval <= 0) return 0;
    const long long att = attempts_run(e);
    if (att - e->attempts_at_order < e->cfg.reorder_interval) return 0;
    e->attempts_at_order = att;
    return reorder(e);
}

int prepare(sbk_engine *e)
{
    if (!e) return fail(SBK_EINVAL, "null engine");
    HIP_TRY(hipSetDevice(e->device));
    return finalize(e);
}

}  // namespace

extern "C" {

const char *sbk_last_error(void)
{
    return g_err.c_str();
}

void sbk_default_config(sbk_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->precision = SBK_MIXED;
    c->time_mode = SBK_TIME_FROZEN;
    c->adaptive = 1;
    c->use_graph = 1;
    c->timing = 0;
    c->rtol = 1e-7;   // bin/ode.cpp:515-516
    c->atol = 1e-7;
    c->h0 = 1e-9;     // bin/ode.cpp:509
    c->spatial_order = 1;
    c->reorder_interval = 256;
}

void sbk_default_fields(sbk_fields *f)
{
    if (!f) return;
    std::memset(f, 0, sizeof(*f));
    // bin/trapparameters.cpp:2-10 and bin/ode.cpp:646-653
    f->trap = 1;
    f->kappa = 0.001;
    f->Vdc = 5;
    f->Vrf = 50;
    f->r0 = 0.04;
    f->z0 = 0.04;
    f->freq_rf = 50e5;
    f->coulomb = 0;
    f->coulomb_scale = 1.0;
    f->rq = 1.0;
    f->ro = 1.0;
    f->rw = 1.0;
    f->rw_order = 2;
    f->T_gas = 273.15;
}

int sbk_spatial_order(int n, const double *pos, int *order)
{
    if (n < 0 || (n > 0 && (!pos || !order))) return fail(SBK_EINVAL, "bad arguments (n=%d)", n);
    const std::vector<int> ord = spatial_order(pos, n);
    std::copy(ord.begin(), ord.end(), order);
    return 0;
}

int sbk_device_info(char *name, int name_len, int *cus, int *clock_mhz)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s", p.gcnArchName[0] ? p.gcnArchName : p.name);
    }
    if (cus) *cus = p.multiProcessorCount;
    if (clock_mhz) *clock_mhz = p.clockRate / 1000;
    return 0;
}

int sbk_create(const sbk_config *cfg, sbk_engine **out)
{
    if (!cfg || !out) return fail(SBK_EINVAL, "null argument");
    if (cfg->precision < SBK_FP64_EXACT || cfg->precision > SBK_FP64_RSQ)
        return fail(SBK_EINVAL, "unknown precision %d", cfg->precision);
    if (cfg->persistent < 0 || cfg->persistent > 1)
        return fail(SBK_EINVAL, "persistent must be 0 or 1 (got %d)", cfg->persistent);
    if (cfg->persistent_spin < 0)
        return fail(SBK_EINVAL, "persistent_spin must be >= 0 (got %d)", cfg->persistent_spin);
    if (cfg->time_mode != SBK_TIME_FROZEN && cfg->time_mode != SBK_TIME_STAGE)
        return fail(SBK_EINVAL, "unknown time mode %d", cfg->time_mode);
    if (!(cfg->h0 > 0.0) || !(cfg->rtol >= 0.0) || !(cfg->atol >= 0.0) ||
        (cfg->rtol == 0.0 && cfg->atol == 0.0))
        return fail(SBK_EINVAL, "h0 must be > 0 and rtol/atol >= 0, not both 0");
    if (!(cfg->beta >= 0.0 && cfg->beta < 0.2))
        return fail(SBK_EINVAL, "beta must be in [0, 0.2) (got %g)", cfg->beta);
    std::unique_ptr<sbk_engine> e(new sbk_engine);
    e->cfg = *cfg;
    HIP_TRY(hipGetDevice(&e->device));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, e->device));
    e->cus = std::max(1, p.multiProcessorCount);
    e->coop = p.cooperativeLaunch;
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    *out = e.release();
    return 0;
}

int sbk_destroy(sbk_engine *e)
{
    if (!e) return 0;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_lists(e);
    for (auto &p : e->ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    (void)hipFree(e->d_f64);
    (void)hipFree(e->d_qf);
    (void)hipFree(e->d_blocks);
    (void)hipFree(e->d_errb);
    (void)hipFree(e->d_cp);
    (void)hipFree(e->d_params);
    (void)hipFree(e->d_state);
    (void)hipFree(e->d_src);
    (void)hipFree(e->d_swift);
    (void)hipFree(e->d_xc);
    (void)hipFree(e->d_xc_count);
    (void)hipFree(e->d_bar);
    (void)hipFree(e->d_timeout);
    if (e->h_state) (void)hipHostFree(e->h_state);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

// Parameters the field terms divide by, or choose a form with.
static int check_fields(const sbk_fields *f)
{
    if (f->Uw != 0.0 && f->rw_order != 1 && f->rw_order != 2)
        return fail(SBK_EINVAL, "rw_order must be 1 or 2 (got %d)", f->rw_order);
    if ((f->Uw != 0.0 && !(f->rw > 0.0)) || (f->Uo != 0.0 && !(f->ro > 0.0)) ||
        (f->Uq != 0.0 && !(f->rq > 0.0)) || (f->V0 != 0.0 && !(f->d0 > 0.0)))
        return fail(SBK_EINVAL, "electrode radii rq, ro, rw and d0 must be > 0 when used");
    if (f->K0 != 0.0 && f->p_mbar != 0.0 && !(f->T_gas > 0.0))
        return fail(SBK_EINVAL, "T_gas must be > 0 for the buffer gas");
    if (f->swift_n < 0 || (f->swift_n > 0 && (!f->swift || f->swift_n < 2 || !(f->swift_dt > 0.0) ||
                                             !(f->swift_r > 0.0) ||
                                             (f->swift_kind != 1 && f->swift_kind != 2))))
        return fail(SBK_EINVAL, "SWIFT needs >= 2 samples, swift_dt > 0, swift_r > 0 and "
                                "swift_kind 1 or 2");
    if (f->emap_nr < 0 || (f->emap_nr > 0 && (!f->emap || f->emap_nr < 2 || f->emap_nz < 2 ||
                                              !(f->emap_dr > 0.0) || !(f->emap_dz > 0.0) ||
                                              !std::isfinite(f->emap_z0) ||
                                              (long long)f->emap_nr * f->emap_nz > (1LL << 26))))
        return fail(SBK_EINVAL, "a field map needs 2 <= emap_nr, 2 <= emap_nz, "
                                "emap_nr * emap_nz <= 2^26, emap_dr > 0, emap_dz > 0");
    if ((f->C4 != 0.0 || f->C6 != 0.0) && f->V0 == 0.0)
        return fail(SBK_EINVAL, "C4 and C6 scale the Penning potential: set V0 and d0");
    return 0;
}

int sbk_add_cloud(sbk_engine *e, int n, const double *pos, const double *vel, const double *mass,
                  const int *charge, const sbk_fields *f, double lifetime, int *cloud_id)
{
    if (!e || !f) return fail(SBK_EINVAL, "null argument");
    if (int rc = check_fields(f)) return rc;
    if (e->finalized) return fail(SBK_ESTATE, "clouds must be added before the first run");
    if (n < 0 || (n > 0 && (!pos || !vel || !mass || !charge)))
        return fail(SBK_EINVAL, "bad cloud arrays (n=%d)", n);
    if (e->ntot + (long long)n > (1LL << 30)) return fail(SBK_EINVAL, "too many ions");
    HostCloud c;
    c.n = n;
    c.f = *f;
    c.lifetime = lifetime;
    c.pos.assign(pos, pos + 3 * (size_t)n);
    c.vel.assign(vel, vel + 3 * (size_t)n);
    c.mass.assign(mass, mass + n);
    c.charge.assign(charge, charge + n);
    for (int i = 0; i < n; ++i)
        if (!(c.mass[i] > 0.0)) return fail(SBK_EINVAL, "ion %d has mass <= 0", i);
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        if (!std::isfinite(c.pos[i]) || !std::isfinite(c.vel[i]))
            return fail(SBK_EINVAL, "ion %zu has a non-finite position or velocity", i / 3);
    if (f->swift_n > 0) c.swift.assign(f->swift, f->swift + f->swift_n);
    c.f.swift = nullptr;   // the samples live in c.swift from here on
    if (f->emap_nr > 0) c.emap.assign(f->emap, f->emap + 2 * (size_t)f->emap_nr * f->emap_nz);
    c.f.emap = nullptr;    // and the map in c.emap
    e->clouds.push_back(std::move(c));
    e->ntot += n;
    if (cloud_id) *cloud_id = (int)e->clouds.size() - 1;
    return 0;
}

int sbk_set_fields(sbk_engine *e, int cloud, const sbk_fields *f)
{
    if (!e || !f || cloud < 0 || cloud >= (int)e->clouds.size())
        return fail(SBK_EINVAL, "bad cloud id");
    if (int rc = check_fields(f)) return rc;
    HostCloud &c = e->clouds[cloud];
    const bool coulomb_changed = (c.f.coulomb != 0) != (f->coulomb != 0);
    const bool scale_changed = c.f.coulomb_scale != f->coulomb_scale;
    c.f = *f;
    c.swift.assign(f->swift, f->swift + (f->swift_n > 0 ? f->swift_n : 0));
    c.f.swift = nullptr;
    c.emap.assign(f->emap, f->emap + (f->emap_nr > 0 ? 2 * (size_t)f->emap_nr * f->emap_nz : 0));
    c.f.emap = nullptr;
    if (!e->finalized) return 0;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));   // the sample buffer may be reallocated
    if (int rc = upload_swift(e)) return rc;
    // every cloud's parameters: the samples of the others may have moved
    if (int rc = upload_params(e)) return rc;
    // the stored stage-1 derivative was computed with the old fields
    int rc = read_states(e);
    if (rc) return rc;
    sbk::CloudState st = e->h_state[cloud];
    st.k1_valid = 0;
    HIP_TRY(hipMemcpyAsync(e->d_state + cloud, &st, sizeof(st), hipMemcpyHostToDevice, e->stream));
    if (scale_changed && c.n > 0) {
        std::vector<double> qm(c.n), kc(c.n), qd(c.n);
        std::vector<float> qf(c.n);
        ion_constants(c, qm.data(), kc.data(), qd.data(), qf.data());
        HIP_TRY(hipMemcpyAsync(e->d_kc + c.off, kc.data(), sizeof(double) * c.n,
                               hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));   // kc goes out of scope
    }
    if (coulomb_changed) free_lists(e);
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int sbk_set_step(sbk_engine *e, int cloud, double h)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (cloud < 0 || cloud >= (int)e->clouds.size() || !(h > 0.0))
        return fail(SBK_EINVAL, "bad cloud id or step");
    rc = read_states(e);
    if (rc) return rc;
    sbk::CloudState s = e->h_state[cloud];
    s.h = h;
    s.hnext = h;
    HIP_TRY(hipMemcpyAsync(e->d_state + cloud, &s, sizeof(s), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 0;
}

int sbk_attempts(sbk_engine *e, int count)
{
    int rc = prepare(e);
    if (rc) return rc;
    return run_attempts(e, count);
}

int sbk_synchronize(sbk_engine *e)
{
    int rc = prepare(e);
    if (rc) return rc;
    return read_states(e);
}

static int check_errors(sbk_engine *e)
{
    for (size_t i = 0; i < e->clouds.size(); ++i)
        if (e->h_state[i].error)
            return fail(SBK_EUNDERFLOW, "stepsize underflow created (cloud %zu)", i);
    return 0;
}

int sbk_step(sbk_engine *e)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (sbk::launch_begin(e->dev, INFINITY, 1, 1, e->stream))
        return fail(SBK_EHIP, "begin kernel launch failed");
    for (long long guard = 0; guard < 100000; ++guard) {
        rc = run_attempts(e, 1);
        if (rc) return rc;
        rc = read_states(e);
        if (rc) return rc;
        rc = check_errors(e);
        if (rc) return rc;
        bool any = false;
        for (size_t i = 0; i < e->clouds.size(); ++i) any |= e->h_state[i].active != 0;
        if (!any) return 0;
    }
    return fail(SBK_ESTATE, "no accepted step after 100000 attempts");
}

// Attempts for the next batch of a run: as many as the active cloud that
// needs the most still wants at its current step size (to its end time, or
// its step budget), plus one for a rejection, within `cap`.  Launches after
// every cloud has finished do nothing, yet each costs its 16 kernels.
static int next_batch(const sbk_engine *e, int cap)
{
    double need = 0.0;
    for (size_t i = 0; i < e->clouds.size(); ++i) {
        const sbk::CloudState &s = e->h_state[i];
        if (!s.active) continue;
        double n = (std::isfinite(s.t_end) && s.h != 0.0) ? std::ceil((s.t_end - s.t) / std::fabs(s.h))
                                                          : (double)cap;
        if (s.budget > 0) n = std::min(n, (double)s.budget);
        need = std::max(need, n);
    }
    return (int)std::max(1.0, std::min((double)cap, need + 1.0));
}

static int advance_impl(sbk_engine *e, double value, int set_end, long long max_attempts)
{
    set_list_mask(e, {});   // every cloud starts the run
    e->list_mask_pairs = coulomb_pairs(e, false);
    if (sbk::launch_begin(e->dev, value, -1, set_end, e->stream))
        return fail(SBK_EHIP, "begin kernel launch failed");
    int rc = 0;
    long long launched = 0;
    int batch = 4, cap = 4;   // the cap doubles to 64 as the run goes on
    while (launched < max_attempts) {
        const int k = (int)std::min<long long>(batch, max_attempts - launched);
        rc = run_attempts(e, k);
        if (rc) break;
        launched += k;
        rc = read_states(e);
        if (rc) break;
        rc = check_errors(e);
        if (rc) break;
        bool any = false;
        for (size_t i = 0; i < e->clouds.size(); ++i) any |= e->h_state[i].active != 0;
        if (!any) break;
        rc = maybe_reorder(e);
        if (rc) break;
        narrow_list(e);
        cap = std::min(cap * 2, 64);
        batch = next_batch(e, cap);
    }
    // other callers (sbk_attempts, the evaluations) see the list of all clouds
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    set_list_mask(e, {});
    return rc;
}

int sbk_advance(sbk_engine *e, double duration, long long max_attempts)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (!(duration >= 0.0)) return fail(SBK_EINVAL, "duration must be >= 0");
    return advance_impl(e, duration, 1, max_attempts);
}

int sbk_advance_until(sbk_engine *e, double t_end, long long max_attempts)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (std::isnan(t_end)) return fail(SBK_EINVAL, "t_end is NaN");
    return advance_impl(e, t_end, 2, max_attempts);
}

int sbk_reorder(sbk_engine *e)
{
    int rc = prepare(e);
    if (rc) return rc;
    rc = read_states(e);
    if (rc) return rc;
    e->attempts_at_order = attempts_run(e);
    return reorder(e);
}

int sbk_get_state(sbk_engine *e, int cloud, double *pos, double *vel)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (cloud < 0 || cloud >= (int)e->clouds.size()) return fail(SBK_EINVAL, "bad cloud id");
    const HostCloud &c = e->clouds[cloud];
    std::vector<double> buf(6 * (size_t)c.n);
    for (int k = 0; k < 3; ++k) {
        HIP_TRY(hipMemcpyAsync(buf.data() + k * (size_t)c.n, e->dev.P0[k] + c.off,
                               sizeof(double) * c.n, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(buf.data() + (3 + k) * (size_t)c.n, e->dev.V0[k] + c.off,
                               sizeof(double) * c.n, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int i = 0; i < c.n; ++i) {
        const size_t o = 3 * (size_t)caller_index(c, i);
        for (int k = 0; k < 3; ++k) {
            if (pos) pos[o + k] = buf[k * (size_t)c.n + i];
            if (vel) vel[o + k] = buf[(3 + k) * (size_t)c.n + i];
        }
    }
    return 0;
}

int sbk_set_crossing_plane(sbk_engine *e, int cloud, int enable, double x_plane,
                           int *list_capacity)
{
    if (!e || cloud < 0 || cloud >= (int)e->clouds.size())
        return fail(SBK_EINVAL, "bad cloud id");
    if (enable && !std::isfinite(x_plane)) return fail(SBK_EINVAL, "the crossing plane must be finite");
    HostCloud &c = e->clouds[cloud];
    c.xcross_on = enable ? 1 : 0;
    c.xcross = enable ? x_plane : 0.0;
    if (e->finalized) {
        HIP_TRY(hipSetDevice(e->device));
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (int rc = ensure_crossings(e)) return rc;
        if (int rc = upload_params(e)) return rc;
    }
    if (list_capacity) {
        long long n = 0;
        for (auto &h : e->clouds) n += h.n;
        *list_capacity = e->d_xc ? e->xc_cap : crossing_capacity(n);
    }
    return 0;
}

int sbk_take_crossings(sbk_engine *e, int capacity, int *n, int *cloud, int *ion, long long *step,
                       double *t, double *pos, double *vel)
{
    if (!e || !n) return fail(SBK_EINVAL, "null argument");
    if (e->finalized && e->d_xc) {
        HIP_TRY(hipSetDevice(e->device));
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (int rc = drain_crossings(e)) return rc;
    }
    std::vector<HostCrossing> &v = e->xc_host;
    // the reference prints a step's crossings in ion order (bin/ion_fly.cpp:110-114)
    std::stable_sort(v.begin(), v.end(), [](const HostCrossing &a, const HostCrossing &b) {
        if (a.cloud != b.cloud) return a.cloud < b.cloud;
        if (a.step != b.step) return a.step < b.step;
        return a.ion < b.ion;
    });
    const int count = (int)v.size();
    if (capacity == 0 && !cloud && !ion && !step && !t && !pos && !vel) {
        *n = count;   // size query: the entries stay
        return 0;
    }
    if (capacity < count)
        return fail(SBK_EINVAL, "capacity %d is below the %d crossings held", capacity, count);
    for (int i = 0; i < count; ++i) {
        const HostCrossing &h = v[i];
        if (cloud) cloud[i] = h.cloud;
        if (ion) ion[i] = h.ion;
        if (step) step[i] = h.step;
        if (t) t[i] = h.t;
        for (int k = 0; k < 3; ++k) {
            if (pos) pos[3 * (size_t)i + k] = h.p[k];
            if (vel) vel[3 * (size_t)i + k] = h.v[k];
        }
    }
    *n = count;
    v.clear();
    return 0;
}

int sbk_get_status(sbk_engine *e, int cloud, sbk_cloud_status *st)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (!st || cloud < 0 || cloud >= (int)e->clouds.size()) return fail(SBK_EINVAL, "bad cloud id");
    rc = read_states(e);
    if (rc) return rc;
    const sbk::CloudState &s = e->h_state[cloud];
    st->lifetime = s.t;
    st->h = s.h;
    st->hnext = s.hnext;
    st->errold = s.errold;
    st->last_err = s.last_err;
    st->t_end = s.t_end;
    st->reject = s.reject;
    st->active = s.active;
    st->accepted_last = s.accepted_last;
    st->error = s.error;
    st->steps = s.steps;
    st->attempts = s.attempts;
    st->n = e->clouds[cloud].n;
    st->coulomb_evals = s.evals;
    st->t_prev = s.t_prev;
    ItemList *L = nullptr;
    st->slices = (get_list(e, e->cfg.precision, true, &L) == 0) ? L->slices[cloud] : 0;
    return 0;
}

int sbk_get_stats(sbk_engine *e, sbk_stats *out)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (!out) return fail(SBK_EINVAL, "null stats");
    rc = read_states(e);
    if (rc) return rc;
    std::memset(out, 0, sizeof(*out));
    out->coulomb_ms = e->coulomb_ms;
    out->coulomb_launches = e->coulomb_launches;
    out->reorders = e->reorders - e->reorders_at_reset;
    out->persistent_launches = e->pers_count;
    for (size_t i = 0; i < e->clouds.size(); ++i) {
        const double n = e->clouds[i].n;
        const long long att = e->h_state[i].attempts - e->attempts_at_reset[i];
        const long long stp = e->h_state[i].steps - e->steps_at_reset[i];
        out->attempts += att;
        out->steps += stp;
        out->ion_steps += (double)stp * n;
        const long long ev = e->h_state[i].evals - e->evals_at_reset[i];
        if (e->clouds[i].f.coulomb) out->pair_interactions += (double)ev * n * (n - 1.0);
    }
    return 0;
}

int sbk_reset_stats(sbk_engine *e)
{
    int rc = prepare(e);
    if (rc) return rc;
    rc = read_states(e);
    if (rc) return rc;
    for (size_t i = 0; i < e->clouds.size(); ++i) {
        e->attempts_at_reset[i] = e->h_state[i].attempts;
        e->steps_at_reset[i] = e->h_state[i].steps;
        e->evals_at_reset[i] = e->h_state[i].evals;
    }
    e->coulomb_ms = 0.0;
    e->coulomb_launches = 0;
    e->pers_count = 0;
    e->reorders_at_reset = e->reorders;
    return 0;
}

static int eval_common(sbk_engine *e, int cloud, int precision, int with_fields, double *out)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (!out || cloud < 0 || cloud >= (int)e->clouds.size()) return fail(SBK_EINVAL, "bad cloud id");
    if (precision < SBK_FP64_EXACT || precision > SBK_FP64_RSQ)
        return fail(SBK_EINVAL, "unknown precision %d", precision);
    const HostCloud &c = e->clouds[cloud];
    ItemList *L = nullptr;
    rc = get_list(e, precision, false, &L);
    if (rc) return rc;
    if ((rc = check_partials(e, L))) return rc;
    sbk::EngineDev d = e->dev;
    d.slices = L->d_slices;
    const bool need = !with_fields || c.f.coulomb;
    if (need && L->nitems > 0) {
        sbk::CoulombArgs a = coulomb_args(e, *L, d.P0, false);
        rc = launch_coulomb_timed(e, precision, a);
        if (rc) return rc;
    }
    if (sbk::launch_eval_derivs(d, with_fields, e->stream))
        return fail(SBK_EHIP, "derivs kernel launch failed");
    std::vector<double> buf(3 * (size_t)c.n);
    for (int k = 0; k < 3; ++k)
        HIP_TRY(hipMemcpyAsync(buf.data() + k * (size_t)c.n, d.derivs_out + k * e->stride + c.off,
                               sizeof(double) * c.n, hipMemcpyDeviceToHost, e->stream));
    rc = read_states(e);
    if (rc) return rc;
    for (int i = 0; i < c.n; ++i) {
        const size_t o = 3 * (size_t)caller_index(c, i);
        for (int k = 0; k < 3; ++k) out[o + k] = buf[k * (size_t)c.n + i];
    }
    return 0;
}

int sbk_eval_force(sbk_engine *e, int cloud, double *derivs)
{
    if (!e) return fail(SBK_EINVAL, "null engine");
    return eval_common(e, cloud, e->cfg.precision, 1, derivs);
}

int sbk_eval_coulomb(sbk_engine *e, int cloud, int precision, double *out)
{
    return eval_common(e, cloud, precision, 0, out);
}

int sbk_set_launch(sbk_engine *e, int tpt, int groups, int slices, int local)
{
    if (!e) return fail(SBK_EINVAL, "null engine");
    if (tpt != 0 && tpt != 1 && tpt != 2 && tpt != 4 && tpt != 8)
        return fail(SBK_EINVAL, "targets_per_thread must be 0 (automatic), 1, 2, 4 or 8");
    if (groups != 1 && groups != 4) return fail(SBK_EINVAL, "source_groups must be 1 or 4");
    if (slices < 0) return fail(SBK_EINVAL, "slices must be >= 0");
    if (local < -1 || local > 1) return fail(SBK_EINVAL, "local_coords must be -1, 0 or 1");
    if (tpt != e->tpt || groups != e->groups || slices != e->slices_req || local != e->local_req) {
        e->tpt = tpt;
        e->groups = groups;
        e->slices_req = slices;
        e->local_req = local;
        destroy_graph(e);
    }
    return 0;
}

int sbk_set_coulomb_kernel(sbk_engine *e, int kernel)
{
    if (!e) return fail(SBK_EINVAL, "null engine");
    if (kernel != SBK_KERNEL_AUTO && kernel != SBK_KERNEL_TILED && kernel != SBK_KERNEL_STREAM)
        return fail(SBK_EINVAL, "unknown Coulomb kernel %d", kernel);
    if (kernel != e->kernel_req) {
        e->kernel_req = kernel;
        destroy_graph(e);
    }
    return 0;
}

int sbk_get_launch(sbk_engine *e, int precision, int *tpt_out, int *groups_out, int *slices_out,
                   int *local_out, int *kernel_out)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (precision < SBK_FP64_EXACT || precision > SBK_FP64_RSQ)
        return fail(SBK_EINVAL, "unknown precision %d", precision);
    ItemList *L = nullptr;
    if ((rc = get_list(e, precision, false, &L))) return rc;
    const bool exact = precision == SBK_FP64_EXACT;
    if (tpt_out) *tpt_out = exact ? 1 : targets_per_thread(e, precision);
    if (groups_out) *groups_out = exact ? 1 : e->groups;
    if (slices_out) *slices_out = L->max_slices;
    if (local_out) *local_out = local_coords(e, precision);
    if (kernel_out)
        *kernel_out = coulomb_kernel(e, precision) == sbk::KK_STREAM ? SBK_KERNEL_STREAM : SBK_KERNEL_TILED;
    return 0;
}

// Mean time of `reps` Coulomb launches on the current positions; with_sum
// adds the pass that reads every slice partial back (what each stage kernel
// pays), so launch shapes with many slices are charged for their partials.
static int time_coulomb_pass(sbk_engine *e, int precision, int reps, bool with_sum, double *ms_out)
{
    ItemList *L = nullptr;
    int rc = get_list(e, precision, false, &L);
    if (rc) return rc;
    if ((rc = check_partials(e, L))) return rc;
    sbk::CoulombArgs a = coulomb_args(e, *L, e->dev.P0, false);
    if (precision == SBK_FP64 || precision == SBK_FP64_RSQ) a.eps_d = 1e-40;
    a.local_coords = local_coords(e, precision);
    sbk::EngineDev d = e->dev;
    d.slices = L->d_slices;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, e->stream));
    for (int i = 0; i < reps; ++i) {
        if (sbk::launch_coulomb(kp(precision), targets_per_thread(e, precision), e->groups,
                                coulomb_kernel(e, precision), a, e->stream))
            return fail(SBK_EHIP, "Coulomb kernel launch failed");
        if (with_sum && sbk::launch_eval_derivs(d, 0, e->stream))
            return fail(SBK_EHIP, "partial-sum kernel launch failed");
    }
    HIP_TRY(hipEventRecord(t1, e->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    *ms_out = ms / reps;
    return 0;
}

int sbk_time_coulomb(sbk_engine *e, int precision, int reps, double *ms_per_launch)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (reps <= 0 || !ms_per_launch) return fail(SBK_EINVAL, "reps must be > 0");
    if (precision < SBK_FP64_EXACT || precision > SBK_FP64_RSQ)
        return fail(SBK_EINVAL, "unknown precision %d", precision);
    return time_coulomb_pass(e, precision, reps, false, ms_per_launch);
}

int sbk_autotune(sbk_engine *e, int precision, int *tpt_out, int *groups_out, int *slices_out,
                 int *local_out, int *kernel_out, const char *table)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (precision == SBK_FP64_EXACT) {
        if (tpt_out) *tpt_out = 1;
        if (groups_out) *groups_out = 1;
        if (slices_out) *slices_out = 1;
        if (local_out) *local_out = 0;
        if (kernel_out) *kernel_out = SBK_KERNEL_TILED;
        return 0;
    }
    const int tpts[3] = {1, 2, 4};
    const int qs[2] = {1, 4};
    const int rounds[6] = {0, 1, 2, 3, 4, 8};   // 0: automatic split
    const bool may_local = e->spatial && (precision == SBK_MIXED || precision == SBK_MIXED_KAHAN);
    const bool may_stream = precision == SBK_MIXED || precision == SBK_MIXED_KAHAN;
    int best_t = e->tpt, best_q = e->groups, best_s = e->slices_req, best_l = e->local_req,
        best_k = e->kernel_req;
    double best = 1e300;
    char dev_name[128] = {0};
    sbk_device_info(dev_name, sizeof(dev_name), nullptr, nullptr);
    FILE *fp = table ? std::fopen(table, "a") : nullptr;
    const int old_t = e->tpt, old_q = e->groups, old_s = e->slices_req, old_l = e->local_req,
              old_k = e->kernel_req;
    {
        // about 50 ms of launches first: the first shape timed on an idle
        // device otherwise reads several per cent slow while its clocks rise
        double ms0 = 0.0;
        rc = time_coulomb_pass(e, precision, 1, true, &ms0);
        if (!rc) {
            const int warm = (int)std::min(200.0, std::max(1.0, 50.0 / std::max(ms0, 1e-3)));
            rc = time_coulomb_pass(e, precision, warm, true, &ms0);
        }
    }
    for (int kk = SBK_KERNEL_TILED; kk <= (may_stream ? SBK_KERNEL_STREAM : SBK_KERNEL_TILED) && !rc; ++kk)
    for (int l = 0; l <= (may_local ? 1 : 0) && !rc; ++l) {
        e->local_req = l;
        e->kernel_req = kk;
        for (int t : tpts) {
            for (int q : qs) {
                if (kk == SBK_KERNEL_STREAM && q != 4) continue;   // four streams per work-group
                long long blocks = 0;
                const int blk = sbk::kBlock / q * t;
                for (auto &c : e->clouds)
                    if (c.n > 0) blocks += (c.n + blk - 1) / blk;
                const long long capacity =
                    (long long)sbk::coulomb_blocks_per_cu(precision, t, e->unit_charges ? 1 : 0, q,
                                                          local_coords(e, precision),
                                                          kk == SBK_KERNEL_STREAM ? sbk::KK_STREAM
                                                                                  : sbk::KK_TILED) *
                    e->cus;
                for (int k : rounds) {
                    const int slices = k == 0 ? 0
                                              : (int)std::max(1LL, (k * capacity + blocks / 2) /
                                                                       std::max(1LL, blocks));
                    e->tpt = t;
                    e->groups = q;
                    e->slices_req = slices;
                    double ms = 0.0;
                    rc = time_coulomb_pass(e, precision, 1, true, &ms);   // warm-up, list build
                    if (rc) break;
                    rc = time_coulomb_pass(e, precision, 3, true, &ms);
                    if (rc) break;
                    if (fp)
                        std::fprintf(fp, "%s\t%d\t%zu\t%d\t%d\t%d\t%d\t%d\t%d\t%.6f\n", dev_name,
                                     e->ntot, e->clouds.size(), precision, t, q, slices,
                                     local_coords(e, precision), coulomb_kernel(e, precision), ms);
                    if (ms < best) {
                        best = ms;
                        best_t = t;
                        best_q = q;
                        best_s = slices;
                        best_l = l;
                        best_k = kk;
                    }
                }
                if (rc) break;
            }
            if (rc) break;
        }
    }
    if (fp) std::fclose(fp);
    e->tpt = old_t;
    e->groups = old_q;
    e->slices_req = old_s;
    e->local_req = old_l;
    e->kernel_req = old_k;
    if (rc) return rc;
    rc = sbk_set_launch(e, best_t, best_q, best_s, best_l);
    if (!rc) rc = sbk_set_coulomb_kernel(e, best_k);
    if (rc) return rc;
    if (tpt_out) *tpt_out = best_t;
    if (groups_out) *groups_out = best_q;
    if (slices_out) *slices_out = best_s;
    if (local_out) *local_out = local_coords(e, precision);
    if (kernel_out) *kernel_out = coulomb_kernel(e, precision);
    return 0;
}

int sbk_load_tuning(sbk_engine *e, int precision, const char *table)
{
    int rc = prepare(e);
    if (rc) return rc;
    if (!table) return fail(SBK_EINVAL, "null table path");
    FILE *fp = std::fopen(table, "r");
    if (!fp) return fail(SBK_ESTATE, "no tuning table at %s", table);
    char dev_name[128] = {0};
    sbk_device_info(dev_name, sizeof(dev_name), nullptr, nullptr);
    char line[512], dev[256];
    int ions, clouds, prec, t, q, sl, lc;
    double ms, best = 1e300;
    int bt = 0, bq = 0, bs = 0, bl = 0, bk = SBK_KERNEL_TILED;
    while (std::fgets(line, sizeof(line), fp)) {
        char *rest = nullptr;
        int used = 0;
        if (std::sscanf(line, "%255s %d %d %d %d %d %d%n", dev, &ions, &clouds, &prec, &t, &q,
                        &sl, &used) != 7)
            continue;
        rest = line + used;
        double f1 = 0.0, f2 = 0.0, f3 = 0.0;
        int kern = SBK_KERNEL_TILED;
        const int nf = std::sscanf(rest, "%lf %lf %lf", &f1, &f2, &f3);
        if (nf == 3) {            // ... local_coords kernel ms
            lc = (int)f1;
            kern = (int)f2;
            ms = f3;
        } else if (nf == 2) {     // older lines without the kernel column (tiled)
            lc = (int)f1;
            ms = f2;
        } else if (nf == 1) {     // older lines without the local_coords column
            lc = 0;
            ms = f1;
        } else {
            continue;
        }
        if (std::strcmp(dev, dev_name) || ions != e->ntot || clouds != (int)e->clouds.size() ||
            prec != precision)
            continue;
        if (ms < best) {
            best = ms;
            bt = t;
            bq = q;
            bs = sl;
            bl = lc;
            bk = kern;
        }
    }
    std::fclose(fp);
    if (best == 1e300) return fail(SBK_ESTATE, "no matching line in %s", table);
    rc = sbk_set_launch(e, bt, bq, bs, bl);
    return rc ? rc : sbk_set_coulomb_kernel(e, bk);
}

}  // extern "C"
