// sbk_cunbody.cpp -- the reference's CUNBODY force-call ABI on the MI355X.
//
// Replaces libcunbody1.a: same three extern "C" entry points
// (bin/cunbody.h:11-33), same contract as their call sites
// (bin/cunbody.cpp:65-66,88 and bin/force.cpp:48-55,84-93): host FP64
// arrays in and out, _begin enqueues and returns, _end with the same
// arguments waits and fills ai.  Device buffers are owned here, cached by
// capacity, and reused across calls on the calling thread's current device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/simbuca_hip.h"
#include "sbk_internal.h"

namespace {

struct CunbodyState {
    int device = -1;
    hipStream_t stream = nullptr;
    int cap_i = 0, cap_j = 0;
    double *d_buf = nullptr;        // xi SoA [3][cap_i], xj SoA [3][cap_j], mj [cap_j], ai [3][cap_i]
    float *d_mf = nullptr;          // mj as FP32
    double *h_buf = nullptr;        // pinned staging of the same layout
    sbk::CoulombItem *d_items = nullptr;
    int cap_items = 0;
    double *d_cp = nullptr;         // slice partials [slices][3][cap_i]
    long long cap_cp = 0;
    int nitems = 0;
    int pending = 0;
    int ni = 0, nj = 0;
    int precision = SBK_FP64;
    int cus = 256;
    int items_key[6] = {-1, -1, -1, -1, -1, -1};   // (ni, nj, tpt, q, precision, unit) of the device list
    std::vector<float> h_mf;                // mj as FP32, kept until the next call's copy
};

CunbodyState g;

// Errors in this ABI cannot be returned (the prototypes are void); they end
// the process with a message, as bin/nbody_kernel.cu:37-47 does.
void die(const char *what, hipError_t e)
{
    std::fprintf(stderr, "cunbody1 (libsimbuca_hip): %s: %s\n", what,
                 e == hipSuccess ? "error" : hipGetErrorString(e));
    std::exit(EXIT_FAILURE);
}

#define CHK(expr)                           \
    do {                                    \
        hipError_t e_ = (expr);             \
        if (e_ != hipSuccess) die(#expr, e_); \
    } while (0)

void ensure(int ni, int nj)
{
    int dev = 0;
    CHK(hipGetDevice(&dev));
    if (dev != g.device) {
        if (g.stream) {
            // buffers belong to the previous device
            (void)hipSetDevice(g.device);
            (void)hipStreamSynchronize(g.stream);
            (void)hipFree(g.d_buf);
            (void)hipFree(g.d_mf);
            (void)hipFree(g.d_items);
            (void)hipFree(g.d_cp);
            (void)hipHostFree(g.h_buf);
            (void)hipStreamDestroy(g.stream);
            (void)hipSetDevice(dev);
        }
        CunbodyState fresh;
        fresh.precision = g.precision;
        g = fresh;
        g.device = dev;
        hipDeviceProp_t p;
        CHK(hipGetDeviceProperties(&p, dev));
        g.cus = std::max(1, p.multiProcessorCount);
        CHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    }
    if (ni > g.cap_i || nj > g.cap_j) {
        CHK(hipStreamSynchronize(g.stream));
        (void)hipFree(g.d_buf);
        (void)hipFree(g.d_mf);
        if (g.h_buf) (void)hipHostFree(g.h_buf);
        g.cap_i = std::max(ni, g.cap_i);
        g.cap_j = std::max(nj, g.cap_j);
        const size_t doubles = 6 * (size_t)g.cap_i + 4 * (size_t)g.cap_j;
        CHK(hipMalloc(&g.d_buf, sizeof(double) * doubles));
        CHK(hipMalloc(&g.d_mf, sizeof(float) * g.cap_j));
        CHK(hipHostMalloc(&g.h_buf, sizeof(double) * doubles, hipHostMallocDefault));
    }
}

// Work items of one call: targets [0, ni) in blocks of kBlock/q*tpt, sources
// [0, nj) in S slices of whole LDS tiles.  S follows the engine's automatic
// split (about eight rounds of resident work-groups, at least 512 sources per
// slice, 64 for sets too small to fill the device), capped at max_slices.
std::vector<sbk::CoulombItem> cunbody_items(int ni, int nj, int tpt, int q, int unit, int precision,
                                            int cus, int max_slices, int *slices)
{
    const int blk = sbk::kBlock / q * tpt;
    const int ib = (ni + blk - 1) / blk;
    int S = 1;
    if (precision != SBK_FP64_EXACT) {
        const long long capacity =
            (long long)sbk::coulomb_blocks_per_cu(precision, tpt, unit, q, 0, sbk::KK_TILED) * cus;
        const long long target = 8 * capacity;
        const long long cap = std::max(1, nj / ((long long)ib * 8 < capacity ? 64 : 512));
        S = (int)std::max(1LL, std::min<long long>((target + ib / 2) / ib, cap));
        S = std::min(S, max_slices);
    }
    // slices of whole LDS tiles for all q wave groups, as the engine's
    const int quantum = sbk::kSubTile * q;
    int chunk = (nj + S - 1) / S;
    chunk = ((chunk + quantum - 1) / quantum) * quantum;
    S = std::max(1, (nj + chunk - 1) / chunk);
    std::vector<sbk::CoulombItem> items;
    for (int s = 0; s < S; ++s)
        for (int b = 0; b < ib; ++b) {
            sbk::CoulombItem w{};
            w.cloud = 0;
            w.i0 = b * blk;
            w.i1 = std::min(ni, (b + 1) * blk);
            w.j0 = s * chunk;
            w.j1 = std::min(nj, (s + 1) * chunk);
            w.slice = s;
            items.push_back(w);
        }
    *slices = S;
    return items;
}

// Targets per thread by size, as the engine's automatic rule: 4 for large
// sets, 2 when the pairs fill the device, 1 for small sets.
int cunbody_tpt(int ni, int nj, int precision)
{
    if (precision == SBK_FP64_EXACT) return 1;
    return ni >= 32768 ? 4 : ((double)ni * nj >= 2.5e8 ? 2 : 1);
}

void build_items(int ni, int nj, int tpt, int q, int unit)
{
    const int key[6] = {ni, nj, tpt, q, g.precision, unit};
    if (std::equal(key, key + 6, g.items_key)) return;   // the list on the device fits
    int S = 1;
    std::vector<sbk::CoulombItem> items =
        cunbody_items(ni, nj, tpt, q, unit, g.precision, g.cus, 1 << 30, &S);
    if ((int)items.size() > g.cap_items) {
        (void)hipFree(g.d_items);
        g.cap_items = (int)items.size();
        CHK(hipMalloc(&g.d_items, sizeof(sbk::CoulombItem) * g.cap_items));
    }
    g.nitems = (int)items.size();
    CHK(hipMemcpyAsync(g.d_items, items.data(), sizeof(sbk::CoulombItem) * items.size(),
                       hipMemcpyHostToDevice, g.stream));
    CHK(hipStreamSynchronize(g.stream));   // `items` is pageable and goes out of scope
    std::copy(key, key + 6, g.items_key);
}

// Coulomb arguments of one call on SoA device arrays.  The reference passes
// one position set as xj and, copied, as xi (bin/cunbody.cpp:65-66): then
// each ion meets itself.  The FP64-exact sum skips that pair by index; the
// tiled sums keep it finite with a floor on the softening (the self pair then
// adds exactly 0), so eps2 = 0 works.
sbk::CoulombArgs cunbody_args(const double *dxi, long long si, const double *dxj, long long sj,
                              const double *dm, const float *dmf, double *cp, long long stride,
                              const sbk::CoulombItem *items, int nitems, double eps2, bool alias,
                              int unit, int precision)
{
    sbk::CoulombArgs a{};
    a.tx = dxi;
    a.ty = dxi + si;
    a.tz = dxi + 2 * si;
    a.sx = dxj;
    a.sy = dxj + sj;
    a.sz = dxj + 2 * sj;
    a.sqd = dm;
    a.sqf = dmf;
    a.kc = nullptr;
    a.out = cp;
    a.stride = stride;
    a.items = items;
    a.nitems = nitems;
    a.self_skip = alias ? 1 : 0;
    a.unit_weights = unit;
    a.state = nullptr;
    a.eps_d = precision == SBK_FP64_EXACT ? eps2 : std::max(eps2, 1e-40);   // 1e-40: bin/force.cpp:136
    a.eps_f = std::max((float)eps2, 1e-23f);   // FP32 needs eps >= 1e-23 to keep r=0 finite
    return a;
}

bool same_positions(const double (*xi)[3], const double (*xj)[3], int ni, int nj)
{
    return ni == nj && (xi == xj || std::memcmp(xi, xj, sizeof(double) * 3 * (size_t)ni) == 0);
}

__global__ void reduce_slices(const double *cp, long long stride, int slices, int ni, double *ai)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    for (int k = 0; k < 3; ++k) {
        double s = cp[k * stride + i];
        for (int q = 1; q < slices; ++q) s += cp[(long long)q * 3 * stride + k * stride + i];
        ai[k * (long long)ni + i] = s;
    }
}

// ---------------------------------------------------------------------------
// The same call without any device allocation (sbk_cunbody_force_static):
// every device buffer is a region of one file-scope array, `scratch`.
// Layout, in doubles, for at most kStaticMaxIons targets and sources:
//   xj AoS staging [3 max], xi AoS staging [3 max], xj SoA [3][max],
//   xi SoA [3][max], mj [max], mj FP32 [max/2 doubles], ai AoS [max][3],
//   slice partials [kStaticMaxSlices][3][max], work items.
// ---------------------------------------------------------------------------
constexpr int kStaticMaxIons = 32768;
constexpr int kStaticMaxSlices = 16;
constexpr int kStaticMaxItems = (kStaticMaxIons / 64) * kStaticMaxSlices;   // 64-target blocks at least
constexpr size_t kMax = kStaticMaxIons;
constexpr size_t kOffXjA = 0, kOffXiA = kOffXjA + 3 * kMax, kOffXj = kOffXiA + 3 * kMax,
                 kOffXi = kOffXj + 3 * kMax, kOffM = kOffXi + 3 * kMax, kOffMf = kOffM + kMax,
                 kOffAi = kOffMf + kMax / 2, kOffCp = kOffAi + 3 * kMax,
                 kOffItems = kOffCp + (size_t)kStaticMaxSlices * 3 * kMax,
                 kScratchDoubles = kOffItems + (size_t)kStaticMaxItems * sizeof(sbk::CoulombItem) / 8;
static_assert(sizeof(sbk::CoulombItem) % 8 == 0, "items are stored in whole doubles");

static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` per device (the array's name is only a host-side
// stand-in); looked up once per device, with that device current.
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};

int static_scratch(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return sbk::record_error(SBK_EINVAL, "sbk_cunbody_force_static: device index out of range");
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[256];
            std::snprintf(msg, sizeof(msg), "hipGetSymbolAddress(scratch) failed: %s",
                          hipGetErrorString(e));
            return sbk::record_error(SBK_EHIP, msg);
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// soa[k * stride + i] = aos[3 i + k]; also the FP32 copy of the weights
__global__ void static_stage(const double *xj_aos, const double *xi_aos, const double *mj,
                             double *xj, double *xi, float *mf, int ni, int nj, long long stride)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nj) {
        for (int k = 0; k < 3; ++k) xj[k * stride + i] = xj_aos[3 * (long long)i + k];
        mf[i] = (float)mj[i];
    }
    if (i < ni)
        for (int k = 0; k < 3; ++k) xi[k * stride + i] = xi_aos[3 * (long long)i + k];
}

// ai[i][k] = sum over slices, in slice order, of the partials
__global__ void static_reduce(const double *cp, long long stride, int slices, int ni, double *ai)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    for (int k = 0; k < 3; ++k) {
        double s = cp[k * stride + i];
        for (int q = 1; q < slices; ++q) s += cp[(long long)q * 3 * stride + k * stride + i];
        ai[3 * (long long)i + k] = s;
    }
}

#define STATIC_HIP(expr)                                                                    \
    do {                                                                                    \
        const hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                             \
            char m_[256];                                                                   \
            std::snprintf(m_, sizeof(m_), "%s failed: %s", #expr, hipGetErrorString(e_));   \
            return sbk::record_error(SBK_EHIP, m_);                                         \
        }                                                                                   \
    } while (0)

int cunbody_force_static(double xj[][3], double mj[], double xi[][3], double eps2, double ai[][3],
                         int ni, int nj)
{
    char msg[256];
    if (ni < 0 || nj < 0 || (ni > 0 && (!xi || !ai)) || (nj > 0 && (!xj || !mj))) {
        std::snprintf(msg, sizeof(msg), "sbk_cunbody_force_static: bad arguments (ni=%d nj=%d)", ni, nj);
        return sbk::record_error(SBK_EINVAL, msg);
    }
    if (ni > kStaticMaxIons || nj > kStaticMaxIons) {
        std::snprintf(msg, sizeof(msg),
                      "sbk_cunbody_force_static: at most %d targets and %d sources (got ni=%d, nj=%d)",
                      kStaticMaxIons, kStaticMaxIons, ni, nj);
        return sbk::record_error(SBK_EINVAL, msg);
    }
    if (ni == 0) return 0;
    if (nj == 0) {   // no sources: every sum is empty
        for (int i = 0; i < ni; ++i) ai[i][0] = ai[i][1] = ai[i][2] = 0.0;
        return 0;
    }
    int dev = 0;
    STATIC_HIP(hipGetDevice(&dev));
    double *base = nullptr;
    if (int rc = static_scratch(dev, &base)) return rc;
    const int precision = g.precision;
    const bool exact = precision == SBK_FP64_EXACT;
    const int tpt = cunbody_tpt(ni, nj, precision);
    const int q = exact ? 1 : 4;
    int unit = 1;
    for (int j = 0; j < nj && unit; ++j) unit = mj[j] == 1.0;
    int cus = 256;
    {
        hipDeviceProp_t p;
        STATIC_HIP(hipGetDeviceProperties(&p, dev));
        cus = std::max(1, p.multiProcessorCount);
    }
    int S = 1;
    const std::vector<sbk::CoulombItem> items =
        cunbody_items(ni, nj, tpt, q, unit, precision, cus, kStaticMaxSlices, &S);
    if (S > kStaticMaxSlices || (int)items.size() > kStaticMaxItems)
        return sbk::record_error(SBK_ESTATE, "sbk_cunbody_force_static: work list exceeds the scratch");
    double *xjA = base + kOffXjA, *xiA = base + kOffXiA, *dxj = base + kOffXj, *dxi = base + kOffXi;
    double *dm = base + kOffM, *dai = base + kOffAi, *cp = base + kOffCp;
    float *dmf = reinterpret_cast<float *>(base + kOffMf);
    sbk::CoulombItem *ditems = reinterpret_cast<sbk::CoulombItem *>(base + kOffItems);
    // null stream throughout: blocking copies, kernels in order after them
    STATIC_HIP(hipMemcpy(xjA, xj, sizeof(double) * 3 * (size_t)nj, hipMemcpyHostToDevice));
    STATIC_HIP(hipMemcpy(xiA, xi, sizeof(double) * 3 * (size_t)ni, hipMemcpyHostToDevice));
    STATIC_HIP(hipMemcpy(dm, mj, sizeof(double) * (size_t)nj, hipMemcpyHostToDevice));
    STATIC_HIP(hipMemcpy(ditems, items.data(), sizeof(sbk::CoulombItem) * items.size(),
                         hipMemcpyHostToDevice));
    const int n = std::max(ni, nj);
    hipLaunchKernelGGL(static_stage, dim3((n + 255) / 256), dim3(256), 0, nullptr,
                       (const double *)xjA, (const double *)xiA, (const double *)dm, dxj, dxi, dmf,
                       ni, nj, (long long)kMax);
    STATIC_HIP(hipGetLastError());
    const sbk::CoulombArgs a = cunbody_args(dxi, (long long)kMax, dxj, (long long)kMax, dm, dmf, cp,
                                            (long long)kMax, ditems, (int)items.size(), eps2,
                                            same_positions(xi, xj, ni, nj), unit, precision);
    if (sbk::launch_coulomb(precision, tpt, q, sbk::KK_TILED, a, nullptr)) {
        std::snprintf(msg, sizeof(msg), "sbk_cunbody_force_static: Coulomb kernel launch failed: %s",
                      hipGetErrorString(hipGetLastError()));
        return sbk::record_error(SBK_EHIP, msg);
    }
    hipLaunchKernelGGL(static_reduce, dim3((ni + 255) / 256), dim3(256), 0, nullptr,
                       (const double *)cp, (long long)kMax, S, ni, dai);
    STATIC_HIP(hipGetLastError());
    STATIC_HIP(hipMemcpy(ai, dai, sizeof(double) * 3 * (size_t)ni, hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

extern "C" {

int sbk_cunbody_set_precision(int precision)
{
    const int old = g.precision;
    if (precision == SBK_FP64_EXACT || precision == SBK_FP64 || precision == SBK_MIXED ||
        precision == SBK_MIXED_KAHAN || precision == SBK_FP64_RSQ)
        g.precision = precision;
    return old;
}

void cunbody1_force_begin(double xj[][3], double mj[], double xi[][3], double eps2, double ai[][3],
                          int ni, int nj)
{
    (void)ai;
    if (ni < 0 || nj < 0 || (ni > 0 && !xi) || (nj > 0 && (!xj || !mj))) {
        std::fprintf(stderr, "cunbody1_force_begin: bad arguments (ni=%d nj=%d)\n", ni, nj);
        std::exit(EXIT_FAILURE);
    }
    ensure(std::max(ni, 1), std::max(nj, 1));
    // a begin without its end: let that call's copies finish before the
    // staging buffers are refilled
    if (g.pending) CHK(hipStreamSynchronize(g.stream));
    g.ni = ni;
    g.nj = nj;
    g.pending = 1;
    if (ni == 0) return;
    const size_t ci = g.cap_i, cj = g.cap_j;
    double *hx = g.h_buf;                 // xi [3][ci]
    double *hj = g.h_buf + 3 * ci;        // xj [3][cj]
    double *hm = hj + 3 * cj;             // mj [cj]
    double *hai = hm + cj;                // ai [3][ni], filled by the D2H copy
    if (nj == 0) {
        // no sources: every sum is empty
        std::fill(hai, hai + 3 * (size_t)ni, 0.0);
        return;
    }
    for (int i = 0; i < ni; ++i)
        for (int k = 0; k < 3; ++k) hx[k * ci + i] = xi[i][k];
    for (int j = 0; j < nj; ++j) {
        for (int k = 0; k < 3; ++k) hj[k * cj + j] = xj[j][k];
        hm[j] = mj[j];
    }
    double *dx = g.d_buf, *dj = g.d_buf + 3 * ci, *dm = dj + 3 * cj, *dai = dm + cj;
    CHK(hipMemcpyAsync(dx, hx, sizeof(double) * 3 * ci, hipMemcpyHostToDevice, g.stream));
    CHK(hipMemcpyAsync(dj, hj, sizeof(double) * 4 * cj, hipMemcpyHostToDevice, g.stream));
    // the FP32 weights come from a buffer that lives until the next call,
    // which follows the end of this one (cunbody1_force_end synchronizes)
    g.h_mf.resize(nj);
    for (int j = 0; j < nj; ++j) g.h_mf[j] = (float)mj[j];
    CHK(hipMemcpyAsync(g.d_mf, g.h_mf.data(), sizeof(float) * nj, hipMemcpyHostToDevice, g.stream));
    const bool exact = g.precision == SBK_FP64_EXACT;
    const int tpt = cunbody_tpt(ni, nj, g.precision);
    const int q = exact ? 1 : 4;
    int unit = 1;
    for (int j = 0; j < nj && unit; ++j) unit = mj[j] == 1.0;
    build_items(ni, nj, tpt, q, unit);
    const int blk = sbk::kBlock / q * tpt;
    const int S = g.nitems / ((ni + blk - 1) / blk);   // items = slices x target blocks
    const long long stride = ci;
    if ((long long)S * 3 * stride > g.cap_cp) {
        CHK(hipStreamSynchronize(g.stream));   // earlier launches may still write the old buffer
        (void)hipFree(g.d_cp);
        g.cap_cp = (long long)S * 3 * stride;
        CHK(hipMalloc(&g.d_cp, sizeof(double) * g.cap_cp));
    }
    const sbk::CoulombArgs a =
        cunbody_args(dx, ci, dj, cj, dm, g.d_mf, g.d_cp, stride, g.d_items, g.nitems, eps2,
                     same_positions(xi, xj, ni, nj), unit, g.precision);
    if (sbk::launch_coulomb(g.precision, tpt, q, sbk::KK_TILED, a, g.stream)) die("Coulomb kernel launch", hipGetLastError());
    hipLaunchKernelGGL(reduce_slices, dim3((ni + 255) / 256), dim3(256), 0, g.stream,
                       (const double *)g.d_cp, stride, S, ni, dai);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(hai, dai, sizeof(double) * 3 * ni, hipMemcpyDeviceToHost, g.stream));
}

void cunbody1_force_end(double xj[][3], double mj[], double xi[][3], double eps2, double ai[][3],
                        int ni, int nj)
{
    (void)xj; (void)mj; (void)xi; (void)eps2;
    if (!g.pending || ni != g.ni || nj != g.nj) {
        std::fprintf(stderr, "cunbody1_force_end: no matching cunbody1_force_begin\n");
        std::exit(EXIT_FAILURE);
    }
    g.pending = 0;
    if (ni == 0) return;
    CHK(hipStreamSynchronize(g.stream));
    const size_t ci = g.cap_i, cj = g.cap_j;
    const double *hai = g.h_buf + 3 * ci + 4 * cj;
    for (int i = 0; i < ni; ++i)
        for (int k = 0; k < 3; ++k) ai[i][k] = hai[k * (size_t)ni + i];
}

void cunbody1_force(double xj[][3], double mj[], double xi[][3], double eps2, double ai[][3], int ni,
                    int nj)
{
    cunbody1_force_begin(xj, mj, xi, eps2, ai, ni, nj);
    cunbody1_force_end(xj, mj, xi, eps2, ai, ni, nj);
}

int sbk_cunbody_force_static(double xj[][3], double mj[], double xi[][3], double eps2,
                             double ai[][3], int ni, int nj)
{
    return cunbody_force_static(xj, mj, xi, eps2, ai, ni, nj);
}

}  // extern "C"
