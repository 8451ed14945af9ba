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

void build_items(int ni, int nj, int tpt, int q, int unit)
{
    const int key[6] = {ni, nj, tpt, q, g.precision, unit};
    if (std::equal(key, key + 6, g.items_key)) return;   // the list on the device fits
    const int blk = sbk::kBlock / q * tpt;
    const int ib = (ni + blk - 1) / blk;
    int S = 1;
    if (g.precision != SBK_FP64_EXACT) {
        // the engine's automatic split: about eight rounds of resident
        // work-groups, at least 512 sources per slice (64 for sets too small
        // to fill the device otherwise)
        const long long capacity =
            (long long)sbk::coulomb_blocks_per_cu(g.precision, tpt, unit, q, 0) * g.cus;
        const long long target = 8 * capacity;
        const long long cap = std::max(1, nj / ((long long)ib * 8 < capacity ? 64 : 512));
        S = (int)std::max(1LL, std::min<long long>((target + ib / 2) / ib, cap));
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
    // targets per thread by size, as the engine's automatic rule: 4 for
    // large sets, 2 when the pairs fill the device, 1 for small sets
    const bool exact = g.precision == SBK_FP64_EXACT;
    const int tpt = exact ? 1 : (ni >= 32768 ? 4 : ((double)ni * nj >= 2.5e8 ? 2 : 1));
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
    sbk::CoulombArgs a{};
    a.tx = dx;
    a.ty = dx + ci;
    a.tz = dx + 2 * ci;
    a.sx = dj;
    a.sy = dj + cj;
    a.sz = dj + 2 * cj;
    a.sqd = dm;
    a.sqf = g.d_mf;
    a.kc = nullptr;
    a.out = g.d_cp;
    a.stride = stride;
    a.items = g.d_items;
    a.nitems = g.nitems;
    // The reference passes one position set as xj and, copied, as xi
    // (bin/cunbody.cpp:65-66): then each ion meets itself.  The FP64-exact sum
    // skips that pair by index; the tiled sums keep it finite with a floor on
    // the softening (the self pair then adds exactly 0), so eps2 = 0 works.
    const bool alias = ni == nj && (xi == xj || std::memcmp(xi, xj, sizeof(double) * 3 * (size_t)ni) == 0);
    a.self_skip = alias ? 1 : 0;
    a.unit_weights = unit;
    a.state = nullptr;
    a.eps_d = exact ? eps2 : std::max(eps2, 1e-40);   // 1e-40: the guard of bin/force.cpp:136
    a.eps_f = std::max((float)eps2, 1e-23f);   // FP32 needs eps >= 1e-23 to keep r=0 finite
    if (sbk::launch_coulomb(g.precision, tpt, q, a, g.stream)) die("Coulomb kernel launch", hipGetLastError());
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

}  // extern "C"
