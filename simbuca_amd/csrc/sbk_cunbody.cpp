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
    int nitems = 0;
    int pending = 0;
    int ni = 0, nj = 0;
    int precision = SBK_FP64;
    int cus = 256;
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
            (void)hipHostFree(g.h_buf);
            (void)hipStreamDestroy(g.stream);
            (void)hipSetDevice(dev);
        }
        g = CunbodyState{g.device, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, 0, 0,
                         0, 0, g.precision, 256};
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

void build_items(int ni, int nj, int tpt, int q)
{
    const int blk = sbk::kBlock / q * tpt;
    const int ib = (ni + blk - 1) / blk;
    int S = 1;
    if (g.precision != SBK_FP64_EXACT) {
        const long long target = 8LL * g.cus;
        S = (int)std::max(1LL, std::min<long long>((target + ib - 1) / ib, std::max(1, nj / 1024)));
    }
    int chunk = (nj + S - 1) / S;
    chunk = ((chunk + sbk::kSubTile - 1) / sbk::kSubTile) * sbk::kSubTile;
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
}

double *partials = nullptr;
long long partials_cap = 0;

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
        precision == SBK_MIXED_KAHAN)
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
    g.ni = ni;
    g.nj = nj;
    g.pending = 1;
    if (ni == 0) return;
    const size_t ci = g.cap_i, cj = g.cap_j;
    double *hx = g.h_buf;                 // xi [3][ci]
    double *hj = g.h_buf + 3 * ci;        // xj [3][cj]
    double *hm = hj + 3 * cj;             // mj [cj]
    for (int i = 0; i < ni; ++i)
        for (int k = 0; k < 3; ++k) hx[k * ci + i] = xi[i][k];
    for (int j = 0; j < nj; ++j) {
        for (int k = 0; k < 3; ++k) hj[k * cj + j] = xj[j][k];
        hm[j] = mj[j];
    }
    double *dx = g.d_buf, *dj = g.d_buf + 3 * ci, *dm = dj + 3 * cj, *dai = dm + cj;
    CHK(hipMemcpyAsync(dx, hx, sizeof(double) * 3 * ci, hipMemcpyHostToDevice, g.stream));
    CHK(hipMemcpyAsync(dj, hj, sizeof(double) * 4 * cj, hipMemcpyHostToDevice, g.stream));
    std::vector<float> mf(nj);
    for (int j = 0; j < nj; ++j) mf[j] = (float)mj[j];
    CHK(hipMemcpyAsync(g.d_mf, mf.data(), sizeof(float) * nj, hipMemcpyHostToDevice, g.stream));
    const int tpt = (g.precision == SBK_FP64_EXACT) ? 1 : 2;
    const int q = (g.precision == SBK_FP64_EXACT) ? 1 : 4;
    build_items(ni, nj, tpt, q);   // synchronizes: mf is consumed before it goes out of scope
    const int blk = sbk::kBlock / q * tpt;
    const int S = g.nitems / ((ni + blk - 1) / blk);   // items = slices x target blocks
    const long long stride = ci;
    if ((long long)S * 3 * stride > partials_cap) {
        (void)hipFree(partials);
        partials_cap = (long long)S * 3 * stride;
        CHK(hipMalloc(&partials, sizeof(double) * partials_cap));
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
    a.out = partials;
    a.stride = stride;
    a.items = g.d_items;
    a.nitems = g.nitems;
    a.self_skip = 0;
    a.unit_weights = 1;
    for (int j = 0; j < nj && a.unit_weights; ++j) a.unit_weights = mj[j] == 1.0;
    a.state = nullptr;
    a.eps_d = eps2;
    a.eps_f = std::max((float)eps2, 1e-23f);   // FP32 needs eps >= 1e-23 to keep r=0 finite
    if (sbk::launch_coulomb(g.precision, tpt, q, a, g.stream)) die("Coulomb kernel launch", hipGetLastError());
    hipLaunchKernelGGL(reduce_slices, dim3((ni + 255) / 256), dim3(256), 0, g.stream,
                       (const double *)partials, stride, S, ni, dai);
    CHK(hipGetLastError());
    double *hai = hm + cj;   // staging for ai [3][ni]
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
