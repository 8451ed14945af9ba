// sbk_nbody.cpp -- the reference's NBODY backend ABI on the MI355X.
//
// The reference's caller declares (bin/nbody.cpp:12-18)
//   initialize(int, bool), finalize(), simulation_Async(float4*, float3*, int, int, int, int)
// and the _begin/_end variants, and drives them from _nbody::Initialization /
// force_begin / force_end (bin/nbody.cpp:24-76,216-271): positions go in as
// FP32 float4 (w ignored), accelerations come back as FP32 float3
//   acc_i = sum_j (x_j - x_i) (|x_j - x_i|^2 + 1e-7)^{-3/2}
// (softening of bin/nbody_kernel.cu:19, no charge, bin/nbody_kernel.cu:182-214),
// and the caller records an event on stream 0 after the call and polls it
// (bin/nbody.cpp:243-245,265-268).  Everything here is therefore enqueued on
// the null stream.  The launch-shape hints p, q, ntiles of the reference's
// table are accepted and ignored: the tiled kernel chooses its own shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/simbuca_nbody.h"
#include "sbk_internal.h"

namespace {

struct NbodyState {
    int n_max = 0;
    int cus = 256;
    double *d_pos = nullptr;       // SoA x, y, z [3][n_max]
    float4 *d_in = nullptr;
    float3 *d_out = nullptr;
    double *d_cp = nullptr;        // partial sums [slices][3][n_max]
    int cp_slices = 0;
    sbk::CoulombItem *d_items = nullptr;
    int nitems = 0, items_for = -1, slices = 0, cap_items = 0;
};

NbodyState g;

constexpr int kTpt = 4, kGroups = 4;

void die(const char *what, hipError_t e)
{
    std::fprintf(stderr, "NBODY (libsimbuca_hip): %s: %s\n", what,
                 e == hipSuccess ? "error" : hipGetErrorString(e));
    std::exit(EXIT_FAILURE);
}

#define CHK(expr)                             \
    do {                                      \
        hipError_t e_ = (expr);               \
        if (e_ != hipSuccess) die(#expr, e_); \
    } while (0)

__global__ void split_positions(const float4 *in, double *pos, int n, int stride)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = in[i];
    pos[i] = p.x;
    pos[stride + i] = p.y;
    pos[2 * stride + i] = p.z;
}

__global__ void gather_slices(const double *cp, int stride, int slices, int n, float3 *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = cp[(long long)k * stride + i];
        for (int q = 1; q < slices; ++q) s[k] += cp[((long long)q * 3 + k) * stride + i];
    }
    out[i] = make_float3((float)s[0], (float)s[1], (float)s[2]);
}

void build_items(int n)
{
    if (g.items_for == n) return;
    const int blk = sbk::kBlock / kGroups * kTpt;
    const int ib = (n + blk - 1) / blk;
    const long long capacity =
        (long long)sbk::coulomb_blocks_per_cu(sbk::KP_MIXED, kTpt, 1, kGroups, 0, sbk::KK_TILED) * g.cus;
    int S = (int)std::max(1LL, std::min<long long>((8 * capacity + ib / 2) / ib, std::max(1, n / 512)));
    // slices of whole LDS tiles for all wave groups
    const int quantum = sbk::kSubTile * kGroups;
    int chunk = (n + S - 1) / S;
    chunk = ((chunk + quantum - 1) / quantum) * quantum;
    S = (n + chunk - 1) / chunk;
    std::vector<sbk::CoulombItem> items;
    for (int s = 0; s < S; ++s)
        for (int b = 0; b < ib; ++b) {
            sbk::CoulombItem w{};
            w.i0 = b * blk;
            w.i1 = std::min(n, (b + 1) * blk);
            w.j0 = s * chunk;
            w.j1 = std::min(n, (s + 1) * chunk);
            w.slice = s;
            items.push_back(w);
        }
    CHK(hipStreamSynchronize(nullptr));   // earlier work may still read the old lists
    if ((int)items.size() > g.cap_items) {
        (void)hipFree(g.d_items);
        g.cap_items = (int)items.size();
        CHK(hipMalloc(&g.d_items, sizeof(sbk::CoulombItem) * g.cap_items));
    }
    CHK(hipMemcpy(g.d_items, items.data(), sizeof(sbk::CoulombItem) * items.size(),
                  hipMemcpyHostToDevice));
    if (S > g.cp_slices) {
        (void)hipFree(g.d_cp);
        g.cp_slices = S;
        CHK(hipMalloc(&g.d_cp, sizeof(double) * 3 * (size_t)g.n_max * S));
    }
    g.nitems = (int)items.size();
    g.slices = S;
    g.items_for = n;
}

void release()
{
    (void)hipDeviceSynchronize();
    (void)hipFree(g.d_pos);
    (void)hipFree(g.d_in);
    (void)hipFree(g.d_out);
    (void)hipFree(g.d_cp);
    (void)hipFree(g.d_items);
    g = NbodyState{};
}

}  // namespace

extern "C" {

void initialize(int n_max, bool withCharge)
{
    (void)withCharge;   // the reference ignores it too (bin/nbody_kernel.cu:116-171)
    if (n_max <= 0) {
        std::fprintf(stderr, "NBODY initialize: n_max must be > 0 (got %d)\n", n_max);
        std::exit(EXIT_FAILURE);
    }
    if (g.n_max) release();
    int dev = 0;
    CHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    CHK(hipGetDeviceProperties(&p, dev));
    g.cus = std::max(1, p.multiProcessorCount);
    g.n_max = n_max;
    CHK(hipMalloc(&g.d_pos, sizeof(double) * 3 * (size_t)n_max));
    CHK(hipMalloc(&g.d_in, sizeof(float4) * (size_t)n_max));
    CHK(hipMalloc(&g.d_out, sizeof(float3) * (size_t)n_max));
}

float finalize(void)
{
    release();
    return 0.0f;
}

void simulation_Async(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q, int ntiles)
{
    (void)p; (void)q; (void)ntiles;
    if (!g.n_max) {
        std::fprintf(stderr, "NBODY simulation_Async: initialize() was not called\n");
        std::exit(EXIT_FAILURE);
    }
    if (nbody < 0 || nbody > g.n_max || (nbody > 0 && (!h_p || !res))) {
        std::fprintf(stderr, "NBODY simulation_Async: nbody=%d outside [0, %d]\n", nbody, g.n_max);
        std::exit(EXIT_FAILURE);
    }
    if (nbody == 0) return;
    build_items(nbody);
    CHK(hipMemcpyAsync(g.d_in, h_p, sizeof(float4) * nbody, hipMemcpyHostToDevice, nullptr));
    hipLaunchKernelGGL(split_positions, dim3((nbody + 255) / 256), dim3(256), 0, nullptr,
                       (const float4 *)g.d_in, g.d_pos, nbody, g.n_max);
    CHK(hipGetLastError());
    sbk::CoulombArgs a{};
    a.tx = a.sx = g.d_pos;
    a.ty = a.sy = g.d_pos + g.n_max;
    a.tz = a.sz = g.d_pos + 2 * (size_t)g.n_max;
    a.sqd = nullptr;
    a.sqf = nullptr;
    a.kc = nullptr;
    a.out = g.d_cp;
    a.stride = g.n_max;
    a.items = g.d_items;
    a.nitems = g.nitems;
    a.self_skip = 0;
    a.unit_weights = 1;          // w of the float4 is ignored (bin/nbody_kernel.cu:199-210)
    a.state = nullptr;
    a.eps_f = 1e-7f;             // softeningSquared, bin/nbody_kernel.cu:19
    a.eps_d = 1e-7;
    if (sbk::launch_coulomb(sbk::KP_MIXED, kTpt, kGroups, sbk::KK_TILED, a, nullptr))
        die("Coulomb kernel launch", hipGetLastError());
    hipLaunchKernelGGL(gather_slices, dim3((nbody + 255) / 256), dim3(256), 0, nullptr,
                       (const double *)g.d_cp, g.n_max, g.slices, nbody, g.d_out);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(res, g.d_out, sizeof(float3) * nbody, hipMemcpyDeviceToHost, nullptr));
}

void simulation_Async_begin(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q, int ntiles)
{
    simulation_Async(h_p, res, nbody, p, q, ntiles);
}

void simulation_Async_end(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q, int ntiles)
{
    (void)h_p; (void)res; (void)nbody; (void)p; (void)q; (void)ntiles;
    CHK(hipStreamSynchronize(nullptr));
}

}  // extern "C"
