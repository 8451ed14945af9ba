// sbk_kernels.hip -- the one-phase kernels of the engine (Coulomb sums,
// DP5 stages, controller, commit, set-up and re-sort helpers) and their
// launchers.  The device functions they call live in sbk_device.h.
#include "sbk_device.h"

namespace sbk {

__global__ __launch_bounds__(kBlock) void coulomb_fp64_exact(CoulombArgs a)
{
    coulomb_fp64_exact_item(a, xcd_item(blockIdx.x, gridDim.x));
}

template <int R, int MODE, bool UNIT, int Q, bool LOC>
__global__ __launch_bounds__(kBlock) void coulomb_tiled(CoulombArgs a)
{
    coulomb_tiled_item<R, MODE, UNIT, Q, LOC>(a, xcd_item(blockIdx.x, gridDim.x));
}

template <int R, int MODE, bool UNIT, bool LOC>
__global__ __launch_bounds__(kBlock) void coulomb_stream(CoulombArgs a)
{
    coulomb_stream_sets<R, MODE, UNIT, LOC>(a, xcd_item(blockIdx.x, gridDim.x));
}

template <int S>
__global__ __launch_bounds__(256) void stage_kernel(EngineDev d)
{
    stage_block<S>(d, blockIdx.x);
}

__global__ __launch_bounds__(256) void controller_kernel(EngineDev d)
{
    controller_cloud(d, blockIdx.x);
}

__global__ __launch_bounds__(256) void commit_kernel(EngineDev d)
{
    commit_block(d, blockIdx.x);
}

__global__ void begin_kernel(EngineDev d, double duration, int budget, int set_end)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.nclouds) return;
    CloudState &cs = d.state[c];
    if (set_end == 1) cs.t_end = cs.t + duration;
    if (set_end == 2) cs.t_end = duration;   // an absolute end time
    cs.budget = budget;
    cs.accepted_last = 0;
    cs.active = (d.params[c].n > 0) && !cs.error && (cs.t < cs.t_end) && budget != 0;
}

// derivs_out = force() at (P0, V0, lifetime) when with_fields, else the
// Coulomb sum f alone.
__global__ __launch_bounds__(256) void eval_derivs_kernel(EngineDev d, int with_fields)
{
    const IonBlock blk = d.blocks[blockIdx.x];
    const long long g = blk.start + threadIdx.x;
    if (g >= blk.end) return;
    const int c = blk.cloud;
    const CloudParams &p = d.params[c];
    double Fx, Fy, Fz;
    if (with_fields) {
        coulomb_total(d, p, c, g, Fx, Fy, Fz);
        const double t = d.state[c].t;
        double ax, ay, az;
        ion_accel(p, t, t, d.qm[g], d.P0[0][g], d.P0[1][g], d.P0[2][g], d.V0[0][g],
                  d.V0[1][g], d.V0[2][g], Fx, Fy, Fz, ax, ay, az);
        Fx = ax; Fy = ay; Fz = az;
    } else {
        CloudParams q = p;
        q.coulomb = 1;
        coulomb_total(d, q, c, g, Fx, Fy, Fz);
    }
    d.derivs_out[g] = Fx;
    d.derivs_out[d.stride + g] = Fy;
    d.derivs_out[2 * d.stride + g] = Fz;
}

// out[g] = in[src[g]]: applies a new ion order to one per-ion array.
template <typename T>
__global__ __launch_bounds__(256) void gather_kernel(const int *src, const T *in, T *out, int n)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = in[src[g]];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <int R, bool UNIT, int Q>
static const void *tiled_fn_q(int precision, int loc)
{
    switch (precision) {
    case KP_FP64:
        return loc ? nullptr : reinterpret_cast<const void *>(&coulomb_tiled<R, KP_FP64, false, Q, false>);
    case KP_FP64_RSQ:
        return loc ? nullptr : reinterpret_cast<const void *>(&coulomb_tiled<R, KP_FP64_RSQ, false, Q, false>);
    case KP_MIXED:
        return loc ? reinterpret_cast<const void *>(&coulomb_tiled<R, KP_MIXED, UNIT, Q, true>)
                   : reinterpret_cast<const void *>(&coulomb_tiled<R, KP_MIXED, UNIT, Q, false>);
    case KP_MIXED_KAHAN:
        return loc ? reinterpret_cast<const void *>(&coulomb_tiled<R, KP_MIXED_KAHAN, UNIT, Q, true>)
                   : reinterpret_cast<const void *>(&coulomb_tiled<R, KP_MIXED_KAHAN, UNIT, Q, false>);
    default: return nullptr;
    }
}

template <int R>
static const void *tiled_fn(int precision, int unit, int q, int loc)
{
    if (q == 4)
        return unit ? tiled_fn_q<R, true, 4>(precision, loc) : tiled_fn_q<R, false, 4>(precision, loc);
    if (q == 1)
        return unit ? tiled_fn_q<R, true, 1>(precision, loc) : tiled_fn_q<R, false, 1>(precision, loc);
    return nullptr;
}

template <int R>
static const void *stream_fn(int precision, int unit, int loc)
{
#define SBK_STREAM(M, U, L) reinterpret_cast<const void *>(&coulomb_stream<R, M, U, L>)
    if (precision == KP_MIXED)
        return unit ? (loc ? SBK_STREAM(KP_MIXED, true, true) : SBK_STREAM(KP_MIXED, true, false))
                    : (loc ? SBK_STREAM(KP_MIXED, false, true) : SBK_STREAM(KP_MIXED, false, false));
    if (precision == KP_MIXED_KAHAN)
        return unit ? (loc ? SBK_STREAM(KP_MIXED_KAHAN, true, true) : SBK_STREAM(KP_MIXED_KAHAN, true, false))
                    : (loc ? SBK_STREAM(KP_MIXED_KAHAN, false, true) : SBK_STREAM(KP_MIXED_KAHAN, false, false));
#undef SBK_STREAM
    return nullptr;
}

static const void *coulomb_fn(int precision, int tpt, int unit, int q, int loc, int kernel)
{
    if (kernel == KK_STREAM) {
        // four source streams per work-group over one target set of 64 tpt
        if (q != 4) return nullptr;
        switch (tpt) {
        case 1: return stream_fn<1>(precision, unit, loc);
        case 2: return stream_fn<2>(precision, unit, loc);
        case 4: return stream_fn<4>(precision, unit, loc);
        default: return nullptr;
        }
    }
    if (precision == KP_FP64_EXACT)
        return loc ? nullptr : reinterpret_cast<const void *>(&coulomb_fp64_exact);
    switch (tpt) {
    case 1: return tiled_fn<1>(precision, unit, q, loc);
    case 2: return tiled_fn<2>(precision, unit, q, loc);
    case 4: return tiled_fn<4>(precision, unit, q, loc);
    case 8: return tiled_fn<8>(precision, unit, q, loc);
    default: return nullptr;
    }
}

int coulomb_blocks_per_cu(int precision, int tpt, int unit, int q, int loc, int kernel)
{
    const void *f = coulomb_fn(precision, tpt, unit, q, loc, kernel);
    if (!f) return 1;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kBlock, 0) != hipSuccess || nb < 1)
        return 1;
    return nb;
}

int launch_coulomb(int precision, int tpt, int q, int kernel, const CoulombArgs &a, void *stream)
{
    if (a.nitems <= 0) return 0;
    const void *f = coulomb_fn(precision, tpt, a.unit_weights, q, a.local_coords, kernel);
    if (!f) return -1;
    void *args[] = {const_cast<CoulombArgs *>(&a)};
    if (hipLaunchKernel(f, dim3(a.nitems), dim3(kBlock), args, 0, (hipStream_t)stream) !=
        hipSuccess)
        return -1;
    return check_launch();
}

int launch_stage(int stage, const EngineDev &d, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (d.nblocks <= 0) return 0;
    const dim3 grid((unsigned)d.nblocks), block(256);
    switch (stage) {
    case 1: hipLaunchKernelGGL(stage_kernel<1>, grid, block, 0, s, d); break;
    case 2: hipLaunchKernelGGL(stage_kernel<2>, grid, block, 0, s, d); break;
    case 3: hipLaunchKernelGGL(stage_kernel<3>, grid, block, 0, s, d); break;
    case 4: hipLaunchKernelGGL(stage_kernel<4>, grid, block, 0, s, d); break;
    case 5: hipLaunchKernelGGL(stage_kernel<5>, grid, block, 0, s, d); break;
    case 6: hipLaunchKernelGGL(stage_kernel<6>, grid, block, 0, s, d); break;
    case 7: hipLaunchKernelGGL(stage_kernel<7>, grid, block, 0, s, d); break;
    default: return -1;
    }
    return check_launch();
}

int launch_controller(const EngineDev &d, void *stream)
{
    hipLaunchKernelGGL(controller_kernel, dim3(d.nclouds), dim3(256), 0, (hipStream_t)stream, d);
    return check_launch();
}

int launch_commit(const EngineDev &d, void *stream)
{
    if (d.nblocks <= 0) return 0;
    hipLaunchKernelGGL(commit_kernel, dim3((unsigned)d.nblocks), dim3(256), 0,
                       (hipStream_t)stream, d);
    return check_launch();
}

int launch_begin(const EngineDev &d, double duration, int budget, int set_end, void *stream)
{
    hipLaunchKernelGGL(begin_kernel, dim3((d.nclouds + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, d, duration, budget, set_end);
    return check_launch();
}

int launch_gather(const int *src, const double *in, double *out, int n, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(gather_kernel<double>, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, src, in, out, n);
    return check_launch();
}

int launch_gather_f32(const int *src, const float *in, float *out, int n, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(gather_kernel<float>, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, src, in, out, n);
    return check_launch();
}

int launch_eval_derivs(const EngineDev &d, int with_fields, void *stream)
{
    if (d.nblocks <= 0) return 0;
    hipLaunchKernelGGL(eval_derivs_kernel, dim3((unsigned)d.nblocks), dim3(256), 0,
                       (hipStream_t)stream, d, with_fields);
    return check_launch();
}

}  // namespace sbk
