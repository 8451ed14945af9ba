// sbk_persistent.hip -- persistent attempts: one cooperative launch runs a
// batch of DP5 attempts with grid barriers between the phases (the barrier
// and the phase functions are in sbk_device.h).  Opt-in: on MI355X the graph
// replay of the one-phase kernels is faster.
#include "sbk_device.h"

namespace sbk {

template <int R, int MODE, bool UNIT, bool LOC, int KIND>
__global__ __launch_bounds__(kBlock) void attempts_kernel(PersistArgs p)
{
    const EngineDev &d = p.d;
    __shared__ int s_any;
    for (int att = 0; att < p.attempts; ++att) {
        if (!persistent_stage<1, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<2, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<3, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<4, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<5, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<6, R, MODE, UNIT, LOC, KIND>(p)) return;
        if (!persistent_stage<7, R, MODE, UNIT, LOC, KIND>(p)) return;
        for (int c = blockIdx.x; c < d.nclouds; c += gridDim.x) {
            __syncthreads();
            controller_cloud(d, c);
        }
        if (!grid_barrier(p)) return;
        for (int b = blockIdx.x; b < d.nblocks; b += gridDim.x) {
            __syncthreads();
            commit_block(d, b);
        }
        if (!grid_barrier(p)) return;
        // another attempt only while some cloud is active (the same answer in
        // every work-group: the states were last written before this barrier)
        if (threadIdx.x == 0) s_any = 0;
        __syncthreads();
        for (int c = threadIdx.x; c < d.nclouds; c += blockDim.x)
            if (d.state[c].active) s_any = 1;
        __syncthreads();
        if (!s_any) {
            // what the remaining attempts' controllers would write for
            // inactive clouds; nothing else reads the states before the end
            if (att + 1 < p.attempts && threadIdx.x == 0)
                for (int c = blockIdx.x; c < d.nclouds; c += gridDim.x) d.state[c].accepted_last = 0;
            return;
        }
    }
}

template <int R>
static const void *persistent_fn_r(int precision, int unit, int loc)
{
#define SBK_PERS(M, U, L, K) reinterpret_cast<const void *>(&attempts_kernel<R, M, U, L, K>)
    switch (precision) {
    case KP_FP64: return loc ? nullptr : SBK_PERS(KP_FP64, false, false, 1);
    case KP_FP64_RSQ: return loc ? nullptr : SBK_PERS(KP_FP64_RSQ, false, false, 1);
    case KP_MIXED:
        return unit ? (loc ? SBK_PERS(KP_MIXED, true, true, 2) : SBK_PERS(KP_MIXED, true, false, 2))
                    : (loc ? SBK_PERS(KP_MIXED, false, true, 2) : SBK_PERS(KP_MIXED, false, false, 2));
    case KP_MIXED_KAHAN:
        return unit ? (loc ? SBK_PERS(KP_MIXED_KAHAN, true, true, 2)
                           : SBK_PERS(KP_MIXED_KAHAN, true, false, 2))
                    : (loc ? SBK_PERS(KP_MIXED_KAHAN, false, true, 2)
                           : SBK_PERS(KP_MIXED_KAHAN, false, false, 2));
    default: return nullptr;
    }
#undef SBK_PERS
}

static const void *persistent_fn(int precision, int tpt, int unit, int loc)
{
    if (precision == KP_FP64_EXACT)
        return (tpt == 1 && !loc)
                   ? reinterpret_cast<const void *>(&attempts_kernel<1, KP_FP64_EXACT, false, false, 0>)
                   : nullptr;
    if (tpt == 1) return persistent_fn_r<1>(precision, unit, loc);
    if (tpt == 2) return persistent_fn_r<2>(precision, unit, loc);
    if (tpt == 4) return persistent_fn_r<4>(precision, unit, loc);
    return nullptr;
}

int persistent_blocks_per_cu(int precision, int tpt, int unit, int loc)
{
    const void *f = persistent_fn(precision, tpt, unit, loc);
    if (!f) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kBlock, 0) != hipSuccess) return 0;
    return nb;
}

int launch_attempts_persistent(int precision, int tpt, const PersistArgs &p, int grid, void *stream)
{
    const void *f = persistent_fn(precision, tpt, p.a.unit_weights, p.a.local_coords);
    if (!f || grid <= 0) return (int)hipErrorInvalidValue;
    void *args[] = {const_cast<PersistArgs *>(&p)};
    // the launch's own status: a refused cooperative launch enqueues nothing
    return (int)hipLaunchCooperativeKernel(f, dim3(grid), dim3(kBlock), args, 0,
                                           (hipStream_t)stream);
}

}  // namespace sbk
