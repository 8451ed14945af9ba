// sbk_device.h -- device functions of the Simbuca force and Dormand-Prince
// 5(4) engine for CDNA4 (gfx950), shared by the kernel translation units
// (sbk_kernels.hip: the one-phase kernels; sbk_persistent.hip: persistent
// attempts).  Each __global__ kernel is defined in exactly one of them.
//
// Built with -ffp-contract=off: every fused multiply-add below is written
// out with fma()/fmaf(), everything else rounds exactly as written, which is
// what lets the FP64-exact Coulomb sum and the stage combinations reproduce
// the reference's CPU arithmetic (bin/force.cpp:34-175, bin/ode.cpp:61-218)
// bit for bit.
#ifndef SBK_DEVICE_H
#define SBK_DEVICE_H

#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>

#include <type_traits>

#include "sbk_internal.h"

namespace sbk {


// bin/globals.h:6
constexpr double kPi = 3.14159265358979323;

// ---------------------------------------------------------------------------
// Work-group -> work-item mapping.  Work-groups are dealt round-robin to the
// 8 XCDs; this gives each XCD one contiguous run of items, so the items of
// one cloud and source slice (which reuse the same sources) share an L2.
// The engine's lists hold 8 equally long runs cut at equal work, the
// shorter ones padded with empty items (i0 == i1) that return at once.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_item(int b, int nb)
{
    const int xcd = b & 7, idx = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd * q + (xcd < r ? xcd : r) + idx;
}

// ---------------------------------------------------------------------------
// Coulomb direct sum, FP64 in the reference's per-target order.
// For target g: f = sum over sources s != g, in increasing s, of
// ((r*invDist^3)*k_g)*q_s with r = x_s - x_g, invDist = 1/sqrt(|r|^2)
// (bin/force.cpp:158-170).  One thread per target, no source split.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void coulomb_fp64_exact_item(const CoulombArgs &a, int item)
{
    __shared__ double lx[kBlock], ly[kBlock], lz[kBlock], lq[kBlock];
    const CoulombItem it = a.items[item];
    // an empty item pads an XCD's run of the list; the checks are uniform
    // over the work-group, so either every wave reaches the barriers or none
    if (it.i0 >= it.i1) return;
    if (a.state && (!a.state[it.cloud].active || (a.skip_valid_k1 && a.state[it.cloud].k1_valid)))
        return;
    const int tid = threadIdx.x;
    const int g = it.i0 + tid;
    const bool valid = g < it.i1;
    const double bx = valid ? a.tx[g] : 0.0;
    const double by = valid ? a.ty[g] : 0.0;
    const double bz = valid ? a.tz[g] : 0.0;
    const double k = (valid && a.kc) ? a.kc[g] : 1.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int t0 = it.j0; t0 < it.j1; t0 += kBlock) {
        const int cnt = min(kBlock, it.j1 - t0);
        __syncthreads();
        if (tid < cnt) {
            lx[tid] = a.sx[t0 + tid];
            ly[tid] = a.sy[t0 + tid];
            lz[tid] = a.sz[t0 + tid];
            lq[tid] = a.sqd[t0 + tid];
        }
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            if (a.self_skip && t0 + j == g) continue;
            const double rx = lx[j] - bx;
            const double ry = ly[j] - by;
            const double rz = lz[j] - bz;
            const double distSqr = rx * rx + ry * ry + rz * rz + a.eps_d;
            const double invDist = 1. / sqrt(distSqr);
            const double invDistCube = (invDist * invDist * invDist);
            fx += rx * invDistCube * k * lq[j];
            fy += ry * invDistCube * k * lq[j];
            fz += rz * invDistCube * k * lq[j];
        }
    }
    if (valid) {
        double *o = a.out + (long long)it.slice * 3 * a.stride + g;
        o[0] = fx;
        o[a.stride] = fy;
        o[2 * a.stride] = fz;
    }
}


// ---------------------------------------------------------------------------
// Coulomb direct sum, tiled.  A work-group owns (kBlock/Q)*R targets (R per
// thread, strided so a wave's loads are unit stride) and one slice of the
// sources.  Sources pass through LDS kBlock at a time; inside a tile every
// lane of a wave reads the same source (an LDS broadcast) and updates its R
// targets.  With Q > 1 the work-group's Q wave groups share one set of
// targets and each takes 1/Q of every tile; their sums are combined in LDS in
// a fixed order at the end, so a cloud yields Q times more work-groups
// without more partial sums in memory.
//
// MIXED: separations are formed in FP64 and rounded to FP32; |r|^2, rsqrt
// and the sum over kSubTile sources are FP32; the sub-tile sums are folded
// into FP64 (KAHAN: compensated FP32) accumulators.
// FP64: the same tiling in FP64 with 1/sqrt.
// FP64_RSQ: as FP64, with |r|^-3 from the hardware reciprocal-square-root
// estimate (v_rsq_f64) and a second-order correction (rsq3_f64) instead of
// the IEEE square root and division.
// LOC (mixed modes only): the work-group takes the centre o of its targets'
// bounding box (half-diagonal h) as origin, and each source also enters LDS
// as the FP32 offset x_s - o (one FP64 subtraction per source per
// work-group).  A source is far when its distance from the targets' bounding
// box is at least kNearFactor h; then every separation r from it is at least
// kNearFactor h long, its offset obeys |x_s - o| <= |r| + h, and the FP32
// difference of two offsets has an error of at most
// 2^-24 (|x_s - o| + |x_t - o| + |r|) <= 2^-24 (2 + 2/kNearFactor) |r|,
// against 2^-24 |r| for the FP64-formed separation.  (The box distance marks
// about half as many sources near as a sphere of radius (1 + kNearFactor) h
// around o would, with the same bound.)  The wave that stages a 64-source sub-tile stores its far
// sources first (a ballot and a lane count), so each sub-tile runs its far
// sources on the offset path, 3 instead of 6 vector instructions per pair,
// and the rest on the FP64-formed separation.  With ions in a spatial order
// most sources are far; in any other order nearly none is, so the accuracy
// never depends on the order, only the speed does.
// The self pair has r = 0 and contributes exactly 0 through the eps term.
// ---------------------------------------------------------------------------
constexpr float kNearFactor = 0.25f;

// LOC far test: the squared distance from the source offset (x, y, z) to the
// box of half-extents (hx, hy, hz) around the origin, against
// near2 = (kNearFactor h)^2.
__device__ __forceinline__ bool far_from_box(float x, float y, float z, float hx, float hy,
                                             float hz, float near2)
{
    const float ex = fmaxf(fabsf(x) - hx, 0.f);
    const float ey = fmaxf(fabsf(y) - hy, 0.f);
    const float ez = fmaxf(fabsf(z) - hz, 0.f);
    return fmaf(ex, ex, fmaf(ey, ey, ez * ez)) >= near2;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// x^(-3/2) for normal x > 0 to within a few ulp, from the hardware
// reciprocal-square-root estimate y = v_rsq_f64(x) = x^(-1/2) (1 + d):
// with r = 1 - x y^2 (one fma, exact up to the rounding of y^2),
// x^(-3/2) = y^3 (1 - r)^(-3/2) = y^3 (1 + 3r/2 + 15r^2/8 + 35r^3/16 + ...).
// Keeping the terms to r^2 leaves a relative error of 35|r|^3/16 with
// |r| ~ 2|d|: below 2^-60 for an estimate good to 2^-22, so the result is
// set by the roundings of y^2, y^3 and the last fma.
__device__ __forceinline__ double rsq3_f64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double y2 = y * y;
    const double w0 = y2 * y;
    const double r = fma(-x, y2, 1.0);
    const double c = r * fma(r, 1.875, 1.5);
    return fma(w0, c, w0);
}

template <int R, int MODE, bool UNIT, int Q, bool LOC>
__device__ __forceinline__ void coulomb_tiled_item(const CoulombArgs &a, int item)
{
    constexpr bool F64 = MODE == KP_FP64 || MODE == KP_FP64_RSQ;
    static_assert(!LOC || !F64, "local coordinates are a mixed-precision variant");
    constexpr int L = kBlock / Q;          // lanes that share one target set
    constexpr int SPAN = kBlock / Q;       // sources of a tile handled by one wave group
    __shared__ double lx[kBlock], ly[kBlock], lz[kBlock];
    __shared__ float lqf[(LOC || UNIT) ? 1 : kBlock];
    __shared__ double lqd[F64 ? kBlock : 1];
    __shared__ float4 lf[LOC ? kBlock : 1];                 // LOC: (x_s - o, weight)
    __shared__ int lfar[LOC ? kBlock / kSubTile : 1];       // LOC: far sources per sub-tile

    const CoulombItem it = a.items[item];
    // an empty item pads an XCD's run of the list; the checks are uniform
    // over the work-group, so either every wave reaches the barriers or none
    if (it.i0 >= it.i1) return;
    if (a.state && (!a.state[it.cloud].active || (a.skip_valid_k1 && a.state[it.cloud].k1_valid)))
        return;
    const int tid = threadIdx.x;
    const int lane = tid % L, grp = tid / L;

    double px[R], py[R], pz[R];
    double Ax[R], Ay[R], Az[R];
    float Sx[R], Sy[R], Sz[R], Cx[R], Cy[R], Cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int g = it.i0 + r * L + lane;
        const bool v = g < it.i1;
        px[r] = v ? a.tx[g] : 0.0;
        py[r] = v ? a.ty[g] : 0.0;
        pz[r] = v ? a.tz[g] : 0.0;
        Ax[r] = Ay[r] = Az[r] = 0.0;
        Sx[r] = Sy[r] = Sz[r] = 0.f;
        Cx[r] = Cy[r] = Cz[r] = 0.f;
    }
    const float epsf = a.eps_f;
    const double epsd = a.eps_d;

    // LOC: origin = centre of the targets' bounding box (the same on every
    // thread), targets as FP32 offsets from it
    double ox = 0.0, oy = 0.0, oz = 0.0;
    float fx[R], fy[R], fz[R];
    float hx = 0.f, hy = 0.f, hz = 0.f, near2 = 0.f;   // LOC: target box and far test
    if (LOC) {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (it.i0 + r * L + lane < it.i1) {
                mn[0] = fmin(mn[0], px[r]); mx[0] = fmax(mx[0], px[r]);
                mn[1] = fmin(mn[1], py[r]); mx[1] = fmax(mx[1], py[r]);
                mn[2] = fmin(mn[2], pz[r]); mx[2] = fmax(mx[2], pz[r]);
            }
        }
        const int wave = tid >> 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = wave_min(mn[k]);
            mx[k] = wave_max(mx[k]);
        }
        if ((tid & 63) == 0) {
            lx[wave] = mn[0]; ly[wave] = mn[1]; lz[wave] = mn[2];
            lx[4 + wave] = mx[0]; ly[4 + wave] = mx[1]; lz[4 + wave] = mx[2];
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) {
            mn[0] = fmin(mn[0], lx[w]); mx[0] = fmax(mx[0], lx[4 + w]);
            mn[1] = fmin(mn[1], ly[w]); mx[1] = fmax(mx[1], ly[4 + w]);
            mn[2] = fmin(mn[2], lz[w]); mx[2] = fmax(mx[2], lz[4 + w]);
        }
        ox = 0.5 * (mn[0] + mx[0]);
        oy = 0.5 * (mn[1] + mx[1]);
        oz = 0.5 * (mn[2] + mx[2]);
        hx = (float)(0.5 * (mx[0] - mn[0]));
        hy = (float)(0.5 * (mx[1] - mn[1]));
        hz = (float)(0.5 * (mx[2] - mn[2]));
        near2 = kNearFactor * kNearFactor * (hx * hx + hy * hy + hz * hz);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            fx[r] = (float)(px[r] - ox);
            fy[r] = (float)(py[r] - oy);
            fz[r] = (float)(pz[r] - oz);
        }
    }

    for (int t0 = it.j0; t0 < it.j1; t0 += kBlock) {
        const int cnt = min(kBlock, it.j1 - t0);
        __syncthreads();
        if (LOC) {
            // offsets; wave w stages sub-tile w, tests each source and stores
            // the sub-tile's far sources first: a far source at its rank among
            // the far lanes below it, a near one after all far ones.  Valid
            // lanes are a prefix of the wave, so the valid sources fill
            // [0, valid count) of the sub-tile.
            bool far_src = false;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            float4 s4 = make_float4(0.f, 0.f, 0.f, 1.f);
            if (tid < cnt) {
                sx = a.sx[t0 + tid];
                sy = a.sy[t0 + tid];
                sz = a.sz[t0 + tid];
                s4 = make_float4((float)(sx - ox), (float)(sy - oy), (float)(sz - oz),
                                 UNIT ? 1.f : a.sqf[t0 + tid]);
                far_src = far_from_box(s4.x, s4.y, s4.z, hx, hy, hz, near2);
            }
            const unsigned long long far_mask = __ballot(far_src);
            const int nfar = __popcll(far_mask);
            const int lane64 = tid & 63;
            const int far_below = (int)__builtin_amdgcn_mbcnt_hi(
                (unsigned)(far_mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)far_mask, 0u));
            const int slot = (tid & ~63) + (far_src ? far_below : nfar + (lane64 - far_below));
            if (tid < cnt) {
                lx[slot] = sx;
                ly[slot] = sy;
                lz[slot] = sz;
                lf[slot] = s4;
            }
            if (lane64 == 0) lfar[tid >> 6] = nfar;
        } else if (tid < cnt) {
            lx[tid] = a.sx[t0 + tid];
            ly[tid] = a.sy[t0 + tid];
            lz[tid] = a.sz[t0 + tid];
            if (F64)
                lqd[tid] = a.sqd[t0 + tid];
            else if (!UNIT)
                lqf[tid] = a.sqf[t0 + tid];
        }
        __syncthreads();
        // this wave group's share of the tile
        const int b0 = grp * SPAN, b1 = min(cnt, b0 + SPAN);

        if (F64) {
            for (int j = b0; j < b1; ++j) {
                const double sx = lx[j], sy = ly[j], sz = lz[j], q = lqd[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double dx = sx - px[r], dy = sy - py[r], dz = sz - pz[r];
                    const double d2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, epsd)));
                    double w;
                    if (MODE == KP_FP64_RSQ) {
                        w = q * rsq3_f64(d2);
                    } else {
                        const double inv = 1.0 / sqrt(d2);
                        w = q * (inv * inv) * inv;
                    }
                    Ax[r] = fma(dx, w, Ax[r]);
                    Ay[r] = fma(dy, w, Ay[r]);
                    Az[r] = fma(dz, w, Az[r]);
                }
            }
        } else {
            for (int jj = b0; jj < b1; jj += kSubTile) {
                float ax[R], ay[R], az[R];
#pragma unroll
                for (int r = 0; r < R; ++r) ax[r] = ay[r] = az[r] = 0.f;
                const int m = min(kSubTile, b1 - jj);
                auto body = [&](int j, auto far_tag) {
                    constexpr bool FAR = decltype(far_tag)::value;
                    double sx = 0.0, sy = 0.0, sz = 0.0;
                    float4 s = make_float4(0.f, 0.f, 0.f, 1.f);
                    if (FAR) {
                        s = lf[j];
                    } else {
                        sx = lx[j]; sy = ly[j]; sz = lz[j];
                        if (!UNIT) s.w = LOC ? lf[j].w : lqf[j];
                    }
                    const float q = s.w;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const float dx = FAR ? s.x - fx[r] : (float)(sx - px[r]);
                        const float dy = FAR ? s.y - fy[r] : (float)(sy - py[r]);
                        const float dz = FAR ? s.z - fz[r] : (float)(sz - pz[r]);
                        float d2 = fmaf(dz, dz, epsf);
                        d2 = fmaf(dy, dy, d2);
                        d2 = fmaf(dx, dx, d2);
                        const float inv = __builtin_amdgcn_rsqf(d2);
                        // UNIT: q == 1 exactly, and (1*inv)*(inv*inv) == inv*(inv*inv)
                        const float w = UNIT ? inv * (inv * inv) : (q * inv) * (inv * inv);
                        ax[r] = fmaf(dx, w, ax[r]);
                        ay[r] = fmaf(dy, w, ay[r]);
                        az[r] = fmaf(dz, w, az[r]);
                    }
                };
                // sources [j0, j1) of the sub-tile on one path
                auto run = [&](int j0, int j1, auto far_tag) {
                    constexpr int U = 8;
                    if (j1 - j0 == kSubTile) {
#pragma unroll U
                        for (int j = 0; j < kSubTile; ++j) body(j0 + j, far_tag);
                    } else {
                        int j = j0;
                        for (; j + U <= j1; j += U) {
#pragma unroll
                            for (int k = 0; k < U; ++k) body(j + k, far_tag);
                        }
                        for (; j < j1; ++j) body(j, far_tag);
                    }
                };
                if constexpr (LOC) {
                    // far sources first (nfar <= m: both count the sub-tile's valid sources)
                    const int nfar = __builtin_amdgcn_readfirstlane(lfar[jj / kSubTile]);
                    if (nfar > 0) run(jj, jj + nfar, std::true_type{});
                    if (nfar < m) run(jj + nfar, jj + m, std::false_type{});
                } else {
                    run(jj, jj + m, std::false_type{});
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (MODE == KP_MIXED_KAHAN) {
                        float y, t;
                        y = ax[r] - Cx[r]; t = Sx[r] + y; Cx[r] = (t - Sx[r]) - y; Sx[r] = t;
                        y = ay[r] - Cy[r]; t = Sy[r] + y; Cy[r] = (t - Sy[r]) - y; Sy[r] = t;
                        y = az[r] - Cz[r]; t = Sz[r] + y; Cz[r] = (t - Sz[r]) - y; Sz[r] = t;
                    } else {
                        Ax[r] += (double)ax[r];
                        Ay[r] += (double)ay[r];
                        Az[r] += (double)az[r];
                    }
                }
            }
        }
    }

    if (MODE == KP_MIXED_KAHAN) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Ax[r] = (double)Sx[r] - (double)Cx[r];
            Ay[r] = (double)Sy[r] - (double)Cy[r];
            Az[r] = (double)Sz[r] - (double)Cz[r];
        }
    }
    if (Q > 1) {
        // combine the wave groups in a fixed order (group 0 first), one target
        // set at a time through the (now free) source tile arrays
#pragma unroll
        for (int r = 0; r < R; ++r) {
            __syncthreads();
            lx[tid] = Ax[r];
            ly[tid] = Ay[r];
            lz[tid] = Az[r];
            __syncthreads();
            if (grp == 0) {
                for (int q = 1; q < Q; ++q) {
                    Ax[r] += lx[q * L + lane];
                    Ay[r] += ly[q * L + lane];
                    Az[r] += lz[q * L + lane];
                }
            }
        }
        if (grp != 0) return;
    }

#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int g = it.i0 + r * L + lane;
        if (g < it.i1) {
            const double k = a.kc ? a.kc[g] : 1.0;
            double *o = a.out + (long long)it.slice * 3 * a.stride + g;
            o[0] = k * Ax[r];
            o[a.stride] = k * Ay[r];
            o[2 * a.stride] = k * Az[r];
        }
    }
}


// ---------------------------------------------------------------------------
// Coulomb direct sum, streamed (mixed modes).  Same work items, targets,
// pair arithmetic, local coordinates and accumulation as coulomb_tiled with
// Q = 4, but the four waves of a work-group never wait for one another
// inside the source loop: every wave holds all 64 R targets of the item
// (target r*64 + lane), computes the same target box by itself, and streams
// its own sub-tiles (wave w takes sub-tiles w, w+4, w+8, ... of the slice)
// through a private LDS region, loading the next sub-tile into registers
// while it sums the current one.  The only barrier is the fixed-order
// combination of the four waves' sums at the end.  A wave that has more
// near sources therefore no longer holds the others back at every tile, and
// the global-load latency of the next sub-tile overlaps the sums.
// ---------------------------------------------------------------------------
// LDS of the streamed kernel: one region per work-group, whichever
// instantiation of coulomb_stream_item runs the item (a non-template holder,
// so the instantiations one kernel calls share it).
union StreamLds {
    struct {
        double x[kBlock], y[kBlock], z[kBlock];   // FP64 sources (near path); reused
        float4 f[kBlock];                         // (offset or 0, weight)
    } tile;
    double sums[4][kBlock / 64][3][64];           // the streams' sums: [set][wave][component][lane]
};
__device__ __forceinline__ StreamLds &stream_lds()
{
    __shared__ StreamLds s;
    return s;
}

template <int R, int MODE, bool UNIT, bool LOC>
__device__ __forceinline__ void coulomb_stream_item(const CoulombArgs &a, const CoulombItem &it)
{
    static_assert(MODE == KP_MIXED || MODE == KP_MIXED_KAHAN, "a mixed-precision kernel");
    constexpr int NW = kBlock / 64;      // waves per work-group: 4 source streams
    constexpr int W = 64;                // sources per sub-tile = lanes per wave
    static_assert(R <= 4, "one LDS region of sums per target set");
    StreamLds &lds = stream_lds();
    double *const lx = lds.tile.x, *const ly = lds.tile.y, *const lz = lds.tile.z;
    float4 *const lf = lds.tile.f;

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    double *const wx = lx + wave * W;
    double *const wy = ly + wave * W;
    double *const wz = lz + wave * W;
    float4 *const wf = lf + wave * W;

    double px[R], py[R], pz[R];
    double Ax[R], Ay[R], Az[R];
    float Sx[R], Sy[R], Sz[R], Cx[R], Cy[R], Cz[R];
    float fx[R], fy[R], fz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int g = it.i0 + r * W + lane;
        const bool v = g < it.i1;
        px[r] = v ? a.tx[g] : 0.0;
        py[r] = v ? a.ty[g] : 0.0;
        pz[r] = v ? a.tz[g] : 0.0;
        Ax[r] = Ay[r] = Az[r] = 0.0;
        Sx[r] = Sy[r] = Sz[r] = 0.f;
        Cx[r] = Cy[r] = Cz[r] = 0.f;
        fx[r] = fy[r] = fz[r] = 0.f;
    }
    const float epsf = a.eps_f;

    // LOC: origin = centre of the targets' bounding box, reduced inside the
    // wave (each wave holds every target, so all four find the same box)
    double ox = 0.0, oy = 0.0, oz = 0.0;
    float hx = 0.f, hy = 0.f, hz = 0.f, near2 = 0.f;
    if (LOC) {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (it.i0 + r * W + lane < it.i1) {
                mn[0] = fmin(mn[0], px[r]); mx[0] = fmax(mx[0], px[r]);
                mn[1] = fmin(mn[1], py[r]); mx[1] = fmax(mx[1], py[r]);
                mn[2] = fmin(mn[2], pz[r]); mx[2] = fmax(mx[2], pz[r]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = wave_min(mn[k]);
            mx[k] = wave_max(mx[k]);
        }
        ox = 0.5 * (mn[0] + mx[0]);
        oy = 0.5 * (mn[1] + mx[1]);
        oz = 0.5 * (mn[2] + mx[2]);
        hx = (float)(0.5 * (mx[0] - mn[0]));
        hy = (float)(0.5 * (mx[1] - mn[1]));
        hz = (float)(0.5 * (mx[2] - mn[2]));
        near2 = kNearFactor * kNearFactor * (hx * hx + hy * hy + hz * hz);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            fx[r] = (float)(px[r] - ox);
            fy[r] = (float)(py[r] - oy);
            fz[r] = (float)(pz[r] - oz);
        }
    }

    // this wave's first sub-tile, loaded ahead
    int base = it.j0 + wave * W;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    float nq = 1.f;
    if (base + lane < it.j1) {
        nx = a.sx[base + lane];
        ny = a.sy[base + lane];
        nz = a.sz[base + lane];
        if (!UNIT) nq = a.sqf[base + lane];
    }
    for (; base < it.j1; base += NW * W) {
        const int m = min(W, it.j1 - base);
        const double sx = nx, sy = ny, sz = nz;
        const float sq = nq;
        {
            const int jn = base + NW * W + lane;
            if (jn < it.j1) {
                nx = a.sx[jn];
                ny = a.sy[jn];
                nz = a.sz[jn];
                if (!UNIT) nq = a.sqf[jn];
            }
        }
        // every lane's reads of the previous sub-tile precede these writes:
        // LDS instructions of one wave complete in order, and the fence keeps
        // the compiler from moving the writes above the reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const bool valid = lane < m;
        int nfar = 0;
        if (LOC) {
            // far sources first: a far source at its rank among the far lanes
            // below it, a near one after all far ones (valid lanes are a prefix)
            float4 s4 = make_float4(0.f, 0.f, 0.f, 1.f);
            bool far_src = false;
            if (valid) {
                s4 = make_float4((float)(sx - ox), (float)(sy - oy), (float)(sz - oz), UNIT ? 1.f : sq);
                far_src = far_from_box(s4.x, s4.y, s4.z, hx, hy, hz, near2);
            }
            const unsigned long long far_mask = __ballot(far_src);
            nfar = __popcll(far_mask);
            const int far_below = (int)__builtin_amdgcn_mbcnt_hi(
                (unsigned)(far_mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)far_mask, 0u));
            const int slot = far_src ? far_below : nfar + (lane - far_below);
            if (valid) {
                wx[slot] = sx;
                wy[slot] = sy;
                wz[slot] = sz;
                wf[slot] = s4;
            }
        } else if (valid) {
            wx[lane] = sx;
            wy[lane] = sy;
            wz[lane] = sz;
            if (!UNIT) wf[lane] = make_float4(0.f, 0.f, 0.f, sq);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        float ax[R], ay[R], az[R];
#pragma unroll
        for (int r = 0; r < R; ++r) ax[r] = ay[r] = az[r] = 0.f;
        auto body = [&](int j, auto far_tag) {
            constexpr bool FAR = decltype(far_tag)::value;
            double ssx = 0.0, ssy = 0.0, ssz = 0.0;
            float4 s = make_float4(0.f, 0.f, 0.f, 1.f);
            if (FAR) {
                s = wf[j];
            } else {
                ssx = wx[j]; ssy = wy[j]; ssz = wz[j];
                if (!UNIT) s.w = wf[j].w;
            }
            const float q = s.w;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = FAR ? s.x - fx[r] : (float)(ssx - px[r]);
                const float dy = FAR ? s.y - fy[r] : (float)(ssy - py[r]);
                const float dz = FAR ? s.z - fz[r] : (float)(ssz - pz[r]);
                float d2 = fmaf(dz, dz, epsf);
                d2 = fmaf(dy, dy, d2);
                d2 = fmaf(dx, dx, d2);
                const float inv = __builtin_amdgcn_rsqf(d2);
                const float w = UNIT ? inv * (inv * inv) : (q * inv) * (inv * inv);
                ax[r] = fmaf(dx, w, ax[r]);
                ay[r] = fmaf(dy, w, ay[r]);
                az[r] = fmaf(dz, w, az[r]);
            }
        };
        auto run = [&](int j0, int j1, auto far_tag) {
            constexpr int U = 8;
            if (j1 - j0 == W) {
#pragma unroll U
                for (int j = 0; j < W; ++j) body(j0 + j, far_tag);
            } else {
                int j = j0;
                for (; j + U <= j1; j += U) {
#pragma unroll
                    for (int k = 0; k < U; ++k) body(j + k, far_tag);
                }
                for (; j < j1; ++j) body(j, far_tag);
            }
        };
        if constexpr (LOC) {
            const int nf = __builtin_amdgcn_readfirstlane(nfar);
            if (nf > 0) run(0, nf, std::true_type{});
            if (nf < m) run(nf, m, std::false_type{});
        } else {
            run(0, m, std::false_type{});
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (MODE == KP_MIXED_KAHAN) {
                float y, t;
                y = ax[r] - Cx[r]; t = Sx[r] + y; Cx[r] = (t - Sx[r]) - y; Sx[r] = t;
                y = ay[r] - Cy[r]; t = Sy[r] + y; Cy[r] = (t - Sy[r]) - y; Sy[r] = t;
                y = az[r] - Cz[r]; t = Sz[r] + y; Cz[r] = (t - Sz[r]) - y; Sz[r] = t;
            } else {
                Ax[r] += (double)ax[r];
                Ay[r] += (double)ay[r];
                Az[r] += (double)az[r];
            }
        }
    }

    if (MODE == KP_MIXED_KAHAN) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            Ax[r] = (double)Sx[r] - (double)Cx[r];
            Ay[r] = (double)Sy[r] - (double)Cy[r];
            Az[r] = (double)Sz[r] - (double)Cz[r];
        }
    }
    // the four streams' sums of every target set in LDS at once (one barrier
    // after the last sub-tile, one after the stores); wave r then adds set r's
    // in the fixed order wave 0, 1, 2, 3 and writes it
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        lds.sums[r][wave][0][lane] = Ax[r];
        lds.sums[r][wave][1][lane] = Ay[r];
        lds.sums[r][wave][2][lane] = Az[r];
    }
    __syncthreads();
    if (wave >= R) return;
    const int g = it.i0 + wave * W + lane;
    if (g < it.i1) {
        double s[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s[c] = lds.sums[wave][0][c][lane];
#pragma unroll
            for (int q = 1; q < NW; ++q) s[c] += lds.sums[wave][q][c][lane];
        }
        const double k = a.kc ? a.kc[g] : 1.0;
        double *o = a.out + (long long)it.slice * 3 * a.stride + g;
        o[0] = k * s[0];
        o[a.stride] = k * s[1];
        o[2 * a.stride] = k * s[2];
    }
}


// A streamed item with R target sets per thread: the last block of a cloud
// may fill fewer than R sets of 64 targets, and the instantiation for the
// sets it fills runs it, so the empty sets cost nothing (the sums of the
// filled ones are the same, bit for bit).
template <int R, int MODE, bool UNIT, bool LOC>
__device__ __forceinline__ void coulomb_stream_sets(const CoulombArgs &a, int item)
{
    const CoulombItem it = a.items[item];
    // uniform over the work-group (one item), so either every wave reaches
    // the closing barrier or none does; an empty item pads an XCD's run
    if (it.i0 >= it.i1) return;
    if (a.state && (!a.state[it.cloud].active || (a.skip_valid_k1 && a.state[it.cloud].k1_valid)))
        return;
    const int sets = (it.i1 - it.i0 + 63) >> 6;
    if constexpr (R >= 2) {
        if (sets <= 1) return coulomb_stream_item<1, MODE, UNIT, LOC>(a, it);
    }
    if constexpr (R >= 4) {
        if (sets <= 2) return coulomb_stream_item<2, MODE, UNIT, LOC>(a, it);
        if (sets <= 3) return coulomb_stream_item<3, MODE, UNIT, LOC>(a, it);
    }
    coulomb_stream_item<R, MODE, UNIT, LOC>(a, it);
}


// ---------------------------------------------------------------------------
// Dormand-Prince 5(4) tableau, bin/ode.cpp:76-86.  kA[s][j] is the
// coefficient of k_j in the input of stage s (kA[7] = b), kE the error
// weights e = b - b*.  Zero entries are skipped, as the reference's
// expressions omit them.
// ---------------------------------------------------------------------------
__device__ constexpr double kC[8] = {0, 0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
__device__ constexpr double kA[8][8] = {
    {0}, {0},
    {0, 1.0 / 5.0},
    {0, 3.0 / 40.0, 9.0 / 40.0},
    {0, 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
    {0, 19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
    {0, 9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
    {0, 35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0},
};
__device__ constexpr double kE[8] = {0, 71.0 / 57600.0, 0.0, -71.0 / 16695.0, 71.0 / 1920.0,
                                     -17253.0 / 339200.0, 22.0 / 525.0, -1.0 / 40.0};

// Sum_{j=1..S} coef[j]*k_j evaluated left to right over the non-zero terms,
// where k_S is `own` and k_j (j<S) is read from K[j][comp][g].
template <int S, int NEXT>
__device__ __forceinline__ double combo(double *const (*K)[3], int comp, long long g, double own)
{
    double acc = 0.0;
    bool first = true;
#pragma unroll
    for (int j = 1; j <= S; ++j) {
        const double c = kA[NEXT][j];
        if (c == 0.0) continue;
        const double kj = (j == S) ? own : K[j][comp][g];
        const double term = c * kj;
        acc = first ? term : acc + term;
        first = false;
    }
    return acc;
}

// SWIFT amplitude at t: samples at t0 + k dt, linear interpolation, 0 outside
__device__ __forceinline__ double swift_amp(const CloudParams &p, double t)
{
    const double s = (t - p.swift_t0) / p.swift_dt;
    if (!(s >= 0.0) || s > (double)(p.swift_n - 1)) return 0.0;
    int i = (int)s;
    if (i > p.swift_n - 2) i = p.swift_n - 2;
    const double w = s - (double)i;
    const double a = p.swift[i], b = p.swift[i + 1];
    return a + w * (b - a);
}

// Axisymmetric field map at (r, z): bilinear interpolation of the (E_r, E_z)
// grid, 0 outside it.  The cell index is clamped to the last cell, so the
// four nodes read are always inside the map.
__device__ __forceinline__ void emap_field(const CloudParams &p, double r, double z, double &er,
                                           double &ez)
{
    er = 0.0;
    ez = 0.0;
    const double u = r / p.emap_dr, v = (z - p.emap_z0) / p.emap_dz;
    if (!(u >= 0.0) || u > (double)(p.emap_nr - 1) || !(v >= 0.0) || v > (double)(p.emap_nz - 1))
        return;
    int i = (int)u, j = (int)v;
    if (i > p.emap_nr - 2) i = p.emap_nr - 2;
    if (j > p.emap_nz - 2) j = p.emap_nz - 2;
    const double wu = u - (double)i, wv = v - (double)j;
    const long long nz = p.emap_nz;
    const double *e00 = p.emap + 2 * ((long long)i * nz + j);
    const double *e01 = e00 + 2, *e10 = e00 + 2 * nz, *e11 = e10 + 2;
    double out[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double lo = e00[k] + wu * (e10[k] - e00[k]);
        const double hi = e01[k] + wu * (e11[k] - e01[k]);
        out[k] = lo + wv * (hi - lo);
    }
    er = out[0];
    ez = out[1];
}

// force() at one ion (bin/force.cpp:58-105 order), then the extension terms,
// then the Coulomb subtraction.  t_life is IonCloud::lifetime (drives the
// trap ramp, bin/ion_fly.cpp:235-239); ts is the field evaluation time.
__device__ __forceinline__ void ion_accel(const CloudParams &p, double t_life, double ts,
                                          double qm, double x, double y, double z, double vx,
                                          double vy, double vz, double Fx, double Fy, double Fz,
                                          double &ax, double &ay, double &az)
{
    ax = 0;
    ay = 0;
    az = 0;
    if (p.trap) {
        double Vrf = p.Vrf, Vdc = p.Vdc;
        if (p.ramp) {
            const double ramp = (t_life - p.ramp_t0) / (p.ramp_t1 - p.ramp_t0);
            Vrf = p.Vrf * (1 - ramp) + p.newVrf * ramp;
            Vdc = p.Vdc * (1 - ramp) + p.newVdc * ramp;
        }
        const double kVdc = p.kappa * Vdc;
        const double crf = cos(p.omega_rf * ts);
        ax += qm * (kVdc / p.z02 - Vrf * crf / p.r02) * x;
        ay += qm * (kVdc / p.z02 + Vrf * crf / p.r02) * y;
        az += qm * (-2.0 * kVdc / p.z02) * z;
    }
    if (p.tof) ax += qm * p.E_kick;
    if (p.cpen != 0.0) {
        ax += qm * p.cpen * x;
        ay += qm * p.cpen * y;
        az += qm * (-2.0 * p.cpen) * z;
    }
    if (p.Bz != 0.0) {
        ax += qm * p.Bz * vy;
        ay -= qm * p.Bz * vx;
    }
    // drive phases 2 pi (f t + rate t^2 / 2) + phi; rate = 0 adds an exact 0
    if (p.Ed != 0.0)
        ax += qm * (p.Ed * cos(2.0 * kPi * p.fd * ts + kPi * p.fd_rate * ts * ts + p.phid));
    if (p.Uq_rq2 != 0.0) {
        const double cq = p.Uq_rq2 * cos(2.0 * kPi * p.fq * ts + kPi * p.fq_rate * ts * ts + p.phiq);
        ax += qm * cq * x;
        ay -= qm * cq * y;
    }
    if (p.delta != 0.0) {
        ax -= p.delta * vx;
        ay -= p.delta * vy;
        az -= p.delta * vz;
    }
    if (p.B2 != 0.0) {
        // magnetic bottle B2 ((z^2 - (x^2+y^2)/2) z_hat - z (x, y, 0)); a += (q e/m) v x B
        const double b2z = p.B2 * z;
        const double bx = -b2z * x, by = -b2z * y;
        const double bz = p.B2 * (z * z - 0.5 * (x * x + y * y));
        ax += qm * (vy * bz - vz * by);
        ay += qm * (vz * bx - vx * bz);
        az += qm * (vx * by - vy * bx);
    }
    if (p.Uo_ro4 != 0.0) {
        // octupole drive: -grad of -(Uo/(4 ro^4)) cos(psi) Re (x+iy)^4
        const double co = p.Uo_ro4 * cos(2.0 * kPi * p.fo * ts + p.phio);
        ax += qm * co * (x * (x * x - 3.0 * y * y));
        ay += qm * co * (y * (y * y - 3.0 * x * x));
    }
    if (p.Uw_rwl != 0.0) {
        // rotating wall of order 1 (uniform field) or 2 (quadrupole), turning at fw
        const double psi = 2.0 * kPi * p.fw * ts + p.phiw;
        const double ang = (p.rw_order == 1) ? psi : 2.0 * psi;
        const double cw = p.Uw_rwl * cos(ang), sw = p.Uw_rwl * sin(ang);
        if (p.rw_order == 1) {
            ax += qm * cw;
            ay += qm * sw;
        } else {
            ax += qm * (cw * x + sw * y);
            ay += qm * (sw * x - cw * y);
        }
    }
    if (p.Eax != 0.0) az += qm * (p.Eax * cos(2.0 * kPi * p.fax * ts + p.phiax));
    if (p.inv_K != 0.0) {
        // buffer gas by mobility: a -= (q e/m) v / K
        const double gk = qm * p.inv_K;
        ax -= gk * vx;
        ay -= gk * vy;
        az -= gk * vz;
    }
    if (p.swift_n > 0) {
        // SWIFT: sampled waveform as a dipole along x or a quadrupole
        const double u = swift_amp(p, ts);
        if (p.swift_kind == 1) {
            ax += qm * (u / p.swift_r);
        } else {
            const double cs = u / (p.swift_r * p.swift_r);
            ax += qm * cs * x;
            ay -= qm * cs * y;
        }
    }
    if (p.c4k != 0.0 || p.c6k != 0.0) {
        // anharmonic trap: -grad of c4k (z^4 - 3 z^2 r^2 + 3 r^4/8) and
        // c6k (z^6 - 15/2 z^4 r^2 + 45/8 z^2 r^4 - 5/16 r^6)
        const double r2 = x * x + y * y, z2 = z * z;
        if (p.c4k != 0.0) {
            const double cr = p.c4k * (6.0 * z2 - 1.5 * r2);
            const double cz = p.c4k * (6.0 * r2 - 4.0 * z2);
            ax += qm * cr * x;
            ay += qm * cr * y;
            az += qm * cz * z;
        }
        if (p.c6k != 0.0) {
            const double z4 = z2 * z2, r4 = r2 * r2, zr = z2 * r2;
            const double cr = p.c6k * (15.0 * z4 - 22.5 * zr + 1.875 * r4);
            const double cz = p.c6k * (30.0 * zr - 6.0 * z4 - 11.25 * r4);
            ax += qm * cr * x;
            ay += qm * cr * y;
            az += qm * cz * z;
        }
    }
    if (p.emap_nr > 0) {
        const double r = sqrt(x * x + y * y);
        double er, ez;
        emap_field(p, r, z, er, ez);
        const double g = qm * p.emap_scale;
        if (r > 0.0) {
            ax += g * (er * (x / r));
            ay += g * (er * (y / r));
        }
        az += g * ez;
    }
    ax -= Fx;
    ay -= Fy;
    az -= Fz;
}

__device__ __forceinline__ void coulomb_total(const EngineDev &d, const CloudParams &p, int c,
                                              long long g, double &Fx, double &Fy, double &Fz)
{
    Fx = Fy = Fz = 0.0;
    if (!p.coulomb) return;
    const int ns = d.slices[c];
    const long long st = d.stride;
    Fx = d.CP[g];
    Fy = d.CP[st + g];
    Fz = d.CP[2 * st + g];
    // unrolled so that the loads of several slices are in flight at once;
    // the sum keeps its slice order
#pragma unroll 8
    for (int s = 1; s < ns; ++s) {
        const double *cp = d.CP + (long long)s * 3 * st + g;
        Fx += cp[0];
        Fy += cp[st];
        Fz += cp[2 * st];
    }
}

// Stage S of one attempt for every active ion: derivatives at the stage
// state, then the next stage state (bin/ode.cpp:103-175), or for S == 7 the
// new state and its scaled error (bin/ode.cpp:177-205).
template <int S>
__device__ __forceinline__ void stage_block(const EngineDev &d, int block)
{
    const IonBlock blk = d.blocks[block];
    const int c = blk.cloud;
    const CloudState &cs = d.state[c];
    if (!cs.active) return;                       // uniform over the block
    const bool valid = blk.start + (int)threadIdx.x < blk.end;
    const long long g = valid ? blk.start + threadIdx.x : blk.start;
    const CloudParams &p = d.params[c];
    const double h = cs.h, t = cs.t;
    const double ts = d.time_mode ? t + kC[S] * h : t;

    double *const *pos_s = (S == 1) ? d.P0 : d.PS;
    const double x = pos_s[0][g], y = pos_s[1][g], z = pos_s[2][g];
    const double vx = d.KV[S][0][g], vy = d.KV[S][1][g], vz = d.KV[S][2][g];
    double a3[3];
    if (S == 1 && cs.k1_valid) {
        // derivative at (y_n, t_n) kept from the previous attempt
#pragma unroll
        for (int k = 0; k < 3; ++k) a3[k] = d.KA[1][k][g];
    } else {
        double Fx, Fy, Fz;
        coulomb_total(d, p, c, g, Fx, Fy, Fz);
        ion_accel(p, t, ts, d.qm[g], x, y, z, vx, vy, vz, Fx, Fy, Fz, a3[0], a3[1], a3[2]);
        if (valid) {
#pragma unroll
            for (int k = 0; k < 3; ++k) d.KA[S][k][g] = a3[k];
        }
    }
    const double v3[3] = {vx, vy, vz};
    const double s3[3] = {x, y, z};

    if constexpr (S < 7) {
        if (!valid) return;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            d.PS[k][g] = d.P0[k][g] + h * combo<S, S + 1>(d.KV, k, g, v3[k]);
            d.KV[S + 1][k][g] = d.V0[k][g] + h * combo<S, S + 1>(d.KA, k, g, a3[k]);
        }
    } else {
        __shared__ double red[256];
        // y_{n+1} equals the stage-7 state: a_7j = b_j, same evaluation order.
        double err = 0.0;
        bool first = true;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            double *const(*K)[3] = half ? d.KA : d.KV;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double own = half ? a3[k] : v3[k];
                double acc = 0.0;
                bool f1 = true;
#pragma unroll
                for (int j = 1; j <= 7; ++j) {
                    if (kE[j] == 0.0) continue;
                    const double kj = (j == 7) ? own : K[j][k][g];
                    const double term = kE[j] * kj;
                    acc = f1 ? term : acc + term;
                    f1 = false;
                }
                const double yerr = h * acc;
                const double old = half ? d.V0[k][g] : d.P0[k][g];
                const double neu = half ? v3[k] : s3[k];
                const double sk = d.atol + d.rtol * fmax(fabs(old), fabs(neu));
                const double term = (yerr / sk) * (yerr / sk);
                err = first ? term : err + term;
                first = false;
            }
        }
        red[threadIdx.x] = valid ? err : 0.0;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) d.ERRB[block] = red[0];
    }
}


// Per-cloud error norm (bin/ode.cpp:215), controller (bin/ode.cpp:426-457)
// and the step bookkeeping of bin/ode.cpp:470-493.  One work-group per cloud;
// the reduction order is fixed, so results are run-to-run reproducible.
__device__ __forceinline__ void controller_cloud(const EngineDev &d, int c)
{
    __shared__ double red[256];
    CloudState &cs = d.state[c];
    const int active = cs.active;
    if (!active) {
        if (threadIdx.x == 0) cs.accepted_last = 0;
        return;
    }
    const CloudParams &p = d.params[c];
    double sum = 0.0;
    for (int b = threadIdx.x; b < p.nblk; b += 256) sum += d.ERRB[p.blk0 + b];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double err = sqrt(red[0] / (6 * p.n));
    cs.last_err = err;
    cs.attempts++;
    cs.evals += cs.k1_valid ? 6 : 7;
    int accept = 1;
    if (d.adaptive) {
        const double beta = d.beta, alpha = 0.2 - beta * 0.75, safe = 0.9, minscale = 0.2,
                     maxscale = 10.0;
        double scale;
        if (err <= 1.0) {
            if (err == 0.0)
                scale = maxscale;
            else {
                scale = safe * pow(err, -alpha) * pow(cs.errold, beta);
                if (scale < minscale) scale = minscale;
                if (scale > maxscale) scale = maxscale;
            }
            if (cs.reject)
                cs.hnext = cs.h * fmin(scale, 1.0);
            else
                cs.hnext = cs.h * scale;
            cs.errold = fmax(err, 1.0e-4);
            cs.reject = 0;
        } else {
            scale = fmax(safe * pow(err, -alpha), minscale);
            cs.h *= scale;
            cs.reject = 1;
            accept = 0;
            if (fabs(cs.h) <= fabs(d.P0[0][p.ion0]) * DBL_EPSILON) {
                cs.error = 1;
                cs.active = 0;
            }
        }
    }
    cs.accepted_last = accept;
    // after a reject the next attempt restarts from the same (y_n, t_n): its
    // stage-1 derivative is unchanged; after an accept it is the stage-7 one
    // when that is exact (commit_kernel copies it)
    cs.k1_valid = accept ? p.fsal : 1;
    if (accept) {
        cs.t_prev = cs.t;
        cs.t = cs.t + cs.h;
        cs.h = cs.hnext;
        cs.steps++;
        if (cs.budget > 0) cs.budget--;
        cs.active = (cs.t < cs.t_end) && cs.budget != 0;
    }
}


// pos <- pos2, vel <- vel2 for clouds whose attempt was accepted
// (bin/ode.cpp:488-489).
// With a crossing plane, an ion whose x lies strictly on one side of it
// before the step and strictly on the other after it (bin/ion_fly.cpp:91-93,
// 110-114) is appended to the crossing list with its new state.
__device__ __forceinline__ void commit_block(const EngineDev &d, int block)
{
    const IonBlock blk = d.blocks[block];
    const CloudState &cs = d.state[blk.cloud];
    if (!cs.accepted_last) return;
    const long long g = blk.start + threadIdx.x;
    if (g >= blk.end) return;
    const CloudParams &p = d.params[blk.cloud];
    if (p.xcross_on) {
        const double x0 = d.P0[0][g], x1 = d.PS[0][g], xp = p.xcross;
        if ((x0 < xp && x1 > xp) || (x0 > xp && x1 < xp)) {
            const int i = atomicAdd(d.xc_count, 1);
            if (i < d.xc_cap) {
                XCrossing &c = d.xc[i];
                c.step = cs.steps;
                c.cloud = blk.cloud;
                c.slot = (int)g;
                c.t = cs.t;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    c.p[k] = d.PS[k][g];
                    c.v[k] = d.KV[7][k][g];
                }
            }
        }
    }
    const bool fsal = p.fsal;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d.P0[k][g] = d.PS[k][g];
        d.V0[k][g] = d.KV[7][k][g];
        if (fsal) d.KA[1][k][g] = d.KA[7][k][g];
    }
}



// ---------------------------------------------------------------------------
// Persistent attempts.  A cooperative launch (every work-group resident at
// once) runs whole attempts with a grid barrier between the phases that read
// what another work-group wrote: Coulomb items -> stage s (7 times), then the
// controller and the commit.  Each phase deals its items (Coulomb work items,
// ion blocks, clouds) to the work-groups round-robin and calls the same
// device functions the one-phase kernels call, so results are the graph
// path's bit for bit.  It replaces the 16 launches of an attempt with one
// launch per batch, but on MI355X a grid barrier (about 2 us at 16
// work-groups, 13 us at 512: the cross-XCD release/acquire and the arrivals
// on one address) costs more than a graph's kernel boundary (1.7 us), so the
// engine uses it only on request.
//
// The barrier: thread 0 of each work-group reads the generation, adds itself
// to the arrivals (agent scope, release) and either, as the last one,
// acquires, resets the arrivals and bumps the generation (release), or polls
// the generation with relaxed loads and acquires once it moved.  Every wait
// is bounded: past spin_limit it sets the timeout flag, which makes every
// later barrier return false at once, so all work-groups leave and the
// kernel ends; the host reports the error.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool grid_barrier(const PersistArgs &p)
{
    __shared__ int s_ok;
    __syncthreads();   // the work-group's writes of this phase are done
    if (threadIdx.x == 0) {
        int ok = __hip_atomic_load(p.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (ok) {
            // the generation cannot move before this work-group arrives; the
            // release of the arrival orders this load before it
            const unsigned gen =
                __hip_atomic_load(p.bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned arrived =
                __hip_atomic_fetch_add(p.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            if (arrived == gridDim.x) {
                // every other arrival's release, then the waiters' release
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __hip_atomic_store(p.bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(p.bar + 1, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                // relaxed polls: an acquire per poll would invalidate the caches
                // the work-groups still computing are using
                unsigned k = 0;
                while (__hip_atomic_load(p.bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                       gen) {
                    if (++k > p.spin_limit ||
                        __hip_atomic_load(p.timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_store(p.timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                // one acquire: every work-group's writes of the phase are visible
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// KIND 0: FP64-exact items; 1: tiled items (Q = 4); 2: streamed items
template <int R, int MODE, bool UNIT, bool LOC, int KIND>
__device__ __forceinline__ void persistent_item(const CoulombArgs &a, int k)
{
    if constexpr (KIND == 0)
        coulomb_fp64_exact_item(a, k);
    else if constexpr (KIND == 1)
        coulomb_tiled_item<R, MODE, UNIT, 4, LOC>(a, k);
    else
        coulomb_stream_sets<R, MODE, UNIT, LOC>(a, k);
}

// Coulomb items of stage S, barrier, stage S, barrier; false once a barrier failed
template <int S, int R, int MODE, bool UNIT, bool LOC, int KIND>
__device__ __forceinline__ bool persistent_stage(const PersistArgs &p)
{
    const EngineDev &d = p.d;
    if (p.any_coulomb) {
        CoulombArgs a = p.a;
        double *const *pos = (S == 1) ? d.P0 : d.PS;
        a.tx = a.sx = pos[0];
        a.ty = a.sy = pos[1];
        a.tz = a.sz = pos[2];
        a.skip_valid_k1 = (S == 1);
        for (int k = blockIdx.x; k < a.nitems; k += gridDim.x) {
            __syncthreads();   // the previous item's LDS is free
            persistent_item<R, MODE, UNIT, LOC, KIND>(a, k);
        }
        if (!grid_barrier(p)) return false;
    }
    for (int b = blockIdx.x; b < d.nblocks; b += gridDim.x) {
        __syncthreads();
        stage_block<S>(d, b);
    }
    return grid_barrier(p);
}

// status of the last launch on this thread
static inline int check_launch()
{
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace sbk

#endif
