/*
 * simbuca_nbody.h -- C ABI of libsimbuca_nbody.so: the reference's NBODY
 * backend interface on the MI355X.  A library of its own, so that these
 * generic names stay out of libsimbuca_hip.so (which it uses for the Coulomb
 * kernels; link both).
 */
#ifndef SIMBUCA_NBODY_H
#define SIMBUCA_NBODY_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Export set 1b -- reference NBODY interface, as its caller declares it
 * (bin/nbody.cpp:12-18; the reference's own kernel file defines different,
 * incompatible signatures at bin/nbody_kernel.cu:116,326).
 *
 * Positions go in as FP32 {x, y, z, w} (w ignored), accelerations come back
 * as FP32 {x, y, z}: acc_i = sum_j (x_j - x_i) (|x_j - x_i|^2 + 1e-7)^{-3/2}
 * (bin/nbody_kernel.cu:19,182-214; no charges, no ke).  The caller scales
 * by -coulomb_scale e^2 ke / m_i (bin/force.cpp:97-104).  Work is enqueued
 * on stream 0, so an event the caller records on stream 0 after
 * simulation_Async completes with it (bin/nbody.cpp:243-245,265-268).
 * p, q, ntiles are the reference's launch-table hints and are ignored.
 * initialize() allocates for up to n_max bodies on the device current at
 * that call; later calls use those buffers with the same device current.
 * Errors print a message and end the process with EXIT_FAILURE.
 * ---------------------------------------------------------------------- */
typedef struct sbk_float4 { float x, y, z, w; } sbk_float4;   /* layout of float4 */
typedef struct sbk_float3 { float x, y, z; } sbk_float3;      /* layout of float3 */
void  initialize(int n_max, bool withCharge);
float finalize(void);
void  simulation_Async(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q, int ntiles);
void  simulation_Async_begin(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q,
                             int ntiles);
void  simulation_Async_end(sbk_float4 *h_p, sbk_float3 *res, int nbody, int p, int q,
                           int ntiles);

#ifdef __cplusplus
}
#endif
#endif
