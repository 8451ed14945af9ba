// sbk_internal.h -- device data layout shared by the HIP kernels
// (sbk_kernels.hip) and the host engine (sbk_engine.cpp).
//
// HBM layout of an engine holding clouds c = 0..C-1 with n_c ions each,
// concatenated at offsets off_c (global ion index g):
//   P0[3][stride]    positions at t_n        (IonCloud::pos,  bin/ioncloud.h:27)
//   V0[3][stride]    velocities at t_n       (IonCloud::vel,  bin/ioncloud.h:29)
//   PS[3][stride]    stage positions         (IonCloud::pos2, bin/ioncloud.h:28)
//   KV[s][3][stride] stage velocities, s=2..7 (k_s[0..2] of bin/ode.cpp:110-173;
//                    KV[1] aliases V0, and KV[7] is the new velocity vel2)
//   KA[s][3][stride] stage accelerations, s=1..7 (k_s[3..5], bin/ode.cpp:104-180)
//   CP[slice][3][stride] Coulomb partial sums per source slice
//   ERRB[nblocks]    per-ion-block sum of squared scaled errors of the embedded pair
// Per-ion kernels run over a table of ion blocks (at most 256 ions of one
// cloud each), so a block never straddles two clouds and stage 7 can reduce
// its error terms in LDS before the per-cloud controller sums the blocks.
// All FP64, structure of arrays, so every per-ion kernel reads and writes
// unit-stride, fully coalesced lines.
#ifndef SBK_INTERNAL_H
#define SBK_INTERNAL_H

#include <stdint.h>

namespace sbk {

// Coulomb tile width: sources staged in LDS per pass = threads per block.
constexpr int kBlock = 256;
// Sources accumulated in FP32 before folding into the cross-tile sum.
constexpr int kSubTile = 64;

struct CloudParams {
    // reference trap term (bin/force.cpp:40-44,70-80)
    int trap, tof, coulomb, ramp;
    double kappa, Vdc, Vrf, r02, z02, omega_rf, E_kick;
    double ramp_t0, ramp_t1, newVrf, newVdc;
    // extensions
    double cpen;              // V0 / (2 d0^2)
    double Bz;
    double Ed, fd, phid;
    double Uq_rq2, fq, phiq;  // Uq / rq^2
    double delta;
    double fd_rate, fq_rate;  // drive frequency sweeps [Hz/s]
    double Uo_ro4, fo, phio;  // octupole drive, Uo / ro^4
    int rw_order, pad3;       // rotating wall: order 1 or 2, amplitude Uw / rw^order
    double Uw_rwl, fw, phiw;
    double Eax, fax, phiax;   // axial dipole drive
    double B2;                // magnetic bottle
    double inv_K;             // buffer gas: 1 / mobility (0: off)
    const double *swift;      // SWIFT samples on the device (swift_n > 0)
    int swift_n, swift_kind;
    double swift_t0, swift_dt, swift_r;
    double c4k, c6k;          // anharmonic trap: V0 C4 / (2 d0^4), V0 C6 / (2 d0^6)
    const double *emap;       // field map on the device (emap_nr > 0): (E_r, E_z) per node
    int emap_nr, emap_nz;
    double emap_dr, emap_z0, emap_dz, emap_scale;
    // layout
    int off, n, slices, pad;
    int blk0, nblk;           // this cloud's range in the ion-block table
    int fsal;                 // an accepted attempt's stage-7 derivative is exactly the
                              // next step's stage-1 derivative (same state and time)
    int xcross_on;            // record the ions whose x crosses xcross in an accepted step
    int ion0;                 // slot (global index) of the caller's ion 0: its x scales the
                              // step-size underflow test (bin/ode.cpp:482)
    double xcross;
};

// One ion that crossed the plane x = CloudParams::xcross in an accepted step
// (bin/ion_fly.cpp:110-114), with its state after the step.
struct XCrossing {
    long long step;           // the cloud's accepted-step count after the step
    int cloud, slot;          // slot: the ion's index in the engine's arrays
    double t;                 // lifetime after the step
    double p[3], v[3];
};

struct IonBlock {
    int cloud, start, end, pad;
};

struct CloudState {
    double t, h, hnext, errold, last_err, t_end;
    int reject, active, accepted_last, error;
    long long steps, attempts;
    int budget;               // accepted steps still wanted (-1: unlimited)
    int k1_valid;             // KA[1] holds the derivative at (P0, V0, t)
    long long evals;          // Coulomb evaluations done (6 or 7 per attempt)
    double t_prev;            // start time of the last accepted step
};

// One Coulomb work item: targets [i0,i1) against sources [j0,j1) of one
// cloud, written into partial-sum slice `slice`.
struct CoulombItem {
    int cloud, i0, i1, j0, j1, slice, pad0, pad1;
};

struct CoulombArgs {
    const double *tx, *ty, *tz;   // targets
    const double *sx, *sy, *sz;   // sources
    const double *sqd;            // source weight (charge), FP64
    const float *sqf;             // source weight, FP32
    const double *kc;             // per-target prefactor (nullptr: 1)
    double *out;                  // out[slice*3*stride + comp*stride + g]
    long long stride;
    const CoulombItem *items;
    int nitems;
    int self_skip;                // FP64-exact: skip source index == target index
    int unit_weights;             // every source weight is exactly 1 (charge +1, mj = 1)
    const CloudState *state;      // nullptr: every item active
    int skip_valid_k1;            // stage 1: skip clouds whose KA[1] is still valid
    int local_coords;             // mixed modes: far sub-tiles use FP32 offsets from the
                                  // target box centre (pays off with ions in spatial order)
    float eps_f;
    double eps_d;
};

struct EngineDev {
    double *P0[3], *V0[3], *PS[3];
    double *KV[8][3];
    double *KA[8][3];
    double *CP;
    double *ERRB;
    double *derivs_out;           // [3][stride] scratch for sbk_eval_force/coulomb
    long long stride;
    const double *qm;             // charge*e/mass (bin/force.cpp:71)
    const IonBlock *blocks;
    int nblocks;
    const CloudParams *params;
    CloudState *state;
    const int *slices;            // per-cloud slice count of the current item list
    int ntot, nclouds;
    int time_mode, adaptive;
    double rtol, atol;
    double beta;                  // PI control of the step controller (0: the reference's)
    XCrossing *xc;                // crossing list (nullptr: no cloud records crossings)
    int *xc_count;                // entries written since the list was last emptied
    int xc_cap;                   // entries the list holds
    int pad_xc;
};

// Persistent attempts: one cooperative launch of gridDim work-groups runs up
// to `attempts` whole attempts (7 x (Coulomb items -> stage), controller,
// commit), separated by grid barriers, and stops early once no cloud is
// active.  The barrier counters live in bar[0] (arrivals) and bar[1]
// (generation); a wait that spins past spin_limit sets *timeout, after which
// every barrier returns at once and every work-group leaves.
struct PersistArgs {
    EngineDev d;
    CoulombArgs a;                // items, weights, prefactors, partials (positions set per stage)
    unsigned *bar;
    int *timeout;
    unsigned spin_limit;
    int attempts;
    int any_coulomb;
    int pad;
};

// sets the message sbk_last_error() returns; returns code (sbk_engine.cpp)
int record_error(int code, const char *msg);

// host-side launchers (sbk_kernels.hip)
enum KernelPrecision {
    KP_FP64_EXACT = 0, KP_FP64 = 1, KP_MIXED = 2, KP_MIXED_KAHAN = 3, KP_FP64_RSQ = 4
};

// Coulomb kernels: KK_TILED (coulomb_tiled / coulomb_fp64_exact: the work-group's
// wave groups share each LDS source tile) or KK_STREAM (coulomb_stream, mixed
// modes, q = 4: each wave streams its own sub-tiles with no barrier inside the
// source loop).  Both take the same work items.
enum CoulombKernel { KK_TILED = 0, KK_STREAM = 1 };

// q: wave groups of a work-group that split the sources of one target set (1 or 4)
int launch_coulomb(int precision, int targets_per_thread, int q, int kernel, const CoulombArgs &a,
                   void *stream);
// resident work-groups per CU of the Coulomb kernel a launch would use
int coulomb_blocks_per_cu(int precision, int targets_per_thread, int unit_weights, int q,
                          int local_coords, int kernel);
int launch_stage(int stage, const EngineDev &d, void *stream);
int launch_controller(const EngineDev &d, void *stream);
int launch_commit(const EngineDev &d, void *stream);
// set_end: 0 keep t_end, 1 t_end = t + duration, 2 t_end = duration (absolute)
int launch_begin(const EngineDev &d, double duration, int budget, int set_end, void *stream);
int launch_eval_derivs(const EngineDev &d, int with_fields, void *stream);
// Persistent attempts (cooperative launch) for precisions with 1, 2 or 4 targets
// per thread and 4 source groups (FP64-exact: 1 and 1); mixed precisions use the
// streamed Coulomb item.  Returns the launch's hipError_t (hipErrorInvalidValue
// when no such kernel exists); a refused launch has enqueued nothing.
int launch_attempts_persistent(int precision, int targets_per_thread, const PersistArgs &p,
                               int grid, void *stream);
// resident work-groups per CU of that kernel (0: none)
int persistent_blocks_per_cu(int precision, int targets_per_thread, int unit_weights,
                             int local_coords);
// out[g] = in[src[g]] for g < n
int launch_gather(const int *src, const double *in, double *out, int n, void *stream);
int launch_gather_f32(const int *src, const float *in, float *out, int n, void *stream);

}  // namespace sbk

#endif
