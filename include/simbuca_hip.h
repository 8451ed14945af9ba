/*
 * simbuca_hip.h -- C ABI of libsimbuca_hip.so, the MI355X (gfx950) force and
 * Dormand-Prince 5(4) engine for ion-cloud simulation.
 *
 * Export sets:
 *
 *  1. cunbody1_force{,_begin,_end}: the exact prototypes the reference binds
 *     for its GPU Coulomb backend (bin/cunbody.h:11-33, called from
 *     bin/cunbody.cpp:66,88 and bin/force.cpp:50,86).  Linking this library
 *     in place of libcunbody1.a makes the reference's force() run its
 *     direct sum on the MI355X.
 *
 *  1b. The reference's NBODY interface: include/simbuca_nbody.h, in
 *     libsimbuca_nbody.so.
 *
 *  2. sbk_*: a device-resident, step-level engine.  It replaces the
 *     reference's Dormand_Prince_5 + error_succes + step() commit
 *     (bin/ode.cpp:61-218,426-494) and the force() it calls
 *     (bin/force.cpp:34-175), keeping the whole ion state in HBM across
 *     stages, steps and retries.  Many independent clouds of different
 *     sizes share one engine; each has its own time, step size and
 *     controller.
 *
 *  3. sbk_writer_*, sbk_import: the reference's output files (host only).
 *
 *  4. sbk_init_cloud, sbk_mass_table_*, sbk_parse_species, sbk_run_ini,
 *     sbk_ini_describe: the reference's simulation set-up -- its cloud
 *     initializer, mass tables and .ini simulation files -- on the engine.
 *
 * All sbk_* functions return 0 on success and a negative SBK_E* code on
 * failure; sbk_last_error() gives the message.  Host arrays are owned by
 * the caller and only read/written during the call.  An engine is used by
 * one host thread at a time and runs on the HIP device that was current
 * when it was created.
 */
#ifndef SIMBUCA_HIP_H
#define SIMBUCA_HIP_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Export set 1 -- reference CUNBODY interface (bin/cunbody.h:11-33).
 *
 * ai[i] = sum_{j<nj} mj[j] (xj[j] - xi[i]) / (|xj[j] - xi[i]|^2 + eps2)^{3/2}
 * for i < ni.  The reference passes mj = 1, eps2 = (double)1e-23f and
 * ni = nj = N (bin/cunbody.cpp:31-35,66) and scales the result by
 * -coulomb_scale e^2 ke / m_i itself (bin/force.cpp:87-93).
 * _begin enqueues the work and returns; _end (same arguments) blocks until
 * ai is written.  cunbody1_force does both.  nj = 0 gives ai = 0; ni = 0
 * writes nothing.  Errors print a message to stderr and end the process
 * with EXIT_FAILURE, as the reference backends do
 * (bin/nbody_kernel.cu:37-47).  The device buffers are one process-wide
 * set on the device current at the call, so calls come from one host
 * thread at a time, as the reference's force() makes them.
 * ---------------------------------------------------------------------- */
void cunbody1_force(double xj[][3], double mj[], double xi[][3], double eps2, double ai[][3],
                    int ni, int nj);
void cunbody1_force_begin(double xj[][3], double mj[], double xi[][3], double eps2,
                          double ai[][3], int ni, int nj);
void cunbody1_force_end(double xj[][3], double mj[], double xi[][3], double eps2,
                        double ai[][3], int ni, int nj);

/* Arithmetic used by cunbody1_force* (no counterpart in the reference,
 * whose backend precision is fixed inside a closed binary): SBK_FP64
 * (default), SBK_FP64_EXACT, SBK_FP64_RSQ, SBK_MIXED or SBK_MIXED_KAHAN.
 * Returns the previous setting; other values leave it unchanged. */
int sbk_cunbody_set_precision(int precision);

/* cunbody1_force with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, result and precision setting as
 * cunbody1_force, but every device buffer (staging, slice partials, work
 * items) lives in one static device array of the library, about 17 MB per
 * GPU, held until the process ends; ni and nj are at most 32768 and larger
 * sets return SBK_EINVAL without touching ai.  Runs on the null stream and
 * returns when ai is written; 0 or an SBK_E* code (sbk_last_error() gives
 * the message).  All calls on one GPU share its array, so calls that overlap
 * on one GPU (from several host threads) race. */
int sbk_cunbody_force_static(double xj[][3], double mj[], double xi[][3], double eps2,
                             double ai[][3], int ni, int nj);

/* Export set 1b, the reference's NBODY interface (initialize, finalize,
 * simulation_Async*), is in include/simbuca_nbody.h and its own library,
 * libsimbuca_nbody.so, so that this library exports no generic names. */

/* ------------------------------------------------------------------------
 * Export set 2 -- device-resident engine.
 * ---------------------------------------------------------------------- */

enum {
    SBK_OK = 0,
    SBK_EINVAL = -1,     /* bad argument */
    SBK_EHIP = -2,       /* HIP runtime error */
    SBK_ESTATE = -3,     /* call not valid in the engine's current state */
    SBK_EUNDERFLOW = -4, /* "stepsize underflow created" (bin/ode.cpp:482-483) */
    SBK_ENOMEM = -5,
    SBK_EIO = -6         /* a file could not be opened, read or parsed */
};

/* Coulomb arithmetic. */
enum {
    SBK_FP64_EXACT = 0,  /* FP64, per-target serial order of bin/force.cpp:158-170:
                            bit-identical to the reference CPU sum */
    SBK_FP64 = 1,        /* FP64, tiled and split over sources (fast FP64 baseline) */
    SBK_MIXED = 2,       /* FP64 separations rounded to FP32, FP32 rsqrt and tile sums,
                            FP64 accumulation across tiles */
    SBK_MIXED_KAHAN = 3, /* as SBK_MIXED with compensated FP32 accumulation across tiles */
    SBK_FP64_RSQ = 4     /* as SBK_FP64, with |r|^-3 from the hardware reciprocal-square-root
                            estimate and a second-order correction instead of IEEE sqrt and
                            division (a few ulp per pair; no counterpart in the reference) */
};

/* Time at which force() is evaluated inside a step. */
enum {
    SBK_TIME_FROZEN = 0, /* every stage sees t_n, as bin/force.cpp:36 does */
    SBK_TIME_STAGE = 1   /* stage s sees t_n + c_s h */
};

/* Field model of one cloud.  The first block is the reference's
 * (bin/trapparameters.h:6-27, bin/ode.h:42-56, bin/force.cpp:70-80); the
 * rest are Penning-trap terms that this fork of the reference lacks. */
typedef struct sbk_fields {
    int    trap;                /* RF+DC quadrupole of bin/force.cpp:70-75 */
    double kappa, Vdc, Vrf, r0, z0, freq_rf;
    int    tof;                 /* constant E_kick along x, bin/force.cpp:77-80 */
    double E_kick;
    int    coulomb;             /* direct-sum Coulomb, bin/force.cpp:48,153 */
    double coulomb_scale;
    double V0, d0;              /* ideal Penning quadrupole V0/(2 d0^2)(z^2-(x^2+y^2)/2); V0=0: off */
    double Bz;                  /* uniform magnetic field along z [T] */
    double Ed, fd, phid;        /* dipole excitation along x: Ed cos(2 pi fd t + phid) [V/m] */
    double Uq, rq, fq, phiq;    /* quadrupole excitation (Uq/rq^2) cos(2 pi fq t + phiq)(x,-y,0) */
    double delta;               /* buffer-gas friction a -= delta v [1/s] */
    int    ramp;                /* linear Vrf/Vdc ramp of bin/ion_fly.cpp:235-239 */
    double ramp_t0, ramp_t1, newVrf, newVdc;
    /* Penning excitation menu and field extras (none exists in the reference
     * fork; each is defined from its potential and is off when its amplitude
     * is 0).  Phases are psi = 2 pi (f t + rate t^2/2) + phi. */
    double Uo, ro, fo, phio;    /* octupole: potential -(Uo/(4 ro^4)) cos(psi) Re (x+iy)^4,
                                   E = (Uo/ro^4) cos(psi) (x^3-3xy^2, y^3-3x^2y, 0) */
    int    rw_order;            /* rotating wall of order l = 1 (dipole) or 2 (quadrupole): */
    double Uw, rw, fw, phiw;    /* potential -(Uw/(l rw^l)) Re[(x+iy)^l e^{-i l psi}], turning at fw;
                                   E = (Uw/rw) (cos psi, sin psi, 0) for l = 1,
                                   E = (Uw/rw^2) (x cos 2psi + y sin 2psi, x sin 2psi - y cos 2psi, 0) for l = 2 */
    double Eax, fax, phiax;     /* axial dipole excitation along z: Eax cos(psi) [V/m] */
    double fd_rate, fq_rate;    /* linear sweeps of the dipole and quadrupole drive frequencies [Hz/s] */
    double B2;                  /* magnetic bottle: B += B2 ((z^2-(x^2+y^2)/2) z_hat - z (x x_hat + y y_hat)) [T/m^2] */
    double K0, p_mbar, T_gas;   /* buffer gas by reduced ion mobility K0 [m^2/(V s)] at pressure p_mbar [mbar]
                                   and temperature T_gas [K]: a -= (q e/m) v / K, K = K0 (1013.25/p_mbar)(T_gas/273.15);
                                   off when K0 or p_mbar is 0 */
    /* SWIFT: a sampled waveform U(t) [V], samples at t0 + k dt (k < swift_n),
     * linearly interpolated and 0 outside them; applied as a dipole along x,
     * E = (U/swift_r) (1, 0, 0) (swift_kind 1), or as a quadrupole,
     * E = (U/swift_r^2) (x, -y, 0) (swift_kind 2).  The engine copies the
     * samples when the fields are set; off when swift_n is 0. */
    const double *swift;
    int    swift_n, swift_kind;
    double swift_t0, swift_dt, swift_r;
    /* Non-ideal trap (the upstream c4, c6 terms; used with V0 != 0): the trap
     * potential is (V0/2) sum_k C_k (rho/d0)^k P_k(cos theta) for k = 2, 4, 6
     * with C_2 = 1 (the ideal term above), that is
     *   + V0 C4 / (2 d0^4) (z^4 - 3 z^2 r^2 + 3 r^4/8)
     *   + V0 C6 / (2 d0^6) (z^6 - 15/2 z^4 r^2 + 45/8 z^2 r^4 - 5/16 r^6),  r^2 = x^2 + y^2. */
    double C4, C6;
    /* Axisymmetric electric-field map (the upstream field-map files): E_r and
     * E_z in V/m on a grid r_i = i emap_dr (i < emap_nr), z_j = emap_z0 +
     * j emap_dz (j < emap_nz), stored emap[2 (i emap_nz + j)] = E_r,
     * emap[2 (i emap_nz + j) + 1] = E_z; bilinearly interpolated, 0 outside
     * the grid, times emap_scale; E = (E_r x/r, E_r y/r, E_z).  The engine
     * copies the grid when the fields are set; off when emap_nr is 0. */
    const double *emap;
    int    emap_nr, emap_nz;
    double emap_dr, emap_z0, emap_dz, emap_scale;
} sbk_fields;

typedef struct sbk_config {
    int    precision;   /* SBK_FP64_EXACT .. SBK_FP64_RSQ */
    int    time_mode;   /* SBK_TIME_FROZEN or SBK_TIME_STAGE */
    int    adaptive;    /* 1: error_succes controller; 0: fixed step (bin/ode.cpp:478-486) */
    int    use_graph;   /* 1: replay one captured hipGraph per attempt */
    int    timing;      /* 1: time every Coulomb launch with HIP events */
    double rtol, atol;  /* the reference hard-codes 1e-7/1e-7 (bin/ode.cpp:515-516) */
    double h0;          /* initial step (the [ode] timestep of the reference) */
    int    spatial_order;    /* mixed precisions: 1 keeps each cloud's ions in a
                                spatial order on the device, so most Coulomb
                                sub-tiles form separations from FP32 offsets
                                (faster, same accuracy); results are returned
                                in the caller's order either way */
    int    reorder_interval; /* attempt launches (one attempt of every active
                                cloud) between re-sorts of that order as the
                                ions move (0: sort once, at the first run) */
    double beta;        /* PI control of the step controller: 0 (the reference's
                           value, default) or e.g. 0.04..0.08, the reference's
                           documented switch (bin/ode.cpp:427-432); alpha =
                           0.2 - 0.75 beta.  Must be in [0, 0.2). */
    int    persistent;  /* 1: each batch of attempts is one cooperative launch, with grid barriers
                           between the phases, instead of the graph replay (16 launches per
                           attempt); same results bit for bit.  0 (the default): the graph or
                           separate launches.  On MI355X a grid barrier (2 us at 16 work-groups,
                           13 us at 512) costs more than a graph's kernel boundary (1.7 us), so
                           the graph is faster at every size measured.  Needs 4 source groups
                           and 1, 2 or 4 targets per thread (SBK_EINVAL otherwise) and a cooperative
                           queue, which the device may refuse while another process holds it
                           (SBK_EHIP); timing = 1 uses separate launches. */
    int    persistent_spin; /* polls of a grid-barrier wait before the launch gives up and
                           reports SBK_EHIP (0: the default, 4194304, seconds of waiting) */
} sbk_config;

typedef struct sbk_cloud_status {
    double lifetime;    /* IonCloud::lifetime */
    double h, hnext, errold, last_err, t_end;
    int    reject, active, accepted_last, error;
    long long steps, attempts;
    int    n;
    int    slices;      /* source slices of the Coulomb work split */
    long long coulomb_evals;  /* Coulomb evaluations done (7 or 6 per attempt) */
    double t_prev;      /* start time of the last accepted step (lifetime before it) */
} sbk_cloud_status;

typedef struct sbk_stats {
    double    coulomb_ms;        /* summed HIP-event time of Coulomb launches (timing=1) */
    long long coulomb_launches;  /* launches counted in coulomb_ms */
    double    pair_interactions; /* sum over clouds of Coulomb evaluations * n(n-1) */
    long long attempts;          /* sum over clouds */
    long long steps;             /* accepted steps, summed over clouds */
    double    ion_steps;         /* sum over clouds of steps*n */
    long long reorders;          /* re-sorts of the device ion order (spatial_order) */
    long long persistent_launches; /* batches of attempts run as one persistent launch */
} sbk_stats;

typedef struct sbk_engine sbk_engine;

void sbk_default_config(sbk_config *cfg);
void sbk_default_fields(sbk_fields *f);
int  sbk_create(const sbk_config *cfg, sbk_engine **out);
int  sbk_destroy(sbk_engine *e);

/* Adds a cloud (n ions, AoS [n][3] positions in m and velocities in m/s,
 * masses in kg, charges in units of e) at time `lifetime`.  Clouds can be
 * added until the first call that runs work on the device. */
int  sbk_add_cloud(sbk_engine *e, int n, const double *pos, const double *vel,
                   const double *mass, const int *charge, const sbk_fields *f,
                   double lifetime, int *cloud_id);
int  sbk_set_fields(sbk_engine *e, int cloud, const sbk_fields *f);
int  sbk_set_step(sbk_engine *e, int cloud, double h);

/* Runs attempts until every cloud's lifetime reaches lifetime+duration
 * (the while-loop of move_particles, bin/ion_fly.cpp:83-101), or
 * max_attempts attempts were launched.
 * An attempt evaluates the Coulomb sum 7 times, as the reference does, except
 * that the stage-1 derivative is reused when it is provably identical: after
 * a rejected attempt (same y_n and t_n), and after an accepted one when the
 * stage-7 derivative was taken at the new state and time (stage-time mode,
 * or no time-dependent field; never with the trap ramp).  Results are
 * bit-identical either way; those attempts cost 6 evaluations. */
int  sbk_advance(sbk_engine *e, double duration, long long max_attempts);
/* As sbk_advance, to the absolute time t_end: every cloud steps while its
 * lifetime < t_end (the loop condition of move_particles). */
int  sbk_advance_until(sbk_engine *e, double t_end, long long max_attempts);
/* One accepted step for every cloud (step(), bin/ode.cpp:459-494). */
int  sbk_step(sbk_engine *e);
/* Enqueues `count` attempts for the active clouds without waiting. */
int  sbk_attempts(sbk_engine *e, int count);
int  sbk_synchronize(sbk_engine *e);
/* Re-sorts the device order of every cloud by its current positions now
 * (spatial_order engines; otherwise does nothing).  sbk_advance does this by
 * itself every reorder_interval attempts.  No counterpart in the reference,
 * whose ion order never matters for speed. */
int  sbk_reorder(sbk_engine *e);

/* The print_at_x detector of move_particles (bin/ion_fly.cpp:91-93,110-114)
 * on the device: with enable != 0, every accepted step of `cloud` records
 * each ion whose x lies strictly on one side of x_plane before the step and
 * strictly on the other after it, with its state after the step.  The list
 * lives on the device and holds *list_capacity entries (max(65,536, ions of
 * the engine), at most 2^24); sbk_take_crossings empties it.  A run that
 * records more between two takes fails with SBK_ESTATE (the extra entries
 * are lost); at most one entry per ion and accepted step is recorded, so
 * taking at least every list_capacity / n steps is enough.  enable = 0
 * stops recording.  May be called before or after the first run. */
int  sbk_set_crossing_plane(sbk_engine *e, int cloud, int enable, double x_plane,
                            int *list_capacity);
/* Takes the recorded crossings of every cloud, ordered by cloud, then step,
 * then ion (the order the reference prints them in), and empties the list.
 * Per entry: cloud, the caller's ion index, the cloud's accepted-step count
 * after the step, the lifetime after the step, and pos/vel [3] after it.
 * capacity = 0 with every output array NULL only sets *n to the number of
 * entries held and keeps them; otherwise capacity must be >= that number.
 * Any output array may be NULL. */
int  sbk_take_crossings(sbk_engine *e, int capacity, int *n, int *cloud, int *ion,
                        long long *step, double *t, double *pos, double *vel);

int  sbk_get_state(sbk_engine *e, int cloud, double *pos, double *vel);
int  sbk_get_status(sbk_engine *e, int cloud, sbk_cloud_status *st);
int  sbk_get_stats(sbk_engine *e, sbk_stats *st);
int  sbk_reset_stats(sbk_engine *e);

/* force() of bin/force.cpp:34-175 at the cloud's current positions,
 * velocities and lifetime: derivs[n][3] (an acceleration). */
int  sbk_eval_force(sbk_engine *e, int cloud, double *derivs);
/* The Coulomb sum f_j of bin/force.cpp:150-170 alone (derivs_j -= f_j),
 * computed with `precision`: out[n][3]. */
int  sbk_eval_coulomb(sbk_engine *e, int cloud, int precision, double *out);

/* Coulomb launch shape.  targets_per_thread in {1,2,4,8}, or 0 (the default:
 * 4 when the pair-weighted mean cloud size is >= 32768 ions, else 2 when the
 * clouds hold >= 2.5e8 pairs in total or local coordinates are on, else 1);
 * source_groups in
 * {1,4}: wave groups of a 256-thread work-group that share one set of
 * targets and split each source tile; slices = number of source slices per
 * target block (0 = automatic); local_coords (spatial_order mixed engines):
 * 1 lets far sources form separations from FP32 offsets, 0 keeps
 * FP64-formed separations everywhere, -1 (automatic) chooses 1 when the
 * clouds hold >= 8e7 pairs in total and their pair-weighted mean size is at
 * least 2048 ions.  Applies to
 * SBK_FP64, SBK_FP64_RSQ and SBK_MIXED*; SBK_FP64_EXACT always uses one thread per target
 * and one slice. */
int  sbk_set_launch(sbk_engine *e, int targets_per_thread, int source_groups, int slices,
                    int local_coords);
/* Coulomb kernel of the mixed precisions with 4 source groups and 1, 2 or 4
 * targets per thread: SBK_KERNEL_TILED (the work-group's waves share each
 * LDS source tile, a barrier per tile), SBK_KERNEL_STREAM (each wave streams
 * its own source sub-tiles and loads the next one ahead, with no barrier in
 * the source loop), or SBK_KERNEL_AUTO (the default).  Results are the same
 * to rounding; only the speed differs.  Other launches use the tiled kernel. */
enum { SBK_KERNEL_AUTO = -1, SBK_KERNEL_TILED = 0, SBK_KERNEL_STREAM = 1 };
int  sbk_set_coulomb_kernel(sbk_engine *e, int kernel);
/* The launch the engine runs for `precision` with its current clouds and
 * settings, the automatic choices resolved: targets per thread, source
 * groups, the largest slice count of any cloud in the work list, local
 * coordinates (0/1) and the Coulomb kernel (SBK_KERNEL_TILED or
 * SBK_KERNEL_STREAM).  Any output pointer may be null.  No counterpart in
 * the reference, whose launch parameters were fixed by its .dat tables. */
int  sbk_get_launch(sbk_engine *e, int precision, int *targets_per_thread, int *source_groups,
                    int *max_slices, int *local_coords, int *kernel);
/* Times `reps` launches of the Coulomb kernel of the current state's
 * positions for every cloud; ms_per_launch receives the mean. */
int  sbk_time_coulomb(sbk_engine *e, int precision, int reps, double *ms_per_launch);
/* Tries the launch shapes for this engine's clouds and keeps the fastest,
 * timing each Coulomb launch together with one pass that reads its slice
 * partials back (as every stage does); returns the chosen shape and Coulomb
 * kernel (SBK_KERNEL_TILED or SBK_KERNEL_STREAM; the streamed kernel is
 * tried for the mixed precisions).  One line per shape tried ("device ions
 * clouds precision targets_per_thread source_groups slices local_coords
 * kernel ms") is appended to `table_path` if non-null. */
int  sbk_autotune(sbk_engine *e, int precision, int *targets_per_thread, int *source_groups,
                  int *slices, int *local_coords, int *kernel, const char *table_path);

/* Applies the fastest shape recorded in an sbk_autotune table for this
 * device, ion count, cloud count and precision, without timing anything
 * (lines without the kernel column are read as the tiled kernel, lines
 * without the local_coords column as local_coords 0):
 * the counterpart of the reference's per-GPU launch table read at start-up
 * (bin/nbody.cpp:194-203,275-321).  Returns SBK_ESTATE when the table has
 * no matching line (the automatic shape stays in use). */
int  sbk_load_tuning(sbk_engine *e, int precision, const char *table_path);

/* The device order a spatial_order engine gives one cloud: order[i] is the
 * index (0..n-1, into the AoS [n][3] positions) of the ion placed in slot i.
 * Host only; recursive bisection along the longest axis, with the first part
 * a multiple of the largest power of two A <= 1024 with 2A <= part size, so
 * aligned runs of 64..1024 ions are spatially compact. */
int  sbk_spatial_order(int n, const double *pos, int *order);

/* ---- Output files in the reference's formats (host only; no GPU) ----------
 * The counterpart of IonCloud's printing (bin/ioncloud.cpp:85-257, 38-58),
 * for a cloud read back with sbk_get_state.  Files:
 *   <prefix>_pAll.txt (or <prefix>_p<k>.txt, k = 1..n, with
 *   per_particle_files): "ID name t(ms) x y z(mm) vx vy vz(m/s) E(eV)" per ion
 *   and print (IonCloud::PrintParticle);
 *   <prefix>_cloud<k>.txt, k = rank of the species name in sorted order:
 *   "k name N t(ms) x0 y0 z0 sx sy sz(mm) E(eV)" per print (IonCloud::PrintCloud,
 *   whose per-species sums divide by the total ion count N and whose energy adds
 *   vz*vy, both kept so the numbers equal the reference's).
 * digits = 6 gives the reference's "%.6e"; 17 makes restart files that
 * sbk_import reads back to within one ulp (positions pass through mm). */
typedef struct sbk_writer sbk_writer;
int  sbk_writer_open(const char *prefix, int n, const char *const *names, const double *mass_amu,
                     int per_particle_files, int digits, sbk_writer **out);
int  sbk_writer_particles(sbk_writer *w, double lifetime, const double *pos, const double *vel);
int  sbk_writer_cloud(sbk_writer *w, double lifetime, const double *pos, const double *vel);
/* Renames ion k and sets the mass (kg) its energies are printed with, as the
 * reference's dissociation operation does (bin/operation.cpp:58-70).  The
 * species files keep the species the writer opened with. */
int  sbk_writer_set_ion(sbk_writer *w, int k, const char *name, double mass_kg);
/* IonCloud::PrintX (bin/ioncloud.cpp:120-146): one line "ID name t(ms) x y z(mm)
 * vx vy vz(m/s)" for ion k in <prefix>_x.txt.  As in the reference, the call
 * that opens the file writes only its header. */
int  sbk_writer_x(sbk_writer *w, int k, double lifetime, const double *pos, const double *vel);
/* IonCloud::Delete: a last print of every ion (if pos/vel are non-null), the
 * closing block, and frees the writer. */
int  sbk_writer_close(sbk_writer *w, double lifetime, const double *pos, const double *vel);
/* Reads the last complete print of <prefix>_pAll.txt written as above: n ions
 * in ID order, names[n][name_len], pos/vel [n][3] in m and m/s, lifetime in s.
 * The reference's own importer (bin/parser.cpp:9-261) expects 14 columns its
 * writer never writes; this reads the writer's 11.  With capacity 0 and every
 * output pointer but n NULL, only *n is set (a size query). */
int  sbk_import(const char *prefix, int capacity, int *n, char *names, int name_len,
                double *pos, double *vel, double *lifetime);

/* ---- The reference's simulation set-up (host only, except sbk_run_ini) ----
 * sbk_init_cloud: InitCloud (bin/initialization.cpp:117-206).  n ions in the
 * ellipsoid offset + semiaxis * (unit ball), velocities of Maxwell-Boltzmann
 * speed around energy_eV[0] with isotropic directions, drawn from MT19937
 * seeded with `seed` (bin/mtrand.h:96-104) in the reference's order, so the
 * same arguments give the reference's cloud bit for bit.  With n_energy = 2,
 * vz takes its own speed around energy_eV[1]; 1 or 3 energies use one speed.
 * mass_amu[n] in u; pos, vel [n][3] in m and m/s. */
int  sbk_init_cloud(int n, unsigned long seed, const double *semiaxis, const double *offset,
                    int n_energy, const double *energy_eV, const double *mass_amu,
                    double *pos, double *vel);
/* The reference's mass tables: the AME file (IonTable, bin/IonTable.cpp:14-91,
 * names "<A><El>" as "133Cs") and the PDG particle file (PDGTable,
 * bin/PDGTable.cpp:79-183, masses converted by 1000/931.494 u per GeV).
 * Either path may be NULL.  Lookups take the first row of the name, AME
 * before PDG. */
typedef struct sbk_mass_table sbk_mass_table;
int  sbk_mass_table_open(const char *ame_path, const char *pdg_path, sbk_mass_table **out);
int  sbk_mass_table_lookup(const sbk_mass_table *t, const char *name, double *mass_amu);
void sbk_mass_table_close(sbk_mass_table *t);
/* A [composition] name (bin/initialization.cpp:333-391): charge from the
 * suffix "+", "-", "<k>+", "<k>-" (1 without one), mass from the tables
 * (t may be NULL) or else the name's leading digits ("133Xx+" -> 133 u). */
int  sbk_parse_species(const char *name, const sbk_mass_table *t, double *mass_amu, int *charge);
/* Runs a reference simulation file (main + DoSimulation, bin/main.cpp:13-31,
 * bin/initialization.cpp:208-670): sections [cloud], [composition], [ode]
 * (Dormand-Prince), [coulomb], [trap] (ideal), [output] (print_at_x
 * included) and [operation0..31] (normal, tof, trap_ramp, nop,
 * dissociation), with the reference's defaults, print schedule and output
 * files, on this engine at the given precision and time_mode.  [import]
 * reads the writer's own prints (the reference's importer cannot); an
 * [output] key "digits" (default 6) sets the printed precision.  Not provided
 * (an error): ODE methods other than Dormand-Prince.  t may be NULL (names then
 * need a leading mass number).  verbose prints each operation on stderr. */
int  sbk_run_ini(const char *ini_path, const sbk_mass_table *t, int precision, int time_mode,
                 int verbose);
/* What sbk_run_ini would run (host only, no GPU): the file's settings after
 * the reference's defaults and rules, as one JSON object in json[len]: N,
 * temp, semiaxis, offset, species [{name, count, mass_amu, charge}] in ion
 * order, timestep, adaptive, coulomb, coulomb_scale, trap {Vrf, Vdc, kappa,
 * r0, z0, freq_rf, E_kick, newVrf, newVdc}, prefix, print_interval,
 * separate, print_after_operation, operations [{type, duration,
 * print_timestep, ode_timestep}], skipped_line (first line that is neither
 * a section nor a pair, 0 if none). */
int  sbk_ini_describe(const char *ini_path, const sbk_mass_table *t, char *json, int len);
/* The reference's INIReader on a file (bin/ini.c, bin/inireader.cpp):
 * sbk_ini_value is Get(section, name): the raw value (repeated names and
 * continuation lines joined by '\n'; names and sections match without regard
 * to case), or SBK_EINVAL when it is not set.  sbk_ini_names lists, one per
 * line, the sections (section NULL, GetSections) or the field names of a
 * section (GetFields), in sorted order with their case kept. */
int  sbk_ini_value(const char *ini_path, const char *section, const char *name, char *out, int len);
int  sbk_ini_names(const char *ini_path, const char *section, char *out, int len);

const char *sbk_last_error(void);
/* Device name, compute-unit count and clock (MHz) of the current device. */
int  sbk_device_info(char *name, int name_len, int *cus, int *clock_mhz);

#ifdef __cplusplus
}
#endif
#endif
