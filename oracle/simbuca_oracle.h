/*
 * simbuca_oracle.h -- CPU FP64 restatement of the reference's force and
 * Dormand-Prince 5(4) path.  TEST INFRASTRUCTURE ONLY: this is the checker
 * that the HIP engine is compared against; it is never linked into, or
 * called by, the product library.
 *
 * Reference files restated (paths relative to the reference tree):
 *   bin/globals.h:4-15          physical constants
 *   bin/ion.cpp:3-11            mass_amu * amu in long double
 *   bin/force.cpp:34-82         trap quadrupole (RF+DC) and TOF terms
 *   bin/force.cpp:147-175       FP64 direct-sum Coulomb (non-MPI branch)
 *   bin/ode.cpp:61-218          Dormand_Prince_5 stages and error norm
 *   bin/ode.cpp:426-457         error_succes step-size controller
 *   bin/ode.cpp:459-494         step() retry loop and commit
 *   bin/ion_fly.cpp:83-101      move_particles loop incl. trap ramp
 *
 * Extensions that the fork does not contain (Penning B field, ideal
 * Penning quadrupole and its C4/C6 anharmonic terms, field maps,
 * dipole/quadrupole RF excitation, buffer-gas friction, stage-time
 * evaluation) are defined here from their closed forms; their tests pin
 * them against analytic solutions.
 */
#ifndef SIMBUCA_ORACLE_H
#define SIMBUCA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_fields {
    /* reference trap term, force.cpp:70-75 */
    int    trap;
    double kappa, Vdc, Vrf, r0, z0, freq_rf;
    /* reference TOF kick, force.cpp:77-80 */
    int    tof;
    double E_kick;
    /* reference Coulomb switch and scale, force.cpp:48,153 */
    int    coulomb;
    double coulomb_scale;
    /* ideal Penning quadrupole potential V0/(2 d0^2) (z^2 - (x^2+y^2)/2); off if V0 == 0 */
    double V0, d0;
    /* uniform magnetic field along z [T] */
    double Bz;
    /* dipole RF excitation along x: Ed cos(2 pi fd t + phid) [V/m] */
    double Ed, fd, phid;
    /* quadrupole RF excitation: (Uq/rq^2) cos(2 pi fq t + phiq) (x, -y, 0) [V/m] */
    double Uq, rq, fq, phiq;
    /* buffer-gas friction a -= delta v [1/s] */
    double delta;
    /* trap ramp of move_particles (ion_fly.cpp:235-239) */
    int    ramp;
    double ramp_t0, ramp_t1, newVrf, newVdc;
    /* octupole excitation (Uo/ro^4) cos(psi) (x^3 - 3xy^2, y^3 - 3x^2y, 0) [V/m] */
    double Uo, ro, fo, phio;
    /* rotating wall of order rw_order (1: (Uw/rw)(cos psi, sin psi, 0);
     * 2: (Uw/rw^2)(x cos 2psi + y sin 2psi, x sin 2psi - y cos 2psi, 0)) */
    int    rw_order;
    double Uw, rw, fw, phiw;
    /* axial dipole excitation Eax cos(psi) along z */
    double Eax, fax, phiax;
    /* frequency sweep rates of the dipole and quadrupole drives [Hz/s] */
    double fd_rate, fq_rate;
    /* magnetic bottle B2 [T/m^2] */
    double B2;
    /* buffer gas by reduced mobility K0 at p_mbar and T_gas */
    double K0, p_mbar, T_gas;
    /* SWIFT waveform: samples at swift_t0 + k swift_dt, linear interpolation,
     * dipole (kind 1, E = U/r along x) or quadrupole (kind 2, (U/r^2)(x, -y, 0)) */
    const double *swift;
    int    swift_n, swift_kind;
    double swift_t0, swift_dt, swift_r;
    /* non-ideal trap: the potential adds V0 C4/(2 d0^4) r^4 P4 and V0 C6/(2 d0^6) r^6 P6
     * (Legendre solid harmonics; used with V0 != 0) */
    double C4, C6;
    /* axisymmetric field map: (E_r, E_z) on r_i = i dr, z_j = z0 + j dz, emap[2 (i nz + j) + {0,1}],
     * bilinear, 0 outside the grid, times emap_scale */
    const double *emap;
    int    emap_nr, emap_nz;
    double emap_dr, emap_z0, emap_dz, emap_scale;
} orc_fields;

enum { ORC_TIME_FROZEN = 0, ORC_TIME_STAGE = 1 };

typedef struct orc_cloud orc_cloud;

/* Ion mass in kg exactly as Ion::Ion computes it (long double product). */
double orc_mass_kg(double mass_amu);

/* Restatement of force(): derivs[n][3] (an acceleration) at time t. */
void orc_force(int n, const double *pos, const double *vel, const double *mass,
               const int *charge, const orc_fields *f, double t, double *derivs);

/* Coulomb sum f_j of force.cpp:150-170 for the listed target rows only
 * (derivs_j -= f_j in the reference); out[nrows][3]. */
void orc_coulomb_rows(int n, const double *pos, const double *mass, const int *charge,
                      double coulomb_scale, const int *rows, int nrows, double *out);

/* Generic direct sum of the cunbody1_force interface:
 * ai[i] = sum_j mj[j] (xj[j]-xi[i]) / (|xj[j]-xi[i]|^2 + eps2)^{3/2}. */
void orc_cunbody_force(const double *xj, const double *mj, const double *xi, double eps2,
                       double *ai, int ni, int nj);

/* Controller of ode.cpp:426-457 on a scalar error. Returns 1 on accept. */
int orc_controller(double err, double *h, double *hnext, double *errold, int *reject);
/* the same with PI control: beta = 0 is orc_controller (bin/ode.cpp:427-432) */
int orc_controller_beta(double err, double beta, double *h, double *hnext, double *errold,
                        int *reject);

orc_cloud *orc_create(int n, const double *pos, const double *vel, const double *mass,
                      const int *charge, const orc_fields *f, double lifetime, double h0,
                      double rtol, double atol, int adaptive, int time_mode);
void   orc_destroy(orc_cloud *c);
/* PI control of the cloud's step controller (0 at creation) */
void   orc_set_beta(orc_cloud *c, double beta);
void   orc_set_fields(orc_cloud *c, const orc_fields *f);
/* One Dormand_Prince_5 attempt from (pos, vel); leaves y_{n+1} in pos2/vel2. */
double orc_dp5_attempt(orc_cloud *c);
/* step(): retries until accepted, commits. 0 = ok, -1 = step-size underflow. */
int    orc_step(orc_cloud *c);
/* move_particles(): steps while lifetime < t_end. Records h before each
 * step into h_hist (if non-null, up to cap). Returns steps taken or -1. */
long   orc_advance(orc_cloud *c, double t_end, long max_steps, double *h_hist, long cap);
void   orc_get_state(const orc_cloud *c, double *pos, double *vel, double *scalars);
/* scalars: [lifetime, h, hnext, errold, reject, steps, attempts, last_err] */

#ifdef __cplusplus
}
#endif
#endif
