/*
 * simbuca_oracle.c -- FP64 CPU restatement of the reference force and
 * Dormand-Prince 5(4) path.  TEST INFRASTRUCTURE: used by tests/, the smoke
 * check and the CPU-baseline leg of the benchmark as the checker, never by
 * the product library.
 *
 * Every expression keeps the reference's evaluation order so that results
 * are bit-identical to the reference built with g++ on x86-64 (SSE2 doubles,
 * no FMA contraction: build with -ffp-contract=off).  Citations are to the
 * reference tree (bin/...).
 */
#include "simbuca_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* bin/globals.h:6-15 */
static const double pi        = 3.14159265358979323;
static const double el_charge = 1.60217733e-19;
static const double amu       = 1.6605402E-27;
static const double ke        = 8.987551787E9;

/* bin/ode.cpp:76-86, Dormand-Prince 5(4) tableau as double quotients */
static const double DPc2 = 1.0 / 5.0, DPc3 = 3.0 / 10.0, DPc4 = 4.0 / 5.0, DPc5 = 8.0 / 9.0,
                    DPc6 = 1.0, DPc7 = 1.0;
static const double DPa21 = 1.0 / 5.0;
static const double DPa31 = 3.0 / 40.0, DPa32 = 9.0 / 40.0;
static const double DPa41 = 44.0 / 45.0, DPa42 = -56.0 / 15.0, DPa43 = 32.0 / 9.0;
static const double DPa51 = 19372.0 / 6561.0, DPa52 = -25360.0 / 2187.0, DPa53 = 64448.0 / 6561.0,
                    DPa54 = -212.0 / 729.0;
static const double DPa61 = 9017.0 / 3168.0, DPa62 = -355.0 / 33.0, DPa63 = 46732.0 / 5247.0,
                    DPa64 = 49.0 / 176.0, DPa65 = -5103.0 / 18656.0;
static const double DPa71 = 35.0 / 384.0, DPa73 = 500.0 / 1113.0, DPa74 = 125.0 / 192.0,
                    DPa75 = -2187.0 / 6784.0, DPa76 = 11.0 / 84.0;
static const double DPb1 = 35.0 / 384.0, DPb3 = 500.0 / 1113.0, DPb4 = 125.0 / 192.0,
                    DPb5 = -2187.0 / 6784.0, DPb6 = 11.0 / 84.0;
static const double DPe1 = 71.0 / 57600.0, DPe3 = -71.0 / 16695.0, DPe4 = 71.0 / 1920.0,
                    DPe5 = -17253.0 / 339200.0, DPe6 = 22.0 / 525.0, DPe7 = -1.0 / 40.0;

struct orc_cloud {
    int n;
    double (*pos)[3], (*vel)[3], (*pos2)[3], (*vel2)[3];
    double *mass;
    int *charge;
    double lifetime;
    double (*k[7])[6];
    double (*derivs)[3];
    double h, hnext, errold, last_err;
    int reject;
    double rtol, atol, beta;
    int adaptive, time_mode;
    orc_fields f;
    double ramp_Vrf0, ramp_Vdc0;
    long steps, attempts;
};

/* bin/ion.cpp:3-6: mass = _mass_amu*amu evaluated in long double, then
 * narrowed to double by IonCloud::CopyParticlesToVectors (ioncloud.cpp:411). */
double orc_mass_kg(double mass_amu)
{
    long double m = (long double)mass_amu * amu;
    return (double)m;
}

/* SWIFT amplitude at t: samples at t0 + k dt, linear interpolation, 0 outside */
static double swift_amp(const orc_fields *f, double t)
{
    double s = (t - f->swift_t0) / f->swift_dt;
    if (!(s >= 0.0) || s > (double)(f->swift_n - 1))
        return 0.0;
    int i = (int)s;
    if (i > f->swift_n - 2)
        i = f->swift_n - 2;
    double w = s - (double)i;
    double a = f->swift[i], b = f->swift[i + 1];
    return a + w * (b - a);
}

/* field map at (r, z): bilinear interpolation of the (E_r, E_z) grid, 0 outside */
static void emap_field(const orc_fields *f, double r, double z, double *er, double *ez)
{
    *er = 0.0;
    *ez = 0.0;
    double u = r / f->emap_dr, v = (z - f->emap_z0) / f->emap_dz;
    if (!(u >= 0.0) || u > (double)(f->emap_nr - 1) || !(v >= 0.0) || v > (double)(f->emap_nz - 1))
        return;
    int i = (int)u, j = (int)v;
    if (i > f->emap_nr - 2) i = f->emap_nr - 2;
    if (j > f->emap_nz - 2) j = f->emap_nz - 2;
    double wu = u - (double)i, wv = v - (double)j;
    long long nz = f->emap_nz;
    const double *e00 = f->emap + 2 * ((long long)i * nz + j);
    const double *e01 = e00 + 2, *e10 = e00 + 2 * nz, *e11 = e10 + 2;
    double out[2];
    for (int k = 0; k < 2; k++) {
        double lo = e00[k] + wu * (e10[k] - e00[k]);
        double hi = e01[k] + wu * (e11[k] - e01[k]);
        out[k] = lo + wv * (hi - lo);
    }
    *er = out[0];
    *ez = out[1];
}

/* bin/force.cpp:34-82 (trap/TOF) followed by the extension terms and then
 * bin/force.cpp:147-175 (Coulomb).  derivs is an acceleration. */
void orc_force(int n, const double *pos_, const double *vel_, const double *mass,
               const int *charge, const orc_fields *f, double t, double *derivs_)
{
    const double (*pos)[3] = (const double (*)[3])pos_;
    const double (*vel)[3] = (const double (*)[3])vel_;
    double (*derivs)[3] = (double (*)[3])derivs_;

    double kVdc = f->kappa * f->Vdc;                 /* force.cpp:40 */
    double Vrf = f->Vrf;
    double r02 = f->r0 * f->r0;
    double z02 = f->z0 * f->z0;
    double omega_rf = f->freq_rf * 2 * pi;           /* force.cpp:44 */
    double cpen = (f->V0 != 0.0) ? f->V0 / (2.0 * f->d0 * f->d0) : 0.0;
    /* drive phases 2 pi (f t + rate t^2 / 2) + phi; rate = 0 adds an exact 0 */
    double cdip = (f->Ed != 0.0)
                      ? f->Ed * cos(2.0 * pi * f->fd * t + pi * f->fd_rate * t * t + f->phid) : 0.0;
    double cqex = (f->Uq != 0.0) ? (f->Uq / (f->rq * f->rq)) *
                                       cos(2.0 * pi * f->fq * t + pi * f->fq_rate * t * t + f->phiq)
                                 : 0.0;
    /* octupole, rotating wall, axial drive, buffer gas (extensions, see the header) */
    double coct = (f->Uo != 0.0)
                      ? (f->Uo / (f->ro * f->ro * f->ro * f->ro)) * cos(2.0 * pi * f->fo * t + f->phio)
                      : 0.0;
    double cw = 0.0, sw = 0.0;
    if (f->Uw != 0.0) {
        double psi = 2.0 * pi * f->fw * t + f->phiw;
        double amp = (f->rw_order == 1) ? f->Uw / f->rw : f->Uw / (f->rw * f->rw);
        double ang = (f->rw_order == 1) ? psi : 2.0 * psi;
        cw = amp * cos(ang);
        sw = amp * sin(ang);
    }
    double cax = (f->Eax != 0.0) ? f->Eax * cos(2.0 * pi * f->fax * t + f->phiax) : 0.0;
    double invK = (f->K0 != 0.0 && f->p_mbar != 0.0)
                      ? (f->p_mbar / 1013.25) * (273.15 / f->T_gas) / f->K0 : 0.0;
    double uswift = (f->swift_n > 0) ? swift_amp(f, t) : 0.0;
    double c4k = (f->V0 != 0.0 && f->C4 != 0.0)
                     ? f->V0 * f->C4 / (2.0 * f->d0 * f->d0 * f->d0 * f->d0) : 0.0;
    double c6k = (f->V0 != 0.0 && f->C6 != 0.0)
                     ? f->V0 * f->C6 / (2.0 * f->d0 * f->d0 * f->d0 * f->d0 * f->d0 * f->d0) : 0.0;

    for (int j_ = 0; j_ < n; j_++) {
        double x = pos[j_][0], y = pos[j_][1], z = pos[j_][2];
        double vx = vel[j_][0], vy = vel[j_][1], vz = vel[j_][2];
        double qDmassa = charge[j_] * el_charge / mass[j_];
        derivs[j_][0] = 0;
        derivs[j_][1] = 0;
        derivs[j_][2] = 0;
        if (f->trap) {                                /* force.cpp:70-75 */
            derivs[j_][0] += qDmassa * (kVdc / z02 - Vrf * cos(omega_rf * t) / r02) * x;
            derivs[j_][1] += qDmassa * (kVdc / z02 + Vrf * cos(omega_rf * t) / r02) * y;
            derivs[j_][2] += qDmassa * (-2.0 * kVdc / z02) * z;
        }
        if (f->tof)                                   /* force.cpp:77-80 */
            derivs[j_][0] += qDmassa * f->E_kick;
        /* extensions (absent from the fork) */
        if (cpen != 0.0) {
            derivs[j_][0] += qDmassa * cpen * x;
            derivs[j_][1] += qDmassa * cpen * y;
            derivs[j_][2] += qDmassa * (-2.0 * cpen) * z;
        }
        if (f->Bz != 0.0) {
            derivs[j_][0] += qDmassa * f->Bz * vy;
            derivs[j_][1] -= qDmassa * f->Bz * vx;
        }
        if (cdip != 0.0)
            derivs[j_][0] += qDmassa * cdip;
        if (cqex != 0.0) {
            derivs[j_][0] += qDmassa * cqex * x;
            derivs[j_][1] -= qDmassa * cqex * y;
        }
        if (f->delta != 0.0) {
            derivs[j_][0] -= f->delta * vx;
            derivs[j_][1] -= f->delta * vy;
            derivs[j_][2] -= f->delta * vz;
        }
        if (f->B2 != 0.0) {
            double b2z = f->B2 * z;
            double bx = -b2z * x, by = -b2z * y;
            double bz = f->B2 * (z * z - 0.5 * (x * x + y * y));
            derivs[j_][0] += qDmassa * (vy * bz - vz * by);
            derivs[j_][1] += qDmassa * (vz * bx - vx * bz);
            derivs[j_][2] += qDmassa * (vx * by - vy * bx);
        }
        if (coct != 0.0) {
            derivs[j_][0] += qDmassa * coct * (x * (x * x - 3.0 * y * y));
            derivs[j_][1] += qDmassa * coct * (y * (y * y - 3.0 * x * x));
        }
        if (f->Uw != 0.0) {
            if (f->rw_order == 1) {
                derivs[j_][0] += qDmassa * cw;
                derivs[j_][1] += qDmassa * sw;
            } else {
                derivs[j_][0] += qDmassa * (cw * x + sw * y);
                derivs[j_][1] += qDmassa * (sw * x - cw * y);
            }
        }
        if (cax != 0.0)
            derivs[j_][2] += qDmassa * cax;
        if (invK != 0.0) {
            double g = qDmassa * invK;
            derivs[j_][0] -= g * vx;
            derivs[j_][1] -= g * vy;
            derivs[j_][2] -= g * vz;
        }
        if (f->swift_n > 0) {
            if (f->swift_kind == 1) {
                derivs[j_][0] += qDmassa * (uswift / f->swift_r);
            } else {
                double cs = uswift / (f->swift_r * f->swift_r);
                derivs[j_][0] += qDmassa * cs * x;
                derivs[j_][1] -= qDmassa * cs * y;
            }
        }
        if (c4k != 0.0 || c6k != 0.0) {
            /* -grad of c4k r^4 P4 and c6k r^6 P6 (see the header) */
            double r2 = x * x + y * y, z2 = z * z;
            if (c4k != 0.0) {
                double cr = c4k * (6.0 * z2 - 1.5 * r2);
                double cz = c4k * (6.0 * r2 - 4.0 * z2);
                derivs[j_][0] += qDmassa * cr * x;
                derivs[j_][1] += qDmassa * cr * y;
                derivs[j_][2] += qDmassa * cz * z;
            }
            if (c6k != 0.0) {
                double z4 = z2 * z2, r4 = r2 * r2, zr = z2 * r2;
                double cr = c6k * (15.0 * z4 - 22.5 * zr + 1.875 * r4);
                double cz = c6k * (30.0 * zr - 6.0 * z4 - 11.25 * r4);
                derivs[j_][0] += qDmassa * cr * x;
                derivs[j_][1] += qDmassa * cr * y;
                derivs[j_][2] += qDmassa * cz * z;
            }
        }
        if (f->emap_nr > 0) {
            double r = sqrt(x * x + y * y), er, ez;
            emap_field(f, r, z, &er, &ez);
            double g = qDmassa * f->emap_scale;
            if (r > 0.0) {
                derivs[j_][0] += g * (er * (x / r));
                derivs[j_][1] += g * (er * (y / r));
            }
            derivs[j_][2] += g * ez;
        }
    }

    if (f->coulomb) {                                 /* force.cpp:147-175 */
        for (int j_ = 0; j_ < n; j_++) {
            double fx = 0.0, fy = 0.0, fz = 0.0;
            double kTqiTqjDmj = f->coulomb_scale * ke * el_charge * charge[j_] * el_charge / mass[j_];
            double bjx = pos[j_][0], bjy = pos[j_][1], bjz = pos[j_][2];
            for (int i_ = 0; i_ < n; i_++) {
                if (i_ != j_) {
                    double rx = pos[i_][0] - bjx;
                    double ry = pos[i_][1] - bjy;
                    double rz = pos[i_][2] - bjz;
                    double distSqr = rx * rx + ry * ry + rz * rz;
                    double invDist = 1. / sqrt(distSqr);
                    double invDistCube = (invDist * invDist * invDist);
                    fx += rx * invDistCube * kTqiTqjDmj * charge[i_];
                    fy += ry * invDistCube * kTqiTqjDmj * charge[i_];
                    fz += rz * invDistCube * kTqiTqjDmj * charge[i_];
                }
            }
            derivs[j_][0] -= fx;
            derivs[j_][1] -= fy;
            derivs[j_][2] -= fz;
        }
    }
}

/* force.cpp:150-170 for selected targets: returns f_j (the sum that the
 * reference subtracts from derivs_j). */
void orc_coulomb_rows(int n, const double *pos_, const double *mass, const int *charge,
                      double coulomb_scale, const int *rows, int nrows, double *out)
{
    const double (*pos)[3] = (const double (*)[3])pos_;
    for (int r = 0; r < nrows; r++) {
        int j_ = rows[r];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        double kTqiTqjDmj = coulomb_scale * ke * el_charge * charge[j_] * el_charge / mass[j_];
        double bjx = pos[j_][0], bjy = pos[j_][1], bjz = pos[j_][2];
        for (int i_ = 0; i_ < n; i_++) {
            if (i_ != j_) {
                double rx = pos[i_][0] - bjx;
                double ry = pos[i_][1] - bjy;
                double rz = pos[i_][2] - bjz;
                double distSqr = rx * rx + ry * ry + rz * rz;
                double invDist = 1. / sqrt(distSqr);
                double invDistCube = (invDist * invDist * invDist);
                fx += rx * invDistCube * kTqiTqjDmj * charge[i_];
                fy += ry * invDistCube * kTqiTqjDmj * charge[i_];
                fz += rz * invDistCube * kTqiTqjDmj * charge[i_];
            }
        }
        out[3 * r + 0] = fx;
        out[3 * r + 1] = fy;
        out[3 * r + 2] = fz;
    }
}

/* Interface semantics of cunbody1_force (bin/cunbody.h:11-33, as called at
 * bin/cunbody.cpp:65-66,88): softened direct sum with source weights mj.
 * The reference passes the same positions as xj and, copied, as xi; then
 * the pair j == i is skipped (its numerator is 0, so with eps2 > 0 it adds
 * nothing, and with eps2 == 0 it would be 0/0). */
void orc_cunbody_force(const double *xj_, const double *mj, const double *xi_, double eps2,
                       double *ai_, int ni, int nj)
{
    const double (*xj)[3] = (const double (*)[3])xj_;
    const double (*xi)[3] = (const double (*)[3])xi_;
    double (*ai)[3] = (double (*)[3])ai_;
    const int alias = ni == nj && (xj_ == xi_ || memcmp(xj_, xi_, sizeof(double) * 3 * (size_t)ni) == 0);
    for (int i = 0; i < ni; i++) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = 0; j < nj; j++) {
            if (alias && j == i) continue;
            double rx = xj[j][0] - xi[i][0];
            double ry = xj[j][1] - xi[i][1];
            double rz = xj[j][2] - xi[i][2];
            double r2 = rx * rx + ry * ry + rz * rz + eps2;
            double inv = 1.0 / sqrt(r2);
            double inv3 = (inv * inv * inv);
            ax += rx * inv3 * mj[j];
            ay += ry * inv3 * mj[j];
            az += rz * inv3 * mj[j];
        }
        ai[i][0] = ax;
        ai[i][1] = ay;
        ai[i][2] = az;
    }
}

/* bin/ode.cpp:426-457 (Numerical Recipes 3rd ed. Controller::success).
 * The reference fixes beta = 0 and documents beta = 0.04..0.08 as the switch
 * to PI control (bin/ode.cpp:427-432); alpha = 0.2 - 0.75 beta follows it. */
int orc_controller_beta(double err, double beta, double *h, double *hnext, double *errold,
                        int *reject)
{
    const double alpha = 0.2 - beta * 0.75, safe = 0.9, minscale = 0.2, maxscale = 10.0;
    double scale;
    if (err <= 1.0) {
        if (err == 0.0)
            scale = maxscale;
        else {
            scale = safe * pow(err, -alpha) * pow(*errold, beta);
            if (scale < minscale) scale = minscale;
            if (scale > maxscale) scale = maxscale;
        }
        if (*reject)
            *hnext = *h * fmin(scale, 1.0);
        else
            *hnext = *h * scale;
        *errold = fmax(err, 1.0e-4);
        *reject = 0;
        return 1;
    }
    scale = fmax(safe * pow(err, -alpha), minscale);
    *h *= scale;
    *reject = 1;
    return 0;
}

int orc_controller(double err, double *h, double *hnext, double *errold, int *reject)
{
    return orc_controller_beta(err, 0.0, h, hnext, errold, reject);
}

void orc_set_beta(orc_cloud *c, double beta)
{
    c->beta = beta;
}

static void *xalloc(size_t bytes)
{
    void *p = calloc(1, bytes ? bytes : 1);
    if (!p) abort();
    return p;
}

orc_cloud *orc_create(int n, const double *pos, const double *vel, const double *mass,
                      const int *charge, const orc_fields *f, double lifetime, double h0,
                      double rtol, double atol, int adaptive, int time_mode)
{
    orc_cloud *c = (orc_cloud *)xalloc(sizeof(orc_cloud));
    size_t v3 = sizeof(double) * 3 * (size_t)n;
    c->n = n;
    c->pos = xalloc(v3);
    c->vel = xalloc(v3);
    c->pos2 = xalloc(v3);
    c->vel2 = xalloc(v3);
    c->derivs = xalloc(v3);
    for (int s = 0; s < 7; s++) c->k[s] = xalloc(2 * v3);
    c->mass = xalloc(sizeof(double) * (size_t)n);
    c->charge = xalloc(sizeof(int) * (size_t)n);
    memcpy(c->pos, pos, v3);
    memcpy(c->vel, vel, v3);
    memcpy(c->pos2, pos, v3);
    memcpy(c->vel2, vel, v3);
    memcpy(c->mass, mass, sizeof(double) * (size_t)n);
    memcpy(c->charge, charge, sizeof(int) * (size_t)n);
    c->lifetime = lifetime;
    /* _ode_vars::Init_ode, ode.cpp:539-546 */
    c->errold = 1.0e-4;
    c->reject = 0;
    c->h = h0;
    c->hnext = h0;
    c->rtol = rtol;
    c->atol = atol;
    c->adaptive = adaptive;
    c->time_mode = time_mode;
    c->f = *f;
    c->ramp_Vrf0 = f->Vrf;
    c->ramp_Vdc0 = f->Vdc;
    return c;
}

void orc_destroy(orc_cloud *c)
{
    if (!c) return;
    free(c->pos); free(c->vel); free(c->pos2); free(c->vel2); free(c->derivs);
    for (int s = 0; s < 7; s++) free(c->k[s]);
    free(c->mass); free(c->charge);
    free(c);
}

void orc_set_fields(orc_cloud *c, const orc_fields *f)
{
    c->f = *f;
    c->ramp_Vrf0 = f->Vrf;
    c->ramp_Vdc0 = f->Vdc;
}

static double stage_time(const orc_cloud *c, double cs)
{
    return c->time_mode == ORC_TIME_STAGE ? c->lifetime + cs * c->h : c->lifetime;
}

static void eval(orc_cloud *c, double cs)
{
    orc_force(c->n, &c->pos2[0][0], &c->vel2[0][0], c->mass, c->charge, &c->f,
              stage_time(c, cs), &c->derivs[0][0]);
}

/* bin/ode.cpp:61-218 */
double orc_dp5_attempt(orc_cloud *c)
{
    int n = c->n;
    double h = c->h;
    double (*k1)[6] = c->k[0], (*k2)[6] = c->k[1], (*k3)[6] = c->k[2], (*k4)[6] = c->k[3];
    double (*k5)[6] = c->k[4], (*k6)[6] = c->k[5], (*k7)[6] = c->k[6];
    double (*pos)[3] = c->pos, (*vel)[3] = c->vel, (*pos2)[3] = c->pos2, (*vel2)[3] = c->vel2;
    double (*d)[3] = c->derivs;
    double err = 0.0, sk;
    double yerr[6];
    c->attempts++;

    for (int j = 0; j < n; j++)
        for (int i = 0; i < 3; i++) k1[j][i] = vel[j][i];
    memcpy(pos2, pos, sizeof(double) * 3 * (size_t)n);
    memcpy(vel2, vel, sizeof(double) * 3 * (size_t)n);

    eval(c, 0.0);                                     /* stage 1 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k1[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++) pos2[j][i] = pos[j][i] + h * (DPa21 * k1[j][i]);
        for (int i = 0; i < 3; i++) k2[j][i] = vel2[j][i] = vel[j][i] + h * (DPa21 * k1[j][3 + i]);
    }
    eval(c, DPc2);                                    /* stage 2 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k2[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPa31 * k1[j][i] + DPa32 * k2[j][i]);
        for (int i = 0; i < 3; i++)
            k3[j][i] = vel2[j][i] = vel[j][i] + h * (DPa31 * k1[j][3 + i] + DPa32 * k2[j][3 + i]);
    }
    eval(c, DPc3);                                    /* stage 3 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k3[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPa41 * k1[j][i] + DPa42 * k2[j][i] + DPa43 * k3[j][i]);
        for (int i = 0; i < 3; i++)
            k4[j][i] = vel2[j][i] = vel[j][i] + h * (DPa41 * k1[j][3 + i] + DPa42 * k2[j][3 + i] +
                                                     DPa43 * k3[j][3 + i]);
    }
    eval(c, DPc4);                                    /* stage 4 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k4[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPa51 * k1[j][i] + DPa52 * k2[j][i] + DPa53 * k3[j][i] +
                                          DPa54 * k4[j][i]);
        for (int i = 0; i < 3; i++)
            k5[j][i] = vel2[j][i] = vel[j][i] + h * (DPa51 * k1[j][3 + i] + DPa52 * k2[j][3 + i] +
                                                     DPa53 * k3[j][3 + i] + DPa54 * k4[j][3 + i]);
    }
    eval(c, DPc5);                                    /* stage 5 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k5[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPa61 * k1[j][i] + DPa62 * k2[j][i] + DPa63 * k3[j][i] +
                                          DPa64 * k4[j][i] + DPa65 * k5[j][i]);
        for (int i = 0; i < 3; i++)
            k6[j][i] = vel2[j][i] =
                vel[j][i] + h * (DPa61 * k1[j][3 + i] + DPa62 * k2[j][3 + i] + DPa63 * k3[j][3 + i] +
                                 DPa64 * k4[j][3 + i] + DPa65 * k5[j][3 + i]);
    }
    eval(c, DPc6);                                    /* stage 6 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k6[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPa71 * k1[j][i] + DPa73 * k3[j][i] + DPa74 * k4[j][i] +
                                          DPa75 * k5[j][i] + DPa76 * k6[j][i]);
        for (int i = 0; i < 3; i++)
            k7[j][i] = vel2[j][i] =
                vel[j][i] + h * (DPa71 * k1[j][3 + i] + DPa73 * k3[j][3 + i] + DPa74 * k4[j][3 + i] +
                                 DPa75 * k5[j][3 + i] + DPa76 * k6[j][3 + i]);
    }
    eval(c, DPc7);                                    /* stage 7 */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 3; i++) k7[j][3 + i] = d[j][i];
        for (int i = 0; i < 3; i++)
            pos2[j][i] = pos[j][i] + h * (DPb1 * k1[j][i] + DPb3 * k3[j][i] + DPb4 * k4[j][i] +
                                          DPb5 * k5[j][i] + DPb6 * k6[j][i]);
        for (int i = 0; i < 3; i++)
            vel2[j][i] = vel[j][i] + h * (DPb1 * k1[j][3 + i] + DPb3 * k3[j][3 + i] +
                                          DPb4 * k4[j][3 + i] + DPb5 * k5[j][3 + i] +
                                          DPb6 * k6[j][3 + i]);
    }
    /* ode.cpp:189-205: serial, ion-major then component order */
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < 6; i++)
            yerr[i] = h * (DPe1 * k1[j][i] + DPe3 * k3[j][i] + DPe4 * k4[j][i] + DPe5 * k5[j][i] +
                           DPe6 * k6[j][i] + DPe7 * k7[j][i]);
        for (int i = 0; i < 3; i++) {
            sk = c->atol + c->rtol * fmax(fabs(pos[j][i]), fabs(pos2[j][i]));
            err = err + (yerr[i] / sk) * (yerr[i] / sk);
        }
        for (int i = 0; i < 3; i++) {
            sk = c->atol + c->rtol * fmax(fabs(vel[j][i]), fabs(vel2[j][i]));
            err = err + (yerr[3 + i] / sk) * (yerr[3 + i] / sk);
        }
    }
    err = sqrt(err / (6 * n));                        /* ode.cpp:215 */
    c->last_err = err;
    return err;
}

/* bin/ode.cpp:459-494 (DP5 branch) */
int orc_step(orc_cloud *c)
{
    for (;;) {
        double err = orc_dp5_attempt(c);
        if (c->adaptive) {
            if (orc_controller_beta(err, c->beta, &c->h, &c->hnext, &c->errold, &c->reject)) break;
            if (fabs(c->h) <= fabs(c->pos[0][0]) * DBL_EPSILON) return -1;
        } else
            break;
    }
    memcpy(c->pos, c->pos2, sizeof(double) * 3 * (size_t)c->n);
    memcpy(c->vel, c->vel2, sizeof(double) * 3 * (size_t)c->n);
    c->lifetime = c->lifetime + c->h;
    c->h = c->hnext;
    c->steps++;
    return 0;
}

/* bin/ion_fly.cpp:157-255, without printing: ramp then step while
 * lifetime < endtime.  The ramp interval is [ramp_t0, ramp_t1]. */
long orc_advance(orc_cloud *c, double t_end, long max_steps, double *h_hist, long cap)
{
    long taken = 0;
    while (c->lifetime < t_end && taken < max_steps) {
        if (c->n == 0) break;
        if (c->f.ramp) {
            double ramp = (c->lifetime - c->f.ramp_t0) / (c->f.ramp_t1 - c->f.ramp_t0);
            c->f.Vrf = c->ramp_Vrf0 * (1 - ramp) + c->f.newVrf * ramp;
            c->f.Vdc = c->ramp_Vdc0 * (1 - ramp) + c->f.newVdc * ramp;
        }
        if (h_hist && taken < cap) h_hist[taken] = c->h;
        if (orc_step(c) != 0) return -1;
        taken++;
    }
    return taken;
}

void orc_get_state(const orc_cloud *c, double *pos, double *vel, double *s)
{
    size_t v3 = sizeof(double) * 3 * (size_t)c->n;
    if (pos) memcpy(pos, c->pos, v3);
    if (vel) memcpy(vel, c->vel, v3);
    if (s) {
        s[0] = c->lifetime;
        s[1] = c->h;
        s[2] = c->hnext;
        s[3] = c->errold;
        s[4] = c->reject;
        s[5] = (double)c->steps;
        s[6] = (double)c->attempts;
        s[7] = c->last_err;
    }
}
