// ref_harness.cpp -- drives the reference's own CPU force()/step() (compiled
// from the reference sources with -D__CPUNBODY_ON__, see oracle/Makefile) on
// inputs read from a binary file.  TEST INFRASTRUCTURE: it produces golden
// vectors and the reference CPU baseline timing; nothing in the product
// links it.
//
// It fills IonCloud (bin/ioncloud.h:17-71) and _ode_vars (bin/ode.h:58-133)
// the way InitializePoolVectors/CopyParticlesToVectors (bin/ioncloud.cpp:
// 383-414) and move_particles' first call (bin/ion_fly.cpp:26-62) do, then
// calls force() (bin/force.cpp:34) or step() (bin/ode.cpp:459).
//
// Input (little endian):
//   int32  magic 0x53424B31, mode, n, nsteps, trap, tof, coulomb, adaptive
//   double lifetime, h0, rtol, atol, kappa, Vdc, Vrf, r0, z0, freq_rf,
//          E_kick, coulomb_scale
//   double mass_amu[n]; int32 charge[n]; double pos[n][3]; double vel[n][3]
// Modes: 1 force -> derivs[n][3]
//        2 steps -> pos[n][3] vel[n][3] scalars[8] hist[nsteps][2]
//        3 time  -> one JSON line on stdout timing nsteps force() calls
//        4 time  -> one JSON line on stdout timing nsteps step() calls
//        5, 6 output -> the reference's own output files under the prefix
//                 given as the second argument (5: <prefix>_pAll.txt, 6: one
//                 <prefix>_p<k>.txt per ion; both with <prefix>_cloud<k>.txt):
//                 PrintParticles + PrintCloud at lifetime, then nsteps times
//                 pos += vel * 1e-7, vel *= 0.75, lifetime += 1e-7 and print
//                 again, then Delete().  Ion k is named "<mass_amu>u" ("%g").
#include "ode.h"
#include "force.h"

#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>
#include <vector>

static void die(const char *m)
{
    fprintf(stderr, "ref_harness: %s\n", m);
    exit(2);
}

static void rd(FILE *f, void *p, size_t bytes)
{
    if (bytes && fread(p, 1, bytes, f) != bytes) die("short read");
}

static void wr(FILE *f, const void *p, size_t bytes)
{
    if (bytes && fwrite(p, 1, bytes, f) != bytes) die("short write");
}

static double now()
{
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

int main(int argc, char **argv)
{
    if (argc != 3) die("usage: ref_harness <input.bin> <output.bin|->");
    FILE *in = fopen(argv[1], "rb");
    if (!in) die("cannot open input");
    int hdr[8];
    rd(in, hdr, sizeof(hdr));
    if (hdr[0] != 0x53424B31) die("bad magic");
    int mode = hdr[1], n = hdr[2], nsteps = hdr[3];
    double p[12];
    rd(in, p, sizeof(p));
    std::vector<double> mass_amu(n), pos(3 * n), vel(3 * n);
    std::vector<int> charge(n);
    rd(in, &mass_amu[0], sizeof(double) * n);
    rd(in, &charge[0], sizeof(int) * n);
    rd(in, &pos[0], sizeof(double) * 3 * n);
    rd(in, &vel[0], sizeof(double) * 3 * n);
    fclose(in);

    if (mode == 5 || mode == 6) {
        // bin/ioncloud.cpp:29-33 (Create), 185 (use_particles_files), 270-280
        // (AddParticle, which opens the files), 383-414 (pool vectors)
        IonCloud *oc = new IonCloud;
        oc->Create(argv[2]);
        oc->use_particles_files(mode == 6);
        for (int i = 0; i < n; i++) {
            Ion ion((long double)mass_amu[i], charge[i]);
            char name[64];
            snprintf(name, sizeof(name), "%gu", mass_amu[i]);
            ion.SetName(name);
            oc->AddParticle(Particle(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], vel[3 * i],
                                     vel[3 * i + 1], vel[3 * i + 2]),
                            ion);
        }
        oc->InitializePoolVectors();
        oc->CopyParticlesToVectors();
        oc->lifetime = p[0];
        oc->PrintParticles();
        oc->PrintCloud();
        for (int s = 0; s < nsteps; s++) {
            for (int i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) {
                    oc->pos[i][c] += oc->vel[i][c] * 1e-7;
                    oc->vel[i][c] *= 0.75;
                }
            oc->lifetime += 1e-7;
            oc->PrintParticles();
            oc->PrintCloud();
        }
        oc->Delete();
        fflush(stdout);
        _exit(0);
    }

    IonCloud *cloud = new IonCloud;
    _ode_vars *odev = new _ode_vars;
    cloud->ions = new vector<Ion>;
    for (int i = 0; i < n; i++) cloud->ions->push_back(Ion((long double)mass_amu[i], charge[i]));
    cloud->nrparticles = n;
    cloud->lifetime = p[0];
    cloud->mass.resize(n);
    cloud->charge.resize(n);
    cloud->pos = new double[n][3];
    cloud->pos2 = new double[n][3];
    cloud->vel = new double[n][3];
    cloud->vel2 = new double[n][3];
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) {
            cloud->pos[i][c] = cloud->pos2[i][c] = pos[3 * i + c];
            cloud->vel[i][c] = cloud->vel2[i][c] = vel[3 * i + c];
        }
        cloud->mass[i] = (*cloud->ions)[i].Getmass();
        cloud->charge[i] = (*cloud->ions)[i].Getcharge();
    }

    odev->Init_ode(5, p[1], hdr[7] != 0);
    odev->rk_rtol = p[2];
    odev->rk_atol = p[3];
    odev->Initpoolvectors(n);
    odev->forcev.trap = hdr[4] != 0;
    odev->forcev.tof = hdr[5] != 0;
    odev->forcev.coulomb_interaction = hdr[6] != 0;
    odev->forcev.trap_param.kappa = p[4];
    odev->forcev.trap_param.Vdc = p[5];
    odev->forcev.trap_param.Vrf = p[6];
    odev->forcev.trap_param.r0 = p[7];
    odev->forcev.trap_param.z0 = p[8];
    odev->forcev.trap_param.freq_rf = p[9];
    odev->forcev.trap_param.E_kick = p[10];
    odev->forcev.coulomb_scale = p[11];

    if (mode == 1) {
        force(*cloud, *odev);
        FILE *out = fopen(argv[2], "wb");
        if (!out) die("cannot open output");
        wr(out, &odev->forcev.derivs[0][0], sizeof(double) * 3 * n);
        fclose(out);
    } else if (mode == 2) {
        std::vector<double> hist(2 * (size_t)nsteps);
        for (int s = 0; s < nsteps; s++) {
            hist[2 * s] = odev->h;
            step(*cloud, *odev);
            hist[2 * s + 1] = cloud->lifetime;
        }
        double sc[8] = {cloud->lifetime, odev->h, odev->hnext, odev->errold,
                        (double)odev->reject, (double)odev->countstep, odev->err, 0.0};
        FILE *out = fopen(argv[2], "wb");
        if (!out) die("cannot open output");
        wr(out, &cloud->pos[0][0], sizeof(double) * 3 * n);
        wr(out, &cloud->vel[0][0], sizeof(double) * 3 * n);
        wr(out, sc, sizeof(sc));
        if (nsteps) wr(out, &hist[0], sizeof(double) * 2 * nsteps);
        fclose(out);
    } else if (mode == 3 || mode == 4) {
        double t0 = now();
        for (int s = 0; s < nsteps; s++) {
            if (mode == 3)
                force(*cloud, *odev);
            else
                step(*cloud, *odev);
        }
        double dt = now() - t0;
        double chk = 0.0;
        for (int i = 0; i < n; i++) chk += odev->forcev.derivs[i][0];
        printf("{\"mode\": %d, \"n\": %d, \"calls\": %d, \"seconds\": %.6f, \"attempts\": %d, "
               "\"checksum\": %.17g}\n",
               mode, n, nsteps, dt, odev->countstep, chk);
    } else
        die("bad mode");
    fflush(stdout);
    _exit(0);
}
