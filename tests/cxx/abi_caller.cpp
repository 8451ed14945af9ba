// abi_caller.cpp -- a C++ caller of libsimbuca_hip.so built with g++ against
// include/simbuca_hip.h only (no HIP headers), the way the reference links a
// backend.
//
//   abi_caller cunbody <in> <out>  force() composed as bin/force.cpp:34-105 does
//                                  with the CUNBODY backend: cunbody1_force_begin,
//                                  the trap/TOF loop on the host, cunbody1_force_end,
//                                  then derivs -= coulomb_scale e^2 ke / m * ai
//   abi_caller engine <in> <out>   nsteps calls of sbk_step (FP64-exact, frozen time)
//
// Input/output use the binary layout of oracle/ref_harness.cpp (header of 8
// int32, 12 doubles, mass_amu[n], charge[n], pos[n][3], vel[n][3]).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "simbuca_hip.h"

static void die(const char *m)
{
    std::fprintf(stderr, "abi_caller: %s (%s)\n", m, sbk_last_error());
    std::exit(2);
}

static void rd(FILE *f, void *p, size_t n)
{
    if (n && std::fread(p, 1, n, f) != n) die("short read");
}

int main(int argc, char **argv)
{
    if (argc != 4) die("usage: abi_caller cunbody|engine <in> <out>");
    FILE *in = std::fopen(argv[2], "rb");
    if (!in) die("cannot open input");
    int hdr[8];
    double p[12];
    rd(in, hdr, sizeof(hdr));
    rd(in, p, sizeof(p));
    const int n = hdr[2], nsteps = hdr[3];
    std::vector<double> amu(n), pos(3 * n), vel(3 * n), mass(n);
    std::vector<int> charge(n);
    rd(in, amu.data(), 8 * n);
    rd(in, charge.data(), 4 * n);
    rd(in, pos.data(), 24 * n);
    rd(in, vel.data(), 24 * n);
    std::fclose(in);
    const double el_charge = 1.60217733e-19, amu_kg = 1.6605402E-27, ke = 8.987551787E9,
                 pi = 3.14159265358979323;
    for (int i = 0; i < n; ++i) mass[i] = (double)((long double)amu[i] * amu_kg);
    const int trap = hdr[4], tof = hdr[5], coulomb = hdr[6], adaptive = hdr[7];
    const double lifetime = p[0], h0 = p[1], rtol = p[2], atol = p[3], kappa = p[4], Vdc = p[5],
                 Vrf = p[6], r0 = p[7], z0 = p[8], freq_rf = p[9], E_kick = p[10],
                 coulomb_scale = p[11];
    std::vector<double> out;

    if (!std::strcmp(argv[1], "cunbody")) {
        // _cunbody::AllocArrays / begin / end (bin/cunbody.cpp:31-41,65-66,88)
        std::vector<double> m(n, 1.0), pos_tmp(pos), ai(3 * n, 0.0), derivs(3 * n, 0.0);
        const double eps = (double)1e-23f;
        auto xj = reinterpret_cast<double(*)[3]>(pos.data());
        auto xi = reinterpret_cast<double(*)[3]>(pos_tmp.data());
        auto a = reinterpret_cast<double(*)[3]>(ai.data());
        if (coulomb) cunbody1_force_begin(xj, m.data(), xi, eps, a, n, n);
        const double kVdc = kappa * Vdc, r02 = r0 * r0, z02 = z0 * z0, omega_rf = freq_rf * 2 * pi;
        for (int j = 0; j < n; ++j) {            // bin/force.cpp:58-82
            const double qD = charge[j] * el_charge / mass[j];
            if (trap) {
                derivs[3 * j + 0] += qD * (kVdc / z02 - Vrf * std::cos(omega_rf * lifetime) / r02) * pos[3 * j];
                derivs[3 * j + 1] += qD * (kVdc / z02 + Vrf * std::cos(omega_rf * lifetime) / r02) * pos[3 * j + 1];
                derivs[3 * j + 2] += qD * (-2.0 * kVdc / z02) * pos[3 * j + 2];
            }
            if (tof) derivs[3 * j] += qD * E_kick;
        }
        if (coulomb) {
            cunbody1_force_end(xj, m.data(), xi, eps, a, n, n);
            const double el2ke = el_charge * el_charge * ke;
            for (int j = 0; j < n; ++j) {         // bin/force.cpp:87-93
                const double s = coulomb_scale * el2ke / mass[j];
                for (int k = 0; k < 3; ++k) derivs[3 * j + k] -= s * ai[3 * j + k];
            }
        }
        out = derivs;
    } else if (!std::strcmp(argv[1], "engine")) {
        sbk_config cfg;
        sbk_default_config(&cfg);
        cfg.precision = SBK_FP64_EXACT;
        cfg.time_mode = SBK_TIME_FROZEN;
        cfg.adaptive = adaptive;
        cfg.rtol = rtol;
        cfg.atol = atol;
        cfg.h0 = h0;
        sbk_engine *e = nullptr;
        if (sbk_create(&cfg, &e)) die("sbk_create");
        sbk_fields f;
        sbk_default_fields(&f);
        f.trap = trap; f.kappa = kappa; f.Vdc = Vdc; f.Vrf = Vrf; f.r0 = r0; f.z0 = z0;
        f.freq_rf = freq_rf; f.tof = tof; f.E_kick = E_kick; f.coulomb = coulomb;
        f.coulomb_scale = coulomb_scale;
        if (sbk_add_cloud(e, n, pos.data(), vel.data(), mass.data(), charge.data(), &f, lifetime,
                          nullptr))
            die("sbk_add_cloud");
        for (int s = 0; s < nsteps; ++s)
            if (sbk_step(e)) die("sbk_step");
        out.assign(6 * n + 2, 0.0);
        if (sbk_get_state(e, 0, out.data(), out.data() + 3 * n)) die("sbk_get_state");
        sbk_cloud_status st;
        if (sbk_get_status(e, 0, &st)) die("sbk_get_status");
        out[6 * n] = st.lifetime;
        out[6 * n + 1] = (double)st.attempts;
        sbk_destroy(e);
    } else
        die("unknown mode");
    FILE *o = std::fopen(argv[3], "wb");
    if (!o || std::fwrite(out.data(), 8, out.size(), o) != out.size()) die("cannot write output");
    std::fclose(o);
    return 0;
}
