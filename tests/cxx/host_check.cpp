// host_check.cpp -- the host-only code of the project under AddressSanitizer
// and UndefinedBehaviorSanitizer (built with g++ by tests/cxx/Makefile, run
// by tests/test_host_sanitizers.py; no GPU):
//
//   * the library's host sources that need no device: the output writer and
//     importer (sbk_output.cpp) and the simulation set-up (sbk_frontend.cpp:
//     INI reader, InitCloud, mass tables, composition, sbk_ini_describe);
//   * the CPU oracle (oracle/simbuca_oracle.c): force, Coulomb rows, the
//     CUNBODY interface sum, DP5 steps with every field term on.
//
// The few engine entry points the front end links against are stubs here
// (the engine itself needs a GPU); sbk_run_ini is not called.
//
//   host_check <scratch dir> [<ini file>...]
// Prints "host_check ok" and exits 0 when every check passes; a sanitizer
// report ends the process with an error first.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "simbuca_hip.h"
#include "../../oracle/simbuca_oracle.h"
#include "../../simbuca_amd/csrc/sbk_plan.h"

// ---- stubs of the engine (sbk_engine.cpp needs the device) ----------------
namespace sbk {
static std::string last_error;
int record_error(int code, const char *msg)
{
    last_error = msg ? msg : "";
    return code;
}
}  // namespace sbk

extern "C" {
const char *sbk_last_error(void) { return sbk::last_error.c_str(); }
void sbk_default_config(sbk_config *c) { std::memset(c, 0, sizeof(*c)); }
void sbk_default_fields(sbk_fields *f)
{
    std::memset(f, 0, sizeof(*f));
    f->trap = 1;
    f->kappa = 0.001;
    f->Vdc = 5;
    f->Vrf = 50;
    f->r0 = 0.04;
    f->z0 = 0.04;
    f->freq_rf = 50e5;
}
static int no_device() { return sbk::record_error(SBK_ESTATE, "no device in host_check"); }
int sbk_create(const sbk_config *, sbk_engine **) { return no_device(); }
int sbk_destroy(sbk_engine *) { return 0; }
int sbk_add_cloud(sbk_engine *, int, const double *, const double *, const double *, const int *,
                  const sbk_fields *, double, int *) { return no_device(); }
int sbk_set_fields(sbk_engine *, int, const sbk_fields *) { return no_device(); }
int sbk_set_step(sbk_engine *, int, double) { return no_device(); }
int sbk_advance_until(sbk_engine *, double, long long) { return no_device(); }
int sbk_get_state(sbk_engine *, int, double *, double *) { return no_device(); }
int sbk_get_status(sbk_engine *, int, sbk_cloud_status *) { return no_device(); }
int sbk_set_crossing_plane(sbk_engine *, int, int, double, int *) { return no_device(); }
int sbk_take_crossings(sbk_engine *, int, int *, int *, int *, long long *, double *, double *,
                       double *) { return no_device(); }
}

static int failures = 0;
#define CHECK(cond)                                                                           \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            std::fprintf(stderr, "host_check: %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond,   \
                         sbk_last_error());                                                   \
            ++failures;                                                                       \
        }                                                                                     \
    } while (0)

static void write_file(const std::string &path, const std::string &text)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
}

static void check_init_cloud()
{
    const int n = 513;
    std::vector<double> amu(n), pos(3 * n), vel(3 * n);
    for (int i = 0; i < n; ++i) amu[i] = (i % 3) ? 133.0 : 0.000548579909;
    const double semi[3] = {1e-3, 2e-3, 5e-4}, off[3] = {0, 1e-4, 0};
    for (int ne = 1; ne <= 3; ++ne) {
        const double e[3] = {0.1, 0.2, 0.3};
        CHECK(sbk_init_cloud(n, 7UL, semi, off, ne, e, amu.data(), pos.data(), vel.data()) == 0);
        for (int i = 0; i < 3 * n; ++i) CHECK(std::isfinite(pos[i]) && std::isfinite(vel[i]));
    }
    CHECK(sbk_init_cloud(0, 0UL, nullptr, nullptr, 1, semi, nullptr, nullptr, nullptr) == 0);
    CHECK(sbk_init_cloud(-1, 0UL, semi, off, 1, semi, amu.data(), pos.data(), vel.data()) != 0);
    CHECK(sbk_init_cloud(1, 0UL, semi, off, 4, semi, amu.data(), pos.data(), vel.data()) != 0);
}

static void check_species(const std::string &dir)
{
    double m = 0;
    int q = 0;
    const char *good[] = {"133Xx+", "85Yy2+", "40Zz-", "7Ww3-", "12", "133Xx", "2-+"};
    for (const char *s : good) CHECK(sbk_parse_species(s, nullptr, &m, &q) == 0 && m > 0);
    const char *bad[] = {"", "+", "-", "Xx+", "1+", "--"};   // "1+": charge 1 of an empty name
    for (const char *s : bad) CHECK(sbk_parse_species(s, nullptr, &m, &q) != 0);
    // a synthetic AME-format table: short, long, '#', 124-character rows
    std::string rows;
    rows += std::string(200, 'x') + "\n";
    for (int k = 0; k < 4; ++k) {
        std::string line(k == 3 ? 124 : 117 + k, ' ');
        std::memcpy(&line[16], " 40", 3);
        std::memcpy(&line[20], k == 1 ? "Q " : "Qx", 2);
        std::memcpy(&line[96], " 39", 3);
        std::memcpy(&line[100], k == 2 ? "  962383.12#" : "  962383.1237", k == 2 ? 12 : 13);
        rows += line + "\n";
    }
    rows += "short\n";
    write_file(dir + "/t.mas", rows);
    sbk_mass_table *t = nullptr;
    CHECK(sbk_mass_table_open((dir + "/t.mas").c_str(), nullptr, &t) == 0);
    CHECK(sbk_mass_table_lookup(t, "40Qx", &m) == 0 && std::fabs(m - 39.9623831237) < 1e-9);
    CHECK(sbk_mass_table_lookup(t, "40Q", &m) == 0);
    CHECK(sbk_mass_table_lookup(t, "41Qx", &m) != 0);
    CHECK(sbk_parse_species("40Qx2+", t, &m, &q) == 0 && q == 2);
    sbk_mass_table_close(t);
    // a synthetic PDG-format table, with decay lines and an antiparticle
    write_file(dir + "/p.txt",
               "# comment\n"
               "    1 aa                 11  1 100 Lepton     -3 5.1099891e-04 0.0e+00 -100 -1 -100 -1   -1   2\n"
               "         1   102   0.5   2       21        1   g d\n"
               "\n"
               "         2   102   0.5   2       22        1   gamma d\n"
               "    2 aa_bar            -11      1     0\n"
               "    3 bb                 13  1 100 Lepton     -3 1.05658e-01 0.0e+00 -100 -1 -100 -1   -1   0\n");
    CHECK(sbk_mass_table_open(nullptr, (dir + "/p.txt").c_str(), &t) == 0);
    double ma = 0, mb = 0;
    CHECK(sbk_mass_table_lookup(t, "aa", &ma) == 0 && sbk_mass_table_lookup(t, "aa_bar", &mb) == 0 &&
          ma == mb);
    CHECK(sbk_mass_table_lookup(t, "bb", &m) == 0);
    CHECK(sbk_mass_table_lookup(t, "g", &m) != 0);
    sbk_mass_table_close(t);
    CHECK(sbk_mass_table_open((dir + "/missing").c_str(), nullptr, &t) != 0);
    CHECK(sbk_mass_table_open(nullptr, nullptr, &t) == 0);
    sbk_mass_table_close(t);
}

static void check_ini(const std::string &dir, int argc, char **argv)
{
    std::vector<char> json(1 << 20);
    for (int i = 2; i < argc; ++i) {   // the golden simulation files
        CHECK(sbk_ini_describe(argv[i], nullptr, json.data(), (int)json.size()) == 0);
        CHECK(std::strlen(json.data()) > 10);
        char tiny[8];
        CHECK(sbk_ini_describe(argv[i], nullptr, tiny, sizeof(tiny)) != 0);
    }
    const std::string base =
        "[cloud]\nNParticles = 10\n[composition]\n133Xx+ = 0.3\n85Yy+ = 1\n[ode]\nmethod = DP\n"
        "[output]\nprefix = p\nprint_at_x = on\n[operation0]\ntype = normal\n"
        "[operation1]\ntype = dissociation\nreactant = 85Yy+\nproduct = 9Q+\n";
    const std::string cases[] = {
        base,
        "\xEF\xBB\xBF" + base,
        base + "  continued value\n",
        base + std::string(450, 'k') + " = 1\n",
        base + "[" + std::string(300, 's') + "]\nx = 1\n",
        std::string(250, ' ') + "\n" + base,
        base + "no separator here\n[unclosed\n;\n#\n=\n:\n",
        "",
        "[cloud]\n",
        base + "[operation2]\ntype = normal\node_timestep = 1\n",
    };
    for (const std::string &c : cases) {
        write_file(dir + "/c.ini", c);
        (void)sbk_ini_describe((dir + "/c.ini").c_str(), nullptr, json.data(), (int)json.size());
        (void)sbk_ini_value((dir + "/c.ini").c_str(), "cloud", "NParticles", json.data(), (int)json.size());
        (void)sbk_ini_names((dir + "/c.ini").c_str(), nullptr, json.data(), (int)json.size());
        (void)sbk_ini_names((dir + "/c.ini").c_str(), "composition", json.data(), (int)json.size());
    }
    char tiny[4];
    write_file(dir + "/c.ini", base);
    CHECK(sbk_ini_value((dir + "/c.ini").c_str(), "cloud", "NParticles", tiny, sizeof(tiny)) == 0);
    CHECK(sbk_ini_value((dir + "/c.ini").c_str(), "output", "prefix", tiny, 1) != 0);
    CHECK(sbk_ini_value((dir + "/c.ini").c_str(), "cloud", "nothing", tiny, sizeof(tiny)) != 0);
    CHECK(sbk_ini_names((dir + "/c.ini").c_str(), nullptr, tiny, sizeof(tiny)) != 0);
    CHECK(sbk_ini_describe((dir + "/missing.ini").c_str(), nullptr, json.data(), (int)json.size()) != 0);
}

static void check_writer(const std::string &dir)
{
    const int n = 5;
    const char *names[n] = {"133Xx+", "85Yy+", "133Xx+", "40Zz2+", "85Yy+"};
    const double amu[n] = {133, 85, 133, 40, 85};
    std::vector<double> pos(3 * n), vel(3 * n);
    for (int i = 0; i < 3 * n; ++i) {
        pos[i] = 1e-4 * (i - 7);
        vel[i] = 10.0 * (i % 4);
    }
    for (int layout = 0; layout < 2; ++layout) {
        const std::string prefix = dir + (layout ? "/pp" : "/all");
        sbk_writer *w = nullptr;
        CHECK(sbk_writer_open(prefix.c_str(), n, names, amu, layout, 17, &w) == 0);
        CHECK(sbk_writer_particles(w, 1e-6, pos.data(), vel.data()) == 0);
        CHECK(sbk_writer_cloud(w, 1e-6, pos.data(), vel.data()) == 0);
        CHECK(sbk_writer_x(w, 1, 1e-6, pos.data(), vel.data()) == 0);
        CHECK(sbk_writer_x(w, 2, 1e-6, pos.data(), vel.data()) == 0);
        CHECK(sbk_writer_x(w, n, 1e-6, pos.data(), vel.data()) != 0);
        CHECK(sbk_writer_set_ion(w, 0, "9Q+", 1e-26) == 0);
        CHECK(sbk_writer_set_ion(w, 0, "two words", 1e-26) != 0);
        CHECK(sbk_writer_set_ion(w, -1, "x", 1e-26) != 0);
        CHECK(sbk_writer_close(w, 2e-6, pos.data(), vel.data()) == 0);
    }
    int m = 0;
    CHECK(sbk_import((dir + "/all").c_str(), 0, &m, nullptr, 64, nullptr, nullptr, nullptr) == 0 && m == n);
    std::vector<char> nm((size_t)n * 64);
    std::vector<double> p2(3 * n), v2(3 * n);
    double t = 0;
    CHECK(sbk_import((dir + "/all").c_str(), n, &m, nm.data(), 64, p2.data(), v2.data(), &t) == 0);
    CHECK(t == 2e-6 || std::fabs(t - 2e-6) < 1e-20);
    CHECK(std::strcmp(nm.data(), "9Q+") == 0);
    CHECK(sbk_import((dir + "/all").c_str(), n - 1, &m, nm.data(), 64, p2.data(), v2.data(), &t) != 0);
    write_file(dir + "/bad_pAll.txt", "# x\n0 a 1 2\n");
    CHECK(sbk_import((dir + "/bad").c_str(), n, &m, nm.data(), 64, p2.data(), v2.data(), &t) != 0);
    sbk_writer *w = nullptr;
    CHECK(sbk_writer_open((dir + "/none").c_str(), 0, nullptr, nullptr, 0, 6, &w) == 0);
    CHECK(sbk_writer_close(w, 0.0, nullptr, nullptr) == 0);
}

static void check_oracle()
{
    const int n = 97;
    std::vector<double> pos(3 * n), vel(3 * n), mass(n), der(3 * n);
    std::vector<int> q(n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = 1e-4 * std::sin(1.7 * i + 2.3 * k);
            vel[3 * i + k] = 50.0 * std::cos(0.9 * i + k);
        }
        mass[i] = orc_mass_kg(i % 2 ? 133.0 : 85.0);
        q[i] = (i % 5) ? 1 : 2;
    }
    std::vector<double> swift(16);
    for (int k = 0; k < 16; ++k) swift[k] = std::sin(0.3 * k);
    orc_fields f;
    std::memset(&f, 0, sizeof(f));
    f.trap = 1; f.kappa = 0.001; f.Vdc = 5; f.Vrf = 50; f.r0 = 0.04; f.z0 = 0.04; f.freq_rf = 1e6;
    f.tof = 1; f.E_kick = 3; f.coulomb = 1; f.coulomb_scale = 1;
    f.V0 = 10; f.d0 = 0.011; f.Bz = 7; f.Ed = 10; f.fd = 940; f.Uq = 0.1; f.rq = 0.011; f.fq = 8e5;
    f.delta = 100; f.Uo = 0.1; f.ro = 0.01; f.fo = 1e5; f.rw_order = 2; f.Uw = 0.1; f.rw = 0.01;
    f.fw = 1e5; f.Eax = 1; f.fax = 1e4; f.fd_rate = 1e9; f.B2 = 100; f.K0 = 1e-4; f.p_mbar = 1e-6;
    f.T_gas = 300; f.swift = swift.data(); f.swift_n = 16; f.swift_kind = 2; f.swift_t0 = 0;
    f.swift_dt = 1e-8; f.swift_r = 0.01;
    orc_force(n, pos.data(), vel.data(), mass.data(), q.data(), &f, 3e-8, der.data());
    for (double d : der) CHECK(std::isfinite(d));
    const int rows[3] = {0, 50, n - 1};
    std::vector<double> out(9);
    orc_coulomb_rows(n, pos.data(), mass.data(), q.data(), 1.0, rows, 3, out.data());
    std::vector<double> ai(3 * n), mj(n, 1.0);
    orc_cunbody_force(pos.data(), mj.data(), pos.data(), 1e-23, ai.data(), n, n);
    for (int mode = 0; mode < 2; ++mode) {
        orc_cloud *c = orc_create(n, pos.data(), vel.data(), mass.data(), q.data(), &f, 0.0, 1e-9, 1e-7,
                                  1e-7, 1, mode);
        CHECK(c != nullptr);
        for (int s = 0; s < 5; ++s) CHECK(orc_step(c) == 0);
        std::vector<double> h(64);
        CHECK(orc_advance(c, 2e-8, 1000, h.data(), (long)h.size()) >= 0);
        std::vector<double> p2(3 * n), v2(3 * n), sc(8);
        orc_get_state(c, p2.data(), v2.data(), sc.data());
        CHECK(sc[0] >= 2e-8);
        orc_destroy(c);
    }
}


// The Coulomb work list (sbk_plan.cpp) over cloud mixes and launch shapes:
// check_plan must hold for every list (each item's slice is below its
// cloud's slice count, which is at most max_slices, the number of partial
// planes the engine allocates; every pair covered once), and must reject a
// list with an item writing one plane past max_slices or a pair missing.
static void check_planner()
{
    std::mt19937 mt(7u);
    std::vector<std::vector<int>> mixes = {{1}, {0}, {500}, {10000}, {65536}, {1, 0, 2, 63, 64, 65},
                                           {4096, 64, 4095, 257}};
    for (int k = 0; k < 6; ++k) {   // ensembles with log-uniform sizes in [64, 4096]
        std::uniform_real_distribution<double> u(std::log(64.0), std::log(4096.0));
        std::vector<int> v(k == 0 ? 256 : 1 + (int)(mt() % 40));
        for (int &n : v) n = (int)std::lround(std::exp(u(mt)));
        mixes.push_back(v);
    }
    long long lists = 0;
    for (const auto &sizes : mixes) {
        std::vector<sbk::PlanCloud> pc(sizes.size());
        int off = 0;
        for (size_t i = 0; i < sizes.size(); ++i) {
            pc[i].n = sizes[i];
            pc[i].off = off;
            pc[i].on = (i % 5) != 3;   // some clouds without Coulomb
            off += sizes[i];
        }
        for (int q : {1, 4})
            for (int tpt : {1, 2, 4, 8})
                for (int sreq : {0, 1, 3, 7, 64, 100000})
                    for (long long cap : {256LL, 1024LL, 4096LL})
                        for (int exact : {0, 1})
                        for (int split : {0, 1}) {
                            const int blk = sbk::kBlock / q * tpt;
                            sbk::Plan p = sbk::plan_coulomb(pc, exact ? sbk::kBlock : blk,
                                                            sbk::kSubTile * q, sreq, cap, exact != 0,
                                                            split ? blk / 2 : 0);
                            const std::string bad = sbk::check_plan(pc, p);
                            if (!bad.empty()) std::fprintf(stderr, "plan: %s\n", bad.c_str());
                            CHECK(bad.empty());
                            if (exact) CHECK(p.max_slices == 1);
                            // the item-cost cap: a cloud's first item (a whole first block
                            // against its first, longest slice) is at most 1/kItemShare of a
                            // resident work-group's share, or its slice cannot be halved
                            if (!exact) {
                                long long pairs = 0;
                                for (const auto &c : pc)
                                    if (c.on && c.n > 0) pairs += (long long)c.n * c.n;
                                for (const auto &u : p.items) {
                                    if (u.i0 == u.i1 || u.slice != 0 || u.i0 != pc[u.cloud].off) continue;
                                    const long long tb = u.i1 - u.i0, chunk = u.j1 - u.j0;
                                    CHECK(chunk < 2 * sbk::kMinCapChunk ||
                                          sbk::kItemShare * tb * chunk * cap <= pairs);
                                }
                            }
                            // the XCD runs hold about equal work: none exceeds an eighth of
                            // the list's cost by more than its costliest item
                            {
                                auto cost = [](const sbk::CoulombItem &u) {
                                    return (long long)((u.i1 - u.i0 + 63) / 64) * (u.j1 - u.j0);
                                };
                                long long total = 0, top = 0;
                                for (const auto &u : p.items) {
                                    total += cost(u);
                                    top = std::max(top, cost(u));
                                }
                                for (int x = 0; x < 8 && !p.items.empty(); ++x) {
                                    long long run = 0;
                                    for (int k = p.xrun[x]; k < p.xrun[x + 1]; ++k) run += cost(p.items[k]);
                                    CHECK(8 * run <= total + 8 * top);
                                }
                            }
                            lists++;
                            if (!p.items.empty() && lists % 7 == 0) {
                                sbk::Plan q1 = p;
                                q1.items[q1.items.size() / 2].slice = q1.max_slices;
                                CHECK(!sbk::check_plan(pc, q1).empty());
                                sbk::Plan q2 = p;
                                q2.items.pop_back();
                                CHECK(!sbk::check_plan(pc, q2).empty());
                                sbk::Plan q3 = p;   // a grid that misses the longest run's last item
                                q3.grid -= 8;
                                CHECK(!sbk::check_plan(pc, q3).empty());
                            }
                        }
    }
    CHECK(lists > 1000);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_check <scratch dir> [<ini file>...]\n");
        return 2;
    }
    const std::string dir = argv[1];
    check_init_cloud();
    check_species(dir);
    check_ini(dir, argc, argv);
    check_writer(dir);
    check_oracle();
    check_planner();
    if (failures) {
        std::fprintf(stderr, "host_check: %d failure(s)\n", failures);
        return 1;
    }
    std::printf("host_check ok\n");
    return 0;
}
