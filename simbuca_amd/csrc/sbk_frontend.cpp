// sbk_frontend.cpp -- the reference's simulation set-up on top of the engine:
// its INI simulation file, its cloud initializer and mass tables, and its
// operation sequence (export set 4 of include/simbuca_hip.h).
//
//   sbk_init_cloud     InitCloud (bin/initialization.cpp:117-206): the draws of
//                      the reference's Mersenne Twister (bin/mtrand.cpp,
//                      MT19937 seeded by init_genrand, u = x / 2^32) in the
//                      reference's order, with its arithmetic;
//   sbk_mass_table_*   the AME table reader (bin/IonTable.cpp:14-91, fixed
//                      columns) and the PDG particle table (bin/PDGTable.cpp);
//   sbk_parse_species  the [composition] name rules;
//   sbk_ini_describe,  DoSimulation + operations::launch + move_particles
//   sbk_run_ini        (bin/initialization.cpp:208-650, bin/operation.cpp:33-127,
//                      bin/ion_fly.cpp:17-122): the same sections and keys,
//                      the same composition rules, the same print schedule and
//                      log lines, with the engine doing the steps and
//                      sbk_writer the files.
// Host code; only sbk_run_ini touches the GPU (through the engine).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/simbuca_hip.h"
#include "sbk_internal.h"

namespace {

// bin/globals.h:4,6,10
const double kPi = 3.14159265358979323;
const double kEvToJoule = 1.60217646e-19;
const double kAmu = 1.6605402E-27;

int ffail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return sbk::record_error(code, buf);
}

// ---------------------------------------------------------------------------
// INI files as the reference reads them (bin/ini.c, bin/inireader.cpp): lines
// "name = value" or "name: value" under "[section]"; ';' and '#' start a
// comment line, ';' after whitespace an inline comment; an indented line
// continues the previous value (joined with '\n'), as do repeated names.
// Lookups lower-case section and name; section and field sets keep the case.
// ---------------------------------------------------------------------------
std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

std::string rstrip(const std::string &s)
{
    size_t e = s.size();
    while (e > 0 && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(0, e);
}

size_t lskip(const std::string &s, size_t i = 0)
{
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    return i;
}

// position of c, or of a ';' that follows whitespace, or npos
size_t find_or_comment(const std::string &s, size_t from, char c)
{
    bool was_space = false;
    for (size_t i = from; i < s.size(); ++i) {
        if (s[i] == c || (was_space && s[i] == ';')) return i;
        was_space = std::isspace((unsigned char)s[i]) != 0;
    }
    return std::string::npos;
}

struct Ini {
    std::map<std::string, std::string> values;
    std::set<std::string> sections;
    std::map<std::string, std::set<std::string>> fields;
    int error_line = 0;   // first line that is neither a section nor a pair

    // inih reads 200-byte chunks (INI_MAX_LINE, bin/ini.h:70): a longer line
    // is parsed as several lines, as there; section names keep 49 characters
    bool parse(const char *path)
    {
        FILE *f = std::fopen(path, "r");
        if (!f) return false;
        char buf[200];
        std::string section, prev_name;
        int lineno = 0;
        while (std::fgets(buf, sizeof(buf), f)) {
            ++lineno;
            std::string line(buf);
            const bool bom = lineno == 1 && line.size() >= 3 && (unsigned char)line[0] == 0xEF &&
                             (unsigned char)line[1] == 0xBB && (unsigned char)line[2] == 0xBF;
            const std::string r = rstrip(line);
            const size_t st = lskip(r, bom ? 3 : 0);
            const char first = st < r.size() ? r[st] : '\0';
            if (first == ';' || first == '#') continue;
            if (!prev_name.empty() && first && st > 0) {   // continuation line
                add(section, prev_name, r.substr(st));
                continue;
            }
            if (first == '[') {
                const size_t end = find_or_comment(r, st + 1, ']');
                if (end != std::string::npos && r[end] == ']') {
                    section = r.substr(st + 1, std::min<size_t>(end - st - 1, 49));
                    prev_name.clear();
                } else if (!error_line) {
                    error_line = lineno;
                }
                continue;
            }
            if (!first) continue;
            size_t end = find_or_comment(r, st, '=');
            if (end == std::string::npos || r[end] != '=') end = find_or_comment(r, st, ':');
            if (end == std::string::npos || (r[end] != '=' && r[end] != ':')) {
                if (!error_line) error_line = lineno;
                continue;
            }
            const std::string name = rstrip(r.substr(st, end - st));
            std::string value = r.substr(lskip(r, end + 1));
            const size_t c = find_or_comment(value, 0, '\0');
            if (c != std::string::npos) value = value.substr(0, c);
            value = rstrip(value);
            add(section, name, value);
            prev_name = name.substr(0, 49);
        }
        std::fclose(f);
        return true;
    }

    void add(const std::string &section, const std::string &name, const std::string &value)
    {
        std::string &v = values[lower(section + "=" + name)];
        if (!v.empty()) v += "\n";
        v += value;
        sections.insert(section);
        fields[lower(section)].insert(name);
    }

    bool has(const std::string &section) const { return sections.count(section) != 0; }

    std::string get(const std::string &s, const std::string &n, const std::string &dflt) const
    {
        auto it = values.find(lower(s + "=" + n));
        return it == values.end() ? dflt : it->second;
    }

    double real(const std::string &s, const std::string &n, double dflt) const
    {
        const std::string v = get(s, n, "");
        char *end = nullptr;
        const double x = std::strtod(v.c_str(), &end);
        return end > v.c_str() ? x : dflt;
    }

    bool boolean(const std::string &s, const std::string &n, bool dflt) const
    {
        const std::string v = lower(get(s, n, ""));
        if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
        if (v == "false" || v == "no" || v == "off" || v == "0") return false;
        return dflt;
    }
};

// ---------------------------------------------------------------------------
// Mass tables: AME nuclides, then PDG particles
// ---------------------------------------------------------------------------
struct Species {
    std::string name;   // AME: A followed by the element symbol ("133Cs"); PDG: its name
    double mass_amu;
};

}  // namespace

struct sbk_mass_table {
    std::vector<Species> ame, pdg;
};

namespace {

// bin/globals.h:11
const double kMassMeV = 931.494;

int field_int(const std::string &line, size_t pos, size_t len)
{
    return std::atoi(line.substr(pos, len).c_str());
}

// IonTable::IonTable (bin/IonTable.cpp:14-91): rows are the lines longer than
// 115 characters other than 124; '#' (estimated values) reads as a blank;
// fixed columns A [16,19), element [20,22) less its first blank, integer
// mass [96,99) and micro-u [100,113).
int read_ame(const char *path, std::vector<Species> &rows)
{
    std::ifstream in(path);
    if (!in) return ffail(SBK_EIO, "cannot open mass table %s", path);
    std::string line;
    while (std::getline(in, line)) {
        if (!(line.size() > 115 && line.size() != 124)) continue;
        std::replace(line.begin(), line.end(), '#', ' ');
        const int A = field_int(line, 16, 3);
        const int whole = field_int(line, 96, 3);
        const double micro = std::atof(line.substr(100, 13).c_str());
        std::string el = line.substr(20, 2);
        const size_t blank = el.find(' ');
        if (blank != std::string::npos) el.erase(blank, 1);
        rows.push_back(Species{std::to_string(A) + el, whole + micro * 1e-6});
    }
    if (rows.empty()) return ffail(SBK_EIO, "no nuclide rows in %s", path);
    return 0;
}

// PDGTable::ReadPDGTable (bin/PDGTable.cpp:79-183): '#' lines are comments;
// a particle line is "i name KF AP class_no class Q mass(GeV) width ... Ndec"
// followed by Ndec decay lines; an antiparticle line "i name -KF ..." takes
// the mass of the first particle with code KF.  The composition converts
// GeV to u by 1000 / 931.494 (bin/initialization.cpp:362).
int read_pdg(const char *path, std::vector<Species> &rows)
{
    std::ifstream in(path);
    if (!in) return ffail(SBK_EIO, "cannot open PDG table %s", path);
    std::vector<std::pair<int, double>> by_code;   // (KF, GeV) in file order
    std::string line;
    int pending_decays = 0;
    long lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line[0] == '#') continue;
        std::istringstream ss(line);
        if (pending_decays > 0) {
            std::string tok;
            if (ss >> tok) --pending_decays;
            continue;
        }
        int index = 0, kf = 0;
        std::string name;
        if (!(ss >> index >> name >> kf)) continue;
        double gev = -99999;   // GetPDGMass of an unknown code
        if (kf < 0) {
            for (const auto &c : by_code)
                if (c.first == -kf) {
                    gev = c.second;
                    break;
                }
        } else {
            int anti = 0, class_no = 0, charge = 0, ints[5] = {0, 0, 0, 0, 0}, ndec = 0;
            std::string class_name;
            double width = 0;
            if (!(ss >> anti >> class_no >> class_name >> charge >> gev >> width >> ints[0] >> ints[1] >>
                  ints[2] >> ints[3] >> ints[4] >> ndec))
                return ffail(SBK_EIO, "%s:%ld: not a PDG particle line", path, lineno);
            pending_decays = ndec;
        }
        by_code.emplace_back(kf, gev);
        rows.push_back(Species{name, gev * 1000.0 / kMassMeV});
    }
    if (rows.empty()) return ffail(SBK_EIO, "no particle rows in %s", path);
    return 0;
}

const Species *find(const std::vector<Species> &rows, const std::string &name)
{
    for (const Species &r : rows)   // the first row of that name, as the reference
        if (r.name == name) return &r;
    return nullptr;
}

// A species name of a composition (bin/initialization.cpp:333-391): an
// optional charge suffix "+", "-", "<k>+" or "<k>-" (charge 1 otherwise),
// then the mass from the AME table, the PDG table, or else from the name's
// leading digits.
int species(const std::string &original, const sbk_mass_table *table, double *mass_amu, int *charge)
{
    std::string name = original;
    if (name.empty()) return ffail(SBK_EINVAL, "empty species name");
    const char last = name.back();
    name.pop_back();
    int q = 1;
    if (last == '+' || last == '-') {
        const size_t k = name.find_last_not_of("0123456789");
        const std::string digits = (k == std::string::npos) ? name : name.substr(k + 1);
        q = digits.empty() ? 1 : std::atoi(digits.c_str());
        name = (k == std::string::npos) ? std::string() : name.substr(0, k + 1);
        if (last == '-') q = -q;
    } else {
        name.push_back(last);
    }
    const Species *s = table ? find(table->ame, name) : nullptr;
    if (!s && table) s = find(table->pdg, name);
    double m;
    if (s) {
        m = s->mass_amu;
    } else {
        const size_t k = name.find_first_not_of("0123456789");
        m = std::atof(name.substr(0, k).c_str());
    }
    if (!(m > 0.0))
        return ffail(SBK_EINVAL, "species %s: no mass (in no table, no leading mass number)",
                     original.c_str());
    *mass_amu = m;
    *charge = q;
    return 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// InitCloud
// ---------------------------------------------------------------------------
int sbk_init_cloud(int n, unsigned long seed, const double *semiaxis, const double *offset,
                   int n_energy, const double *energy_eV, const double *mass_amu, double *pos,
                   double *vel)
{
    if (n < 0 || (n > 0 && (!semiaxis || !offset || !energy_eV || !mass_amu || !pos || !vel)) ||
        n_energy < 1 || n_energy > 3)
        return ffail(SBK_EINVAL, "bad arguments");
    // MTRand (bin/mtrand.h:100-104): 32-bit MT19937 output divided by 2^32
    std::mt19937 mt((std::uint_fast32_t)(seed & 0xFFFFFFFFUL));
    auto u = [&mt]() { return static_cast<double>(mt()) * (1. / 4294967296.); };
    for (int i = 0; i < n; ++i) {
        double x, y, z;
        for (;;) {   // a point of the unit ball, by rejection
            x = u() * 2. - 1.;
            y = u() * 2. - 1.;
            z = u() * 2. - 1.;
            if (x * x + y * y + z * z <= 1) break;
        }
        pos[3 * (size_t)i] = offset[0] + semiaxis[0] * x;
        pos[3 * (size_t)i + 1] = offset[1] + semiaxis[1] * y;
        pos[3 * (size_t)i + 2] = offset[2] + semiaxis[2] * z;
        // Ion::Getmass (long double, bin/ion.cpp:3-6) read into a double
        const double mass = (double)((long double)mass_amu[i] * kAmu);
        const double theta = u() * 2.0 * kPi;
        const double phi = std::acos(u() * 2.0 - 1.0);
        // The reference program, built with g++ -O2, takes the sines and
        // cosines of its two-energy branch from sincos() and those of the
        // other branch from sin() and cos(); the two can differ by an ulp.
        // The same calls here keep the cloud bit-identical.
        double sin_t, cos_t, sin_p, cos_p;
        if (n_energy == 2) {
            sincos(theta, &sin_t, &cos_t);
            sincos(phi, &sin_p, &cos_p);
        } else {
            sin_t = std::sin(theta);
            cos_t = std::cos(theta);
            sin_p = std::sin(phi);
            cos_p = std::cos(phi);
        }
        double vmb = std::sqrt(-2.0 * (energy_eV[0] * kEvToJoule) / (mass) * std::log(1 - std::sqrt(u())));
        double *v = vel + 3 * (size_t)i;
        v[0] = vmb * cos_t * sin_p;
        v[1] = vmb * sin_t * sin_p;
        if (n_energy == 2)   // a separate axial energy
            vmb = std::sqrt(-2.0 * (energy_eV[1] * kEvToJoule) / (mass) * std::log(1 - std::sqrt(u())));
        v[2] = vmb * cos_p;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Mass tables
// ---------------------------------------------------------------------------
int sbk_mass_table_open(const char *ame_path, const char *pdg_path, sbk_mass_table **out)
{
    if (!out) return ffail(SBK_EINVAL, "bad arguments");
    std::unique_ptr<sbk_mass_table> t(new sbk_mass_table);
    if (ame_path)
        if (int rc = read_ame(ame_path, t->ame)) return rc;
    if (pdg_path)
        if (int rc = read_pdg(pdg_path, t->pdg)) return rc;
    *out = t.release();
    return 0;
}

int sbk_mass_table_lookup(const sbk_mass_table *t, const char *name, double *mass_amu)
{
    if (!t || !name || !mass_amu) return ffail(SBK_EINVAL, "bad arguments");
    const Species *s = find(t->ame, name);
    if (!s) s = find(t->pdg, name);
    if (!s) return ffail(SBK_EINVAL, "%s is in neither mass table", name);
    *mass_amu = s->mass_amu;
    return 0;
}

void sbk_mass_table_close(sbk_mass_table *t)
{
    delete t;
}

int sbk_parse_species(const char *name, const sbk_mass_table *table, double *mass_amu, int *charge)
{
    if (!name || !mass_amu || !charge) return ffail(SBK_EINVAL, "bad arguments");
    return species(name, table, mass_amu, charge);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Simulation files: the plan (host) and the run (engine)
// ---------------------------------------------------------------------------
namespace {

struct Operation {
    std::string name;   // normal, tof, trap_ramp, free, dissociation
    double time = 1e-3, print_time = 0.0, timestep = 0.0;
    // dissociation (bin/initialization.cpp:529-536, bin/operation.cpp:58-70)
    double fraction = 1.0, product_mass = 0.0;
    std::string reactant = "none", product = "none";
};

// What DoSimulation reads from a simulation file (bin/initialization.cpp:208-546)
struct Plan {
    int N = 0;
    double temp = 1e-3, semi[3] = {0.001, 0.001, 0.001}, off[3] = {0, 0, 0};
    std::vector<std::string> names;   // per ion
    std::vector<double> mass_amu;
    std::vector<int> charge;
    double timestep = 1e-9;
    bool adaptive = false;
    sbk_fields f;
    std::string import_prefix;        // [import]: the cloud is a print read back
    std::vector<double> pos, vel;     // the imported state
    double lifetime0 = 0.0;           // lifetime at the start of the run
    std::string prefix;
    int digits = 6;
    double print_interval = 1e-6;
    bool separate = false, print_after_operation = false, print_at_x = false;
    double print_x = 0.0;
    std::vector<Operation> ops;
    int skipped_line = 0;   // first line that is neither a section nor a pair
};

// [cloud] and [composition] (bin/initialization.cpp:308-396): species in
// sorted order, n = fraction * N, the last species takes the rest
int read_cloud(Ini &ini, const sbk_mass_table *table, Plan &P)
{
    if (!ini.has("cloud")) return ffail(SBK_EINVAL, "no [cloud] section");
    const double nreal = ini.real("cloud", "NParticles", 1024);
    if (!(nreal >= 0.0 && nreal < 1073741824.0))
        return ffail(SBK_EINVAL, "[cloud] NParticles must be in [0, 2^30)");
    P.N = (int)nreal;
    P.temp = ini.real("cloud", "temp", 1e-3);
    P.semi[0] = ini.real("cloud", "sx", 0.001);
    P.semi[1] = ini.real("cloud", "sy", 0.001);
    P.semi[2] = ini.real("cloud", "sz", 0.001);
    P.off[0] = ini.real("cloud", "x0", 0.0);
    P.off[1] = ini.real("cloud", "y0", 0.0);
    P.off[2] = ini.real("cloud", "z0", 0.0);
    if (!ini.has("composition")) return ffail(SBK_EINVAL, "no [composition] section");
    const std::set<std::string> &comps = ini.fields["composition"];
    for (auto it = comps.begin(); it != comps.end(); ++it) {
        const double fraction = ini.real("composition", *it, 1.0);
        if (fraction > 1) return ffail(SBK_EINVAL, "wrong composition of the cloud in fraction");
        double m = 0.0;
        int q = 1;
        if (int rc = species(*it, table, &m, &q)) return rc;
        if (!(fraction == fraction)) return ffail(SBK_EINVAL, "[composition] %s: not a number", it->c_str());
        int n = (int)(fraction * P.N);
        const int left = P.N - (int)P.names.size();
        if (std::next(it) == comps.end()) n = left;
        // the reference compares n with an unsigned count, so a negative n
        // (a negative fraction) takes every ion left
        if (n < 0 || n > left) n = left;
        for (int i = 0; i < n; ++i) {
            P.names.push_back(*it);
            P.mass_amu.push_back(m);
            P.charge.push_back(q);
        }
    }
    if ((int)P.names.size() != P.N)
        return ffail(SBK_EINVAL, "composition gives %zu of %d ions", P.names.size(), P.N);
    return 0;
}

int make_plan(const char *ini_path, const sbk_mass_table *table, Plan &P)
{
    if (!ini_path) return ffail(SBK_EINVAL, "null path");
    Ini ini;
    if (!ini.parse(ini_path)) return ffail(SBK_EIO, "cannot read %s", ini_path);
    // the reference reads on past such lines (it does not check the parse error)
    P.skipped_line = ini.error_line;

    if (ini.has("import")) {
        // [import] (:465-484): the cloud is the last print of <prefix>_pAll.txt.
        // The reference's ImportData (bin/parser.cpp:9-261) reads 14 columns
        // its writer never writes and looks names up without their charge
        // suffix; here the writer's 11 columns are read back and each name
        // goes through the composition rules.  [cloud] and [composition]
        // are then not used.
        P.import_prefix = ini.get("import", "prefix", "out");
        int n = 0;
        if (int rc = sbk_import(P.import_prefix.c_str(), 0, &n, nullptr, 64, nullptr, nullptr, nullptr))
            return rc;
        const int name_len = 256;
        std::vector<char> buf((size_t)n * name_len);
        P.pos.resize(3 * (size_t)n);
        P.vel.resize(3 * (size_t)n);
        if (int rc = sbk_import(P.import_prefix.c_str(), n, &n, buf.data(), name_len, P.pos.data(),
                                P.vel.data(), &P.lifetime0))
            return rc;
        P.N = n;
        for (int i = 0; i < n; ++i) {
            const std::string name(&buf[(size_t)i * name_len]);
            double m = 0.0;
            int q = 1;
            if (int rc = species(name, table, &m, &q)) return rc;
            P.names.push_back(name);
            P.mass_amu.push_back(m);
            P.charge.push_back(q);
        }
    } else if (int rc = read_cloud(ini, table, P)) {
        return rc;
    }


    // [ode], [coulomb], [trap], [output] (:398-507)
    if (!ini.has("ode")) return ffail(SBK_EINVAL, "no ODE configuration provided.");
    const std::string method = ini.get("ode", "method", "default");
    if (method != "DP" && method != "Dormand-Prince" && method != "5")
        return ffail(SBK_EINVAL, "[ode] method %s: only Dormand-Prince (DP) is provided", method.c_str());
    P.timestep = ini.real("ode", "timestep", 1e-9);
    P.adaptive = ini.boolean("ode", "adaptive", false);
    sbk_default_fields(&P.f);
    P.f.coulomb = 0;
    P.f.coulomb_scale = 0.0;
    if (ini.has("coulomb")) {
        P.f.coulomb = ini.boolean("coulomb", "enable", false) ? 1 : 0;
        P.f.coulomb_scale = ini.real("coulomb", "scalefactor", 1);
    }
    // another [trap] type leaves the trap's defaults, as in the reference
    // (only "ideal" reads the section; trap_type stays 0)
    if (ini.has("trap") && ini.get("trap", "type", "ideal") == "ideal") {
        P.f.Vrf = ini.real("trap", "Vrf", 50);
        P.f.Vdc = ini.real("trap", "Vdc", 5);
        P.f.kappa = ini.real("trap", "kappa", 0.025);
        P.f.r0 = ini.real("trap", "r0", 0.04);
        P.f.z0 = ini.real("trap", "z0", 0.04);
        P.f.freq_rf = ini.real("trap", "freq_rf", 0.04);
    }
    if (!ini.has("output")) return ffail(SBK_EINVAL, "no Output file configuration provided");
    P.prefix = ini.get("output", "prefix", "out");
    P.print_interval = ini.real("output", "timestep", 1e-6);
    P.separate = ini.boolean("output", "separate", false);
    P.print_after_operation = ini.boolean("output", "print_after_operation", false);
    // the writer's precision, "%.<digits>e" (not a reference key; 17 makes exact restart files)
    P.digits = (int)ini.real("output", "digits", 6);
    if (P.digits < 1 || P.digits > 17) return ffail(SBK_EINVAL, "[output] digits must be 1..17");
    P.print_at_x = ini.boolean("output", "print_at_x", false);
    P.print_x = ini.real("output", "print_x", 0.0);

    // [operation0..31] (:509-546).  Ekick and the ramp's Vrf/Vdc are one
    // setting of the trap each: the last operation that names one sets it
    // for every operation.
    for (int i = 0; i < 32; ++i) {
        const std::string s = "operation" + std::to_string(i);
        if (!ini.has(s)) continue;
        Operation op;
        const std::string type = ini.get(s, "type", "none");
        op.time = ini.real(s, "duration", 1e-3);
        op.print_time = ini.real(s, "print_timestep", 0.0);
        op.timestep = ini.real(s, "ode_timestep", 0.0);
        if (type == "normal" || type == "none" || type == "n" || type == "trap") {
            op.name = "normal";
        } else if (type == "tof" || type == "TOF") {
            op.name = "tof";
            P.f.E_kick = ini.real(s, "Ekick", 0.0);
        } else if (type == "trap ramp" || type == "trap_ramp") {
            op.name = "trap_ramp";
            P.f.newVrf = ini.real(s, "Vrf", 0.0);
            P.f.newVdc = ini.real(s, "Vdc", 0.0);
        } else if (type == "dissociation") {
            op.name = "dissociation";
            op.fraction = ini.real(s, "fraction", 1.0);
            op.reactant = ini.get(s, "reactant", "none");
            op.product = ini.get(s, "product", "none");
            op.product_mass = ini.real(s, "mass", 0.0);   // kg: it replaces Ion::mass
            if (op.product.empty() || op.product.find_first_of(" \t\n") != std::string::npos)
                return ffail(SBK_EINVAL, "[%s] product must be one word", s.c_str());
        } else if (type == "nop") {
            op.name = "free";
        } else {
            continue;   // the reference skips other types
        }
        if (op.timestep && !(op.print_time > 0))
            return ffail(SBK_EINVAL, "[%s] ode_timestep sets the step to print_timestep, which is not > 0",
                         s.c_str());
        P.ops.push_back(op);
    }
    if (P.ops.empty()) return ffail(SBK_EINVAL, "No operation loaded");
    return 0;
}

void json_str(std::string &o, const std::string &s)
{
    o += '"';
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof(b), "\\u%04x", c);
            o += b;
        } else {
            o += c;
        }
    }
    o += '"';
}

void json_num(std::string &o, double v)
{
    char b[32];
    std::snprintf(b, sizeof(b), "%.17g", v);
    o += b;
}

struct EngineGuard {
    sbk_engine *e = nullptr;
    ~EngineGuard() { sbk_destroy(e); }
};

struct WriterGuard {
    sbk_writer *w = nullptr;
    ~WriterGuard()
    {
        if (w) sbk_writer_close(w, 0.0, nullptr, nullptr);
    }
};

// <prefix>_logfile.txt: the lines the reference's loggers write for a run
// (LogFile, bin/logfile.cpp, numbers as an ostream prints a double: %g)
struct RunLog {
    FILE *f = nullptr;
    time_t start = 0;
    ~RunLog()
    {
        if (f) std::fclose(f);
    }
    int open(const std::string &prefix)
    {
        const std::string fn = prefix + "_logfile.txt";
        f = std::fopen(fn.c_str(), "w");
        if (!f) return ffail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
        start = time(nullptr);
        struct tm tmv;
        char buf[64];
        localtime_r(&start, &tmv);
        std::fprintf(f, "logfile created on: \t%s", asctime_r(&tmv, buf) ? buf : "?\n");
        return 0;
    }
    void line(const char *fmt, ...)
    {
        if (!f) return;
        va_list ap;
        va_start(ap, fmt);
        std::vfprintf(f, fmt, ap);
        va_end(ap);
        std::fflush(f);
    }
};

int make_folder(const std::string &prefix)
{
    // mkpath of the prefix's folder (bin/initialization.cpp:562-570)
    const size_t slash = prefix.find_last_of("/\\");
    if (slash == std::string::npos || slash == 0) return 0;
    const std::string folder = prefix.substr(0, slash);
    for (size_t k = 1; k <= folder.size(); ++k)
        if (k == folder.size() || folder[k] == '/') {
            const std::string d = folder.substr(0, k);
            if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST)
                return ffail(SBK_EIO, "cannot create %s: %s", d.c_str(), std::strerror(errno));
        }
    return 0;
}

}  // namespace

namespace {

int copy_out(const std::string &v, char *out, int len)
{
    if (!out || len <= (int)v.size()) return ffail(SBK_EINVAL, "the value needs %zu bytes", v.size() + 1);
    std::memcpy(out, v.c_str(), v.size() + 1);
    return 0;
}

}  // namespace

extern "C" int sbk_ini_value(const char *ini_path, const char *section, const char *name, char *out,
                             int len)
{
    if (!ini_path || !section || !name) return ffail(SBK_EINVAL, "bad arguments");
    Ini ini;
    if (!ini.parse(ini_path)) return ffail(SBK_EIO, "cannot read %s", ini_path);
    auto it = ini.values.find(lower(std::string(section) + "=" + name));
    if (it == ini.values.end()) return ffail(SBK_EINVAL, "[%s] %s is not set", section, name);
    return copy_out(it->second, out, len);
}

extern "C" int sbk_ini_names(const char *ini_path, const char *section, char *out, int len)
{
    if (!ini_path) return ffail(SBK_EINVAL, "bad arguments");
    Ini ini;
    if (!ini.parse(ini_path)) return ffail(SBK_EIO, "cannot read %s", ini_path);
    std::string v;
    if (!section) {
        for (const std::string &x : ini.sections) v += x + "\n";
    } else {
        auto it = ini.fields.find(lower(section));
        if (it != ini.fields.end())
            for (const std::string &x : it->second) v += x + "\n";
    }
    return copy_out(v, out, len);
}

extern "C" int sbk_ini_describe(const char *ini_path, const sbk_mass_table *table, char *json, int len)
{
    Plan P;
    if (int rc = make_plan(ini_path, table, P)) return rc;
    std::string o = "{\"N\": " + std::to_string(P.N) + ", \"temp\": ";
    json_num(o, P.temp);
    o += ", \"semiaxis\": [";
    for (int k = 0; k < 3; ++k) {
        if (k) o += ", ";
        json_num(o, P.semi[k]);
    }
    o += "], \"offset\": [";
    for (int k = 0; k < 3; ++k) {
        if (k) o += ", ";
        json_num(o, P.off[k]);
    }
    o += "], \"species\": [";
    for (int i = 0; i < P.N; ++i) {
        if (i && P.names[i] == P.names[i - 1]) continue;
        int count = 0;
        for (int j = i; j < P.N && P.names[j] == P.names[i]; ++j) ++count;
        if (i) o += ", ";
        o += "{\"name\": ";
        json_str(o, P.names[i]);
        o += ", \"count\": " + std::to_string(count) + ", \"mass_amu\": ";
        json_num(o, P.mass_amu[i]);
        o += ", \"charge\": " + std::to_string(P.charge[i]) + "}";
    }
    o += "], \"timestep\": ";
    json_num(o, P.timestep);
    o += std::string(", \"adaptive\": ") + (P.adaptive ? "true" : "false");
    o += std::string(", \"coulomb\": ") + (P.f.coulomb ? "true" : "false") + ", \"coulomb_scale\": ";
    json_num(o, P.f.coulomb_scale);
    const std::pair<const char *, double> trap[] = {
        {"Vrf", P.f.Vrf},     {"Vdc", P.f.Vdc},       {"kappa", P.f.kappa},   {"r0", P.f.r0},
        {"z0", P.f.z0},       {"freq_rf", P.f.freq_rf}, {"E_kick", P.f.E_kick}, {"newVrf", P.f.newVrf},
        {"newVdc", P.f.newVdc}};
    o += ", \"trap\": {";
    for (size_t k = 0; k < sizeof(trap) / sizeof(trap[0]); ++k) {
        if (k) o += ", ";
        o += std::string("\"") + trap[k].first + "\": ";
        json_num(o, trap[k].second);
    }
    o += "}, \"import\": ";
    if (P.import_prefix.empty())
        o += "null";
    else
        json_str(o, P.import_prefix);
    o += ", \"lifetime\": ";
    json_num(o, P.lifetime0);
    o += ", \"prefix\": ";
    json_str(o, P.prefix);
    o += ", \"digits\": " + std::to_string(P.digits);
    o += ", \"print_interval\": ";
    json_num(o, P.print_interval);
    o += std::string(", \"separate\": ") + (P.separate ? "true" : "false");
    o += std::string(", \"print_after_operation\": ") + (P.print_after_operation ? "true" : "false");
    o += std::string(", \"print_at_x\": ") + (P.print_at_x ? "true" : "false") + ", \"print_x\": ";
    json_num(o, P.print_x);
    o += ", \"operations\": [";
    for (size_t k = 0; k < P.ops.size(); ++k) {
        if (k) o += ", ";
        o += "{\"type\": ";
        json_str(o, P.ops[k].name);
        o += ", \"duration\": ";
        json_num(o, P.ops[k].time);
        o += ", \"print_timestep\": ";
        json_num(o, P.ops[k].print_time);
        o += ", \"ode_timestep\": ";
        json_num(o, P.ops[k].timestep);
        if (P.ops[k].name == "dissociation") {
            o += ", \"fraction\": ";
            json_num(o, P.ops[k].fraction);
            o += ", \"reactant\": ";
            json_str(o, P.ops[k].reactant);
            o += ", \"product\": ";
            json_str(o, P.ops[k].product);
            o += ", \"mass\": ";
            json_num(o, P.ops[k].product_mass);
        }
        o += "}";
    }
    o += "], \"skipped_line\": " + std::to_string(P.skipped_line) + "}";
    if (!json || len <= (int)o.size())
        return ffail(SBK_EINVAL, "the description needs %zu bytes", o.size() + 1);
    std::memcpy(json, o.c_str(), o.size() + 1);
    return 0;
}

extern "C" int sbk_run_ini(const char *ini_path, const sbk_mass_table *table, int precision,
                           int time_mode, int verbose)
{
    Plan P;
    if (int rc = make_plan(ini_path, table, P)) return rc;
    if (P.skipped_line && verbose)
        std::fprintf(stderr, "%s:%d: not a section or name = value line (skipped)\n", ini_path, P.skipped_line);
    if (int rc = make_folder(P.prefix)) return rc;
    const int N = P.N;
    sbk_fields f = P.f;
    // the set-up lines of InitIonFly, Init_ode, SetCoulomb, UseScaledCoulomb
    // (bin/ion_fly.cpp:125-132,149-168, bin/ode.cpp:571-585)
    RunLog log;
    if (int rc = log.open(P.prefix)) return rc;
    log.line("Ideal Trap: Vrf = %g, Vdc = %g\n", f.Vrf, f.Vdc);
    log.line("Choose Dormand Prince 5th order. With %s stepsize and abs error: %g, rel error: %g\n",
             P.adaptive ? "adaptive" : "fixed", 1e-7, 1e-7);
    log.line("%s Coulomb Interaction\n", f.coulomb ? "with" : "without");
    log.line("with scaled coulomb Interaction: factor %g.\n", f.coulomb_scale);
    if (!P.import_prefix.empty()) {   // ImportData's lines (bin/parser.cpp:228-238)
        log.line("Data Imported from %s_pAll.txt\n%d particles : \n", P.import_prefix.c_str(), N);
    } else {   // InitCloud's lines (bin/initialization.cpp:202-205)
        log.line("%d particles created. Max-Boltz distribution around %g eV. \n", N, P.temp);
        log.line("spheroid cloud with semi-axis's = {%g;%g;%g} and offset = {%g;%g;%g}  \n", P.semi[0],
                 P.semi[1], P.semi[2], P.off[0], P.off[1], P.off[2]);
    }

    // the cloud: imported, or drawn with seed 0 and three equal energies
    // (:571-593), so InitCloud takes one speed per ion
    std::vector<double> pos(3 * (size_t)N), vel(3 * (size_t)N);
    if (!P.import_prefix.empty()) {
        pos = P.pos;
        vel = P.vel;
    } else {
        const double energy[3] = {P.temp, P.temp, P.temp};
        if (int rc = sbk_init_cloud(N, 0UL, P.semi, P.off, 3, energy, P.mass_amu.data(), pos.data(),
                                    vel.data()))
            return rc;
    }
    std::vector<double> mass_kg(N);
    for (int i = 0; i < N; ++i) mass_kg[i] = (double)((long double)P.mass_amu[i] * kAmu);
    std::vector<std::string> names = P.names;

    std::vector<const char *> cnames(N);
    for (int i = 0; i < N; ++i) cnames[i] = P.names[i].c_str();
    WriterGuard wg;
    if (int rc = sbk_writer_open(P.prefix.c_str(), N, cnames.data(), P.mass_amu.data(), P.separate ? 1 : 0,
                                 P.digits, &wg.w))
        return rc;

    // The engine starts with the first operation that moves the ions: the
    // reference copies masses into its force vectors there, once
    // (bin/ion_fly.cpp:26-62), so a dissociation before it changes the
    // dynamics and one after it only the printed names and energies.
    EngineGuard eg;
    sbk_config cfg;
    sbk_default_config(&cfg);
    cfg.precision = precision;
    cfg.time_mode = time_mode;
    cfg.adaptive = P.adaptive ? 1 : 0;
    cfg.h0 = P.timestep;
    double pending_h = 0.0;   // an ode_timestep setting made before the engine exists
    int xc_cap = 0;   // entries of the engine's crossing list
    auto start_engine = [&]() -> int {
        if (eg.e) return 0;
        if (int rc = sbk_create(&cfg, &eg.e)) return rc;
        if (int rc = sbk_add_cloud(eg.e, N, pos.data(), vel.data(), mass_kg.data(), P.charge.data(), &f,
                                   P.lifetime0, nullptr))
            return rc;
        if (pending_h > 0.0)
            if (int rc = sbk_set_step(eg.e, 0, pending_h)) return rc;
        if (P.print_at_x)
            if (int rc = sbk_set_crossing_plane(eg.e, 0, 1, P.print_x, &xc_cap)) return rc;
        return 0;
    };

    auto print = [&](double t) -> int {
        if (int rc = sbk_writer_particles(wg.w, t, pos.data(), vel.data())) return rc;
        return sbk_writer_cloud(wg.w, t, pos.data(), vel.data());
    };

    // operation::launch and move_particles (bin/operation.cpp:33-83,
    // bin/ion_fly.cpp:17-108): after every step, print once when
    // lifetime - t0 > print_interval * nr_interval, nr_interval counting
    // prints over the whole run and t0 the lifetime the run started at (0
    // unless imported; the reference measures only its first operation from
    // there and the others from 0).  The engine stops at the first step past
    // the next print time (or at the end), which is the step the reference
    // prints at.  With print_at_x the engine records, in every accepted step,
    // the ions whose x crossed print_x (strictly, from either side), and they
    // are written in step and ion order after each advance; an advance takes
    // at most as many attempts as the crossing list has room for.
    double print_interval = P.print_interval;
    int nr_interval = 0;
    sbk_cloud_status st;
    std::memset(&st, 0, sizeof(st));
    st.lifetime = P.lifetime0;
    std::vector<double> xpos, xvel;   // one crossing's state at its ion's slot
    std::vector<int> xc_ion;
    std::vector<double> xc_t, xc_pos, xc_vel;
    auto write_crossings = [&]() -> int {
        int n = 0;
        if (int rc = sbk_take_crossings(eg.e, 0, &n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr))
            return rc;
        if (n == 0) return 0;
        xc_ion.resize(n);
        xc_t.resize(n);
        xc_pos.resize(3 * (size_t)n);
        xc_vel.resize(3 * (size_t)n);
        if (int rc = sbk_take_crossings(eg.e, n, &n, nullptr, xc_ion.data(), nullptr, xc_t.data(),
                                        xc_pos.data(), xc_vel.data()))
            return rc;
        xpos.assign(3 * (size_t)N, 0.0);
        xvel.assign(3 * (size_t)N, 0.0);
        for (int i = 0; i < n; ++i) {
            const int k = xc_ion[i];
            for (int c = 0; c < 3; ++c) {
                xpos[3 * (size_t)k + c] = xc_pos[3 * (size_t)i + c];
                xvel[3 * (size_t)k + c] = xc_vel[3 * (size_t)i + c];
            }
            if (int rc = sbk_writer_x(wg.w, k, xc_t[i], xpos.data(), xvel.data())) return rc;
        }
        return 0;
    };
    for (const Operation &op : P.ops) {
        if (op.print_time) print_interval = op.print_time;
        if (op.timestep) {   // the reference sets h to the print step here (bin/operation.cpp:37-40)
            if (eg.e) {
                if (int rc = sbk_set_step(eg.e, 0, op.print_time)) return rc;
            } else {
                pending_h = op.print_time;
            }
        }
        if (op.name == "dissociation") {
            // MTRand seeded with 0 at each dissociation, one draw per reactant ion
            std::mt19937 mt(0u);
            for (int i = 0; i < N; ++i) {
                if (names[i] != op.reactant) continue;
                if (static_cast<double>(mt()) * (1. / 4294967296.) < op.fraction) {
                    names[i] = op.product;
                    if (!eg.e) mass_kg[i] = op.product_mass;
                    if (int rc = sbk_writer_set_ion(wg.w, i, op.product.c_str(), op.product_mass)) return rc;
                }
            }
            if (verbose) std::fprintf(stderr, "operation dissociation of %s into %s\n", op.reactant.c_str(),
                                      op.product.c_str());
            continue;
        }
        if (int rc = start_engine()) return rc;
        sbk_engine *e = eg.e;
        sbk_fields g = f;
        g.trap = (op.name == "normal" || op.name == "trap_ramp") ? 1 : 0;
        g.tof = (op.name == "tof") ? 1 : 0;
        g.ramp = (op.name == "trap_ramp") ? 1 : 0;
        const double t_start = st.lifetime;
        const double endtime = op.time + t_start;
        g.ramp_t0 = t_start;
        g.ramp_t1 = t_start + op.time;
        if (int rc = sbk_set_fields(e, 0, &g)) return rc;
        log.line("Normal trap operation for %g sec,\n", op.time);   // normal_operation, bin/ion_fly.cpp:139-143
        if (verbose)
            std::fprintf(stderr, "operation %s for %g s from t = %g s\n", op.name.c_str(), op.time, t_start);
        // move_particles leaves at once when the cloud is empty (bin/ion_fly.cpp:87-88)
        while (N > 0 && st.lifetime < endtime) {
            const double t_print = print_interval * nr_interval;
            double target = st.lifetime;
            if (!(target - P.lifetime0 > t_print)) {
                // the first lifetime with lifetime - t0 > t_print: the least
                // such double, as x - t0 rounds monotonically in x
                target = P.lifetime0 + t_print;
                while (target - P.lifetime0 > t_print) target = std::nextafter(target, -INFINITY);
                while (!(target - P.lifetime0 > t_print)) target = std::nextafter(target, INFINITY);
            }
            if (target <= st.lifetime) target = std::nextafter(st.lifetime, INFINITY);   // one step
            if (target > endtime) target = endtime;
            const long long max_attempts =
                P.print_at_x ? std::max(1, xc_cap / std::max(N, 1)) : (1LL << 62);
            const double before = st.lifetime;
            if (int rc = sbk_advance_until(e, target, max_attempts)) return rc;
            if (int rc = sbk_get_status(e, 0, &st)) return rc;
            // an advance that stopped with the cloud idle, no error and no time gained
            // would repeat forever
            if (st.lifetime == before && !st.active && !st.error)
                return ffail(SBK_ESTATE, "the engine made no progress at lifetime %.17g s", before);
            if (P.print_at_x)
                if (int rc = write_crossings()) return rc;
            const bool due = st.lifetime - P.lifetime0 > t_print;
            if (due) {
                if (int rc = sbk_get_state(e, 0, pos.data(), vel.data())) return rc;
                if (int rc = print(st.lifetime)) return rc;
                nr_interval++;
            }
        }
        if (g.ramp && st.steps > 0 && st.t_prev >= t_start) {
            // the trap keeps the values of the last step's start (bin/ion_fly.cpp:95-99)
            const double ramp = (st.t_prev - g.ramp_t0) / (g.ramp_t1 - g.ramp_t0);
            f.Vrf = g.Vrf * (1 - ramp) + g.newVrf * ramp;
            f.Vdc = g.Vdc * (1 - ramp) + g.newVdc * ramp;
        }
        if (P.print_after_operation) {
            if (int rc = sbk_get_state(e, 0, pos.data(), vel.data())) return rc;
            if (int rc = print(st.lifetime)) return rc;
        }
    }
    if (eg.e)
        if (int rc = sbk_get_state(eg.e, 0, pos.data(), vel.data())) return rc;
    sbk_writer *w = wg.w;
    wg.w = nullptr;
    log.line("Close everything...\n");   // ExitIonFly, then Delete (bin/ioncloud.cpp:69)
    log.line("#simulation ended after %g s\n", (double)(time(nullptr) - log.start));
    return sbk_writer_close(w, st.lifetime, pos.data(), vel.data());
}
