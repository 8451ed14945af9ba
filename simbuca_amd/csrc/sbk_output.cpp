// sbk_output.cpp -- the reference's output files, and an importer for them
// (export set "output" of include/simbuca_hip.h).  Host code only.
//
// Files and columns follow IonCloud (bin/ioncloud.cpp):
//   <prefix>_pAll.txt, or <prefix>_p<k>.txt per ion (CreateFile, :196-243):
//       one line per ion and print, "ID name t(ms) x y z(mm) vx vy vz(m/s) E(eV)"
//       (PrintParticle, :85-113), E = m |v|^2 / 2 in eV with m the ion's
//       long-double mass (bin/ion.cpp:3-6);
//   <prefix>_cloud<k>.txt per species, k = rank of the name in sorted order
//       (:244-257): "k name N t(ms) x0 y0 z0 sx sy sz(mm) E(eV)" (PrintCloud,
//       :148-183), where the reference divides the species' sums by the
//       total ion count N and adds vz*vy instead of vz^2 to the energy.  Both
//       are kept, so that the numbers equal the reference's;
//   the closing block of Delete() (:38-58): a last print of every ion, then
//       "#***", "#simulation ended after <s> s", "# #Particles left= <N>",
//       "#End time of simulation: <date>".
// Numbers are written as std::scientific with the default precision (6),
// i.e. "%.6e".  The writer can also print with more digits, for restart
// files that reproduce the state to within one ulp.
//
// The reference's importer (ImportData, bin/parser.cpp:9-261) reads 14
// columns that its writer never writes.  sbk_import reads the writer's own
// 11 columns instead: the last complete print of <prefix>_pAll.txt.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/simbuca_hip.h"
#include "sbk_internal.h"

namespace {

// bin/globals.h:7,9
const double kAmu = 1.6605402E-27;
const double kJouleToEv = 6.24150e18;

int out_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return sbk::record_error(code, buf);
}

struct FileCloser {
    void operator()(FILE *f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

std::string now_text()
{
    const time_t t = time(nullptr);
    char buf[64];
    const char *s = ctime_r(&t, buf);
    return s ? std::string(s) : std::string("?\n");
}

void write_header(FILE *f, const char *columns)
{
    std::fprintf(f, "#Simulation started: %s", now_text().c_str());
    std::fprintf(f, "#------------------------------------------------------------------------------#\n");
    std::fprintf(f, "#                 Simbuca ion-cloud engine (MI355X, libsimbuca_hip)            #\n");
    std::fprintf(f, "#                 Dormand-Prince 5(4) with adaptive step control               #\n");
    std::fprintf(f, "#------------------------------------------------------------------------------#\n");
    std::fprintf(f, "%s\n", columns);
    std::fprintf(f, "#*******************************************************************************\n");
}

}  // namespace

struct sbk_writer {
    std::string prefix;
    int n = 0;
    int digits = 6;
    int per_particle = 0;
    std::vector<std::string> names;
    std::vector<long double> mass;       // kg, as Ion keeps it
    std::vector<std::string> species;    // sorted distinct names
    File all;
    std::vector<File> parts;
    std::vector<File> clouds;
    File xfile;                          // <prefix>_x.txt, opened by the first crossing
    time_t start = 0;
};

namespace {

// "%.<digits>e" as std::scientific with that precision
void put(FILE *f, int digits, double v)
{
    std::fprintf(f, " %.*e", digits, v);
}

void particle_line(const sbk_writer *w, FILE *f, int k, double lifetime, const double *pos,
                   const double *vel)
{
    const double *p = pos + 3 * (size_t)k, *v = vel + 3 * (size_t)k;
    // IonCloud::PrintParticle: long-double mass times the double |v|^2
    const double energy =
        (double)(w->mass[k] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * 0.5 * kJouleToEv);
    std::fprintf(f, "%d %s", k, w->names[k].c_str());
    put(f, w->digits, lifetime * 1000.0);
    put(f, w->digits, p[0] * 1000.0);
    put(f, w->digits, p[1] * 1000.0);
    put(f, w->digits, p[2] * 1000.0);
    put(f, w->digits, v[0]);
    put(f, w->digits, v[1]);
    put(f, w->digits, v[2]);
    put(f, w->digits, energy);
    std::fputc('\n', f);
}

int check_writer(const sbk_writer *w, const double *pos, const double *vel)
{
    if (!w) return out_fail(SBK_EINVAL, "null writer");
    if (w->n > 0 && (!pos || !vel)) return out_fail(SBK_EINVAL, "null state arrays");
    return 0;
}

}  // namespace

extern "C" {

int sbk_writer_open(const char *prefix, int n, const char *const *names, const double *mass_amu,
                    int per_particle_files, int digits, sbk_writer **out)
{
    if (!prefix || !out || n < 0 || (n > 0 && (!names || !mass_amu)))
        return out_fail(SBK_EINVAL, "bad arguments");
    if (digits < 1 || digits > 17) return out_fail(SBK_EINVAL, "digits must be 1..17");
    std::unique_ptr<sbk_writer> w(new sbk_writer);
    w->prefix = prefix;
    w->n = n;
    w->digits = digits;
    w->per_particle = per_particle_files ? 1 : 0;
    w->start = time(nullptr);
    std::set<std::string> types;
    for (int i = 0; i < n; ++i) {
        if (!names[i] || !names[i][0] || std::strpbrk(names[i], " \t\n"))
            return out_fail(SBK_EINVAL, "ion %d: a name must be one non-empty word", i);
        w->names.emplace_back(names[i]);
        // bin/ion.cpp:3-6: mass = mass_amu * amu in long double
        w->mass.push_back((long double)mass_amu[i] * kAmu);
        types.insert(w->names.back());
    }
    w->species.assign(types.begin(), types.end());
    const char *cols = "#index \t mass \t t (ms) \t x y z (mm) vx vy vz (m/s) \t Energy(eV)";
    if (w->per_particle) {
        for (int i = 0; i < n; ++i) {
            const std::string fn = w->prefix + "_p" + std::to_string(i + 1) + ".txt";
            File f(std::fopen(fn.c_str(), "w"));
            if (!f) return out_fail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
            write_header(f.get(), cols);
            w->parts.push_back(std::move(f));
        }
    } else if (n > 0) {
        const std::string fn = w->prefix + "_pAll.txt";
        w->all.reset(std::fopen(fn.c_str(), "w"));
        if (!w->all) return out_fail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
        write_header(w->all.get(), cols);
    }
    for (size_t k = 0; k < w->species.size(); ++k) {
        const std::string fn = w->prefix + "_cloud" + std::to_string(k) + ".txt";
        File f(std::fopen(fn.c_str(), "w"));
        if (!f) return out_fail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
        write_header(f.get(),
                     "#index \t type \t #particules \t t (ms) \t x0 y0 z0 (mm) sx sy sz (m/s) \t Energy(eV)");
        w->clouds.push_back(std::move(f));
    }
    *out = w.release();
    return 0;
}

int sbk_writer_particles(sbk_writer *w, double lifetime, const double *pos, const double *vel)
{
    if (int rc = check_writer(w, pos, vel)) return rc;
    for (int k = 0; k < w->n; ++k)
        particle_line(w, w->per_particle ? w->parts[k].get() : w->all.get(), k, lifetime, pos, vel);
    return 0;
}

int sbk_writer_cloud(sbk_writer *w, double lifetime, const double *pos, const double *vel)
{
    if (int rc = check_writer(w, pos, vel)) return rc;
    const double N = (double)w->n;
    for (size_t id = 0; id < w->species.size(); ++id) {
        const std::string &type = w->species[id];
        double x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0, energy = 0.0;
        for (int i = 0; i < w->n; ++i) {
            if (w->names[i] != type) continue;
            const double *p = pos + 3 * (size_t)i, *v = vel + 3 * (size_t)i;
            x0 += p[0];
            y0 += p[1];
            z0 += p[2];
            // the reference's sum, vz*vy included (bin/ioncloud.cpp:160)
            energy = (double)(energy + 0.5 * kJouleToEv * w->mass[i] *
                                           (v[0] * v[0] + v[1] * v[1] + v[2] * v[1]));
        }
        x0 /= N;   // over all ions, as the reference divides
        y0 /= N;
        z0 /= N;
        for (int i = 0; i < w->n; ++i) {
            if (w->names[i] != type) continue;
            const double *p = pos + 3 * (size_t)i;
            x1 += (p[0] - x0) * (p[0] - x0);
            y1 += (p[1] - y0) * (p[1] - y0);
            z1 += (p[2] - z0) * (p[2] - z0);
        }
        x1 = std::sqrt(x1 / N);
        y1 = std::sqrt(y1 / N);
        z1 = std::sqrt(z1 / N);
        FILE *f = w->clouds[id].get();
        std::fprintf(f, "%zu %s %d", id, type.c_str(), w->n);
        put(f, w->digits, lifetime * 1e3);
        put(f, w->digits, 1000.0 * x0);
        put(f, w->digits, 1000.0 * y0);
        put(f, w->digits, 1000.0 * z0);
        put(f, w->digits, 1000.0 * x1);
        put(f, w->digits, 1000.0 * y1);
        put(f, w->digits, 1000.0 * z1);
        put(f, w->digits, energy);
        std::fputc('\n', f);
    }
    return 0;
}

int sbk_writer_set_ion(sbk_writer *w, int k, const char *name, double mass_kg)
{
    if (!w || k < 0 || k >= w->n || !name) return out_fail(SBK_EINVAL, "bad arguments");
    if (!name[0] || std::strpbrk(name, " \t\n"))
        return out_fail(SBK_EINVAL, "ion %d: a name must be one non-empty word", k);
    // the species files keep the species of the opening (IonCloud::ion_types
    // is filled by AddParticle only), so a renamed ion leaves its old sums
    w->names[k] = name;
    w->mass[k] = (long double)mass_kg;
    return 0;
}

int sbk_writer_x(sbk_writer *w, int k, double lifetime, const double *pos, const double *vel)
{
    if (int rc = check_writer(w, pos, vel)) return rc;
    if (k < 0 || k >= w->n) return out_fail(SBK_EINVAL, "bad ion index");
    if (!w->xfile) {
        // IonCloud::PrintX (bin/ioncloud.cpp:120-146): the call that opens
        // the file writes its header only
        const std::string fn = w->prefix + "_x.txt";
        w->xfile.reset(std::fopen(fn.c_str(), "w"));
        if (!w->xfile) return out_fail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
        std::fprintf(w->xfile.get(), "#Print at X position");
        write_header(w->xfile.get(), "#index \t type \t t (ms) \t x y z (mm) vx vy vz (m/s)");
        return 0;
    }
    const double *p = pos + 3 * (size_t)k, *v = vel + 3 * (size_t)k;
    FILE *f = w->xfile.get();
    std::fprintf(f, "%d %s", k, w->names[k].c_str());
    put(f, w->digits, lifetime * 1000.0);
    put(f, w->digits, p[0] * 1000.0);
    put(f, w->digits, p[1] * 1000.0);
    put(f, w->digits, p[2] * 1000.0);
    put(f, w->digits, v[0]);
    put(f, w->digits, v[1]);
    put(f, w->digits, v[2]);
    std::fputc('\n', f);
    return 0;
}

int sbk_writer_close(sbk_writer *w, double lifetime, const double *pos, const double *vel)
{
    if (!w) return 0;
    int rc = 0;
    if (w->n > 0 && pos && vel) rc = sbk_writer_particles(w, lifetime, pos, vel);
    // the reference keeps its start time as a double, so the elapsed
    // seconds print in the stream's scientific format
    const double secs = (double)(time(nullptr) - w->start);
    const std::string end = now_text();
    auto trailer = [&](FILE *f) {
        std::fprintf(f, "#*******************************************************************************\n");
        std::fprintf(f, "#simulation ended after %.6e s\n", secs);
        std::fprintf(f, "# #Particles left= %d\n", w->n);
        std::fprintf(f, "#End time of simulation: %s", end.c_str());
    };
    if (w->per_particle) {
        for (auto &f : w->parts) trailer(f.get());
    } else if (w->all) {
        trailer(w->all.get());
    }
    delete w;   // closes every file
    return rc;
}

int sbk_import(const char *prefix, int capacity, int *n_out, char *names, int name_len,
               double *pos, double *vel, double *lifetime)
{
    if (!prefix || !n_out || capacity < 0 || name_len < 2)
        return out_fail(SBK_EINVAL, "bad arguments");
    const std::string fn = std::string(prefix) + "_pAll.txt";
    File f(std::fopen(fn.c_str(), "r"));
    if (!f) return out_fail(SBK_EIO, "cannot open %s: %s", fn.c_str(), std::strerror(errno));
    struct Row {
        int id;
        std::string name;
        double t, x, y, z, vx, vy, vz;
    };
    std::vector<Row> rows;
    char line[4096];
    long lineno = 0;
    while (std::fgets(line, sizeof(line), f.get())) {
        ++lineno;
        const char *s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '#' || *s == '\n' || *s == '\0') continue;
        Row r;
        char nm[256];
        double e;
        if (std::sscanf(s, "%d %255s %lf %lf %lf %lf %lf %lf %lf %lf", &r.id, nm, &r.t, &r.x, &r.y,
                        &r.z, &r.vx, &r.vy, &r.vz, &e) != 10)
            return out_fail(SBK_EIO, "%s:%ld: expected 11 columns (ID name t x y z vx vy vz E)",
                            fn.c_str(), lineno);
        r.name = nm;
        rows.push_back(r);
    }
    // the last print: walk back from the end until an ID repeats
    std::set<int> seen;
    size_t first = rows.size();
    while (first > 0 && !seen.count(rows[first - 1].id)) {
        seen.insert(rows[first - 1].id);
        --first;
    }
    const int n = (int)(rows.size() - first);
    *n_out = n;
    if (n == 0) return out_fail(SBK_EIO, "%s holds no ion lines", fn.c_str());
    if (capacity == 0 && !names && !pos && !vel && !lifetime) return 0;   // a size query
    if (n > capacity) return out_fail(SBK_ENOMEM, "%s holds %d ions, capacity %d", fn.c_str(), n, capacity);
    std::vector<const Row *> last;
    for (size_t i = first; i < rows.size(); ++i) last.push_back(&rows[i]);
    std::sort(last.begin(), last.end(), [](const Row *a, const Row *b) { return a->id < b->id; });
    for (int i = 0; i < n; ++i) {
        const Row &r = *last[i];
        if (names) {
            std::snprintf(names + (size_t)i * name_len, name_len, "%s", r.name.c_str());
        }
        // mm -> m as the reference's importer does (bin/parser.cpp:104-106)
        if (pos) {
            pos[3 * (size_t)i] = r.x / 1000.;
            pos[3 * (size_t)i + 1] = r.y / 1000.;
            pos[3 * (size_t)i + 2] = r.z / 1000.;
        }
        if (vel) {
            vel[3 * (size_t)i] = r.vx;
            vel[3 * (size_t)i + 1] = r.vy;
            vel[3 * (size_t)i + 2] = r.vz;
        }
    }
    if (lifetime) *lifetime = last[0]->t * 1e-3;   // ms -> s (bin/parser.cpp:225)
    return 0;
}

}  // extern "C"
