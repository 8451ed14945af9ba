// sbk_main.cpp -- simbuca_hip, the command-line program: runs a reference
// simulation file on the GPU, as the reference's `simbuca <file.ini>` runs it
// on the CPU (bin/main.cpp:13-31).
//
//   simbuca_hip [--precision fp64_exact|fp64|fp64_rsq|mixed|mixed_kahan]
//               [--time frozen|stage] [--ame FILE] [--pdg FILE] [--quiet]
//               [--describe] FILE.ini
//
// --describe prints the settings the file resolves to (JSON) and runs nothing.
// The mass tables default to the reference's own paths relative to the
// working directory, ../libraries/ame/mass.mas12 and
// ../libraries/pdg/pdg_table.txt, and are skipped when absent (species names
// then need a leading mass number).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/simbuca_hip.h"

namespace {

bool exists(const std::string &p)
{
    struct stat sb;
    return stat(p.c_str(), &sb) == 0 && S_ISREG(sb.st_mode);
}

int usage()
{
    std::fprintf(stderr,
                 "usage: simbuca_hip [--precision fp64_exact|fp64|fp64_rsq|mixed|mixed_kahan] [--time frozen|stage]\n"
                 "                   [--ame FILE] [--pdg FILE] [--quiet] [--describe] FILE.ini\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv)
{
    int precision = SBK_FP64, time_mode = SBK_TIME_FROZEN, verbose = 1;
    std::string ame = "../libraries/ame/mass.mas12", pdg = "../libraries/pdg/pdg_table.txt", ini;
    bool ame_given = false, pdg_given = false, describe = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](std::string &out) {
            if (i + 1 >= argc) return false;
            out = argv[++i];
            return true;
        };
        std::string v;
        if (a == "--precision") {
            if (!value(v)) return usage();
            if (v == "fp64_exact") precision = SBK_FP64_EXACT;
            else if (v == "fp64") precision = SBK_FP64;
            else if (v == "mixed") precision = SBK_MIXED;
            else if (v == "mixed_kahan") precision = SBK_MIXED_KAHAN;
            else if (v == "fp64_rsq") precision = SBK_FP64_RSQ;
            else return usage();
        } else if (a == "--time") {
            if (!value(v)) return usage();
            if (v == "frozen") time_mode = SBK_TIME_FROZEN;
            else if (v == "stage") time_mode = SBK_TIME_STAGE;
            else return usage();
        } else if (a == "--ame") {
            if (!value(ame)) return usage();
            ame_given = true;
        } else if (a == "--pdg") {
            if (!value(pdg)) return usage();
            pdg_given = true;
        } else if (a == "--quiet") {
            verbose = 0;
        } else if (a == "--describe") {
            describe = true;
        } else if (!a.empty() && a[0] == '-') {
            return usage();
        } else if (ini.empty()) {
            ini = a;
        } else {
            return usage();
        }
    }
    if (ini.empty()) return usage();
    const char *ame_path = (ame_given || exists(ame)) ? ame.c_str() : nullptr;
    const char *pdg_path = (pdg_given || exists(pdg)) ? pdg.c_str() : nullptr;
    sbk_mass_table *t = nullptr;
    if (sbk_mass_table_open(ame_path, pdg_path, &t) != 0) {
        std::fprintf(stderr, "simbuca_hip: %s\n", sbk_last_error());
        return 1;
    }
    if (describe) {
        std::vector<char> json(1 << 22);
        const int rc = sbk_ini_describe(ini.c_str(), t, json.data(), (int)json.size());
        sbk_mass_table_close(t);
        if (rc != 0) {
            std::fprintf(stderr, "simbuca_hip: %s\n", sbk_last_error());
            return 1;
        }
        std::printf("%s\n", json.data());
        return 0;
    }
    const int rc = sbk_run_ini(ini.c_str(), t, precision, time_mode, verbose);
    sbk_mass_table_close(t);
    if (rc != 0) {
        std::fprintf(stderr, "simbuca_hip: %s\n", sbk_last_error());
        return 1;
    }
    if (verbose) std::fprintf(stderr, "Program ended\n");
    return 0;
}
