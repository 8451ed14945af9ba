// ref_init.cpp -- drives the reference's own InitCloud and mass tables
// (bin/initialization.cpp:33-43,117-206; bin/IonTable.cpp; bin/PDGTable.cpp),
// compiled from the reference sources by `make -C oracle refsim`.  TEST
// INFRASTRUCTURE: it writes golden vectors; nothing in the product links it.
//
// The reference builds its tables at start-up from ../libraries/ame/mass.mas12
// and ../libraries/pdg/pdg_table.txt, so run it from a directory whose parent
// holds libraries/.
//
//   ref_init init <n> <seed> <n_energy> <e...> <sx sy sz> <x0 y0 z0> <mass_amu...>
//       -> n lines "x y z vx vy vz" (%.17g), the cloud InitCloud draws
//   ref_init mass <name...>
//       -> one line per name: "<name> ame <mass>", "<name> pdg <mass in u>"
//          or "<name> none"
//   ref_init ini <file> [<section> <name>]...
//       -> the reference's INIReader (bin/inireader.cpp) on the file:
//          "S <hex>" per section (GetSections), then per queried pair
//          "F <hex>" per field of the section (GetFields) and "V <hex>" for
//          Get(section, name) or "V -" when it is not set; strings hex-encoded
#include "ioncloud.h"
#include "ion.h"
#include "IonTable.h"
#include "PDGTable.h"
#include "trapparameters.h"
#include "inireader.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

// defined in bin/initialization.cpp (its header declares an older signature)
void InitCloud(int nparticles, std::vector<double> eV_max_boltz, int seed, double semiaxis[3],
               double offset[3], std::vector<Ion> Ions, IonCloud &_cloud, _trap_param &trap_param);
extern IonTable Table;
extern PDGTable pdgTable;
extern std::vector<double> x_, y_, z_, vx_, vy_, vz_;

static std::string to_hex(const std::string &v)
{
    std::string o;
    char b[3];
    for (size_t i = 0; i < v.size(); ++i) {
        snprintf(b, sizeof(b), "%02x", (unsigned char)v[i]);
        o += b;
    }
    return o.empty() ? std::string("=") : o;
}

static void usage()
{
    fprintf(stderr, "usage: ref_init init n seed n_energy e... sx sy sz x0 y0 z0 mass...\n"
                    "       ref_init mass name...\n"
                    "       ref_init ini file [section name]...\n");
    exit(2);
}

int main(int argc, char **argv)
{
    if (argc < 2) usage();
    if (!strcmp(argv[1], "mass")) {
        for (int i = 2; i < argc; ++i) {
            const std::string name = argv[i];
            if (Table.TestIonName(name))
                printf("%s ame %.17g\n", argv[i], Table.ResearchMass(name));
            else if (pdgTable.TestPDGName(name))
                printf("%s pdg %.17g\n", argv[i],
                       pdgTable.GetPDGMass(pdgTable.GetPDGId(name)) * 1000.0 / 931.494);
            else
                printf("%s none\n", argv[i]);
        }
        return 0;
    }
    if (!strcmp(argv[1], "ini") && argc >= 3 && argc % 2 == 1) {
        INIReader reader(argv[2]);
        std::set<std::string> secs = reader.GetSections();
        for (std::set<std::string>::iterator it = secs.begin(); it != secs.end(); ++it)
            printf("S %s\n", to_hex(*it).c_str());
        const std::string missing("\x01");
        for (int i = 3; i + 1 < argc; i += 2) {
            std::set<std::string> f = reader.GetFields(argv[i]);
            for (std::set<std::string>::iterator it = f.begin(); it != f.end(); ++it)
                printf("F %s\n", to_hex(*it).c_str());
            const std::string v = reader.Get(argv[i], argv[i + 1], missing);
            printf("V %s\n", v == missing ? "-" : to_hex(v).c_str());
        }
        return 0;
    }
    if (strcmp(argv[1], "init") || argc < 5) usage();
    int a = 2;
    const int n = atoi(argv[a++]);
    const int seed = atoi(argv[a++]);
    const int ne = atoi(argv[a++]);
    if (argc != 5 + ne + 6 + n) usage();
    std::vector<double> energy;
    for (int k = 0; k < ne; ++k) energy.push_back(atof(argv[a++]));
    double semi[3], off[3];
    for (int k = 0; k < 3; ++k) semi[k] = atof(argv[a++]);
    for (int k = 0; k < 3; ++k) off[k] = atof(argv[a++]);
    std::vector<Ion> ions;
    for (int i = 0; i < n; ++i) ions.push_back(Ion((long double)atof(argv[a++]), 1));
    IonCloud cloud;   // AddParticle opens its output file in the working directory
    cloud.use_particles_files(false);
    cloud.Create("ref_init");
    _trap_param trap;
    InitCloud(n, energy, seed, semi, off, ions, cloud, trap);
    for (int i = 0; i < n; ++i)
        printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", x_[i], y_[i], z_[i], vx_[i], vy_[i], vz_[i]);
    return 0;
}
