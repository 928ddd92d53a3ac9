// Reference-compatible extern "C" entry points (include/mbx/external_call.h) over bblock::System.
// One global system per process, as in the reference (external_call.cpp:42-45).  Exceptions
// are reported on stderr and turn into a process exit, since a C / Fortran caller cannot
// catch them (the reference lets them escape, which terminates the process the same way).
#include "../../include/mbx/external_call.h"

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/mbx/system.h"

namespace {
bblock::System* g_sys = nullptr;

template <class F>
void guarded(const char* name, F f) {
    try {
        f();
    } catch (const std::exception& e) {
        std::cerr << name << ": " << e.what() << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

bblock::System& sys(const char* name) {
    if (!g_sys) {
        std::cerr << name << ": the system is not initialized" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    return *g_sys;
}

template <class Names>
void init(double* coords, int* nat_monomers, Names at_names, Names monomers, int nmon, const char* json) {
    delete g_sys;
    g_sys = new bblock::System();
    int count = 0;
    for (int i = 0; i < nmon; ++i) {
        std::vector<double> xyz(coords + 3 * count, coords + 3 * (count + nat_monomers[i]));
        std::vector<std::string> names;
        for (int a = 0; a < nat_monomers[i]; ++a) names.push_back(std::string(at_names[count + a]));
        g_sys->AddMonomer(xyz, names, std::string(monomers[i]));
        count += nat_monomers[i];
    }
    g_sys->Initialize();
    g_sys->SetUpFromJson(const_cast<char*>(json));
}
}  // namespace

extern "C" {

void initialize_system_(double* coords, int* nat_monomers, char at_names[][5], char monomers[][5], int* nmon,
                        char json_file[20]) {
    guarded(__func__, [&] { init(coords, nat_monomers, at_names, monomers, *nmon, json_file); });
}

void initialize_system_py_(double* coords, int* nat_monomers, char** at_names, char** monomers, int* nmon,
                           char* json_file) {
    guarded(__func__, [&] { init(coords, nat_monomers, at_names, monomers, *nmon, json_file); });
}

void get_energy_(double* coords, int* nat, double* energy) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        *energy = s.Energy(false);
    });
}

void get_energy_g_(double* coords, int* nat, double* energy, double* grads) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_g_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        *energy = s.Energy(true);
        std::vector<double> g = s.GetRealGrads();
        std::copy(g.begin(), g.end(), grads);
    });
}

void get_energy_pbc_(double* coords, int* nat, double* box, double* energy) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_pbc_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        s.SetPBC(std::vector<double>(box, box + 9));
        *energy = s.Energy(false);
    });
}

void get_energy_pbc_g_(double* coords, int* nat, double* box, double* energy, double* grads) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_pbc_g_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        s.SetPBC(std::vector<double>(box, box + 9));
        *energy = s.Energy(true);
        std::vector<double> g = s.GetRealGrads();
        std::copy(g.begin(), g.end(), grads);
    });
}

static void decomp(bblock::System& s, double* E1b, double* E2b, double* E3b, double* E4b, double* Edisp,
                   double* Ebuck, double* Eelec) {
    *E1b = 0.0;
    *E2b = s.TwoBodyEnergy(false);
    *E3b = s.ThreeBodyEnergy(false);
    *E4b = 0.0;
    *Edisp = 0.0;
    *Ebuck = 0.0;
    *Eelec = 0.0;
}

void get_energy_decomp_(double* coords, int* nat, double* E1b, double* E2b, double* E3b, double* E4b, double* Edisp,
                        double* Ebuck, double* Eelec) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_decomp_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        decomp(s, E1b, E2b, E3b, E4b, Edisp, Ebuck, Eelec);
    });
}

void get_energy_decomp_pbc_(double* coords, int* nat, double* box, double* E1b, double* E2b, double* E3b,
                            double* E4b, double* Edisp, double* Ebuck, double* Eelec) {
    guarded(__func__, [&] {
        bblock::System& s = sys("get_energy_decomp_pbc_");
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        s.SetPBC(std::vector<double>(box, box + 9));
        decomp(s, E1b, E2b, E3b, E4b, Edisp, Ebuck, Eelec);
    });
}

void get_virial_(double* virial) {
    guarded(__func__, [&] {
        std::vector<double> v = sys("get_virial_").GetVirial();
        if (v.size() == 9)
            std::copy(v.begin(), v.end(), virial);
        else
            std::fill(virial, virial + 9, 0.0);
    });
}

void set_real_xyz_(double* coords, int* nat) {
    guarded(__func__, [&] { sys("set_real_xyz_").SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat))); });
}

void set_box_(int* length, double* box) {
    guarded(__func__, [&] {
        sys("set_box_").SetPBC(*length > 0 ? std::vector<double>(box, box + *length) : std::vector<double>());
    });
}

void finalize_system_(void) {
    delete g_sys;
    g_sys = nullptr;
}

}  // extern "C"
