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
std::vector<bblock::System*> g_rank_sys;  // initialize_system2_: one system per rank (external_call.cpp:43)

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

bblock::System& rank_sys(const char* name, int rank) {
    if (rank < 0 || (size_t)rank >= g_rank_sys.size()) {
        std::cerr << name << ": no system for rank " << rank << " (initialize_system2_ made "
                  << g_rank_sys.size() << ")" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    return *g_rank_sys[rank];
}

template <class Names>
bblock::System* build(double* coords, int* nat_monomers, Names at_names, Names monomers, int nmon, const char* json) {
    bblock::System* s = new bblock::System();
    int count = 0;
    for (int i = 0; i < nmon; ++i) {
        std::vector<double> xyz(coords + 3 * count, coords + 3 * (count + nat_monomers[i]));
        std::vector<std::string> names;
        for (int a = 0; a < nat_monomers[i]; ++a) names.push_back(std::string(at_names[count + a]));
        s->AddMonomer(xyz, names, std::string(monomers[i]));
        count += nat_monomers[i];
    }
    s->Initialize();
    s->SetUpFromJson(const_cast<char*>(json));
    return s;
}

template <class Names>
void init(double* coords, int* nat_monomers, Names at_names, Names monomers, int nmon, const char* json) {
    delete g_sys;
    g_sys = nullptr;
    g_sys = build(coords, nat_monomers, at_names, monomers, nmon, json);
}

void energy_g(bblock::System& s, double* coords, int nat, double* energy, double* grads) {
    s.SetRealXyz(std::vector<double>(coords, coords + 3 * nat));
    *energy = s.Energy(true);
    std::vector<double> g = s.GetRealGrads();
    std::copy(g.begin(), g.end(), grads);
}
}  // namespace

extern "C" {

void initialize_system_(double* coords, int* nat_monomers, char at_names[][5], char monomers[][5], int* nmon,
                        char json_file[20]) {
    guarded(__func__, [&] { init(coords, nat_monomers, at_names, monomers, *nmon, json_file); });
}

void initialize_system2_(double* coords, int* nat_monomers, char at_names[][5], char monomers[][5], int* nmon,
                         char json_file[20], int* nranks) {
    guarded(__func__, [&] {
        for (int k = 0; k < *nranks; ++k)
            g_rank_sys.push_back(build(coords, nat_monomers, at_names, monomers, *nmon, json_file));
    });
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

void get_energy2_(double* coords, int* nat, double* energy, int* rank) {
    guarded(__func__, [&] {
        bblock::System& s = rank_sys("get_energy2_", *rank);
        s.SetRealXyz(std::vector<double>(coords, coords + 3 * (*nat)));
        *energy = s.Energy(false);
    });
}

void get_energy_g_(double* coords, int* nat, double* energy, double* grads) {
    guarded(__func__, [&] { energy_g(sys("get_energy_g_"), coords, *nat, energy, grads); });
}

void get_energy_g2_(double* coords, int* nat, double* energy, double* grads, int* rank) {
    guarded(__func__, [&] { energy_g(rank_sys("get_energy_g2_", *rank), coords, *nat, energy, grads); });
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
    *E1b = s.OneBodyEnergy(false);
    *E2b = s.TwoBodyEnergy(false);
    *E3b = s.ThreeBodyEnergy(false);
    *E4b = s.FourBodyEnergy(false);  // external_call.cpp:170 (0 unless a 4B cutoff is set)
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

void get_xyz_(double* coords) {
    guarded(__func__, [&] {
        std::vector<double> xyz = sys("get_xyz_").GetXyz();
        std::copy(xyz.begin(), xyz.end(), coords);
    });
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

void finalize_system2_(void) {
    for (bblock::System* s : g_rank_sys) delete s;
    g_rank_sys.clear();
}

}  // extern "C"
