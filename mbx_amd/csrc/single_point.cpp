// single_point: the 1B, 2B and 3B energies, virials and, on request, gradients of the
// first system of an NRG file -- the reference's src/tests/single_point.cpp (ReadNrg, then
// SetUpFromJson with the optional json file, then Energy(true) and the per-term block of its
// PRINT_VIRIAL / PRINT_GRADS output) for the terms this library computes.
//
//   single_point <file.nrg> [mbx.json] [--grads] [--write <out.nrg>]
//                                       --write: every system read, written back with WriteNrg
//                                       (the reference writes input_out.nrg unconditionally)
//   single_point --dump <file.nrg>     every system, molecule, monomer and atom the file holds
//                                       (the parser alone: no device is used)
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mbx/read_nrg.h"
#include "mbx/system.h"

namespace {

int dump(const char* path) {
    std::ifstream ifs(path);
    if (!ifs) throw std::runtime_error(std::string("could not open the NRG file ") + path);
    std::vector<tools::NrgSystem> systems;
    tools::ParseNrg(ifs, systems);
    std::printf("systems %zu\n", systems.size());
    for (size_t s = 0; s < systems.size(); ++s) {
        const tools::NrgSystem& sys = systems[s];
        std::printf("system %zu monomers %zu molecules %zu\n", s, sys.monomers.size(), sys.molecules.size());
        for (const auto& mol : sys.molecules) {
            std::printf("molecule");
            for (size_t m : mol) std::printf(" %zu", m);
            std::printf("\n");
        }
        for (const tools::NrgMonomer& m : sys.monomers) {
            std::printf("monomer %s %zu\n", m.id.c_str(), m.atoms.size());
            for (size_t a = 0; a < m.atoms.size(); ++a)
                std::printf("%s %.17g %.17g %.17g\n", m.atoms[a].c_str(), m.xyz[3 * a], m.xyz[3 * a + 1],
                            m.xyz[3 * a + 2]);
        }
    }
    return 0;
}

void print_virial(const char* term, const std::vector<double>& v) {
    std::printf(" %s Virial [kcal/mol]", term);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s %.15e", i ? " ," : "", v[i]);
    std::printf("\n");
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc == 3 && std::strcmp(argv[1], "--dump") == 0) return dump(argv[2]);
        const char* nrg = nullptr;
        char* json = nullptr;
        const char* out = nullptr;
        bool grads = false;
        for (int i = 1; i < argc; ++i) {
            if (std::strcmp(argv[i], "--grads") == 0)
                grads = true;
            else if (std::strcmp(argv[i], "--write") == 0 && i + 1 < argc)
                out = argv[++i];
            else if (!nrg)
                nrg = argv[i];
            else if (!json)
                json = argv[i];
            else
                nrg = nullptr, i = argc;
        }
        if (!nrg) {
            std::fprintf(stderr, "usage: %s <file.nrg> [mbx.json] [--grads] [--write <out.nrg>]\n       %s --dump <file.nrg>\n", argv[0],
                         argv[0]);
            return 1;
        }
        std::vector<bblock::System> systems;
        tools::ReadNrg(nrg, systems);
        if (out) tools::WriteNrg(out, systems);
        bblock::System& sys = systems[0];
        sys.SetUpFromJson(json);
        const double e = sys.Energy(true);
        std::printf("Energy= %.15e\n", e);
        const std::vector<double> g = sys.GetRealGrads();
        const double e1 = sys.OneBodyEnergy(true);
        std::printf(" 1B Energy [kcal/mol] %.15e\n", e1);
        print_virial("1B", sys.GetVirial());
        const double e2 = sys.TwoBodyEnergy(true);
        std::printf(" 2B Energy [kcal/mol] %.15e\n", e2);
        print_virial("2B", sys.GetVirial());
        const double e3 = sys.ThreeBodyEnergy(true);
        std::printf(" 3B Energy [kcal/mol] %.15e\n", e3);
        print_virial("3B", sys.GetVirial());
        if (grads)  // the gradients of Energy(true): 1B + 2B + 3B, input atom order
            for (size_t i = 0; i < g.size() / 3; ++i)
                std::printf("%.15e %.15e %.15e\n", g[3 * i], g[3 * i + 1], g[3 * i + 2]);
    } catch (const std::exception& ex) {
        std::cerr << " ** Error ** : " << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
