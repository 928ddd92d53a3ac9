// Test driver for the C++ drop-in boundary (include/mbx/system.h, include/mbx/read_nrg.h): every
// call is written the way the reference's own callers write it (src/main/single_point.cpp:38-80,
// src/unittests/unittest-system.cpp:67-123, 644-718) -- a std::vector<bblock::System> filled by
// tools::ReadNrg, SetUpFromJson, Energy / OneBodyEnergy / TwoBodyEnergy / ThreeBodyEnergy,
// GetGrads / GetRealGrads / GetVirial, GetPairList, SetXyz / GetXyz, copies -- and the results
// are printed as one JSON object for tests/test_gpu_system_api.py to check.
//
//   system_api terms <file.nrg> [mbx.json]           per-term energies (without and with
//                                                     gradients), gradients, virials, sites
//   system_api pairlist <file.nrg> <cutoff> <nmax>   GetPairList(nmax, cutoff, 0, nmon)
//   system_api copies <file.nrg> [mbx.json]          the same energy from copies and moves
//   system_api xyz <file.nrg> <mbx.json> <mon> <dx dy dz>
//                                                     SetRealXyz with monomer <mon> shifted, then
//                                                     GetRealXyz / GetXyz / SetXyz round trip
//   system_api jsonbox <file.nrg> <box.json> <nobox.json>
//                                                     SetUpFromJson twice, the second without a box
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "mbx/read_nrg.h"
#include "mbx/system.h"

namespace {

struct Out {
    bool first = true;
    void key(const char* k) {
        std::printf("%s\"%s\": ", first ? "{" : ", ", k);
        first = false;
    }
    void num(const char* k, double v) {
        key(k);
        std::printf("%.17g", v);
    }
    void vec(const char* k, const std::vector<double>& v) {
        key(k);
        std::printf("[");
        for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
        std::printf("]");
    }
    void ivec(const char* k, const std::vector<size_t>& v) {
        key(k);
        std::printf("[");
        for (size_t i = 0; i < v.size(); ++i) std::printf("%s%zu", i ? ", " : "", v[i]);
        std::printf("]");
    }
    void svec(const char* k, const std::vector<std::string>& v) {
        key(k);
        std::printf("[");
        for (size_t i = 0; i < v.size(); ++i) std::printf("%s\"%s\"", i ? ", " : "", v[i].c_str());
        std::printf("]");
    }
    void str(const char* k, const std::string& v) {
        key(k);
        std::printf("\"%s\"", v.c_str());
    }
    void end() { std::printf("%s}\n", first ? "{" : ""); }
};

std::vector<bblock::System> read(const char* nrg) {
    std::vector<bblock::System> systems;
    tools::ReadNrg(nrg, systems);
    return systems;
}

// The TestEnergy sections of unittest-system.cpp:67-123 for one term.
template <class F>
void term(Out& o, bblock::System& s, const std::string& name, F energy) {
    const double e0 = energy(false);
    const double e1 = energy(true);
    o.num((name + "_nograd").c_str(), e0);
    o.num(name.c_str(), e1);
    o.vec((name + "_grads").c_str(), s.GetGrads());
    o.vec((name + "_real_grads").c_str(), s.GetRealGrads());
    o.vec((name + "_virial").c_str(), s.GetVirial());
}

int terms(int argc, char** argv) {
    std::vector<bblock::System> systems = read(argv[2]);
    bblock::System& s = systems[0];
    s.SetUpFromJson(argc > 3 ? argv[3] : nullptr);
    Out o;
    o.num("nsystems", (double)systems.size());
    o.num("nsites", (double)s.GetNumSites());
    o.num("natoms", (double)s.GetNumRealSites());
    o.svec("atom_names", s.GetAtomNames());
    o.svec("real_atom_names", s.GetRealAtomNames());
    term(o, s, "total", [&](bool g) { return s.Energy(g); });
    term(o, s, "1b", [&](bool g) { return s.OneBodyEnergy(g); });
    term(o, s, "2b", [&](bool g) { return s.TwoBodyEnergy(g); });
    term(o, s, "3b", [&](bool g) { return s.ThreeBodyEnergy(g); });
    o.vec("xyz", s.GetXyz());
    o.vec("real_xyz", s.GetRealXyz());
    o.end();
    return 0;
}

int pairlist(int argc, char** argv) {
    if (argc < 5) return 2;
    std::vector<bblock::System> systems = read(argv[2]);
    for (size_t i = 0; i < systems.size(); ++i) systems[i].SetUpFromJson();  // unittest-system.cpp:659-661
    bblock::System& s = systems[0];
    const double cutoff = std::atof(argv[3]);
    const size_t nmax = (size_t)std::atoi(argv[4]);
    Out o;
    o.ivec("nmers", s.GetPairList(nmax, cutoff, 0, s.GetNumMon()));
    bool threw = false;
    try {
        s.GetPairList(4, cutoff, 0, s.GetNumMon());  // unittest-system.cpp:706-716
    } catch (const std::exception&) {
        threw = true;
    }
    o.num("nmax4_threw", threw ? 1.0 : 0.0);
    o.end();
    return 0;
}

int copies(int argc, char** argv) {
    std::vector<bblock::System> systems = read(argv[2]);
    systems[0].SetUpFromJson(argc > 3 ? argv[3] : nullptr);
    const double e_orig = systems[0].Energy(true);
    const std::vector<double> g_orig = systems[0].GetRealGrads();
    bblock::System copy(systems[0]);        // copy after a computation: a context of its own
    bblock::System assigned;
    assigned = systems[0];                  // copy assignment
    std::vector<bblock::System> grown(systems);  // copies of every system in a vector
    grown.push_back(systems[0]);            // and a reallocation that moves them
    bblock::System moved(std::move(copy));  // move construction
    Out o;
    o.num("orig", e_orig);
    o.vec("orig_grads", g_orig);
    o.num("orig_again", systems[0].Energy(true));
    o.num("assigned", assigned.Energy(true));
    o.vec("assigned_grads", assigned.GetRealGrads());
    o.num("grown_last", grown.back().Energy(true));
    o.num("moved", moved.Energy(true));
    o.vec("moved_grads", moved.GetRealGrads());
    o.end();
    return 0;
}

int xyz(int argc, char** argv) {
    if (argc < 8) return 2;
    std::vector<bblock::System> systems = read(argv[2]);
    bblock::System& s = systems[0];
    s.SetUpFromJson(argv[3]);
    const size_t mon = (size_t)std::atoi(argv[4]);
    const double d[3] = {std::atof(argv[5]), std::atof(argv[6]), std::atof(argv[7])};
    std::vector<double> r = s.GetRealXyz();
    Out o;
    o.vec("wrapped_initial", r);
    const size_t first = s.GetFirstInd(mon), nat = s.GetMonNumAt(mon);
    for (size_t a = first; a < first + nat; ++a)
        for (int k = 0; k < 3; ++k) r[3 * a + k] += (a == first ? d[k] : -d[k]);  // split the monomer
    s.SetRealXyz(r);
    o.vec("set", r);
    o.vec("get_real", s.GetRealXyz());  // wrapped by SetPBC(box_) first
    const std::vector<double> all = s.GetXyz();
    o.vec("get_all", all);
    s.SetXyz(all);
    o.vec("after_setxyz", s.GetRealXyz());
    o.num("energy", s.Energy(true));
    o.vec("grads", s.GetRealGrads());
    bool threw = false;
    try {
        s.SetXyz(std::vector<double>(3, 0.0));
    } catch (const std::exception&) {
        threw = true;
    }
    o.num("bad_size_threw", threw ? 1.0 : 0.0);
    o.end();
    return 0;
}

int jsonbox(int argc, char** argv) {
    if (argc < 5) return 2;
    std::vector<bblock::System> systems = read(argv[2]);
    bblock::System& s = systems[0];
    s.SetUpFromJson(argv[3]);
    s.SetUpFromJson(argv[4]);  // no "box" key: the box set before is kept (system.cpp:1200-1206)
    Out o;
    o.vec("box", s.GetBox());
    o.num("energy", s.Energy(true));
    o.vec("grads", s.GetRealGrads());
    o.end();
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s terms|pairlist|copies|xyz|jsonbox <file.nrg> ...\n", argv[0]);
        return 2;
    }
    try {
        const std::string cmd = argv[1];
        if (cmd == "terms") return terms(argc, argv);
        if (cmd == "pairlist") return pairlist(argc, argv);
        if (cmd == "copies") return copies(argc, argv);
        if (cmd == "xyz") return xyz(argc, argv);
        if (cmd == "jsonbox") return jsonbox(argc, argv);
        std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
