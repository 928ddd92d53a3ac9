// bblock::System -- the reference's public System API for the many-body short-range path,
// backed by the MI355X engine (include/mbxg.h).
//
// Mirrors src/bblock/system.h of MBX v1.3.3 for the members a caller of the 1B/2B/3B path uses:
// same names, by-value arguments and return values, the same internal monomer ordering
// (types from fewest to most monomers, input order within a type; sys_tools.cpp:44-131), the
// same site layout (systools::SetUpMonomers, sys_tools.cpp:133-255: h2o 4 sites for 3 atoms,
// ch4 5/5, co2 3/3, ...) and CUException errors.  Energy() returns the 1B + 2B + 3B terms, plus
// the 4B term when a 4B cutoff is set; the dispersion, Buckingham and electrostatics terms of the
// reference's Energy (system.cpp:1995-2138) are outside this engine and contribute nothing here.
//
// A System is copyable, as the reference's is (tools::ReadNrg fills a std::vector<System>).
// Its device context (one HIP stream and the device buffers) is created on the first call that
// needs the GPU, separately for every copy, and never shared between copies.
#pragma once

#include <cstddef>
#include <exception>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

struct mbxg_ctx;

namespace bblock {

// custom_exceptions.h of the reference: a std::exception carrying a formatted message.
class CUException : public std::exception {
   public:
    CUException(const char* func, const char* file, int line, const std::string& text);
    const char* what() const noexcept override { return msg_.c_str(); }

   private:
    std::string msg_;
};

class System {
   public:
    System();
    ~System();
    System(const System& other);
    System& operator=(const System& other);
    System(System&& other) noexcept;
    System& operator=(System&& other) noexcept;

    // system.h:564, system.cpp:848-870
    void AddMonomer(std::vector<double> xyz, std::vector<std::string> atoms, std::string id, size_t islocal = 1,
                    int tag = 0);
    // system.h:578, system.cpp:872: a molecule = the indices (input order) of its monomers
    void AddMolecule(std::vector<size_t> molec);
    // system.h:713, system.cpp:1000.  Unknown monomer ids throw "No data in the dataset for
    // monomer: <id>" (sys_tools.cpp:232-241).
    void Initialize();
    // system.h:770-782, system.cpp:1198-1608.  Keys read under "MBX": box, twobody_cutoff,
    // threebody_cutoff, max_n_eval_2b, max_n_eval_3b, ignore_2b_poly, ignore_3b_poly.  A key
    // that is absent keeps the current value (the box included).
    void SetUpFromJson(char* json_file = 0);
    void SetUpFromJson(std::string json_text);
    // system.h:782 takes an nlohmann::json; any JSON object with dump() (nlohmann's included) is
    // accepted here through its text, so this header does not depend on a JSON library.
    template <class Json, class = decltype(std::declval<const Json&>().dump())>
    void SetUpFromJson(const Json& j) {
        SetUpFromJson(std::string(j.dump()));
    }

    // system.h:1057, 1088, 1100
    double Energy(bool do_grads);
    double OneBodyEnergy(bool do_grads);
    double TwoBodyEnergy(bool do_grads, bool use_ghost = 0);
    double ThreeBodyEnergy(bool do_grads, bool use_ghost = 0);
    // system.h:1112: the 4B energy (0 unless Set4bCutoff / "fourbody_cutoff" > 0)
    double FourBodyEnergy(bool do_grads, bool use_ghost = 0);

    // system.h:278, 285, 450: gradients of the last energy call in input order, of every site
    // (virtual sites included, which have none from these terms) or of the real atoms; virial
    std::vector<double> GetGrads();
    std::vector<double> GetRealGrads();
    std::vector<double> GetVirial();
    // system.h:262-270, system.cpp:392-401: coordinates in input order, wrapped into the cell
    // under PBC first (SetPBC(box_)); GetXyz includes the virtual sites (the water M-site,
    // systools::SetVSites, sys_tools.cpp:874-915)
    std::vector<double> GetXyz();
    std::vector<double> GetRealXyz();
    // system.h:534, 542: set every site's coordinates (virtual sites are recomputed, as the
    // reference's SetPBC does) or the real atoms'
    void SetXyz(std::vector<double> xyz);
    void SetRealXyz(std::vector<double> xyz);
    // system.h:925
    void SetPBC(std::vector<double> box = {});
    std::vector<double> GetBox();

    void Set2bCutoff(double cutoff2b);
    void Set3bCutoff(double cutoff3b);
    double Get2bCutoff();
    double Get3bCutoff();
    void SetNMaxEval2b(size_t nmax);
    void SetNMaxEval3b(size_t nmax);
    size_t GetMaxEval2b();
    size_t GetMaxEval3b();
    void Set2bIgnorePoly(std::vector<std::vector<std::string>> ignore_2b);
    void Set3bIgnorePoly(std::vector<std::vector<std::string>> ignore_3b);
    // system.h:408, 444, 699, 705, 842: the 4B term's cutoff and ignore list
    void Set4bCutoff(double cutoff4b);
    double Get4bCutoff();
    void Set4bIgnorePoly(std::vector<std::vector<std::string>> ignore_4b);
    void Add4bIgnorePoly(std::string mon1, std::string mon2, std::string mon3, std::string mon4);
    std::vector<std::vector<std::string>> Get4bIgnorePoly();

    // system.h:246, system.cpp:313: n-mers (input monomer indices) with a member in [istart, iend)
    std::vector<size_t> GetPairList(size_t nmax, double cutoff, size_t istart, size_t iend, bool use_ghost = false);

    size_t GetNumMon();
    // system.cpp:136 and :387 (the count is molecules_.size(), as system.cpp:1032 sets it at Initialize)
    size_t GetNumMol();
    std::vector<size_t> GetMolecule(size_t n);
    // system.h:120-125: sites (virtual ones included) and real atoms
    size_t GetNumSites();
    size_t GetNumRealSites();
    std::string GetMonId(size_t n);
    // system.h:151, 207, 353 -- here in input order (real atoms only): atoms of monomer n, its first
    // atom's index, every atom's name
    size_t GetMonNumAt(size_t n);
    size_t GetFirstInd(size_t n);
    std::vector<std::string> GetRealAtomNames();
    // system.h:345: every site's name in input order, "virt" for virtual sites
    std::vector<std::string> GetAtomNames();

    // The n-mer counts of the last energy call: listed / evaluated dimers, listed / evaluated trimers.
    std::vector<long> GetNmerCounts();

   private:
    void check_initialized(const char* func, int line);
    double run_terms(int terms, bool do_grads, bool use_ghost);
    mbxg_ctx* device();  // the context, created (topology, box) on first use
    void apply_box();
    void wrap_xyz();     // FixMonomerCoordinates of xyz_ under PBC (SetPBC's side effect)
    std::vector<double> real_xyz_input_order() const;  // xyz_ in input atom order, as stored

    mbxg_ctx* ctx_ = nullptr;
    bool initialized_ = false;
    // input order
    std::vector<double> in_xyz_;
    std::vector<std::string> in_atoms_;
    std::vector<std::string> in_mon_;
    std::vector<std::vector<size_t>> molecules_;
    std::vector<size_t> in_local_;
    std::vector<int> in_tag_;  // per atom
    std::vector<size_t> in_nat_, in_first_;
    std::vector<size_t> in_sites_, in_first_site_;  // per monomer, set at Initialize
    int tag_counter_ = 0;
    // internal order (OrderMonomers): internal monomer i is input monomer order_[i]
    std::vector<size_t> order_, first_;  // first_: first real atom of internal monomer i
    std::vector<int> topo_kind_, topo_nat_, topo_tag_;
    std::vector<unsigned char> topo_local_;
    std::vector<double> xyz_;            // internal order, real atoms
    std::vector<double> grad_;           // internal order, real atoms
    std::vector<double> virial_;
    std::vector<double> box_;
    double cutoff2b_ = 100.0, cutoff3b_ = 5.0;
    size_t max2b_ = 1024, max3b_ = 1024;
    double cutoff4b_ = 0.0;  // system.cpp:89
    std::vector<std::vector<std::string>> ignore2b_, ignore3b_, ignore4b_;
    std::vector<long> counts_ = std::vector<long>(4, 0);
    size_t numat_ = 0, numsites_ = 0;
};

}  // namespace bblock
