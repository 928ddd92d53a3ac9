// bblock::System over the mbxg engine (see include/mbx/system.h).
//
// Bookkeeping restated from the reference (paths relative to its src/):
//   AddMonomer tags          bblock/system.cpp:848-870
//   SetUpMonomers sites      bblock/sys_tools.cpp:133-255 (sites and atoms per monomer id)
//   OrderMonomers            bblock/sys_tools.cpp:44-131 (types from fewest to most monomers)
//   SetUpFromJson keys       bblock/system.cpp:1198-1581 (defaults :78, :85, :98-100)
//   SetPBC box forms         bblock/system.cpp:712-763, tools/math_tools.cpp:83-152
//   SetXyz / SetRealXyz      bblock/system.cpp:812-846
//   ResetOrder(Real)3N       bblock/sys_tools.cpp:843-872 (GetGrads / GetRealGrads / GetXyz / GetRealXyz)
//   SetVSites                bblock/sys_tools.cpp:874-915 (the water M-site of GetXyz)
//   GetPairList              bblock/system.cpp:313-380
#include "../../include/mbx/system.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

#include "../../include/mbxg.h"
#include "json_min.hpp"

namespace bblock {

CUException::CUException(const char* func, const char* file, int line, const std::string& text) {
    std::ostringstream os;
    os << "\n\n[ERROR]  Function: " << func << " in " << file << ":" << line << "\n[ERROR]  " << text << "\n";
    msg_ = os.str();
}

#define MBX_THROW(text) throw CUException(__func__, __FILE__, __LINE__, (text))

namespace {

int kind_of(const std::string& id) {
    if (id == "h2o") return MBXG_H2O;
    if (id == "ch4") return MBXG_CH4;
    if (id == "co2") return MBXG_CO2;
    if (id == "f-") return MBXG_F;
    if (id == "cl-") return MBXG_CL;
    if (id == "br-") return MBXG_BR;
    if (id == "i-") return MBXG_I;
    if (id == "li+") return MBXG_LI;
    if (id == "na+") return MBXG_NA;
    if (id == "k+") return MBXG_K;
    if (id == "rb+") return MBXG_RB;
    if (id == "cs+") return MBXG_CS;
    return MBXG_OTHER;
}

// systools::SetUpMonomers (sys_tools.cpp:133-255): sites (real atoms and virtual sites) and real
// atoms of each monomer id the reference knows; false for any other id.
bool site_info(const std::string& id, size_t* sites, size_t* nat) {
    struct Row {
        const char* id;
        size_t sites, nat;
    };
    static const Row rows[] = {
        {"h2o", 4, 3},       {"dp2", 2, 2},       {"dp1", 1, 1},      {"dp1p", 1, 1},       {"ch4", 5, 5},
        {"co2_archive", 3, 3}, {"nh3", 4, 4},     {"nh3pbe0d3bj", 4, 4}, {"h4_dummy", 4, 4}, {"co2", 3, 3},
        {"co2cm5100", 3, 3}, {"co2cm595", 3, 3},  {"co2cm590", 3, 3}, {"co2cm5875", 3, 3},  {"co2cm585", 3, 3},
        {"co2cm580", 3, 3},  {"n2o5", 7, 7},      {"ar", 1, 1},       {"he", 1, 1},         {"dummy", 1, 1},
        {"f-", 1, 1},        {"cl-", 1, 1},       {"br-", 1, 1},      {"i-", 1, 1},         {"li+", 1, 1},
        {"na+", 1, 1},       {"k+", 1, 1},        {"rb+", 1, 1},      {"cs+", 1, 1},        {"h2", 2, 2},
        {"mbpbe", 4, 3},
    };
    for (const Row& r : rows)
        if (id == r.id) {
            *sites = r.sites;
            *nat = r.nat;
            return true;
        }
    return false;
}

// The water M-site (sys_tools.h:77-87): gamma1 * O + gamma2 * (H1 + H2)
constexpr double kGammaM = 0.426706882;
constexpr double kGamma1 = 1.0 - kGammaM;
constexpr double kGamma2 = kGammaM / 2;

void check(int rc) {
    if (rc != MBXG_OK) throw CUException("mbxg", __FILE__, __LINE__, mbxg_last_error());
}

// BoxABCabcToBoxVec (tools/math_tools.cpp:132-152)
std::vector<double> abc_to_vec(const std::vector<double>& box) {
    const double A = box[0], B = box[1], C = box[2];
    const double alpha = box[3] / 180.0 * M_PI, beta = box[4] / 180.0 * M_PI, gamma = box[5] / 180.0 * M_PI;
    std::vector<double> out(9, 0.0);
    out[0] = A;
    out[3] = B * cos(gamma);
    out[4] = B * sin(gamma);
    out[6] = C * cos(beta);
    const double tmp = (cos(alpha) - cos(beta) * cos(gamma)) / sin(gamma);
    out[7] = C * tmp;
    out[8] = C * sqrt(1.0 - cos(beta) * cos(beta) - tmp * tmp);
    return out;
}

std::vector<std::vector<std::string>> read_ignore(const mbxjson::Value& v) {
    std::vector<std::vector<std::string>> out;
    if (v.kind != mbxjson::Value::Array) return out;
    for (const auto& e : v.arr) {
        std::vector<std::string> p;
        for (const auto& s : e.arr) p.push_back(s.as_string());
        std::sort(p.begin(), p.end());
        out.push_back(p);
    }
    return out;
}

bool ignored(const std::vector<std::vector<std::string>>& list, std::vector<std::string> key) {
    std::sort(key.begin(), key.end());
    return std::find(list.begin(), list.end(), key) != list.end();
}

}  // namespace

System::System() : virial_(9, 0.0) {}

System::~System() {
    if (ctx_) mbxg_destroy(ctx_);
}

// A copy holds the same state and no device context (it makes its own on first use).
System::System(const System& o)
    : initialized_(o.initialized_),
      in_xyz_(o.in_xyz_),
      in_atoms_(o.in_atoms_),
      in_mon_(o.in_mon_),
      molecules_(o.molecules_),
      in_local_(o.in_local_),
      in_tag_(o.in_tag_),
      in_nat_(o.in_nat_),
      in_first_(o.in_first_),
      in_sites_(o.in_sites_),
      in_first_site_(o.in_first_site_),
      tag_counter_(o.tag_counter_),
      order_(o.order_),
      first_(o.first_),
      topo_kind_(o.topo_kind_),
      topo_nat_(o.topo_nat_),
      topo_tag_(o.topo_tag_),
      topo_local_(o.topo_local_),
      xyz_(o.xyz_),
      grad_(o.grad_),
      virial_(o.virial_),
      box_(o.box_),
      cutoff2b_(o.cutoff2b_),
      cutoff3b_(o.cutoff3b_),
      max2b_(o.max2b_),
      max3b_(o.max3b_),
      cutoff4b_(o.cutoff4b_),
      ignore2b_(o.ignore2b_),
      ignore3b_(o.ignore3b_),
      ignore4b_(o.ignore4b_),
      counts_(o.counts_),
      numat_(o.numat_),
      numsites_(o.numsites_) {}

System& System::operator=(const System& o) {
    if (this != &o) {
        System tmp(o);
        *this = std::move(tmp);
    }
    return *this;
}

System::System(System&& o) noexcept : System() { *this = std::move(o); }

System& System::operator=(System&& o) noexcept {
    if (this == &o) return *this;
    if (ctx_) mbxg_destroy(ctx_);
    ctx_ = o.ctx_;
    o.ctx_ = nullptr;
    initialized_ = o.initialized_;
    in_xyz_ = std::move(o.in_xyz_);
    in_atoms_ = std::move(o.in_atoms_);
    in_mon_ = std::move(o.in_mon_);
    molecules_ = std::move(o.molecules_);
    in_local_ = std::move(o.in_local_);
    in_tag_ = std::move(o.in_tag_);
    in_nat_ = std::move(o.in_nat_);
    in_first_ = std::move(o.in_first_);
    in_sites_ = std::move(o.in_sites_);
    in_first_site_ = std::move(o.in_first_site_);
    tag_counter_ = o.tag_counter_;
    order_ = std::move(o.order_);
    first_ = std::move(o.first_);
    topo_kind_ = std::move(o.topo_kind_);
    topo_nat_ = std::move(o.topo_nat_);
    topo_tag_ = std::move(o.topo_tag_);
    topo_local_ = std::move(o.topo_local_);
    xyz_ = std::move(o.xyz_);
    grad_ = std::move(o.grad_);
    virial_ = std::move(o.virial_);
    box_ = std::move(o.box_);
    cutoff2b_ = o.cutoff2b_;
    cutoff3b_ = o.cutoff3b_;
    cutoff4b_ = o.cutoff4b_;
    max2b_ = o.max2b_;
    max3b_ = o.max3b_;
    ignore2b_ = std::move(o.ignore2b_);
    ignore3b_ = std::move(o.ignore3b_);
    ignore4b_ = std::move(o.ignore4b_);
    counts_ = std::move(o.counts_);
    numat_ = o.numat_;
    numsites_ = o.numsites_;
    return *this;
}

void System::AddMonomer(std::vector<double> xyz, std::vector<std::string> atoms, std::string id, size_t islocal,
                        int tag) {
    if (initialized_)
        MBX_THROW("The system has already been initialized. Adding a new monomer is not possible");
    if (xyz.size() != 3 * atoms.size()) MBX_THROW("Number of coordinates is not consistent with number of atoms");
    in_first_.push_back(in_atoms_.size());
    in_nat_.push_back(atoms.size());
    in_xyz_.insert(in_xyz_.end(), xyz.begin(), xyz.end());
    in_atoms_.insert(in_atoms_.end(), atoms.begin(), atoms.end());
    in_mon_.push_back(id);
    in_local_.push_back(islocal);
    for (size_t i = 0; i < atoms.size(); ++i) in_tag_.push_back(tag == 0 ? tag_counter_ + (int)i : tag + (int)i);
    tag_counter_ += (int)atoms.size();
}

void System::AddMolecule(std::vector<size_t> molec) { molecules_.push_back(std::move(molec)); }

void System::Initialize() {
    if (initialized_) MBX_THROW("The system has already been initialized. Reinitialization is not possible");
    const size_t nmon = in_mon_.size();
    if (nmon == 0) MBX_THROW("Cannot initialize the system with less than 3 coordinates, and current size of xyz is 0");
    in_sites_.assign(nmon, 0);
    in_first_site_.assign(nmon, 0);
    numsites_ = 0;
    for (size_t i = 0; i < nmon; ++i) {
        size_t ns = 0, na = 0;
        if (!site_info(in_mon_[i], &ns, &na)) MBX_THROW("No data in the dataset for monomer: " + in_mon_[i]);
        if (na != in_nat_[i])
            MBX_THROW("Monomer " + std::to_string(i) + " (" + in_mon_[i] + ") has " + std::to_string(in_nat_[i]) +
                      " atoms; " + in_mon_[i] + " has " + std::to_string(na));
        in_sites_[i] = ns;
        in_first_site_[i] = numsites_;
        numsites_ += ns;
    }
    // OrderMonomers: distinct types in order of appearance, then repeatedly the type with the
    // fewest monomers (first such on ties), each in input order.
    std::vector<std::pair<std::string, size_t>> types;
    for (const auto& m : in_mon_) {
        auto it = std::find_if(types.begin(), types.end(), [&](const auto& p) { return p.first == m; });
        if (it == types.end())
            types.push_back({m, 1});
        else
            ++it->second;
    }
    order_.clear();
    while (!types.empty()) {
        size_t best = 0;
        for (size_t t = 1; t < types.size(); ++t)
            if (types[t].second < types[best].second) best = t;
        for (size_t i = 0; i < nmon; ++i)
            if (in_mon_[i] == types[best].first) order_.push_back(i);
        types.erase(types.begin() + best);
    }
    first_.assign(nmon, 0);
    numat_ = 0;
    topo_kind_.assign(nmon, 0);
    topo_nat_.assign(nmon, 0);
    topo_tag_.assign(nmon, 0);
    topo_local_.assign(nmon, 0);
    for (size_t i = 0; i < nmon; ++i) {
        const size_t k = order_[i];
        first_[i] = numat_;
        numat_ += in_nat_[k];
        topo_kind_[i] = kind_of(in_mon_[k]);
        topo_nat_[i] = (int)in_nat_[k];
        topo_tag_[i] = in_tag_[in_first_[k]];
        topo_local_[i] = in_local_[k] ? 1 : 0;
    }
    xyz_.assign(3 * numat_, 0.0);
    grad_.assign(3 * numat_, 0.0);
    for (size_t i = 0; i < nmon; ++i) {
        const size_t k = order_[i];
        std::copy(in_xyz_.begin() + 3 * in_first_[k], in_xyz_.begin() + 3 * (in_first_[k] + in_nat_[k]),
                  xyz_.begin() + 3 * first_[i]);
    }
    initialized_ = true;
    wrap_xyz();  // SetPBC(box_) (system.cpp:1037)
}

mbxg_ctx* System::device() {
    if (!ctx_) {
        mbxg_ctx* c = nullptr;
        check(mbxg_create(&c, 0));
        const int rc = mbxg_set_topology(c, (int)topo_kind_.size(), topo_kind_.data(), topo_nat_.data(),
                                         topo_tag_.data(), topo_local_.data());
        if (rc != MBXG_OK) {
            const std::string msg = mbxg_last_error();
            mbxg_destroy(c);
            MBX_THROW(msg);
        }
        ctx_ = c;
        // steady-state repeats of one call (same terms, same sizes) replay a captured HIP graph
        check(mbxg_set_graph(ctx_, 1));
        apply_box();
    }
    return ctx_;
}

void System::apply_box() {
    if (!ctx_) return;
    check(mbxg_set_box(ctx_, box_.size() == 9 ? box_.data() : nullptr));
}

// FixMonomerCoordinates of the stored coordinates under PBC (SetPBC, system.cpp:742-747), on the
// device with the wrap of every compute.
void System::wrap_xyz() {
    if (!initialized_ || box_.size() != 9 || numat_ == 0) return;
    check(mbxg_fix_coordinates(device(), xyz_.data(), xyz_.data()));
}

void System::SetUpFromJson(char* json_file) {
    std::string text = "{}";
    if (json_file) {
        std::ifstream f(json_file);
        if (f) {
            std::stringstream ss;
            ss << f.rdbuf();
            text = ss.str();
        } else {
            std::cerr << "**WARNING** json file " << json_file << " could not be read. Using defaults.\n";
        }
    }
    SetUpFromJson(text);
}

void System::SetUpFromJson(std::string json_text) {
    mbxjson::Value root;
    try {
        root = mbxjson::parse(json_text);
    } catch (const std::exception& e) {
        std::cerr << "**WARNING** json parse error (" << e.what() << "). Using defaults.\n";
        root = mbxjson::parse("{}");
    }
    const mbxjson::Value* mbx = root.kind == mbxjson::Value::Object && root.has("MBX") ? &root.at("MBX") : nullptr;
    if (mbx && mbx->kind != mbxjson::Value::Object) mbx = nullptr;
    auto num = [&](const char* k, double dflt) { return mbx && mbx->has(k) ? mbx->at(k).as_number() : dflt; };
    // a key that is absent keeps the current value (system.cpp:1200-1206 keeps box_)
    std::vector<double> box = box_;
    if (mbx && mbx->has("box")) {
        box.clear();
        for (const auto& v : mbx->at("box").arr) box.push_back(v.as_number());
    }
    cutoff2b_ = num("twobody_cutoff", cutoff2b_);
    cutoff3b_ = num("threebody_cutoff", cutoff3b_);
    cutoff4b_ = num("fourbody_cutoff", cutoff4b_);  // system.cpp:1287-1295
    max2b_ = (size_t)num("max_n_eval_2b", (double)max2b_);
    max3b_ = (size_t)num("max_n_eval_3b", (double)max3b_);
    if (mbx && mbx->has("ignore_2b_poly")) ignore2b_ = read_ignore(mbx->at("ignore_2b_poly"));
    if (mbx && mbx->has("ignore_3b_poly")) ignore3b_ = read_ignore(mbx->at("ignore_3b_poly"));
    if (!initialized_ && !in_mon_.empty()) Initialize();
    SetPBC(box);
}

void System::SetPBC(std::vector<double> box) {
    if (box.size() != 9 && box.size() != 6 && box.size() != 0)
        MBX_THROW("Box size of " + std::to_string(box.size()) + " is not acceptable.");
    std::vector<double> b = box.size() == 6 ? abc_to_vec(box) : box;
    if (b.size() == 9) {
        const double det = b[0] * (b[4] * b[8] - b[7] * b[5]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                           b[2] * (b[3] * b[7] - b[4] * b[6]);
        if (std::fabs(det) < 1E-36) MBX_THROW("Matrix has no inverse. Determinant is 0.");
    }
    box_ = b;
    apply_box();
    wrap_xyz();
}

std::vector<double> System::GetBox() { return box_; }

void System::check_initialized(const char* func, int line) {
    if (!initialized_)
        throw CUException(func, __FILE__, line, "System has not been initialized. Energy calculation not possible.");
}

double System::run_terms(int terms, bool do_grads, bool use_ghost) {
    mbxg_ctx* c = device();
    // use_ghost: local-ghost n-mers from local query monomers (sys_tools.cpp:502-537)
    check(mbxg_set_use_ghost(c, use_ghost ? 1 : 0));
    std::fill(grad_.begin(), grad_.end(), 0.0);
    std::fill(virial_.begin(), virial_.end(), 0.0);
    // ignore_2b_poly / ignore_3b_poly (system.cpp:2606-2619, 2940-2953): those n-mers are listed
    // but not evaluated
    int fam = 0;
    if (!ignored(ignore2b_, {"h2o", "h2o"})) fam |= MBXG_FAM_MBPOL_2B;
    if (!ignored(ignore2b_, {"ch4", "h2o"})) fam |= MBXG_FAM_CH4H2O_2B;
    if (!ignored(ignore2b_, {"ch4", "ch4"})) fam |= MBXG_FAM_CH4CH4_2B;
    if (!ignored(ignore3b_, {"h2o", "h2o", "h2o"})) fam |= MBXG_FAM_MBPOL_3B;
    if (!ignored(ignore3b_, {"ch4", "h2o", "h2o"})) fam |= MBXG_FAM_CH4H2O2_3B;
    if (!ignored(ignore2b_, {"f-", "h2o"})) fam |= MBXG_FAM_F_H2O_2B;
    if (!ignored(ignore2b_, {"cl-", "h2o"})) fam |= MBXG_FAM_CL_H2O_2B;
    if (!ignored(ignore2b_, {"br-", "h2o"})) fam |= MBXG_FAM_BR_H2O_2B;
    if (!ignored(ignore2b_, {"i-", "h2o"})) fam |= MBXG_FAM_I_H2O_2B;
    if (!ignored(ignore2b_, {"co2", "co2"})) fam |= MBXG_FAM_CO2CO2_2B;
    if (!ignored(ignore3b_, {"f-", "h2o", "h2o"})) fam |= MBXG_FAM_F_H2O2_3B;
    if (!ignored(ignore3b_, {"cl-", "h2o", "h2o"})) fam |= MBXG_FAM_CL_H2O2_3B;
    if (!ignored(ignore3b_, {"br-", "h2o", "h2o"})) fam |= MBXG_FAM_BR_H2O2_3B;
    if (!ignored(ignore3b_, {"i-", "h2o", "h2o"})) fam |= MBXG_FAM_I_H2O2_3B;
    if (!ignored(ignore3b_, {"co2", "co2", "co2"})) fam |= MBXG_FAM_CO2X3_3B;
    if (!ignored(ignore4b_, {"h2o", "h2o", "h2o", "h2o"})) fam |= MBXG_FAM_H2OX4_4B;  // system.cpp:3143-3155
    static const char* const kAlkali[5] = {"li+", "na+", "k+", "rb+", "cs+"};
    for (int a = 0; a < 5; ++a) {
        if (!ignored(ignore2b_, {kAlkali[a], "h2o"})) fam |= MBXG_FAM_LI_H2O_2B << a;
        if (!ignored(ignore3b_, {kAlkali[a], "h2o", "h2o"})) fam |= MBXG_FAM_LI_H2O2_3B << a;
    }
    check(mbxg_set_families(c, fam));
    check(mbxg_set_cutoffs(c, cutoff2b_, cutoff3b_));
    // Get4B runs only with a cutoff above machine epsilon (system.cpp:2084, 2762)
    check(mbxg_set_cutoff4b(c, cutoff4b_ > std::numeric_limits<double>::epsilon() ? cutoff4b_ : 0.0));
    check(mbxg_set_coords(c, xyz_.data()));
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    check(mbxg_compute(c, terms, do_grads ? 1 : 0));
    check(mbxg_get_results(c, nullptr, nullptr, do_grads ? grad_.data() : nullptr, virial_.data()));
    check(mbxg_get_term_energies(c, e));
    if (terms & (MBXG_2B | MBXG_3B)) {
        int64_t n[4];
        check(mbxg_get_counts(c, n));
        for (int k = 0; k < 4; ++k) counts_[k] = (long)n[k];
    }
    // SetPBC wraps the working coordinates each call (system.cpp:2015 -> sys_tools.cpp:257-320)
    check(mbxg_get_coords(c, xyz_.data()));
    if (!do_grads) std::fill(virial_.begin(), virial_.end(), 0.0);
    return e[0] + e[1] + e[2] + e[3];
}

// System::Energy (system.cpp:1995-2138) for the terms this engine computes: Get1B + Get2B + Get3B
// (+ Get4B with a 4B cutoff, :2083-2086).
double System::Energy(bool do_grads) {
    if (!initialized_) SetUpFromJson(std::string("{}"));
    check_initialized(__func__, __LINE__);
    return run_terms(MBXG_1B | MBXG_2B | MBXG_3B | MBXG_4B, do_grads, false);
}

// System::OneBodyEnergy (system.cpp:2140-2158): the 1B energy of the local monomers (Get1B,
// :2269-2401), gradients into grad_ and the virial.
double System::OneBodyEnergy(bool do_grads) {
    check_initialized(__func__, __LINE__);
    return run_terms(MBXG_1B, do_grads, false);
}

double System::TwoBodyEnergy(bool do_grads, bool use_ghost) {
    check_initialized(__func__, __LINE__);
    return run_terms(MBXG_2B, do_grads, use_ghost);
}

double System::ThreeBodyEnergy(bool do_grads, bool use_ghost) {
    check_initialized(__func__, __LINE__);
    return run_terms(MBXG_3B, do_grads, use_ghost);
}

// system.cpp:2747-2767: Get4B when the cutoff exceeds machine epsilon, 0 otherwise.
double System::FourBodyEnergy(bool do_grads, bool use_ghost) {
    if (!initialized_)
        throw CUException(__func__, __FILE__, __LINE__,
                          "System has not been initialized. 4B Energy calculation not possible.");
    return run_terms(MBXG_4B, do_grads, use_ghost);
}

std::vector<double> System::GetRealGrads() {
    std::vector<double> out(3 * numat_);
    for (size_t i = 0; i < order_.size(); ++i) {
        const size_t k = order_[i];
        std::copy(grad_.begin() + 3 * first_[i], grad_.begin() + 3 * (first_[i] + in_nat_[k]),
                  out.begin() + 3 * in_first_[k]);
    }
    return out;
}

// ResetOrder3N over all sites: the real atoms' gradients, zeros for the virtual sites (the 1B,
// 2B and 3B terms put no gradient on them)
std::vector<double> System::GetGrads() {
    std::vector<double> out(3 * numsites_, 0.0);
    for (size_t i = 0; i < order_.size(); ++i) {
        const size_t k = order_[i];
        std::copy(grad_.begin() + 3 * first_[i], grad_.begin() + 3 * (first_[i] + in_nat_[k]),
                  out.begin() + 3 * in_first_site_[k]);
    }
    return out;
}

std::vector<double> System::GetVirial() { return virial_; }

std::vector<double> System::GetRealXyz() {
    wrap_xyz();  // SetPBC(box_) (system.cpp:400)
    return real_xyz_input_order();
}

// The stored real-atom coordinates in input order, as they are (ResetOrderReal3N).
std::vector<double> System::real_xyz_input_order() const {
    std::vector<double> out(3 * numat_);
    for (size_t i = 0; i < order_.size(); ++i) {
        const size_t k = order_[i];
        std::copy(xyz_.begin() + 3 * first_[i], xyz_.begin() + 3 * (first_[i] + in_nat_[k]),
                  out.begin() + 3 * in_first_[k]);
    }
    return out;
}

// ResetOrder3N of every site (system.cpp:397): real atoms as GetRealXyz, then each water's M-site
// as SetPBC's SetVSites recomputes it; other virtual sites read 0 (no term here places them).
// GetXyz does not wrap (system.cpp:397 reads xyz_ as the last SetPBC left it).
std::vector<double> System::GetXyz() {
    const std::vector<double> real = real_xyz_input_order();
    std::vector<double> out(3 * numsites_, 0.0);
    for (size_t k = 0; k < in_mon_.size(); ++k) {
        const double* r = real.data() + 3 * in_first_[k];
        double* o = out.data() + 3 * in_first_site_[k];
        std::copy(r, r + 3 * in_nat_[k], o);
        if (in_mon_[k] == "h2o" || in_mon_[k] == "mbpbe")
            for (int j = 0; j < 3; ++j) o[9 + j] = kGamma1 * r[j] + kGamma2 * (r[3 + j] + r[6 + j]);
    }
    return out;
}

void System::SetRealXyz(std::vector<double> xyz) {
    check_initialized(__func__, __LINE__);
    if (xyz.size() != 3 * numat_)
        MBX_THROW("Sizes " + std::to_string(xyz.size()) + " and " + std::to_string(3 * numat_) + " don't match.");
    for (size_t i = 0; i < order_.size(); ++i) {
        const size_t k = order_[i];
        std::copy(xyz.begin() + 3 * in_first_[k], xyz.begin() + 3 * (in_first_[k] + in_nat_[k]),
                  xyz_.begin() + 3 * first_[i]);
    }
}

// system.cpp:812-828: every site's coordinates; the real atoms are kept (virtual sites are
// functions of them, recomputed where needed)
void System::SetXyz(std::vector<double> xyz) {
    check_initialized(__func__, __LINE__);
    if (xyz.size() != 3 * numsites_)
        MBX_THROW("Sizes " + std::to_string(xyz.size()) + " and " + std::to_string(3 * numsites_) + " don't match.");
    for (size_t i = 0; i < order_.size(); ++i) {
        const size_t k = order_[i];
        std::copy(xyz.begin() + 3 * in_first_site_[k], xyz.begin() + 3 * (in_first_site_[k] + in_nat_[k]),
                  xyz_.begin() + 3 * first_[i]);
    }
}

void System::Set2bCutoff(double c) { cutoff2b_ = c; }
void System::Set3bCutoff(double c) { cutoff3b_ = c; }
double System::Get2bCutoff() { return cutoff2b_; }
double System::Get3bCutoff() { return cutoff3b_; }
void System::SetNMaxEval2b(size_t n) { max2b_ = n; }
void System::SetNMaxEval3b(size_t n) { max3b_ = n; }
size_t System::GetMaxEval2b() { return max2b_; }
size_t System::GetMaxEval3b() { return max3b_; }

void System::Set2bIgnorePoly(std::vector<std::vector<std::string>> v) {
    for (auto& p : v) std::sort(p.begin(), p.end());
    ignore2b_ = v;
}

void System::Set3bIgnorePoly(std::vector<std::vector<std::string>> v) {
    for (auto& p : v) std::sort(p.begin(), p.end());
    ignore3b_ = v;
}

void System::Set4bCutoff(double c) { cutoff4b_ = c; }
double System::Get4bCutoff() { return cutoff4b_; }
std::vector<std::vector<std::string>> System::Get4bIgnorePoly() { return ignore4b_; }

void System::Set4bIgnorePoly(std::vector<std::vector<std::string>> v) {
    for (auto& p : v) std::sort(p.begin(), p.end());
    ignore4b_ = v;
}

// system.cpp:981-988: added once, names sorted
void System::Add4bIgnorePoly(std::string mon1, std::string mon2, std::string mon3, std::string mon4) {
    std::vector<std::string> p = {mon1, mon2, mon3, mon4};
    std::sort(p.begin(), p.end());
    if (std::find(ignore4b_.begin(), ignore4b_.end(), p) == ignore4b_.end()) ignore4b_.push_back(p);
}

std::vector<size_t> System::GetPairList(size_t nmax, double cutoff, size_t istart, size_t iend, bool use_ghost) {
    if (nmax != 2 && nmax != 3)
        MBX_THROW("nmax value of " + std::to_string(nmax) + " is not acceptable. Possible values are 2 or 3.");
    check_initialized(__func__, __LINE__);
    mbxg_ctx* c = device();
    check(mbxg_set_use_ghost(c, use_ghost ? 1 : 0));
    check(mbxg_set_cutoffs(c, nmax == 2 ? cutoff : 0.0, nmax == 3 ? cutoff : 0.0));
    check(mbxg_set_coords(c, xyz_.data()));
    check(mbxg_compute(c, (nmax == 2 ? MBXG_2B : MBXG_3B) | MBXG_LISTS_ONLY, 0));
    int64_t n4[4];
    check(mbxg_get_counts(c, n4));
    const int64_t n = nmax == 2 ? n4[0] : n4[2];
    std::vector<int32_t> l((size_t)(nmax * (n > 0 ? n : 1)));
    check(mbxg_get_lists(c, nmax == 2 ? l.data() : nullptr, n, nmax == 3 ? l.data() : nullptr, n));
    check(mbxg_get_coords(c, xyz_.data()));
    std::vector<size_t> out;
    for (int64_t t = 0; t < n; ++t) {
        bool in = false;
        size_t mon[3];
        for (size_t q = 0; q < nmax; ++q) {
            mon[q] = order_[(size_t)l[nmax * t + q]];
            in = in || (mon[q] >= istart && mon[q] < iend);
        }
        if (in)
            for (size_t q = 0; q < nmax; ++q) out.push_back(mon[q]);
    }
    return out;
}

size_t System::GetNumMon() { return in_mon_.size(); }
size_t System::GetNumMol() { return molecules_.size(); }
std::vector<size_t> System::GetMolecule(size_t n) {
    if (n >= molecules_.size()) MBX_THROW("Molecule " + std::to_string(n) + " does not exist");
    return molecules_[n];
}
size_t System::GetNumSites() { return numsites_; }
size_t System::GetNumRealSites() { return numat_; }

std::string System::GetMonId(size_t n) {
    if (n >= in_mon_.size()) MBX_THROW("monomer index out of range");
    return in_mon_[n];
}

size_t System::GetMonNumAt(size_t n) {
    if (n >= in_nat_.size()) MBX_THROW("monomer index out of range");
    return in_nat_[n];
}

size_t System::GetFirstInd(size_t n) {
    if (n >= in_first_.size()) MBX_THROW("monomer index out of range");
    return in_first_[n];
}

std::vector<std::string> System::GetRealAtomNames() { return in_atoms_; }

std::vector<std::string> System::GetAtomNames() {
    std::vector<std::string> out(numsites_, "virt");
    for (size_t k = 0; k < in_mon_.size(); ++k)
        std::copy(in_atoms_.begin() + in_first_[k], in_atoms_.begin() + in_first_[k] + in_nat_[k],
                  out.begin() + in_first_site_[k]);
    return out;
}

std::vector<long> System::GetNmerCounts() { return counts_; }

}  // namespace bblock
