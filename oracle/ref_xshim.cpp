// Test infrastructure only: C entry points over the reference's further MB-nrg holders.
//
// Ours; compiled with the reference's own holder and polynomial sources, read in place from the
// reference checkout by oracle/Makefile, into oracle/_ref/libmbxref_x.so.  Only tests/ load it.
//
// Wrapped reference entry points (paths relative to the reference src/), all E+G with virial:
//   mbnrg_A1_B1C2X2_deg6_v1::eval          potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp:3789 (halide, water)
//   mbnrg_A1_B1C2X2_B1C2X2_deg4_v1::eval   potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp:12955
//   x2b_A1B2_A1B2_v1x::eval                potential/2b/x2b_A1B2_A1B2_v1x.cpp:18485 (co2, co2)
//   mbnrg_A1B2_A1B2_A1B2_deg4_v1::eval     potential/3b/mbnrg_3b_A1B2_A1B2_A1B2_deg4_v1.cpp:1223
//   mbnrg_A1B2_A1B2_A1B2_A1B2_deg4_v1::eval potential/4b/mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.cpp:657
//   h2o_ion::x2b_h2o_ion_v2x::eval         potential/2b/x2b-A1B2Z2_C-v1x.cpp:2546 (water, alkali ion)
//   x3b_h2o_ion_v1x_deg4_filtered::operator() potential/3b/x3b-h2o-ion-v1x-deg4-filtered.cpp:206
//                                          (water, water, alkali ion)
// Each takes n n-mers, the monomers' coordinates in the holder's argument order (the ion with 3
// doubles, a water or CO2 with 9), returns the summed energy and adds the gradients (+=) and the
// virial (+=, 9).
#include <string>
#include <vector>

#include "potential/1b/x1b_A1B2_deg4_v1x.h"
#include "potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.h"
#include "potential/2b/x2b-A1B2Z2_C-v1x.h"
#include "potential/3b/x3b-h2o-ion-v1x-deg4-filtered.h"
#include "potential/2b/x2b_A1B2_A1B2_v1x.h"
#include "potential/3b/mbnrg_3b_A1B2_A1B2_A1B2_deg4_v1.h"
#include "potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.h"
#include "potential/4b/mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.h"

namespace {
void add_virial(const std::vector<double>& v, double* out) {
    for (int i = 0; i < 9; ++i) out[i] += v[i];
}
}  // namespace

extern "C" {

double ref_x2b_halide_h2o(const char* ion, const double* xi, const double* xw, size_t n, double* gi, double* gw,
                          double* vir) {
    mbnrg_A1_B1C2X2_deg6::mbnrg_A1_B1C2X2_deg6_v1 pot(ion, "h2o");
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(xi, xw, gi, gw, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x3b_halide_h2o2(const char* ion, const double* xi, const double* xw1, const double* xw2, size_t n,
                           double* gi, double* g1, double* g2, double* vir) {
    mbnrg_A1_B1C2X2_B1C2X2_deg4::mbnrg_A1_B1C2X2_B1C2X2_deg4_v1 pot(ion, "h2o", "h2o");
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(xi, xw1, xw2, gi, g1, g2, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x2b_co2(const double* x1, const double* x2, size_t n, double* g1, double* g2, double* vir) {
    x2b_A1B2_A1B2_deg5::x2b_A1B2_A1B2_v1x pot("co2", "co2");
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(x1, x2, g1, g2, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x3b_co2(const double* x1, const double* x2, const double* x3, size_t n, double* g1, double* g2, double* g3,
                   double* vir) {
    mbnrg_A1B2_A1B2_A1B2_deg4::mbnrg_A1B2_A1B2_A1B2_deg4_v1 pot("co2", "co2", "co2");
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(x1, x2, x3, g1, g2, g3, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x4b_h2o(const double* x1, const double* x2, const double* x3, const double* x4, size_t n, double* g1,
                   double* g2, double* g3, double* g4, double* vir) {
    mbnrg_A1B2_A1B2_A1B2_A1B2_deg4::mbnrg_A1B2_A1B2_A1B2_A1B2_deg4_v1 pot("h2o", "h2o", "h2o", "h2o");
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(x1, x2, x3, x4, g1, g2, g3, g4, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x2b_alkali_h2o(const char* ion, const double* xw, const double* xi, size_t n, double* gw, double* gi,
                          double* vir) {
    h2o_ion::x2b_h2o_ion_v2x pot("h2o", ion);
    std::vector<double> v(9, 0.0);
    const double e = pot.eval(xw, xi, gw, gi, n, &v);
    add_virial(v, vir);
    return e;
}

double ref_x3b_alkali_h2o2(const char* ion, const double* xw1, const double* xw2, const double* xi, size_t n,
                           double* gw1, double* gw2, double* gi, double* vir) {
    x3b_h2o_ion_v1x_deg4_filtered pot(ion);
    std::vector<double> v(9, 0.0);
    const double e = pot(xw1, xw2, xi, gw1, gw2, gi, n, &v);
    add_virial(v, vir);
    return e;
}

// x1b_A1B2_v1x("co2")::eval with gradients and virial (potential/1b/x1b_A1B2_deg4_v1x.cpp:160):
// energies[n], grad[9 n] (overwritten), virial[9] (accumulated).
void ref_co2_1b(const double* xyz, double* grad, size_t n, double* energies, double* virial) {
    x1b_A1B2_deg4::x1b_A1B2_v1x pot("co2");
    std::vector<double> v(virial, virial + 9);
    std::vector<double> e = pot.eval(xyz, grad, n, &v);
    std::copy(e.begin(), e.end(), energies);
    std::copy(v.begin(), v.end(), virial);
}

}  // extern "C"
