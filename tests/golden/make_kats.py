"""Copy the reference's known-answer values for the MB-pol 2B/3B path into JSON fixtures.

Parses the reference's Catch2 test sources as TEXT and keeps the numeric data of the named
TEST_CASE blocks: every ``std::vector<double> NAME = {...};`` and ``double NAME = value;``.
Nothing is compiled, imported or executed.  Output: tests/golden/kats.json.

    poly_2b    src/unittests/unittest-poly-2b.cpp         TEST_CASE(":poly_2b_v6x:struct")
    pair_list  src/unittests/unittest-system.cpp          TEST_CASE("Pair List") (expected n-mer sets)
    system_h2o_2_ch4_1  src/unittests/setup_h2o_2_ch4_1.h  (System-level 1B/2B/3B energies, gradients,
                                                            virials of two waters and a methane)
    poly_3b    src/unittests/unittest-poly-3b.cpp         TEST_CASE("poly-3b-v2x::class")
    holder_2b  src/unittests/unittest-poly-holder-2b.cpp  TEST_CASE("x2b_v9x::struct")
    holder_3b  src/unittests/unittest-poly-holder-3b.cpp  TEST_CASE("x3b-v2x::class")
    and the MB-nrg CH4 families' polynomial, holder and energy-routine cases, the alkali-ion-water
    holder, polynomial and energy-routine cases and the Na+ / Cs+ water setups' per-term energies
    (src/tests/unittests/setup_h2o_{1,2}_{na,cs}_1.h), the MB-pol
    1B (ps::pot_nasa) case, the CH4 and CO2 1B polynomial and holder cases (see CASES), and the
    ch4, h2o and co2 rows of unittest-energy1b.cpp, the h2o sections of unittest-energy2b.cpp /
    unittest-energy3b.cpp, and the three-water gas-phase n-body energies of
    src/tests/unittests/unittest-gas-nbodyterms-mbpol.cpp.

Usage:  python tests/golden/make_kats.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "poly_2b": ("src/unittests/unittest-poly-2b.cpp", ':poly_2b_v6x:struct'),
    "poly_3b": ("src/unittests/unittest-poly-3b.cpp", 'poly-3b-v2x::class'),
    "holder_2b": ("src/unittests/unittest-poly-holder-2b.cpp", 'x2b_v9x::struct'),
    "holder_3b": ("src/unittests/unittest-poly-holder-3b.cpp", 'x3b-v2x::class'),
    # MB-nrg CH4 families: polynomials, holders, and the dispatching energy routines
    "poly_ch4h2o_2b": ("src/unittests/unittest-poly-2b.cpp", ':poly_2b_A1B2Z2_C1D4_deg3_exp0_v1x:struct'),
    "poly_ch4ch4_2b": ("src/unittests/unittest-poly-2b.cpp", 'poly_2b_A1B4_A1B4_deg4_exp0_v1x:struct'),
    "poly_ch4h2o2_3b": ("src/unittests/unittest-poly-3b.cpp", 'poly_3b_A1B4_C1D2_C1D2_deg3_v1.h'),
    "holder_ch4h2o_2b": ("src/unittests/unittest-poly-holder-2b.cpp", 'x2b_A1B2Z2_C1D4_v1x::struct'),
    "holder_ch4ch4_2b": ("src/unittests/unittest-poly-holder-2b.cpp", 'x2b_A1B4_A1B4_v1x::struct'),
    "holder_ch4h2o2_3b": ("src/unittests/unittest-poly-holder-3b.cpp", 'mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1::class'),
    "energy_ch4ch4_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:ch4-ch4'),
    "energy_h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:h2o-h2o'),
    "energy_h2o_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:h2o-h2o-h2o'),
    "energy_ch4h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:ch4-h2o'),
    "energy_ch4h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:ch4-h2o-h2o'),
    # MB-pol 1B (Partridge-Schwenke water monomer)
    "ps_1b": ("src/unittests/unittest-ps.cpp", 'ps::energy'),
    # MB-nrg CH4 1B: polynomial and holder
    "poly_ch4_1b": ("src/unittests/unittest-poly-1b.cpp", 'poly_1b_A1B4_deg5_exp0_v1x::struct'),
    "holder_ch4_1b": ("src/unittests/unittest-poly-holder-1b.cpp", 'x1b_A1B4_deg5_exp0::struct'),
    # MB-nrg CO2 1B: polynomial and holder
    "poly_co2_1b": ("src/unittests/unittest-poly-1b.cpp", 'poly_1b_A1B2_deg4_v1x::struct'),
    "holder_co2_1b": ("src/unittests/unittest-poly-holder-1b.cpp", 'x1b_A1B2_deg4_v1x::struct'),
    # halide-water 2B / 3B, CO2 2B / 3B and water 4B energy routines ("#n": the n-th section of
    # that name; the first co2-co2 section uses the co2_archive model, the first co2-co2-co2 one is
    # commented out)
    "energy_f_h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:f--h2o'),
    "energy_cl_h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:cl-h2o'),
    "energy_br_h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:br-h2o'),
    "energy_i_h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:i-h2o'),
    "energy_co2co2_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:co2-co2#2'),
    "energy_f_h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:f-h2o-h2o'),
    "energy_cl_h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:cl-h2o-h2o'),
    "energy_br_h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:br-h2o-h2o'),
    "energy_i_h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:i-h2o-h2o'),
    "energy_co2x3_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:co2-co2-co2#2'),
    "energy_h2ox4_4b": ("src/unittests/unittest-energy4b.cpp", 'SECTION:h2o-h2o-h2o-h2o'),
    "poly_co2co2_2b": ("src/unittests/unittest-poly-2b.cpp", ':poly_2b_A1B2_A1B2_v1x:struct'),
}


# Alkali-metal-ion-water families: x2b_h2o_ion_v2x and x3b_h2o_ion_v1x_deg4_filtered holder cases
# (every ion), the 3B polynomial case, and the li+ / cs+ sections of the 3B energy routine.
for _ion in ("li", "na", "k", "rb", "cs"):
    CASES["holder_%s_h2o_2b" % _ion] = ("src/unittests/unittest-poly-holder-2b.cpp", "SECTION:h2o-%s" % _ion)
    CASES["holder_%s_h2o2_3b" % _ion] = ("src/unittests/unittest-poly-holder-3b.cpp", "SECTION:h2o-h2o-%s" % _ion)
CASES["poly_alkali_h2o2_3b"] = ("src/unittests/unittest-poly-3b.cpp", "poly-3b-h2o-ion-v1x_deg4_filtered::class")
CASES["energy_li_h2o2_3b"] = ("src/unittests/unittest-energy3b.cpp", "SECTION:li-h2o-h2o")
CASES["energy_cs_h2o2_3b"] = ("src/unittests/unittest-energy3b.cpp", "SECTION:cs-h2o-h2o")

# System-level data of the reference's ion-water setups (src/tests/unittests/setup_h2o_<n>_<ion>_1.h):
# per-term energies and virials of one or two waters with one Na+ or Cs+.
SETUPS = ["setup_h2o_1_na_1", "setup_h2o_2_na_1", "setup_h2o_1_cs_1", "setup_h2o_2_cs_1"]


def block(text, name):
    if name.startswith("SECTION:"):
        sec, _, nth = name[len("SECTION:"):].partition("#")
        i = -1
        for _ in range(int(nth or 1)):
            i = text.index('SECTION("%s")' % sec, i + 1)
    else:
        i = text.index('TEST_CASE("%s")' % name)
    j = text.index("{", i) + 1
    depth = 1
    k = j
    while depth:
        depth += {"{": 1, "}": -1}.get(text[k], 0)
        k += 1
    return text[j:k - 1]


def values(body):
    body = re.sub(r"//[^\n]*", "", body)
    out = {}
    for m in re.finditer(r"std::vector<double>\s+(\w+)\s*=\s*\{([^}]*)\}\s*;", body):
        out[m.group(1)] = [float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
    for m in re.finditer(r"(?<![\w>])double\s+(\w+)\s*=\s*([-+0-9.eE]+)\s*;", body):
        out[m.group(1)] = float(m.group(2))
    for m in re.finditer(r"size_t\s+(\w+)\s*=\s*(\d+)\s*;", body):
        out[m.group(1)] = int(m.group(2))
    for m in re.finditer(r"std::string\s+(mon\d)\s*=\s*\"([^\"]*)\"\s*;", body):
        out[m.group(1)] = m.group(2)
    return out


def gas_nbody(test_text, setup_text):
    """src/tests/unittests/unittest-gas-nbodyterms-mbpol.cpp over setup_h2o_3.h: the three waters'
    real coordinates (O H H of each monomer, the setup's coords at first_ind = 4 n) and the 1B / 2B
    polynomial / 3B polynomial energies the test expects (its REQUIRE ... Approx values)."""
    src = re.sub(r"\\\n", "\n", setup_text)
    m = re.search(r"std::vector<double>\s+coords\s*\{([^}]*)\}", src)
    coords = [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    assert re.search(r"first_ind\[n\]\s*=\s*4\s*\*\s*n", src)
    xyz = []
    for n in range(3):
        xyz += coords[12 * n:12 * n + 9]
    out = {"xyz": xyz}
    for key, sec in (("e1b", "One body (polynomials)"), ("e2b", "Two body (polynomials)"),
                     ("e3b", "Three body (polynomials)")):
        b = block(test_text, "SECTION:" + sec)
        out[key] = float(re.search(r"Approx\(([-+0-9.eE]+)\)", b).group(1))
    return out


def energy1b(text, mons=("ch4", "h2o", "co2")):
    """unittest-energy1b.cpp keeps one row per monomer kind in nested vectors (xyz, grad, virial,
    grad_expected, virial_expected), indexed by monid, plus nmon and energies_expected: the rows
    of the named kinds."""
    src = re.sub(r"//[^\n]*", "", text)

    def vec(name, nested):
        m = re.search(r"std::vector<(?:std::vector<double> ?|double|size_t|std::string)> ?%s\s*=\s*\{" % name, src)
        i = m.end() - 1
        depth, j = 0, i
        while True:
            depth += {"{": 1, "}": -1}.get(src[j], 0)
            if depth == 0:
                break
            j += 1
        body = src[i + 1:j]
        if not nested:
            return [v.strip().strip('"') for v in body.split(",") if v.strip()]
        return [[float(v) for v in row.split(",") if v.strip()] for row in re.findall(r"\{([^{}]*)\}", body)]

    monid = vec("monid", False)
    nmon = [int(v) for v in vec("nmon", False)]
    energies = [float(v) for v in vec("energies_expected", False)]
    rows = {n: vec(n, True) for n in ("xyz", "grad", "virial", "grad_expected", "virial_expected")}
    out = {}
    for mon in mons:
        k = monid.index(mon)
        out[mon] = {"nmon": nmon[k], "energy_expected": energies[k], **{n: r[k] for n, r in rows.items()}}
    return out


def pair_list(text):
    """unittest-system.cpp TEST_CASE("Pair List"): the expected dimer / trimer sets of
    GetPairList at 10 and 5 Angstrom (input monomer indices, in the order written)."""
    b = re.sub(r"//[^\n]*", "", block(re.sub(r"//[^\n]*", "", text), "Pair List"))
    out = {}
    for name in ("expected_d", "expected_t", "expected_d2", "expected_t2"):
        m = re.search(r"std::vector<size_t>\s+%s\s*=\s*\{([^}]*)\}" % name, b)
        out[name] = [int(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    out["cutoffs"] = [float(v) for v in re.findall(r"GetPairList\(2,\s*([0-9.]+),", b)]
    out["input"] = re.search(r'ReadNrg\("([^"]+)"', b).group(1)
    return out


def test_data(text):
    """The TestEnergyData of a setup_*.h macro (unittests/setup_h2o_2_ch4_1.h): numbers, vectors
    and strings of every testData.<field> assignment."""
    src = re.sub(r"\\\n", "\n", text)
    out = {}
    for m in re.finditer(r"testData\.(\w+)\s*=\s*(\{[^}]*\}|R\"\((.*?)\)\"|[-+0-9.eE]+)\s*;", src, re.S):
        name, val = m.group(1), m.group(2)
        if m.group(3) is not None:
            out[name] = m.group(3)
        elif val.startswith("{"):
            items = [v.strip() for v in val[1:-1].replace("\n", " ").split(",") if v.strip()]
            out[name] = [v.strip('"') if v.startswith('"') else float(v) for v in items]
        else:
            out[name] = float(val)
    return out


def setup_macro(text):
    """The variables a setup_*.h macro declares (src/tests/unittests/setup_h2o_2_na_1.h): numbers,
    double vectors (with or without '='), and the monomer names."""
    src = re.sub(r"//[^\n]*", "", re.sub(r"\\\n", "\n", text))
    out = {}
    for m in re.finditer(r"std::vector<double>\s+(\w+)\s*=?\s*\{([^}]*)\}\s*;", src):
        out[m.group(1)] = [float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
    for m in re.finditer(r"(?<![\w>])double\s+(\w+)\s*=\s*([-+0-9.eE]+)\s*;", src):
        out[m.group(1)] = float(m.group(2))
    m = re.search(r"std::vector<std::string>\s+monomer_names\s*=\s*\{([^}]*)\}", src)
    out["monomer_names"] = [v.strip().strip('"') for v in m.group(1).split(",") if v.strip()]
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    res = {}
    for key, (path, name) in CASES.items():
        res[key] = values(block(open(os.path.join(ref, path)).read(), name))
        res[key]["_source"] = ('%s SECTION("%s")' % (path, name[8:].split("#")[0]) if name.startswith("SECTION:")
                               else '%s TEST_CASE("%s")' % (path, name))
    path = "src/unittests/unittest-energy1b.cpp"
    res["energy1b"] = energy1b(open(os.path.join(ref, path)).read())
    res["energy1b"]["_source"] = path + ' TEST_CASE("energy1b::get_1b_energy_grad"), rows ch4, h2o and co2'
    path = "src/tests/unittests/unittest-gas-nbodyterms-mbpol.cpp"
    res["gas_nbody_h2o_3"] = gas_nbody(open(os.path.join(ref, path)).read(),
                                       open(os.path.join(ref, "src/tests/unittests/setup_h2o_3.h")).read())
    res["gas_nbody_h2o_3"]["_source"] = path + " (over setup_h2o_3.h)"
    path = "src/unittests/unittest-system.cpp"
    res["pair_list"] = pair_list(open(os.path.join(ref, path)).read())
    res["pair_list"]["_source"] = path + ' TEST_CASE("Pair List")'
    path = "src/unittests/setup_h2o_2_ch4_1.h"
    res["system_h2o_2_ch4_1"] = test_data(open(os.path.join(ref, path)).read())
    res["system_h2o_2_ch4_1"]["_source"] = path + " (TestEnergyData of unittest-system.cpp:215-240)"
    for name in SETUPS:
        path = "src/tests/unittests/%s.h" % name
        d = setup_macro(open(os.path.join(ref, path)).read())
        keep = ("one_body_energy", "two_body_energy", "three_body_energy", "real_coords", "monomer_names",
                "virial_1b", "virial_2b", "virial_3b")
        res[name] = {k: d[k] for k in keep if k in d}  # setup_h2o_1_na_1.h has no virial_3b
        res[name]["_source"] = path
    
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(res, f)
    for k, v in res.items():
        print(k, sorted(x for x in v if not x.startswith("_")))


if __name__ == "__main__":
    main()
