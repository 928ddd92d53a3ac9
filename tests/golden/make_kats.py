"""Copy the reference's known-answer values for the MB-pol 2B/3B path into JSON fixtures.

Parses the reference's Catch2 test sources as TEXT and keeps the numeric data of the named
TEST_CASE blocks: every ``std::vector<double> NAME = {...};`` and ``double NAME = value;``.
Nothing is compiled, imported or executed.  Output: tests/golden/kats.json.

    poly_2b    src/unittests/unittest-poly-2b.cpp         TEST_CASE(":poly_2b_v6x:struct")
    poly_3b    src/unittests/unittest-poly-3b.cpp         TEST_CASE("poly-3b-v2x::class")
    holder_2b  src/unittests/unittest-poly-holder-2b.cpp  TEST_CASE("x2b_v9x::struct")
    holder_3b  src/unittests/unittest-poly-holder-3b.cpp  TEST_CASE("x3b-v2x::class")
    and the MB-nrg CH4 families' polynomial, holder and energy-routine cases, and the MB-pol
    1B (ps::pot_nasa) case (see CASES).

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
    "energy_ch4h2o_2b": ("src/unittests/unittest-energy2b.cpp", 'SECTION:ch4-h2o'),
    "energy_ch4h2o2_3b": ("src/unittests/unittest-energy3b.cpp", 'SECTION:ch4-h2o-h2o'),
    # MB-pol 1B (Partridge-Schwenke water monomer)
    "ps_1b": ("src/unittests/unittest-ps.cpp", 'ps::energy'),
}


def block(text, name):
    if name.startswith("SECTION:"):
        i = text.index('SECTION("%s")' % name[len("SECTION:"):])
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
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    res = {}
    for key, (path, name) in CASES.items():
        res[key] = values(block(open(os.path.join(ref, path)).read(), name))
        res[key]["_source"] = ('%s SECTION("%s")' % (path, name[8:]) if name.startswith("SECTION:")
                               else '%s TEST_CASE("%s")' % (path, name))
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(res, f)
    for k, v in res.items():
        print(k, sorted(x for x in v if not x.startswith("_")))


if __name__ == "__main__":
    main()
