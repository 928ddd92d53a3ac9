"""CPU tests of the NRG reader (include/mbx/read_nrg.h: tools::ParseNrg / ReadNrg, the grammar of
the reference's src/io_tools/read_nrg.cpp:43-298), through `single_point --dump` (the parser
alone; no device).

Pins: the reference's own multi-system test input (unittests_inputs/
input_h2o_co2_ch4_dp1_unittest_system.nrg, copied to tests/golden) against an independent
tokenisation of the same text; the reference's inputs (systems.npz) written out and read back bit
for bit, in lower- and upper-case keywords, with several monomers per molecule and several
systems per file; each error message of the reference, with its line number."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import nrg_files


def _tokens(path):
    """Systems of an NRG file by plain tokenisation (each MONOMER line opens a monomer, each
    coordinate line is name + 3 floats)."""
    systems = []
    for line in open(path):
        w = line.split()
        if not w:
            continue
        k = w[0].lower()
        if k == "system":
            systems.append({"monomers": [], "molecules": []})
        elif k == "molecule":
            systems[-1]["molecules"].append([])
        elif k == "monomer":
            systems[-1]["molecules"][-1].append(len(systems[-1]["monomers"]))
            systems[-1]["monomers"].append((w[1].lower(), [], []))
        elif k not in ("endmon", "endmol", "endsys"):
            systems[-1]["monomers"][-1][1].append(w[0])
            systems[-1]["monomers"][-1][2].append([float(v) for v in w[1:4]])
    return systems


def _same(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g["molecules"] == w["molecules"]
        assert len(g["monomers"]) == len(w["monomers"])
        for (gi, gn, gx), (wi, wn, wx) in zip(g["monomers"], w["monomers"]):
            assert gi == wi and gn == wn
            assert np.array_equal(np.asarray(gx).reshape(-1, 3), np.asarray(wx, np.float64).reshape(-1, 3))


def test_reference_unittest_input():
    path = os.path.join(GOLDEN, "input_h2o_co2_ch4_dp1.nrg")
    got = nrg_files.dump(path)
    assert len(got) == 3  # the reference's test: complete, without point charges, point charges only
    assert [len(s["monomers"]) for s in got] == [10, 8, 2]  # counted with awk on the file
    _same(got, _tokens(path))
    assert sorted({m[0] for s in got for m in s["monomers"]}) == ["ch4", "co2", "dp1", "h2o"]


@pytest.mark.parametrize("case", ["h2o_256_cubic", "clathrate"])
@pytest.mark.parametrize("per_molecule,case_fn", [(1, str.upper), (3, str.lower)])
def test_roundtrip(systems, tmp_path, case, per_molecule, case_fn):
    mons = [str(m) for m in systems[case + "/mon"]]
    nat = systems[case + "/nat"]
    names = [str(a) for a in systems[case + "/atoms"]]
    xyz = systems[case + "/xyz"]
    half = len(mons) // 2
    cut = int(np.sum(nat[:half]))
    parts = [(mons[:half], nat[:half], names[:cut], xyz[:3 * cut]), (mons[half:], nat[half:], names[cut:], xyz[3 * cut:])]
    path = str(tmp_path / "in.nrg")
    nrg_files.write_nrg(path, parts, per_molecule=per_molecule, keyword_case=case_fn)
    got = nrg_files.dump(path)
    assert len(got) == 2
    for g, (m, n, a, x) in zip(got, parts):
        assert [v[0] for v in g["monomers"]] == m
        assert [len(v[1]) for v in g["monomers"]] == list(n)
        assert sum((v[1] for v in g["monomers"]), []) == a
        assert np.array_equal(np.concatenate([v[2] for v in g["monomers"]]).reshape(-1), x)
        assert g["molecules"] == [list(range(i, min(i + per_molecule, len(m)))) for i in range(0, len(m), per_molecule)]


GOOD = ["SYSTEM", "MOLECULE", "MONOMER h2o", "O 0 0 0", "H 0.9572 0 0", "H -0.24 0.927 0", "ENDMON", "ENDMOL", "ENDSYS"]


@pytest.mark.parametrize("edit,message,line", [
    ({0: "MOLECULE"}, "Expected SYSTEM keyword", 1),
    ({1: "MONOMER h2o"}, "Expected MOLECULE keyword", 2),
    ({2: "MONOMER"}, "along with the monomer id", 3),
    ({3: "O"}, "Expected coordinates or ENDMON", 4),
    ({4: "H 0.9572 x 0"}, "Expected numerical coordinates", 5),
    ({5: "H 0.24 0.927"}, "Expected numerical coordinates", 6),
    ({7: "garbage"}, "along with the monomer id", 8),
    ({8: "garbage"}, "Expected MOLECULE keyword", 9),
])
def test_errors(tmp_path, edit, message, line):
    lines = [edit.get(i, l) for i, l in enumerate(GOOD)]
    path = str(tmp_path / "bad.nrg")
    open(path, "w").write("\n".join(lines) + "\n")
    r = nrg_files.run("--dump", path)
    assert r.returncode == 1
    assert message in r.stderr and ("line %d of the NRG file" % line) in r.stderr, r.stderr


def test_truncated_and_trailing(tmp_path):
    path = str(tmp_path / "cut.nrg")
    open(path, "w").write("\n".join(GOOD[:6]) + "\n")
    r = nrg_files.run("--dump", path)
    assert r.returncode == 1 and "Expected coordinates or ENDMON" in r.stderr
    # an empty line after ENDSYS ends the list of systems: what follows is not read
    open(path, "w").write("\n".join(GOOD + [""] + GOOD[:3]) + "\n")
    assert len(nrg_files.dump(path)) == 1
    # SYSTEM blocks back to back are all read
    open(path, "w").write("\n".join(GOOD + GOOD) + "\n")
    assert len(nrg_files.dump(path)) == 2


def test_missing_file(tmp_path):
    r = nrg_files.run(str(tmp_path / "none.nrg"))
    assert r.returncode == 1 and "Could not open NRG file" in r.stderr
