"""Write the reference's own input systems and published per-term energies as compact fixtures.

Runs only where the reference checkout exists; its output (``systems.npz``) is
committed.  The data copied are the reference's test/benchmark inputs
(NRG coordinate files and their mbx.json box/cutoffs) and the energies printed
in the committed ``output`` files -- data, not code.  Nothing of the reference
is imported or executed.

    timing/<case>/{input.nrg, mbx.json, output}
    src/unittests/unittests_inputs/{2048_water.nrg, 2048_water.json}
    src/unittests/unittests_inputs/input_h2o_co2_ch4_dp1_unittest_system.nrg (copied as it is:
        the reference's multi-system NRG test input, for the NRG reader's tests)

Usage:  python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "h2o_256_cubic": "timing/256_h2o_cubic_box",
    "h2o_2048_cubic": "timing/2048_h2o_cubic_box",
    "h2o_2048_droplet": "timing/2048_h2o_droplet",
    "ch4_57_h2o_513_rect": "timing/57_ch4_513_h2o_rectangular",
    "ch4_342_h2o_228_rect": "timing/342_ch4_228_h2o_rectangular",
    "clathrate": "timing/clathrate",
}


def read_nrg(path):
    """NRG reader: returns (monomer ids, atoms per monomer, atom names, xyz[3N])."""
    mons, nat, names, xyz = [], [], [], []
    cur = None
    for line in open(path):
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "MONOMER":
            cur = tok[1]
            mons.append(cur)
            nat.append(0)
        elif tok[0] in ("ENDMON", "ENDMOL", "ENDSYS", "SYSTEM", "MOLECULE"):
            continue
        elif cur is not None and len(tok) == 4:
            names.append(tok[0])
            xyz.extend(float(v) for v in tok[1:])
            nat[-1] += 1
    return mons, nat, names, xyz


def read_json(path):
    txt = open(path).read()
    return json.loads(txt)["MBX"]


def read_output(path):
    out = {}
    for line in open(path):
        m = re.match(r"^(1B|2B|3B|4B|Disp|Elec|Total)\s*=\s*([-+0-9.eE]+)", line.strip())
        if m:
            out[m.group(1)] = float(m.group(2))
    return out


NRG_FILES = {"input_h2o_co2_ch4_dp1.nrg": "src/unittests/unittests_inputs/input_h2o_co2_ch4_dp1_unittest_system.nrg"}


def copy_nrg(ref):
    for name, src in NRG_FILES.items():
        shutil.copyfile(os.path.join(ref, src), os.path.join(HERE, name))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    copy_nrg(ref)
    data = {}
    for key, d in CASES.items():
        mons, nat, names, xyz = read_nrg(os.path.join(ref, d, "input.nrg"))
        cfg = read_json(os.path.join(ref, d, "mbx.json"))
        out = read_output(os.path.join(ref, d, "output"))
        data[key + "/mon"] = np.array(mons)
        data[key + "/nat"] = np.array(nat, np.int32)
        data[key + "/atoms"] = np.array(names)
        data[key + "/xyz"] = np.array(xyz)
        data[key + "/box"] = np.array(cfg.get("box", []), dtype=np.float64)
        data[key + "/cutoffs"] = np.array([cfg["twobody_cutoff"], cfg["threebody_cutoff"]])
        data[key + "/published"] = np.array([out.get("1B", np.nan), out.get("2B", np.nan), out.get("3B", np.nan)])
    mons, nat, names, xyz = read_nrg(os.path.join(ref, "src/unittests/unittests_inputs/2048_water.nrg"))
    cfg = read_json(os.path.join(ref, "src/unittests/unittests_inputs/2048_water.json"))
    data["h2o_2048_unittest/mon"] = np.array(mons)
    data["h2o_2048_unittest/nat"] = np.array(nat, np.int32)
    data["h2o_2048_unittest/atoms"] = np.array(names)
    data["h2o_2048_unittest/xyz"] = np.array(xyz)
    data["h2o_2048_unittest/box"] = np.array(cfg["box"], dtype=np.float64)
    data["h2o_2048_unittest/cutoffs"] = np.array([cfg["twobody_cutoff"], cfg["threebody_cutoff"]])
    data["h2o_2048_unittest/published"] = np.array([np.nan, np.nan, np.nan])
    np.savez_compressed(os.path.join(HERE, "systems.npz"), **data)
    print("wrote", os.path.join(HERE, "systems.npz"))


if __name__ == "__main__":
    main()
