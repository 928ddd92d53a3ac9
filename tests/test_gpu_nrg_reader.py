"""GPU: an NRG file through the reference's single-point flow (tools::ReadNrg -> SetUpFromJson ->
Energy(true), OneBodyEnergy / TwoBodyEnergy / ThreeBodyEnergy; src/tests/single_point.cpp) with
the prebuilt mbx_amd/single_point.  The reference's inputs (systems.npz) are written out as NRG
and mbx.json files; the printed energies must equal the reference's published per-term energies
(11 significant digits) and the gradients the oracle's (1e-8 kcal/mol/A)."""
import numpy as np
import pytest

import nrg_files
import oracle

pytestmark = pytest.mark.gpu


def _files(systems, tmp_path, case):
    mons = [str(m) for m in systems[case + "/mon"]]
    nat = systems[case + "/nat"]
    names = [str(a) for a in systems[case + "/atoms"]]
    xyz = systems[case + "/xyz"]
    c2, c3 = (float(v) for v in systems[case + "/cutoffs"])
    nrg, js = str(tmp_path / "input.nrg"), str(tmp_path / "mbx.json")
    nrg_files.write_nrg(nrg, [(mons, nat, names, xyz)])
    nrg_files.write_json(js, systems[case + "/box"], c2, c3)
    return nrg, js, mons, xyz, c2, c3


def test_single_point_water_published(systems, tmp_path):
    nrg, js, _, _, _, _ = _files(systems, tmp_path, "h2o_256_cubic")
    print("single_point h2o_256_cubic", flush=True)
    r = nrg_files.run(nrg, js)
    assert r.returncode == 0, r.stderr
    e, _ = nrg_files.energies(r.stdout)
    pub = systems["h2o_256_cubic/published"]
    assert e["1B"] == pytest.approx(pub[0], rel=5e-11)
    assert e["2B"] == pytest.approx(pub[1], rel=5e-11)
    assert e["3B"] == pytest.approx(pub[2], rel=5e-11)
    # Energy(true): one combined compute of the three terms vs three separate ones
    assert e["total"] == pytest.approx(e["1B"] + e["2B"] + e["3B"], rel=1e-12)


def test_single_point_clathrate_grads(systems, tmp_path):
    nrg, js, mons, xyz, c2, c3 = _files(systems, tmp_path, "clathrate")
    print("single_point clathrate", flush=True)
    r = nrg_files.run(nrg, js, "--grads")
    assert r.returncode == 0, r.stderr
    e, g = nrg_files.energies(r.stdout)
    pub = systems["clathrate/published"]
    # the CH4 1B polynomial cancels terms of ~2e7 per monomer: its FP64 floor is ~1e-8 here
    assert e["1B"] == pytest.approx(pub[0], abs=2e-8)
    assert e["2B"] == pytest.approx(pub[1], rel=5e-11)
    assert e["3B"] == pytest.approx(pub[2], rel=5e-11)
    kinds = np.array([0 if m == "h2o" else 1 for m in mons])
    box = systems["clathrate/box"]
    box = box if box.size else None
    ref = oracle.system(xyz, kinds, box, c2, c3)
    _, g1, _ = oracle.one_body(xyz, kinds, box)
    assert g.shape == (len(xyz) // 3, 3)
    assert np.abs(g.reshape(-1) - (ref["grad"] + g1)).max() <= 1e-8


def test_write_nrg_roundtrip(tmp_path):
    """tools::WriteNrg (through single_point --write) on the reference's multi-system test input:
    the file written reads back as the same systems, molecules, monomer ids and atom names, and
    the coordinates to the 8 digits after the point the reference's format keeps."""
    from conftest import GOLDEN
    import os
    src = os.path.join(GOLDEN, "input_h2o_co2_ch4_dp1.nrg")
    out = str(tmp_path / "out.nrg")
    print("write_nrg roundtrip", flush=True)
    r = nrg_files.run(src, "--write", out)
    # the file is written right after reading; the energy of system 0 then fails loudly on its dp1
    # monomers (no holder for them here)
    assert r.returncode == 1 and "has no GPU holder" in r.stderr, r.stderr
    a, b = nrg_files.dump(src), nrg_files.dump(out)
    assert len(a) == len(b) == 3
    for sa, sb in zip(a, b):
        assert sa["molecules"] == sb["molecules"]
        assert [(m[0], m[1]) for m in sa["monomers"]] == [(m[0], m[1]) for m in sb["monomers"]]
        xa = np.concatenate([m[2] for m in sa["monomers"]])
        xb = np.concatenate([m[2] for m in sb["monomers"]])
        np.testing.assert_allclose(xb, xa, rtol=1e-8, atol=0)
    assert open(out).readline().strip() == "SYSTEM 0"
