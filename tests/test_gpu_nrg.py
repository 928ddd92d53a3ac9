"""GPU parity of the MB-nrg CH4 families (CH4-H2O 2B, CH4-CH4 2B, CH4-H2O-H2O 3B) mixed with
MB-pol, against the oracle (orc_system) and the reference's published per-term energies.

Bar: lists bit-exact as sets; per-family evaluated counts equal; E2B / E3B within 1e-10
relative; gradients within 1e-8 kcal/mol/A; virial within 1e-8."""
import numpy as np
import pytest

from conftest import sorted_tuples

import oracle
from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu

E_REL = 1e-10
G_ABS = 1e-8
MIXED = ["ch4_57_h2o_513_rect", "ch4_342_h2o_228_rect", "clathrate"]
# the reference runs' own ignore lists (timing/<case>/mbx.json)
PUBLISHED = {
    "ch4_57_h2o_513_rect": ("mbpol_2b", "ch4h2o_2b", "ch4ch4_2b", "mbpol_3b"),
    "ch4_342_h2o_228_rect": ("mbpol_2b", "mbpol_3b"),
    "clathrate": engine.FAMILIES,
}


def mixed_case(systems, name):
    kinds = np.array([0 if m == "h2o" else 1 for m in systems[name + "/mon"]])
    box = systems[name + "/box"]
    c2, c3 = (float(v) for v in systems[name + "/cutoffs"])
    return systems[name + "/xyz"], kinds, (box if box.size else None), c2, c3, systems[name + "/published"]


def mixed_engine(xyz, kinds, box, c2, c3, families=engine.FAMILIES):
    e = engine.Engine()
    e.set_topology(kinds)  # tags: each monomer's index (the oracle gets the same below)
    e.set_box(box)
    e.set_cutoffs(c2, c3)
    e.set_families(families)
    e.set_coords(xyz)
    e.compute()
    return e


@pytest.mark.parametrize("name", MIXED)
def test_mixed_all_families(systems, name):
    xyz, kinds, box, c2, c3, _ = mixed_case(systems, name)
    print("mixed", name, flush=True)
    e = mixed_engine(xyz, kinds, box, c2, c3)
    e2, e3, g, v = e.results()
    r = oracle.system(xyz, kinds, box, c2, c3, tags=np.arange(len(kinds)))
    c = e.counts()
    assert (c[0], c[1], c[2], c[3]) == (r["n2"], r["n2_eval"], r["n3"], r["n3_eval"])
    assert e2 == pytest.approx(r["e2b"], rel=E_REL)
    assert e3 == pytest.approx(r["e3b"], rel=E_REL)
    assert np.abs(g - r["grad"]).max() <= G_ABS
    np.testing.assert_allclose(v, r["virial"], rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("name", MIXED)
def test_mixed_published(systems, name):
    """The reference's printed 2B / 3B energies (11 significant digits) with its ignore lists."""
    xyz, kinds, box, c2, c3, pub = mixed_case(systems, name)
    e = mixed_engine(xyz, kinds, box, c2, c3, PUBLISHED[name])
    e2, e3, _, _ = e.results()
    assert e2 == pytest.approx(pub[1], rel=5e-11)
    assert e3 == pytest.approx(pub[2], rel=5e-11)


def test_mixed_lists(systems):
    xyz, kinds, box, c2, c3, _ = mixed_case(systems, "ch4_342_h2o_228_rect")
    e = engine.Engine()
    e.set_topology(kinds)
    e.set_box(box)
    e.set_cutoffs(c2, c3)
    e.set_coords(xyz)
    e.compute(engine.TERM_2B | engine.TERM_3B | engine.LISTS_ONLY)
    d, t = e.lists()
    nat = np.array([3 if k == 0 else 5 for k in kinds])
    fixed = oracle.fix_monomer_coordinates(xyz, box, nat)
    first = fixed.reshape(-1, 3)[np.concatenate([[0], np.cumsum(nat)[:-1]])]
    assert np.array_equal(sorted_tuples(d), sorted_tuples(oracle.clusters(2, c2, first, box)))
    assert np.array_equal(sorted_tuples(t), sorted_tuples(oracle.clusters(3, c3, first, box)))


def _methane(rng, center):
    """CH4 with C-H 1.09 A, ideal tetrahedron, random orientation."""
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) / np.sqrt(3.0) * 1.09
    rot = configs.random_rotations(rng, 1)[0]
    return np.concatenate([center, (tet @ rot.T + center).reshape(-1)])


def test_mixed_gas_cluster_and_energy_only():
    """A gas-phase CH4/H2O cluster at close range (every family evaluated, switching regions
    crossed), with gradients and energy only."""
    rng = np.random.default_rng(21)
    pts = configs.water_box(40, 14.0, (4, 4, 3), seed=4)[0].reshape(40, 3, 3)[:, 0]
    kinds = np.array([1 if i % 4 == 0 else 0 for i in range(40)])
    parts = []
    for i, k in enumerate(kinds):
        if k == 1:
            parts.append(_methane(rng, pts[i]))
        else:
            parts.append(configs.place_waters(rng, pts[i:i + 1]))
    xyz = np.concatenate(parts)
    e = mixed_engine(xyz, kinds, None, 9.0, 6.5)
    r = oracle.system(xyz, kinds, None, 9.0, 6.5, tags=np.arange(len(kinds)))
    # a small cluster's totals sit far below their cancelling terms: 1e-10 relative, or 1e-16 of
    # the terms' magnitudes (oracle conditioning; below one unit roundoff), whichever is looser
    s = oracle.system(xyz, kinds, None, 9.0, 6.5, tags=np.arange(len(kinds)), conditioning=True)
    close = lambda a, b, sc: abs(a - b) <= max(E_REL * abs(b), 1e-16 * sc)
    e2, e3, g, v = e.results()
    assert close(e2, r["e2b"], s["e2b"]) and close(e3, r["e3b"], s["e3b"]), (e2, e3, r["e2b"], r["e3b"])
    assert np.abs(g - r["grad"]).max() <= G_ABS
    e.compute(do_grads=False)
    f2, f3, _, _ = e.results(grad=False)
    assert close(f2, r["e2b"], s["e2b"]) and close(f3, r["e3b"], s["e3b"]), (f2, f3)


@pytest.mark.parametrize("split", [1, 2, 4, 8])
def test_ch4h2o2_splits(systems, split):
    """Every wavefront split of the blocked CH4-H2O-H2O polynomial (split_ch4h2o2_3b) gives the
    oracle's 3B energy and gradients on the CH4-rich box."""
    xyz, kinds, box, c2, c3, _ = mixed_case(systems, "ch4_342_h2o_228_rect")
    print("ch4h2o2 split", split, flush=True)
    e = engine.Engine()
    e.set_topology(kinds)
    e.set_box(box)
    e.set_cutoffs(c2, c3)
    e.set_families(("ch4h2o2_3b",))
    e.set_tuning("split_ch4h2o2_3b", split)
    assert e.tuning("split_ch4h2o2_3b") == split
    e.set_coords(xyz)
    e.compute()
    _, e3, g, _ = e.results()
    r = oracle.system(xyz, kinds, box, c2, c3, tags=np.arange(len(kinds)), families=("ch4h2o2_3b",))
    assert e3 != 0.0 and e3 == pytest.approx(r["e3b"], rel=E_REL)
    assert np.abs(g - r["grad"]).max() <= G_ABS


def test_autotune_ch4h2o2(systems):
    """mbxg_autotune also picks the CH4-H2O-H2O split when that family is evaluated (its whole
    evaluation is timed as a stage of its own), and the picked schedule keeps the oracle's values."""
    xyz, kinds, box, c2, c3, _ = mixed_case(systems, "ch4_342_h2o_228_rect")
    print("autotune ch4h2o2", flush=True)
    e = engine.Engine()
    e.set_topology(kinds)
    e.set_box(box)
    e.set_cutoffs(c2, c3)
    e.set_coords(xyz)
    e.compute()
    e.set_timing(True)
    e.compute()
    assert e.timings()["ch4h2o2_3b"] > 0.0
    e.set_timing(False)
    picked = e.autotune(reps=1)
    assert all(picked[k] in (1, 2, 4, 8) for k in ("split_2b", "split_3b", "split_ch4h2o2_3b"))
    e2, e3, g, _ = e.results()
    r = oracle.system(xyz, kinds, box, c2, c3, tags=np.arange(len(kinds)))
    assert e2 == pytest.approx(r["e2b"], rel=E_REL)
    assert e3 == pytest.approx(r["e3b"], rel=E_REL)
    assert np.abs(g - r["grad"]).max() <= G_ABS
