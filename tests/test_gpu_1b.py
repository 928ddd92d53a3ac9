"""GPU parity of the MB-pol 1B term (mbxg_eval_1b: the Partridge-Schwenke water monomer,
ps::pot_nasa, potential/1b/ps.cpp:346) against the oracle's restatement, which is pinned to the
reference's unittest-ps.cpp known answers, to the compiled reference and to the published 1B
energy of timing/256_h2o_cubic_box (tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle
from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu


def test_1b_water_box_against_oracle(systems):
    xyz = systems["h2o_256_cubic/xyz"]
    box = systems["h2o_256_cubic/box"]
    print("1b 256 waters", flush=True)
    e = engine.water_engine(len(xyz) // 9, box, 9.0, 7.0)
    e.set_coords(xyz)
    e1, g, v = e.eval_1b()
    w = xyz.reshape(-1, 9)
    er, gr, vr = oracle.ps_1b(w)
    assert e1 == pytest.approx(er.sum(), rel=1e-12)
    assert e1 == pytest.approx(3.0996035371e+02, rel=5e-11)  # timing/256_h2o_cubic_box/output
    assert np.abs(g - gr.reshape(-1)).max() <= 1e-10
    np.testing.assert_allclose(v[[0, 1, 2, 4, 5, 8]], vr, rtol=1e-10, atol=1e-9)
    assert v[3] == v[1] and v[6] == v[2] and v[7] == v[5]
    e0, g0, v0 = e.eval_1b(do_grads=False)
    assert e0 == e1 and g0 is None and not v0.any()


def test_1b_ghosts_and_kat():
    """Only local waters count; the reference's two-water known answer."""
    import json
    import os

    from conftest import GOLDEN

    k = json.load(open(os.path.join(GOLDEN, "kats.json")))["ps_1b"]
    xyz = np.array(k["xyz"])
    e = engine.Engine()
    e.set_topology([engine.H2O, engine.H2O], islocal=[1, 0])
    e.set_box(None)
    e.set_cutoffs(9.0, 7.0)
    e.set_coords(xyz)
    e1, g, _ = e.eval_1b()
    assert e1 == pytest.approx(k["energies_expected"][0], rel=1e-6)
    np.testing.assert_allclose(g[:9], k["grad_expected"][:9], rtol=1e-6, atol=1e-6)
    assert not g[9:].any()
    rng = np.random.default_rng(8)
    w = configs.water_geometries(rng, 100).reshape(100, 9) + rng.uniform(-5, 5, (100, 1)).repeat(9, 1)
    e = engine.Engine()
    e.set_topology([engine.H2O] * 100)
    e.set_box(None)
    e.set_cutoffs(9.0, 7.0)
    e.set_coords(w)
    e1, g, _ = e.eval_1b()
    er, gr, _ = oracle.ps_1b(w)
    assert e1 == pytest.approx(er.sum(), rel=1e-12)
    assert np.abs(g - gr.reshape(-1)).max() <= 1e-10


# ---------------------------------------------------------------------------------------------
# MB-nrg CH4 1B and the 1B term inside a combined compute


def _ch4_conditioning(xyz, kinds, box):
    """Sum over the CH4 monomers of sum_k |a_k| M_k: the FP64 floor of the CH4 1B energy
    (about 2e7 per monomer for ~3 kcal/mol; see tests/test_oracle_nrg.py)."""
    nat = np.where(np.asarray(kinds) == 0, 3, 5)
    x = oracle.fix_monomer_coordinates(xyz, box, nat) if box is not None else np.asarray(xyz, np.float64)
    first = np.concatenate([[0], np.cumsum(nat)[:-1]])
    w = [x[3 * f:3 * f + 15] for f, k in zip(first, kinds) if k == 1]
    if not w:
        return 0.0
    a = np.abs(oracle.table("ch4_1b").coef)
    return float(oracle.nrg_1b("ch4_1b", np.array(w), a=a, grad=False)[0].sum())


MIXED = ["ch4_57_h2o_513_rect", "ch4_342_h2o_228_rect", "clathrate"]


@pytest.mark.parametrize("name", MIXED)
def test_1b_mixed_published_and_oracle(systems, name):
    """The CH4 / H2O reference systems: E1B against the published 1B energy (timing/<case>/output,
    11 significant digits) and the oracle; the 1B gradient and virial against the oracle."""
    kinds = np.array([0 if m == "h2o" else 1 for m in systems[name + "/mon"]])
    xyz = systems[name + "/xyz"]
    box = systems[name + "/box"]
    box = box if box.size else None
    pub = systems[name + "/published"]
    print("1b", name, flush=True)
    e = engine.Engine()
    e.set_topology(kinds)
    e.set_box(box)
    e.set_cutoffs(9.0, 0.0)
    e.set_coords(xyz)
    e1, g, v = e.eval_1b()
    er, gr, vr = oracle.one_body(xyz, kinds, box)
    floor = 2e-16 * _ch4_conditioning(xyz, kinds, box)
    assert abs(e1 - er) <= max(1e-12 * abs(er), floor)
    assert abs(e1 - pub[0]) <= max(5e-11 * abs(pub[0]), 5 * floor)
    assert np.abs(g - gr).max() <= 1e-8
    np.testing.assert_allclose(v, vr, rtol=1e-9, atol=1e-8)


def test_1b_inside_compute(systems):
    """MBXG_1B | MBXG_2B | MBXG_3B in one compute (what System::Energy runs): each energy equals
    its own compute, the gradient and virial equal their sums; graph replays give the same."""
    name = "ch4_57_h2o_513_rect"
    kinds = np.array([0 if m == "h2o" else 1 for m in systems[name + "/mon"]])
    xyz = systems[name + "/xyz"]
    box = systems[name + "/box"]
    c2, c3 = (float(v) for v in systems[name + "/cutoffs"])
    print("1b inside compute", flush=True)
    e = engine.Engine()
    e.set_topology(kinds)
    e.set_box(box)
    e.set_cutoffs(c2, c3)
    e.set_coords(xyz)
    e.compute(engine.TERM_1B)
    e1_only = e.energies()[0]
    _, _, g1, v1 = e.results()
    e.compute(engine.TERM_2B | engine.TERM_3B)
    _, e2_only, e3_only = e.energies()
    _, _, g23, v23 = e.results()
    e.set_graph(True)
    for rep in range(4):  # eager (sizes), eager, then graph replays
        e.compute(engine.TERM_1B | engine.TERM_2B | engine.TERM_3B)
        e1, e2, e3 = e.energies()
        _, _, g, v = e.results()
        assert e1 == pytest.approx(e1_only, rel=1e-13), rep
        assert e2 == pytest.approx(e2_only, rel=1e-12) and e3 == pytest.approx(e3_only, rel=1e-12), rep
        assert np.abs(g - (g1 + g23)).max() <= 1e-10, rep
        np.testing.assert_allclose(v, v1 + v23, rtol=1e-11, atol=1e-9)
    e.set_graph(False)
    with pytest.raises(engine.MbxgError):  # the 1B term is not evaluated over batches
        e.set_batch([0, 10, len(kinds)])
        e.compute(engine.TERM_1B)


def test_1b_co2_kat_and_oracle():
    """The CO2 1B holder (x1b_A1B2_v1x): the reference's three-monomer known answer
    (unittest-poly-holder-1b.cpp:52-83), then 125 distorted monomers against the oracle."""
    import json
    import os

    from conftest import GOLDEN

    k = json.load(open(os.path.join(GOLDEN, "kats.json")))["holder_co2_1b"]
    print("1b co2 kat", flush=True)
    e = engine.Engine()
    e.set_topology([engine.CO2] * 3)
    e.set_box(None)
    e.set_cutoffs(9.0, 7.0)
    e.set_coords(np.array(k["xyz"]))
    e1, g, v = e.eval_1b()
    assert e1 == pytest.approx(sum(k["energies_expected"]), abs=3e-9)
    np.testing.assert_allclose(g, k["grad_expected"], rtol=0, atol=1e-8)
    np.testing.assert_allclose(v, k["virial_expected"], rtol=0, atol=1e-8)
    rng = np.random.default_rng(4)
    w = []
    for _ in range(125):
        q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
        c = rng.uniform(-8, 8, 3)
        th = np.pi - rng.uniform(0.0, 0.2)
        r1, r2 = 1.16 + rng.uniform(-0.05, 0.05, 2)
        w.append(np.concatenate([c, c + q @ [r1, 0, 0], c + q @ [r2 * np.cos(th), r2 * np.sin(th), 0]]))
    w = np.array(w)
    e = engine.Engine()
    e.set_topology([engine.CO2] * len(w))
    e.set_box(None)
    e.set_cutoffs(9.0, 7.0)
    e.set_coords(w)
    e1, g, v = e.eval_1b()
    er, gr, vr = oracle.nrg_1b("co2_1b", w)
    cond = oracle.nrg_1b("co2_1b", w, a=np.abs(oracle.table("co2_1b").coef), grad=False)[0]
    assert abs(e1 - er.sum()) <= 1e-12 * abs(er.sum()) + 4e-16 * cond.sum()
    assert np.abs(g - gr.reshape(-1)).max() <= 1e-10 + 4e-15 * cond.max()
    np.testing.assert_allclose(v[[0, 1, 2, 4, 5, 8]], vr, rtol=1e-10, atol=1e-9)


def test_1b_mixed_co2_halides_in_compute(systems):
    """1B + 2B + 3B in one compute over waters, CH4, CO2 and ions: the 1B part equals the
    oracle's one_body (ions have none), and the total gradient is 1B + 2B + 3B."""
    import test_gpu_xfam as X

    xyz, names, box = X.mixed(systems)
    kinds = [engine.KIND_OF[n] for n in names]
    print("1b mixed co2 + halides", flush=True)
    e = X.run_gpu(xyz, names, box, 9.0, 4.5, terms=engine.TERM_1B | engine.TERM_2B | engine.TERM_3B)
    te = e.term_energies()
    _, _, g, _ = e.results()
    e1, g1, _ = oracle.one_body(xyz, kinds, box)
    # 72 CH4 monomers: their 1B terms cancel ~2e7 each, a floor of ~2e-16 * 1.4e9 ~ 3e-7
    assert te[0] == pytest.approx(e1, rel=1e-10, abs=1e-6)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, tags=np.arange(len(names)), cond=True)
    nat = np.array([engine.NAT_OF[k] for k in kinds])
    gtol = X.G_ABS + 1e-7 + 10 * X.COND * np.repeat(r["cond_mon"], 3 * nat)
    assert np.all(np.abs(g - (g1 + r["grad"])) <= gtol)
