"""CPU tests: pin the oracle's further MB-nrg families (orc_xholder_eval, the name dispatch and the
tetramer lists) to the reference.

Families: halide-water 2B (mbnrg_A1_B1C2X2_deg6_v1: f-, cl-, br-, i-), halide-water-water 3B
(mbnrg_A1_B1C2X2_B1C2X2_deg4_v1), CO2-CO2 2B (x2b_A1B2_A1B2_v1x, the "co2" model), CO2 x3 3B
(mbnrg_A1B2_A1B2_A1B2_deg4_v1), water x4 4B (mbnrg_A1B2_A1B2_A1B2_A1B2_deg4_v1), and the
alkali-metal-ion-water 2B (x2b_h2o_ion_v2x) and 3B (x3b_h2o_ion_v1x_deg4_filtered: li+, na+, k+,
rb+, cs+).

Pins:
  * energy routines -- the reference's known answers of unittest-energy2b.cpp (f--h2o, cl-h2o,
                       br-h2o, i-h2o, the second co2-co2 section), unittest-energy3b.cpp
                       (f/cl/br/i-h2o-h2o, co2-co2-co2, li-h2o-h2o, cs-h2o-h2o) and
                       unittest-energy4b.cpp (h2o x4):
                       energy, gradients and virial, with the names given in every order the
                       reference tests (the name sorts of get_2b/3b/4b_energy);
  * alkali holders  -- unittest-poly-holder-2b.cpp x2b_h2o_ion_v2x::struct and
                       unittest-poly-holder-3b.cpp x3b-3b-h2o-ion-v1x_deg4_filtered::class, every
                       ion (energy, gradients, virial); the Na+ and Cs+ water setups'
                       System-level 2B / 3B energies (src/tests/unittests/setup_h2o_{1,2}_{na,cs}_1.h);
  * polynomial      -- unittest-poly-2b.cpp :poly_2b_A1B2_A1B2_v1x:struct and unittest-poly-3b.cpp
                       poly-3b-h2o-ion-v1x_deg4_filtered::class (21-digit value and gradient) and
                       linearity / Euler's identity for every table;
  * holders         -- the compiled reference holders (oracle/_ref/libmbxref_x.so) on perturbed
                       n-mers, to rounding (1e-15 of the sum of the cancelling terms' magnitudes);
                       five-point central differences of the oracle's energy;
  * lists           -- AddClusters for tetramers (the center-3-neighbour rule) against the
                       compiled AddClusters, periodic and gas phase, and against a brute force.
"""
import itertools
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, sorted_tuples

import oracle

KATS = json.load(open(os.path.join(GOLDEN, "kats.json")))
refx_only = pytest.mark.skipif(not oracle.refx_available(),
                               reason="oracle/_ref/libmbxref_x.so not built (no reference checkout)")
ref_only = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference checkout)")

FAMS = {"f_h2o_2b": "energy_f_h2o_2b", "cl_h2o_2b": "energy_cl_h2o_2b", "br_h2o_2b": "energy_br_h2o_2b",
        "i_h2o_2b": "energy_i_h2o_2b", "co2co2_2b": "energy_co2co2_2b", "f_h2o2_3b": "energy_f_h2o2_3b",
        "cl_h2o2_3b": "energy_cl_h2o2_3b", "br_h2o2_3b": "energy_br_h2o2_3b", "i_h2o2_3b": "energy_i_h2o2_3b",
        "co2x3_3b": "energy_co2x3_3b", "h2ox4_4b": "energy_h2ox4_4b", "li_h2o2_3b": "energy_li_h2o2_3b",
        "cs_h2o2_3b": "energy_cs_h2o2_3b"}
ALKALI = ("li", "na", "k", "rb", "cs")
# holder known answers, in the holder's own argument order (water(s), then the ion)
HOLDER_KATS = {"%s_h2o_2b" % a: "holder_%s_h2o_2b" % a for a in ALKALI}
HOLDER_KATS.update({"%s_h2o2_3b" % a: "holder_%s_h2o2_3b" % a for a in ALKALI})
ALL = sorted(set(FAMS) | set(HOLDER_KATS))
TOL = 1e-6  # the reference's own margin (unittest-energy2b.cpp:45)


def kat_case(name):
    k = KATS[FAMS[name]]
    order = sum(1 for m in ("mon1", "mon2", "mon3", "mon4") if m in k)
    names = [k["mon%d" % (i + 1)] for i in range(order)]
    xs = [np.array(k["xyz%d" % (i + 1)]).reshape(k["nm"], -1) for i in range(order)]
    return k, names, xs


def conditioning(name, ws):
    """sum_k |a_k| M_k(x) sw: the magnitude of the polynomial's cancelling terms (x > 0)."""
    return abs(oracle.xnmer(name, ws, a=np.abs(oracle.xholder(name).coef))[0])


def run_names(names, xs):
    """The n-mers of a known-answer case through the dispatch: (E, grads per listed monomer, virial)."""
    order, nm = len(names), len(xs[0])
    hn, perm = oracle.holder_order(names)
    E, V = 0.0, np.zeros(9)
    G = [np.zeros_like(x) for x in xs]
    for j in range(nm):
        e, gs, v, _ = oracle.xnmer(hn, [xs[p][j] for p in perm])
        E += e
        V += v
        for s, p in enumerate(perm):
            G[p][j] += gs[s]
    return hn, E, G, V


@pytest.mark.parametrize("name", sorted(FAMS))
def test_energy_kats(name):
    k, names, xs = kat_case(name)
    hn, E, G, V = run_names(names, xs)
    assert hn == name
    want = k.get("energy_expected", k.get("expected_energy"))
    # the answers are printed with 9 to 11 digits and carry the reference's own rounding of
    # strongly cancelling sums (CO2-CO2: E = -0.073 from 2e7 of cancelling terms): 1e-8 relative
    # (the reference's own margin is 1e-6 absolute)
    assert abs(E - want) <= 1e-8 * abs(want) + 1e-12, (E, want)
    for i in range(len(names)):
        np.testing.assert_allclose(G[i].reshape(-1), k["grad%d_expected" % (i + 1)], rtol=0, atol=TOL)
    np.testing.assert_allclose(V, k["virial_expected"], rtol=0, atol=TOL)
    # every order of the names reaches the same holder and terms (the reference's KATs call
    # several orders); identical monomers then enter the symmetric polynomial in another order,
    # which changes only the rounding
    hn, perm = oracle.holder_order(names)
    cond = sum(conditioning(name, [xs[q][j] for q in perm]) for j in range(len(xs[0])))
    for p in itertools.permutations(range(len(names))):
        hp, Ep, Gp, Vp = run_names([names[i] for i in p], [xs[i] for i in p])
        assert hp == name and abs(Ep - E) <= 1e-15 * cond
        for a, b in enumerate(p):
            np.testing.assert_allclose(Gp[a], G[b], rtol=0, atol=1e-12 * max(1.0, cond))


def test_dispatch_has_no_term_for_other_mixtures():
    for names in (["f-", "co2"], ["cl-", "cl-"], ["h2o", "co2"], ["co2", "co2", "h2o"], ["f-", "f-", "h2o"],
                  ["h2o", "h2o", "h2o", "co2"], ["co2"] * 4):
        assert oracle.holder_order(names) == (None, None)


def test_poly_co2co2_kat():
    k = KATS["poly_co2co2_2b"]
    t = oracle.table("co2co2_2b")  # the known answer comes with coefficients of its own
    e, g = oracle.poly_eval(t, k["x"], a=k["a"])
    ea, _ = oracle.poly_eval(t, np.abs(k["x"]), a=np.abs(k["a"]))
    assert abs(e - k["e_expected"]) <= 1e-15 * ea + 1e-13 * abs(e)
    np.testing.assert_allclose(g, k["g_expected"], rtol=1e-10, atol=1e-14 * ea)


@pytest.mark.parametrize("poly", ["halide_h2o_2b", "halide_h2o2_3b", "co2co2_2b", "co2x3_3b", "h2ox4_4b",
                                  "alkali_h2o_2b", "alkali_h2o2_3b"])
def test_poly_linear_and_euler(poly):
    t = oracle.table(poly)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.05, 1.0, t.nvars)
    a, b = rng.normal(size=t.ncoef), rng.normal(size=t.ncoef)
    ea, _ = oracle.poly_eval(t, x, a=a)
    eb, _ = oracle.poly_eval(t, x, a=b)
    eab, _ = oracle.poly_eval(t, x, a=2.0 * a - 3.0 * b)
    scale = oracle.poly_eval(t, x, a=np.abs(a) + np.abs(b))[0]
    assert abs(eab - (2.0 * ea - 3.0 * eb)) <= 1e-14 * scale
    basis = oracle.poly_basis(t, x)
    assert abs(float(np.dot(a, basis)) - ea) <= 1e-14 * scale


def holder_case(name):
    """A holder known answer: (case, n-mers as lists of monomers in the holder's order)."""
    k = KATS[HOLDER_KATS[name]]
    order = 2 if name.endswith("_2b") else 3
    n = k.get("ndim", k.get("ntrim"))
    xs = [np.array(k["xyz%d" % (i + 1)]).reshape(n, -1) for i in range(order)]
    return k, [[x[j] for x in xs] for j in range(n)]


@pytest.mark.parametrize("name", sorted(HOLDER_KATS))
def test_alkali_holder_kats(name):
    k, nmers = holder_case(name)
    order = len(nmers[0])
    E, V = 0.0, np.zeros(9)
    G = [np.zeros((len(nmers), len(m))) for m in nmers[0]]
    for j, ws in enumerate(nmers):
        e, gs, v, _ = oracle.xnmer(name, ws)
        E += e
        V += v
        for i in range(order):
            G[i][j] = gs[i]
    want = k["energy_expected"]
    assert abs(E - want) <= 1e-9 * abs(want) + 1e-11, (E, want)  # printed with 11 digits
    for i in range(order):
        np.testing.assert_allclose(G[i].reshape(-1), k["grad%d_expected" % (i + 1)], rtol=0, atol=TOL)
    np.testing.assert_allclose(V, k["virial_expected"], rtol=0, atol=TOL)


def test_alkali_poly_kat():
    k = KATS["poly_alkali_h2o2_3b"]
    t = oracle.table("alkali_h2o2_3b")  # the known answer comes with coefficients of its own
    e, g = oracle.poly_eval(t, k["x"], a=k["a"])
    ea, _ = oracle.poly_eval(t, np.abs(k["x"]), a=np.abs(k["a"]))
    assert abs(e - k["e_expected"]) <= 1e-15 * ea + 1e-13 * abs(e)
    np.testing.assert_allclose(g, k["g_expected"], rtol=1e-10, atol=1e-14 * ea)


@pytest.mark.parametrize("setup", ["setup_h2o_1_na_1", "setup_h2o_2_na_1", "setup_h2o_1_cs_1", "setup_h2o_2_cs_1"])
def test_alkali_water_system_energies(setup):
    """One or two waters with Na+ or Cs+ through the System-level orchestration (gas phase): the
    2B energy (water-water MB-pol plus the ion-water dimers) and the 3B energy (the ion-water-water
    trimer) the reference's setups hold, printed with 11 digits; their virials with 6."""
    d = KATS[setup]
    r = oracle.xsystem(d["real_coords"], d["monomer_names"], None, 100.0, 100.0, cond=True)
    nw = d["monomer_names"].count("h2o")
    assert r["counts"][2] == (nw * (nw + 1) // 2,) * 2 and r["counts"][3] == ((1, 1) if nw == 2 else (0, 0))
    # half a unit of the printed 11th digit, plus the rounding any evaluation order leaves (the
    # Cs+ trimer's cancelling terms sum to 8e5 for E3B = 0.2)
    assert abs(r["e2b"] - d["two_body_energy"]) <= 6e-11 * abs(d["two_body_energy"]) + 1e-15 * r["cond"][2]
    assert abs(r["e3b"] - d["three_body_energy"]) <= 6e-11 * abs(d["three_body_energy"]) + 1e-15 * r["cond"][3]
    vir = np.array(d["virial_2b"]) + np.array(d.get("virial_3b", np.zeros(9)))
    np.testing.assert_allclose(r["virial"], vir, rtol=1e-5, atol=1e-5)


def perturbed(name, rng, sigma=0.05):
    if name in HOLDER_KATS:
        _, nmers = holder_case(name)
        ws = nmers[rng.integers(len(nmers))]
        return [w + rng.normal(0.0, sigma, w.shape) for w in ws]
    k, names, xs = kat_case(name)
    hn, perm = oracle.holder_order(names)
    j = rng.integers(len(xs[0]))
    return [xs[p][j] + rng.normal(0.0, sigma, xs[p][j].shape) for p in perm]


@refx_only
@pytest.mark.parametrize("name", ALL)
def test_holders_vs_compiled_reference(name):
    rng = np.random.default_rng(11)
    evaluated = 0
    for _ in range(12):
        ws = perturbed(name, rng)
        e, gs, v, ev = oracle.xnmer(name, ws)
        er, gr, vr = oracle.ref_xnmer(name, ws)
        evaluated += ev
        cond = conditioning(name, ws)
        assert abs(e - er) <= 1e-15 * cond + 1e-300, (e, er, cond)
        gmax = max(np.abs(g).max() for g in gr) + 1e-300
        for a, b in zip(gs, gr):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-9 * gmax + 1e-15 * cond)
        np.testing.assert_allclose(v, vr, rtol=0, atol=1e-9 * np.abs(vr).max() + 1e-15 * cond)
    assert evaluated > 0


@pytest.mark.parametrize("name", ALL)
def test_holder_finite_differences(name):
    rng = np.random.default_rng(5)
    for _ in range(50):  # an evaluated n-mer (some known-answer cases hold skipped ones)
        ws = perturbed(name, rng, sigma=0.02)
        e, gs, _, ev = oracle.xnmer(name, ws)
        if ev:
            break
    assert ev == 1
    # five-point stencil; the energy's own rounding (about eps times the sum of the cancelling
    # terms' magnitudes, 2e7 for CO2-CO2) enters the difference quotient divided by h
    h = 1e-4
    noise = 4 * 2.2e-16 * conditioning(name, ws) / h
    for m, w in enumerate(ws):
        for i in rng.choice(len(w), size=min(4, len(w)), replace=False):
            def E(dx):
                ww = [x.copy() for x in ws]
                ww[m][i] += dx
                return oracle.xnmer(name, ww)[0]
            fd = (-E(2 * h) + 8 * E(h) - 8 * E(-h) + E(-2 * h)) / (12 * h)
            gmax = max(np.abs(g).max() for g in gs)
            assert abs(fd - gs[m][i]) <= 1e-6 * gmax + noise, (m, i, fd, gs[m][i])


def test_skip_rule_and_switch_ends():
    """Beyond the outer radius on every pair the n-mer is skipped; inside the inner radius on
    every pair the switch combination is 1 (2B: sw12; 3B: s12 s13 + s12 s23 + s13 s23 - 2 s12 s13 s23;
    4B: the center-3-neighbour combination), and 3 for the alkali-ion 3B holder, whose combination
    s_ab s_ac + s_ab s_bc + s_ac s_bc has no cubic term (x3b-h2o-ion-v1x-deg4-filtered.cpp:306)."""
    for name in ALL:
        h = oracle.xholder(name).desc
        terms = h["switch"]["terms"]
        want = 3.0 if h["poly"] == "alkali_h2o2_3b" else 1.0
        assert abs(sum(c for c, _ in terms) - want) < 1e-15  # all switches 1
        ws = perturbed(name, np.random.default_rng(2), sigma=0.0)
        far = [w + 100.0 * i for i, w in enumerate(ws)]
        e, _, _, ev = oracle.xnmer(name, far)
        assert ev == 0 and e == 0.0


def _brute_tetramers(first, box, cutoff):
    """Every 4-subset that has a member within the cutoff (minimum image) of the other three:
    AddClusters lists it once, from the lowest-tagged such member (sys_tools.cpp:594-634).  The
    cutoff must stay below a quarter of the box, so that two neighbours of one monomer are as
    far apart as their minimum images."""
    n = len(first)
    d = first[:, None, :] - first[None, :, :]
    if box is not None:
        B = np.array(box, float).reshape(3, 3)
        f = d @ np.linalg.inv(B)
        f -= np.round(f)
        d = f @ B
    near = (d * d).sum(-1) < cutoff * cutoff
    out = []
    for s in itertools.combinations(range(n), 4):
        if any(all(near[q, o] for o in s if o != q) for q in s):
            out.append(s)
    return np.array(out, np.int64).reshape(-1, 4)


def _tetramer_box(n, edge, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-edge / 2, edge / 2, (n, 3)), [edge, 0, 0, 0, edge, 0, 0, 0, edge]


def test_tetramers_are_the_star_subsets():
    """Each 4-subset with a member within the cutoff of the other three is listed exactly once
    (periodic cube and gas phase)."""
    first, box = _tetramer_box(48, 11.0, 3)
    for b in (box, None):
        got = oracle.clusters(4, 2.7, first, b)
        assert len(got) > 10
        st = sorted_tuples(got)
        assert len(np.unique(st, axis=0)) == len(st)  # no tetramer twice
        assert np.array_equal(st, sorted_tuples(_brute_tetramers(first, b, 2.7)))


@ref_only
def test_tetramers_vs_compiled_reference():
    first, box = _tetramer_box(40, 12.5, 9)
    for b in (box, None):
        got = sorted_tuples(oracle.clusters(4, 5.0, first, b))
        want = sorted_tuples(oracle.ref_add_clusters(4, 5.0, first, b))
        assert len(got) == len(want) and np.array_equal(got, want)


@pytest.mark.parametrize("ion", ALKALI)
def test_alkali_3b_skips_two_pairs_beyond(ion):
    """x3b_h2o_ion_v1x_deg4_filtered skips a trimer when two of its three O-O / O-X separations
    exceed r3f = 4.5 A (x3b-h2o-ion-v1x-deg4-filtered.cpp:295); with one beyond it is evaluated
    (and its switch combination, one product of two switches, is not zero)."""
    name = "%s_h2o2_3b" % ion
    _, nmers = holder_case(name)
    w1, w2, x = nmers[0]
    o1 = w1[:3]
    def place(w, target):  # rigid shift of a monomer so its first atom sits at target
        return w - w[:3].tolist() * (len(w) // 3) + np.tile(target, len(w) // 3)
    u = np.array([1.0, 0.0, 0.0])
    v = np.array([0.0, 1.0, 0.0])
    # one pair beyond (a-b at 5 A, a-x and b-x near 3.5 A): evaluated
    e1, _, _, ev1 = oracle.xnmer(name, [w1, place(w2, o1 + 5.0 * u), place(x, o1 + 2.5 * u + 2.5 * v)])
    assert ev1 == 1 and e1 != 0.0
    # two pairs beyond (x at 6 A from both waters, the waters 3 A apart): skipped
    e2, _, _, ev2 = oracle.xnmer(name, [w1, place(w2, o1 + 3.0 * u), place(x, o1 + 1.5 * u + 6.0 * v)])
    assert ev2 == 0 and e2 == 0.0
