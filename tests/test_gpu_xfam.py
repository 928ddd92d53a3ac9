"""GPU parity of the further MB-nrg families and the 4B term against the oracle (oracle.xsystem,
oracle.clusters(4)).

Families: halide-water 2B and halide-water-water 3B (f-, cl-, br-, i-), CO2-CO2 2B, CO2 x3 3B,
water x4 4B, alkali-metal-ion-water 2B and 3B (li+, na+, k+, rb+, cs+), alone and mixed with
MB-pol and the CH4 families; the reference's Na+ / Cs+ water setups against their published
2B / 3B energies.

Bar: tetramer lists bit-exact as sets; listed and evaluated counts equal; each order's energy
within 1e-10 relative plus 1e-15 of the sum of the cancelling terms' magnitudes (cond: the CO2
and water 4B polynomials cancel terms up to 1e7-1e9 times their sum, which no evaluation order
can beat); each gradient component within 1e-8 kcal/mol/A plus 1e-14 of the summed cond of the
n-mers holding its monomer (a variable's derivative is at most a few times the variable)."""
import numpy as np
import pytest

from conftest import sorted_tuples, water_case

import oracle
from mbx_amd import engine

pytestmark = pytest.mark.gpu

E_REL = 1e-10
COND = 1e-15
G_ABS = 1e-8
IONS = ("f-", "cl-", "br-", "i-")
ALKALI = ("li+", "na+", "k+", "rb+", "cs+")
XF_ALK_2B = ("li_h2o_2b", "na_h2o_2b", "k_h2o_2b", "rb_h2o_2b", "cs_h2o_2b")
XF_ALK_3B = ("li_h2o2_3b", "na_h2o2_3b", "k_h2o2_3b", "rb_h2o2_3b", "cs_h2o2_3b")
XF_2B = ("f_h2o_2b", "cl_h2o_2b", "br_h2o_2b", "i_h2o_2b", "co2co2_2b") + XF_ALK_2B
XF_3B = ("f_h2o2_3b", "cl_h2o2_3b", "br_h2o2_3b", "i_h2o2_3b", "co2x3_3b") + XF_ALK_3B


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def halide_water(systems, nion=16, every=None, ions=IONS):
    """The 256-water cubic box with nion waters replaced by ions (those of `ions` in turn) at
    their oxygen positions."""
    xyz, box, _, _, _ = water_case(systems, "h2o_256_cubic")
    w = xyz.reshape(-1, 3, 3)
    step = every or len(w) // nion
    pick = {i * step: ions[i % len(ions)] for i in range(nion)}
    names, out = [], []
    for i in range(len(w)):
        if i in pick:
            names.append(pick[i])
            out.append(w[i, 0])
        else:
            names.append("h2o")
            out.extend(w[i])
    return np.concatenate(out), names, box


def co2_box(n_side=5, spacing=4.5, seed=7):
    """CO2 molecules (C, O, O; C-O 1.16 A) on a jittered cubic lattice, random orientations, in a
    periodic cube of edge n_side * spacing."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                c = (np.array([i, j, k]) + 0.5) * spacing + rng.uniform(-0.2, 0.2, 3)
                u = _rot(rng)[:, 0] * 1.16
                out += [c, c + u, c - u]
    edge = n_side * spacing
    return np.concatenate(out), ["co2"] * n_side ** 3, [edge, 0, 0, 0, edge, 0, 0, 0, edge]


def mixed(systems, seed=3):
    """The clathrate (CH4 + water) with some waters turned into ions (halides and alkali metals)
    and some into CO2 (carbon at the oxygen, a random axis)."""
    rng = np.random.default_rng(seed)
    xyz, mon, box = systems["clathrate/xyz"], list(systems["clathrate/mon"]), systems["clathrate/box"]
    nat = [3 if m == "h2o" else 5 for m in mon]
    first = np.concatenate([[0], np.cumsum(nat)[:-1]])
    waters = [i for i, m in enumerate(mon) if m == "h2o"]
    pick = rng.choice(waters, 30, replace=False)
    conv = {int(p): ((IONS + ALKALI)[i % 9] if i < 18 else "co2") for i, p in enumerate(pick)}
    names, out = [], []
    for i, m in enumerate(mon):
        a = xyz[3 * first[i]:3 * (first[i] + nat[i])].reshape(-1, 3)
        if i in conv and conv[i] == "co2":
            u = _rot(rng)[:, 0] * 1.16
            names.append("co2")
            out += [a[0], a[0] + u, a[0] - u]
        elif i in conv:
            names.append(conv[i])
            out.append(a[0])
        else:
            names.append(m)
            out.extend(a)
    return np.concatenate(out), names, box


def droplet_core(systems, radius=9.0):
    """The waters of the 2048-water droplet whose oxygen lies within radius of its centroid."""
    w = systems["h2o_2048_droplet/xyz"].reshape(-1, 9)
    o = w[:, :3]
    keep = np.linalg.norm(o - o.mean(0), axis=1) < radius
    return w[keep].reshape(-1), ["h2o"] * int(keep.sum()), None


def run_gpu(xyz, names, box, c2, c3, c4=0.0, terms=engine.TERM_2B | engine.TERM_3B, families=engine.FAMILIES):
    e = engine.Engine()
    e.set_topology([engine.KIND_OF[n] for n in names])
    e.set_box(None if box is None or len(box) == 0 else np.asarray(box, float))
    e.set_cutoffs(c2, c3)
    e.set_cutoff4b(c4)
    e.set_families(families)
    e.set_coords(np.asarray(xyz, float))
    e.compute(terms)
    e.names = list(names)
    return e


def check(e, r, orders=(2, 3)):
    """The engine's energies, gradient, virial and counts against an oracle.xsystem result."""
    te = e.term_energies()
    _, _, g, v = e.results()
    c = e.counts()
    xc = e.xcounts()
    ev = {2: c[1] + sum(xc[f] for f in XF_2B), 3: c[3] + sum(xc[f] for f in XF_3B), 4: xc["h2ox4_4b"]}
    listed = {2: c[0], 3: c[2], 4: xc["tetramers"]}
    cond = r["cond"]
    for o in orders:
        got, want = te[o - 1], r["e%db" % o]
        tol = E_REL * abs(want) + COND * cond[o]
        print("order %d: listed %d evaluated %d  E %.12g  oracle %.12g  |dE| %.3g  tol %.3g"
              % (o, listed[o], ev[o], got, want, abs(got - want), tol), flush=True)
        assert (listed[o], ev[o]) == r["counts"][o]
        assert abs(got - want) <= tol
    nat = np.array([3 if n in ("h2o", "co2") else 5 if n == "ch4" else 1 for n in e.names])
    gtol = G_ABS + 10 * COND * np.repeat(r["cond_mon"], 3 * nat)
    dg = np.abs(g - r["grad"])
    print("max |dgrad| %.3g  max |dgrad|/tol %.3g" % (dg.max(), (dg / gtol).max()), flush=True)
    assert np.all(dg <= gtol)
    vtol = G_ABS + 10 * COND * sum(cond[o] for o in orders) * 10.0
    np.testing.assert_allclose(v, r["virial"], rtol=1e-8, atol=vtol)


def test_halide_water_box(systems):
    xyz, names, box = halide_water(systems)
    print("halide_water 256 box, 16 ions", flush=True)
    e = run_gpu(xyz, names, box, 9.0, 4.5)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, tags=np.arange(len(names)), cond=True)
    check(e, r)
    xc = e.xcounts()
    assert all(xc[f] > 0 for f in XF_2B[:4] + XF_3B[:4])


def test_halide_water_ignore_lists(systems):
    """Families left out of the mask (the System's ignore lists) are listed but not evaluated."""
    xyz, names, box = halide_water(systems)
    skip = ("cl_h2o_2b", "i_h2o2_3b", "mbpol_3b")
    e = run_gpu(xyz, names, box, 9.0, 4.5, families=[f for f in engine.FAMILIES if f not in skip])
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, tags=np.arange(len(names)), cond=True,
                       ignore=[("cl-", "h2o"), ("h2o", "h2o", "i-"), ("h2o", "h2o", "h2o")])
    check(e, r)
    assert e.xcounts()["cl_h2o_2b"] == 0 and e.xcounts()["i_h2o2_3b"] == 0


def test_halide_water_cluster(systems):
    xyz, names, _ = halide_water(systems, nion=8)
    print("halide_water 256 gas phase, 8 ions", flush=True)
    e = run_gpu(xyz, names, None, 9.0, 4.5)
    r = oracle.xsystem(xyz, names, None, 9.0, 4.5, tags=np.arange(len(names)), cond=True)
    check(e, r)


@pytest.mark.parametrize("periodic", [True, False])
def test_co2(periodic):
    xyz, names, box = co2_box()
    box = box if periodic else None
    print("co2 125, periodic", periodic, flush=True)
    e = run_gpu(xyz, names, box, 9.0, 4.5)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, tags=np.arange(len(names)), cond=True)
    check(e, r)
    xc = e.xcounts()
    assert xc["co2co2_2b"] > 0 and xc["co2x3_3b"] > 0


def test_alkali_water_box(systems):
    """Alkali-metal ions in the 256-water box, 3B cutoff 7 A: the x3b_h2o_ion_v1x_deg4_filtered
    skip rule (two of three separations beyond 4.5 A) drops most listed ion-water-water trimers."""
    xyz, names, box = halide_water(systems, nion=20, ions=ALKALI)
    print("alkali_water 256 box, 20 ions", flush=True)
    e = run_gpu(xyz, names, box, 9.0, 7.0)
    r = oracle.xsystem(xyz, names, box, 9.0, 7.0, tags=np.arange(len(names)), cond=True)
    check(e, r)
    xc = e.xcounts()
    assert all(xc[f] > 0 for f in XF_ALK_2B + XF_ALK_3B)


def test_alkali_water_cluster_and_ignore_lists(systems):
    xyz, names, _ = halide_water(systems, nion=10, ions=ALKALI)
    print("alkali_water 256 gas phase, 10 ions", flush=True)
    e = run_gpu(xyz, names, None, 9.0, 6.0)
    r = oracle.xsystem(xyz, names, None, 9.0, 6.0, tags=np.arange(len(names)), cond=True)
    check(e, r)
    skip = ("na_h2o_2b", "cs_h2o2_3b")
    e = run_gpu(xyz, names, None, 9.0, 6.0, families=[f for f in engine.FAMILIES if f not in skip])
    r = oracle.xsystem(xyz, names, None, 9.0, 6.0, tags=np.arange(len(names)), cond=True,
                       ignore=[("h2o", "na+"), ("cs+", "h2o", "h2o")])
    check(e, r)
    assert e.xcounts()["na_h2o_2b"] == 0 and e.xcounts()["cs_h2o2_3b"] == 0


@pytest.mark.parametrize("setup", ["setup_h2o_1_na_1", "setup_h2o_2_na_1", "setup_h2o_1_cs_1", "setup_h2o_2_cs_1"])
def test_alkali_water_setups(setup):
    """The reference's Na+ / Cs+ water setups (src/tests/unittests/setup_h2o_<n>_<ion>_1.h): the
    published 2B and 3B energies (11 digits; plus the rounding the Cs+ trimer's cancelling terms
    leave), gas phase, with the oracle's gradients."""
    import json
    import os

    from conftest import GOLDEN

    d = json.load(open(os.path.join(GOLDEN, "kats.json")))[setup]
    names, xyz = d["monomer_names"], np.array(d["real_coords"])
    print(setup, names, flush=True)
    e = run_gpu(xyz, names, None, 100.0, 100.0)
    r = oracle.xsystem(xyz, names, None, 100.0, 100.0, tags=np.arange(len(names)), cond=True)
    check(e, r)
    te = e.term_energies()
    assert abs(te[1] - d["two_body_energy"]) <= 6e-11 * abs(d["two_body_energy"]) + COND * r["cond"][2]
    assert abs(te[2] - d["three_body_energy"]) <= 6e-11 * abs(d["three_body_energy"]) + COND * r["cond"][3]


def test_mixed_all_families(systems):
    """MB-pol, CH4-H2O, CH4-CH4, CH4-H2O-H2O, halide-water, alkali-ion-water and CO2 families
    in one box; n-mers whose names have no holder (co2-h2o, ion-ion, ...) are listed and
    skipped."""
    xyz, names, box = mixed(systems)
    print("mixed clathrate + ions + co2", flush=True)
    e = run_gpu(xyz, names, box, 9.0, 4.5)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, tags=np.arange(len(names)), cond=True)
    check(e, r)


def _tetramers_oracle(xyz, names, box, c4):
    nat = np.array([oracle.NAT_OF_NAME[n] for n in names])
    first = np.concatenate([[0], np.cumsum(nat)[:-1]])
    x = np.asarray(xyz, float)
    if box is not None:
        x = oracle.fix_monomer_coordinates(x, box, nat)
    fx = x.reshape(-1, 3)[first]
    return sorted_tuples(oracle.clusters(4, c4, fx.reshape(-1), box))


@pytest.mark.parametrize("case", ["box", "droplet"])
def test_water_4b(systems, case):
    if case == "box":
        xyz, box, _, _, _ = water_case(systems, "h2o_256_cubic")
        names = ["h2o"] * (len(xyz) // 9)
    else:
        xyz, names, box = droplet_core(systems)
    c4 = 4.0
    print("water 4B", case, len(names), flush=True)
    e = run_gpu(xyz, names, box, 9.0, 4.5, c4, engine.TERM_2B | engine.TERM_3B | engine.TERM_4B)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, c4, tags=np.arange(len(names)), cond=True)
    check(e, r, (2, 3, 4))
    assert r["counts"][4][1] > 100
    got = sorted_tuples(e.tetramers())
    want = _tetramers_oracle(xyz, names, box, c4)
    assert len(got) == len(want) and np.array_equal(got, want)


def test_4b_alone_and_masked(systems):
    """TERM_4B alone gives the 4B energy only; with the water 4B family out of the mask the
    tetramers are still listed and nothing is evaluated."""
    xyz, box, _, _, _ = water_case(systems, "h2o_256_cubic")
    names = ["h2o"] * (len(xyz) // 9)
    r = oracle.xsystem(xyz, names, box, 0.0, 0.0, 4.0, tags=np.arange(len(names)), cond=True)
    e = run_gpu(xyz, names, box, 9.0, 4.5, 4.0, engine.TERM_4B)
    te = e.term_energies()
    assert te[1] == 0.0 and te[2] == 0.0
    check(e, r, (4,))
    e = run_gpu(xyz, names, box, 9.0, 4.5, 4.0, engine.TERM_4B,
                families=[f for f in engine.FAMILIES if f != "h2ox4_4b"])
    assert e.term_energies()[3] == 0.0
    assert e.xcounts()["tetramers"] == r["counts"][4][0] and e.xcounts()["h2ox4_4b"] == 0


def test_4b_with_ions(systems):
    """Tetramers over a box with ions: only the all-water ones have a holder."""
    xyz, names, box = halide_water(systems)
    e = run_gpu(xyz, names, box, 9.0, 4.5, 4.0, engine.TERM_2B | engine.TERM_3B | engine.TERM_4B)
    r = oracle.xsystem(xyz, names, box, 9.0, 4.5, 4.0, tags=np.arange(len(names)), cond=True)
    check(e, r, (2, 3, 4))
    assert r["counts"][4][0] > r["counts"][4][1]
    got = sorted_tuples(e.tetramers())
    assert np.array_equal(got, _tetramers_oracle(xyz, names, box, 4.0))


def test_repeat_compute(systems):
    """The energies are reduced in a fixed order: a second compute repeats them bit for bit (the
    gradient's atomic adds may land in another order: to rounding)."""
    xyz, names, box = mixed(systems)
    e = run_gpu(xyz, names, box, 9.0, 4.5, 4.0, engine.TERM_2B | engine.TERM_3B | engine.TERM_4B)
    a = e.term_energies().copy()
    ga = e.results()[2].copy()
    e.compute(engine.TERM_2B | engine.TERM_3B | engine.TERM_4B)
    assert np.array_equal(a, e.term_energies())
    assert np.abs(ga - e.results()[2]).max() <= 1e-12 * max(1.0, np.abs(ga).max())
