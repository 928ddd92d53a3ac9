"""CPU tests: pin the oracle's geometry and list functions.

  * orc_invert_unit_cell, orc_fix_monomer_coordinates, orc_close_image -- bit for bit against the
    compiled reference (InvertUnitCell, systools::FixMonomerCoordinates, GetCloseDimerImage from
    oracle/_ref) on cubic, rectangular, hexagonal, triclinic and skewed boxes, and against the
    reference's own known answers (unittest-systools.cpp TEST_CASEs "systools::FixMonomerCoordinates"
    :252-410, "systools::GetCloseDimerImage" :412-589, "systools::GetCloseTrimerImage" :591-832);
  * orc_clusters -- the AddClusters counts of unittest-systools.cpp:852-1044 (27 points of a cube,
    gas phase) and the explicit pair / triple sets of unittest-system.cpp:644-718 (the reference
    test compares them sorted within and across n-mers, as here), also through the compiled
    AddClusters;
  * the water M-site of System::GetXyz -- unittest-systools.cpp:1477-1505 (SetVSites), through
    the library's get_xyz_ entry point (a gas-phase system, which needs no device).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, sorted_tuples

import oracle

KATS = json.load(open(os.path.join(GOLDEN, "kats.json")))
ref_only = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference checkout)")

TOL = 1e-6  # the reference tests' margin (unittest-systools.cpp:51)

CUBIC = [10.0, 0.0, 0.0, 0.0, 10.0, 0.0, 0.0, 0.0, 10.0]
HEXAGONAL = [10.0, 0.0, 0.0, 5.0, 8.660254037844386, 0.0, 0.0, 0.0, 20.0]  # unittest-systools.cpp:305
TRICLINIC = [10.0, 0.0, 0.0, 10.0, 15.0, 0.0, 5.0, 5.0, 20.0]  # unittest-systools.cpp:501


def _boxes():
    rng = np.random.default_rng(7)
    skew = [12.3, 0.0, 0.0, 3.1, 11.7, 0.0, -2.2, 4.4, 13.9]
    rand = list(np.array([[15.0, 0, 0], [2.0, 14.0, 0], [1.0, -3.0, 16.0]]).reshape(-1) +
                np.concatenate([[0, 0, 0], rng.uniform(-1, 1, 2), [0], rng.uniform(-1, 1, 3)]))
    return {"cubic": CUBIC, "rectangular": [19.7, 0, 0, 0, 21.6, 0, 0, 0, 30.0], "hexagonal": HEXAGONAL,
            "triclinic": TRICLINIC, "skewed": skew, "random": rand}


BOXES = _boxes()


def _shift(base, *terms):
    """base (atoms x 3) plus, per atom and component, ((c1 v1 + c2 v2) + c3 v3) in the reference
    test's operation order (a `+=` of one expression); terms = (c1, v1), (c2, v2), (c3, v3)."""
    base = np.array(base, np.float64).reshape(-1, 3)
    s = np.zeros(3)
    for j in range(3):
        acc = terms[0][0] * terms[0][1][j]
        for c, v in terms[1:]:
            acc = acc + c * v[j]
        s[j] = acc
    return (base + s).reshape(-1)


# ---------------------------------------------------------------------------------------------
# InvertUnitCell, FixMonomerCoordinates, GetCloseDimerImage against the compiled reference


@ref_only
@pytest.mark.parametrize("name", sorted(BOXES))
def test_invert_unit_cell_bit_exact(name):
    box = np.array(BOXES[name])
    assert np.array_equal(oracle.invert_unit_cell(box), oracle.ref_invert_unit_cell(box))


@ref_only
@pytest.mark.parametrize("name", sorted(BOXES))
def test_fix_monomer_coordinates_bit_exact(name):
    """Monomers of 1-7 atoms whose atoms sit several cells away from the first one, first atoms
    spread over [-2.5, 2.5] cells, plus first atoms exactly at half-cell fractions."""
    box = np.array(BOXES[name])
    b = box.reshape(3, 3)
    rng = np.random.default_rng(11)
    nat = rng.integers(1, 8, 400)
    xyz = []
    for n in nat:
        f0 = rng.uniform(-2.5, 2.5, 3)
        first = f0 @ b
        xyz.append(first)
        for _ in range(n - 1):
            xyz.append(first + rng.normal(0, 1.2, 3) + rng.integers(-2, 3, 3) @ b)
    for q in range(3):  # fractional coordinates at +-0.5: the floor(x + 0.5) boundary
        for s in (-0.5, 0.5, 1.5):
            f = np.zeros(3)
            f[q] = s
            xyz.append(f @ b)
    nat = np.concatenate([nat, np.ones(9, np.int64)])
    xyz = np.concatenate(xyz)
    got = oracle.fix_monomer_coordinates(xyz, box, nat)
    want = oracle.ref_fix_monomer_coordinates(xyz, box, nat)
    assert np.array_equal(got, want)


@ref_only
@pytest.mark.parametrize("name", sorted(BOXES))
def test_close_image_bit_exact(name):
    box = np.array(BOXES[name])
    b = box.reshape(3, 3)
    rng = np.random.default_rng(5)
    for _ in range(300):
        n1, n2 = rng.integers(1, 6, 2)
        x1 = (rng.uniform(-1, 1, 3) @ b + rng.normal(0, 1, (n1, 3))).reshape(-1)
        x2 = (rng.uniform(-3, 3, 3) @ b + rng.normal(0, 1, (n2, 3))).reshape(-1)
        assert np.array_equal(oracle.close_image(box, x1, x2), oracle.ref_close_image(box, x1, x2))


# ---------------------------------------------------------------------------------------------
# The reference's known answers


def test_fix_monomer_coordinates_kat_cubic():
    """unittest-systools.cpp:253-284."""
    c1 = [4.9, 4.9, 4.9, -4.1, 4.9, 4.9, 4.9, -4.1, 4.9, 4.9, 4.9, -4.1]
    f1 = [4.9, 4.9, 4.9, 5.9, 4.9, 4.9, 4.9, 5.9, 4.9, 4.9, 4.9, 5.9]
    c2 = [-4.9, -4.9, -4.9, 4.1, -4.9, -4.9, -4.9, 4.1, -4.9, -4.9, -4.9, 4.1]
    f2 = [-4.9, -4.9, -4.9, -5.9, -4.9, -4.9, -4.9, -5.9, -4.9, -4.9, -4.9, -5.9]
    c3 = [-5.1, -5.1, -5.1, 5.1, 5.1, 5.1]
    f3 = [4.9, 4.9, 4.9, -4.9, -4.9, -4.9]
    for c, f, nat in ((c1, f1, [4]), (c2, f2, [4]), (c3, f3, [1, 1])):
        np.testing.assert_allclose(oracle.fix_monomer_coordinates(c, CUBIC, nat), f, rtol=0, atol=TOL)


def test_fix_monomer_coordinates_kat_hexagonal():
    """unittest-systools.cpp:304-369 (A = B = 10, C = 20, gamma = 60 degrees)."""
    v1, v2, v3 = HEXAGONAL[0:3], HEXAGONAL[3:6], HEXAGONAL[6:9]
    seven = [0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, -1, 0, 0, 0, -1, 0, 0, 0, -1]
    s1 = _shift(seven, (0.51, v1), (0.51, v2), (0.51, v3))
    s1f = _shift(seven, (-0.49, v1), (-0.49, v2), (-0.49, v3))
    s2 = _shift(seven, (-0.51, v1), (-0.51, v2), (-0.51, v3))
    s2f = _shift(seven, (0.49, v1), (0.49, v2), (0.49, v3))
    m1 = _shift(seven, (-2.51, v1), (-2.51, v2))
    m1f = _shift(seven, (0.49, v1), (0.49, v2))
    m2 = _shift([0, 0, 0, -1, 1, 1], (2.51, v1), (2.51, v2), (-0.51, v3))
    m2f = _shift([0, 0, 0, -1, 1, 1], (-0.49, v1), (-0.49, v2), (0.49, v3))
    cases = [(s1, s1f, [7]), (s2, s2f, [7]), (np.concatenate([m1, m2]), np.concatenate([m1f, m2f]), [7, 2])]
    for c, f, nat in cases:
        np.testing.assert_allclose(oracle.fix_monomer_coordinates(c, HEXAGONAL, nat), f, rtol=0, atol=TOL)


def _image_cases(box):
    """The monoatomic, polyatomic and two-dimer cases of unittest-systools.cpp:412-589 for one
    box: (m1, m2 per dimer, expected m2 per dimer)."""
    v1, v2, v3 = box[0:3], box[3:6], box[6:9]
    lo = ((-0.49, v1), (-0.49, v2), (-0.49, v3))
    far = ((1.51, v1), (3.51, v2), (2.51, v3))
    out = []
    for b1, b2 in (([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]),
                   ([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0], [2.0, 2.0, 2.0, 2.0, 3.0, 2.0])):
        out.append((_shift(b1, *lo), _shift(b2, *far), _shift(b2, *lo)))
    return out


@pytest.mark.parametrize("box", [CUBIC, TRICLINIC], ids=["cubic", "triclinic"])
def test_close_dimer_image_kat(box):
    for m1, m2, want in _image_cases(box):
        np.testing.assert_allclose(oracle.close_image(box, m1, m2), want, rtol=0, atol=TOL)
        # "multiple dimers" (:474-498): the same dimer twice, each imaged on its own
        for _ in range(2):
            np.testing.assert_allclose(oracle.close_image(box, m1, m2), want, rtol=0, atol=TOL)


@pytest.mark.parametrize("box", [CUBIC, TRICLINIC], ids=["cubic", "triclinic"])
def test_close_trimer_image_kat(box):
    """unittest-systools.cpp:591-832: GetCloseTrimerImage images monomers 2 and 3 against monomer 1
    (sys_tools.cpp:438-442)."""
    v1, v2, v3 = box[0:3], box[3:6], box[6:9]
    lo = ((-0.49, v1), (-0.49, v2), (-0.49, v3))
    far = ((1.51, v1), (3.51, v2), (2.51, v3))
    near = ((-0.51, v1), (-0.51, v2), (-0.51, v3))
    neg = ((-1.51, v1), (-3.51, v2), (-2.51, v3))
    for b1, b2, b3 in (([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [-1.0, -1.0, -1.0]),
                       ([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0], [2.0, 2.0, 2.0, 2.0, 3.0, 2.0],
                        [1.0, 1.0, 1.0, 2.0, 3.0, 2.0, 2.0, 2.0, 3.0, 3.0, 2.0, 2.0])):
        m1 = _shift(b1, *lo)
        np.testing.assert_allclose(oracle.close_image(box, m1, _shift(b2, *far)), _shift(b2, *lo), rtol=0, atol=TOL)
        np.testing.assert_allclose(oracle.close_image(box, m1, _shift(b3, *neg)), _shift(b3, *near), rtol=0,
                                   atol=TOL)


# ---------------------------------------------------------------------------------------------
# AddClusters


def _cube27():
    """unittest-systools.cpp:855-885: a monoatomic monomer at every point of a 3 x 3 x 3 grid of
    spacing 1 (corners, edge midpoints, face centres, centre); atom tags i + 1."""
    pts = [[i - 1.0, j - 1.0, k - 1.0] for i in range(3) for j in range(3) for k in range(3)]
    return np.array(pts).reshape(-1), np.arange(1, 28, dtype=np.int32)


def _expected_counts(xyz, cutoff):
    """The test's own expectation (:924-945, :993-1030): pairs / triples by the L1 distance,
    which on this grid selects the same pairs as the Euclidean one below 1.01."""
    p = xyz.reshape(-1, 3)
    n = len(p)
    d = np.abs(p[:, None, :] - p[None, :, :]).sum(-1) < cutoff
    nd = sum(d[i, j] for i in range(n) for j in range(i + 1, n))
    nt = 0
    for i in range(n - 2):
        for j in range(i + 1, n - 1):
            for k in range(j + 1, n):
                a, b, c = d[i, j], d[i, k], d[j, k]
                nt += (a and b) or (a and c) or (b and c)
    return nd, nt


@pytest.mark.parametrize("cutoff", [0.5, 10.0, 1.01])
def test_add_clusters_counts_kat(cutoff):
    xyz, tags = _cube27()
    n = 27
    d = oracle.clusters(2, cutoff, xyz, tags=tags)
    t = oracle.clusters(3, cutoff, xyz, tags=tags)
    if cutoff == 0.5:
        assert len(d) == 0 and len(t) == 0
    elif cutoff == 10.0:
        assert len(d) == n * (n - 1) // 2 and len(t) == n * (n - 1) * (n - 2) // 6
    else:
        nd, nt = _expected_counts(xyz, cutoff)
        assert (len(d), len(t)) == (nd, nt)
    if oracle.ref_available():
        assert np.array_equal(sorted_tuples(d), sorted_tuples(oracle.ref_add_clusters(2, cutoff, xyz)))
        assert np.array_equal(sorted_tuples(t), sorted_tuples(oracle.ref_add_clusters(3, cutoff, xyz)))


def _pair_list_system():
    """The first system of the reference's input_h2o_co2_ch4_dp1 NRG file: first atoms in
    internal order (OrderMonomers: ch4, dp1, co2, h2o), the internal -> input index map and the
    input-order atom tags of the first atoms."""
    from nrg_files import read_first_system

    mons = read_first_system(os.path.join(GOLDEN, "input_h2o_co2_ch4_dp1.nrg"))
    names = [m[0] for m in mons]
    counts = {}
    for nm in names:
        counts[nm] = counts.get(nm, 0) + 1
    types = list(dict.fromkeys(names))
    order = []
    while types:
        best = min(types, key=lambda t: (counts[t], types.index(t)))
        order += [i for i, nm in enumerate(names) if nm == best]
        types.remove(best)
    first_atom = np.cumsum([0] + [len(m[1]) for m in mons])[:-1]
    first = np.array([mons[i][2][:3] for i in order], np.float64).reshape(-1)
    tags = np.array([first_atom[i] for i in order], np.int32)
    return first, np.array(order), tags


@pytest.mark.parametrize("which", [0, 1])
def test_pair_list_sets_kat(which):
    """unittest-system.cpp:644-718: GetPairList(2|3, 10.0 / 5.0, 0, nmon) of the gas-phase system."""
    k = KATS["pair_list"]
    cutoff = k["cutoffs"][which]
    want_d = np.array(k["expected_d" if which == 0 else "expected_d2"]).reshape(-1, 2)
    want_t = np.array(k["expected_t" if which == 0 else "expected_t2"]).reshape(-1, 3)
    first, order, tags = _pair_list_system()
    d = order[oracle.clusters(2, cutoff, first, tags=tags)]
    t = order[oracle.clusters(3, cutoff, first, tags=tags)]
    assert np.array_equal(sorted_tuples(d), sorted_tuples(want_d))
    assert np.array_equal(sorted_tuples(t), sorted_tuples(want_t))


# ---------------------------------------------------------------------------------------------
# The water M-site of GetXyz (SetVSites), through the library's C ABI


def test_get_xyz_msite_kat():
    """unittest-systools.cpp:1478-1505: the M-site of one water, via initialize_system_py_ /
    get_xyz_ (gas phase: no device is touched)."""
    from mbx_amd import engine

    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libmbxg.so not built")
    w = [9.9259685460e-01, 2.1439525102e+00, 1.6639051316e+00, 8.0501784870e-01, 2.7574021582e+00,
         2.3791378598e+00, 1.6203025410e+00, 1.4960194683e+00, 2.0291893372e+00]
    want = w + [1.0864993964e+00, 2.1365953595e+00, 1.8944371375e+00]
    L = engine.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    cpp = ctypes.POINTER(ctypes.c_char_p)
    L.initialize_system_py_.argtypes = [dp, ip, cpp, cpp, ip, ctypes.c_char_p]
    L.get_xyz_.argtypes = [dp]
    c = np.array(w)
    L.initialize_system_py_(c.ctypes.data_as(dp), (ctypes.c_int * 1)(3), (ctypes.c_char_p * 3)(b"O", b"H", b"H"),
                            (ctypes.c_char_p * 1)(b"h2o"), ctypes.byref(ctypes.c_int(1)), b"")
    out = np.zeros(12)
    L.get_xyz_(out.ctypes.data_as(dp))
    L.finalize_system_()
    np.testing.assert_allclose(out, want, rtol=0, atol=TOL)
