"""The one-pass list build (count pass that stages each family's evaluated n-mers per query and
wavefront, scan, compaction into the evaluate lists), which every unsynchronised compute that
does not ask for the listed lists takes, against the two-pass build of a synchronised compute:
identical evaluate lists (so bit-identical energies; gradients and virials to rounding, their
atomic accumulation order varies), listed lists written on demand by mbxg_get_lists, in eager
and graph mode.  Staging overflow (a compute with
more n-mers per wavefront than the last synchronised one) is covered by
test_gpu_parity.py::test_lists_outgrow_buffers."""
import numpy as np
import pytest

from conftest import sorted_tuples

import oracle
from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu


def _state(e):
    e2, e3, g, v = e.results()
    d, t = e.lists()
    return e2, e3, g, v, d, t


def _same(a, b):
    assert a[0] == b[0] and a[1] == b[1]
    np.testing.assert_allclose(a[2], b[2], rtol=0, atol=1e-11)
    np.testing.assert_allclose(a[3], b[3], rtol=1e-12, atol=1e-10)
    assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


@pytest.mark.parametrize("graph", [False, True])
def test_one_pass_equals_two_pass(graph):
    xyz, box, c2, c3 = configs.build(configs.C1)
    rng = np.random.default_rng(7)
    moved = xyz + rng.normal(0.0, 0.05, xyz.shape)
    print("one-pass build, graph", graph, flush=True)
    e = engine.water_engine(configs.C1["nmon"], box, c2, c3)
    e.set_graph(graph)
    e.set_coords(xyz)
    e.compute()  # synchronised: two passes, sizes the staging
    first = _state(e)
    # a fresh context's synchronised compute of the moved coordinates is the two-pass answer
    f = engine.water_engine(configs.C1["nmon"], box, c2, c3)
    f.set_coords(moved)
    f.compute()
    ref = _state(f)
    for k in range(3):  # steady state: one pass (graph mode: captured, then replayed)
        e.set_coords(moved)
        e.compute()
        _same(_state(e), ref)
        # results read without the lists first, then the lists
        e.set_coords(xyz)
        e.compute()
        e2, e3, g, v = e.results()
        assert e2 == first[0] and e3 == first[1]
        np.testing.assert_allclose(g, first[2], rtol=0, atol=1e-11)
        d, t = e.lists()
        assert np.array_equal(d, first[4]) and np.array_equal(t, first[5])
    # and the lists are the oracle's
    fx = oracle.fix_first(xyz, box)
    assert np.array_equal(sorted_tuples(first[4]), sorted_tuples(oracle.clusters(2, c2, fx, box)))
    assert np.array_equal(sorted_tuples(first[5]), sorted_tuples(oracle.clusters(3, c3, fx, box)))


def test_paired_polynomials_match_separate_launches():
    """The MB-pol 3B and 2B polynomials in one launch (fuse_poly = 1, the default) against
    separate launches: the same per-n-mer schedules, so energies and gradients agree to rounding
    (the 2B parts are summed over split_3b wavefronts instead of split_2b: gradients of
    cancelling terms differ by a few 1e-10); both meet the oracle's bar."""
    xyz, box, c2, c3 = configs.build(configs.C1)
    out = {}
    for fuse in (1, 0):
        print("fuse_poly", fuse, flush=True)
        e = engine.water_engine(configs.C1["nmon"], box, c2, c3)
        e.set_tuning("fuse_poly", fuse)
        assert e.tuning("fuse_poly") == fuse
        e.set_coords(xyz)
        e.compute()
        e.set_timing(True)
        e.compute()
        tm = e.timings()
        assert (tm["poly_2b"] == 0.0) == (fuse == 1) and tm["poly_3b"] > 0.0
        out[fuse] = e.results()
    (a2, a3, ag, av), (b2, b3, bg, bv) = out[1], out[0]
    assert a2 == pytest.approx(b2, rel=1e-12) and a3 == pytest.approx(b3, rel=1e-12)
    np.testing.assert_allclose(ag, bg, rtol=0, atol=5e-9)
    np.testing.assert_allclose(av, bv, rtol=1e-10, atol=1e-8)
    r = oracle.water_system(xyz, box, c2, c3)
    for e2, e3, g in ((a2, a3, ag), (b2, b3, bg)):
        assert e2 == pytest.approx(r["e2b"], rel=1e-10) and e3 == pytest.approx(r["e3b"], rel=1e-10)
        assert np.abs(g - r["grad"]).max() <= 1e-8


@pytest.mark.parametrize("case", ["box", "droplet", "clathrate"])
def test_fused_prep_matches_separate_launches(systems, case):
    """FixMonomerCoordinates, the cloud and its binning in one single-workgroup launch
    (fuse_prep = 1, the default for clouds of at most 65536 slots) against the separate launches:
    the same working coordinates bit for bit, the same lists in the same order (each query's
    matches are sorted, so the slots' order inside a cell does not matter) and so bit-identical
    energies; gradients to rounding.  Eager and graph replays."""
    from conftest import water_case

    if case == "clathrate":
        xyz, box = systems["clathrate/xyz"], systems["clathrate/box"]
        kinds = [engine.H2O if m == "h2o" else engine.CH4 for m in systems["clathrate/mon"]]
        c2, c3 = 9.0, 6.0
    else:
        xyz, box, c2, c3, _ = water_case(systems, "h2o_2048_cubic" if case == "box" else "h2o_2048_droplet")
        kinds = [engine.H2O] * (len(xyz) // 9)
    out = {}
    for fuse in (1, 0):
        print("fuse_prep", fuse, case, flush=True)
        e = engine.Engine()
        e.set_topology(kinds)
        e.set_box(box)
        e.set_cutoffs(c2, c3)
        e.set_tuning("fuse_prep", fuse)
        assert e.tuning("fuse_prep") == fuse
        e.set_graph(True)
        e.set_coords(xyz)
        for _ in range(3):  # sizing compute, eager, graph replay
            e.compute()
        out[fuse] = _state(e) + (e.coords(),)
    a, b = out[1], out[0]
    _same(a[:6], b[:6])
    assert np.array_equal(a[6], b[6])


@pytest.mark.parametrize("split", [1, 2, 4, 8])
def test_fused_chain_matches_separate_kernel(systems, split):
    """The MB-pol 3B chain rule in the paired polynomial launch's 3B blocks (fuse_chain = 1, the
    default) against its own kernel (fuse_chain = 0), for every split of the 64-trimer blocks:
    bit-identical energies (the same per-trimer arithmetic and the same records of 64 trimers),
    gradients and virial to the rounding of the atomics' order.  Eager and graph replays; and the
    oracle's energies and gradients."""
    from conftest import water_case

    xyz, box, c2, c3, _ = water_case(systems, "h2o_2048_cubic")
    kinds = [engine.H2O] * (len(xyz) // 9)
    out = {}
    for fuse in (1, 0):
        print("fuse_chain", fuse, "split", split, flush=True)
        e = engine.Engine()
        e.set_topology(kinds)
        e.set_box(box)
        e.set_cutoffs(c2, c3)
        e.set_tuning("split_2b", split)
        e.set_tuning("split_3b", split)
        e.set_tuning("fuse_chain", fuse)
        assert e.tuning("fuse_chain") == fuse
        e.set_graph(True)
        e.set_coords(xyz)
        for _ in range(3):  # sizing compute, eager, graph replay
            e.compute()
        out[fuse] = _state(e)
    _same(out[1], out[0])
    if split == 4:
        r = oracle.water_system(xyz, box, c2, c3)
        e2, e3, g, _ = out[1][:4]
        assert e2 == pytest.approx(r["e2b"], rel=1e-10) and e3 == pytest.approx(r["e3b"], rel=1e-10)
        assert np.abs(g - r["grad"]).max() <= 1e-8
