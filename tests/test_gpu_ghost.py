"""GPU parity of the local/ghost decomposition (mbxg_set_use_ghost; the reference's use_ghost
argument of TwoBodyEnergy / ThreeBodyEnergy, reaching systools::AddClusters,
sys_tools.cpp:472-537): lists against the oracle's AddClusters restatement (itself pinned to the
compiled reference in test_oracle.py), and energies / gradients against the oracle's MB-pol
holders summed over those lists (each n-mer evaluated once, ghost atoms receiving gradients)."""
import numpy as np
import pytest

from conftest import sorted_tuples

import oracle
from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu


def _cluster():
    xyz, _ = configs.water_box(120, 16.0, (5, 5, 5), seed=31)
    first = xyz.reshape(-1, 9)[:, :3]
    islocal = (first[:, 0] < np.median(first[:, 0])).astype(np.uint8)
    return xyz, first, islocal


def _oracle_terms(xyz, lists, order):
    """Energy and per-atom gradient of the listed n-mers by the oracle's MB-pol holders."""
    w = xyz.reshape(-1, 9)
    x = np.concatenate([w[lists[:, k]] for k in range(order)], axis=1)
    e, g = oracle.mbpol_nmers("mbpol_2b" if order == 2 else "mbpol_3b", x)
    grad = np.zeros_like(w)
    for k in range(order):
        np.add.at(grad, lists[:, k], g[:, 9 * k:9 * k + 9])
    return e.sum(), grad.reshape(-1)


@pytest.mark.parametrize("graph", [False, True])
def test_use_ghost_lists_energies_gradients(graph):
    xyz, first, islocal = _cluster()
    c2, c3 = 9.0, 6.0
    print("use_ghost, graph", graph, flush=True)
    e = engine.Engine()
    e.set_topology([engine.H2O] * len(islocal), islocal=islocal)
    e.set_box(None)
    e.set_cutoffs(c2, c3)
    e.set_graph(graph)
    e.set_use_ghost(True)
    e.set_coords(xyz)
    for _ in range(3):  # synchronised, then steady state (one-pass lists; captured and replayed in graph mode)
        e.compute()
        d, t = e.lists()
        want_d = oracle.clusters(2, c2, first, None, islocal=islocal, use_ghost=True)
        want_t = oracle.clusters(3, c3, first, None, islocal=islocal, use_ghost=True)
        assert np.array_equal(sorted_tuples(d), sorted_tuples(want_d))
        assert np.array_equal(sorted_tuples(t), sorted_tuples(want_t))
        assert np.any(islocal[d] == 0) and np.any(islocal[t] == 0)
        e2, e3, g, _ = e.results()
        r2, g2 = _oracle_terms(xyz, want_d, 2)
        r3, g3 = _oracle_terms(xyz, want_t, 3)
        assert e2 == pytest.approx(r2, rel=1e-10) and e3 == pytest.approx(r3, rel=1e-10)
        assert np.abs(g - (g2 + g3)).max() <= 1e-8
        e.set_coords(xyz)
    # back to the ghost-free search: the lists only hold local monomers
    e.set_use_ghost(False)
    e.compute()
    d, t = e.lists()
    assert np.all(islocal[d] == 1) and np.all(islocal[t] == 1)
    loc = np.flatnonzero(islocal)
    ld = oracle.clusters(2, c2, first[loc], None)
    assert np.array_equal(sorted_tuples(d), sorted_tuples(loc[ld]))
