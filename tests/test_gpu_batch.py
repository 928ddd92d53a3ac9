"""GPU parity of batches: many configurations in one compute (mbxg_set_batch).

Each configuration's energies, virial and gradient must equal the oracle's for that
configuration alone (energies 1e-10 relative, gradients 1e-8 kcal/mol/A), its lists must be
exactly its own lists, and the batched result must equal plain per-configuration computes on
the GPU to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import sorted_tuples

import oracle
from mbx_amd import configs, engine

pytestmark = pytest.mark.gpu

E_REL = 1e-10
G_ABS = 1e-8


def _batch_engine(parts, box, c2, c3):
    sizes = [len(x) // 9 for x in parts]
    off = np.concatenate([[0], np.cumsum(sizes)])
    e = engine.water_engine(int(off[-1]), box, c2, c3)
    e.set_batch(off)
    e.set_coords(np.concatenate(parts))
    e.compute()
    return e, off


def close(got, want, scale):
    """Energies are sums of n-mer energies, each a sum of strongly cancelling monomials; a small
    cluster's total can be many orders below its terms (a 100-water cluster: E3B = -0.14 from
    sum |a_k M_k| = 1e7).  No FP64 evaluation order is better than ~eps * scale there, so the
    bar is 1e-10 relative to the total or 1e-16 * scale (just under one unit roundoff of the
    terms' magnitude), whichever is looser; scale from oracle.water_conditioning."""
    return abs(got - want) <= max(E_REL * abs(want), 1e-16 * scale, 1e-13)


def _check(parts, box, c2, c3):
    e, off = _batch_engine(parts, box, c2, c3)
    e2, e3, vir = e.batch_results()
    _, _, g, _ = e.results()
    d, t = e.lists()
    want_d, want_t = [], []
    for k, x in enumerate(parts):
        r = oracle.water_system(x, box, c2, c3)
        s2, s3 = oracle.water_conditioning(x, box, c2, c3)
        assert close(e2[k], r["e2b"], s2), (k, e2[k], r["e2b"], s2)
        assert close(e3[k], r["e3b"], s3), (k, e3[k], r["e3b"], s3)
        gk = g[9 * off[k]:9 * off[k + 1]]
        assert np.abs(gk - r["grad"]).max() <= G_ABS, k
        np.testing.assert_allclose(vir[k], r["virial"], rtol=1e-8, atol=1e-8)
        first = oracle.fix_first(x, box)
        want_d.append(oracle.clusters(2, c2, first, box) + off[k])
        want_t.append(oracle.clusters(3, c3, first, box) + off[k])
    assert np.array_equal(sorted_tuples(d), sorted_tuples(np.concatenate(want_d)))
    assert np.array_equal(sorted_tuples(t), sorted_tuples(np.concatenate(want_t)))
    return e, off


def test_batch_gas_clusters():
    parts, sizes = configs.cluster_batch(24, nmin=16, nmax=160, seed=7)
    print("batch gas", list(sizes), flush=True)
    _check(parts, None, 9.0, 7.0)


def test_batch_periodic_beads():
    """Beads of one periodic box: the centroid plus N(0, 0.05 A) per atom, same box."""
    xyz, box, c2, c3 = configs.build(configs.C1)
    rng = np.random.default_rng(3)
    parts = [xyz + rng.normal(0.0, 0.05, xyz.shape) for _ in range(4)]
    print("batch beads", flush=True)
    _check(parts, box, c2, c3)


def test_batch_equals_plain():
    parts, sizes = configs.cluster_batch(16, nmin=16, nmax=128, seed=8)
    print("batch vs plain", list(sizes), flush=True)
    e, off = _batch_engine(parts, None, 9.0, 7.0)
    e2, e3, vir = e.batch_results()
    _, _, g, _ = e.results()
    for k, x in enumerate(parts):
        p = engine.water_engine(len(x) // 9, None, 9.0, 7.0)
        p.set_coords(x)
        p.compute()
        pe2, pe3, pg, pv = p.results()
        # same kernels, different summation order of the per-n-mer energies
        s2 = np.abs(p.nmer_energies(2)[1]).sum()
        s3 = np.abs(p.nmer_energies(3)[1]).sum()
        assert abs(e2[k] - pe2) <= max(1e-12 * abs(pe2), 1e-14 * s2, 1e-14)
        assert abs(e3[k] - pe3) <= max(1e-12 * abs(pe3), 1e-14 * s3, 1e-14)
        assert np.abs(g[9 * off[k]:9 * off[k + 1]] - pg).max() <= 1e-10
        np.testing.assert_allclose(vir[k], pv, rtol=1e-10, atol=1e-10)
    # totals of the batch = sum over configurations
    t2, t3, _, _ = e.results()
    assert t2 == pytest.approx(e2.sum(), rel=1e-11)
    assert t3 == pytest.approx(e3.sum(), rel=1e-11)


def test_batch_reset_to_single():
    parts, _ = configs.cluster_batch(3, nmin=16, nmax=40, seed=9)
    e, off = _batch_engine(parts, None, 9.0, 7.0)
    e.set_batch([0, int(off[-1])])
    e.compute()
    x = np.concatenate(parts)
    r = oracle.water_system(x, None, 9.0, 7.0)
    e2, e3, _, _ = e.results()
    assert e2 == pytest.approx(r["e2b"], rel=E_REL)
    assert e3 == pytest.approx(r["e3b"], rel=E_REL)


def test_batch_pimd_mixed_beads():
    """Beads of a periodic CH4 / H2O box in one batch: every bead's 2B / 3B energies and
    gradients (all five families) equal the oracle's for that bead alone."""
    xyz, kinds, box = configs.mixed_box()
    parts = configs.beads(xyz, 2)
    print("pimd beads", flush=True)
    n = len(kinds)
    e = engine.Engine()
    e.set_topology(np.concatenate([kinds, kinds]))
    e.set_box(box)
    e.set_cutoffs(9.0, 7.0)
    e.set_batch([0, n, 2 * n])
    e.set_coords(np.concatenate(parts))
    e.compute()
    e2, e3, vir = e.batch_results()
    _, _, g, _ = e.results()
    na = len(xyz)
    for k, x in enumerate(parts):
        r = oracle.system(x, kinds, box, 9.0, 7.0, tags=np.arange(n))
        assert e2[k] == pytest.approx(r["e2b"], rel=E_REL)
        assert e3[k] == pytest.approx(r["e3b"], rel=E_REL)
        assert np.abs(g[k * na:(k + 1) * na] - r["grad"]).max() <= G_ABS
        np.testing.assert_allclose(vir[k], r["virial"], rtol=1e-8, atol=1e-8)
