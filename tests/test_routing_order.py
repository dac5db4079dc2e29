"""mhmhip_routing_order (O(n log n)) against the oracle's statement-by-statement
restatement of L11_routing_order (src/mRM/mo_mrm_net_startup.f90:728-859,
O(n^2) sweeps), on synthetic D8 networks and random trees.  Bit-exact."""
import numpy as np
import pytest

import oracle_lib as O
from mhm_amd import _native as N, synth


def _oracle_order(fromN, toN, n_nodes):
    nl = fromN.shape[0]
    fr = np.zeros(n_nodes, np.int32)
    to = np.zeros(n_nodes, np.int32)
    fr[:nl] = fromN
    to[:nl] = toN
    ro = np.zeros(n_nodes, np.int32)
    pm = np.zeros(n_nodes, np.int32)
    O.lib().or_routing_order(n_nodes, nl, O.iptr(fr), O.iptr(to), O.iptr(ro), O.iptr(pm))
    return ro[:nl], pm[:nl]


@pytest.mark.parametrize("shape,seed", [((8, 8), 3), ((16, 16), 4), ((40, 70), 9), ((64, 64), 1),
                                        ((1, 30), 5), ((30, 1), 6)])
def test_d8_networks(shape, seed):
    d = synth.make_domain(*shape, seed=seed)
    ro, pm = _oracle_order(d.fromN, d.toN, d.n_cells)
    assert np.array_equal(ro, d.rOrder)
    assert np.array_equal(pm, d.netPerm)


@pytest.mark.parametrize("seed", range(12))
def test_random_forests(seed):
    """Random trees with arbitrary node numbering and several outlets."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    n_out = int(rng.integers(1, 4))
    order = rng.permutation(n)
    parent = np.full(n, -1)
    for i in range(n_out, n):  # attach each node to an earlier one in `order`
        parent[order[i]] = order[rng.integers(0, i)]
    nodes = [k for k in range(n) if parent[k] >= 0]  # link per non-outlet node, in node order
    fromN = np.array([k + 1 for k in nodes], np.int32)
    toN = np.array([parent[k] + 1 for k in nodes], np.int32)
    ro, pm = _oracle_order(fromN, toN, n)
    ro2, pm2 = N.routing_order(fromN, toN, n)
    nl = fromN.shape[0]
    assert np.array_equal(ro, ro2[:nl])
    assert np.array_equal(pm, pm2[:nl])


def test_refined_domain_maps_every_l1_cell_to_its_coarse_node():
    """The synthetic basin with a routing grid f times coarser than L1 (as in
    the reference's check case_04): each L11 node holds the f x f L1 cells of
    its block, column-major numbering on both grids, and the total area is
    kept."""
    from mhm_amd import synth
    d11 = synth.make_domain(5, 6, seed=3)
    for f in (2, 3):
        d1 = synth.refine_domain(d11, f, 4)
        assert d1.n_cells == f * f * d11.n_cells and d1.n_nodes == d11.n_cells
        counts = np.bincount(d1.l1_l11 - 1, minlength=d11.n_cells)
        assert np.all(counts == f * f)
        # L1 cell (i, j) lies in node (i // f, j // f)
        k = np.arange(d1.n_cells)
        i, j = k % d1.nrows, k // d1.nrows
        assert np.array_equal(d1.l1_l11 - 1, (j // f) * d11.nrows + i // f)
        assert d1.area_km2.sum() == pytest.approx(d11.area_km2.sum(), rel=1e-12)
        assert np.array_equal(d1.fromN, d11.fromN) and np.array_equal(d1.gauge_nodes, d11.gauge_nodes)
