"""The tiled routing layout (mhmhip_plan_network, the layout bind_network
builds), checked on the host against the network it came from.

The routing kernel trusts these tables for every index it forms, so each
invariant it relies on is checked here:
  * device indices are a permutation of the nodes;
  * every node is the real slot of exactly one tile, at its device index;
  * a slot's upstream slots are contiguous inside its bin, one per upstream
    node in the reference's routing rank (netPerm order), either the node
    itself (same tile, one level deeper: lag - 1) or a proxy for it;
  * every proxy has exactly one tile root with its boundary row, whose tile
    lies one tile-level deeper; domain sinks head depth-0 tiles;
  * tiles stay within 256 slots and 16 levels (lag 0..15);
  * the tile depth grows by one per tile boundary along every path;
  * the longest chain of tiles is as short as the height limit allows.
Networks: synthetic D8 basins and hand-made edge cases (single node, a
chain longer than the height limit, a 15-child star, several outlets).
"""
import numpy as np
import pytest

from mhm_amd import _native as N, synth

SLOTS, HEIGHT = 256, 16
SF_REAL, SF_PROXY, SF_TROOT = 1 << 16, 1 << 17, 1 << 18


class Net:
    def __init__(self, n, from_n, to_n):
        self.n_cells = n
        self.fromN = np.asarray(from_n, np.int32)
        self.toN = np.asarray(to_n, np.int32)
        _, perm = N.routing_order(self.fromN, self.toN, n)
        self.netPerm = perm[:len(from_n)]
        self.n_outlets = n - len(from_n)


def _networks():
    out = [("d8_40x40", synth.make_domain(40, 40, seed=4)),
           ("d8_97x31", synth.make_domain(97, 31, seed=9)),
           ("single", Net(1, [], []))]
    # chain of 100 nodes: node i+1 drains into node i (1-based), sink 1
    out.append(("chain100", Net(100, np.arange(2, 101), np.arange(1, 100))))
    # star: 15 leaves into node 1
    out.append(("star15", Net(16, np.arange(2, 17), np.ones(15, int))))
    # two trees: 0..599 chain-ish binary tree, 600..999 random forest with 3 sinks
    rng = np.random.default_rng(2)
    fr, to = [], []
    for v in range(2, 601):
        fr.append(v)
        to.append(v // 2)
    for v in range(604, 1001):
        fr.append(v)
        to.append(int(rng.integers(601, v)))
    out.append(("forest", Net(1000, fr, to)))
    return out


@pytest.mark.parametrize("name,net", _networks(), ids=[n for n, _ in _networks()])
def test_tile_layout_invariants(name, net):
    plan, pos, sl = N.plan_network(net, slots=True)
    n = net.n_cells
    assert sorted(pos.tolist()) == list(range(n))
    node_at = np.empty(n, np.int64)
    node_at[pos] = np.arange(n)
    # upstream nodes in routing rank (0-based node ids)
    links = net.netPerm - 1
    up = [[] for _ in range(n)]
    parent = np.full(n, -1)
    for i in links:
        up[net.toN[i] - 1].append(net.fromN[i] - 1)
        parent[net.fromN[i] - 1] = net.toN[i] - 1
    depth = np.zeros(n, np.int64)
    for v in range(n):  # depth by walking (small networks)
        u, k = v, 0
        while parent[u] >= 0:
            u, k = parent[u], k + 1
        depth[v] = k
    assert plan.dmax == depth.max()
    assert plan.n_slots == plan.n_bins * SLOTS and sl.shape == (plan.n_slots, 4)

    p, bnd, a, b = sl[:, 0], sl[:, 1], sl[:, 2], sl[:, 3]
    real = (a & SF_REAL) != 0
    proxy = (a & SF_PROXY) != 0
    troot = (a & SF_TROOT) != 0
    assert not np.any(real & proxy)
    assert sorted(p[real].tolist()) == list(range(n))
    assert np.all(p[~real] == -1)
    slot_of_node = np.empty(n, np.int64)
    slot_of_node[node_at[p[real]]] = np.where(real)[0]
    lag = a & 0xFFF
    nch = (a >> 12) & 0xF
    c0 = b & 0xFFFF
    tau = b >> 16
    assert lag.max() < HEIGHT
    assert tau.max() == plan.tau_max
    for s in np.where(real)[0]:
        v = node_at[p[s]]
        assert nch[s] == len(up[v])
        binbase = (s // SLOTS) * SLOTS
        for k, c in enumerate(up[v]):
            cs = binbase + c0[s] + k
            assert binbase <= cs < binbase + SLOTS
            assert lag[cs] == lag[s] - 1 and tau[cs] == tau[s]
            if real[cs]:
                assert node_at[p[cs]] == c
            else:
                assert proxy[cs]
                rs = slot_of_node[c]  # the upstream tile's root slot
                assert troot[rs] and bnd[rs] == bnd[cs] and tau[rs] == tau[s] + 1
        if parent[v] < 0:
            assert tau[s] == 0 and not troot[s]
    # boundary rows: one proxy and one tile root each
    rows_p = np.sort(bnd[proxy])
    rows_r = np.sort(bnd[troot])
    assert np.array_equal(rows_p, np.arange(plan.n_bnd))
    assert np.array_equal(rows_r, np.arange(plan.n_bnd))
    assert plan.n_tiles == plan.n_bnd + net.n_outlets
    # tile depth along every path = tile boundaries crossed
    for v in range(n):
        if parent[v] >= 0:
            sv, sp = slot_of_node[v], slot_of_node[parent[v]]
            assert tau[sv] == tau[sp] + (1 if troot[sv] else 0)


def test_deep_chain_is_cut_by_height():
    plan, _, sl = N.plan_network(Net(100, np.arange(2, 101), np.arange(1, 100)), slots=True)
    assert plan.tau_max == (100 - 1) // HEIGHT
    assert plan.n_tiles == -(-100 // HEIGHT)


@pytest.mark.parametrize("n,seed", [(40, 4), (97, 9), (256, 1)])
def test_tile_chain_is_shortest(n, seed):
    """Along the deepest path every tile but the top one spends a level on the
    proxy of the next tile up, so no partition has fewer than
    ceil((dmax + 1 - H) / (H - 1)) tile boundaries on it; on D8 basins the
    slot limit never binds and the partition meets this bound."""
    plan, _, _ = N.plan_network(synth.make_domain(n, n, seed=seed))
    bound = -(-max(0, plan.dmax + 1 - HEIGHT) // (HEIGHT - 1))
    assert plan.tau_max == bound


@pytest.mark.parametrize("name,net", _networks(), ids=[n for n, _ in _networks()])
def test_tile_chain_lower_bound(name, net):
    plan, _, _ = N.plan_network(net)
    assert plan.tau_max >= -(-max(0, plan.dmax + 1 - HEIGHT) // (HEIGHT - 1))
