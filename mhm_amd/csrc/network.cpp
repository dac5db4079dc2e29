// network.cpp -- host-side network preparation.
//
// mhmhip_routing_order reproduces L11_routing_order
// (src/mRM/mo_mrm_net_startup.f90:728-859) without its O(nLinks^2) sweeps.
// The reference numbers headwater links first, in link-index order, then
// sweeps the links in index order again and again, numbering a link as soon
// as every upstream link carries a number -- including links numbered
// earlier in the same sweep.  Here a link becomes "ready" when its last
// upstream link is numbered; if that happened at a smaller index position
// of the current sweep it is still numbered in this sweep (it lies ahead of
// the cursor), otherwise in the next one.  Two min-heaps keyed by link index
// give exactly the reference's numbering in O(nLinks log nLinks).
#include <algorithm>
#include <cstdint>
#include <functional>
#include <queue>
#include <vector>

#include "../../include/mhmhip.h"
#include "network.h"


extern "C" int mhmhip_routing_order(int64_t n_nodes, int64_t n_links, const int32_t *fromN,
                                    const int32_t *toN, int32_t *rOrder, int32_t *netPerm) {
  if (n_nodes < 1 || n_links < 0 || n_links > n_nodes || (n_links > 0 && (!fromN || !toN)) ||
      !rOrder || !netPerm)
    return MHMHIP_EINVAL;
  for (int64_t i = 0; i < n_nodes; ++i) netPerm[i] = -9999;
  for (int64_t i = 0; i < n_links; ++i) rOrder[i] = 1;
  if (n_nodes > n_links) rOrder[n_nodes - 1] = -9999;
  if (n_nodes <= 1 || n_links == 0) return MHMHIP_OK;
  // out-link of each node, pending upstream count of each link
  std::vector<int64_t> out_link(n_nodes, -1);
  for (int64_t i = 0; i < n_links; ++i) {
    const int64_t f = fromN[i] - 1, t = toN[i] - 1;
    if (f < 0 || f >= n_nodes || t < 0 || t >= n_nodes) return MHMHIP_EINVAL;
    if (out_link[f] >= 0) return MHMHIP_EINVAL;  // a node drains through one link only
    out_link[f] = i;
  }
  std::vector<int32_t> pending(n_links, 0);
  for (int64_t j = 0; j < n_links; ++j) {
    const int64_t dn = out_link[toN[j] - 1];
    if (dn >= 0 && dn != j) pending[dn]++;
  }
  using MinHeap = std::priority_queue<int64_t, std::vector<int64_t>, std::greater<int64_t>>;
  MinHeap cur, next;
  int64_t kk = 0;
  std::vector<int32_t> order(n_links, 0);
  auto release = [&](int64_t j, int64_t cursor, bool in_sweep) {
    const int64_t dn = out_link[toN[j] - 1];
    if (dn < 0 || dn == j) return;
    if (--pending[dn] == 0) {
      if (in_sweep && dn > cursor)
        cur.push(dn);
      else
        next.push(dn);
    }
  };
  // headwaters (reference lines 798-805)
  for (int64_t i = 0; i < n_links; ++i)
    if (pending[i] == 0) order[i] = (int32_t)(++kk);
  for (int64_t i = 0; i < n_links; ++i)
    if (order[i] > 0) release(i, -1, false);
  // sweeps (reference lines 808-831)
  while (!next.empty()) {
    std::swap(cur, next);
    while (!cur.empty()) {
      const int64_t i = cur.top();
      cur.pop();
      order[i] = (int32_t)(++kk);
      release(i, i, true);
    }
  }
  if (kk != n_links) return MHMHIP_EINVAL;  // a cycle: the reference would never stop
  for (int64_t i = 0; i < n_links; ++i) {
    rOrder[i] = order[i];
    netPerm[order[i] - 1] = (int32_t)(i + 1);
  }
  return MHMHIP_OK;
}

namespace mhmhip {

// Device layout of every domain's forest.  Depth = links to the sink.  The
// upstream nodes of a node are kept in the reference's routing rank, so
// their qTR sum runs in the order of L11_routing's accumulation.
//
// Tile partition, bottom-up (reverse breadth-first order): a node's open
// tile takes upstream open tiles within kTileSlots slots and
// in.tile_height levels; an upstream open tile that is closed becomes a tile of its own
// and a proxy slot in the node's tile.  Which ones are closed is chosen to
// keep the longest chain of tiles short (the rule is at the partition).
// Tiles are numbered breadth-first over the tile tree (tile depth tau
// nondecreasing), laid out breadth-first inside (upstream slots of a slot
// contiguous, in routing rank) and packed next-fit into bins of kTileSlots
// slots, lag-major inside a bin; device indices follow the real slots of the
// bins in order.
int build_layout(const LayoutIn &in, LayoutOut &out, std::string &err) {
  if (in.tile_height < 2 || in.tile_height > kTileHeightMax) {
    err = "bind_network: tile height out of range";
    return MHMHIP_EINVAL;
  }
  const int64_t nN = in.n_nodes;
  out.pos.assign(nN, -1);
  out.node_at.assign(nN, -1);
  out.child.assign(nN, 0);
  out.packed.assign(nN, 0);
  out.link_of_node.assign(nN, -1);
  out.dom_dmax.assign(in.n_domains, 0);
  out.dmax = 0;
  out.child_list.clear();
  out.slots.clear();
  out.bins.clear();
  out.tau_max = 0;
  out.bin_span = 0;
  out.n_bnd = 0;
  out.n_tiles = 0;
  int64_t link_off = 0;
  int64_t next_pos = 0;
  for (int d = 0; d < in.n_domains; ++d) {
    const int64_t n0 = in.dom_node_start[d], n1 = in.dom_node_start[d + 1], nd = n1 - n0;
    if (n0 != next_pos) {
      err = "bind_network: domain node ranges must be contiguous from 0";
      return MHMHIP_EINVAL;
    }
    const int64_t nOut = in.n_outlets[d];
    const int64_t nL = nd - nOut;
    if (nOut < 1 || nL < 0) {
      err = "bind_network: domain " + std::to_string(d) + " has an invalid outlet count";
      return MHMHIP_EINVAL;
    }
    const int32_t *perm = in.netPerm + link_off, *fr = in.fromN + link_off, *to = in.toN + link_off;
    std::vector<int64_t> rank(nL, -1);
    for (int64_t k = 0; k < nL; ++k) {
      const int64_t i = perm[k] - 1;
      if (i < 0 || i >= nL || rank[i] >= 0) {
        err = "bind_network: netPerm of domain " + std::to_string(d) + " is not a permutation";
        return MHMHIP_EINVAL;
      }
      rank[i] = k;
    }
    std::vector<int64_t> out_link(nd, -1);
    std::vector<std::vector<int64_t>> up(nd);
    for (int64_t i = 0; i < nL; ++i) {
      const int64_t f = fr[i] - 1, t = to[i] - 1;
      if (f < 0 || f >= nd || t < 0 || t >= nd || f == t || out_link[f] >= 0) {
        err = "bind_network: bad link " + std::to_string(i + 1) + " in domain " + std::to_string(d);
        return MHMHIP_EINVAL;
      }
      out_link[f] = i;
      up[t].push_back(i);
    }
    for (int64_t n = 0; n < nd; ++n) {
      std::sort(up[n].begin(), up[n].end(), [&](int64_t a, int64_t b) { return rank[a] < rank[b]; });
      if (up[n].size() > 15) {
        err = "bind_network: more than 15 upstream links at one node";
        return MHMHIP_EINVAL;
      }
    }
    // the last-ranked link's downstream node gets its own qOUT
    // (mo_mrm_routing.f90:466-467); every link's source node does too (445).
    // A domain of one node takes the single-node shortcut instead, Qmod =
    // qOUT (285-289): its node adding its own qOUT gives qTIN(IT) = 0 + qOUT,
    // the same value.
    const int64_t last_t = nL > 0 ? (int64_t)to[perm[nL - 1] - 1] - 1 : (nd == 1 ? 0 : -1);
    std::vector<int64_t> q;
    q.reserve(nd);
    std::vector<int32_t> depth(nd, -1);
    for (int64_t n = 0; n < nd; ++n)
      if (out_link[n] < 0) {
        q.push_back(n);
        depth[n] = 0;
      }
    for (size_t h = 0; h < q.size(); ++h) {
      const int64_t n = q[h];
      for (int64_t i : up[n]) {
        const int64_t c = fr[i] - 1;
        depth[c] = depth[n] + 1;
        q.push_back(c);
      }
    }
    if ((int64_t)q.size() != nd) {
      err = "bind_network: domain " + std::to_string(d) + " has a cycle or unreachable nodes";
      return MHMHIP_EINVAL;
    }
    int32_t dm = 0;
    for (int64_t n = 0; n < nd; ++n) dm = std::max(dm, depth[n]);
    if (dm >= (1 << 20)) {
      err = "bind_network: network deeper than 2^20 levels";
      return MHMHIP_EINVAL;
    }
    auto child_node = [&](int64_t i) { return (int64_t)fr[i] - 1; };
    // -- tile partition --
    // Bottom-up, each node's open tile is chosen to minimise first oc, the
    // number of tile boundaries on its longest upstream path, then oh, the
    // open tile's height in levels (a proxy slot is one level), then osize,
    // its slots.  Upstream tiles whose oc is below the largest (M) can be
    // closed without raising the node's oc: those taller than the tallest
    // oc == M tile are closed, so the height stays that tile's height plus
    // one, and the largest of them are closed while the slots overflow.  When
    // the oc == M tiles cannot stay open (height or slot limit), the node's
    // oc becomes M + 1 and every upstream tile except single leaves is
    // closed, which leaves the lowest possible height (2).  Under the height
    // limit alone this greedy is optimal: no partition has fewer tile
    // boundaries on its longest path, which is the number of launches a
    // window waits for before it reaches the outlet.
    std::vector<int32_t> osize(nd), oh(nd), oc(nd);
    std::vector<char> cut(nd, 0);
    for (int64_t h = nd - 1; h >= 0; --h) {
      const int64_t n = q[h];
      int32_t M = -1, hc = 1, sz = 1;
      for (int64_t i : up[n]) M = std::max(M, oc[child_node(i)]);
      for (int64_t i : up[n]) {
        const int64_t c = child_node(i);
        if (oc[c] == M) hc = std::max(hc, oh[c]);
      }
      bool open = M < 0 || 1 + hc <= in.tile_height;
      if (open) {
        for (int64_t i : up[n]) {
          const int64_t c = child_node(i);
          cut[c] = oc[c] < M && oh[c] > hc;
          sz += cut[c] ? 1 : osize[c];
        }
        while (sz > kTileSlots) {
          int64_t best = -1;
          for (int64_t i : up[n]) {
            const int64_t c = child_node(i);
            if (!cut[c] && oc[c] < M && (best < 0 || osize[c] > osize[best])) best = c;
          }
          if (best < 0) break;
          cut[best] = 1;
          sz += 1 - osize[best];
        }
        open = sz <= kTileSlots;
      }
      if (open) {
        oc[n] = std::max(M, 0);
        oh[n] = M < 0 ? 1 : 1 + hc;
      } else {
        sz = 1;  // 1 + nchild <= 16 slots, height 2
        for (int64_t i : up[n]) {
          const int64_t c = child_node(i);
          cut[c] = oh[c] > 1;
          sz += cut[c] ? 1 : osize[c];
        }
        oc[n] = M + 1;
        oh[n] = 2;
      }
      osize[n] = sz;
    }
    // -- tiles, breadth-first over the tile tree --
    struct TSlot {
      int64_t node;  // domain-local node
      bool proxy;
      int32_t ldepth, first, nc;
    };
    struct Tile {
      std::vector<TSlot> s;
      int32_t tau, height;
      int64_t root;
    };
    std::vector<Tile> tiles;
    std::vector<int64_t> troots;
    std::vector<int32_t> ttau;
    for (int64_t n = 0; n < nd; ++n)
      if (out_link[n] < 0) {
        troots.push_back(n);
        ttau.push_back(0);
      }
    for (size_t ti = 0; ti < troots.size(); ++ti) {
      Tile T;
      T.root = troots[ti];
      T.tau = ttau[ti];
      T.s.push_back({T.root, false, 0, 0, 0});
      for (size_t k = 0; k < T.s.size(); ++k) {
        if (T.s[k].proxy) continue;
        const int64_t n = T.s[k].node;
        const int32_t first = (int32_t)T.s.size(), ld1 = T.s[k].ldepth + 1;
        T.s[k].first = first;
        T.s[k].nc = (int32_t)up[n].size();
        for (int64_t i : up[n]) {
          const int64_t c = child_node(i);
          T.s.push_back({c, cut[c] != 0, ld1, 0, 0});
          if (cut[c]) {
            troots.push_back(c);
            ttau.push_back(T.tau + 1);
          }
        }
      }
      T.height = 0;
      for (const TSlot &u : T.s) T.height = std::max(T.height, u.ldepth + 1);
      if ((int)T.s.size() > kTileSlots || T.height > in.tile_height) {
        err = "bind_network: tile partition failed";
        return MHMHIP_EINVAL;
      }
      tiles.push_back(std::move(T));
    }
    // -- bins (next-fit) and device indices --
    std::vector<int64_t> bnd_row(nd, -1);
    for (int64_t n = 0; n < nd; ++n)
      if (cut[n]) bnd_row[n] = out.n_bnd++;
    size_t t0 = 0;
    while (t0 < tiles.size()) {
      size_t t1 = t0;
      int32_t used = 0;
      BinInfo bi{0, tiles[t0].tau, tiles[t0].tau, 0};
      while (t1 < tiles.size() && used + (int32_t)tiles[t1].s.size() <= kTileSlots) {
        used += (int32_t)tiles[t1].s.size();
        bi.height = std::max(bi.height, tiles[t1].height);
        bi.tau_lo = std::min(bi.tau_lo, tiles[t1].tau);
        bi.tau_hi = std::max(bi.tau_hi, tiles[t1].tau);
        ++t1;
      }
      const size_t sbase = out.slots.size();
      out.slots.resize(sbase + kTileSlots, SlotMeta{-1, -1, 0, 0});
      // Slots of the bin in lag-major order (largest lag first; within a
      // lag, tile order, then breadth-first order): the nodes of one lag,
      // whose qOUT sits in one ring row on every diagonal, get consecutive
      // device indices.  A slot's upstream slots (one lag lower, same tile,
      // consecutive breadth-first) stay contiguous.
      struct Ent {
        int32_t lag, t, k;
      };
      std::vector<Ent> ent;
      std::vector<std::vector<int32_t>> idx(t1 - t0);
      for (size_t t = t0; t < t1; ++t) {
        idx[t - t0].resize(tiles[t].s.size());
        for (size_t k = 0; k < tiles[t].s.size(); ++k)
          ent.push_back({tiles[t].height - 1 - tiles[t].s[k].ldepth, (int32_t)(t - t0), (int32_t)k});
      }
      std::stable_sort(ent.begin(), ent.end(), [](const Ent &a, const Ent &b) { return a.lag > b.lag; });
      for (size_t i = 0; i < ent.size(); ++i) idx[ent[i].t][ent[i].k] = (int32_t)i;
      for (size_t i = 0; i < ent.size(); ++i) {
        const Tile &T = tiles[t0 + ent[i].t];
        const size_t k = (size_t)ent[i].k;
        const TSlot &u = T.s[k];
        SlotMeta &m = out.slots[sbase + i];
        if (u.proxy) {
          m.bnd = (int32_t)bnd_row[u.node];
          m.a = ent[i].lag | SF_PROXY;
        } else {
          const int64_t p = next_pos++;
          out.pos[n0 + u.node] = p;
          m.pos = (int32_t)p;
          m.a = ent[i].lag | (u.nc << 12) | SF_REAL;
          if (k == 0 && cut[u.node]) {
            m.bnd = (int32_t)bnd_row[u.node];
            m.a |= SF_TROOT;
          }
        }
        m.b = (u.proxy || u.nc == 0 ? 0 : idx[ent[i].t][u.first]) | (T.tau << 16);
        out.tau_max = std::max(out.tau_max, T.tau);
      }
      out.bins.push_back(bi);
      out.n_tiles += (int64_t)(t1 - t0);
      t0 = t1;
    }
    // -- node metadata and child list --
    for (int64_t n = 0; n < nd; ++n) {
      const int64_t p = out.pos[n0 + n];
      out.node_at[p] = n0 + n;
      int32_t flags = 0;
      if (out_link[n] >= 0) {
        flags |= NF_LINK | NF_QOUT;
        out.link_of_node[p] = link_off + out_link[n];
      }
      if (n == last_t) flags |= NF_QOUT;
      const int32_t nc = (int32_t)up[n].size();
      out.packed[p] = depth[n] | (nc << 20) | flags;
    }
    for (int64_t p = n0; p < n1; ++p) {
      const int64_t n = out.node_at[p] - n0;
      out.child[p] = (int32_t)out.child_list.size();
      for (int64_t i : up[n]) out.child_list.push_back((int32_t)out.pos[n0 + child_node(i)]);
    }
    out.dom_dmax[d] = dm;
    out.dmax = std::max(out.dmax, dm);
    link_off += nL;
  }
  out.n_links_total = link_off;
  // bins of all domains ordered by their shallowest tile, so that the bins a
  // routing launch needs form one contiguous range (device indices keep
  // their per-domain order; slot tables move with their bin)
  const int64_t nb = (int64_t)out.bins.size();
  std::vector<int64_t> ord(nb);
  for (int64_t b = 0; b < nb; ++b) ord[b] = b;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int64_t u, int64_t v) { return out.bins[u].tau_lo < out.bins[v].tau_lo; });
  std::vector<BinInfo> bins(nb);
  std::vector<SlotMeta> slots(out.slots.size());
  out.bin_span = 0;
  for (int64_t b = 0; b < nb; ++b) {
    bins[b] = out.bins[ord[b]];
    std::copy(out.slots.begin() + ord[b] * kTileSlots, out.slots.begin() + (ord[b] + 1) * kTileSlots,
              slots.begin() + b * kTileSlots);
    out.bin_span = std::max(out.bin_span, bins[b].tau_hi - bins[b].tau_lo);
  }
  out.bins.swap(bins);
  out.slots.swap(slots);
  return MHMHIP_OK;
}

// Bins [lo, hi) that can hold a tile of depth tau in [tau_a, tau_b]: bins are
// ordered by their shallowest tile depth and span at most bin_span depths.
void tile_bin_range(const LayoutOut &lay, int32_t tau_a, int32_t tau_b, int64_t &lo, int64_t &hi) {
  auto first_ge = [&](int32_t tau) {
    return (int64_t)(std::lower_bound(lay.bins.begin(), lay.bins.end(), tau,
                                      [](const BinInfo &b, int32_t v) { return b.tau_lo < v; }) -
                     lay.bins.begin());
  };
  lo = first_ge(tau_a - lay.bin_span);
  hi = first_ge(tau_b + 1);
}

}  // namespace mhmhip

extern "C" int mhmhip_plan_network(int64_t n_nodes, int32_t n_outlets, const int32_t *netPerm,
                                   const int32_t *fromN, const int32_t *toN, mhmhip_plan *plan,
                                   int64_t *pos, int32_t *slots, int64_t slots_cap) {
  if (n_nodes < 1 || n_outlets < 1 || n_outlets > n_nodes || !plan ||
      (n_nodes > n_outlets && (!netPerm || !fromN || !toN)))
    return MHMHIP_EINVAL;
  const int64_t start[2] = {0, n_nodes};
  mhmhip::LayoutIn in{n_nodes, 1, start, &n_outlets, netPerm, fromN, toN};
  mhmhip::LayoutOut lay;
  std::string err;
  const int rc = mhmhip::build_layout(in, lay, err);
  if (rc != MHMHIP_OK) return rc;
  plan->dmax = lay.dmax;
  plan->tau_max = lay.tau_max;
  plan->n_tiles = lay.n_tiles;
  plan->n_bins = (int64_t)lay.bins.size();
  plan->n_bnd = lay.n_bnd;
  plan->n_slots = (int64_t)lay.slots.size();
  if (pos)
    for (int64_t i = 0; i < n_nodes; ++i) pos[i] = lay.pos[i];
  if (slots) {
    if (slots_cap < plan->n_slots) return MHMHIP_EINVAL;
    for (size_t k = 0; k < lay.slots.size(); ++k) {
      slots[4 * k + 0] = lay.slots[k].pos;
      slots[4 * k + 1] = lay.slots[k].bnd;
      slots[4 * k + 2] = lay.slots[k].a;
      slots[4 * k + 3] = lay.slots[k].b;
    }
  }
  return MHMHIP_OK;
}
