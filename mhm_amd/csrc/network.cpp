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

// Breadth-first layout of every domain's forest: roots are the domain's sink
// nodes in node order, the upstream nodes of a node are appended in the
// reference's routing rank, so they are contiguous and their qTR sum runs in
// the order of L11_routing's accumulation.
int build_layout(const LayoutIn &in, LayoutOut &out, std::string &err) {
  const int64_t nN = in.n_nodes;
  out.pos.assign(nN, -1);
  out.node_at.assign(nN, -1);
  out.child.assign(nN, 0);
  out.packed.assign(nN, 0);
  out.link_of_node.assign(nN, -1);
  out.dom_dmax.assign(in.n_domains, 0);
  out.dmax = 0;
  int64_t link_off = 0;
  int64_t next_pos = 0;
  for (int d = 0; d < in.n_domains; ++d) {
    const int64_t n0 = in.dom_node_start[d], n1 = in.dom_node_start[d + 1], nd = n1 - n0;
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
    // (mo_mrm_routing.f90:466-467); every link's source node does too (445)
    const int64_t last_t = nL > 0 ? (int64_t)to[perm[nL - 1] - 1] - 1 : -1;
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
    const int64_t base = next_pos;
    for (int64_t h = 0; h < nd; ++h) out.pos[n0 + q[h]] = base + h;
    for (int64_t h = 0; h < nd; ++h) {
      const int64_t n = q[h];
      const int64_t p = base + h;
      out.node_at[p] = n0 + n;
      int32_t flags = 0;
      if (out_link[n] >= 0) {
        flags |= NF_LINK | NF_QOUT;
        out.link_of_node[p] = link_off + out_link[n];
      }
      if (n == last_t) flags |= NF_QOUT;
      const int32_t nc = (int32_t)up[n].size();
      out.child[p] = nc ? (int32_t)out.pos[n0 + fr[up[n][0]] - 1] : 0;
      if (depth[n] >= (1 << 20)) {
        err = "bind_network: network deeper than 2^20 levels";
        return MHMHIP_EINVAL;
      }
      out.packed[p] = depth[n] | (nc << 20) | flags;
      dm = std::max(dm, depth[n]);
    }
    out.dom_dmax[d] = dm;
    out.dmax = std::max(out.dmax, dm);
    next_pos += nd;
    link_off += nL;
  }
  out.n_links_total = link_off;
  return MHMHIP_OK;
}

}  // namespace mhmhip
