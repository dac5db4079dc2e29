// network.h -- tiled device layout of the packed L11 networks.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "internal.h"

namespace mhmhip {

struct LayoutIn {
  int64_t n_nodes;
  int32_t n_domains;
  const int64_t *dom_node_start;  // n_domains + 1
  const int32_t *n_outlets;       // per domain
  const int32_t *netPerm;         // packed per-domain links, 1-based local
  const int32_t *fromN;
  const int32_t *toN;
  int32_t tile_height = kTileHeight;  // levels per tile, 2..kTileHeightMax
};

struct LayoutOut {
  std::vector<int64_t> pos;           // reference node (global 0-based) -> device index
  std::vector<int64_t> node_at;       // device index -> reference node
  std::vector<int32_t> child;         // NodeMeta.child (offset into child_list)
  std::vector<int32_t> child_list;    // upstream device indices in routing rank
  std::vector<int32_t> packed;        // NodeMeta.packed
  std::vector<int64_t> link_of_node;  // device index -> packed link index (or -1)
  std::vector<int32_t> dom_dmax;
  int32_t dmax;
  int64_t n_links_total;
  // tiles (kTileSlots slots per bin)
  std::vector<SlotMeta> slots;        // [n_bins * kTileSlots]; a without NF_* flags
  std::vector<BinInfo> bins;
  int32_t tau_max;
  int32_t bin_span;                   // largest tau_hi - tau_lo of a bin
  int64_t n_bnd;
  int64_t n_tiles;
};

// Validates the packed networks and builds the device layout: the tile
// partition of every domain's trees, tile-major device indices (domains stay
// contiguous), node metadata and the child list in routing rank.
int build_layout(const LayoutIn &in, LayoutOut &out, std::string &err);

// Bins [lo, hi) that can hold a tile of depth tau in [tau_a, tau_b].
void tile_bin_range(const LayoutOut &lay, int32_t tau_a, int32_t tau_b, int64_t &lo, int64_t &hi);

}  // namespace mhmhip
