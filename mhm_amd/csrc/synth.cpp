// synth.cpp -- host-only tools for synthetic test domains.
//
// syn_priority_flood: D8 drainage of a DEM towards one outlet by priority
// flood (cells are claimed in order of their filled elevation, each draining
// to the neighbour that claimed it), so every cell reaches the outlet and
// the flow directions are valid ESRI D8 codes as read by
// src/mRM/mo_mrm_net_startup.f90:1768-1797 (moveDownOneCell).
// Grids are column-major (row index fastest), as the reference's L0 arrays.
#include <cmath>
#include <cstdint>
#include <queue>
#include <vector>

namespace {
struct Item {
  double z;
  int64_t order;
  int64_t idx;
  bool operator>(const Item &o) const { return z > o.z || (z == o.z && order > o.order); }
};
// neighbour offsets (drow, dcol) and the D8 code of the move from the
// neighbour back to the current cell (E=1, SE=2, S=4, SW=8, W=16, NW=32, N=64, NE=128)
const int kDr[8] = {0, 1, 1, 1, 0, -1, -1, -1};
const int kDc[8] = {1, 1, 0, -1, -1, -1, 0, 1};
int code_of(int dr, int dc) {
  if (dr == 0 && dc == 1) return 1;
  if (dr == 1 && dc == 1) return 2;
  if (dr == 1 && dc == 0) return 4;
  if (dr == 1 && dc == -1) return 8;
  if (dr == 0 && dc == -1) return 16;
  if (dr == -1 && dc == -1) return 32;
  if (dr == -1 && dc == 0) return 64;
  return 128;
}
}  // namespace

extern "C" int syn_priority_flood(int64_t nrows, int64_t ncols, const double *dem, int64_t outlet,
                                  double eps, double *filled, int64_t *down, int32_t *fdir) {
  const int64_t n = nrows * ncols;
  if (nrows < 1 || ncols < 1 || outlet < 0 || outlet >= n) return -1;
  std::vector<char> seen(n, 0);
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  int64_t order = 0;
  for (int64_t i = 0; i < n; ++i) {
    down[i] = -1;
    fdir[i] = 0;
  }
  filled[outlet] = dem[outlet];
  seen[outlet] = 1;
  pq.push({filled[outlet], order++, outlet});
  while (!pq.empty()) {
    const Item it = pq.top();
    pq.pop();
    const int64_t r = it.idx % nrows, c = it.idx / nrows;
    for (int k = 0; k < 8; ++k) {
      const int64_t rr = r + kDr[k], cc = c + kDc[k];
      if (rr < 0 || rr >= nrows || cc < 0 || cc >= ncols) continue;
      const int64_t j = rr + cc * nrows;
      if (seen[j]) continue;
      seen[j] = 1;
      filled[j] = std::max(dem[j], filled[it.idx] + eps);
      down[j] = it.idx;
      fdir[j] = code_of((int)(r - rr), (int)(c - cc));
      pq.push({filled[j], order++, j});
    }
  }
  return 0;
}

// Depth of every cell in the drainage tree (number of links to the outlet).
extern "C" int syn_tree_depth(int64_t n, const int64_t *down, int32_t *depth) {
  std::vector<int64_t> stack;
  for (int64_t i = 0; i < n; ++i) depth[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    int64_t j = i;
    stack.clear();
    while (j >= 0 && depth[j] < 0) {
      stack.push_back(j);
      j = down[j];
      if ((int64_t)stack.size() > n) return -1;
    }
    int32_t d = (j >= 0) ? depth[j] : -1;
    for (auto it = stack.rbegin(); it != stack.rend(); ++it) depth[*it] = ++d;
  }
  return 0;
}
