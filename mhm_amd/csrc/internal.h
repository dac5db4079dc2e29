// internal.h -- device data layout and kernel argument blocks shared by the
// kernels (kernels.hip) and the host side of the C-ABI (api.hip).
//
// Layout in HBM (one context):
//   cells are stored in routing order (breadth-first from each outlet,
//   domain after domain), so a cell and its L11 node share one index p.
//   A parameter set s occupies the slice [s*sstride, s*sstride + n_cells)
//   of every row; sstride = n_cells rounded up to 64 doubles.
//   P[row][ld]   effective parameters          (ld = n_sets * sstride)
//   S[row][ld]   states
//   X[row][ld]   fluxes of the last step
//   F[row][column][sstride]  forcing, shared by all parameter sets: the
//                variables the PET case reads, one row each (frow); day d
//                (iMeteoTS) in column (d - 1) mod n_forcing_days
//   ring[R + kTileHeight][ld]  qOUT per step (L1 total runoff in m3/s),
//                consumed by the routing: step t of node p in row
//                (t mod R) + lag[p] (ring_row), lag 0 for the diagonal kernels
#pragma once

#include <stdint.h>

#include "../../include/mhmhip.h"

namespace mhmhip {

constexpr int kMaxH = 6;

// parameter rows (device).  Divisions and scalings the reference performs
// every step on step-invariant operands are done once on the host
// (bit-identical): DMAXS = degDayMax*c2TSTu, DNPS = degDayNoPre*c2TSTu,
// K*S = c2TSTu/k*, and the latitude terms of extraterr_rad_approx.
enum PRow {
  PR_FS = 0, PR_SEALTHR, PR_JARVIS, PR_K0S, PR_K1S, PR_K2S, PR_KPS, PR_ALPHA, PR_KARST,
  PR_UNSATTHR, PR_FASP, PR_HARSAM, PR_TANLAT, PR_SINLAT, PR_COSLAT, PR_DMAXS, PR_DNPS,
  PR_INCR, PR_TTH, PR_NSCALAR
};
__host__ __device__ constexpr int pr_sat(int nh) { return PR_NSCALAR; }
__host__ __device__ constexpr int pr_froots(int nh) { return PR_NSCALAR + nh; }
__host__ __device__ constexpr int pr_fc(int nh) { return PR_NSCALAR + 2 * nh; }
__host__ __device__ constexpr int pr_wp(int nh) { return PR_NSCALAR + 3 * nh; }
__host__ __device__ constexpr int pr_exp(int nh) { return PR_NSCALAR + 4 * nh; }
__host__ __device__ constexpr int pr_maxinter(int nh) { return PR_NSCALAR + 5 * nh; }
__host__ __device__ constexpr int pr_petlai(int nh) { return PR_NSCALAR + 5 * nh + 12; }
__host__ __device__ constexpr int pr_prietay(int nh) { return PR_NSCALAR + 5 * nh + 24; }
__host__ __device__ constexpr int pr_aero(int nh) { return PR_NSCALAR + 5 * nh + 36; }
__host__ __device__ constexpr int pr_surf(int nh) { return PR_NSCALAR + 5 * nh + 48; }
__host__ __device__ constexpr int pr_count(int nh) { return PR_NSCALAR + 5 * nh + 60; }

enum SRow { SR_INTER = 0, SR_SNOW, SR_SEAL, SR_UNSAT, SR_SAT, SR_NSCALAR };
__host__ __device__ constexpr int sr_count(int nh) { return SR_NSCALAR + nh; }

enum XRow {
  XR_PET = 0, XR_TEMP, XR_PREC, XR_THROUGHFALL, XR_AETCAN, XR_DEGDAY, XR_MELT, XR_PEFF,
  XR_RAIN, XR_SNOW, XR_RUNOFFSEAL, XR_AETSEAL, XR_FAST, XR_SLOW, XR_PERC, XR_BASEFLOW,
  XR_TOTAL, XR_NSCALAR
};
__host__ __device__ constexpr int xr_infil(int nh) { return XR_NSCALAR; }
__host__ __device__ constexpr int xr_aet(int nh) { return XR_NSCALAR + nh; }
__host__ __device__ constexpr int xr_count(int nh) { return XR_NSCALAR + 2 * nh; }

// Per-step scalars on the device (host-evaluated from the step clock).
struct StepDev {
  double dr, tand, sind, cosd;  // extraterr_rad_approx day-of-year terms
  // monthly namelist factors of this step, day or night already selected
  // (mo_meteo_handler.f90:1111-1117, 1192-1198, 1273-1279; mo_mhm.f90:473)
  double f_pet, f_prec, f_temp, evap;
  double revap;                 // RN(1 / evap)
  int32_t day;                  // 0-based forcing column
  int32_t month0;               // 0..11
  int32_t isday;                // 6 < hour <= 18
  int32_t init_sm;              // tt == 1 .and. .not. read_states
  int32_t doy;
  int32_t meteo_same;           // same forcing column, day and factors as the step before
  int32_t wrow;                 // month0 * 24 + hour: the row of the meteo weights
};

struct MonthTables {
  double evap[12];
  double fday_pet[12], fnight_pet[12];
  double fday_prec[12], fnight_prec[12];
  double fday_temp[12], fnight_temp[12];
};

// Output aggregation of L1 states and fluxes (mHM_updateDataset,
// src/mHM/mo_write_fluxes_states.f90:283-438): row r of the sums holds output
// kind[r] (outputFlxState index 1..21) of horizon hor[r]; every step adds
// each row's value to its sum (updateVariable, mo_nc_output.f90:140-148).
constexpr int kOutRowsMax = 64;
struct OutAcc {
  double *sum;               // [n][ld], null: off
  int32_t n;
  int8_t kind[kOutRowsMax];
  int8_t hor[kOutRowsMax];
};

struct CellArgs {
  const double *P;
  double *S;
  double *X;
  const double *F;
  double *ring;            // nullptr: no routing input wanted
  const double *area;      // [n_cells] L1 cell area [km2] (routing input)
  const uint8_t *lag;      // [n_cells] ring lag of the node (tiled routing), null: 0
  double tst;              // HourSecs * timestep [s]
  double rtst;             // RN(1 / tst)
  const StepDev *steps;    // device step table (fused kernel)
  int64_t n_cells;
  int64_t sstride;
  int64_t ld;
  int64_t forcing_days;
  int64_t ring_len;
  int32_t pet_case;
  int32_t jarvis;
  int32_t hourly;
  int32_t ring_rep;        // routing sub-steps per model step: qOUT goes to that
                           // many consecutive ring rows (1: one row per step)
  int32_t raw_total;       // 1: the ring gets L1_total_runoff per cell (routing on a
                           // coarser grid aggregates it, runoff_agg_kernel), 0: qOUT
  double nT;               // steps per day as real (nTstepDay_dp)
  double rnT;              // RN(1 / nT)
  OutAcc out;              // output sums (out.sum null: none)
  // meteo weights [3][12 * 24][sstride] (pre, temp, pet), null: the day/night
  // factors disaggregate
  const double *W;
};

// The row of forcing variable v in F for a PET case (-1: not read by that
// case, mo_meteo_handler.f90:1058-1097): pre and tavg, then the case's own
// inputs in variable order.
__host__ __device__ constexpr int forcing_row(int v, int pet_case) {
  return v == MHMHIP_F_PRE ? 0
         : v == MHMHIP_F_TAVG ? 1
         : pet_case <= 0 ? (v == MHMHIP_F_PET ? 2 : -1)
         : pet_case == 1 ? (v == MHMHIP_F_TMIN ? 2 : v == MHMHIP_F_TMAX ? 3 : -1)
         : v == MHMHIP_F_NETRAD ? 2
         : pet_case == 3 ? (v == MHMHIP_F_ABSVAPPRESS ? 3 : v == MHMHIP_F_WINDSPEED ? 4 : -1)
                         : -1;
}
__host__ __device__ constexpr int forcing_rows(int pet_case) {
  return pet_case == 1 ? 4 : pet_case == 3 ? 5 : 3;
}

// Streamed forcing: the forcing row (in F) of the k-th variable of a staged
// chunk, k < n.
struct FRowMap {
  int32_t row[MHMHIP_F_COUNT];
  int32_t n;
};

// Routing node metadata, packed in 8 bytes per node.
//   child: offset of the node's upstream nodes in the child list (device
//          indices, in the reference's routing rank)
//   packed: depth (bits 0..19) | nchild (20..23) | flags (24..28)
enum NodeFlag {
  NF_LINK = 1 << 24,    // node has an outgoing link (computes qTR)
  NF_QOUT = 1 << 25,    // node adds its own qOUT to qTIN
  NF_CUT = 1 << 26,     // downstream node is a non-headwater inflow gauge
  NF_GAUGE = 1 << 27,   // node carries one or more gauges
  NF_INFLOW = 1 << 28,  // node is an inflow gauge node
};
struct NodeMeta {
  int32_t child;
  int32_t packed;
};
__host__ __device__ inline int nm_depth(int32_t y) { return y & 0xFFFFF; }
__host__ __device__ inline int nm_nchild(int32_t y) { return (y >> 20) & 0xF; }
constexpr int32_t NF_MASK = NF_LINK | NF_QOUT | NF_CUT | NF_GAUGE | NF_INFLOW;

// One L1 cell of the meteo-grid mapping (meteo_l2_kernel): its L2 block
// (0-based rows r0..r1, columns c0..c1; empty when r0 > r1) and the number n
// of valid L2 cells in it when aggregating, else the source L2 cell (src, -1:
// nodata); dst: the cell's index in a forcing row.
struct MeteoL2Cell {
  int32_t r0, r1, c0, c1;
  int32_t n, src;
  int64_t dst;
};

// Tiled routing.  Each domain's tree is cut into connected tiles of at most
// kTileSlots slots and kTileHeight levels; a tile holds its nodes plus one
// proxy slot per upstream node that roots another tile.  Tiles are packed
// into bins of kTileSlots slots, one workgroup per (bin, parameter set).
#ifndef MHMHIP_TILE_SLOTS
#define MHMHIP_TILE_SLOTS 256
#endif
#ifndef MHMHIP_TILE_HEIGHT
#define MHMHIP_TILE_HEIGHT 16
#endif
#ifndef MHMHIP_TILE_WINDOW
#define MHMHIP_TILE_WINDOW 64
#endif
constexpr int kTileSlots = MHMHIP_TILE_SLOTS;
constexpr int kTileHeight = MHMHIP_TILE_HEIGHT;  // default levels per tile
constexpr int kTileHeightMax = 64;                // mhmhip_set_tile_height's limit
constexpr int kTileWindowMax = MHMHIP_TILE_WINDOW;  // steps per window, at most
static_assert(kTileHeight >= 1 && kTileHeight <= kTileHeightMax && kTileHeightMax <= 256,
              "slot lags are stored in 8 bits");

// Ring row of step t (relative to the call) of a node with lag g, in a ring
// of ring_len logical rows (plus kTileHeight rows past them).  The tiled
// routing's slot with lag g handles window step j on diagonal j + g, so
// writing step t of the node g rows late puts everything a bin reads on one
// diagonal into one row.  The diagonal kernels use g = 0.
__host__ __device__ inline int64_t ring_row(int64_t t, int32_t g, int64_t ring_len) {
  return t % ring_len + g;
}

// One slot of a bin (16 bytes).
//   pos: device node index of a real slot, -1 for proxies and padding
//   bnd: boundary row read by a proxy / written by a tile root, else -1
//   a:   lag (bits 0..11) | nchild (12..15) | SF_* | NF_* flags of the node
//   b:   bin-local slot of the first upstream slot (bits 0..15) | tile depth (16..31)
// lag = tile height - 1 - local depth: the slot handles window step j on the
// bin's internal diagonal j + lag.
enum SlotFlag {
  SF_REAL = 1 << 16,   // a node of this tile
  SF_PROXY = 1 << 17,  // stands for the root of an upstream tile
  SF_TROOT = 1 << 18,  // roots a tile that drains into another tile
};
struct SlotMeta {
  int32_t pos;
  int32_t bnd;
  int32_t a;
  int32_t b;
};
struct BinInfo {
  int32_t height;    // tallest tile in the bin
  int32_t tau_lo;    // smallest / largest tile depth in the bin
  int32_t tau_hi;
  int32_t out_rows;  // tile roots + gauge nodes of the bin (LDS output rows)
};
// LDS a tiled launch may use to stage its tile-root series and gauge sums;
// past it the kernel stores them directly
constexpr size_t kTileStageLds = 64 * 1024;

struct RouteArgs {
  const NodeMeta *meta;    // [n_nodes]
  const int32_t *child_list;  // upstream device indices, NodeMeta.child offsets
  const double *C1;        // [ld]
  const double *C2;        // [ld]
  double *qTIN;            // [ld]   qTIN(:,IT1)
  double *pub;             // [2][ld] qTR by step parity
  double *qOUT;            // [ld]   last qOUT (reference L11_qOUT), may be null
  double *qacc;            // [ld]   Qmod summed over the routed steps, null: off
  const double *ring;      // [ring_len + kTileHeight][ld] (ring_row)
  double *gauge_out;       // [nsteps][n_sets][n_gauges] relative to t_base
  const int32_t *gauge_node;  // [n_gauges] device node index
  const int32_t *inflow_node; // [n_inflow] device node index
  const int32_t *inflow_hw;   // [n_inflow]
  const double *inflow_q;     // [nsteps][n_inflow] relative to t_base
  int64_t n_nodes;
  int64_t sstride;
  int64_t ring_len;
  int64_t t_lo, t_hi;      // active global steps [t_lo, t_hi)
  int64_t t_base;          // global step of gauge_out/inflow_q row 0
  int64_t acc_from;        // qacc adds the steps t >= acc_from only
  int32_t n_sets;
  int32_t n_gauges;
  int32_t n_inflow;
  int32_t dmax;            // maximum depth over the nodes this launch covers
  double tst;              // HourSecs * timestep [s]
  int32_t rep;             // routing sub-steps per model step (inflow_q rows are
                           // per model step: row (t - t_base) / rep)
  int32_t pad;
};

// MPR-lite on the device: the synthetic domains' stand-in regionalisation
// (one set's parameter rows from its gamma vector and the cell inputs).
constexpr int kMprNGamma = MHMHIP_MPRLITE_NGAMMA;
struct MprLiteArgs {
  double *P;               // parameter rows of set 0 (rows ld apart)
  const double *in;        // input rows (sstride apart): forest, impervious,
                           // pervious fraction, texture, aspect [deg], and the
                           // tan, sin, cos of the latitude
  int64_t ld, sstride, n_cells;
  int32_t nh, pet_case;
  double c2;               // c2TSTu (timestep / 24)
  double hd[MHMHIP_MAX_HORIZONS];  // horizon lower boundaries [mm]
  double lai[36];          // monthly LAI of forest, impervious, pervious
  double gamma[kMprNGamma];
};

// Extra arguments of the tiled routing kernel.
struct TileArgs {
  const SlotMeta *slot;    // [n_bins * kTileSlots]
  const BinInfo *bin;      // [n_bins]
  const int32_t *node_gauge;  // [n_nodes] first gauge at the node (device index), -1: none
  const int32_t *slot_out;    // [n_bins * kTileSlots] output rows: tile root (bits 0..15),
                              // gauge (16..31), 0xFFFF: none
  double *bnd;             // [2][n_sets][n_bnd][bstride] tile-root outflow by window parity
  int64_t n_bnd;
  int32_t tau_max;         // deepest tile
  int32_t B;               // steps per window
  int32_t nwin;            // windows in [t_lo, t_hi)
  int32_t bstride;         // boundary series length (kTileWindowMax)
  int32_t out_rows;        // largest BinInfo.out_rows
  int32_t pad;
};

}  // namespace mhmhip
