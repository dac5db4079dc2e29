// api.hip -- host side of the mhmhip C-ABI: device memory, data layout,
// argument checking and the time loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhmhip.h"
#include "internal.h"
#include "launch.h"
#include "network.h"

using namespace mhmhip;

namespace {

// FORCES mo_constants values used by the host-evaluated terms.
const double kTwoPiD = 6.283185307179586476925286766559005768394;
const double kDeg2Rad = 3.141592653589793238462643383279502884197 / 180.0;
const double kYearDays = 365.0;
const double kHourSecs = 3600.0;

thread_local std::string g_create_error;

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

// Device scratch of mhmhip_run_noalloc: the step table, gauge series and
// runoff ring of one call, in a file-scope array that every GPU the library
// runs on holds (its own copy) until the process ends.  Calls on one GPU
// share it: two such calls that overlap on one GPU race.  In host code the
// name stands for a host-side shadow; the device address of each GPU's copy
// comes from hipGetSymbolAddress with that GPU current (g_scratch).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of this GPU's scratch, resolved once per device with
// the device current (the caller has made it current).
hipError_t scratch_resolve(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch[dev] = static_cast<unsigned char *>(p);
  return hipSuccess;
}
}  // namespace

// Workgroup size of the one-step cell kernel (mhmhip_mhm_step, the plain
// time loop): two waves per group schedule the HBM-bound step better than four.
constexpr int kStepBlock = 128;

struct mhmhip_ctx {
  mhmhip_config cfg{};
  int nh = 0;
  int n_sets = 1;
  int n_domains = 1;
  int64_t n_cells = 0, sstride = 0, ld = 0;
  double c2TSTu = 1.0, nT = 24.0;
  std::vector<int64_t> dom_cell_start;
  MonthTables mt{};
  hipStream_t stream = nullptr;
  hipStream_t rstream = nullptr;  // routing stream of a captured run (default priority)
  std::string err;

  double *P = nullptr, *S = nullptr, *X = nullptr, *F = nullptr;
  int64_t n_prows = 0, n_srows = 0, n_xrows = 0;
  // forcing rows: frow[v] is variable v's row of F (-1: the PET case does
  // not read it), n_frows rows in all
  int32_t frow[MHMHIP_F_COUNT] = {};
  int32_t n_frows = 0;
  double *W = nullptr;  // meteo weights [3][12 * 24][sstride], null: off
  // MPR-lite inputs (mhmhip_mprlite_inputs): host copies in reference cell
  // order, and the kernel's input rows laid out as the cells are (rebuilt
  // from the host copies after a network binding moves the cells)
  std::vector<double> mpr_cells;  // [6][n_cells]
  double mpr_hd[MHMHIP_MAX_HORIZONS] = {};
  double mpr_lai[36] = {};
  double *d_mpr = nullptr;        // [8][sstride]
  bool mpr_stale = true;
  // whether P, S, X and F hold anything written since creation (a network
  // bound before any data re-lays nothing)
  bool p_dirty = false, s_dirty = false, x_dirty = false, f_dirty = false;
  std::vector<int64_t> pos;  // reference cell -> device index (identity until a network is bound)
  // L11 routing nodes: their count, device index of every reference node and
  // per-domain ranges (the cells' until a network is bound).  On a routing
  // grid coarser than L1 (coarse) the cells are grouped by node on the device
  // and their runoff is aggregated per node (d_agg_start: the device cell
  // range of each device node; d_cring: L1_total_runoff of a launch's steps).
  int64_t n_nodes = 0;
  std::vector<int64_t> npos;
  std::vector<int64_t> dom_node_start;
  bool coarse = false;
  int32_t tile_h = kTileHeight;  // levels per routing tile (mhmhip_set_tile_height)
  int32_t *d_agg_start = nullptr;
  double *d_cring = nullptr;
  int64_t cring_rows = 0;
  // Nodes of single-node domains: the reference skips L11_routing there
  // (mo_mrm_routing.f90:285-289) and leaves their qTIN as set (0 at
  // init); the device's qTIN holds Qmod for them, so their qTIN per set is
  // kept here.
  std::vector<int64_t> lone_nodes;
  std::vector<double> lone_qtin;
  // adaptive routing (processCase(8) = 2): the routing step [s] of each
  // (set, domain), 0 for processCase 1; the Qmod sums of the last
  // mrm_step's sub-steps and their count (> 1 when they apply)
  std::vector<double> ts_rout;
  // processCase(8) = 1 coefficients per (set, domain): the routing1
  // parameters last given to set_routing_params (reg_rout recomputes C1/C2
  // from them at every routed step with read_states = 0,
  // mo_mrm_routing.f90:211-218) and whether C1/C2 were last loaded from a
  // restart (mo_mrm_restart.f90:548-555), which only read_states = 1 keeps
  // routing steps longer than the model step: RunToRout [mm] per (set, cell)
  // on the device, InflowDischarge summed on the host, the inflow row of
  // the next mrm_step kept on the host, and the averaged inflow of a
  // routing step on the device
  double *d_runacc = nullptr;
  std::vector<double> inflow_acc, h_inflow_row;
  double *d_inflow_avg = nullptr;
  std::vector<double> gam5;        // [set][domain][5]
  std::vector<char> gam_set, coef_restart;
  double *d_qsub = nullptr;
  int32_t qsub_loops = 0;
  // pinned staging rows (kStageSlots x sstride doubles): an upload fills the
  // next slot once that slot's previous copy has left it (stage_ev), so
  // several row uploads are in flight at once
  static constexpr int kStageSlots = 4;
  double *stage = nullptr;
  hipEvent_t stage_ev[kStageSlots] = {};
  int stage_next = 0;

  // Forcing columns.  Day d (1-based, iMeteoTS) lives in column
  // (d - 1) mod n_forcing_days; col_day[c] is the day column c holds (c + 1
  // until a streamed chunk replaces it) and col_seq[c] the streamed chunk
  // that wrote it (-1: written synchronously).  Streamed chunks go up on the
  // copy stream (cstream) in the order they are given; chunk_ev[k] marks the
  // end of chunk k's upload, and a launch reading a column first waits for
  // the chunk that wrote it (waited_seq: the newest chunk the compute stream
  // has waited for).  A chunk overwrites its columns only after the compute
  // work queued before it (which may still read them) is done.
  std::vector<int32_t> col_day;
  std::vector<int64_t> col_seq;
  std::vector<hipEvent_t> chunk_ev;
  int64_t waited_seq = -1;
  hipStream_t cstream = nullptr;
  // double-buffered staging of streamed chunks: pinned host copies of the
  // caller's rows and their device copies before the gather into routing
  // order; fstage_ev[i] marks the end of the upload out of slot i
  double *fstage_h[2] = {nullptr, nullptr}, *fstage_d[2] = {nullptr, nullptr};
  int64_t fstage_cap = 0;  // doubles per slot
  hipEvent_t fstage_ev[2] = {nullptr, nullptr};
  int fstage_next = 0;
  int32_t *d_pos = nullptr;       // device index of every reference cell (gather)
  bool d_pos_stale = true;

  // routing
  bool net_bound = false;
  LayoutOut lay;
  NodeMeta *d_meta = nullptr;
  double *d_area = nullptr, *d_C1 = nullptr, *d_C2 = nullptr, *d_qTIN = nullptr,
         *d_pub = nullptr, *d_qOUT = nullptr;
  int64_t *d_dom_start = nullptr;
  int32_t *d_dom_dmax = nullptr;
  int32_t *d_child_list = nullptr;
  int64_t n_child_dev = 0;
  // tiled routing
  SlotMeta *d_slots = nullptr;
  BinInfo *d_bins = nullptr;
  double *d_bnd = nullptr;
  int64_t n_slots_dev = 0, n_bins_dev = 0, n_bnd_dev = 0;
  std::vector<double> h_length, h_slope, h_fFPimp, h_area;
  std::vector<int64_t> dom_link_start;
  std::vector<int32_t> h_nout;
  int32_t n_gauges = 0, n_inflow = 0;
  int32_t *d_gauge_node = nullptr, *d_inflow_node = nullptr, *d_inflow_hw = nullptr;
  int32_t *d_node_gauge = nullptr;  // [n_cells] first gauge at the node, -1: none
  uint8_t *d_lag = nullptr;         // [n_cells] ring lag of each node (its tile slot's lag)
  int32_t *d_slot_out = nullptr;    // [n_slots] LDS output rows of the tiled kernel
  // L1 output sums (mhmhip_output_accumulate): rows in the reference's
  // output order, (kind, horizon) each
  double *d_out = nullptr;
  int64_t out_rows_cap = 0;
  OutAcc out_cfg{};
  bool out_on = false;
  int64_t out_count = 0;
  double *d_qacc = nullptr;         // [ld] Qmod summed over routed steps (output aggregation)
  bool qacc_on = false;
  int64_t qacc_count = 0;
  int32_t tile_out_rows = 0;        // most output rows of a bin
  std::vector<std::pair<int32_t, int32_t>> gauge_dup;  // (gauge, earlier gauge at its node)
  std::vector<int32_t> h_inflow_node_dev, h_inflow_hw;
  double *d_inflow_q = nullptr;
  int64_t inflow_q_rows = 0;
  double *d_ring = nullptr;
  int64_t ring_len = 0;
  double *d_gauge_buf = nullptr;
  int64_t gauge_buf_rows = 0;
  StepDev *d_steps = nullptr;
  int64_t steps_cap = 0;

  int64_t t_cell = 0, t_route = 0;  // global step counters
  int64_t device_bytes = 0;
  // Kernels index steps relative to the start of each call; the parity slot
  // of pub holding qTR(:,IT1) is normalised to 1 before a call (step -1).
  int qtr_slot = 1;
  // positions [lev_start[k], lev_start[k+1]) hold the nodes at depth k when
  // the layout keeps depths contiguous (lets a diagonal launch only its
  // active band); empty for the tile-major layout
  std::vector<int64_t> lev_start;

  // time loop captured as a hipGraph, replayed while its key matches
  struct GraphKey {
    int32_t nsteps = -1, mode = -1, S = -1, block = -1, wg = -1, prof = -1, B = -1, out_n = -1,
            rep = -1;
    const void *ring = nullptr, *gauge = nullptr, *steps = nullptr, *inflow = nullptr,
               *qacc = nullptr, *out = nullptr, *cring = nullptr;
    int64_t ring_len = 0;
    bool operator==(const GraphKey &o) const {
      return nsteps == o.nsteps && mode == o.mode && S == o.S && block == o.block && wg == o.wg &&
             prof == o.prof && B == o.B && ring == o.ring && gauge == o.gauge && steps == o.steps &&
             inflow == o.inflow && qacc == o.qacc && ring_len == o.ring_len && out == o.out &&
             out_n == o.out_n && rep == o.rep && cring == o.cring;
    }
  };
  GraphKey gkey;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // one routed step of mhmhip_mrm_step (the coarser grid's aggregation, the
  // inflow patch and the 1 + tile-depth tile launches) captured once and
  // replayed every step; rebuilt when a buffer it reads or writes changes
  struct StepKey {
    const void *ring = nullptr, *gauge = nullptr, *inflow = nullptr, *qacc = nullptr;
    bool operator==(const StepKey &o) const {
      return ring == o.ring && gauge == o.gauge && inflow == o.inflow && qacc == o.qacc;
    }
  };
  StepKey skey;
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;

  // kernel timing (mhmhip_set_profiling)
  struct Pending {
    int kind;
    hipEvent_t a, b;
    double units;
  };
  // With profiling on, a captured graph holds event-record nodes around the
  // timed launches; these events belong to the graph and are read after
  // every replay.
  bool capturing = false;
  std::vector<Pending> graph_events;
  std::vector<hipEvent_t> graph_sync;  // cross-stream edges of the captured graph
  int64_t graph_all[MHMHIP_K_COUNT] = {};
  bool prof = false;
  int prof_stride = 1;         // time every prof_stride-th launch of each kernel
  int64_t prof_seq[MHMHIP_K_COUNT] = {};  // launches seen per kernel while profiling
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[MHMHIP_K_COUNT] = {};
  int64_t prof_n[MHMHIP_K_COUNT] = {};      // timed launches
  int64_t prof_all[MHMHIP_K_COUNT] = {};    // all launches
  double prof_units[MHMHIP_K_COUNT] = {};

  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  int hipfail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? MHMHIP_ENOMEM : MHMHIP_EHIP;
  }
  template <typename T>
  int dalloc(T **p, int64_t n, const char *what) {
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) return hipfail(e, what);
    device_bytes += sizeof(T) * n;
    e = hipMemsetAsync(*p, 0, sizeof(T) * n, stream);
    if (e != hipSuccess) return hipfail(e, what);
    return MHMHIP_OK;
  }
  template <typename T>
  void dfree(T *&p, int64_t n) {
    if (p) {
      (void)hipFree(p);
      device_bytes -= sizeof(T) * (n <= 0 ? 1 : n);
    }
    p = nullptr;
  }
};

#define CHECK_HIP(ctx, expr, what)                  \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return (ctx)->hipfail(e_, what); \
  } while (0)

#define CHECK_RC(expr)         \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != MHMHIP_OK) return rc_; \
  } while (0)

namespace {
void drop_graph(mhmhip_ctx *x);
}

// Copies are ordered on the context's (non-blocking) stream behind the
// clears and kernels queued there, then joined.
static hipError_t sync_copy(mhmhip_ctx *x, void *dst, const void *src, size_t n,
                            hipMemcpyKind k) {
  hipError_t e = hipMemcpyAsync(dst, src, n, k, x->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(x->stream);
}

// ---------------------------------------------------------------------------
extern "C" int mhmhip_create(const mhmhip_config *cfg, mhmhip_ctx **out) {
  if (!cfg || !out) {
    g_create_error = "create: null argument";
    return MHMHIP_EINVAL;
  }
  *out = nullptr;
  const mhmhip_config &c = *cfg;
  if (c.n_cells < 1 || c.n_cells > ((int64_t)1 << 31) - 64 || c.n_horizons < 1 ||
      c.n_horizons > MHMHIP_MAX_HORIZONS || c.n_sets < 1 || c.n_sets > 65535 ||
      c.timestep < 1 || c.timestep > 24 || 24 % c.timestep != 0 || c.process_soil < 1 ||
      c.process_soil > 4 || c.process_pet < -1 || c.process_pet > 3 || c.process_route < 0 ||
      c.process_route > 1 || c.n_forcing_days < 1 || c.n_domains < 1) {
    g_create_error = "create: configuration out of range";
    return MHMHIP_EINVAL;
  }
  // The kernels address a (cell, set) of a row by a 32-bit byte offset
  // ((set * sstride + cell) * 8) and step between ring rows by ld * 8 bytes:
  // every row of all sets must span less than 4 GiB.
  if ((uint64_t)round_up(c.n_cells, 64) * (uint64_t)c.n_sets * 8u > 0xFFFFFFFFull) {
    g_create_error = "create: n_sets x cells too large (a row of all sets must stay below 4 GiB: "
                     "at most 536870912 (cell, set) slots, cells rounded up to 64 per set)";
    return MHMHIP_EINVAL;
  }
  for (int m = 0; m < 12; ++m)
    if (!(c.evap_coeff[m] > 0.0) || !std::isfinite(c.evap_coeff[m])) {
      // the sealed-storage loss divides by it every step (mo_soil_moisture.f90:151)
      g_create_error = "create: evap_coeff must be positive and finite";
      return MHMHIP_EINVAL;
    }
  std::vector<int64_t> ds(c.n_domains + 1);
  if (c.domain_cell_start) {
    for (int d = 0; d <= c.n_domains; ++d) ds[d] = c.domain_cell_start[d];
    if (ds[0] != 0 || ds[c.n_domains] != c.n_cells) {
      g_create_error = "create: domain_cell_start must run from 0 to n_cells";
      return MHMHIP_EINVAL;
    }
    for (int d = 0; d < c.n_domains; ++d)
      if (ds[d + 1] <= ds[d]) {
        g_create_error = "create: empty or unordered domain";
        return MHMHIP_EINVAL;
      }
  } else {
    if (c.n_domains != 1) {
      g_create_error = "create: domain_cell_start required for several domains";
      return MHMHIP_EINVAL;
    }
    ds[0] = 0;
    ds[1] = c.n_cells;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    g_create_error = "create: no HIP device";
    return MHMHIP_EHIP;
  }
  if (c.device < 0 || c.device >= ndev) {
    g_create_error = "create: device ordinal out of range";
    return MHMHIP_EINVAL;
  }
  if (hipSetDevice(c.device) != hipSuccess) {
    g_create_error = "create: hipSetDevice failed";
    return MHMHIP_EHIP;
  }
  // set-up of mhmhip_run_noalloc: this GPU's scratch address, once per device
  if (scratch_resolve(c.device) != hipSuccess) {
    g_create_error = "create: device scratch address (hipGetSymbolAddress) failed";
    return MHMHIP_EHIP;
  }
  mhmhip_ctx *x = new mhmhip_ctx();
  x->cfg = c;
  x->cfg.domain_cell_start = nullptr;
  x->dom_cell_start = ds;
  x->nh = c.n_horizons;
  x->n_sets = c.n_sets;
  x->n_domains = c.n_domains;
  x->n_cells = c.n_cells;
  x->sstride = round_up(c.n_cells, 64);
  x->ld = x->sstride * c.n_sets;
  x->c2TSTu = (double)c.timestep / 24.0;  // mo_startup.f90:168 c2TSTu = real(timeStep,dp)/24
  x->nT = (double)(24 / c.timestep);      // mo_meteo_handler.f90:327
  for (int m = 0; m < 12; ++m) {
    x->mt.evap[m] = c.evap_coeff[m];
    x->mt.fnight_prec[m] = c.fnight_prec[m];
    x->mt.fnight_pet[m] = c.fnight_pet[m];
    x->mt.fnight_temp[m] = c.fnight_temp[m];
    // mo_meteo_handler.f90:404-406
    x->mt.fday_prec[m] = 1.0 - c.fnight_prec[m];
    x->mt.fday_pet[m] = 1.0 - c.fnight_pet[m];
    x->mt.fday_temp[m] = -1.0 * c.fnight_temp[m];
  }
  x->pos.resize(c.n_cells);
  for (int64_t i = 0; i < c.n_cells; ++i) x->pos[i] = i;
  x->n_nodes = c.n_cells;
  x->npos = x->pos;
  x->dom_node_start = x->dom_cell_start;
  x->col_day.resize(c.n_forcing_days);
  for (int32_t k = 0; k < c.n_forcing_days; ++k) x->col_day[k] = k + 1;
  x->col_seq.assign(c.n_forcing_days, -1);
  int rc = MHMHIP_OK;
  if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) {
    g_create_error = "create: stream creation failed";
    delete x;
    return MHMHIP_EHIP;
  }
  x->n_prows = pr_count(x->nh);
  // the forcing variables get_corrected_pet / get_temp / get_prec read for
  // this PET case (mo_meteo_handler.f90:1058-1097), one row of F each
  for (int v = 0; v < MHMHIP_F_COUNT; ++v) x->frow[v] = forcing_row(v, c.process_pet);
  x->n_frows = forcing_rows(c.process_pet);
  x->n_srows = sr_count(x->nh);
  x->n_xrows = xr_count(x->nh);
  if ((rc = x->dalloc(&x->P, x->n_prows * x->ld, "alloc params")) ||
      (rc = x->dalloc(&x->S, x->n_srows * x->ld, "alloc states")) ||
      (rc = x->dalloc(&x->X, x->n_xrows * x->ld, "alloc fluxes")) ||
      (rc = x->dalloc(&x->F, (int64_t)x->n_frows * c.n_forcing_days * x->sstride,
                      "alloc forcing"))) {
    g_create_error = x->err;
    mhmhip_destroy(x);
    return rc;
  }
  if (hipHostMalloc((void **)&x->stage, sizeof(double) * x->sstride * mhmhip_ctx::kStageSlots,
                    hipHostMallocDefault) != hipSuccess) {
    g_create_error = "create: pinned staging allocation failed";
    mhmhip_destroy(x);
    return MHMHIP_ENOMEM;
  }
  std::fill(x->stage, x->stage + x->sstride * mhmhip_ctx::kStageSlots, 0.0);
  for (int i = 0; i < mhmhip_ctx::kStageSlots; ++i)
    if (hipEventCreateWithFlags(&x->stage_ev[i], hipEventDisableTiming) != hipSuccess) {
      g_create_error = "create: staging event creation failed";
      mhmhip_destroy(x);
      return MHMHIP_EHIP;
    }
  if (hipStreamSynchronize(x->stream) != hipSuccess) {
    g_create_error = "create: initial clear failed";
    mhmhip_destroy(x);
    return MHMHIP_EHIP;
  }
  *out = x;
  return MHMHIP_OK;
}

extern "C" int mhmhip_destroy(mhmhip_ctx *x) {
  if (!x) return MHMHIP_OK;
  (void)hipSetDevice(x->cfg.device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  (void)hipFree(x->P);
  (void)hipFree(x->S);
  (void)hipFree(x->X);
  (void)hipFree(x->F);
  (void)hipFree(x->W);
  (void)hipFree(x->d_agg_start);
  (void)hipFree(x->d_cring);
  (void)hipFree(x->d_meta);
  (void)hipFree(x->d_area);
  (void)hipFree(x->d_C1);
  (void)hipFree(x->d_C2);
  (void)hipFree(x->d_qTIN);
  (void)hipFree(x->d_pub);
  (void)hipFree(x->d_qOUT);
  (void)hipFree(x->d_dom_start);
  (void)hipFree(x->d_dom_dmax);
  (void)hipFree(x->d_child_list);
  (void)hipFree(x->d_slots);
  (void)hipFree(x->d_bins);
  (void)hipFree(x->d_bnd);
  (void)hipFree(x->d_gauge_node);
  (void)hipFree(x->d_node_gauge);
  (void)hipFree(x->d_lag);
  (void)hipFree(x->d_slot_out);
  (void)hipFree(x->d_qacc);
  (void)hipFree(x->d_out);
  (void)hipFree(x->d_qsub);
  (void)hipFree(x->d_runacc);
  (void)hipFree(x->d_inflow_avg);
  (void)hipFree(x->d_inflow_node);
  (void)hipFree(x->d_inflow_hw);
  (void)hipFree(x->d_inflow_q);
  (void)hipFree(x->d_ring);
  (void)hipFree(x->d_gauge_buf);
  (void)hipFree(x->d_steps);
  for (hipEvent_t e : x->stage_ev)
    if (e) (void)hipEventDestroy(e);
  if (x->stage) (void)hipHostFree(x->stage);
  if (x->cstream) (void)hipStreamSynchronize(x->cstream);
  for (int i = 0; i < 2; ++i) {
    if (x->fstage_h[i]) (void)hipHostFree(x->fstage_h[i]);
    (void)hipFree(x->fstage_d[i]);
    if (x->fstage_ev[i]) (void)hipEventDestroy(x->fstage_ev[i]);
  }
  (void)hipFree(x->d_pos);
  (void)hipFree(x->d_mpr);
  for (hipEvent_t e : x->chunk_ev) (void)hipEventDestroy(e);
  if (x->cstream) (void)hipStreamDestroy(x->cstream);
  for (auto &p : x->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : x->ev_pool) (void)hipEventDestroy(e);
  drop_graph(x);
  if (x->rstream) (void)hipStreamDestroy(x->rstream);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return MHMHIP_OK;
}

extern "C" const char *mhmhip_last_error(const mhmhip_ctx *x) {
  return x ? x->err.c_str() : g_create_error.c_str();
}

extern "C" int64_t mhmhip_device_bytes(const mhmhip_ctx *x) { return x ? x->device_bytes : 0; }

extern "C" int mhmhip_sync(mhmhip_ctx *x) {
  if (!x) return MHMHIP_EINVAL;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "sync");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "sync");
  return MHMHIP_OK;
}

// ---------------------------------------------------------------------------
// cell data

namespace {

// The next pinned staging row, free once its previous upload has left it.
// Positions past `used` (the stride's padding) keep the zeros they were
// cleared to; a download through the row clears them again.
int stage_take(mhmhip_ctx *x, double **row) {
  const int i = x->stage_next;
  x->stage_next = (i + 1) % mhmhip_ctx::kStageSlots;
  CHECK_HIP(x, hipEventSynchronize(x->stage_ev[i]), "stage wait");
  *row = x->stage + (int64_t)i * x->sstride;
  return MHMHIP_OK;
}

int stage_send(mhmhip_ctx *x, double *dst, const double *row) {
  const int i = (int)((row - x->stage) / x->sstride);
  CHECK_HIP(x, hipMemcpyAsync(dst, row, sizeof(double) * x->sstride, hipMemcpyHostToDevice,
                              x->stream),
            "upload");
  CHECK_HIP(x, hipEventRecord(x->stage_ev[i], x->stream), "upload");
  return MHMHIP_OK;
}

// upload one device row slice from host column col (reference cell order),
// applying f to every value
template <typename Fn>
int put_row(mhmhip_ctx *x, double *dst, const double *src, Fn f) {
  double *row = nullptr;
  CHECK_RC(stage_take(x, &row));
  for (int64_t c = 0; c < x->n_cells; ++c) row[x->pos[c]] = f(src[c]);
  return stage_send(x, dst, row);
}

// one row of node values in reference node order -> device node order
int put_node_row(mhmhip_ctx *x, double *dst, const double *src) {
  double *row = nullptr;
  CHECK_RC(stage_take(x, &row));
  // past the nodes: zeros (a cell row written later overwrites [0, n_cells))
  std::fill(row + x->n_nodes, row + x->sstride, 0.0);
  for (int64_t k = 0; k < x->n_nodes; ++k) row[x->npos[k]] = src[k];
  return stage_send(x, dst, row);
}

int get_row(mhmhip_ctx *x, const double *srcdev, double *dst) {
  double *row = nullptr;
  CHECK_RC(stage_take(x, &row));
  CHECK_HIP(x, hipMemcpyAsync(row, srcdev, sizeof(double) * x->sstride, hipMemcpyDeviceToHost,
                              x->stream),
            "download");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "download");
  for (int64_t c = 0; c < x->n_cells; ++c) dst[c] = row[x->pos[c]];
  std::fill(row + x->n_cells, row + x->sstride, 0.0);  // the padding a cell row leaves
  return MHMHIP_OK;
}

int var_cols(const mhmhip_ctx *x, int var) {
  switch (var) {
    case MHMHIP_P_SOILMOIST_SAT:
    case MHMHIP_P_FROOTS:
    case MHMHIP_P_SOILMOIST_FC:
    case MHMHIP_P_WILTINGPOINT:
    case MHMHIP_P_SOILMOIST_EXP:
    case MHMHIP_S_SOILMOIST:
    case MHMHIP_X_INFILSOIL:
    case MHMHIP_X_AETSOIL:
      return x->nh;
    case MHMHIP_P_MAXINTER:
    case MHMHIP_P_PETLAICOR:
    case MHMHIP_P_PRIETAYALPHA:
    case MHMHIP_P_AERORESIST:
    case MHMHIP_P_SURFRESIST:
      return 12;
    default:
      return 1;
  }
}

}  // namespace

extern "C" int mhmhip_set_cell_var(mhmhip_ctx *x, int32_t var, int32_t set, const double *host,
                                   int64_t ld, int32_t ncols) {
  if (!x) return MHMHIP_EINVAL;
  if (!host || set < 0 || set >= x->n_sets || ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "set_cell_var: bad pointer, set or leading dimension");
  const bool is_param = var >= 0 && var < MHMHIP_P_COUNT;
  const bool is_state = var >= MHMHIP_S_INTER && var < MHMHIP_S_END;
  if (!is_param && !is_state) return x->fail(MHMHIP_EINVAL, "set_cell_var: unknown or read-only variable");
  (is_param ? x->p_dirty : x->s_dirty) = true;
  if (ncols != var_cols(x, var)) return x->fail(MHMHIP_EINVAL, "set_cell_var: wrong column count");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_cell_var");
  const int64_t off = (int64_t)set * x->sstride;
  auto P = [&](int row) { return x->P + row * x->ld + off; };
  auto id = [](double v) { return v; };
  const int nh = x->nh;
  const double c2 = x->c2TSTu;
  switch (var) {
    case MHMHIP_P_FSEALED: return put_row(x, P(PR_FS), host, id);
    case MHMHIP_P_DEGDAY_INC: return put_row(x, P(PR_INCR), host, id);
    // mo_mhm.f90:463-464 pass deg_day_max*c2TSTu and deg_day_noprec*c2TSTu
    case MHMHIP_P_DEGDAY_MAX: return put_row(x, P(PR_DMAXS), host, [c2](double v) { return v * c2; });
    case MHMHIP_P_DEGDAY_NOPRE: return put_row(x, P(PR_DNPS), host, [c2](double v) { return v * c2; });
    case MHMHIP_P_TEMP_THRESH: return put_row(x, P(PR_TTH), host, id);
    case MHMHIP_P_SEALED_THRESH: return put_row(x, P(PR_SEALTHR), host, id);
    case MHMHIP_P_JARVIS_C1: return put_row(x, P(PR_JARVIS), host, id);
    // mo_mhm.f90:484,488 pass c2TSTu/k0, c2TSTu/k1, c2TSTu/kp, c2TSTu/k2
    case MHMHIP_P_KFASTFLOW: return put_row(x, P(PR_K0S), host, [c2](double v) { return c2 / v; });
    case MHMHIP_P_KSLOWFLOW: return put_row(x, P(PR_K1S), host, [c2](double v) { return c2 / v; });
    case MHMHIP_P_KBASEFLOW: return put_row(x, P(PR_K2S), host, [c2](double v) { return c2 / v; });
    case MHMHIP_P_KPERCO: return put_row(x, P(PR_KPS), host, [c2](double v) { return c2 / v; });
    case MHMHIP_P_ALPHA: return put_row(x, P(PR_ALPHA), host, id);
    case MHMHIP_P_KARSTLOSS: return put_row(x, P(PR_KARST), host, id);
    case MHMHIP_P_UNSAT_THRESH: return put_row(x, P(PR_UNSATTHR), host, id);
    case MHMHIP_P_FASP: return put_row(x, P(PR_FASP), host, id);
    case MHMHIP_P_HARSAMCOEFF: return put_row(x, P(PR_HARSAM), host, id);
    case MHMHIP_P_LATITUDE:
      // mo_pet.f90:113 deg2rad_dp * latitude, then tan/sin/cos of it (331-353)
      CHECK_RC(put_row(x, P(PR_TANLAT), host, [](double v) { return std::tan(kDeg2Rad * v); }));
      CHECK_RC(put_row(x, P(PR_SINLAT), host, [](double v) { return std::sin(kDeg2Rad * v); }));
      return put_row(x, P(PR_COSLAT), host, [](double v) { return std::cos(kDeg2Rad * v); });
    default: break;
  }
  int row0 = -1;
  double *base = x->P;
  switch (var) {
    case MHMHIP_P_SOILMOIST_SAT: row0 = pr_sat(nh); break;
    case MHMHIP_P_FROOTS: row0 = pr_froots(nh); break;
    case MHMHIP_P_SOILMOIST_FC: row0 = pr_fc(nh); break;
    case MHMHIP_P_WILTINGPOINT: row0 = pr_wp(nh); break;
    case MHMHIP_P_SOILMOIST_EXP: row0 = pr_exp(nh); break;
    case MHMHIP_P_MAXINTER: row0 = pr_maxinter(nh); break;
    case MHMHIP_P_PETLAICOR: row0 = pr_petlai(nh); break;
    case MHMHIP_P_PRIETAYALPHA: row0 = pr_prietay(nh); break;
    case MHMHIP_P_AERORESIST: row0 = pr_aero(nh); break;
    case MHMHIP_P_SURFRESIST: row0 = pr_surf(nh); break;
    case MHMHIP_S_INTER: row0 = SR_INTER; base = x->S; break;
    case MHMHIP_S_SNOWPACK: row0 = SR_SNOW; base = x->S; break;
    case MHMHIP_S_SEALSTW: row0 = SR_SEAL; base = x->S; break;
    case MHMHIP_S_UNSATSTW: row0 = SR_UNSAT; base = x->S; break;
    case MHMHIP_S_SATSTW: row0 = SR_SAT; base = x->S; break;
    case MHMHIP_S_SOILMOIST: row0 = SR_NSCALAR; base = x->S; break;
    default: return x->fail(MHMHIP_EINVAL, "set_cell_var: unknown variable");
  }
  // the host values are in the staging rows when this returns; the uploads
  // finish on the context's stream ahead of any later work
  for (int col = 0; col < ncols; ++col)
    CHECK_RC(put_row(x, base + (row0 + col) * x->ld + off, host + col * ld, id));
  return MHMHIP_OK;
}

extern "C" int mhmhip_get_cell_var(mhmhip_ctx *x, int32_t var, int32_t set, double *host,
                                   int64_t ld, int32_t ncols) {
  if (!x) return MHMHIP_EINVAL;
  if (!host || set < 0 || set >= x->n_sets || ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "get_cell_var: bad pointer, set or leading dimension");
  if (ncols != var_cols(x, var)) return x->fail(MHMHIP_EINVAL, "get_cell_var: wrong column count");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "get_cell_var");
  const int nh = x->nh;
  int row0 = -1;
  double *base = nullptr;
  if (var >= MHMHIP_S_INTER && var < MHMHIP_S_END) {
    base = x->S;
    row0 = (var == MHMHIP_S_SOILMOIST) ? SR_NSCALAR : (var - MHMHIP_S_INTER);
  } else if (var >= MHMHIP_X_PET_CALC && var < MHMHIP_X_END) {
    base = x->X;
    if (var == MHMHIP_X_INFILSOIL)
      row0 = xr_infil(nh);
    else if (var == MHMHIP_X_AETSOIL)
      row0 = xr_aet(nh);
    else
      row0 = var - MHMHIP_X_PET_CALC;
  } else {
    return x->fail(MHMHIP_EINVAL, "get_cell_var: only states and fluxes can be read");
  }
  const int64_t off = (int64_t)set * x->sstride;
  for (int col = 0; col < ncols; ++col)
    CHECK_RC(get_row(x, base + (row0 + col) * x->ld + off, host + col * ld));
  return MHMHIP_OK;
}

extern "C" int mhmhip_default_states(mhmhip_ctx *x, const double *hd) {
  if (!x) return MHMHIP_EINVAL;
  if (!hd && x->nh > 1) return x->fail(MHMHIP_EINVAL, "default_states: horizon depths required");
  x->s_dirty = x->x_dirty = true;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "default_states");
  std::vector<double> v(x->n_cells);
  auto fill = [&](double val) { std::fill(v.begin(), v.end(), val); };
  auto id = [](double a) { return a; };
  for (int s = 0; s < x->n_sets; ++s) {
    const int64_t off = (int64_t)s * x->sstride;
    // mo_init_states.f90:273-287 with mo_mpr_constants P1..P5, C1_InitStateSM
    fill(0.0);
    CHECK_RC(put_row(x, x->S + SR_INTER * x->ld + off, v.data(), id));
    CHECK_RC(put_row(x, x->S + SR_SEAL * x->ld + off, v.data(), id));
    fill(15.0);
    CHECK_RC(put_row(x, x->S + SR_SNOW * x->ld + off, v.data(), id));
    fill(10.0);
    CHECK_RC(put_row(x, x->S + SR_UNSAT * x->ld + off, v.data(), id));
    fill(75.0);
    CHECK_RC(put_row(x, x->S + SR_SAT * x->ld + off, v.data(), id));
    for (int h = 0; h < x->nh - 1; ++h) {
      fill((h == 0 ? hd[0] : hd[h] - hd[h - 1]) * 0.25);
      CHECK_RC(put_row(x, x->S + (SR_NSCALAR + h) * x->ld + off, v.data(), id));
    }
    fill((1500.0 - (x->nh >= 2 ? hd[x->nh - 2] : 0.0)) * 0.25);
    CHECK_RC(put_row(x, x->S + (SR_NSCALAR + x->nh - 1) * x->ld + off, v.data(), id));
    fill(0.0);
    for (int r = 0; r < x->n_xrows; ++r)
      CHECK_RC(put_row(x, x->X + r * x->ld + off, v.data(), id));
  }
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "default_states");
  return MHMHIP_OK;
}

namespace {

// Columns day0 - 1 .. day0 - 1 + ndays - 1 written synchronously (resident
// forcing): streamed uploads still in flight finish first, and the columns
// hold those days from now on.
int forcing_columns_direct(mhmhip_ctx *x, int32_t day0, int32_t ndays) {
  if (x->cstream) CHECK_HIP(x, hipStreamSynchronize(x->cstream), "forcing");
  for (int32_t j = 0; j < ndays; ++j) {
    x->col_day[day0 - 1 + j] = day0 + j;
    x->col_seq[day0 - 1 + j] = -1;
  }
  return MHMHIP_OK;
}

// The compute stream waits for the streamed chunks that wrote the columns
// the steps read (chunks complete in order on the copy stream: waiting for
// the newest is enough).
int forcing_wait(mhmhip_ctx *x, const StepDev *st, int64_t n) {
  int64_t newest = -1;
  for (int64_t k = 0; k < n; ++k) newest = std::max(newest, x->col_seq[st[k].day]);
  if (newest > x->waited_seq) {
    CHECK_HIP(x, hipStreamWaitEvent(x->stream, x->chunk_ev[newest], 0), "forcing wait");
    x->waited_seq = newest;
  }
  return MHMHIP_OK;
}

}  // namespace

extern "C" int mhmhip_stream_forcing(mhmhip_ctx *x, int32_t day0, int32_t ndays,
                                     const double *const *host, int64_t ld) {
  if (!x) return MHMHIP_EINVAL;
  const int32_t nfd = x->cfg.n_forcing_days;
  if (!host || day0 < 1 || ndays < 1 || ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "stream_forcing: bad day range, pointers or leading dimension");
  if (nfd < 2 || ndays > nfd || ndays > 65535)
    return x->fail(MHMHIP_EINVAL, "stream_forcing: a chunk must fit the " + std::to_string(nfd) +
                                      " forcing columns (at least 2; two chunks for double "
                                      "buffering)");
  // the given variables this context holds, in variable order (the others
  // are not read by its PET case and are skipped)
  FRowMap map{};
  for (int v = 0; v < MHMHIP_F_COUNT; ++v)
    if (host[v] && x->frow[v] >= 0) map.row[map.n++] = x->frow[v];
  const int32_t nv = map.n;
  if (nv == 0) {
    bool any = false;
    for (int v = 0; v < MHMHIP_F_COUNT; ++v) any = any || host[v] != nullptr;
    if (!any) return x->fail(MHMHIP_EINVAL, "stream_forcing: no variable given");
  }
  x->f_dirty = true;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "stream_forcing");
  if (!x->cstream) {
    CHECK_HIP(x, hipStreamCreateWithFlags(&x->cstream, hipStreamNonBlocking), "copy stream");
    for (int i = 0; i < 2; ++i)
      CHECK_HIP(x, hipEventCreateWithFlags(&x->fstage_ev[i], hipEventDisableTiming),
                "stream_forcing event");
  }
  if (x->d_pos_stale) {
    CHECK_HIP(x, hipStreamSynchronize(x->cstream), "stream_forcing");
    x->dfree(x->d_pos, x->n_cells);
    CHECK_RC(x->dalloc(&x->d_pos, x->n_cells, "alloc cell positions"));
    std::vector<int32_t> p32(x->n_cells);
    for (int64_t c = 0; c < x->n_cells; ++c) p32[c] = (int32_t)x->pos[c];
    CHECK_HIP(x, sync_copy(x, x->d_pos, p32.data(), sizeof(int32_t) * x->n_cells,
                           hipMemcpyHostToDevice), "upload cell positions");
    x->d_pos_stale = false;
  }
  const int64_t need = (int64_t)nv * ndays * x->n_cells;
  const int i = x->fstage_next;
  x->fstage_next ^= 1;
  if (x->fstage_cap < need) {
    // both slots grow together, once nothing is in flight from them
    CHECK_HIP(x, hipStreamSynchronize(x->cstream), "stream_forcing");
    for (int k = 0; k < 2; ++k) {
      if (x->fstage_h[k]) (void)hipHostFree(x->fstage_h[k]);
      x->fstage_h[k] = nullptr;
      x->dfree(x->fstage_d[k], x->fstage_cap);
    }
    x->fstage_cap = 0;
    for (int k = 0; k < 2; ++k) {
      if (hipHostMalloc((void **)&x->fstage_h[k], sizeof(double) * need, hipHostMallocDefault) !=
          hipSuccess)
        return x->fail(MHMHIP_ENOMEM, "stream_forcing: pinned staging allocation failed");
      CHECK_RC(x->dalloc(&x->fstage_d[k], need, "alloc forcing staging"));
    }
    x->fstage_cap = need;
  }
  // the slot's previous upload has left its pinned buffer
  CHECK_HIP(x, hipEventSynchronize(x->fstage_ev[i]), "stream_forcing");
  double *h = x->fstage_h[i];
  int32_t k = 0;
  for (int v = 0; v < MHMHIP_F_COUNT; ++v) {
    if (!host[v] || x->frow[v] < 0) continue;
    for (int32_t j = 0; j < ndays; ++j)
      std::memcpy(h + ((int64_t)k * ndays + j) * x->n_cells, host[v] + (int64_t)j * ld,
                  sizeof(double) * x->n_cells);
    ++k;
  }
  // the columns are overwritten after the compute work queued so far, which
  // may still read them
  hipEvent_t after = nullptr;
  CHECK_HIP(x, hipEventCreateWithFlags(&after, hipEventDisableTiming), "stream_forcing event");
  hipError_t e = hipEventRecord(after, x->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(x->cstream, after, 0);
  (void)hipEventDestroy(after);
  CHECK_HIP(x, e, "stream_forcing order");
  if (nv > 0)
    CHECK_HIP(x, hipMemcpyAsync(x->fstage_d[i], h, sizeof(double) * need, hipMemcpyHostToDevice,
                                x->cstream), "stream_forcing upload");
  CHECK_HIP(x, hipEventRecord(x->fstage_ev[i], x->cstream), "stream_forcing event");
  if (nv > 0)
    CHECK_HIP(x, launch_forcing_gather(x->fstage_d[i], map, ndays, x->n_cells, x->d_pos, x->F, nfd,
                                       day0 - 1, x->sstride, x->cstream),
              "stream_forcing gather");
  hipEvent_t done = nullptr;
  CHECK_HIP(x, hipEventCreateWithFlags(&done, hipEventDisableTiming), "stream_forcing event");
  x->chunk_ev.push_back(done);
  CHECK_HIP(x, hipEventRecord(done, x->cstream), "stream_forcing event");
  const int64_t seq = (int64_t)x->chunk_ev.size() - 1;
  for (int32_t j = 0; j < ndays; ++j) {
    const int32_t col = (day0 - 1 + j) % nfd;
    x->col_day[col] = day0 + j;
    x->col_seq[col] = seq;
  }
  return MHMHIP_OK;
}

extern "C" int mhmhip_set_forcing(mhmhip_ctx *x, int32_t var, int32_t day0, int32_t ndays,
                                  const double *host, int64_t ld) {
  if (!x) return MHMHIP_EINVAL;
  if (!host || var < 0 || var >= MHMHIP_F_COUNT || day0 < 1 || ndays < 1 ||
      (int64_t)day0 - 1 + ndays > x->cfg.n_forcing_days || ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "set_forcing: variable or day range out of bounds");
  if (x->frow[var] < 0) return MHMHIP_OK;  // not read by this PET case
  x->f_dirty = true;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_forcing");
  CHECK_RC(forcing_columns_direct(x, day0, ndays));
  auto id = [](double a) { return a; };
  for (int32_t j = 0; j < ndays; ++j) {
    double *dst =
        x->F + ((int64_t)x->frow[var] * x->cfg.n_forcing_days + (day0 - 1 + j)) * x->sstride;
    CHECK_RC(put_row(x, dst, host + (int64_t)j * ld, id));
  }
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_forcing");
  return MHMHIP_OK;
}

namespace {
// device buffers of one call, freed on every return path
struct CallBufs {
  double *data2 = nullptr;
  uint8_t *mask2 = nullptr;
  MeteoL2Cell *cells = nullptr;
  ~CallBufs() {
    (void)hipFree(data2);
    (void)hipFree(mask2);
    (void)hipFree(cells);
  }
};
}  // namespace

extern "C" int mhmhip_set_forcing_l2(mhmhip_ctx *x, int32_t var, int32_t domain, int32_t day0,
                                     int32_t ndays, const double *data2, int32_t nr2,
                                     int32_t nc2, double cs2, const uint8_t *mask2, int32_t nr1,
                                     int32_t nc1, double cs1, const uint8_t *mask1) {
  if (!x) return MHMHIP_EINVAL;
  if (!data2 || !mask2 || !mask1 || var < 0 || var >= MHMHIP_F_COUNT || domain < 0 ||
      domain >= x->n_domains || day0 < 1 || ndays < 1 ||
      (int64_t)day0 - 1 + ndays > x->cfg.n_forcing_days || nr2 < 1 || nc2 < 1 || nr1 < 1 ||
      nc1 < 1 || (int64_t)nr2 * nc2 >= INT32_MAX || !(cs1 > 0.0) || !(cs2 > 0.0))
    return x->fail(MHMHIP_EINVAL,
                   "set_forcing_l2: variable, domain, day range, grid or cell size out of bounds");
  const int64_t n1 = (int64_t)nr1 * nc1, n2 = (int64_t)nr2 * nc2;
  const int64_t c_lo = x->dom_cell_start[domain], c_hi = x->dom_cell_start[domain + 1];
  int64_t ncells = 0;
  for (int64_t k = 0; k < n1; ++k) ncells += mask1[k] != 0;
  if (ncells != c_hi - c_lo)
    return x->fail(MHMHIP_EINVAL, "set_forcing_l2: mask1 must hold the domain's " +
                                      std::to_string(c_hi - c_lo) + " cells, it holds " +
                                      std::to_string(ncells));
  if (x->frow[var] < 0) return MHMHIP_OK;  // not read by this PET case
  x->f_dirty = true;
  // cellFactorHbyM (mo_meteo_helper.f90:120) picks the mapping
  const double hbym = cs1 / cs2;
  const bool aggregate = hbym > 1.0, equal = !(hbym > 1.0) && !(hbym < 1.0);
  std::vector<MeteoL2Cell> cells((size_t)ncells);
  if (aggregate) {
    // ceiling(i / cellFactor) as in spatial_aggregation_3d: the L2 rows and
    // columns of every L1 row and column (contiguous: the map is monotone)
    const double f = cs1 / cs2;
    auto ceil_div = [f](int32_t i) { return (int64_t)std::ceil((double)i / f); };
    std::vector<int32_t> rlo(nr1, INT32_MAX), rhi(nr1, -1), clo(nc1, INT32_MAX), chi(nc1, -1);
    for (int32_t i = 1; i <= nr2; ++i) {
      const int64_t ic = ceil_div(i);
      if (ic >= 1 && ic <= nr1) {
        rlo[ic - 1] = std::min(rlo[ic - 1], i - 1);
        rhi[ic - 1] = std::max(rhi[ic - 1], i - 1);
      }
    }
    for (int32_t j = 1; j <= nc2; ++j) {
      const int64_t jc = ceil_div(j);
      if (jc >= 1 && jc <= nc1) {
        clo[jc - 1] = std::min(clo[jc - 1], j - 1);
        chi[jc - 1] = std::max(chi[jc - 1], j - 1);
      }
    }
    // nTCells: the valid L2 cells of each L1 cell
    std::vector<int32_t> ntc((size_t)n1, 0);
    for (int32_t j = 1; j <= nc2; ++j) {
      const int64_t jc = ceil_div(j);
      for (int32_t i = 1; i <= nr2; ++i) {
        if (!mask2[(int64_t)(j - 1) * nr2 + (i - 1)]) continue;
        const int64_t ic = ceil_div(i);
        if (ic > nr1 || jc > nc1)
          return x->fail(MHMHIP_EINVAL, "set_forcing_l2: a valid L2 cell lies outside the L1 grid");
        ntc[(size_t)((jc - 1) * nr1 + (ic - 1))] += 1;
      }
    }
    int64_t c = 0;
    for (int32_t j = 0; j < nc1; ++j)
      for (int32_t i = 0; i < nr1; ++i) {
        const int64_t k = (int64_t)j * nr1 + i;
        if (!mask1[k]) continue;
        MeteoL2Cell &m = cells[(size_t)c];
        m.r0 = rlo[i];
        m.r1 = rhi[i];
        m.c0 = clo[j];
        m.c1 = chi[j];
        m.n = ntc[(size_t)k];
        m.src = -1;
        m.dst = x->pos[c_lo + c];
        ++c;
      }
  } else {
    if (equal && (nr1 != nr2 || nc1 != nc2))
      return x->fail(MHMHIP_EINVAL, "set_forcing_l2: equal cell sizes need equal grids");
    const double f = cs2 / cs1;
    auto ceil_div = [f](int32_t i) { return (int64_t)std::ceil((double)i / f); };
    int64_t c = 0;
    for (int32_t j = 0; j < nc1; ++j)
      for (int32_t i = 0; i < nr1; ++i) {
        const int64_t k = (int64_t)j * nr1 + i;
        if (!mask1[k]) continue;
        int64_t src = k;
        if (!equal) {
          // spatial_disaggregation_3d: the L2 cell (ceiling(i / f), ceiling(j / f))
          const int64_t ic = ceil_div(i + 1), jc = ceil_div(j + 1);
          if (ic > nr2 || jc > nc2)
            return x->fail(MHMHIP_EINVAL,
                           "set_forcing_l2: a valid L1 cell lies outside the L2 grid");
          const int64_t k2 = (jc - 1) * nr2 + (ic - 1);
          src = mask2[k2] ? k2 : -1;
        }
        cells[(size_t)c] = MeteoL2Cell{0, -1, 0, -1, 0, (int32_t)src, x->pos[c_lo + c]};
        ++c;
      }
  }
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_forcing_l2");
  CHECK_RC(forcing_columns_direct(x, day0, ndays));
  CallBufs b;
  CHECK_HIP(x, hipMalloc((void **)&b.data2, sizeof(double) * n2 * ndays), "set_forcing_l2 alloc");
  CHECK_HIP(x, hipMalloc((void **)&b.mask2, (size_t)n2), "set_forcing_l2 alloc");
  CHECK_HIP(x, hipMalloc((void **)&b.cells, sizeof(MeteoL2Cell) * std::max<int64_t>(1, ncells)),
            "set_forcing_l2 alloc");
  CHECK_HIP(x, hipMemcpyAsync(b.data2, data2, sizeof(double) * n2 * ndays, hipMemcpyHostToDevice,
                              x->stream), "set_forcing_l2 upload");
  CHECK_HIP(x, hipMemcpyAsync(b.mask2, mask2, (size_t)n2, hipMemcpyHostToDevice, x->stream),
            "set_forcing_l2 upload");
  CHECK_HIP(x, hipMemcpyAsync(b.cells, cells.data(), sizeof(MeteoL2Cell) * ncells,
                              hipMemcpyHostToDevice, x->stream), "set_forcing_l2 upload");
  double *dst = x->F + ((int64_t)x->frow[var] * x->cfg.n_forcing_days + (day0 - 1)) * x->sstride;
  CHECK_HIP(x, launch_meteo_l2(b.data2, b.mask2, nr2, n2, b.cells, ncells, ndays, dst, x->sstride,
                               aggregate, x->stream), "set_forcing_l2 map");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_forcing_l2");
  return MHMHIP_OK;
}

extern "C" int mhmhip_mprlite_inputs(mhmhip_ctx *x, const double *cells, int64_t ld,
                                    const double *hd, const double *lai_class) {
  if (!x) return MHMHIP_EINVAL;
  if (!cells || !hd || !lai_class || ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "mprlite_inputs: bad pointer or leading dimension");
  for (int h = 0; h < x->nh; ++h)
    if (!(hd[h] > (h ? hd[h - 1] : 0.0)))
      return x->fail(MHMHIP_EINVAL, "mprlite_inputs: horizon depths must increase from above 0");
  x->mpr_cells.resize((size_t)6 * x->n_cells);
  for (int k = 0; k < 6; ++k)
    std::copy(cells + k * ld, cells + k * ld + x->n_cells, x->mpr_cells.begin() + k * x->n_cells);
  std::copy(hd, hd + x->nh, x->mpr_hd);
  std::copy(lai_class, lai_class + 36, x->mpr_lai);
  x->mpr_stale = true;
  return MHMHIP_OK;
}

extern "C" int mhmhip_mprlite_params(mhmhip_ctx *x, int32_t set, const double *gamma) {
  if (!x) return MHMHIP_EINVAL;
  if (!gamma || set < 0 || set >= x->n_sets)
    return x->fail(MHMHIP_EINVAL, "mprlite_params: bad pointer or set");
  if (x->mpr_cells.empty())
    return x->fail(MHMHIP_ESTATE, "mprlite_params: no inputs (call mprlite_inputs first)");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "mprlite_params");
  const int64_t n = x->n_cells;
  if (x->mpr_stale) {
    if (!x->d_mpr) CHECK_RC(x->dalloc(&x->d_mpr, 8 * x->sstride, "alloc MPR-lite inputs"));
    auto id = [](double v) { return v; };
    for (int k = 0; k < 5; ++k)
      CHECK_RC(put_row(x, x->d_mpr + k * x->sstride, x->mpr_cells.data() + k * n, id));
    // the latitude as mhmhip_set_cell_var stores it (mo_pet.f90:113, 331-353)
    const double *lat = x->mpr_cells.data() + 5 * n;
    CHECK_RC(put_row(x, x->d_mpr + 5 * x->sstride, lat,
                     [](double v) { return std::tan(kDeg2Rad * v); }));
    CHECK_RC(put_row(x, x->d_mpr + 6 * x->sstride, lat,
                     [](double v) { return std::sin(kDeg2Rad * v); }));
    CHECK_RC(put_row(x, x->d_mpr + 7 * x->sstride, lat,
                     [](double v) { return std::cos(kDeg2Rad * v); }));
    x->mpr_stale = false;
  }
  MprLiteArgs a{};
  a.P = x->P;
  a.in = x->d_mpr;
  a.ld = x->ld;
  a.sstride = x->sstride;
  a.n_cells = n;
  a.nh = x->nh;
  a.pet_case = x->cfg.process_pet;
  a.c2 = x->c2TSTu;
  std::copy(x->mpr_hd, x->mpr_hd + x->nh, a.hd);
  std::copy(x->mpr_lai, x->mpr_lai + 36, a.lai);
  std::copy(gamma, gamma + kMprNGamma, a.gamma);
  x->p_dirty = true;
  CHECK_HIP(x, launch_mprlite(a, set, x->stream), "mprlite_params");
  return MHMHIP_OK;
}

extern "C" int mhmhip_set_meteo_weights(mhmhip_ctx *x, const double *pre_w,
                                        const double *temp_w, const double *pet_w, int64_t ld) {
  if (!x) return MHMHIP_EINVAL;
  const int given = (pre_w != nullptr) + (temp_w != nullptr) + (pet_w != nullptr);
  if (given != 0 && given != 3)
    return x->fail(MHMHIP_EINVAL, "set_meteo_weights: give all three weight arrays or none");
  if (given == 3 && ld < x->n_cells)
    return x->fail(MHMHIP_EINVAL, "set_meteo_weights: leading dimension below the cell count");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_meteo_weights");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_meteo_weights");
  drop_graph(x);  // a captured run holds the kernels' arguments
  if (given == 0) {
    x->dfree(x->W, (int64_t)3 * 12 * 24 * x->sstride);
    x->W = nullptr;
    return MHMHIP_OK;
  }
  if (!x->W) CHECK_RC(x->dalloc(&x->W, (int64_t)3 * 12 * 24 * x->sstride, "alloc meteo weights"));
  // W[v][month * 24 + hour][sstride], v = pre, temp, pet; the host arrays are
  // L1_x_weights(cell, month, hour + 1)
  const double *host[3] = {pre_w, temp_w, pet_w};
  auto id = [](double a) { return a; };
  for (int v = 0; v < 3; ++v)
    for (int m = 0; m < 12; ++m)
      for (int h = 0; h < 24; ++h)
        CHECK_RC(put_row(x, x->W + ((int64_t)v * 288 + m * 24 + h) * x->sstride,
                         host[v] + ld * ((int64_t)m + 12 * h), id));
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_meteo_weights");
  return MHMHIP_OK;
}

// ---------------------------------------------------------------------------
// step scalars

namespace {

int make_step(mhmhip_ctx *x, const mhmhip_step &s, StepDev &d) {
  if (s.tt < 1 || s.month < 1 || s.month > 12 || s.hour < 0 || s.hour > 23 || s.doy < 1 ||
      s.doy > 366 || s.iMeteoTS < 1)
    return x->fail(MHMHIP_EINVAL, "step: tt/month/hour/doy/iMeteoTS out of range");
  const int32_t col = (s.iMeteoTS - 1) % x->cfg.n_forcing_days;
  if (x->col_day[col] != s.iMeteoTS)
    return x->fail(MHMHIP_EINVAL, "step: forcing day " + std::to_string(s.iMeteoTS) +
                                      " (iMeteoTS) is not on the device: its column holds day " +
                                      std::to_string(x->col_day[col]) + " (" +
                                      std::to_string(x->cfg.n_forcing_days) +
                                      " columns; stream the day first)");
  // mo_pet.f90:339-340, evaluated with the same libm as the reference
  const double doy = (double)s.doy;
  d.dr = 1.0 + 0.0330 * std::cos(kTwoPiD * doy / kYearDays);
  const double delta = 0.4090 * std::sin(kTwoPiD * doy / kYearDays - 1.3900);
  d.tand = std::tan(delta);
  d.sind = std::sin(delta);
  d.cosd = std::cos(delta);
  d.day = col;
  d.month0 = s.month - 1;
  d.isday = (s.hour > 6) && (s.hour <= 18);  // mo_meteo_handler.f90:1048
  const int m = s.month - 1;
  d.f_pet = d.isday ? x->mt.fday_pet[m] : x->mt.fnight_pet[m];
  d.f_prec = d.isday ? x->mt.fday_prec[m] : x->mt.fnight_prec[m];
  d.f_temp = d.isday ? x->mt.fday_temp[m] : x->mt.fnight_temp[m];
  d.evap = x->mt.evap[m];
  d.revap = 1.0 / d.evap;
  d.init_sm = (s.tt == 1) && !s.read_states;  // mo_mhm.f90:448
  d.doy = s.doy;
  d.meteo_same = 0;
  d.wrow = m * 24 + s.hour;
  return MHMHIP_OK;
}

CellArgs cell_args(mhmhip_ctx *x, bool ring) {
  CellArgs a{};
  a.P = x->P;
  a.S = x->S;
  a.X = x->X;
  a.F = x->F;
  a.ring = ring ? x->d_ring : nullptr;
  a.area = x->d_area;
  a.tst = kHourSecs * (double)x->cfg.timestep;  // mo_mrm_pre_routing.f90:110
  a.rtst = 1.0 / a.tst;
  a.steps = x->d_steps;
  a.n_cells = x->n_cells;
  a.sstride = x->sstride;
  a.ld = x->ld;
  a.forcing_days = x->cfg.n_forcing_days;
  a.ring_len = x->ring_len;
  a.pet_case = x->cfg.process_pet;
  a.jarvis = (x->cfg.process_soil == 2 || x->cfg.process_soil == 3);
  a.hourly = x->cfg.hourly_forcing;
  a.nT = x->nT;
  a.rnT = 1.0 / x->nT;
  a.ring_rep = 1;
  a.W = x->W;
  if (x->out_on) a.out = x->out_cfg;
  else a.out.sum = nullptr;
  return a;
}

bool routing_on(const mhmhip_ctx *x) { return x->cfg.process_route == 1 && x->net_bound; }

int ensure_ring(mhmhip_ctx *x, int64_t len) {
  if (x->ring_len >= len) return MHMHIP_OK;
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "ring");
  x->dfree(x->d_ring, x->ring_len * x->ld);
  x->ring_len = 0;
  CHECK_RC(x->dalloc(&x->d_ring, len * x->ld, "alloc runoff ring"));
  x->ring_len = len;
  return MHMHIP_OK;
}

int ensure_steps(mhmhip_ctx *x, int64_t n) {
  if (x->steps_cap >= n) return MHMHIP_OK;
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "steps");
  x->dfree(x->d_steps, x->steps_cap);
  x->steps_cap = 0;
  CHECK_RC(x->dalloc(&x->d_steps, n, "alloc step table"));
  x->steps_cap = n;
  return MHMHIP_OK;
}

int ensure_gauge_buf(mhmhip_ctx *x, int64_t rows) {
  const int64_t per = (int64_t)x->n_sets * std::max(1, x->n_gauges);
  if (x->gauge_buf_rows >= rows) return MHMHIP_OK;
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "gauges");
  x->dfree(x->d_gauge_buf, x->gauge_buf_rows * per);
  x->gauge_buf_rows = 0;
  CHECK_RC(x->dalloc(&x->d_gauge_buf, rows * per, "alloc gauge series"));
  x->gauge_buf_rows = rows;
  return MHMHIP_OK;
}

RouteArgs route_args(mhmhip_ctx *x, int64_t t_lo, int64_t t_hi, int64_t t_base) {
  RouteArgs r{};
  r.meta = x->d_meta;
  r.child_list = x->d_child_list;
  r.C1 = x->d_C1;
  r.C2 = x->d_C2;
  r.qTIN = x->d_qTIN;
  r.pub = x->d_pub;
  r.qOUT = x->d_qOUT;
  r.qacc = x->qacc_on ? x->d_qacc : nullptr;
  r.ring = x->d_ring;
  r.gauge_out = x->d_gauge_buf;
  r.gauge_node = x->d_gauge_node;
  r.inflow_node = x->d_inflow_node;
  r.inflow_hw = x->d_inflow_hw;
  r.inflow_q = x->d_inflow_q;
  r.n_nodes = x->n_nodes;
  r.sstride = x->sstride;
  r.ring_len = x->ring_len;
  r.t_lo = t_lo;
  r.t_hi = t_hi;
  r.t_base = t_base;
  r.n_sets = x->n_sets;
  r.n_gauges = x->n_gauges;
  r.n_inflow = x->n_inflow;
  r.dmax = x->lay.dmax;
  r.tst = kHourSecs * (double)x->cfg.timestep;  // mo_mrm_pre_routing.f90:110
  r.rep = 1;
  r.acc_from = 0;
  return r;
}

// Largest tree (nodes of one domain) for the workgroup-resident routing.
int64_t max_domain_nodes(const mhmhip_ctx *x) {
  int64_t m = 0;
  for (int d = 0; d < x->n_domains; ++d)
    m = std::max(m, x->dom_node_start[d + 1] - x->dom_node_start[d]);
  return m;
}

// Routing kernel for a route_mode: 1 grid diagonals, 2 workgroup-resident
// trees, 3 and 0 (auto) tiles.
int route_kind(int route_mode) {
  if (route_mode == 1) return MHMHIP_K_ROUTE_DIAG;
  if (route_mode == 2) return MHMHIP_K_ROUTE_WG;
  return MHMHIP_K_ROUTE_TILE;
}

// Steps per routing window of the tiled kernel: the requested one (0: 8), at
// most kTileWindowMax.  The ring spans (tau_max + 1) windows, so smaller
// windows cost launches and save memory.
constexpr int32_t kTileWindowDefault = 8;
int32_t tile_window(int32_t req) {
  const int32_t b = req > 0 ? req : kTileWindowDefault;
  return std::max(1, std::min(b, (int32_t)kTileWindowMax));
}

// Ring rows the routing of a call of nsteps steps uses: the pipelined
// diagonals read a step up to dmax diagonals after its cell launch, the tiles
// up to tau_max windows after; never more than the call's steps.  For the
// tiles the count is a multiple of B, so that no window's rows wrap.
int64_t ring_need(const mhmhip_ctx *x, int kind, int S, int32_t B, int64_t nsteps) {
  if (kind == MHMHIP_K_ROUTE_TILE) {
    const int64_t n = std::min(((int64_t)x->lay.tau_max + 1) * B + S + 1, nsteps);
    return (n + B - 1) / B * B;
  }
  return std::min((int64_t)S + x->lay.dmax + 2, nsteps);
}

TileArgs tile_args(const mhmhip_ctx *x, int32_t B, int64_t nsteps) {
  TileArgs g{};
  g.slot = x->d_slots;
  g.bin = x->d_bins;
  g.node_gauge = x->d_node_gauge;
  g.slot_out = x->d_slot_out;
  g.out_rows = x->tile_out_rows;
  g.bnd = x->d_bnd;
  g.n_bnd = std::max<int64_t>(1, x->lay.n_bnd);
  g.tau_max = x->lay.tau_max;
  g.B = B;
  g.nwin = (int32_t)((nsteps + B - 1) / B);
  g.bstride = kTileWindowMax;
  return g;
}

hipError_t ev_get(mhmhip_ctx *x, hipEvent_t *e) {
  if (!x->ev_pool.empty()) {
    *e = x->ev_pool.back();
    x->ev_pool.pop_back();
    return hipSuccess;
  }
  return hipEventCreate(e);
}

// Launches f() on the context stream; with profiling on, brackets it with
// events: every prof_stride-th launch of each kernel (in a captured graph,
// the choice is made at capture and every replay times the same launches).
template <typename Fn>
hipError_t timed(mhmhip_ctx *x, int kind, double units, Fn f, hipStream_t s = nullptr) {
  if (!s) s = x->stream;
  if (!x->prof) return f();
  (x->capturing ? x->graph_all : x->prof_all)[kind] += 1;
  // a time loop launches kernels by the hundred or thousand, and each event
  // pair costs the stream a few microseconds: time one launch in prof_stride
  if ((x->prof_seq[kind]++ % x->prof_stride) != 0) return f();
  mhmhip_ctx::Pending p{kind, nullptr, nullptr, units};
  hipError_t e;
  if (x->capturing) {
    if ((e = hipEventCreate(&p.a)) != hipSuccess) return e;
    if ((e = hipEventCreate(&p.b)) != hipSuccess) {
      (void)hipEventDestroy(p.a);
      return e;
    }
    x->graph_events.push_back(p);  // owned by the graph from here on
  } else if ((e = ev_get(x, &p.a)) != hipSuccess || (e = ev_get(x, &p.b)) != hipSuccess) {
    return e;
  }
  // under capture a plain record only marks a dependency; the external flag
  // makes it an event-record node that timestamps every replay
  const unsigned flags = x->capturing ? hipEventRecordExternal : 0u;
  if ((e = hipEventRecordWithFlags(p.a, s, flags)) != hipSuccess) return e;
  if ((e = f()) != hipSuccess) return e;
  if ((e = hipEventRecordWithFlags(p.b, s, flags)) != hipSuccess) return e;
  if (!x->capturing) x->pending.push_back(p);
  return hipSuccess;
}

// After a stream join: fold the recorded event pairs into the totals.
int prof_collect(mhmhip_ctx *x) {
  for (auto &p : x->pending) {
    float ms = 0.f;
    CHECK_HIP(x, hipEventElapsedTime(&ms, p.a, p.b), "profile");
    x->prof_ms[p.kind] += ms;
    x->prof_n[p.kind] += 1;
    x->prof_units[p.kind] += p.units;
    x->ev_pool.push_back(p.a);
    x->ev_pool.push_back(p.b);
  }
  x->pending.clear();
  return MHMHIP_OK;
}

// After a replay of a graph captured with profiling on and a stream join.
int prof_collect_graph(mhmhip_ctx *x) {
  for (auto &p : x->graph_events) {
    float ms = 0.f;
    CHECK_HIP(x, hipEventElapsedTime(&ms, p.a, p.b), "profile");
    x->prof_ms[p.kind] += ms;
    x->prof_n[p.kind] += 1;
    x->prof_units[p.kind] += p.units;
  }
  for (int k = 0; k < MHMHIP_K_COUNT; ++k) x->prof_all[k] += x->graph_all[k];
  return MHMHIP_OK;
}

// Node positions that can be active on diagonal d (relative clock): nodes at
// depths with t = d - (dmax - depth) in [t_lo, t_hi).  Single-domain layouts
// keep depths contiguous; packed domains use the full range.
void active_range(const mhmhip_ctx *x, const RouteArgs &r, int64_t d, int64_t &lo, int64_t &hi) {
  lo = 0;
  hi = x->n_nodes;
  if (x->lev_start.empty()) return;
  const int64_t dm = x->lay.dmax;
  const int64_t dep_lo = std::max<int64_t>(0, r.t_lo - d + dm);
  const int64_t dep_hi = std::min<int64_t>(dm, r.t_hi - 1 - d + dm);
  if (dep_lo > dep_hi) {
    lo = hi = 0;
    return;
  }
  lo = x->lev_start[dep_lo];
  hi = x->lev_start[dep_hi + 1];
}

// routing diagonals [d_lo, d_hi) (tiles) or the drain after t_hi
int route_diagonals(mhmhip_ctx *x, const RouteArgs &r, int64_t d_lo, int64_t d_hi, bool drain,
                    bool wg) {
  if (wg) {
    CHECK_HIP(x, timed(x, MHMHIP_K_ROUTE_WG, 0.0, [&] {
                return launch_route_wg(r, x->d_dom_start, x->d_dom_dmax, x->n_domains, d_lo,
                                       d_hi, drain ? 1 : 0, 256, x->stream);
              }),
              "route_wg");
    return MHMHIP_OK;
  }
  const int64_t end = drain ? d_lo + x->lay.dmax : d_hi;
  for (int64_t d = d_lo; d < end; ++d) {
    int64_t lo, hi;
    active_range(x, r, d, lo, hi);
    if (hi <= lo) continue;
    CHECK_HIP(x, timed(x, MHMHIP_K_ROUTE_DIAG, 0.0,
                       [&] { return launch_route_diag(r, d, lo, hi, x->stream); }),
              "route_diag");
  }
  return MHMHIP_OK;
}

// Tiled routing launches [L_lo, L_hi) of a call.
// Launch L handles window k = L - (tau_max - tau) of each tile: only the bins
// holding tiles of depth tau_max - L .. tau_max - L + nwin - 1 are launched.
int route_tiles(mhmhip_ctx *x, const RouteArgs &r, const TileArgs &g, int32_t L_lo, int32_t L_hi,
                hipStream_t s = nullptr) {
  if (!s) s = x->stream;
  const int64_t nb = (int64_t)x->lay.bins.size();
  for (int32_t L = L_lo; L < L_hi; ++L) {
    int64_t lo, hi;
    tile_bin_range(x->lay, g.tau_max - L, g.tau_max - L + g.nwin - 1, lo, hi);
    if (hi <= lo) continue;
    CHECK_HIP(x, timed(x, MHMHIP_K_ROUTE_TILE, 0.0, [&] {
                return launch_route_tile(r, g, lo, hi, nb, L, s);
              }, s),
              "route_tile");
  }
  return MHMHIP_OK;
}

// Under capture: an edge from what stream `from` has enqueued so far to
// what `to` enqueues next (the event is kept until the graph is dropped).
int capture_edge(mhmhip_ctx *x, hipStream_t from, hipStream_t to, hipEvent_t *out = nullptr) {
  hipEvent_t e = nullptr;
  CHECK_HIP(x, hipEventCreateWithFlags(&e, hipEventDisableTiming), "capture edge");
  x->graph_sync.push_back(e);
  CHECK_HIP(x, hipEventRecord(e, from), "capture edge");
  if (to) CHECK_HIP(x, hipStreamWaitEvent(to, e, 0), "capture edge");
  if (out) *out = e;
  return MHMHIP_OK;
}

// Gauges that share a node with an earlier gauge: copies of that gauge's
// series in the downloaded rows (steps x sets).
void copy_dup_gauges(const mhmhip_ctx *x, double *out, int64_t nsteps) {
  if (x->gauge_dup.empty() || x->n_gauges <= 0) return;
  const int64_t rows = nsteps * x->n_sets;
  for (int64_t i = 0; i < rows; ++i)
    for (const auto &d : x->gauge_dup) out[i * x->n_gauges + d.first] = out[i * x->n_gauges + d.second];
}

// node-steps routed, credited to the routing kernel that did them
void credit_route(mhmhip_ctx *x, int kind, int64_t nsteps) {
  if (x->prof) x->prof_units[kind] += (double)x->n_nodes * x->n_sets * (double)nsteps;
}

// Moves qTR(:,IT1) into parity slot 1, where a call's step -1 reads it.
int normalize_qtr(mhmhip_ctx *x) {
  if (x->qtr_slot == 1) return MHMHIP_OK;
  CHECK_HIP(x, hipMemcpyAsync(x->d_pub + x->ld, x->d_pub, sizeof(double) * x->ld,
                              hipMemcpyDeviceToDevice, x->stream),
            "qTR slot");
  x->qtr_slot = 1;
  return MHMHIP_OK;
}

void drop_graph(mhmhip_ctx *x) {
  if (x->step_exec) (void)hipGraphExecDestroy(x->step_exec);
  if (x->step_graph) (void)hipGraphDestroy(x->step_graph);
  x->step_exec = nullptr;
  x->step_graph = nullptr;
  x->skey = mhmhip_ctx::StepKey();
  if (x->graph_exec) (void)hipGraphExecDestroy(x->graph_exec);
  if (x->graph) (void)hipGraphDestroy(x->graph);
  for (auto &p : x->graph_events) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  x->graph_events.clear();
  for (hipEvent_t e : x->graph_sync) (void)hipEventDestroy(e);
  x->graph_sync.clear();
  for (int k = 0; k < MHMHIP_K_COUNT; ++k) x->graph_all[k] = 0;
  x->graph_exec = nullptr;
  x->graph = nullptr;
  x->gkey = mhmhip_ctx::GraphKey();
}

}  // namespace

extern "C" int mhmhip_mhm_step(mhmhip_ctx *x, const mhmhip_step *st) {
  if (!x) return MHMHIP_EINVAL;
  if (!st) return x->fail(MHMHIP_EINVAL, "mhm_step: null step");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "mhm_step");
  StepDev d;
  CHECK_RC(make_step(x, *st, d));
  CHECK_RC(forcing_wait(x, &d, 1));
  const bool route = routing_on(x);
  if (route) {
    if (x->t_route != x->t_cell)
      return x->fail(MHMHIP_ESTATE, "mhm_step: previous step not routed (call mrm_step)");
    CHECK_RC(ensure_ring(x, 1 + x->tile_h));
  }
  // the step is step 0 of a one-step ring, routed by the tiles (mrm_step); on
  // a coarser routing grid mrm_step aggregates the step's L1_total_runoff
  x->s_dirty = x->x_dirty = true;
  CellArgs a = cell_args(x, route && !x->coarse);
  a.ring_len = 1;
  a.lag = route && !x->coarse ? x->d_lag : nullptr;
  CHECK_HIP(x, timed(x, MHMHIP_K_CELL_STEP, (double)x->n_cells * x->n_sets, [&] {
              return launch_cell_step(x->nh, a, d, 0, kStepBlock, x->n_sets, x->stream);
            }),
            "cell_step");
  x->t_cell++;
  if (x->out_on) x->out_count += 1;
  if (!route) x->t_route = x->t_cell;
  return MHMHIP_OK;
}

// ---------------------------------------------------------------------------
// network and routing

extern "C" int mhmhip_set_tile_height(mhmhip_ctx *x, int32_t height) {
  if (!x) return MHMHIP_EINVAL;
  // a tile's proxies for its upstream tiles sit one level below its root
  if (height < 0 || height == 1 || height > kTileHeightMax)
    return x->fail(MHMHIP_EINVAL, "set_tile_height: height must be 0 or 2.." +
                                      std::to_string(kTileHeightMax));
  // the ring's rows past its length and the bound tiles follow the height
  if (x->net_bound)
    return x->fail(MHMHIP_ESTATE, "set_tile_height: call it before bind_network");
  x->tile_h = height ? height : kTileHeight;
  return MHMHIP_OK;
}

extern "C" int mhmhip_bind_network(mhmhip_ctx *x, const mhmhip_network *net) {
  if (!x) return MHMHIP_EINVAL;
  if (!net || !net->netPerm || !net->fromN || !net->toN || !net->length || !net->slope ||
      !net->fFPimp || !net->L1_area || !net->L1_L11_Id || !net->n_outlets)
    return x->fail(MHMHIP_EINVAL, "bind_network: missing arrays");
  // L1_L11_Id maps every L1 cell to an L11 node of its domain; the routing
  // grid is the L1 grid (identity) or coarser (map_flag, the runoff of a
  // node's cells is aggregated: mo_mrm_pre_routing.f90:110-126).  A routing
  // grid finer than L1 (map_flag = .false., the gathering branch) is refused.
  if (net->n_nodes < 1 || net->n_nodes > x->n_cells)
    return x->fail(MHMHIP_EINVAL, "bind_network: n_nodes must be 1..n_cells (a routing grid finer "
                                  "than the hydrology grid is not supported)");
  if (net->n_nodes != x->n_cells && !net->domain_node_start && x->n_domains > 1)
    return x->fail(MHMHIP_EINVAL, "bind_network: domain_node_start required when n_nodes != n_cells");
  std::vector<int64_t> ds(x->n_domains + 1);
  for (int d = 0; d <= x->n_domains; ++d)
    ds[d] = net->domain_node_start ? net->domain_node_start[d]
                                   : (d == 0 ? 0 : (d == x->n_domains ? net->n_nodes
                                                                      : x->dom_cell_start[d]));
  if (ds[0] != 0 || ds[x->n_domains] != net->n_nodes)
    return x->fail(MHMHIP_EINVAL, "bind_network: domain node ranges must cover the n_nodes nodes");
  bool identity = net->n_nodes == x->n_cells && ds == x->dom_cell_start;
  for (int d = 0; d < x->n_domains; ++d) {
    if (ds[d + 1] <= ds[d])
      return x->fail(MHMHIP_EINVAL, "bind_network: every domain needs a node");
    const int64_t nd = ds[d + 1] - ds[d];
    for (int64_t k = x->dom_cell_start[d]; k < x->dom_cell_start[d + 1]; ++k) {
      const int32_t id = net->L1_L11_Id[k];
      if (id < 1 || id > nd)
        return x->fail(MHMHIP_EINVAL, "bind_network: L1_L11_Id out of its domain's node range");
      identity = identity && id == k - x->dom_cell_start[d] + 1;
    }
  }
  const bool coarse = !identity;
  if (net->n_gauges < 0 || net->n_inflow < 0 ||
      (net->n_gauges > 0 && (!net->gauge_node || !net->gauge_domain)) ||
      (net->n_inflow > 0 && (!net->inflow_node || !net->inflow_domain || !net->inflow_headwater)))
    return x->fail(MHMHIP_EINVAL, "bind_network: gauge lists missing");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "bind_network");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "bind_network");

  LayoutIn in{net->n_nodes, x->n_domains, ds.data(), net->n_outlets, net->netPerm, net->fromN,
              net->toN, x->tile_h};
  LayoutOut lay;
  std::string e;
  int rc = build_layout(in, lay, e);
  if (rc != MHMHIP_OK) return x->fail(rc, e);

  // gauges / inflow gauges -> device node indices and flags
  std::vector<int32_t> gnode(std::max(1, net->n_gauges));
  for (int g = 0; g < net->n_gauges; ++g) {
    const int d = net->gauge_domain[g];
    if (d < 0 || d >= x->n_domains || net->gauge_node[g] < 1 ||
        net->gauge_node[g] > ds[d + 1] - ds[d])
      return x->fail(MHMHIP_EINVAL, "bind_network: gauge node out of range");
    const int64_t p = lay.pos[ds[d] + net->gauge_node[g] - 1];
    gnode[g] = (int32_t)p;
    lay.packed[p] |= NF_GAUGE;
  }
  std::vector<int32_t> inode(std::max(1, net->n_inflow)), ihw(std::max(1, net->n_inflow));
  for (int g = 0; g < net->n_inflow; ++g) {
    const int d = net->inflow_domain[g];
    if (d < 0 || d >= x->n_domains || net->inflow_node[g] < 1 ||
        net->inflow_node[g] > ds[d + 1] - ds[d])
      return x->fail(MHMHIP_EINVAL, "bind_network: inflow node out of range");
    const int64_t p = lay.pos[ds[d] + net->inflow_node[g] - 1];
    inode[g] = (int32_t)p;
    ihw[g] = net->inflow_headwater[g] ? 1 : 0;
    lay.packed[p] |= NF_INFLOW;
    if (!ihw[g]) {
      // links draining into this node have their routed outflow cut (mo_mrm_routing.f90:452-457)
      const int64_t c0 = lay.child[p];
      const int nc = nm_nchild(lay.packed[p]);
      for (int c = 0; c < nc; ++c) lay.packed[lay.child_list[c0 + c]] |= NF_CUT;
    }
  }

  // device cells: in routing order (cell = node) on an L1 = L11 grid, else
  // grouped by their node's device index, reference order inside a node
  std::vector<int64_t> cpos(x->n_cells);
  std::vector<int32_t> agg_start;
  if (!coarse) {
    cpos = lay.pos;
  } else {
    std::vector<int64_t> node_dev(x->n_cells);
    for (int d = 0; d < x->n_domains; ++d)
      for (int64_t k = x->dom_cell_start[d]; k < x->dom_cell_start[d + 1]; ++k)
        node_dev[k] = lay.pos[ds[d] + net->L1_L11_Id[k] - 1];
    std::vector<int64_t> order(x->n_cells);
    for (int64_t k = 0; k < x->n_cells; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return node_dev[a] < node_dev[b]; });
    agg_start.assign(net->n_nodes + 1, 0);
    for (int64_t r = 0; r < x->n_cells; ++r) {
      cpos[order[r]] = r;
      agg_start[node_dev[order[r]] + 1] += 1;
    }
    for (int64_t p = 0; p < net->n_nodes; ++p) agg_start[p + 1] += agg_start[p];
  }
  // re-lay any cell data already on the device into the new cell order
  if (x->cstream) CHECK_HIP(x, hipStreamSynchronize(x->cstream), "bind_network");
  std::vector<int64_t> old_pos = x->pos;
  bool moved = false;
  for (int64_t c = 0; c < x->n_cells && !moved; ++c) moved = old_pos[c] != cpos[c];
  if (moved) {
    std::vector<double> a(x->sstride), b(x->sstride);
    auto relayout = [&](double *base, int64_t rows, int64_t stride, int64_t nsl) -> int {
      for (int64_t r = 0; r < rows; ++r)
        for (int64_t s = 0; s < nsl; ++s) {
          double *row = base + r * stride + s * x->sstride;
          CHECK_HIP(x, sync_copy(x, a.data(), row, sizeof(double) * x->sstride, hipMemcpyDeviceToHost),
                    "relayout");
          std::fill(b.begin(), b.end(), 0.0);
          for (int64_t c = 0; c < x->n_cells; ++c) b[cpos[c]] = a[old_pos[c]];
          CHECK_HIP(x, sync_copy(x, row, b.data(), sizeof(double) * x->sstride, hipMemcpyHostToDevice),
                    "relayout");
        }
      return MHMHIP_OK;
    };
    // buffers never written since creation hold zeros in any order
    if (x->p_dirty) CHECK_RC(relayout(x->P, x->n_prows, x->ld, x->n_sets));
    if (x->s_dirty) CHECK_RC(relayout(x->S, x->n_srows, x->ld, x->n_sets));
    if (x->x_dirty) CHECK_RC(relayout(x->X, x->n_xrows, x->ld, x->n_sets));
    if (x->f_dirty)
      CHECK_RC(relayout(x->F, (int64_t)x->n_frows * x->cfg.n_forcing_days, x->sstride, 1));
    if (x->W) CHECK_RC(relayout(x->W, (int64_t)3 * 12 * 24, x->sstride, 1));
    if (x->d_out) CHECK_RC(relayout(x->d_out, x->out_rows_cap, x->ld, x->n_sets));  // output sums
  }
  x->pos = cpos;
  x->d_pos_stale = true;
  if (moved) x->mpr_stale = true;
  const int64_t old_nodes = x->n_nodes;

  // per-link host copies for reg_rout, packed by domain
  x->h_nout.assign(net->n_outlets, net->n_outlets + x->n_domains);
  x->dom_link_start.assign(x->n_domains + 1, 0);
  for (int d = 0; d < x->n_domains; ++d)
    x->dom_link_start[d + 1] = x->dom_link_start[d] + (ds[d + 1] - ds[d]) - x->h_nout[d];
  const int64_t nLtot = x->dom_link_start[x->n_domains];
  x->h_length.assign(net->length, net->length + nLtot);
  x->h_slope.assign(net->slope, net->slope + nLtot);
  x->h_fFPimp.assign(net->fFPimp, net->fFPimp + nLtot);

  // free old routing buffers and allocate new ones; until the new network is
  // complete the context has none (a failure below leaves routing unbound)
  x->net_bound = false;
  drop_graph(x);
  // the gauge series and inflow discharge buffers are sized by the old
  // network's gauge and inflow counts: a rebound network starts without them
  // (set_inflow must be called again)
  x->dfree(x->d_gauge_buf, x->gauge_buf_rows * (int64_t)x->n_sets * std::max(1, x->n_gauges));
  x->gauge_buf_rows = 0;
  x->dfree(x->d_inflow_q, x->inflow_q_rows * x->n_inflow);
  x->inflow_q_rows = 0;
  x->dfree(x->d_meta, old_nodes);
  x->dfree(x->d_area, x->n_cells);
  x->dfree(x->d_agg_start, old_nodes + 1);
  x->dfree(x->d_cring, x->cring_rows * x->ld);
  x->cring_rows = 0;
  x->dfree(x->d_C1, x->ld);
  x->dfree(x->d_C2, x->ld);
  x->dfree(x->d_qTIN, x->ld);
  x->dfree(x->d_pub, 2 * x->ld);
  x->dfree(x->d_qOUT, x->ld);
  x->dfree(x->d_dom_start, x->n_domains + 1);
  x->dfree(x->d_dom_dmax, x->n_domains);
  x->dfree(x->d_child_list, x->n_child_dev);
  x->dfree(x->d_slots, x->n_slots_dev);
  x->dfree(x->d_slot_out, x->n_slots_dev);
  x->dfree(x->d_bins, x->n_bins_dev);
  x->dfree(x->d_bnd, x->n_bnd_dev);
  x->n_child_dev = x->n_slots_dev = x->n_bins_dev = x->n_bnd_dev = 0;
  x->dfree(x->d_gauge_node, std::max(1, x->n_gauges));
  x->dfree(x->d_node_gauge, old_nodes);
  x->dfree(x->d_lag, old_nodes);
  x->dfree(x->d_qacc, x->ld);  // a new network restarts the output aggregation
  x->dfree(x->d_qsub, x->ld);
  x->dfree(x->d_runacc, x->ld);
  x->dfree(x->d_inflow_avg, std::max(1, x->n_inflow));
  x->inflow_acc.clear();
  x->qacc_on = false;
  x->qacc_count = 0;
  x->dfree(x->d_inflow_node, std::max(1, x->n_inflow));
  x->dfree(x->d_inflow_hw, std::max(1, x->n_inflow));
  const int64_t nn = net->n_nodes;
  x->n_nodes = nn;
  x->npos = lay.pos;
  x->dom_node_start = ds;
  x->coarse = coarse;
  CHECK_RC(x->dalloc(&x->d_meta, nn, "alloc meta"));
  CHECK_RC(x->dalloc(&x->d_area, x->n_cells, "alloc area"));
  if (coarse) {
    CHECK_RC(x->dalloc(&x->d_agg_start, nn + 1, "alloc node cell ranges"));
    CHECK_HIP(x, sync_copy(x, x->d_agg_start, agg_start.data(), sizeof(int32_t) * (nn + 1),
                           hipMemcpyHostToDevice),
              "upload node cell ranges");
  }
  CHECK_RC(x->dalloc(&x->d_C1, x->ld, "alloc C1"));
  CHECK_RC(x->dalloc(&x->d_C2, x->ld, "alloc C2"));
  CHECK_RC(x->dalloc(&x->d_qTIN, x->ld, "alloc qTIN"));
  CHECK_RC(x->dalloc(&x->d_pub, 2 * x->ld, "alloc qTR"));
  CHECK_RC(x->dalloc(&x->d_qOUT, x->ld, "alloc qOUT"));
  CHECK_RC(x->dalloc(&x->d_dom_start, x->n_domains + 1, "alloc domains"));
  CHECK_RC(x->dalloc(&x->d_dom_dmax, x->n_domains, "alloc domains"));
  // the slot flags carry the node flags (gauges, inflow, cut links set above)
  for (SlotMeta &m : lay.slots)
    if (m.a & SF_REAL) m.a |= lay.packed[m.pos] & NF_MASK;
  // output rows of the tiled kernel, per bin: its tile roots, then its gauge
  // nodes (the kernel stages their series in LDS)
  std::vector<int32_t> slot_out(lay.slots.size(), -1);  // both fields 0xFFFF
  int32_t out_rows = 0;
  for (size_t b = 0; b < lay.bins.size(); ++b) {
    int32_t n = 0;
    for (int i = 0; i < kTileSlots; ++i)
      if (lay.slots[b * kTileSlots + i].a & SF_TROOT)
        slot_out[b * kTileSlots + i] = (int32_t)(0xFFFF0000u | (uint32_t)n++);
    for (int i = 0; i < kTileSlots; ++i) {
      const int32_t a = lay.slots[b * kTileSlots + i].a;
      if ((a & SF_REAL) && (a & NF_GAUGE)) {
        int32_t &v = slot_out[b * kTileSlots + i];
        v = (int32_t)(((uint32_t)v & 0xFFFFu) | ((uint32_t)n++ << 16));
      }
    }
    lay.bins[b].out_rows = n;
    out_rows = std::max(out_rows, n);
  }
  const int64_t n_child = std::max<int64_t>(1, (int64_t)lay.child_list.size());
  const int64_t n_bnd = 2 * (int64_t)x->n_sets * std::max<int64_t>(1, lay.n_bnd) * kTileWindowMax;
  CHECK_RC(x->dalloc(&x->d_child_list, n_child, "alloc child list"));
  x->n_child_dev = n_child;
  CHECK_RC(x->dalloc(&x->d_slots, (int64_t)lay.slots.size(), "alloc tiles"));
  CHECK_RC(x->dalloc(&x->d_slot_out, (int64_t)lay.slots.size(), "alloc tiles"));
  x->n_slots_dev = (int64_t)lay.slots.size();
  x->tile_out_rows = out_rows;
  CHECK_RC(x->dalloc(&x->d_bins, (int64_t)lay.bins.size(), "alloc tiles"));
  x->n_bins_dev = (int64_t)lay.bins.size();
  CHECK_RC(x->dalloc(&x->d_bnd, n_bnd, "alloc tile boundaries"));
  x->n_bnd_dev = n_bnd;
  if (!lay.child_list.empty())
    CHECK_HIP(x, sync_copy(x, x->d_child_list, lay.child_list.data(),
                           sizeof(int32_t) * lay.child_list.size(), hipMemcpyHostToDevice),
              "upload child list");
  CHECK_HIP(x, sync_copy(x, x->d_slot_out, slot_out.data(), sizeof(int32_t) * slot_out.size(),
                         hipMemcpyHostToDevice),
            "upload tiles");
  CHECK_HIP(x, sync_copy(x, x->d_slots, lay.slots.data(), sizeof(SlotMeta) * lay.slots.size(),
                         hipMemcpyHostToDevice),
            "upload tiles");
  CHECK_HIP(x, sync_copy(x, x->d_bins, lay.bins.data(), sizeof(BinInfo) * lay.bins.size(),
                         hipMemcpyHostToDevice),
            "upload tiles");
  x->n_gauges = net->n_gauges;
  x->n_inflow = net->n_inflow;
  CHECK_RC(x->dalloc(&x->d_gauge_node, std::max(1, x->n_gauges), "alloc gauges"));
  CHECK_RC(x->dalloc(&x->d_node_gauge, nn, "alloc gauges"));
  CHECK_RC(x->dalloc(&x->d_lag, nn, "alloc ring lags"));
  CHECK_RC(x->dalloc(&x->d_inflow_node, std::max(1, x->n_inflow), "alloc inflow"));
  CHECK_RC(x->dalloc(&x->d_inflow_hw, std::max(1, x->n_inflow), "alloc inflow"));

  std::vector<NodeMeta> meta(nn);
  std::vector<double> area(x->n_cells);
  for (int64_t p = 0; p < nn; ++p) {
    meta[p].child = lay.child[p];
    meta[p].packed = lay.packed[p];
  }
  for (int64_t k = 0; k < x->n_cells; ++k) area[cpos[k]] = net->L1_area[k];  // per L1 cell
  std::vector<int64_t> dstart(x->n_domains + 1);
  for (int d = 0; d <= x->n_domains; ++d) dstart[d] = ds[d];  // BFS keeps domains contiguous
  CHECK_HIP(x, sync_copy(x, x->d_meta, meta.data(), sizeof(NodeMeta) * nn, hipMemcpyHostToDevice), "upload meta");
  CHECK_HIP(x, sync_copy(x, x->d_area, area.data(), sizeof(double) * x->n_cells, hipMemcpyHostToDevice), "upload area");
  CHECK_HIP(x, sync_copy(x, x->d_dom_start, dstart.data(), sizeof(int64_t) * (x->n_domains + 1), hipMemcpyHostToDevice), "upload domains");
  CHECK_HIP(x, sync_copy(x, x->d_dom_dmax, lay.dom_dmax.data(), sizeof(int32_t) * x->n_domains, hipMemcpyHostToDevice), "upload domains");
  CHECK_HIP(x, sync_copy(x, x->d_gauge_node, gnode.data(), sizeof(int32_t) * gnode.size(), hipMemcpyHostToDevice), "upload gauges");
  // the tiled kernel writes a node's first gauge; later gauges at the same
  // node are copies of it
  std::vector<int32_t> node_gauge(nn, -1);
  x->gauge_dup.clear();
  for (int g = 0; g < net->n_gauges; ++g) {
    int32_t &first = node_gauge[gnode[g]];
    if (first < 0) first = g;
    else x->gauge_dup.emplace_back(g, first);
  }
  CHECK_HIP(x, sync_copy(x, x->d_node_gauge, node_gauge.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice), "upload gauges");
  // ring lag of each node: the lag of its real slot
  std::vector<uint8_t> lag(nn, 0);
  for (const SlotMeta &m : lay.slots)
    if (m.a & SF_REAL) lag[m.pos] = (uint8_t)(m.a & 0xFFF);
  CHECK_HIP(x, sync_copy(x, x->d_lag, lag.data(), nn, hipMemcpyHostToDevice), "upload ring lags");
  CHECK_HIP(x, sync_copy(x, x->d_inflow_node, inode.data(), sizeof(int32_t) * inode.size(), hipMemcpyHostToDevice), "upload inflow");
  CHECK_HIP(x, sync_copy(x, x->d_inflow_hw, ihw.data(), sizeof(int32_t) * ihw.size(), hipMemcpyHostToDevice), "upload inflow");
  x->h_inflow_node_dev = inode;
  x->h_inflow_hw = ihw;
  x->lev_start.clear();  // the tile-major layout does not keep depths contiguous
  x->lay = std::move(lay);
  x->lone_nodes.clear();
  for (int d = 0; d < x->n_domains; ++d)
    if (ds[d + 1] - ds[d] == 1) x->lone_nodes.push_back(ds[d]);
  x->lone_qtin.assign(x->lone_nodes.size() * (size_t)x->n_sets, 0.0);
  x->ts_rout.assign((size_t)x->n_sets * x->n_domains, 0.0);
  x->gam5.assign((size_t)x->n_sets * x->n_domains * 5, 0.0);
  x->gam_set.assign((size_t)x->n_sets * x->n_domains, 0);
  x->coef_restart.assign((size_t)x->n_sets * x->n_domains, 0);
  x->qsub_loops = 0;
  x->net_bound = true;
  x->qtr_slot = 1;
  drop_graph(x);
  x->t_route = x->t_cell;
  x->dfree(x->d_ring, x->ring_len * x->ld);
  x->ring_len = 0;
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "bind_network");
  return MHMHIP_OK;
}

namespace {

// C1/C2 of the links of the (set, domain) pairs selected by `pick`, uploaded
// into the set's rows (other domains keep theirs).  coef(d, lo, hi, c1, c2)
// fills the links [lo, hi) of domain d (indices into the packed link arrays)
// and returns false to leave the domain untouched.
template <typename Fn>
int upload_coefs(mhmhip_ctx *x, int32_t set, Fn coef) {
  const int64_t nLtot = x->dom_link_start[x->n_domains];
  const int64_t off = (int64_t)set * x->sstride;
  std::vector<double> n1(x->sstride, 0.0), n2(x->sstride, 0.0);
  CHECK_HIP(x, sync_copy(x, n1.data(), x->d_C1 + off, sizeof(double) * x->sstride, hipMemcpyDeviceToHost), "download C1");
  CHECK_HIP(x, sync_copy(x, n2.data(), x->d_C2 + off, sizeof(double) * x->sstride, hipMemcpyDeviceToHost), "download C2");
  std::vector<double> c1(std::max<int64_t>(nLtot, 1), 0.0), c2(std::max<int64_t>(nLtot, 1), 0.0);
  std::vector<char> touched(std::max<int64_t>(nLtot, 1), 0);
  for (int d = 0; d < x->n_domains; ++d) {
    const int64_t l0 = x->dom_link_start[d], l1 = x->dom_link_start[d + 1];
    if (!coef(d, l0, l1, c1.data(), c2.data())) continue;
    for (int64_t i = l0; i < l1; ++i) touched[i] = 1;
  }
  for (int64_t q = 0; q < x->n_nodes; ++q) {
    const int64_t l = x->lay.link_of_node[q];
    if (l >= 0 && touched[l]) {
      n1[q] = c1[l];
      n2[q] = c2[l];
    }
  }
  CHECK_HIP(x, sync_copy(x, x->d_C1 + off, n1.data(), sizeof(double) * x->sstride, hipMemcpyHostToDevice), "upload C1");
  CHECK_HIP(x, sync_copy(x, x->d_C2 + off, n2.data(), sizeof(double) * x->sstride, hipMemcpyHostToDevice), "upload C2");
  return MHMHIP_OK;
}

// reg_rout (mo_mrm_mpr.f90:61-119) over the links [l0, l1) of one domain,
// TS = real(timestep) [h]
void reg_rout(const mhmhip_ctx *x, const double *p, int64_t l0, int64_t l1, double *c1,
              double *c2) {
  if (l1 <= l0) return;
  const double TS = (double)x->cfg.timestep;
  double ssMax = x->h_slope[l0];
  for (int64_t i = l0 + 1; i < l1; ++i)
    if (x->h_slope[i] > ssMax) ssMax = x->h_slope[i];
  for (int64_t i = l0; i < l1; ++i) {
    double K = p[0] + p[1] * (x->h_length[i] * 0.001) + p[2] * x->h_slope[i] + p[3] * x->h_fFPimp[i];
    double xi = p[4] * (1.0 + x->h_slope[i] / ssMax);
    if (xi > 0.5) xi = 0.5;
    if (xi < 0.005) xi = 0.005;
    if (K > 0.5 * TS / xi) K = 0.5 * TS / xi;
    if (K < 0.5 * TS / (1.0 - xi)) K = 0.5 * TS / (1.0 - xi);
    c1[i] = TS / (K * (1.0 - xi) + 0.5 * TS);
    c2[i] = 1.0 - c1[i] * K / TS;
  }
}

// A routed step with read_states = 0 evaluates reg_rout from the routing1
// parameters (mo_mrm_routing.f90:211-218): C1/C2 loaded from a restart are
// replaced by that.  The parameters must be known; other (set, domain)
// pairs already hold reg_rout's (or mrm_update_param's) values.
int refresh_reg_rout(mhmhip_ctx *x) {
  for (int s = 0; s < x->n_sets; ++s) {
    bool any = false;
    for (int d = 0; d < x->n_domains; ++d) {
      const size_t k = (size_t)s * x->n_domains + d;
      if (!x->coef_restart[k] || x->ts_rout[k] > 0.0) continue;
      if (x->dom_link_start[d + 1] <= x->dom_link_start[d]) continue;
      if (!x->gam_set[k])
        return x->fail(MHMHIP_EINVAL,
                       "read_states = 0 recomputes C1/C2 by reg_rout from the routing "
                       "parameters (mo_mrm_routing.f90:211-218): set them first, or route "
                       "with read_states = 1 to keep the restart's C1/C2");
      any = true;
    }
    if (!any) continue;
    CHECK_RC(upload_coefs(x, s, [&](int d, int64_t l0, int64_t l1, double *c1, double *c2) {
      const size_t k = (size_t)s * x->n_domains + d;
      if (!x->coef_restart[k] || x->ts_rout[k] > 0.0) return false;
      reg_rout(x, &x->gam5[k * 5], l0, l1, c1, c2);
      x->coef_restart[k] = 0;
      return true;
    }));
  }
  return MHMHIP_OK;
}

}  // namespace

extern "C" int mhmhip_set_routing_params(mhmhip_ctx *x, int32_t set, int32_t domain,
                                         const double *p) {
  if (!x) return MHMHIP_EINVAL;
  if (!p || set < 0 || set >= x->n_sets || domain >= x->n_domains)
    return x->fail(MHMHIP_EINVAL, "set_routing_params: bad set or domain");
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "set_routing_params: bind a network first");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_routing_params");
  // The reference re-evaluates reg_rout every step with unchanged inputs
  // (mo_mrm_routing.f90:211-218); evaluating it once gives the same C1, C2.
  CHECK_RC(upload_coefs(x, set, [&](int d, int64_t l0, int64_t l1, double *c1, double *c2) {
    if (domain >= 0 && d != domain) return false;
    const size_t k = (size_t)set * x->n_domains + d;
    x->ts_rout[k] = 0.0;  // processCase 1
    for (int i = 0; i < 5; ++i) x->gam5[k * 5 + i] = p[i];
    x->gam_set[k] = 1;
    x->coef_restart[k] = 0;
    reg_rout(x, p, l0, l1, c1, c2);
    return true;
  }));
  return MHMHIP_OK;
}

namespace {
// the routing steps a domain may take [s]: given_TS (mo_mrm_constants.F90:42-46)
constexpr double kGivenTS[19] = {60.0,    120.0,   180.0,   240.0,   300.0,   360.0,  600.0,
                                 720.0,   900.0,   1200.0,  1800.0,  3600.0,  7200.0, 10800.0,
                                 14400.0, 21600.0, 28800.0, 43200.0, 86400.0};
// FORCES mo_utils locate over given_TS (the first minimum of |x - y|; c if
// LE(x(c), y), else c - 1; 1-based) with LE = lesserequal (a <= b within
// epsilon): restated from the copy in pre-proc/mtCLIM/mo_utils.f90:924-936,
// 1329-1346
bool le_eps(double a, double b) {
  return !(((DBL_EPSILON * std::fabs(b) - std::fabs(a - b)) < 0.0) && (a > b));
}
int locate_ts(double y) {
  int c = 0;
  for (int i = 1; i < 19; ++i)
    if (std::fabs(kGivenTS[i] - y) < std::fabs(kGivenTS[c] - y)) c = i;
  return le_eps(kGivenTS[c], y) ? c + 1 : c;
}

// Routing sub-steps per mrm_step: every (set, domain) with links must agree.
// The routing cadence of the context (every (set, domain) with links must
// agree): tsRoutFactor = TSrout / (timestep HourSecs)
// (mo_mhm_interface_run.f90:482); below 1 each step routes
// rout_loop = max(1, nint(1 / tsRoutFactor)) sub-steps
// (mo_mrm_routing.f90:224); from 1 up, runoff is accumulated and routed every
// nint(tsRoutFactor) steps (mo_mhm_interface_run.f90:494-512).  *loops: the
// sub-steps (1 when accumulating); *factor: tsRoutFactor (1 for
// processCase 1).
int rout_cadence(mhmhip_ctx *x, int32_t *loops, double *factor) {
  int32_t common = 0;
  double cf = -1.0;
  for (int s = 0; s < x->n_sets; ++s)
    for (int d = 0; d < x->n_domains; ++d) {
      if (x->dom_link_start[d + 1] <= x->dom_link_start[d]) continue;
      const double ts = x->ts_rout[(size_t)s * x->n_domains + d];
      int32_t n = 1;
      double f = 1.0;
      if (ts > 0.0) {
        f = ts / ((double)x->cfg.timestep * kHourSecs);
        if (f < 1.0) n = std::max<int32_t>(1, (int32_t)std::floor(1.0 / f + 0.5));
      }
      if ((common != 0 && n != common) || (cf >= 0.0 && f != cf))
        return x->fail(MHMHIP_EINVAL, "routing steps differ between parameter sets or "
                                      "domains of one context");
      common = n;
      cf = f;
    }
  *loops = common > 0 ? common : 1;
  *factor = cf >= 0.0 ? cf : 1.0;
  return MHMHIP_OK;
}

int rout_loops(mhmhip_ctx *x, int32_t *loops) {
  double f = 1.0;
  CHECK_RC(rout_cadence(x, loops, &f));
  if (f > 1.0)
    return x->fail(MHMHIP_EINVAL, "routing step longer than the model step (runoff accumulated "
                                  "over several steps): route step by step with mhmhip_mrm_step");
  return MHMHIP_OK;
}
}  // namespace

namespace {

// mrm_update_param (mo_mrm_mpr.f90:241-329) for the links of the selected
// domains of one set: K = length / celerity [s] (302-311; cel(i) the
// celerity of packed link i), the routing step TSrout = given_TS(locate(
// given_TS, min K)) (312-317), C1 = TSrout / (K (1 - xi) + TSrout / 2),
// C2 = 1 - C1 K / TSrout with xi = |rout_space_weight| = 0 (318-323).
template <typename Cel>
int update_param(mhmhip_ctx *x, int32_t set, int32_t domain, Cel cel) {
  const double xi = std::fabs(0.0);  // rout_space_weight (mo_mrm_constants.F90:33)
  CHECK_RC(upload_coefs(x, set, [&](int d, int64_t l0, int64_t l1, double *c1, double *c2) {
    if ((domain >= 0 && d != domain) || l1 <= l0) return false;
    double kmin = 0.0;
    for (int64_t i = l0; i < l1; ++i) {
      const double K = x->h_length[i] / cel(i);
      if (i == l0 || K < kmin) kmin = K;
    }
    const int ind = std::max(1, locate_ts(kmin));
    const double ts = kGivenTS[ind - 1];
    for (int64_t i = l0; i < l1; ++i) {
      const double K = x->h_length[i] / cel(i);
      c1[i] = ts / (K * (1.0 - xi) + 0.5 * ts);
      c2[i] = 1.0 - c1[i] * K / ts;
    }
    const size_t k = (size_t)set * x->n_domains + d;
    x->ts_rout[k] = ts;
    x->coef_restart[k] = 0;
    return true;
  }));
  drop_graph(x);
  return MHMHIP_OK;
}

}  // namespace

extern "C" int mhmhip_set_routing_celerity(mhmhip_ctx *x, int32_t set, int32_t domain,
                                           double celerity) {
  if (!x) return MHMHIP_EINVAL;
  if (set < 0 || set >= x->n_sets || domain >= x->n_domains || !(celerity > 0.0))
    return x->fail(MHMHIP_EINVAL, "set_routing_celerity: bad set, domain or celerity");
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "set_routing_celerity: bind a network first");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_routing_celerity");
  // processCase(8) = 2: K = L11_length / param(1) (mo_mrm_mpr.f90:302-304)
  return update_param(x, set, domain, [celerity](int64_t) { return celerity; });
}

extern "C" int mhmhip_set_routing_celerity_links(mhmhip_ctx *x, int32_t set, int32_t domain,
                                                 const double *celerity) {
  if (!x) return MHMHIP_EINVAL;
  if (!celerity || set < 0 || set >= x->n_sets || domain >= x->n_domains)
    return x->fail(MHMHIP_EINVAL, "set_routing_celerity_links: bad set, domain or pointer");
  if (!x->net_bound)
    return x->fail(MHMHIP_ESTATE, "set_routing_celerity_links: bind a network first");
  // celerity is given for the domain's links (domain >= 0) or all packed links
  const int64_t base = domain >= 0 ? x->dom_link_start[domain] : 0;
  const int64_t l_lo = domain >= 0 ? x->dom_link_start[domain] : 0;
  const int64_t l_hi = domain >= 0 ? x->dom_link_start[domain + 1] : x->dom_link_start[x->n_domains];
  for (int64_t i = l_lo; i < l_hi; ++i)
    if (!(celerity[i - base] > 0.0) || !std::isfinite(celerity[i - base]))
      return x->fail(MHMHIP_EINVAL, "set_routing_celerity_links: celerity of link " +
                                        std::to_string(i - base + 1) + " is not positive");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_routing_celerity_links");
  // processCase(8) = 3: K = L11_length / L11_celerity (mo_mrm_mpr.f90:306-311)
  return update_param(x, set, domain, [celerity, base](int64_t i) { return celerity[i - base]; });
}

extern "C" int mhmhip_set_inflow(mhmhip_ctx *x, int32_t nsteps, const double *q) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "set_inflow: bind a network first");
  if (x->n_inflow == 0) return MHMHIP_OK;
  if (nsteps < 1 || !q) return x->fail(MHMHIP_EINVAL, "set_inflow: bad series");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_inflow");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_inflow");
  x->dfree(x->d_inflow_q, x->inflow_q_rows * x->n_inflow);
  x->inflow_q_rows = 0;
  CHECK_RC(x->dalloc(&x->d_inflow_q, (int64_t)nsteps * x->n_inflow, "alloc inflow series"));
  CHECK_HIP(x, sync_copy(x, x->d_inflow_q, q, sizeof(double) * nsteps * x->n_inflow, hipMemcpyHostToDevice), "upload inflow");
  x->inflow_q_rows = nsteps;
  x->h_inflow_row.assign(q, q + x->n_inflow);  // the row an mrm_step reads
  return MHMHIP_OK;
}

extern "C" int mhmhip_get_routing_state(mhmhip_ctx *x, int32_t set, double *qOUT, double *qTIN,
                                        double *qTR, double *Qmod, double *C1, double *C2) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "get_routing_state: no network bound");
  if (set < 0 || set >= x->n_sets) return x->fail(MHMHIP_EINVAL, "get_routing_state: bad set");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "get_routing_state");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "get_routing_state");
  const int64_t off = (int64_t)set * x->sstride;
  std::vector<double> buf(x->sstride);
  auto node_out = [&](const double *dev, double *dst) -> int {
    if (!dst) return MHMHIP_OK;
    CHECK_HIP(x, sync_copy(x, buf.data(), dev, sizeof(double) * x->sstride, hipMemcpyDeviceToHost), "download");
    for (int64_t k = 0; k < x->n_nodes; ++k) dst[k] = buf[x->npos[k]];
    return MHMHIP_OK;
  };
  auto link_out = [&](const double *dev, double *dst) -> int {
    if (!dst) return MHMHIP_OK;
    CHECK_HIP(x, sync_copy(x, buf.data(), dev, sizeof(double) * x->sstride, hipMemcpyDeviceToHost), "download");
    for (int64_t q = 0; q < x->n_nodes; ++q)
      if (x->lay.link_of_node[q] >= 0) dst[x->lay.link_of_node[q]] = buf[q];
    return MHMHIP_OK;
  };
  CHECK_RC(node_out(x->d_qOUT + off, qOUT));
  CHECK_RC(node_out(x->d_qTIN + off, qTIN));
  if (Qmod && x->qsub_loops > 1) {
    // the average of the last step's routing sub-steps (qAcc / rout_loop)
    CHECK_RC(node_out(x->d_qsub + off, Qmod));
    for (int64_t k = 0; k < x->n_nodes; ++k) Qmod[k] = Qmod[k] / (double)x->qsub_loops;
  }
  if (qTIN)
    for (size_t i = 0; i < x->lone_nodes.size(); ++i)
      qTIN[x->lone_nodes[i]] = x->lone_qtin[(size_t)set * x->lone_nodes.size() + i];
  if (x->qsub_loops <= 1) CHECK_RC(node_out(x->d_qTIN + off, Qmod));
  CHECK_RC(node_out(x->d_pub + x->qtr_slot * x->ld + off, qTR));
  CHECK_RC(link_out(x->d_C1 + off, C1));
  CHECK_RC(link_out(x->d_C2 + off, C2));
  return MHMHIP_OK;
}

extern "C" int mhmhip_qmod_accumulate(mhmhip_ctx *x, int32_t on) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "qmod_accumulate: no network bound");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "qmod_accumulate");
  x->qacc_on = on != 0;
  if (!x->qacc_on) return MHMHIP_OK;
  if (!x->d_qacc) {
    CHECK_RC(x->dalloc(&x->d_qacc, x->ld, "alloc Qmod sums"));  // zeroed
  } else {
    CHECK_HIP(x, hipMemsetAsync(x->d_qacc, 0, sizeof(double) * x->ld, x->stream), "qmod_accumulate");
  }
  x->qacc_count = 0;
  return MHMHIP_OK;
}

extern "C" int mhmhip_get_qmod_sum(mhmhip_ctx *x, int32_t set, double *sum, int64_t *count) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->d_qacc) return x->fail(MHMHIP_ESTATE, "get_qmod_sum: accumulation never started");
  if (set < 0 || set >= x->n_sets) return x->fail(MHMHIP_EINVAL, "get_qmod_sum: bad set");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "get_qmod_sum");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "get_qmod_sum");
  if (sum) {
    std::vector<double> buf(x->sstride);
    CHECK_HIP(x, sync_copy(x, buf.data(), x->d_qacc + (int64_t)set * x->sstride,
                           sizeof(double) * x->sstride, hipMemcpyDeviceToHost),
              "download");
    for (int64_t k = 0; k < x->n_nodes; ++k) sum[k] = buf[x->npos[k]];
  }
  if (count) *count = x->qacc_count;
  return MHMHIP_OK;
}

namespace {
// outputFlxState indices with one variable per soil horizon
bool out_per_horizon(int k) { return k == 3 || k == 4 || k == 17 || k == 19; }
}  // namespace

extern "C" int mhmhip_output_accumulate(mhmhip_ctx *x, const int32_t *flags) {
  if (!x) return MHMHIP_EINVAL;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "output_accumulate");
  if (!flags) {
    x->out_on = false;
    return MHMHIP_OK;
  }
  if (flags[17])
    return x->fail(MHMHIP_EINVAL, "output_accumulate: outputFlxState(18) (neutrons) is not on "
                                  "this path");
  // the reference's output order (mHM_updateDataset, mo_write_fluxes_states.f90:318-436)
  static const int order[] = {1, 2, 3, 4, 5, 6, 7, 8, 18, 9, 10, 11, 12, 13, 14, 15, 16, 17,
                              19, 20, 21};
  OutAcc o{};
  int n = 0;
  for (int k : order) {
    if (!flags[k - 1]) continue;
    const int nh = out_per_horizon(k) ? x->nh : 1;
    for (int h = 0; h < nh; ++h) {
      if (n >= kOutRowsMax) return x->fail(MHMHIP_EINVAL, "output_accumulate: too many rows");
      o.kind[n] = (int8_t)k;
      o.hor[n] = (int8_t)h;
      ++n;
    }
  }
  if (n == 0) {
    x->out_on = false;
    return MHMHIP_OK;
  }
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "output_accumulate");
  if (x->out_rows_cap < n) {
    x->dfree(x->d_out, x->out_rows_cap * x->ld);
    x->out_rows_cap = 0;
    CHECK_RC(x->dalloc(&x->d_out, (int64_t)n * x->ld, "alloc output sums"));  // zeroed
    x->out_rows_cap = n;
  } else {
    CHECK_HIP(x, hipMemsetAsync(x->d_out, 0, sizeof(double) * n * x->ld, x->stream),
              "output_accumulate");
  }
  o.sum = x->d_out;
  o.n = n;
  x->out_cfg = o;
  x->out_on = true;
  x->out_count = 0;
  return MHMHIP_OK;
}

extern "C" int mhmhip_get_output_sum(mhmhip_ctx *x, int32_t k, int32_t horizon, int32_t set,
                                     double *sum, int64_t *count) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->out_cfg.sum) return x->fail(MHMHIP_ESTATE, "get_output_sum: accumulation never started");
  if (set < 0 || set >= x->n_sets) return x->fail(MHMHIP_EINVAL, "get_output_sum: bad set");
  const int h = out_per_horizon(k) ? horizon - 1 : 0;
  int row = -1;
  for (int r = 0; r < x->out_cfg.n; ++r)
    if (x->out_cfg.kind[r] == k && x->out_cfg.hor[r] == h) row = r;
  if (row < 0 || (out_per_horizon(k) && (horizon < 1 || horizon > x->nh)))
    return x->fail(MHMHIP_EINVAL, "get_output_sum: output " + std::to_string(k) +
                                      " (horizon " + std::to_string(horizon) +
                                      ") is not accumulated");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "get_output_sum");
  if (sum) CHECK_RC(get_row(x, x->out_cfg.sum + (int64_t)row * x->ld + (int64_t)set * x->sstride,
                            sum));
  if (count) *count = x->out_count;
  return MHMHIP_OK;
}

extern "C" int mhmhip_set_routing_state(mhmhip_ctx *x, int32_t set, const double *qTIN,
                                        const double *qTR, const double *C1, const double *C2) {
  if (!x) return MHMHIP_EINVAL;
  if (!x->net_bound) return x->fail(MHMHIP_ESTATE, "set_routing_state: no network bound");
  if (set < 0 || set >= x->n_sets) return x->fail(MHMHIP_EINVAL, "set_routing_state: bad set");
  if ((C1 == nullptr) != (C2 == nullptr))
    return x->fail(MHMHIP_EINVAL, "set_routing_state: give C1 and C2 together");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_routing_state");
  const int64_t off = (int64_t)set * x->sstride;
  if (qTIN) CHECK_RC(put_node_row(x, x->d_qTIN + off, qTIN));
  if (qTIN)
    for (size_t i = 0; i < x->lone_nodes.size(); ++i)
      x->lone_qtin[(size_t)set * x->lone_nodes.size() + i] = qTIN[x->lone_nodes[i]];
  if (qTR) CHECK_RC(put_node_row(x, x->d_pub + x->qtr_slot * x->ld + off, qTR));
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_routing_state");
  if (C1) {
    // the restart's L11_C1/L11_C2 (mo_mrm_restart.f90:548-555), per link
    CHECK_RC(upload_coefs(x, set, [&](int d, int64_t l0, int64_t l1, double *c1, double *c2) {
      for (int64_t i = l0; i < l1; ++i) {
        c1[i] = C1[i];
        c2[i] = C2[i];
      }
      x->coef_restart[(size_t)set * x->n_domains + d] = 1;
      return true;
    }));
  }
  return MHMHIP_OK;
}

extern "C" int mhmhip_mrm_step(mhmhip_ctx *x, const mhmhip_step *st, double *gauge_out) {
  if (!x) return MHMHIP_EINVAL;
  if (!st) return x->fail(MHMHIP_EINVAL, "mrm_step: null step");
  if (!routing_on(x)) return x->fail(MHMHIP_ESTATE, "mrm_step: routing is off or no network bound");
  if (x->t_route != x->t_cell - 1)
    return x->fail(MHMHIP_ESTATE, "mrm_step: call mhm_step for this step first");
  if (x->n_inflow > 0 && x->inflow_q_rows < 1)
    return x->fail(MHMHIP_ESTATE, "mrm_step: inflow discharge not set");
  int32_t loops = 1;
  double factor = 1.0;
  CHECK_RC(rout_cadence(x, &loops, &factor));
  if (loops > 1 && x->qacc_on)
    return x->fail(MHMHIP_EINVAL, "mrm_step: Qmod output sums with routing sub-steps are not supported");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "mrm_step");
  if (!st->read_states) CHECK_RC(refresh_reg_rout(x));
  CHECK_RC(ensure_gauge_buf(x, 1));
  const double *inflow_row = x->d_inflow_q;
  if (factor > 1.0) {
    // routing steps longer than the model step (mo_mhm_interface_run.f90:494-512):
    // RunToRout and InflowDischarge add up; routing happens every
    // nint(tsRoutFactorIn) steps of tt and at the run's last step (tt ==
    // nTimeSteps, a shorter last period: tsRoutFactorIn = mod(tt, nint(...)))
    if (x->qacc_on)
      return x->fail(MHMHIP_EINVAL, "mrm_step: Qmod output sums with routing steps longer than "
                                    "the model step are not supported");
    double fin = factor;
    const int32_t tt = st->tt;
    const bool last = st->n_steps > 0 && tt == st->n_steps;
    const int32_t nn = (int32_t)std::floor(fin + 0.5);
    if (last && tt % nn != 0) fin = (double)(tt % nn);
    const bool route = tt % (int32_t)std::floor(fin + 0.5) == 0 || last;
    if (!x->d_runacc) CHECK_RC(x->dalloc(&x->d_runacc, x->ld, "alloc RunToRout"));  // zeroed
    if (x->n_inflow > 0) {
      if (x->inflow_acc.empty()) x->inflow_acc.assign(x->n_inflow, 0.0);
      for (int32_t g = 0; g < x->n_inflow; ++g)
        x->inflow_acc[g] = x->inflow_acc[g] + x->h_inflow_row[g];
    }
    const double tst = kHourSecs * (double)(x->cfg.timestep * (int32_t)std::floor(fin + 0.5));
    // RunToRout per L1 cell; on a coarser routing grid it is aggregated per
    // node on the routing step and starts again from 0
    CHECK_HIP(x, launch_runoff_sum(x->X + XR_TOTAL * x->ld, x->d_runacc, x->d_ring, x->d_area,
                                   x->d_lag, x->n_cells, x->sstride, x->ld, x->n_sets, tst,
                                   route && !x->coarse, x->stream),
              "runoff_sum");
    if (route && x->coarse) {
      CHECK_HIP(x, launch_runoff_agg(x->d_runacc, 1, x->d_agg_start, x->d_area, x->n_nodes,
                                     x->sstride, x->ld, x->n_sets, x->d_ring, x->d_lag, 1, 1, 0,
                                     1, tst, x->stream),
                "runoff_agg");
      CHECK_HIP(x, hipMemsetAsync(x->d_runacc, 0, sizeof(double) * x->ld, x->stream), "mrm_step");
    }
    if (!route) {
      x->t_route++;
      CHECK_HIP(x, hipStreamSynchronize(x->stream), "mrm_step");
      return MHMHIP_PENDING;
    }
    if (x->n_inflow > 0) {
      // InflowDischarge / tsRoutFactorIn, then the sums start again
      std::vector<double> avg(x->n_inflow);
      for (int32_t g = 0; g < x->n_inflow; ++g) avg[g] = x->inflow_acc[g] / fin;
      std::fill(x->inflow_acc.begin(), x->inflow_acc.end(), 0.0);
      if (!x->d_inflow_avg)
        CHECK_RC(x->dalloc(&x->d_inflow_avg, x->n_inflow, "alloc inflow average"));
      CHECK_HIP(x, sync_copy(x, x->d_inflow_avg, avg.data(), sizeof(double) * x->n_inflow,
                             hipMemcpyHostToDevice), "upload inflow average");
      inflow_row = x->d_inflow_avg;
    }
  }
  const int kind = MHMHIP_K_ROUTE_TILE;
  const TileArgs g = tile_args(x, 1, 1);
  const int64_t ng = (int64_t)x->n_sets * x->n_gauges;
  std::vector<double> gsum(loops > 1 ? ng : 0, 0.0), gone(loops > 1 ? ng : 0);
  if (loops > 1) {
    if (!x->d_qsub) CHECK_RC(x->dalloc(&x->d_qsub, x->ld, "alloc Qmod sub-step sums"));
    CHECK_HIP(x, hipMemsetAsync(x->d_qsub, 0, sizeof(double) * x->ld, x->stream), "mrm_step");
  }
  // this step is step 0 of a one-step ring: mhm_step wrote its qOUT to ring
  // row lag (the tiles' layout), or, on a coarser routing grid, its
  // L1_total_runoff, which L11_runoff_acc aggregates into that row here
  auto enqueue = [&](int32_t it) -> int {
    RouteArgs r = route_args(x, 0, 1, 0);
    r.ring_len = 1;
    r.inflow_q = inflow_row;
    if (loops > 1) r.qacc = x->d_qsub;
    if (x->coarse && factor <= 1.0 && it == 0)
      CHECK_HIP(x, launch_runoff_agg(x->X + XR_TOTAL * x->ld, 1, x->d_agg_start, x->d_area,
                                     x->n_nodes, x->sstride, x->ld, x->n_sets, x->d_ring,
                                     x->d_lag, 1, 1, 0, 1, kHourSecs * (double)x->cfg.timestep,
                                     x->stream),
                "runoff_agg");
    if (x->n_inflow > 0 && it == 0)
      CHECK_HIP(x, launch_inflow_patch(r, x->d_ring, x->d_lag, 0, 1, x->stream), "inflow_patch");
    return route_tiles(x, r, g, 0, g.nwin + g.tau_max);
  };
  // the ordinary step (one routing, the model step's own runoff) replays a
  // captured graph: one launch of 1 + tile-depth tile kernels instead of as
  // many stream launches
  const bool replay = loops == 1 && factor <= 1.0 && !x->prof;
  if (replay) {
    CHECK_RC(normalize_qtr(x));
    mhmhip_ctx::StepKey key;
    key.ring = x->d_ring;
    key.gauge = x->d_gauge_buf;
    key.inflow = inflow_row;
    key.qacc = x->qacc_on ? x->d_qacc : nullptr;
    if (!x->step_exec || !(key == x->skey)) {
      if (x->step_exec) (void)hipGraphExecDestroy(x->step_exec);
      if (x->step_graph) (void)hipGraphDestroy(x->step_graph);
      x->step_exec = nullptr;
      x->step_graph = nullptr;
      CHECK_HIP(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal), "capture");
      x->capturing = true;
      const int rc = enqueue(0);
      x->capturing = false;
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(x->stream, &gr);
      if (rc != MHMHIP_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      CHECK_HIP(x, e, "capture");
      x->step_graph = gr;
      CHECK_HIP(x, hipGraphInstantiate(&x->step_exec, gr, nullptr, nullptr, 0), "instantiate");
      x->skey = key;
    }
    CHECK_HIP(x, hipGraphLaunch(x->step_exec, x->stream), "graph launch");
    credit_route(x, kind, 1);
    x->qtr_slot = 0;
  }
  // loops routings of this step's qOUT (mRM_routing, mo_mrm_routing.f90:224-284):
  // the states advance each time, Qmod is their average
  for (int32_t it = 0; it < (replay ? 0 : loops); ++it) {
    CHECK_RC(normalize_qtr(x));
    CHECK_RC(enqueue(it));
    credit_route(x, kind, 1);
    x->qtr_slot = 0;
    if (loops > 1 && gauge_out && ng > 0) {
      CHECK_HIP(x, sync_copy(x, gone.data(), x->d_gauge_buf, sizeof(double) * ng,
                             hipMemcpyDeviceToHost), "download gauges");
      for (int64_t k = 0; k < ng; ++k) gsum[k] = gsum[k] + gone[k];  // qAcc = qAcc + Qmod
    }
  }
  x->qsub_loops = loops;
  if (x->qacc_on) x->qacc_count += 1;
  x->t_route++;
  if (gauge_out && ng > 0) {
    if (loops > 1) {
      for (int64_t k = 0; k < ng; ++k) gauge_out[k] = gsum[k] / (double)loops;
    } else {
      CHECK_HIP(x, hipMemcpyAsync(gauge_out, x->d_gauge_buf, sizeof(double) * ng,
                                  hipMemcpyDeviceToHost, x->stream),
                "download gauges");
    }
  }
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "mrm_step");
  if (gauge_out) copy_dup_gauges(x, gauge_out, 1);
  return prof_collect(x);
}

namespace {

// This GPU's scratch (resolved when the context was created on it).
int scratch_base(mhmhip_ctx *x, unsigned char **out) {
  CHECK_HIP(x, scratch_resolve(x->cfg.device), "run_noalloc: scratch address");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  *out = g_scratch[x->cfg.device];
  return MHMHIP_OK;
}

// The context's per-call buffers (step table, gauge series, runoff ring),
// replaced by ranges of the device scratch array for one no-allocation run
// and given back on every return path.
struct ScratchSwap {
  mhmhip_ctx *x;
  bool on = false;
  StepDev *steps = nullptr;
  int64_t steps_cap = 0;
  double *gauge = nullptr;
  int64_t gauge_rows = 0;
  double *ring = nullptr;
  int64_t ring_len = 0;
  explicit ScratchSwap(mhmhip_ctx *c) : x(c) {}
  void swap_in(StepDev *st, int64_t st_cap, double *g, int64_t g_rows, double *r, int64_t r_len) {
    steps = x->d_steps;
    steps_cap = x->steps_cap;
    gauge = x->d_gauge_buf;
    gauge_rows = x->gauge_buf_rows;
    ring = x->d_ring;
    ring_len = x->ring_len;
    x->d_steps = st;
    x->steps_cap = st_cap;
    x->d_gauge_buf = g;
    x->gauge_buf_rows = g_rows;
    x->d_ring = r;
    x->ring_len = r_len;
    on = true;
  }
  ~ScratchSwap() {
    if (!on) return;
    x->d_steps = steps;
    x->steps_cap = steps_cap;
    x->d_gauge_buf = gauge;
    x->gauge_buf_rows = gauge_rows;
    x->d_ring = ring;
    x->ring_len = ring_len;
  }
};

int run_impl(mhmhip_ctx *x, const mhmhip_step *steps, int32_t nsteps,
             const mhmhip_run_opts *opts, double *gauge_out, bool noalloc) {
  if (!x) return MHMHIP_EINVAL;
  if (!steps || nsteps < 1) return x->fail(MHMHIP_EINVAL, "run: no steps");
  mhmhip_run_opts o{};
  if (opts) o = *opts;
  if (o.mode != MHMHIP_RUN_PLAIN && o.mode != MHMHIP_RUN_FUSED)
    return x->fail(MHMHIP_EINVAL, "run: unknown mode");
  if (o.block != 0 && o.block != 64 && o.block != 128 && o.block != 256)
    return x->fail(MHMHIP_EINVAL, "run: block must be 64, 128 or 256");
  if (o.route_mode < 0 || o.route_mode > 3) return x->fail(MHMHIP_EINVAL, "run: bad route_mode");
  if (o.steps_per_launch < 0) return x->fail(MHMHIP_EINVAL, "run: bad steps_per_launch");
  if (o.route_window < 0 || o.route_window > kTileWindowMax)
    return x->fail(MHMHIP_EINVAL, "run: route_window must be 0.." + std::to_string(kTileWindowMax));
  const bool route = routing_on(x);
  // rep: routing sub-steps per model step (adaptive routing step shorter than
  // the model step, mo_mrm_routing.f90:224-284): the routing runs rep steps on
  // each model step's qOUT, and a model step's Qmod and gauges are the
  // average of its sub-steps'
  int32_t rep = 1;
  if (route) {
    CHECK_RC(rout_loops(x, &rep));
    if (rep > 1 && x->qacc_on)
      return x->fail(MHMHIP_EINVAL, "run: Qmod output sums with routing sub-steps are not "
                                    "supported");
    x->qsub_loops = 0;
  }
  const int64_t nr = (int64_t)nsteps * rep;  // routing steps of the call
  if (route && x->t_route != x->t_cell)
    return x->fail(MHMHIP_ESTATE, "run: previous step not routed (call mrm_step)");
  // read_restart is a constant of a run in the reference: one value per call
  for (int32_t k = 1; k < nsteps; ++k)
    if ((steps[k].read_states != 0) != (steps[0].read_states != 0))
      return x->fail(MHMHIP_EINVAL, "run: read_states must be the same for every step of a call");
  if (route && x->n_inflow > 0 && x->inflow_q_rows < nsteps)
    return x->fail(MHMHIP_ESTATE, "run: inflow discharge series shorter than the run");
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "run");
  if (route && !steps[0].read_states) CHECK_RC(refresh_reg_rout(x));
  std::vector<StepDev> hs(nsteps);
  for (int32_t k = 0; k < nsteps; ++k) CHECK_RC(make_step(x, steps[k], hs[k]));
  // a step whose forcing column, day and day/night factors (and, with meteo
  // weights, month and hour) equal the step before's has the same
  // disaggregated meteo values (the fused kernel keeps them; its first step
  // of a launch computes them regardless)
  for (int32_t k = 1; k < nsteps; ++k) {
    const StepDev &p = hs[k - 1], &c = hs[k];
    hs[k].meteo_same = p.day == c.day && p.doy == c.doy && p.month0 == c.month0 &&
                       p.f_pet == c.f_pet && p.f_prec == c.f_prec && p.f_temp == c.f_temp &&
                       (!x->W || p.wrow == c.wrow);
  }
  CHECK_RC(forcing_wait(x, hs.data(), nsteps));
  x->s_dirty = x->x_dirty = true;
  const int S = o.mode == MHMHIP_RUN_PLAIN ? 1 : (o.steps_per_launch > 0 ? o.steps_per_launch : 24);
  // default block: 256 for the fused kernel (its per-lane LDS slots fill three
  // groups per CU), 128 for the one-step kernel (8.2e9 against 7.3e9
  // cell-steps/s at 256 on the 1024x1024 basin)
  const int block = o.block > 0 ? o.block : (o.mode == MHMHIP_RUN_PLAIN ? kStepBlock : 256);
  const int kind = route_kind(o.route_mode);
  const bool wg = kind == MHMHIP_K_ROUTE_WG;
  // tiled routing: windows of B steps (the plain path routes in windows too)
  const int32_t B = tile_window(o.route_window);
  const int64_t ring_rows = route ? ring_need(x, kind, S * rep, B, nr) : 0;
  ScratchSwap swap(x);
  if (noalloc) {
    // graph capture and instantiation, the routing stream and the timing
    // events would all allocate
    if (o.use_graph) return x->fail(MHMHIP_EINVAL, "run_noalloc: use_graph is not available");
    if (route && x->coarse)
      return x->fail(MHMHIP_EINVAL, "run_noalloc: routing on a grid coarser than L1 is not "
                                    "available (use mhmhip_run)");
    if (rep > 1 && !x->d_qsub)
      return x->fail(MHMHIP_EINVAL, "run_noalloc: routing sub-steps keep their Qmod sums in a "
                                    "context buffer that a first mhmhip_run allocates");
    if (x->prof) return x->fail(MHMHIP_EINVAL, "run_noalloc: turn profiling off first");
    unsigned char *base = nullptr;
    CHECK_RC(scratch_base(x, &base));
    size_t used = 0;
    auto carve = [&](size_t bytes) {
      const size_t at = used;
      used = (size_t)round_up((int64_t)(used + bytes), 256);
      return base + at;
    };
    const int64_t gper = (int64_t)x->n_sets * std::max(1, x->n_gauges);
    const size_t need = (size_t)round_up((int64_t)sizeof(StepDev) * nsteps, 256) +
                        (route ? (size_t)round_up(8 * nr * gper, 256) +
                                     (size_t)(8 * (ring_rows + x->tile_h) * x->ld)
                               : 0);
    if (need > kScratchBytes)
      return x->fail(MHMHIP_EINVAL,
                     "run_noalloc: the call needs " + std::to_string(need) + " bytes of scratch "
                     "(step table, gauge series, runoff ring), more than the " +
                     std::to_string(kScratchBytes) + " it has (fewer steps, cells, sets or a "
                     "smaller route_window, or mhmhip_run)");
    StepDev *st = reinterpret_cast<StepDev *>(carve(sizeof(StepDev) * nsteps));
    double *gb = route ? reinterpret_cast<double *>(carve(8 * (size_t)nr * gper)) : nullptr;
    double *rg = route ? reinterpret_cast<double *>(carve(8 * (ring_rows + x->tile_h) * x->ld))
                       : nullptr;
    swap.swap_in(st, nsteps, route ? gb : x->d_gauge_buf, route ? nr : x->gauge_buf_rows,
                 route ? rg : x->d_ring, route ? ring_rows + x->tile_h : x->ring_len);
  }
  CHECK_RC(ensure_steps(x, nsteps));
  CHECK_HIP(x, hipMemcpyAsync(x->d_steps, hs.data(), sizeof(StepDev) * nsteps,
                              hipMemcpyHostToDevice, x->stream),
            "upload steps");
  if (route) {
    CHECK_RC(ensure_ring(x, ring_rows + x->tile_h));  // ring_row's rows past R
    CHECK_RC(ensure_gauge_buf(x, nr));
    if (rep > 1 && !x->d_qsub) CHECK_RC(x->dalloc(&x->d_qsub, x->ld, "alloc Qmod sub-step sums"));
  }
  if (route) CHECK_RC(normalize_qtr(x));
  if (route && x->coarse && x->cring_rows < S) {
    // L1_total_runoff of one launch's steps, aggregated per node after it
    x->dfree(x->d_cring, x->cring_rows * x->ld);
    x->cring_rows = 0;
    CHECK_RC(x->dalloc(&x->d_cring, (int64_t)S * x->ld, "alloc cell runoff rows"));
    x->cring_rows = S;
  }
  CellArgs a = cell_args(x, route);
  RouteArgs r = route_args(x, 0, nr, 0);
  if (route) {
    a.ring_len = r.ring_len = ring_rows;  // this call's rows of the ring
    a.ring_rep = r.rep = rep;
    if (kind == MHMHIP_K_ROUTE_TILE) a.lag = x->d_lag;  // qOUT written lag rows late
    if (rep > 1) {
      // the last model step's sub-steps summed into d_qsub: its Qmod
      CHECK_HIP(x, hipMemsetAsync(x->d_qsub, 0, sizeof(double) * x->ld, x->stream), "run");
      r.qacc = x->d_qsub;
      r.acc_from = nr - rep;
    }
  }
  // on a routing grid coarser than L1 the cell kernels write each step's
  // L1_total_runoff to the launch's cell rows, and runoff_agg_kernel turns
  // them into the nodes' qOUT at the ring rows the cell kernels would write
  const uint8_t *node_lag = a.lag;
  if (route && x->coarse) {
    a.ring = x->d_cring;
    a.ring_len = S;
    a.lag = nullptr;
    a.ring_rep = 1;
    a.raw_total = 1;
  }
  const TileArgs g = tile_args(x, B, nr);
  const int32_t L_end = g.nwin + g.tau_max;
  const bool graph = o.use_graph && o.mode == MHMHIP_RUN_FUSED;
  // A captured fused run puts the tiled routing on a stream of its own: a
  // routing launch waits for the cell launch that completes its window, and
  // the cell launches run on while the routing of earlier windows proceeds.
  // When the ring is shorter than the call, a cell launch also waits for the
  // routing launch that last reads the rows it overwrites.
  const bool two = graph && route && kind == MHMHIP_K_ROUTE_TILE;
  if (two && !x->rstream)
    CHECK_HIP(x, hipStreamCreateWithFlags(&x->rstream, hipStreamNonBlocking), "routing stream");
  const bool wrap = ring_rows < nr;
  // the whole loop, in steps relative to this call (replayable as a graph)
  auto enqueue = [&]() -> int {
    int32_t L = 0;  // next tiled routing launch
    hipStream_t rs = two ? x->rstream : x->stream;
    std::vector<hipEvent_t> evL(two && wrap ? L_end : 0, nullptr);  // routing launch L done
    if (two) CHECK_RC(capture_edge(x, x->stream, rs));  // the routing stream joins the capture
    for (int32_t k = 0; k < nsteps; k += S) {
      const int n = std::min<int32_t>(S, nsteps - k);
      // the launch's routing steps [u0, u1)
      const int64_t u0 = (int64_t)k * rep, u1 = (int64_t)(k + n) * rep;
      if (two && wrap && u1 - 1 - ring_rows >= 0) {
        // rows of the routing steps up to u1 - 1 - R are overwritten: the
        // last routing launch reading them is enqueued already
        const int64_t Ln = (u1 - 1 - ring_rows) / B + g.tau_max;
        if (Ln >= L) return x->fail(MHMHIP_EINVAL, "run: ring shorter than the routing lag");
        int64_t Lw = Ln;
        while (Lw >= 0 && !evL[Lw]) --Lw;  // a launch with no bin due reads nothing
        if (Lw >= 0) CHECK_HIP(x, hipStreamWaitEvent(x->stream, evL[Lw], 0), "capture edge");
      }
      if (o.mode == MHMHIP_RUN_PLAIN) {
        CHECK_HIP(x, timed(x, MHMHIP_K_CELL_STEP, (double)x->n_cells * x->n_sets, [&] {
                    return launch_cell_step(x->nh, a, hs[k], k, block, x->n_sets, x->stream);
                  }),
                  "cell_step");
      } else {
        CHECK_HIP(x, timed(x, MHMHIP_K_CELL_FUSED, (double)x->n_cells * x->n_sets * n, [&] {
                    return launch_cell_fused(x->nh, a, k, n, k, block, x->n_sets, x->stream);
                  }),
                  "cell_fused");
      }
      if (!route) continue;
      if (x->coarse)
        CHECK_HIP(x, launch_runoff_agg(x->d_cring, S, x->d_agg_start, x->d_area, x->n_nodes,
                                       x->sstride, x->ld, x->n_sets, x->d_ring, node_lag,
                                       ring_rows, rep, k, n,
                                       kHourSecs * (double)x->cfg.timestep, x->stream),
                  "runoff_agg");
      if (kind == MHMHIP_K_ROUTE_TILE && x->n_inflow > 0)
        CHECK_HIP(x, launch_inflow_patch(r, x->d_ring, x->d_lag, u0, u1, x->stream, rep),
                  "inflow_patch");
      if (kind == MHMHIP_K_ROUTE_TILE) {
        // launch L needs the steps of window L (its deepest tiles' window)
        int32_t L1 = L;
        while (L1 < L_end && std::min<int64_t>((int64_t)(L1 + 1) * B, nr) <= u1) ++L1;
        if (two && L1 > L) {
          CHECK_RC(capture_edge(x, x->stream, rs));
          for (int32_t l = L; l < L1; ++l) {
            CHECK_RC(route_tiles(x, r, g, l, l + 1, rs));
            if (wrap) CHECK_RC(capture_edge(x, rs, nullptr, &evL[l]));
          }
        } else if (!two) {
          CHECK_RC(route_tiles(x, r, g, L, L1));
        }
        L = L1;
      } else {
        CHECK_RC(route_diagonals(x, r, u0, u1, false, wg));
      }
    }
    if (route && kind == MHMHIP_K_ROUTE_TILE) {
      if (two) {
        if (L < L_end) {
          CHECK_RC(capture_edge(x, x->stream, rs));
          CHECK_RC(route_tiles(x, r, g, L, L_end, rs));
        }
        CHECK_RC(capture_edge(x, rs, x->stream));  // join
      } else {
        CHECK_RC(route_tiles(x, r, g, L, L_end));
      }
    }
    if (route && kind != MHMHIP_K_ROUTE_TILE) CHECK_RC(route_diagonals(x, r, nr, 0, true, wg));
    return MHMHIP_OK;
  };
  if (graph) {
    mhmhip_ctx::GraphKey key;
    key.nsteps = nsteps;
    key.mode = o.mode;
    key.S = S;
    key.block = block;
    key.wg = kind;
    key.prof = x->prof ? x->prof_stride : 0;
    key.ring = route ? x->d_ring : nullptr;
    key.gauge = route ? x->d_gauge_buf : nullptr;
    key.steps = x->d_steps;
    key.inflow = x->d_inflow_q;
    key.ring_len = ring_rows;
    key.B = B;
    key.qacc = r.qacc;
    key.out = a.out.sum;
    key.out_n = a.out.sum ? a.out.n : 0;
    key.rep = rep;
    key.cring = route && x->coarse ? x->d_cring : nullptr;
    if (!x->graph_exec || !(key == x->gkey)) {
      drop_graph(x);
      CHECK_HIP(x, hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal), "capture");
      x->capturing = true;
      const int rc = enqueue();
      x->capturing = false;
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(x->stream, &g);
      if (rc != MHMHIP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      CHECK_HIP(x, e, "capture");
      x->graph = g;
      CHECK_HIP(x, hipGraphInstantiate(&x->graph_exec, g, nullptr, nullptr, 0), "instantiate");
      x->gkey = key;
    }
    CHECK_HIP(x, hipGraphLaunch(x->graph_exec, x->stream), "graph launch");
  } else {
    CHECK_RC(enqueue());
  }
  if (route) {
    credit_route(x, kind, nr);
    x->qtr_slot = (int)((nr - 1) & 1);
    if (x->qacc_on) x->qacc_count += nsteps;
    x->qsub_loops = rep;
  }
  x->t_cell += nsteps;
  x->t_route = x->t_cell;
  if (x->out_on) x->out_count += nsteps;
  const int64_t gper = (int64_t)x->n_sets * x->n_gauges;
  std::vector<double> gsub(route && gauge_out && rep > 1 ? nr * gper : 0);
  if (route && gauge_out && x->n_gauges > 0) {
    CHECK_HIP(x, hipMemcpyAsync(rep > 1 ? gsub.data() : gauge_out, x->d_gauge_buf,
                                sizeof(double) * nr * gper, hipMemcpyDeviceToHost, x->stream),
              "download gauges");
  }
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "run");
  if (!gsub.empty()) {
    // qAcc = qAcc + Qmod over the sub-steps, then Qmod = qAcc / rout_loop
    // (mo_mrm_routing.f90:246-284), per gauge
    for (int64_t t = 0; t < nsteps; ++t)
      for (int64_t i = 0; i < gper; ++i) {
        double acc = 0.0;
        for (int32_t j = 0; j < rep; ++j) acc = acc + gsub[(t * rep + j) * gper + i];
        gauge_out[t * gper + i] = acc / (double)rep;
      }
  }
  if (route && gauge_out) copy_dup_gauges(x, gauge_out, nsteps);
  if (graph && x->prof) CHECK_RC(prof_collect_graph(x));
  return prof_collect(x);
}

}  // namespace

extern "C" int mhmhip_run(mhmhip_ctx *x, const mhmhip_step *steps, int32_t nsteps,
                          const mhmhip_run_opts *opts, double *gauge_out) {
  return run_impl(x, steps, nsteps, opts, gauge_out, false);
}

extern "C" int mhmhip_run_noalloc(mhmhip_ctx *x, const mhmhip_step *steps, int32_t nsteps,
                                  const mhmhip_run_opts *opts, double *gauge_out) {
  return run_impl(x, steps, nsteps, opts, gauge_out, true);
}

extern "C" int mhmhip_set_profiling(mhmhip_ctx *x, int32_t on) {
  if (!x) return MHMHIP_EINVAL;
  CHECK_HIP(x, hipSetDevice(x->cfg.device), "set_profiling");
  CHECK_HIP(x, hipStreamSynchronize(x->stream), "set_profiling");
  CHECK_RC(prof_collect(x));
  x->prof = on > 0;
  x->prof_stride = on > 0 ? on : 1;
  for (int k = 0; k < MHMHIP_K_COUNT; ++k) x->prof_seq[k] = 0;
  if (x->prof)
    for (int k = 0; k < MHMHIP_K_COUNT; ++k) {
      x->prof_ms[k] = 0.0;
      x->prof_n[k] = 0;
      x->prof_all[k] = 0;
      x->prof_units[k] = 0.0;
    }
  return MHMHIP_OK;
}

extern "C" int mhmhip_get_profile(mhmhip_ctx *x, int32_t kernel, double *ms, int64_t *launches,
                                  double *units) {
  if (!x) return MHMHIP_EINVAL;
  if (kernel < 0 || kernel >= MHMHIP_K_COUNT) return x->fail(MHMHIP_EINVAL, "get_profile: bad kernel");
  // routing units are counted for all launches (a call's node-steps), so
  // they are scaled to the timed ones; cell units come with each timed launch
  const bool routing = kernel == MHMHIP_K_ROUTE_DIAG || kernel == MHMHIP_K_ROUTE_WG ||
                       kernel == MHMHIP_K_ROUTE_TILE;
  const double frac = routing && x->prof_all[kernel]
                          ? (double)x->prof_n[kernel] / x->prof_all[kernel]
                          : 1.0;
  if (ms) *ms = x->prof_ms[kernel];
  if (launches) *launches = x->prof_n[kernel];
  if (units) *units = x->prof_units[kernel] * frac;
  return MHMHIP_OK;
}
