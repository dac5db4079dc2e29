// launch.h -- host-side launchers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mhmhip.h"
#include "internal.h"

namespace mhmhip {

hipError_t launch_cell_step(int nh, const CellArgs &a, const StepDev &st, int64_t t, int block,
                            int n_sets, hipStream_t s);
hipError_t launch_cell_fused(int nh, const CellArgs &a, int64_t t0, int nsteps, int step0,
                             int block, int n_sets, hipStream_t s);
hipError_t launch_route_diag(const RouteArgs &r, int64_t d, int64_t p_lo, int64_t p_hi,
                             hipStream_t s);
hipError_t launch_route_tile(const RouteArgs &r, const TileArgs &g, int64_t bin_lo, int64_t bin_hi,
                             int64_t n_bins, int32_t L, hipStream_t s);
hipError_t launch_inflow_patch(const RouteArgs &r, double *ring, const uint8_t *lag, int64_t t_lo,
                               int64_t t_hi, hipStream_t s, int32_t rep = 1);
hipError_t launch_meteo_l2(const double *data2, const uint8_t *mask2, int32_t nr2, int64_t n2,
                           const MeteoL2Cell *cells, int64_t ncells, int32_t ndays, double *F,
                           int64_t day_stride, bool aggregate, hipStream_t s);
hipError_t launch_forcing_gather(const double *raw, const FRowMap &m, int32_t ndays,
                                 int64_t n_cells, const int32_t *pos, double *F, int32_t n_cols,
                                 int32_t day0, int64_t sstride, hipStream_t s);
hipError_t launch_runoff_sum(const double *total, double *acc, double *ring, const double *area,
                             const uint8_t *lag, int64_t n_cells, int64_t sstride, int64_t ld,
                             int n_sets, double tst, bool route, hipStream_t s);
hipError_t launch_runoff_agg(const double *tot, int64_t tot_rows, const int32_t *cstart,
                             const double *area, int64_t n_nodes, int64_t sstride, int64_t ld,
                             int n_sets, double *ring, const uint8_t *lag, int64_t ring_len,
                             int32_t rep, int64_t t0, int32_t nsteps, double tst, hipStream_t s);
hipError_t launch_mprlite(const MprLiteArgs &a, int32_t set, hipStream_t s);
hipError_t launch_route_wg(const RouteArgs &r, const int64_t *dom_start, const int32_t *dom_dmax,
                           int n_domains, int64_t d_lo, int64_t d_hi, int drain, int block,
                           hipStream_t s);

}  // namespace mhmhip
