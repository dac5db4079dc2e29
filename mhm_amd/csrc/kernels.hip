// kernels.hip -- HIP kernels for gfx950: the per-cell chain and the routing.
//
// Cell kernels run one lane per (cell, parameter set).  Every row of P, S, X
// and F is a structure-of-arrays slab, so each of the ~70 loads and stores a
// lane makes per step is a fully coalesced 512-byte wavefront access.
//
// Routing is pipelined over (depth, time) diagonals: node p at depth delta
// (distance to its outlet) handles step t on diagonal d = t + dmax - delta.
// Its upstream nodes sit one level deeper, so they produced step t on
// diagonal d-1; a two-slot parity buffer of routed outflow (qTR) keeps step t
// readable while those nodes already write step t+1.  A whole run of T steps
// takes T + dmax diagonals instead of T * dmax level sweeps.
#include <hip/hip_runtime.h>

#include "chain.h"
#include "internal.h"
#include "launch.h"

namespace mhmhip {

template <int NH>
struct CellRegs {
  double fS, thrS, c1, k0s, k1s, k2s, kps, alpha, karst, thrU, fasp, harsam, tanlat, sinlat,
      coslat, dmaxs, dnps, incr, tth, ddthr;
  double sat[NH], fr[NH], fc[NH], wp[NH], bexp[NH];
  double inter, pack, seal, unsat, satS;
  double sm[NH];
};

template <int NH>
struct FluxRegs {
  double pet, temp, prec, thr, ecan, ddf, melt, peff, rain, snow, rseal, aseal, fast, slow, perc,
      bflow, total;
  double infil[NH], aet[NH];
};

template <int NH>
__device__ __forceinline__ void load_params(const CellArgs &a, int64_t v, CellRegs<NH> &c) {
  const double *P = a.P + v;
  const int64_t ld = a.ld;
  c.fS = P[PR_FS * ld];
  c.thrS = P[PR_SEALTHR * ld];
  c.c1 = a.jarvis ? P[PR_JARVIS * ld] : 0.0;
  c.k0s = P[PR_K0S * ld];
  c.k1s = P[PR_K1S * ld];
  c.k2s = P[PR_K2S * ld];
  c.kps = P[PR_KPS * ld];
  c.alpha = P[PR_ALPHA * ld];
  c.karst = P[PR_KARST * ld];
  c.thrU = P[PR_UNSATTHR * ld];
  c.fasp = P[PR_FASP * ld];
  c.harsam = P[PR_HARSAM * ld];
  c.tanlat = P[PR_TANLAT * ld];
  c.sinlat = P[PR_SINLAT * ld];
  c.coslat = P[PR_COSLAT * ld];
  c.dmaxs = P[PR_DMAXS * ld];
  c.dnps = P[PR_DNPS * ld];
  c.incr = P[PR_INCR * ld];
  c.tth = P[PR_TTH * ld];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    c.sat[h] = P[(pr_sat(NH) + h) * ld];
    c.fr[h] = P[(pr_froots(NH) + h) * ld];
    c.fc[h] = P[(pr_fc(NH) + h) * ld];
    c.wp[h] = P[(pr_wp(NH) + h) * ld];
    c.bexp[h] = P[(pr_exp(NH) + h) * ld];
  }
  // snow_accum_melt's branch threshold (mo_snow_accum_melt.f90:127)
  c.ddthr = (c.dmaxs - c.dnps) / c.incr;
}

template <int NH>
__device__ __forceinline__ void load_states(const CellArgs &a, int64_t v, CellRegs<NH> &c) {
  const double *S = a.S + v;
  const int64_t ld = a.ld;
  c.inter = S[SR_INTER * ld];
  c.pack = S[SR_SNOW * ld];
  c.seal = S[SR_SEAL * ld];
  c.unsat = S[SR_UNSAT * ld];
  c.satS = S[SR_SAT * ld];
#pragma unroll
  for (int h = 0; h < NH; ++h) c.sm[h] = S[(SR_NSCALAR + h) * ld];
}

template <int NH>
__device__ __forceinline__ void store_states(const CellArgs &a, int64_t v, const CellRegs<NH> &c) {
  double *S = a.S + v;
  const int64_t ld = a.ld;
  S[SR_INTER * ld] = c.inter;
  S[SR_SNOW * ld] = c.pack;
  S[SR_SEAL * ld] = c.seal;
  S[SR_UNSAT * ld] = c.unsat;
  S[SR_SAT * ld] = c.satS;
#pragma unroll
  for (int h = 0; h < NH; ++h) S[(SR_NSCALAR + h) * ld] = c.sm[h];
}

template <int NH>
__device__ __forceinline__ void store_fluxes(const CellArgs &a, int64_t v, const FluxRegs<NH> &x) {
  double *X = a.X + v;
  const int64_t ld = a.ld;
  X[XR_PET * ld] = x.pet;
  X[XR_TEMP * ld] = x.temp;
  X[XR_PREC * ld] = x.prec;
  X[XR_THROUGHFALL * ld] = x.thr;
  X[XR_AETCAN * ld] = x.ecan;
  X[XR_DEGDAY * ld] = x.ddf;
  X[XR_MELT * ld] = x.melt;
  X[XR_PEFF * ld] = x.peff;
  X[XR_RAIN * ld] = x.rain;
  X[XR_SNOW * ld] = x.snow;
  X[XR_RUNOFFSEAL * ld] = x.rseal;
  X[XR_AETSEAL * ld] = x.aseal;
  X[XR_FAST * ld] = x.fast;
  X[XR_SLOW * ld] = x.slow;
  X[XR_PERC * ld] = x.perc;
  X[XR_BASEFLOW * ld] = x.bflow;
  X[XR_TOTAL * ld] = x.total;
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    X[(xr_infil(NH) + h) * ld] = x.infil[h];
    X[(xr_aet(NH) + h) * ld] = x.aet[h];
  }
}

// Daily PET before disaggregation (get_corrected_pet, mo_meteo_handler.f90:1050-1097)
template <int NH>
__device__ __forceinline__ double pet_day(const CellArgs &a, const CellRegs<NH> &c,
                                          const StepDev &st, int64_t v, int64_t f, double tavg,
                                          double tmax, double tmin) {
  if (a.pet_case == 1)
    return c.fasp * pet_hargreaves(c.harsam, tavg, tmax, tmin, c.tanlat, c.sinlat, c.coslat,
                                   st.dr, st.tand, st.sind, st.cosd);
  const double petin = a.F[((int64_t)MHMHIP_F_PET * a.forcing_days + st.day) * a.sstride + f];
  if (a.pet_case == 0) return c.fasp * petin;
  return a.P[(pr_petlai(NH) + st.month0) * a.ld + v] * petin;
}

// One step of mHM() for one cell after the meteo getters (mo_mhm.f90:455-519)
template <int NH>
__device__ __forceinline__ void chain_step(const CellArgs &a, CellRegs<NH> &c, FluxRegs<NH> &x,
                                           double maxinter, double evap) {
  canopy_interc(x.pet, maxinter, x.prec, c.inter, x.thr, x.ecan);
  snow_accum_melt(c.incr, c.dmaxs, c.dnps, c.ddthr, x.prec, x.temp, c.tth, x.thr, c.pack, x.ddf,
                  x.melt, x.peff, x.rain, x.snow);
  soil_moisture<NH>(a.jarvis != 0, c.fS, c.thrS, x.pet, evap, c.sat, c.fr, c.fc, c.wp, c.bexp,
                    c.c1, x.ecan, x.peff, x.rseal, c.seal, x.infil, c.sm, x.aet, x.aseal);
  runoff_unsat_zone(c.k1s, c.kps, c.k0s, c.alpha, c.karst, x.infil[NH - 1], c.thrU, c.satS,
                    c.unsat, x.slow, x.fast, x.perc);
  runoff_sat_zone(c.k2s, c.satS, x.bflow);
  x.total = total_runoff(c.fS, x.fast, x.slow, x.bflow, x.rseal);
}

// ---------------------------------------------------------------------------
// Plain path: one launch per model step; every state and flux is written.
template <int NH>
__global__ __launch_bounds__(256) void cell_step_kernel(CellArgs a, StepDev st, MonthTables mt,
                                                        int64_t t) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  const int64_t v = (int64_t)blockIdx.y * a.sstride + c;
  CellRegs<NH> r;
  FluxRegs<NH> x;
  load_params<NH>(a, v, r);
  load_states<NH>(a, v, r);
  const int m = st.month0;
  const bool isday = st.isday != 0;
  const double *F = a.F + (int64_t)st.day * a.sstride + c;
  const int64_t fv = a.forcing_days * a.sstride;
  const double pre = F[MHMHIP_F_PRE * fv], tavg = F[MHMHIP_F_TAVG * fv];
  const double tmax = a.pet_case == 1 ? F[MHMHIP_F_TMAX * fv] : 0.0;
  const double tmin = a.pet_case == 1 ? F[MHMHIP_F_TMIN * fv] : 0.0;
  const double pd = pet_day<NH>(a, r, st, v, c, tavg, tmax, tmin);
  x.pet = a.hourly ? pd : disagg_flux(isday, a.nT, pd, mt.fday_pet[m], mt.fnight_pet[m]);
  x.temp = a.hourly ? tavg : disagg_state_add(isday, a.nT, tavg, mt.fday_temp[m], mt.fnight_temp[m]);
  x.prec = a.hourly ? pre : disagg_flux(isday, a.nT, pre, mt.fday_prec[m], mt.fnight_prec[m]);
  if (st.init_sm) {
#pragma unroll
    for (int h = 0; h < NH; ++h) r.sm[h] = 0.5 * r.fc[h];
  }
  const double maxinter = a.P[(pr_maxinter(NH) + m) * a.ld + v];
  chain_step<NH>(a, r, x, maxinter, mt.evap[m]);
  store_states<NH>(a, v, r);
  store_fluxes<NH>(a, v, x);
  if (a.ring) a.ring[(t % a.ring_len) * a.ld + v] = x.total;
}

// ---------------------------------------------------------------------------
// Fused path: nsteps model steps per launch with parameters and states held
// in registers.  Forcing and Hargreaves PET are evaluated once per forcing day
// (the reference recomputes the same value every hour).  Only L1 total runoff
// leaves the lane each step; states and the last step's fluxes are written at
// the end.
template <int NH>
__global__ __launch_bounds__(256) void cell_fused_kernel(CellArgs a, MonthTables mt, int64_t t0,
                                                         int32_t nsteps, int32_t step0) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  const int64_t v = (int64_t)blockIdx.y * a.sstride + c;
  CellRegs<NH> r;
  FluxRegs<NH> x;
  load_params<NH>(a, v, r);
  load_states<NH>(a, v, r);
  const int64_t fv = a.forcing_days * a.sstride;
  int cached_day = -1, cached_doy = -1;
  double pre = 0.0, tavg = 0.0, pd = 0.0;
  for (int32_t s = 0; s < nsteps; ++s) {
    const StepDev st = a.steps[step0 + s];
    const int m = st.month0;
    const bool isday = st.isday != 0;
    if (st.day != cached_day || st.doy != cached_doy) {
      const double *F = a.F + (int64_t)st.day * a.sstride + c;
      pre = F[MHMHIP_F_PRE * fv];
      tavg = F[MHMHIP_F_TAVG * fv];
      const double tmax = a.pet_case == 1 ? F[MHMHIP_F_TMAX * fv] : 0.0;
      const double tmin = a.pet_case == 1 ? F[MHMHIP_F_TMIN * fv] : 0.0;
      pd = pet_day<NH>(a, r, st, v, c, tavg, tmax, tmin);
      cached_day = st.day;
      cached_doy = st.doy;
    }
    x.pet = a.hourly ? pd : disagg_flux(isday, a.nT, pd, mt.fday_pet[m], mt.fnight_pet[m]);
    x.temp = a.hourly ? tavg
                      : disagg_state_add(isday, a.nT, tavg, mt.fday_temp[m], mt.fnight_temp[m]);
    x.prec = a.hourly ? pre : disagg_flux(isday, a.nT, pre, mt.fday_prec[m], mt.fnight_prec[m]);
    if (st.init_sm) {
#pragma unroll
      for (int h = 0; h < NH; ++h) r.sm[h] = 0.5 * r.fc[h];
    }
    const double maxinter = a.P[(pr_maxinter(NH) + m) * a.ld + v];
    chain_step<NH>(a, r, x, maxinter, mt.evap[m]);
    if (a.ring) a.ring[((t0 + s) % a.ring_len) * a.ld + v] = x.total;
  }
  store_states<NH>(a, v, r);
  store_fluxes<NH>(a, v, x);
}

// ---------------------------------------------------------------------------
// One node, one step of L11_runoff_acc + add_inflow + L11_routing
// (mo_mrm_pre_routing.f90:77-143,179-214; mo_mrm_routing.f90:425-472).
__device__ __forceinline__ void route_node(const RouteArgs &r, int64_t p, int32_t set, int64_t t,
                                           NodeMeta m) {
  const int64_t ld = (int64_t)r.n_sets * r.sstride;
  const int64_t v = (int64_t)set * r.sstride + p;
  const int32_t y = m.packed;
  // L11_runoff_acc, map_flag with L1 = L11: qAcc = (0 + q*area) * 1000 / TST
  const double q = r.ring[(t % r.ring_len) * ld + v];
  double qout = ((0.0 + q * r.area[p]) * 1000.0) / r.tst;
  if (y & NF_INFLOW) {
    for (int32_t g = 0; g < r.n_inflow; ++g) {
      if (r.inflow_node[g] == p) {
        const double qi = r.inflow_q[(t - r.t_base) * r.n_inflow + g];
        qout = r.inflow_hw[g] ? qout + qi : qi;
      }
    }
  }
  if (r.qOUT && t == r.t_hi - 1) r.qOUT[v] = qout;
  // qTIN(iNode,IT): upstream qTR in routing rank, then own qOUT
  const double *pub_t = r.pub + (t & 1) * ld + (int64_t)set * r.sstride + m.child;
  double s = 0.0;
  const int nc = nm_nchild(y);
  for (int c = 0; c < nc; ++c) s = s + pub_t[c];
  if (y & NF_QOUT) s = s + qout;
  if (y & NF_LINK) {
    const double qtin1 = r.qTIN[v];
    const double qtr1 = r.pub[((t - 1) & 1) * ld + v];
    double qtr = qtr1 + r.C1[v] * (qtin1 - qtr1) + r.C2[v] * (s - qtin1);
    if (y & NF_CUT) qtr = 0.0;
    r.pub[(t & 1) * ld + v] = qtr;
  }
  r.qTIN[v] = s;
  if (y & NF_GAUGE) {
    for (int32_t g = 0; g < r.n_gauges; ++g)
      if (r.gauge_node[g] == p)
        r.gauge_out[((t - r.t_base) * r.n_sets + set) * r.n_gauges + g] = s;
  }
}

__global__ __launch_bounds__(256) void route_diag_kernel(RouteArgs r, int64_t d) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= r.n_nodes) return;
  const NodeMeta m = r.meta[p];
  const int64_t t = d - (int64_t)(r.dmax - nm_depth(m.packed));
  if (t < r.t_lo || t >= r.t_hi) return;
  route_node(r, p, (int32_t)blockIdx.y, t, m);
}

// Workgroup-resident routing: one workgroup owns one (domain, parameter set)
// tree and walks that domain's own diagonals d in [d_lo, d_end) -- node at
// depth delta handles step t = d - (dmax_dom - delta) -- with a workgroup
// barrier between diagonals instead of a kernel boundary.  drain != 0 walks
// the dmax_dom diagonals that finish the last steps of a run.
__global__ __launch_bounds__(1024) void route_wg_kernel(RouteArgs r, const int64_t *dom_start,
                                                        const int32_t *dom_dmax, int64_t d_lo,
                                                        int64_t d_hi, int32_t drain) {
  const int32_t dom = blockIdx.x;
  const int32_t set = blockIdx.y;
  const int64_t lo = dom_start[dom], hi = dom_start[dom + 1];
  const int32_t dm = dom_dmax[dom];
  const int64_t d_end = drain ? d_lo + dm : d_hi;
  for (int64_t d = d_lo; d < d_end; ++d) {
    for (int64_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
      const NodeMeta m = r.meta[p];
      const int64_t t = d - (int64_t)(dm - nm_depth(m.packed));
      if (t >= r.t_lo && t < r.t_hi) route_node(r, p, set, t, m);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host launchers
template <int NH>
static hipError_t step_nh(const CellArgs &a, const StepDev &st, const MonthTables &mt, int64_t t,
                          int block, int n_sets, hipStream_t s) {
  dim3 grid((unsigned)((a.n_cells + block - 1) / block), (unsigned)n_sets);
  hipLaunchKernelGGL(cell_step_kernel<NH>, grid, dim3(block), 0, s, a, st, mt, t);
  return hipGetLastError();
}

template <int NH>
static hipError_t fused_nh(const CellArgs &a, const MonthTables &mt, int64_t t0, int nsteps,
                           int step0, int block, int n_sets, hipStream_t s) {
  dim3 grid((unsigned)((a.n_cells + block - 1) / block), (unsigned)n_sets);
  hipLaunchKernelGGL(cell_fused_kernel<NH>, grid, dim3(block), 0, s, a, mt, t0, nsteps, step0);
  return hipGetLastError();
}

hipError_t launch_cell_step(int nh, const CellArgs &a, const StepDev &st, const MonthTables &mt,
                            int64_t t, int block, int n_sets, hipStream_t s) {
  switch (nh) {
    case 1: return step_nh<1>(a, st, mt, t, block, n_sets, s);
    case 2: return step_nh<2>(a, st, mt, t, block, n_sets, s);
    case 3: return step_nh<3>(a, st, mt, t, block, n_sets, s);
    case 4: return step_nh<4>(a, st, mt, t, block, n_sets, s);
    case 5: return step_nh<5>(a, st, mt, t, block, n_sets, s);
    case 6: return step_nh<6>(a, st, mt, t, block, n_sets, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_cell_fused(int nh, const CellArgs &a, const MonthTables &mt, int64_t t0,
                             int nsteps, int step0, int block, int n_sets, hipStream_t s) {
  switch (nh) {
    case 1: return fused_nh<1>(a, mt, t0, nsteps, step0, block, n_sets, s);
    case 2: return fused_nh<2>(a, mt, t0, nsteps, step0, block, n_sets, s);
    case 3: return fused_nh<3>(a, mt, t0, nsteps, step0, block, n_sets, s);
    case 4: return fused_nh<4>(a, mt, t0, nsteps, step0, block, n_sets, s);
    case 5: return fused_nh<5>(a, mt, t0, nsteps, step0, block, n_sets, s);
    case 6: return fused_nh<6>(a, mt, t0, nsteps, step0, block, n_sets, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_route_diag(const RouteArgs &r, int64_t d, hipStream_t s) {
  dim3 grid((unsigned)((r.n_nodes + 255) / 256), (unsigned)r.n_sets);
  hipLaunchKernelGGL(route_diag_kernel, grid, dim3(256), 0, s, r, d);
  return hipGetLastError();
}

hipError_t launch_route_wg(const RouteArgs &r, const int64_t *dom_start, const int32_t *dom_dmax,
                           int n_domains, int64_t d_lo, int64_t d_hi, int drain, int block,
                           hipStream_t s) {
  dim3 grid((unsigned)n_domains, (unsigned)r.n_sets);
  hipLaunchKernelGGL(route_wg_kernel, grid, dim3(block), 0, s, r, dom_start, dom_dmax, d_lo, d_hi,
                     drain);
  return hipGetLastError();
}

}  // namespace mhmhip
