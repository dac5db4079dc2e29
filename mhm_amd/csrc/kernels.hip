// kernels.hip -- HIP kernels for gfx950: the per-cell chain and the routing.
//
// Cell kernels run one lane per (cell, parameter set).  Every row of P, S, X
// and F is a structure-of-arrays slab, so each of the ~70 loads and stores a
// lane makes per step is a fully coalesced 512-byte wavefront access.  Rows
// are addressed as a wave-uniform base (SGPRs) plus one 32-bit byte offset
// per lane shared by all rows, which maps onto global_load/store with
// saddr + voffset and keeps per-row 64-bit addresses out of the VGPRs.
//
// Routing is pipelined over (depth, time) diagonals: node p at depth delta
// (distance to its outlet) handles step t on diagonal d = t + dmax - delta.
// Its upstream nodes sit one level deeper, so they produced step t on
// diagonal d-1; a two-slot parity buffer of routed outflow (qTR) keeps step t
// readable while those nodes already write step t+1.  A whole run of T steps
// takes T + dmax diagonals instead of T * dmax level sweeps.  The default
// schedule (route_tile_kernel) runs the same diagonals inside tiles of the
// tree, one workgroup per bin of tiles and window of steps, with the window
// handed from tile to tile between launches; route_diag_kernel and
// route_wg_kernel run them over the whole grid or inside one workgroup.
#include <hip/hip_runtime.h>

#include "chain.h"
#include "internal.h"
#include "launch.h"

namespace mhmhip {

// row[boff / 8] with a uniform row base and a 32-bit per-lane byte offset
__device__ __forceinline__ double ldr(const double *row, uint32_t boff) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(row) + boff);
}
__device__ __forceinline__ void str(double *row, uint32_t boff, double x) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(row) + boff) = x;
}
// Waits for v's load here and makes v count as defined by this (empty)
// statement.  A value loaded in a block that runs once per day or month and
// used every step otherwise leaves the compiler's memory-counter tracking
// unsure whether its load is still in flight where the paths merge, and the
// wait it puts there (every step) also waits for the step before's ring
// store.
__device__ __forceinline__ void settle(double &v) { asm volatile("" : "+v"(v)); }
// The lane's byte offset, made opaque where it is used: an address formed
// from it inside a loop cannot be hoisted out of the loop.  Rows stored in a
// block that runs once per launch (the last step's fluxes) then keep no
// 64-bit per-lane address in VGPRs across the whole time loop.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}
// streaming store for outputs no kernel of the step reads back
#ifndef MHMHIP_NT_STORE
#define MHMHIP_NT_STORE 1
#endif
__device__ __forceinline__ void stn(double *row, uint32_t boff, double x) {
#if MHMHIP_NT_STORE
  __builtin_nontemporal_store(x, reinterpret_cast<double *>(reinterpret_cast<char *>(row) + boff));
#else
  str(row, boff, x);
#endif
}

template <int NH>
struct CellRegs {
  double fS, thrS, c1, k0s, k1s, k2s, kps, alpha, karst, thrU, dmaxs, dnps, incr, tth, ddthr;
  double sat[NH], fr[NH], fc[NH], wp[NH], bexp[NH];
  // correctly rounded reciprocals of the per-cell divisors (div_rcp)
  double rthrS, rsat[NH], rfw[NH];
  double inter, pack, seal, unsat, satS;
  double sm[NH];
};

template <int NH>
struct FluxRegs {
  double pet, temp, prec, thr, ecan, ddf, melt, peff, rain, snow, rseal, aseal, fast, slow, perc,
      bflow, total;
  double infil[NH], aet[NH];
};

template <int NH, bool JV>
__device__ __forceinline__ void load_params(const CellArgs &a, uint32_t bv, CellRegs<NH> &c) {
  auto P = [&](int r) { return ldr(a.P + (int64_t)r * a.ld, bv); };
  c.fS = P(PR_FS);
  c.thrS = P(PR_SEALTHR);
  c.c1 = JV ? P(PR_JARVIS) : 0.0;
  c.k0s = P(PR_K0S);
  c.k1s = P(PR_K1S);
  c.k2s = P(PR_K2S);
  c.kps = P(PR_KPS);
  c.alpha = P(PR_ALPHA);
  c.karst = P(PR_KARST);
  c.thrU = P(PR_UNSATTHR);
  c.dmaxs = P(PR_DMAXS);
  c.dnps = P(PR_DNPS);
  c.incr = P(PR_INCR);
  c.tth = P(PR_TTH);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    c.sat[h] = P(pr_sat(NH) + h);
    c.fr[h] = P(pr_froots(NH) + h);
    c.fc[h] = P(pr_fc(NH) + h);
    c.wp[h] = P(pr_wp(NH) + h);
    c.bexp[h] = P(pr_exp(NH) + h);
  }
  // snow_accum_melt's branch threshold (mo_snow_accum_melt.f90:127)
  c.ddthr = (c.dmaxs - c.dnps) / c.incr;
  c.rthrS = 1.0 / c.thrS;
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    c.rsat[h] = 1.0 / c.sat[h];
    c.rfw[h] = 1.0 / (c.fc[h] - c.wp[h]);
  }
}

template <int NH>
__device__ __forceinline__ void load_states(const CellArgs &a, uint32_t bv, CellRegs<NH> &c) {
  auto S = [&](int r) { return ldr(a.S + (int64_t)r * a.ld, bv); };
  c.inter = S(SR_INTER);
  c.pack = S(SR_SNOW);
  c.seal = S(SR_SEAL);
  c.unsat = S(SR_UNSAT);
  c.satS = S(SR_SAT);
#pragma unroll
  for (int h = 0; h < NH; ++h) c.sm[h] = S(SR_NSCALAR + h);
}

template <int NH>
__device__ __forceinline__ void store_states(const CellArgs &a, uint32_t bv0,
                                             const CellRegs<NH> &c) {
  const uint32_t bv = opaque(bv0);
  auto S = [&](int r, double x) { str(a.S + (int64_t)r * a.ld, bv, x); };
  S(SR_INTER, c.inter);
  S(SR_SNOW, c.pack);
  S(SR_SEAL, c.seal);
  S(SR_UNSAT, c.unsat);
  S(SR_SAT, c.satS);
#pragma unroll
  for (int h = 0; h < NH; ++h) S(SR_NSCALAR + h, c.sm[h]);
}

template <int NH>
__device__ __forceinline__ void store_fluxes_pre(const CellArgs &a, uint32_t bv0,
                                                 const FluxRegs<NH> &x) {
  const uint32_t bv = opaque(bv0);
  auto X = [&](int r, double y) { stn(a.X + (int64_t)r * a.ld, bv, y); };
  X(XR_PET, x.pet);
  X(XR_TEMP, x.temp);
  X(XR_PREC, x.prec);
  X(XR_THROUGHFALL, x.thr);
  X(XR_AETCAN, x.ecan);
  X(XR_DEGDAY, x.ddf);
  X(XR_MELT, x.melt);
  X(XR_PEFF, x.peff);
  X(XR_RAIN, x.rain);
  X(XR_SNOW, x.snow);
  X(XR_RUNOFFSEAL, x.rseal);
  X(XR_AETSEAL, x.aseal);
  X(XR_FAST, x.fast);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    X(xr_infil(NH) + h, x.infil[h]);
    X(xr_aet(NH) + h, x.aet[h]);
  }
}

// the fluxes chain_post sets
template <int NH>
__device__ __forceinline__ void store_fluxes_post(const CellArgs &a, uint32_t bv0,
                                                  const FluxRegs<NH> &x) {
  const uint32_t bv = opaque(bv0);
  auto X = [&](int r, double y) { stn(a.X + (int64_t)r * a.ld, bv, y); };
  X(XR_SLOW, x.slow);
  X(XR_PERC, x.perc);
  X(XR_BASEFLOW, x.bflow);
  X(XR_TOTAL, x.total);
}

template <int NH>
__device__ __forceinline__ void store_fluxes(const CellArgs &a, uint32_t bv, const FluxRegs<NH> &x) {
  store_fluxes_pre<NH>(a, bv, x);
  store_fluxes_post<NH>(a, bv, x);
}

// One step's contribution to the output sums (mHM_updateDataset,
// src/mHM/mo_write_fluxes_states.f90:283-438, updateVariable): each selected
// output's value of the step -- states after it, its fluxes -- in the
// reference's expression and operation order, added to its sum.  Horizon
// sums run from 0 over the horizons in order (sum(..., dim = 2)); the
// unsealed fraction is 1 - fSealed (mo_mhm_interface_run.f90:239).
template <int NH>
__device__ __forceinline__ double pick_h(const double *v, int h) {
  double y = v[0];
#pragma unroll
  for (int k = 1; k < NH; ++k)
    if (h == k) y = v[k];
  return y;
}

template <int NH>
__device__ __forceinline__ void add_outputs(const CellArgs &a, uint32_t bv, const CellRegs<NH> &c,
                                         const FluxRegs<NH> &x) {
  const double fs = c.fS, fns = 1.0 - c.fS;
  for (int r = 0; r < a.out.n; ++r) {
    const int h = a.out.hor[r];
    double v = 0.0;
    switch (a.out.kind[r]) {
      case 1: v = c.inter; break;
      case 2: v = c.pack; break;
      case 3: v = pick_h<NH>(c.sm, h); break;
      case 4: v = pick_h<NH>(c.sm, h) / pick_h<NH>(c.sat, h); break;
      case 5: {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int k = 0; k < NH; ++k) s1 = s1 + c.sm[k];
#pragma unroll
        for (int k = 0; k < NH; ++k) s2 = s2 + c.sat[k];
        v = s1 / s2;
        break;
      }
      case 6: v = c.seal; break;
      case 7: v = c.unsat; break;
      case 8: v = c.satS; break;
      case 9: v = x.pet; break;
      case 10: {
        double s1 = 0.0;
#pragma unroll
        for (int k = 0; k < NH; ++k) s1 = s1 + x.aet[k];
        v = s1 * fns + x.ecan + x.aseal * fs;
        break;
      }
      case 11: v = x.total; break;
      case 12: v = x.rseal * fs; break;
      case 13: v = x.fast * fns; break;
      case 14: v = x.slow * fns; break;
      case 15: v = x.bflow * fns; break;
      case 16: v = x.perc * fns; break;
      case 17: v = pick_h<NH>(x.infil, h) * fns; break;
      case 19: v = pick_h<NH>(x.aet, h) * fns; break;
      case 20: v = x.peff; break;
      case 21: v = x.melt; break;
      default: break;
    }
    double *row = a.out.sum + (int64_t)r * a.ld;
    str(row, bv, ldr(row, bv) + v);
  }
}

// Daily PET before disaggregation (get_corrected_pet, mo_meteo_handler.f90:1050-1097).
// Its per-cell parameters are read here, when a new forcing day starts, rather
// than held in registers across steps.  bc: byte offset of the cell in a
// forcing row; bv: byte offset of (cell, set) in a parameter row.
__device__ __forceinline__ const double *frow_ptr(const CellArgs &a, int v, int32_t day) {
  return a.F + ((int64_t)forcing_row(v, a.pet_case) * a.forcing_days + day) * a.sstride;
}

template <int NH>
__device__ __forceinline__ double pet_day(const CellArgs &a, const StepDev &st, uint32_t bv,
                                          uint32_t bc, double tavg, double tmax, double tmin) {
  auto P = [&](int r) { return ldr(a.P + (int64_t)r * a.ld, bv); };
  if (a.pet_case == 1)
    return P(PR_FASP) * pet_hargreaves(P(PR_HARSAM), tavg, tmax, tmin, P(PR_TANLAT),
                                       P(PR_SINLAT), P(PR_COSLAT), st.dr, st.tand, st.sind,
                                       st.cosd);
  if (a.pet_case >= 2) {
    // Priestley-Taylor and Penman-Monteith are not defined for netrad < 0
    const double rn = vmax0(ldr(frow_ptr(a, MHMHIP_F_NETRAD, st.day), bc));
    if (a.pet_case == 2) return pet_priestly(P(pr_prietay(NH) + st.month0), rn, tavg);
    const double avp = ldr(frow_ptr(a, MHMHIP_F_ABSVAPPRESS, st.day), bc) / 1000.0;
    const double ws = ldr(frow_ptr(a, MHMHIP_F_WINDSPEED, st.day), bc);
    return pet_penman(rn, tavg, avp, P(pr_aero(NH) + st.month0) / ws,
                      P(pr_surf(NH) + st.month0));
  }
  const double petin = ldr(frow_ptr(a, MHMHIP_F_PET, st.day), bc);
  if (a.pet_case == 0) return P(PR_FASP) * petin;
  return P(pr_petlai(NH) + st.month0) * petin;
}

// The step's pet, temp and prec from the daily values: the meteo weights of
// the step's month and hour when given, else the day/night factors
// (get_corrected_pet 1099-1118, get_temp 1183-1200, get_prec 1265-1282).
template <int NH>
__device__ __forceinline__ void disagg_step(const CellArgs &a, const StepDev &st, uint32_t bc,
                                            double pd, double tavg, double pre, double &pet,
                                            double &temp, double &prec) {
  if (a.hourly) {
    pet = pd;
    temp = tavg;
    prec = pre;
  } else if (a.W) {
    const int64_t wv = (int64_t)12 * 24 * a.sstride;
    const double *w = a.W + (int64_t)st.wrow * a.sstride;
    pet = disagg_weights(pd, ldr(w + 2 * wv, bc), 0.0);
    temp = disagg_weights(tavg, ldr(w + wv, bc), kT0);
    prec = disagg_weights(pre, ldr(w, bc), 0.0);
  } else {
    pet = disagg_flux(a.nT, a.rnT, pd, st.f_pet);
    temp = disagg_state_add(a.nT, tavg, st.f_temp);
    prec = disagg_flux(a.nT, a.rnT, pre, st.f_prec);
  }
}

// The step table is read-only during a launch and the same for every lane:
// read through the constant address space, it comes in with scalar loads
// into SGPRs instead of vector loads into VGPRs.
typedef const __attribute__((address_space(4))) StepDev ConstStep;
// read_step copies every field: a field added to StepDev is added there too
static_assert(sizeof(StepDev) == 9 * sizeof(double) + 7 * sizeof(int32_t) + 4,
              "StepDev changed: update read_step");
__device__ __forceinline__ StepDev read_step(const ConstStep *p) {
  StepDev st;
  st.dr = p->dr;
  st.tand = p->tand;
  st.sind = p->sind;
  st.cosd = p->cosd;
  st.f_pet = p->f_pet;
  st.f_prec = p->f_prec;
  st.f_temp = p->f_temp;
  st.evap = p->evap;
  st.revap = p->revap;
  st.day = p->day;
  st.month0 = p->month0;
  st.isday = p->isday;
  st.init_sm = p->init_sm;
  st.doy = p->doy;
  st.meteo_same = p->meteo_same;
  st.wrow = p->wrow;
  return st;
}

// L11_runoff_acc for L1 = L11 with map_flag (mo_mrm_pre_routing.f90:110-126):
// qAcc = 0; qAcc = qAcc + q * area; qAcc = qAcc * 1000 / TST   [m3 s-1]
// (rtst = RN(1/TST): the division as div_rcp, same bits)
__device__ __forceinline__ double runoff_acc(double q, double area, double tst, double rtst) {
  return div_rcp((0.0 + q * area) * 1000.0, tst, rtst);
}

// Every horizon's (sm / sat)**beta depends only on the soil moisture at the
// start of the step: the bases and exponents of these NH powers, into
// slots off .. off + NH - 1.  Bases the reference never raises to a power
// are replaced by 1 and their results left unused.
template <int NH, int N>
__device__ __forceinline__ void horizon_pow_args(const CellRegs<NH> &c, DV<N> &py, DV<N> &px,
                                                 int off) {
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const double q = div_rcp(c.sm[h], c.sat[h], c.rsat[h]);
    py.v[off + h] = c.bexp[h];
    px.v[off + h] = q > 0.0 ? q : 1.0;
  }
}

template <int NH>
__device__ __forceinline__ DV<NH> horizon_pows(const CellRegs<NH> &c) {
  DV<NH> py, px;
  horizon_pow_args<NH, NH>(c, py, px, 0);
  return exp_ylogx_v<NH>(py, px);
}

// One step of mHM() for one cell after the meteo getters (mo_mhm.f90:455-519),
// in two parts around the slow interflow's power: chain_pre (canopy, snow,
// soil moisture with the horizon powers pwh of the step's starting soil
// moisture, fast interflow) and chain_post (slow interflow with pw_slow =
// unsat**(1 + alpha), percolation, baseflow, total runoff).  The canopy's
// (interc / interc_max)**(2/3) is pow23.
// JV: Jarvis soil-moisture stress (processCase(3) = 2, 3) instead of Feddes
// The soil-moisture stress parameters of a cell (root fractions, field
// capacity, wilting point, RN(1 / (fc - wp))), passed on their own: the
// fused kernel keeps them in LDS between steps.
template <int NH>
struct Stress {
  double fr[NH], fc[NH], wp[NH], rfw[NH];
};

template <int NH>
__device__ __forceinline__ Stress<NH> stress_of(const CellRegs<NH> &c) {
  Stress<NH> t;
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    t.fr[h] = c.fr[h];
    t.fc[h] = c.fc[h];
    t.wp[h] = c.wp[h];
    t.rfw[h] = c.rfw[h];
  }
  return t;
}

template <int NH, bool JV>
__device__ __forceinline__ void chain_pre(CellRegs<NH> &c, FluxRegs<NH> &x, double maxinter,
                                          double rmaxinter, double evap, double revap,
                                          const double *pwh, const Stress<NH> &t) {
  const double ratio = canopy_fill(maxinter, rmaxinter, x.prec, c.inter, x.thr);
  canopy_evap(x.pet, maxinter, ratio, ratio > 0.0 ? pow23(ratio) : 0.0, c.inter, x.ecan);
  snow_accum_melt(c.incr, c.dmaxs, c.dnps, c.ddthr, x.prec, x.temp, c.tth, x.thr, c.pack, x.ddf,
                  x.melt, x.peff, x.rain, x.snow);
  soil_moisture<NH, JV>(c.fS, c.thrS, c.rthrS, x.pet, evap, revap, c.sat, t.fr, t.fc, t.wp, t.rfw,
                        c.c1, x.ecan, x.peff, pwh, x.rseal, c.seal, x.infil, c.sm, x.aet, x.aseal);
  runoff_fast(c.k0s, x.infil[NH - 1], c.thrU, c.unsat, x.fast);
}

template <int NH>
__device__ __forceinline__ void chain_post(CellRegs<NH> &c, FluxRegs<NH> &x, double pw_slow) {
  runoff_slow_perc(c.k1s, c.kps, c.karst, pw_slow, c.satS, c.unsat, x.slow, x.perc);
  runoff_sat_zone(c.k2s, c.satS, x.bflow);
  x.total = total_runoff(c.fS, x.fast, x.slow, x.bflow, x.rseal);
}

template <int NH, bool JV>
__device__ __forceinline__ void chain_step(const CellArgs &a, CellRegs<NH> &c, FluxRegs<NH> &x,
                                           double maxinter, double rmaxinter, double evap,
                                           double revap) {
  const DV<NH> pwh = horizon_pows<NH>(c);
  chain_pre<NH, JV>(c, x, maxinter, rmaxinter, evap, revap, pwh.v, stress_of<NH>(c));
  chain_post<NH>(c, x, c.unsat > 0.0 ? exp_ylogx(1.0 + c.alpha, c.unsat) : 0.0);
}

// ---------------------------------------------------------------------------
// Plain path: one launch per model step; every state and flux is written.
// The cell's loads are issued before the log / exp tables are filled, so the
// fill's reads and barrier wait behind them rather than in front.
template <int NH, bool JV, bool OUT>
__global__ __launch_bounds__(256) void cell_step_kernel(CellArgs a, StepDev st, int64_t t) {
  const uint32_t c0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = c0 < a.n_cells;
  const uint32_t c = act ? c0 : 0u;  // an idle lane loads cell 0 and leaves after the fill
  const uint32_t bc = c * 8u;
  const uint32_t bv = (uint32_t)(blockIdx.y * a.sstride + c) * 8u;
  CellRegs<NH> r;
  FluxRegs<NH> x;
  load_params<NH, JV>(a, bv, r);
  load_states<NH>(a, bv, r);
  const double pre = ldr(frow_ptr(a, MHMHIP_F_PRE, st.day), bc);
  const double tavg = ldr(frow_ptr(a, MHMHIP_F_TAVG, st.day), bc);
  const double tmax = a.pet_case == 1 ? ldr(frow_ptr(a, MHMHIP_F_TMAX, st.day), bc) : 0.0;
  const double tmin = a.pet_case == 1 ? ldr(frow_ptr(a, MHMHIP_F_TMIN, st.day), bc) : 0.0;
  mathtab_fill();  // the chain's log and exp read it (every thread, before any returns)
  if (!act) return;
  const double pd = pet_day<NH>(a, st, bv, bc, tavg, tmax, tmin);
  disagg_step<NH>(a, st, bc, pd, tavg, pre, x.pet, x.temp, x.prec);
  if (st.init_sm) {
#pragma unroll
    for (int h = 0; h < NH; ++h) r.sm[h] = 0.5 * r.fc[h];
  }
  const double maxinter = ldr(a.P + (int64_t)(pr_maxinter(NH) + st.month0) * a.ld, bv);
  chain_step<NH, JV>(a, r, x, maxinter, 1.0 / maxinter, st.evap, st.revap);
  store_states<NH>(a, bv, r);
  store_fluxes<NH>(a, bv, x);
  if (OUT) add_outputs<NH>(a, bv, r, x);
  if (a.ring) {
    const double q = a.raw_total ? x.total : runoff_acc(x.total, ldr(a.area, bc), a.tst, a.rtst);
    const int32_t g = a.lag ? a.lag[c] : 0;
    for (int32_t i = 0; i < a.ring_rep; ++i)
      a.ring[ring_row(t * a.ring_rep + i, g, a.ring_len) * a.ld + bv / 8u] = q;
  }
}

// Per-lane LDS slots of the fused kernel (dynamic shared memory,
// lane_slots(NH) doubles per thread, laid out lane after lane; an odd count
// puts the 64-bit accesses of a wave's lanes on distinct bank pairs): the
// values that change once per forcing day, month or day/night switch, and
// the soil-moisture stress parameters read once per step -- in LDS rather
// than in registers across the whole time loop.
enum LaneSlot {
  LS_PRE = 0, LS_TAVG, LS_PD, LS_MAXI, LS_RMAXI, LS_XPET, LS_XTEMP, LS_XPREC, LS_STRESS
};
__host__ __device__ constexpr int lane_slots(int nh) { return (LS_STRESS + 4 * nh) | 1; }
extern __shared__ double s_lane[];

// ---------------------------------------------------------------------------
// Fused path: nsteps model steps per launch with parameters and states held
// in registers.  Forcing and daily PET are read/evaluated once per forcing day
// (the reference recomputes the same value every hour), maximum interception
// once per month.  Only the routing input leaves the lane each step; states
// are written at the end and the last step's fluxes as soon as each is known.
// Per-step fluxes are scoped to the loop body so they do not stay live across
// steps.
template <int NH, bool JV, bool OUT>
#ifndef MHMHIP_FUSED_WAVES
#define MHMHIP_FUSED_WAVES 2
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MHMHIP_FUSED_WAVES)))
void cell_fused_kernel(CellArgs a, int64_t t0, int32_t nsteps,
                                                         int32_t step0) {
  mathtab_fill();  // the chain's log and exp read it (before any thread returns)
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  const uint32_t bc = c * 8u;
  const uint32_t bv = (uint32_t)(blockIdx.y * a.sstride + c) * 8u;
  CellRegs<NH> r;
  load_params<NH, JV>(a, bv, r);
  load_states<NH>(a, bv, r);
  // this lane's LDS slots (lane_slots(NH) doubles, laid out lane by lane)
  const uint32_t lane0 = threadIdx.x * (uint32_t)lane_slots(NH);
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    s_lane[lane0 + LS_STRESS + h] = r.fr[h];
    s_lane[lane0 + LS_STRESS + NH + h] = r.fc[h];
    s_lane[lane0 + LS_STRESS + 2 * NH + h] = r.wp[h];
    s_lane[lane0 + LS_STRESS + 3 * NH + h] = r.rfw[h];
  }
  const double area = a.ring ? ldr(a.area, bc) : 0.0;
  int cached_day = -1, cached_doy = -1, cached_month = -1;
  const ConstStep *steps = (const ConstStep *)(a.steps + step0);
  // ring: the lane's row offset (its lag) is fixed, the step's row uniform
  double *rlane = nullptr;
  int64_t roff = 0;
  if (a.ring) {
    rlane = a.ring + (int64_t)(a.lag ? a.lag[c] : 0) * a.ld + bv / 8u;
    roff = t0 * a.ring_rep % a.ring_len * a.ld;
  }
  // The horizon powers of a step depend only on its starting soil moisture,
  // known when the step before reaches its slow interflow: one call
  // evaluates that step's unsat**(1 + alpha) and the next step's NH horizon
  // powers side by side (independent chains for the scheduler); the first
  // step's come from here, a step that resets the soil moisture recomputes.
  DV<NH> pwh = horizon_pows<NH>(r);
#pragma nounroll
  for (int32_t s = 0; s < nsteps; ++s) {
    const StepDev st = read_step(steps + s);
    // the slot index opaque each step: the slots are read where they are
    // used, not held in registers across the loop
    const uint32_t li = opaque(lane0);
    if (st.day != cached_day || st.doy != cached_doy) {
      // offsets opaque here: no per-lane address of these rows is held
      // across the time loop
      const uint32_t oc = opaque(bc), ov = opaque(bv);
      const double pre = ldr(frow_ptr(a, MHMHIP_F_PRE, st.day), oc);
      const double tavg = ldr(frow_ptr(a, MHMHIP_F_TAVG, st.day), oc);
      const double tmax =
          a.pet_case == 1 ? ldr(frow_ptr(a, MHMHIP_F_TMAX, st.day), oc) : 0.0;
      const double tmin =
          a.pet_case == 1 ? ldr(frow_ptr(a, MHMHIP_F_TMIN, st.day), oc) : 0.0;
      s_lane[li + LS_PRE] = pre;
      s_lane[li + LS_TAVG] = tavg;
      s_lane[li + LS_PD] = pet_day<NH>(a, st, ov, oc, tavg, tmax, tmin);
      cached_day = st.day;
      cached_doy = st.doy;
    }
    if (st.month0 != cached_month) {
      const double maxinter = ldr(a.P + (int64_t)(pr_maxinter(NH) + st.month0) * a.ld,
                                  opaque(bv));
      s_lane[li + LS_MAXI] = maxinter;
      s_lane[li + LS_RMAXI] = 1.0 / maxinter;
      cached_month = st.month0;
    }
    FluxRegs<NH> x;
    if (s == 0 || !st.meteo_same) {
      // day or night factors changed (or a new day): disaggregate again
      const double pd = s_lane[li + LS_PD], tavg = s_lane[li + LS_TAVG];
      const double pre = s_lane[li + LS_PRE];
      disagg_step<NH>(a, st, opaque(bc), pd, tavg, pre, x.pet, x.temp, x.prec);
      s_lane[li + LS_XPET] = x.pet;
      s_lane[li + LS_XTEMP] = x.temp;
      s_lane[li + LS_XPREC] = x.prec;
    } else {
      x.pet = s_lane[li + LS_XPET];
      x.temp = s_lane[li + LS_XTEMP];
      x.prec = s_lane[li + LS_XPREC];
    }
    Stress<NH> t;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      t.fr[h] = s_lane[li + LS_STRESS + h];
      t.fc[h] = s_lane[li + LS_STRESS + NH + h];
      t.wp[h] = s_lane[li + LS_STRESS + 2 * NH + h];
      t.rfw[h] = s_lane[li + LS_STRESS + 3 * NH + h];
    }
    if (st.init_sm) {
#pragma unroll
      for (int h = 0; h < NH; ++h) r.sm[h] = 0.5 * t.fc[h];
      pwh = horizon_pows<NH>(r);
    }
    chain_pre<NH, JV>(r, x, s_lane[li + LS_MAXI], s_lane[li + LS_RMAXI], st.evap, st.revap,
                      pwh.v, t);
    // the last step's fluxes leave as soon as they are known, so that they
    // are not held across the powers
    const bool last = s == nsteps - 1;
    if (last && !OUT) store_fluxes_pre<NH>(a, bv, x);
    DV<NH + 1> py, px;
    py.v[0] = 1.0 + r.alpha;
    px.v[0] = r.unsat > 0.0 ? r.unsat : 1.0;
    horizon_pow_args<NH, NH + 1>(r, py, px, 1);
    const DV<NH + 1> pw = exp_ylogx_v<NH + 1>(py, px);
    chain_post<NH>(r, x, r.unsat > 0.0 ? pw.v[0] : 0.0);
#pragma unroll
    for (int h = 0; h < NH; ++h) pwh.v[h] = pw.v[h + 1];
    if (a.ring) {
      // the step's qOUT for each of its routing sub-steps (rows in routing steps)
      const double q = a.raw_total ? x.total : runoff_acc(x.total, area, a.tst, a.rtst);
      for (int32_t i = 0; i < a.ring_rep; ++i) {
        rlane[roff] = q;
        roff += a.ld;
        if (roff == a.ring_len * a.ld) roff = 0;
      }
    }
    if (last && OUT) store_fluxes_pre<NH>(a, bv, x);
    if (last) store_fluxes_post<NH>(a, bv, x);
    if (OUT) add_outputs<NH>(a, bv, r, x);
  }
  store_states<NH>(a, bv, r);
}

// ---------------------------------------------------------------------------
// One node, one step of L11_runoff_acc + add_inflow + L11_routing
// (mo_mrm_pre_routing.f90:77-143,179-214; mo_mrm_routing.f90:425-472).
__device__ __forceinline__ void route_node(const RouteArgs &r, uint32_t p, int32_t set, int64_t t,
                                           NodeMeta m) {
  const int64_t ld = (int64_t)r.n_sets * r.sstride;
  const uint32_t bv = (uint32_t)(set * r.sstride + p) * 8u;
  const int32_t y = m.packed;
  // qOUT of this node and step (L11_runoff_acc, done by the cell kernel)
  double qout = ldr(r.ring + (t % r.ring_len) * ld, bv);
  if (y & NF_INFLOW) {
    for (int32_t g = 0; g < r.n_inflow; ++g) {
      if (r.inflow_node[g] == (int32_t)p) {
        const double qi = r.inflow_q[(t - r.t_base) / r.rep * r.n_inflow + g];
        qout = r.inflow_hw[g] ? qout + qi : qi;
      }
    }
  }
  if (r.qOUT && t == r.t_hi - 1) str(r.qOUT, bv, qout);
  // qTIN(iNode,IT): upstream qTR in routing rank, then own qOUT
  const double *pub_t = r.pub + (t & 1) * ld + (int64_t)set * r.sstride;
  const int32_t *up = r.child_list + m.child;
  double s = 0.0;
  const int nc = nm_nchild(y);
  for (int c = 0; c < nc; ++c) s = s + pub_t[up[c]];
  if (y & NF_QOUT) s = s + qout;
  if (y & NF_LINK) {
    const double qtin1 = ldr(r.qTIN, bv);
    const double qtr1 = ldr(r.pub + ((t - 1) & 1) * ld, bv);
    double qtr = qtr1 + ldr(r.C1, bv) * (qtin1 - qtr1) + ldr(r.C2, bv) * (s - qtin1);
    if (y & NF_CUT) qtr = 0.0;
    str(r.pub + (t & 1) * ld, bv, qtr);
  }
  str(r.qTIN, bv, s);
  if (r.qacc && t >= r.acc_from) str(r.qacc, bv, ldr(r.qacc, bv) + s);
  if (y & NF_GAUGE) {
    for (int32_t g = 0; g < r.n_gauges; ++g)
      if (r.gauge_node[g] == (int32_t)p)
        r.gauge_out[((t - r.t_base) * r.n_sets + set) * r.n_gauges + g] = s;
  }
}

__global__ __launch_bounds__(256) void route_diag_kernel(RouteArgs r, int64_t d, uint32_t p_lo,
                                                         uint32_t p_hi) {
  const uint32_t p = p_lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p_hi) return;
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
  const uint32_t lo = (uint32_t)dom_start[dom], hi = (uint32_t)dom_start[dom + 1];
  const int32_t dm = dom_dmax[dom];
  const int64_t d_end = drain ? d_lo + dm : d_hi;
  for (int64_t d = d_lo; d < d_end; ++d) {
    for (uint32_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
      const NodeMeta m = r.meta[p];
      const int64_t t = d - (int64_t)(dm - nm_depth(m.packed));
      if (t >= r.t_lo && t < r.t_hi) route_node(r, p, set, t, m);
    }
    __syncthreads();
  }
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// operations, not for its global stores (no other lane of the workgroup
// reads those before the kernel ends).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------------------
// Tiled routing: launch L handles, for every tile of depth tau, the window
// k = L - (tau_max - tau) of B steps.  Tile-root outflow of window k is
// written to the boundary buffer (parity k & 1) in launch L and read by the
// downstream tile's proxy slot in launch L + 1.  Inside the window the
// workgroup walks its own (depth, step) diagonals: slot with lag g handles
// window step j on diagonal j + g, upstream slots (lag g - 1) produced step j
// on the diagonal before, and a two-entry LDS parity buffer keeps it
// readable while they write step j + 1.  Each slot streams its own series --
// qOUT rows of the ring (real slots, inflow gauges already applied by
// inflow_patch_kernel) or the upstream tile root's boundary series (proxies)
// -- from memory, kTileAhead diagonals ahead of use, so LDS holds only the
// parity buffers and the window length is not bounded by it.
// The host keeps a window's ring rows contiguous (ring length and window
// starts are multiples of B).
// Same per-node arithmetic as route_node (mo_mrm_routing.f90:425-472).
constexpr int kTileAhead = 8;

struct TileLane {
  const double *p;      // series element of diagonal 0 (real) or window step 0 (proxy)
  uint32_t stride;      // bytes between consecutive elements of the series (0: idle)
  double *bout;         // tile root: boundary series of this window, else null
  int64_t t0;           // global step of window step 0
  int32_t lag, nb, nc, c0;
  uint32_t bv;          // byte offset of (set, node) in a row
  int32_t a, pos, set;
  int32_t gauge;        // first gauge at the node, -1: none
  int32_t orow_t, orow_g;  // LDS output rows of the tile-root series / gauge sums, -1: none
  bool act, real, proxy;
  double qtin1, qtr1, c1, c2;
  double qlast;         // qOUT of the run's last step (when it falls in this window)
  double qacc;          // Qmod summed over the steps (ACC)
};

// Output row of the tiled kernel's LDS staging: where its B values go.
struct TileOutRow {
  double *base;
  int32_t stride;  // elements between consecutive steps
  int32_t nb;      // steps to write
};

// One sweep over the bin's nd diagonals; NCW bounds the upstream slots summed
// (the wave's largest upstream count rounded up to 1, 3, 5 or 8 -- a node's
// slots past its own count read a +0.0 slot, which leaves the running sum's
// value unchanged).  LDS: qTR parity buffers [2][W], +0.0 at 2W and 3W, so
// that buffer 1's offsets are buffer 0's plus W for upstream and zero slots
// alike.  STAGE: the tile-root series and gauge sums go to LDS rows (out,
// B values each) and leave after the sweep, so that the loop holds no global
// store: stores count in the same counter as the loads in flight, and each
// one would shorten the wait for a fetched value.
template <int NCW, bool STAGE, bool ACC>
__device__ __forceinline__ void tile_sweep(const RouteArgs &r, TileLane &u, double *qtr, int nd,
                                           double *out, int B, int out_cap) {
  constexpr int W = kTileSlots;
  const int s = threadIdx.x;
  uint32_t up[NCW];  // byte offsets in buffer 0
  const uint32_t zoff = (uint32_t)(2 * W) * 8u;
#pragma unroll
  for (int c = 0; c < NCW; ++c) up[c] = c < u.nc ? (uint32_t)(u.c0 + c) * 8u : zoff;
  const bool link = (u.a & NF_LINK) != 0, qflag = (u.a & NF_QOUT) != 0, cut = (u.a & NF_CUT) != 0;
  const bool rare = u.bout != nullptr || u.gauge >= 0;
  const char *qtr_b = reinterpret_cast<const char *>(qtr);
  // Value for diagonal df: a real slot's qOUT sits in ring row df of the
  // window (the ring is written lag rows late), the same row for every real
  // slot of the bin; a proxy's boundary element is its window step df - lag,
  // clamped into the window.  The address is formed from the index each time:
  // written as a conditional advance, the compiler reuses the previous
  // element where the clamp holds it, and each fetch then waits for the one
  // before.
  // (a real slot's index is df itself: no shift, no clamp below nd)
  const int32_t fshift = u.real ? 0 : u.lag;
  const int32_t fmax = u.real ? 0x7FFFFFFF : (u.nb > 0 ? u.nb - 1 : 0);
  auto fetch = [&](int df) -> double {
    const uint32_t jc = (uint32_t)min(max(df - fshift, 0), fmax);
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(u.p) +
                                             (uint64_t)jc * u.stride);
  };
  // steps this slot updates (real and due), and the run's last step in the window
  const uint32_t nupd = u.act && u.real ? (uint32_t)u.nb : 0u;
  const int32_t jlast = (int32_t)min<int64_t>(r.t_hi - 1 - u.t0, 0x7FFFFFFF);
  double qq[kTileAhead];
#pragma unroll
  for (int i = 0; i < kTileAhead; ++i) qq[i] = 0.0;
  double musk = u.qtr1 + u.c1 * (u.qtin1 - u.qtr1);
  // STAGE: a diagonal's outputs are written to LDS on the next diagonal, after
  // its upstream reads are issued, not before the barrier that ends it.
  // Every lane writes both, without a branch: a lane with nothing to keep
  // writes its own scratch entries (after the rows).  Rows are B + 1 apart:
  // the lanes of a wave write different banks.
  double pq = 0.0, ps = 0.0;
  const int scratch_t = out_cap * (B + 1) + s;
  const int scratch_g = scratch_t + W;
  int it = scratch_t, ig = scratch_g;
  auto put_rows = [&]() {
    out[it] = pq;
    out[ig] = ps;
  };
  // Diagonal d reads parity buffer (d + 1) & 1, written on diagonal d - 1 by
  // the upstream slots (one level deeper, one diagonal earlier, same step),
  // and writes buffer d & 1.  A slot writes every diagonal: an inactive
  // slot's value is read only by a parent that is inactive for the same step.
  auto diagonal = [&](int d, int wb, double qo) {
    const int j = d - u.lag;
    const char *rb_b = qtr_b + (wb ^ 1) * W * 8;  // the buffer written on diagonal d - 1
    double v[NCW];
#pragma unroll
    for (int c = 0; c < NCW; ++c) v[c] = *reinterpret_cast<const double *>(rb_b + up[c]);
    if (STAGE) put_rows();
    // upstream qTR in routing rank
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < NCW; ++c) sum = sum + v[c];
    if (NCW == 8 && u.nc > 8) {
      const double *rb = qtr + (wb ^ 1) * W;
      for (int c = 8; c < u.nc; ++c) sum = sum + rb[u.c0 + c];
    }
    if (qflag) sum = sum + qo;
    // musk = qTR(IT1) + C1 (qTIN(IT1) - qTR(IT1)), formed on the diagonal before
    double q = musk + u.c2 * (sum - u.qtin1);
    if (cut) q = 0.0;
    qtr[wb * W + s] = u.proxy ? qo : q;
    const bool upd = (uint32_t)j < nupd;
    if (upd && link) u.qtr1 = q;
    if (upd) u.qtin1 = sum;
    // Qmod of step j, added in step order
    if (ACC && upd && u.t0 + j >= r.acc_from) u.qacc = u.qacc + sum;
    musk = u.qtr1 + u.c1 * (u.qtin1 - u.qtr1);
    if (upd && j == jlast) u.qlast = qo;
    if (STAGE) {
      it = upd && u.orow_t >= 0 ? u.orow_t * (B + 1) + j : scratch_t;
      ig = upd && u.orow_g >= 0 ? u.orow_g * (B + 1) + j : scratch_g;
      pq = q;
      ps = sum;
    } else if (upd && rare) {
      // at most two stores, none in a loop
      if (u.bout) u.bout[j] = q;
      if (u.gauge >= 0)
        r.gauge_out[((u.t0 + j - r.t_base) * r.n_sets + u.set) * r.n_gauges + u.gauge] = sum;
    }
  };
  // The first pass only fetches.  Fetches stay in diagonal order, one per
  // diagonal, in every pass, so that a value is waited for kTileAhead - 1
  // fetches after its own (a prologue whose loads the compiler reorders would
  // make the loop wait for all of them).
  for (int d0 = -kTileAhead; d0 < nd; d0 += kTileAhead) {
#pragma unroll
    for (int i = 0; i < kTileAhead; ++i) {
      const int d = d0 + i;
      if (d >= 0 && d < nd) {
        diagonal(d, i & 1, qq[i]);
        lds_barrier();
      }
      qq[i] = fetch(min(d + kTileAhead, nd - 1));
    }
  }
  if (STAGE) {
    put_rows();
    lds_barrier();
  }
}

// LDS: qTR parity buffers [2][W], +0.0 at 2W and 3W, then (STAGE) the bin's
// output rows: g.out_rows TileOutRow, then g.out_rows x (B + 1) values and
// 2 W scratch values.
// ACC: Qmod summed per node (r.qacc).
template <bool STAGE, bool ACC>
__global__ __launch_bounds__(kTileSlots) void route_tile_kernel(RouteArgs r, TileArgs g,
                                                                int32_t L, int32_t bin0) {
  extern __shared__ double lds[];
  constexpr int W = kTileSlots;
  const int bin = bin0 + blockIdx.x, set = blockIdx.y, s = threadIdx.x;
  const BinInfo bi = g.bin[bin];
  // whole bin idle in this launch: every thread leaves before any barrier
  if (L - (g.tau_max - bi.tau_hi) < 0 || L - (g.tau_max - bi.tau_lo) >= g.nwin) return;
  const SlotMeta m = g.slot[(int64_t)bin * W + s];
  TileLane u;
  const int32_t k = L - (g.tau_max - (m.b >> 16));
  u.a = m.a;
  u.pos = m.pos;
  u.set = set;
  u.real = (m.a & SF_REAL) != 0;
  u.proxy = (m.a & SF_PROXY) != 0;
  u.act = (u.real || u.proxy) && k >= 0 && k < g.nwin;
  u.t0 = r.t_lo + (int64_t)k * g.B;
  u.nb = u.act ? (int)min((int64_t)g.B, r.t_hi - u.t0) : 0;
  u.lag = m.a & 0xFFF;
  u.nc = (m.a >> 12) & 0xF;
  u.c0 = m.b & 0xFFFF;
  const int64_t ld = (int64_t)r.n_sets * r.sstride;
  u.bv = u.real ? (uint32_t)(set * r.sstride + m.pos) * 8u : 0u;
  u.qtin1 = u.qtr1 = u.c1 = u.c2 = 0.0;
  // series of the slot: ring rows t0 mod R + d, d < nd (t0 and R are
  // multiples of B, and the ring has kTileHeight rows past R) or the
  // boundary series of the proxy's tile root; idle slots re-read one valid
  // element
  const int64_t bseries = ((int64_t)(k & 1) * r.n_sets + set) * g.n_bnd;
  if (u.act && u.real) {
    u.p = r.ring + ring_row(u.t0, 0, r.ring_len) * ld + (int64_t)set * r.sstride + m.pos;
    u.stride = (uint32_t)(ld * 8);
    u.qtin1 = ldr(r.qTIN, u.bv);
    if (m.a & NF_LINK) {
      u.qtr1 = ldr(r.pub + ((u.t0 - 1) & 1) * ld, u.bv);
      u.c1 = ldr(r.C1, u.bv);
      u.c2 = ldr(r.C2, u.bv);
    }
  } else if (u.act) {
    u.p = g.bnd + (bseries + m.bnd) * g.bstride;
    u.stride = 8;
  } else {
    u.p = g.bnd;
    u.stride = 0;
  }
  u.bout = (u.act && (m.a & SF_TROOT)) ? g.bnd + (bseries + m.bnd) * g.bstride : nullptr;
  // gauges: the node's first gauge only (the host copies it to the others)
  u.gauge = (u.act && u.real && (m.a & NF_GAUGE)) ? g.node_gauge[m.pos] : -1;
  u.orow_t = u.orow_g = -1;
  u.qlast = 0.0;
  u.qacc = ACC && u.act && u.real ? ldr(r.qacc, u.bv) : 0.0;
  double *qtr = lds;
  TileOutRow *rows = reinterpret_cast<TileOutRow *>(lds + 3 * W + 1);
  double *out = reinterpret_cast<double *>(rows + g.out_rows);
  if (STAGE) {
    // the slot's output rows (host-numbered inside the bin); an idle owner
    // leaves its row empty
    const int32_t so = g.slot_out[(int64_t)bin * W + s];
    const int32_t rt = so & 0xFFFF, rg = (so >> 16) & 0xFFFF;
    if (rt != 0xFFFF) {
      u.orow_t = rt;
      rows[rt] = TileOutRow{u.bout ? u.bout : g.bnd, 1, u.bout ? u.nb : 0};
    }
    if (rg != 0xFFFF) {
      u.orow_g = rg;
      const int64_t gi = ((u.t0 - r.t_base) * r.n_sets + set) * r.n_gauges + u.gauge;
      rows[rg] = TileOutRow{u.gauge >= 0 ? r.gauge_out + gi : g.bnd, r.n_sets * r.n_gauges,
                            u.gauge >= 0 ? u.nb : 0};
    }
  }
  if (s == 0) qtr[2 * W] = qtr[3 * W] = 0.0;
  lds_barrier();
  const int nd = bi.height + g.B - 1;
  if (__any(u.nc > 5)) tile_sweep<8, STAGE, ACC>(r, u, qtr, nd, out, g.B, g.out_rows);
  else if (__any(u.nc > 3)) tile_sweep<5, STAGE, ACC>(r, u, qtr, nd, out, g.B, g.out_rows);
  else if (__any(u.nc > 1)) tile_sweep<3, STAGE, ACC>(r, u, qtr, nd, out, g.B, g.out_rows);
  else tile_sweep<1, STAGE, ACC>(r, u, qtr, nd, out, g.B, g.out_rows);
  if (STAGE) {
    // the sweep ended with a barrier after the last rows: they are complete
    for (int idx = s; idx < bi.out_rows * g.B; idx += W) {
      const int rr = idx / g.B, j = idx - rr * g.B;
      const TileOutRow o = rows[rr];
      if (j < o.nb) o.base[(int64_t)j * o.stride] = out[rr * (g.B + 1) + j];
    }
  }
  if (u.act && u.real) {
    str(r.qTIN, u.bv, u.qtin1);
    if (m.a & NF_LINK) str(r.pub + ((u.t0 + u.nb - 1) & 1) * ld, u.bv, u.qtr1);
    if (r.qOUT && r.t_hi - 1 - u.t0 < u.nb) str(r.qOUT, u.bv, u.qlast);
    if (ACC) str(r.qacc, u.bv, u.qacc);
  }
}

// Forcing on the meteo grid (L2) to the L1 cells of one domain, as
// meteo_forcings_wrapper maps it (mo_meteo_helper.f90:119-143).  One thread
// per (L1 cell, day).  Aggregation (L1 coarser, mo_meteo_spatial_tools.f90:
// 94-200): the valid L2 cells of the cell's block added column by column (j
// outer, i inner, the reference's order, from 0.0) and divided by their
// count (0/0 gives NaN, as there).  Disaggregation (313-377) and equal grids
// (mo_meteo_helper.f90:129-131): the source L2 cell, or nodata (-9999).
__global__ __launch_bounds__(256) void meteo_l2_kernel(const double *data2, const uint8_t *mask2,
                                                       int32_t nr2, int64_t n2,
                                                       const MeteoL2Cell *cells, int64_t ncells,
                                                       int32_t ndays, double *F,
                                                       int64_t day_stride, int32_t aggregate) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const MeteoL2Cell m = cells[c];
  for (int32_t t = blockIdx.y; t < ndays; t += gridDim.y) {
    const double *d2 = data2 + (int64_t)t * n2;
    double v;
    if (aggregate) {
      double sum = 0.0;
      for (int32_t j = m.c0; j <= m.c1; ++j)
        for (int32_t i = m.r0; i <= m.r1; ++i) {
          const int64_t k = (int64_t)j * nr2 + i;
          if (mask2[k]) sum = sum + d2[k];
        }
      v = sum / (double)m.n;
    } else {
      v = m.src >= 0 ? d2[m.src] : -9999.0;
    }
    F[(int64_t)t * day_stride + m.dst] = v;
  }
}

hipError_t launch_meteo_l2(const double *data2, const uint8_t *mask2, int32_t nr2, int64_t n2,
                           const MeteoL2Cell *cells, int64_t ncells, int32_t ndays, double *F,
                           int64_t day_stride, bool aggregate, hipStream_t s) {
  if (ncells <= 0 || ndays <= 0) return hipSuccess;
  const dim3 grid((unsigned)((ncells + 255) / 256), (unsigned)(ndays < 65535 ? ndays : 65535));
  hipLaunchKernelGGL(meteo_l2_kernel, grid, dim3(256), 0, s, data2, mask2, nr2, n2, cells, ncells,
                     ndays, F, day_stride, aggregate ? 1 : 0);
  return hipGetLastError();
}

// Streamed forcing: a chunk of days of the variables in var_mask (bit v:
// forcing variable v), uploaded in reference cell order (raw[k][j][cell],
// k counting the variables given), scattered into the forcing columns in
// device order: F[v][col(j)][pos[cell]], col(j) = (day0 + j) mod n_cols.  One
// thread per (cell, day of the chunk): the reads are coalesced, the writes
// follow the tile layout (cells of a tile close together).
__global__ __launch_bounds__(256) void forcing_gather_kernel(const double *raw, FRowMap m,
                                                             int32_t ndays, int64_t n_cells,
                                                             const int32_t *pos, double *F,
                                                             int32_t n_cols, int32_t day0,
                                                             int64_t sstride) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t j = blockIdx.y;
  if (c >= n_cells || j >= ndays) return;
  const int64_t p = pos[c];
  const int64_t col = (int64_t)((day0 + j) % n_cols);
  for (int32_t k = 0; k < m.n; ++k)
    F[((int64_t)m.row[k] * n_cols + col) * sstride + p] = raw[((int64_t)k * ndays + j) * n_cells + c];
}

hipError_t launch_forcing_gather(const double *raw, const FRowMap &m, int32_t ndays,
                                 int64_t n_cells, const int32_t *pos, double *F, int32_t n_cols,
                                 int32_t day0, int64_t sstride, hipStream_t s) {
  if (m.n < 1 || m.n > MHMHIP_F_COUNT || ndays < 1 || ndays > 65535 || n_cells < 1 ||
      n_cols < 1 || day0 < 0)
    return hipErrorInvalidValue;
  for (int32_t k = 0; k < m.n; ++k)
    if (m.row[k] < 0 || m.row[k] >= MHMHIP_F_COUNT) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n_cells + 255) / 256), (unsigned)ndays);
  hipLaunchKernelGGL(forcing_gather_kernel, grid, dim3(256), 0, s, raw, m, ndays, n_cells, pos, F,
                     n_cols, day0, sstride);
  return hipGetLastError();
}

// L11_runoff_acc with map_flag on a routing grid coarser than the L1 grid
// (mo_mrm_pre_routing.f90:110-126): qAcc(node) = 0, plus qAll(k) * area(k)
// for the node's L1 cells k in increasing reference order, then
// * 1000 / TST.  The cells of device node p are device cells
// [cstart[p], cstart[p + 1]) (grouped by node, reference order inside).  tot
// holds tot_rows rows of L1_total_runoff (row (t0 + j) mod tot_rows for step
// t0 + j); the step's qOUT goes to its rep ring rows at the node's lag, as the
// cell kernels write them on an L1 = L11 grid.  One thread per (node, set),
// blockIdx.z the step.
__global__ __launch_bounds__(256) void runoff_agg_kernel(
    const double *tot, int64_t tot_rows, const int32_t *cstart, const double *area,
    int64_t n_nodes, int64_t sstride, int64_t ld, double *ring, const uint8_t *lag,
    int64_t ring_len, int32_t rep, int64_t t0, double tst, double rtst) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_nodes) return;
  const int64_t set0 = (int64_t)blockIdx.y * sstride;
  const int64_t t = t0 + blockIdx.z;
  const double *src = tot + (t % tot_rows) * ld + set0;
  double sum = 0.0;
  for (int32_t c = cstart[p]; c < cstart[p + 1]; ++c) sum = sum + src[c] * area[c];
  const double q = div_rcp(sum * 1000.0, tst, rtst);
  const int32_t g = lag ? lag[p] : 0;
  for (int32_t i = 0; i < rep; ++i) ring[ring_row(t * rep + i, g, ring_len) * ld + set0 + p] = q;
}

hipError_t launch_runoff_agg(const double *tot, int64_t tot_rows, const int32_t *cstart,
                             const double *area, int64_t n_nodes, int64_t sstride, int64_t ld,
                             int n_sets, double *ring, const uint8_t *lag, int64_t ring_len,
                             int32_t rep, int64_t t0, int32_t nsteps, double tst, hipStream_t s) {
  if (n_nodes < 1 || n_sets < 1 || nsteps < 1 || nsteps > 65535 || tot_rows < 1 || ring_len < 1 ||
      rep < 1 || !cstart || !tot || !ring || !(tst > 0.0))
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n_nodes + 255) / 256), (unsigned)n_sets, (unsigned)nsteps);
  hipLaunchKernelGGL(runoff_agg_kernel, grid, dim3(256), 0, s, tot, tot_rows, cstart, area,
                     n_nodes, sstride, ld, ring, lag, ring_len, rep, t0, tst, 1.0 / tst);
  return hipGetLastError();
}

// Routing steps longer than the model step (processCase(8) = 2, 3 with
// tsRoutFactor >= 1, mo_mhm_interface_run.f90:494-512): every model step adds
// its L1_total_runoff to RunToRout [mm] (acc = acc + q from 0 after each
// routing); on a routing step L11_runoff_acc converts the sum with the routing
// step TS_rout (mo_mrm_pre_routing.f90:110-126: qAcc = 0 + RunToRout * area,
// then * 1000 / (HourSecs * TS_rout)) into the step's ring row (lag, as the
// cell kernels write it) and RunToRout starts again from 0.  One thread per
// (cell, set).
__global__ __launch_bounds__(256) void runoff_sum_kernel(const double *total, double *acc,
                                                         double *ring, const double *area,
                                                         const uint8_t *lag, int64_t n_cells,
                                                         int64_t sstride, int64_t ld, double tst,
                                                         int32_t route) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const int64_t i = (int64_t)blockIdx.y * sstride + c;
  const double sum = acc[i] + total[i];
  if (route) {
    ring[(int64_t)lag[c] * ld + i] = ((0.0 + sum * area[c]) * 1000.0) / tst;
    acc[i] = 0.0;
  } else {
    acc[i] = sum;
  }
}

hipError_t launch_runoff_sum(const double *total, double *acc, double *ring, const double *area,
                             const uint8_t *lag, int64_t n_cells, int64_t sstride, int64_t ld,
                             int n_sets, double tst, bool route, hipStream_t s) {
  if (n_cells < 1 || n_sets < 1 || !lag) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n_cells + 255) / 256), (unsigned)n_sets);
  hipLaunchKernelGGL(runoff_sum_kernel, grid, dim3(256), 0, s, total, acc, ring, area, lag,
                     n_cells, sstride, ld, tst, route ? 1 : 0);
  return hipGetLastError();
}

// Inflow gauges for the tiled routing (add_inflow, mo_mrm_pre_routing.f90:
// 179-214, as route_node applies them): the qOUT of an inflow node, in the
// ring rows of steps [t_lo, t_hi), is replaced by the gauge's discharge, or
// increased by it for a headwater gauge, before the tiles read those rows.
// One thread per (step, set) applies the gauges in their order.
__global__ __launch_bounds__(256) void inflow_patch_kernel(RouteArgs r, double *ring,
                                                           const uint8_t *lag, int64_t t_lo,
                                                           int64_t t_hi, int32_t rep) {
  const int64_t t = t_lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t set = blockIdx.y;
  if (t >= t_hi) return;
  const int64_t ld = (int64_t)r.n_sets * r.sstride;
  for (int32_t g = 0; g < r.n_inflow; ++g) {
    const int32_t p = r.inflow_node[g];
    // each routing sub-step of a model step adds that step's inflow
    const double qi = r.inflow_q[(t - r.t_base) / rep * r.n_inflow + g];
    double &q = ring[ring_row(t, lag[p], r.ring_len) * ld + (int64_t)set * r.sstride + p];
    q = r.inflow_hw[g] ? q + qi : qi;
  }
}

// ---------------------------------------------------------------------------
// host launchers
template <int NH>
static hipError_t step_nh(const CellArgs &a, const StepDev &st, int64_t t, int block, int n_sets,
                          hipStream_t s) {
  dim3 grid((unsigned)((a.n_cells + block - 1) / block), (unsigned)n_sets);
  const bool out = a.out.sum != nullptr;
  if (a.jarvis && out)
    hipLaunchKernelGGL((cell_step_kernel<NH, true, true>), grid, dim3(block), 0, s, a, st, t);
  else if (a.jarvis)
    hipLaunchKernelGGL((cell_step_kernel<NH, true, false>), grid, dim3(block), 0, s, a, st, t);
  else if (out)
    hipLaunchKernelGGL((cell_step_kernel<NH, false, true>), grid, dim3(block), 0, s, a, st, t);
  else
    hipLaunchKernelGGL((cell_step_kernel<NH, false, false>), grid, dim3(block), 0, s, a, st, t);
  return hipGetLastError();
}

template <int NH>
static hipError_t fused_nh(const CellArgs &a, int64_t t0, int nsteps, int step0, int block,
                           int n_sets, hipStream_t s) {
  // the lanes' LDS slots of a workgroup stay within 64 KB
  while ((size_t)block * lane_slots(NH) * sizeof(double) > 65536) block /= 2;
  const size_t shm = (size_t)block * lane_slots(NH) * sizeof(double);
  dim3 grid((unsigned)((a.n_cells + block - 1) / block), (unsigned)n_sets);
  const bool out = a.out.sum != nullptr;
  if (a.jarvis && out)
    hipLaunchKernelGGL((cell_fused_kernel<NH, true, true>), grid, dim3(block), shm, s, a, t0,
                       nsteps, step0);
  else if (a.jarvis)
    hipLaunchKernelGGL((cell_fused_kernel<NH, true, false>), grid, dim3(block), shm, s, a, t0,
                       nsteps, step0);
  else if (out)
    hipLaunchKernelGGL((cell_fused_kernel<NH, false, true>), grid, dim3(block), shm, s, a, t0,
                       nsteps, step0);
  else
    hipLaunchKernelGGL((cell_fused_kernel<NH, false, false>), grid, dim3(block), shm, s, a, t0,
                       nsteps, step0);
  return hipGetLastError();
}

hipError_t launch_cell_step(int nh, const CellArgs &a, const StepDev &st, int64_t t, int block,
                            int n_sets, hipStream_t s) {
  switch (nh) {
    case 1: return step_nh<1>(a, st, t, block, n_sets, s);
    case 2: return step_nh<2>(a, st, t, block, n_sets, s);
    case 3: return step_nh<3>(a, st, t, block, n_sets, s);
    case 4: return step_nh<4>(a, st, t, block, n_sets, s);
    case 5: return step_nh<5>(a, st, t, block, n_sets, s);
    case 6: return step_nh<6>(a, st, t, block, n_sets, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_cell_fused(int nh, const CellArgs &a, int64_t t0, int nsteps, int step0,
                             int block, int n_sets, hipStream_t s) {
  switch (nh) {
    case 1: return fused_nh<1>(a, t0, nsteps, step0, block, n_sets, s);
    case 2: return fused_nh<2>(a, t0, nsteps, step0, block, n_sets, s);
    case 3: return fused_nh<3>(a, t0, nsteps, step0, block, n_sets, s);
    case 4: return fused_nh<4>(a, t0, nsteps, step0, block, n_sets, s);
    case 5: return fused_nh<5>(a, t0, nsteps, step0, block, n_sets, s);
    case 6: return fused_nh<6>(a, t0, nsteps, step0, block, n_sets, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_route_diag(const RouteArgs &r, int64_t d, int64_t p_lo, int64_t p_hi,
                             hipStream_t s) {
  if (p_lo < 0 || p_hi > r.n_nodes || p_hi <= p_lo) return hipErrorInvalidValue;
  dim3 grid((unsigned)((p_hi - p_lo + 255) / 256), (unsigned)r.n_sets);
  hipLaunchKernelGGL(route_diag_kernel, grid, dim3(256), 0, s, r, d, (uint32_t)p_lo,
                     (uint32_t)p_hi);
  return hipGetLastError();
}

hipError_t launch_route_tile(const RouteArgs &r, const TileArgs &g, int64_t bin_lo, int64_t bin_hi,
                             int64_t n_bins, int32_t L, hipStream_t s) {
  // a window's rows do not wrap (ring length and window starts are multiples
  // of B; the ring has kTileHeight rows past its length), and a ring row's
  // bytes fit the kernel's 32-bit stride
  if (bin_lo < 0 || bin_hi > n_bins || bin_hi <= bin_lo || g.B < 1 || g.B > g.bstride ||
      g.bstride > kTileWindowMax || g.nwin < 1 || r.ring_len % g.B != 0 || r.t_lo % g.B != 0 ||
      (uint64_t)r.n_sets * (uint64_t)r.sstride * 8u > 0xFFFFFFFFull)
    return hipErrorInvalidValue;
  dim3 grid((unsigned)(bin_hi - bin_lo), (unsigned)r.n_sets);
  const size_t base = sizeof(double) * (3 * (size_t)kTileSlots + 1);
  const size_t stage = (size_t)g.out_rows * (sizeof(TileOutRow) + sizeof(double) * (g.B + 1)) +
                       sizeof(double) * 2 * kTileSlots;
  const bool st = base + stage <= kTileStageLds;
  const size_t shm = st ? base + stage : base;
  if (st && r.qacc)
    hipLaunchKernelGGL((route_tile_kernel<true, true>), grid, dim3(kTileSlots), shm, s, r, g, L,
                       (int32_t)bin_lo);
  else if (st)
    hipLaunchKernelGGL((route_tile_kernel<true, false>), grid, dim3(kTileSlots), shm, s, r, g, L,
                       (int32_t)bin_lo);
  else if (r.qacc)
    hipLaunchKernelGGL((route_tile_kernel<false, true>), grid, dim3(kTileSlots), shm, s, r, g, L,
                       (int32_t)bin_lo);
  else
    hipLaunchKernelGGL((route_tile_kernel<false, false>), grid, dim3(kTileSlots), shm, s, r, g, L,
                       (int32_t)bin_lo);
  return hipGetLastError();
}

hipError_t launch_inflow_patch(const RouteArgs &r, double *ring, const uint8_t *lag, int64_t t_lo,
                               int64_t t_hi, hipStream_t s, int32_t rep) {
  if (t_hi <= t_lo || t_lo < r.t_base || r.n_inflow < 1 || !lag || rep < 1)
    return hipErrorInvalidValue;
  dim3 grid((unsigned)((t_hi - t_lo + 255) / 256), (unsigned)r.n_sets);
  hipLaunchKernelGGL(inflow_patch_kernel, grid, dim3(256), 0, s, r, ring, lag, t_lo, t_hi, rep);
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

// ---------------------------------------------------------------------------
// MPR-lite on the device: every parameter row of one set from its gamma
// vector (the order of mhmhip.h's MHMHIP_MPRLITE_* indices) and the cells'
// land-cover fractions, soil texture, aspect and latitude -- the synthetic
// domains' stand-in regionalisation, the same operations in the same order
// as its numpy statement (parameters.mpr_lite with static_fields and
// pet_radiation_params), so every row but the root fractions has the host's
// bits (those take c**(z/10) from the chain's power instead of libm's pow).
// The rows hold what mhmhip_set_cell_var stores for those fields: the
// degree-day rows times c2TSTu, the recession rows as c2TSTu / k.  The snow
// and sealed-threshold fields follow the reference's transfer functions
// (src/MPR/mo_multi_param_reg.f90:800-852, 883-894). One lane per cell.
__global__ __launch_bounds__(256) void mprlite_kernel(MprLiteArgs a, int32_t set) {
  mathtab_fill();  // the root fractions' power reads it (before any thread returns)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.n_cells) return;
  const int64_t S = a.sstride;
  const double fF = a.in[c], fI = a.in[S + c], fP = a.in[2 * S + c];
  const double tx = a.in[3 * S + c], asp = a.in[4 * S + c];
  double *P = a.P + (int64_t)set * S + c;
  auto put = [&](int row, double v) { P[(int64_t)row * a.ld] = v; };
  const double *g = a.gamma;
  const int nh = a.nh;
  const double c2 = a.c2;
  put(PR_FS, fI);
  // snow_acc_melt_param: degree-day factors per cover class, then the
  // class-weighted cell values
  const double ddf_f = g[MHMHIP_MPRLITE_DDF_FOREST];
  const double ddf_p = ddf_f + g[MHMHIP_MPRLITE_DDF_PERVIOUS];
  const double ddf_i = ddf_p + g[MHMHIP_MPRLITE_DDF_IMPERVIOUS];
  const double max_f = ddf_f + g[MHMHIP_MPRLITE_MAXDDF_FOREST];
  const double max_i = ddf_i + g[MHMHIP_MPRLITE_MAXDDF_IMPERVIOUS];
  const double max_p = ddf_p + g[MHMHIP_MPRLITE_MAXDDF_PERVIOUS];
  put(PR_TTH, g[MHMHIP_MPRLITE_SNOW_TEMP]);
  put(PR_INCR, g[MHMHIP_MPRLITE_DDF_INCREASE]);
  put(PR_DNPS, (ddf_f * fF + ddf_i * fI + ddf_p * fP) * c2);
  put(PR_DMAXS, (max_f * fF + max_i * fI + max_p * fP) * c2);
  put(PR_SEALTHR, g[MHMHIP_MPRLITE_IMPERVIOUS_STORAGE]);
  // monthly leaf area index: interception capacity, PET-LAI factor and the
  // radiation cases' fields
  for (int m = 0; m < 12; ++m) {
    const double lai = fF * a.lai[m] + fI * a.lai[12 + m] + fP * a.lai[24 + m];
    put(pr_maxinter(nh) + m, g[MHMHIP_MPRLITE_CANOPY] * lai);
    put(pr_petlai(nh) + m, fmin(fmax(0.6 + 0.08 * lai, 0.3), 1.3));
    if (a.pet_case == 2) put(pr_prietay(nh) + m, 1.05 + 0.06 * lai);
    if (a.pet_case == 3) {
      put(pr_aero(nh) + m, 208.0 / (0.6 + 0.1 * lai));
      put(pr_surf(nh) + m, 40.0 + 120.0 / (1.0 + lai));
    }
  }
  // soil water contents per horizon
  const double poro = 0.38 + 0.12 * tx, fcr = 0.55 + 0.25 * (1.0 - tx);
  const double wpr = 0.30 + 0.25 * tx;
  const double sexp = g[MHMHIP_MPRLITE_INFIL_SHAPE] * (1.0 + 0.3 * tx);
  double rc = fF * g[MHMHIP_MPRLITE_ROOT_FOREST] + fI * g[MHMHIP_MPRLITE_ROOT_IMPERVIOUS] +
              fP * g[MHMHIP_MPRLITE_ROOT_PERVIOUS];
  rc = fmin(fmax(rc, 1e-3), 0.999);
  double fr[MHMHIP_MAX_HORIZONS];
  double cum_top = 0.0, fsum = 0.0, top = 0.0;
  for (int h = 0; h < nh; ++h) {
    const double dz = a.hd[h] - top;
    top = a.hd[h];
    const double sat = dz * poro, fc = sat * fcr;
    put(pr_sat(nh) + h, sat);
    put(pr_fc(nh) + h, fc);
    put(pr_wp(nh) + h, fc * wpr);
    put(pr_exp(nh) + h, sexp);
    // root fraction: cumulative 1 - c**(z / 10 mm), differenced, normalised
    const double cum = 1.0 - exp_ylogx(a.hd[h] / 10.0, rc);
    fr[h] = fmax(cum - cum_top, 0.0);
    cum_top = cum;
    fsum = fsum + fr[h];
  }
  const double fden = fmax(fsum, 1e-12);
  for (int h = 0; h < nh; ++h) put(pr_froots(nh) + h, fr[h] / fden);
  put(PR_JARVIS, g[MHMHIP_MPRLITE_JARVIS_C1]);
  // recession constants [d], stored as c2TSTu / k
  const double k1 = g[MHMHIP_MPRLITE_SLOW_KS] * (0.5 + tx);
  put(PR_K0S, c2 / (g[MHMHIP_MPRLITE_FAST_FOREST] * (1.0 + 0.5 * (1.0 - fF))));
  put(PR_K1S, c2 / k1);
  put(PR_KPS, c2 / fmax(g[MHMHIP_MPRLITE_RECHARGE] * (0.5 + tx), 2.0));
  put(PR_K2S, c2 / fmax(g[MHMHIP_MPRLITE_GEOPARAM] * (0.5 + 0.5 * tx), k1));
  put(PR_ALPHA, g[MHMHIP_MPRLITE_EXP_SLOW]);
  put(PR_KARST, 1.0);
  put(PR_UNSATTHR, g[MHMHIP_MPRLITE_INTERFLOW_CAPACITY] * (0.2 + 0.3 * tx));
  // PET aspect correction
  const double thr = g[MHMHIP_MPRLITE_ASPECT_THRESH], mn = g[MHMHIP_MPRLITE_PET_MIN];
  const double mx = g[MHMHIP_MPRLITE_PET_MAX];
  put(PR_FASP, asp < thr ? mn + mx * (thr - asp) / thr : mn + mx * (asp - thr) / (360.0 - thr));
  put(PR_HARSAM, g[MHMHIP_MPRLITE_HARGREAVES]);
  put(PR_TANLAT, a.in[5 * S + c]);
  put(PR_SINLAT, a.in[6 * S + c]);
  put(PR_COSLAT, a.in[7 * S + c]);
}

hipError_t launch_mprlite(const MprLiteArgs &a, int32_t set, hipStream_t s) {
  if (a.n_cells < 1 || a.nh < 1 || a.nh > MHMHIP_MAX_HORIZONS || !a.P || !a.in || set < 0)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(mprlite_kernel, dim3((unsigned)((a.n_cells + 255) / 256)), dim3(256), 0, s, a,
                     set);
  return hipGetLastError();
}

}  // namespace mhmhip
