/*
 * mhm_oracle.c -- CPU restatement of the mHM per-cell process chain and the
 * mRM Muskingum routing (FP64).  TEST INFRASTRUCTURE ONLY: the checker for
 * the HIP path and the timed CPU baseline; never part of the product.
 *
 * Each function restates the cited reference routine statement by statement.
 * Compiled with -ffp-contract=off so that every multiply and add rounds
 * separately, as the reference's unfused Fortran does on x86-64.
 */
#include "mhm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/common/mo_common_constants.f90: eps_dp = epsilon(1.0_dp) */
#define EPS_DP DBL_EPSILON
/* FORCES v0.6.0 mo_constants (absent from the tree; values as published,
 * cross-checked with the older CHS copy pre-proc/mtCLIM/mo_constants.f90:52-165) */
static const double TWOTHIRD_DP = 0.6666666666666666666666666666666666667;
static const double PI_D = 3.141592653589793238462643383279502884197;
static const double TWOPI_D = 6.283185307179586476925286766559005768394;
#define DEG2RAD_DP (3.141592653589793238462643383279502884197 / 180.0)
static const double DAYSECS = 86400.0;
static const double YEARDAYS = 365.0;
static const double HOURSECS = 3600.0;
static const double SOLARCONST_DP = 1367.0;
static const double SPECHEATET_DP = 2.45e06;
static const double PSYCHRO_DP = 0.0646;
static const double RHO0_DP = 1.225;
static const double CP0_DP = 1005.0;
/* src/mHM/mo_mhm_constants.f90:36-48 */
static const double DUFFIE_DR = 0.0330;
static const double DUFFIE_DELTA1 = 0.4090;
static const double DUFFIE_DELTA2 = 1.3900;
static const double TETENS_C1 = 0.6108;
static const double TETENS_C2 = 17.270;
static const double TETENS_C3 = 237.30;
static const double SATPRESSURESLOPE1 = 4098.0;
static const double HARSAMCONST = 17.800;
static const double T0_DP = 273.15; /* FORCES mo_constants: T0_dp */

/* ---------------------------------------------------------------------------
 * src/mHM/mo_canopy_interc.f90:75-133
 */
void or_canopy_interc(double pet, double interc_max, double precip, double *interc,
                      double *throughfall, double *evap_canopy)
{
  double aux_help = *interc + precip;
  if (aux_help >= interc_max) {
    *throughfall = aux_help - interc_max;
    *interc = interc_max;
  } else {
    *throughfall = 0.0;
    *interc = aux_help;
  }
  if (interc_max > EPS_DP) {
    *evap_canopy = pet * pow(*interc / interc_max, TWOTHIRD_DP);
  } else {
    *evap_canopy = 0.0;
  }
  if (*evap_canopy < 0.0) *evap_canopy = 0.0;
  if (*interc > *evap_canopy) {
    *interc = *interc - *evap_canopy;
  } else {
    *evap_canopy = *interc;
    *interc = 0.0;
  }
}

/* ---------------------------------------------------------------------------
 * src/mHM/mo_snow_accum_melt.f90:70-158
 */
void or_snow_accum_melt(double deg_day_incr, double deg_day_max, double deg_day_noprec,
                        double prec, double temperature, double temperature_thresh,
                        double thrfall, double *snow_pack, double *deg_day, double *melt,
                        double *prec_effect, double *rain, double *snow)
{
  double aux_help;
  if (temperature > temperature_thresh) {
    *snow = 0.0;
    *rain = thrfall;
  } else {
    *snow = thrfall;
    *rain = 0.0;
  }
  if (prec <= (deg_day_max - deg_day_noprec) / deg_day_incr) {
    *deg_day = deg_day_noprec + deg_day_incr * prec;
  } else {
    *deg_day = deg_day_max;
  }
  if (temperature > temperature_thresh) {
    if (*snow_pack > 0.0) {
      aux_help = *deg_day * (temperature - temperature_thresh);
      if (aux_help > *snow_pack) {
        *melt = *snow_pack;
        *snow_pack = 0.0;
      } else {
        *melt = aux_help;
        *snow_pack = *snow_pack - aux_help;
      }
    } else {
      *melt = 0.0;
      *snow_pack = 0.0;
    }
  } else {
    *melt = 0.0;
    *snow_pack = *snow_pack + *snow;
  }
  *prec_effect = *melt + *rain;
}

/* ---------------------------------------------------------------------------
 * src/mHM/mo_soil_moisture.f90:333-363
 */
double or_feddes_et_reduction(double sm, double sm_fc, double wp, double froots)
{
  if (sm >= sm_fc) return froots;
  if (sm > wp) return froots * (sm - wp) / (sm_fc - wp);
  return 0.0;
}

/* src/mHM/mo_soil_moisture.f90:405-446 */
double or_jarvis_et_reduction(double sm, double sm_sat, double wp, double froots, double c1)
{
  double theta = (sm - wp) / (sm_sat - wp);
  double r = 0.0;
  if (theta < 0.0) theta = 0.0;
  if (theta > 1.0) theta = 1.0;
  if (theta >= c1) {
    r = froots;
  } else if (theta < c1) {
    r = froots * (theta / c1);
  }
  return r;
}

/* ---------------------------------------------------------------------------
 * src/mHM/mo_soil_moisture.f90:94-290
 */
void or_soil_moisture(int processCase, int nH, double frac_sealed, double water_thresh_sealed,
                      double pet, double evap_coeff, const double *sm_sat,
                      const double *frac_roots, const double *sm_fc, const double *wilting_point,
                      const double *sm_exponen, double jarvis_thresh_c1, double aet_canopy,
                      double prec_effec, double *runoff_sealed, double *storage_sealed,
                      double *infiltration, double *soil_moist, double *aet, double *aet_sealed)
{
  double tmp, prec_effec_soil, frac_runoff, soil_stress_factor = 0.0;
  int hh, j;

  /* impervious cover (179-210) */
  *runoff_sealed = 0.0;
  *aet_sealed = 0.0;
  if (frac_sealed > 0.0) {
    tmp = *storage_sealed + prec_effec;
    if (tmp > water_thresh_sealed) {
      *runoff_sealed = tmp - water_thresh_sealed;
      *storage_sealed = water_thresh_sealed;
    } else {
      *runoff_sealed = 0.0;
      *storage_sealed = tmp;
    }
    if (water_thresh_sealed > EPS_DP) {
      *aet_sealed = (pet / evap_coeff - aet_canopy) * (*storage_sealed / water_thresh_sealed);
      if (*aet_sealed < 0.0) *aet_sealed = 0.0;
    } else {
      *aet_sealed = DBL_MAX; /* huge(1.0_dp) */
    }
    if (*storage_sealed > *aet_sealed) {
      *storage_sealed = *storage_sealed - *aet_sealed;
    } else {
      *aet_sealed = *storage_sealed;
      *storage_sealed = 0.0;
    }
  }

  /* N-layer soil module (215-286) */
  for (hh = 0; hh < nH; ++hh) {
    aet[hh] = 0.0;
    infiltration[hh] = 0.0;
  }
  prec_effec_soil = prec_effec;
  for (hh = 0; hh < nH; ++hh) {
    if (hh != 0) prec_effec_soil = infiltration[hh - 1];
    if (soil_moist[hh] > sm_sat[hh]) {
      infiltration[hh] = prec_effec_soil;
    } else {
      if (soil_moist[hh] > EPS_DP) {
        frac_runoff = exp(sm_exponen[hh] * log(soil_moist[hh] / sm_sat[hh]));
      } else {
        frac_runoff = 0.0;
      }
      tmp = prec_effec_soil * (1.0 - frac_runoff);
      if ((soil_moist[hh] + tmp) > sm_sat[hh]) {
        infiltration[hh] = prec_effec_soil + (soil_moist[hh] - sm_sat[hh]);
        soil_moist[hh] = sm_sat[hh];
      } else {
        infiltration[hh] = prec_effec_soil - tmp;
        soil_moist[hh] = soil_moist[hh] + tmp;
      }
    }
    /* ET demand top to bottom (254-255): sum(aet(1:hh-1), mask=aet>0) */
    aet[hh] = pet - aet_canopy;
    if (hh != 0) {
      double s = 0.0;
      for (j = 0; j < hh; ++j)
        if (aet[j] > 0.0) s = s + aet[j];
      aet[hh] = aet[hh] - s;
    }
    switch (processCase) {
      case 1:
      case 4:
        soil_stress_factor =
            or_feddes_et_reduction(soil_moist[hh], sm_fc[hh], wilting_point[hh], frac_roots[hh]);
        break;
      case 2:
      case 3:
        soil_stress_factor = or_jarvis_et_reduction(soil_moist[hh], sm_sat[hh],
                                                    wilting_point[hh], frac_roots[hh],
                                                    jarvis_thresh_c1);
        break;
      default:
        break;
    }
    aet[hh] = aet[hh] * soil_stress_factor;
    if (aet[hh] < 0.0) aet[hh] = 0.0;
    if (soil_moist[hh] > aet[hh]) {
      soil_moist[hh] = soil_moist[hh] - aet[hh];
    } else {
      aet[hh] = soil_moist[hh] - EPS_DP;
      soil_moist[hh] = EPS_DP;
    }
    if (soil_moist[hh] < EPS_DP) soil_moist[hh] = EPS_DP;
  }
}

/* ---------------------------------------------------------------------------
 * src/mHM/mo_runoff.f90:74-154
 */
static inline double fmin_f(double a, double b) { return (b < a) ? b : a; }

void or_runoff_unsat_zone(double k1, double kp, double k0, double alpha, double karst_loss,
                          double pefec_soil, double unsat_thresh, double *sat_storage,
                          double *unsat_storage, double *slow_interflow,
                          double *fast_interflow, double *perc)
{
  *unsat_storage = *unsat_storage + pefec_soil;
  *fast_interflow = 0.0;
  if (*unsat_storage > unsat_thresh) {
    *fast_interflow =
        fmin_f(k0 * (*unsat_storage - unsat_thresh), *unsat_storage - EPS_DP);
  }
  *unsat_storage = *unsat_storage - *fast_interflow;
  *slow_interflow = 0.0;
  if (*unsat_storage > EPS_DP) {
    *slow_interflow =
        fmin_f(k1 * pow(*unsat_storage, 1.0 + alpha), *unsat_storage - EPS_DP);
  }
  *unsat_storage = *unsat_storage - *slow_interflow;
  *perc = kp * *unsat_storage;
  if (*unsat_storage > *perc) {
    *unsat_storage = *unsat_storage - *perc;
    *sat_storage = *sat_storage + *perc * karst_loss;
  } else {
    *sat_storage = *sat_storage + *unsat_storage * karst_loss;
    *unsat_storage = 0.0;
  }
}

/* src/mHM/mo_runoff.f90:191-212 */
void or_runoff_sat_zone(double k2, double *sat_storage, double *baseflow)
{
  if (*sat_storage > 0.0) {
    *baseflow = k2 * *sat_storage;
    *sat_storage = *sat_storage - *baseflow;
  } else {
    *baseflow = 0.0;
    *sat_storage = 0.0;
  }
}

/* src/mHM/mo_runoff.f90:248-274 */
double or_L1_total_runoff(double fsealed, double fast, double slow, double baseflow,
                          double direct_runoff)
{
  return ((baseflow + slow + fast) * (1.0 - fsealed)) + (direct_runoff * fsealed);
}

/* ---------------------------------------------------------------------------
 * FORCES mo_utils LE (older CHS copy: pre-proc/mtCLIM/mo_utils.f90:922-934)
 */
int or_LE(double a, double b)
{
  int r = 1;
  if (((DBL_EPSILON * fabs(b) - fabs(a - b)) < 0.0) && (a > b)) r = 0;
  return r;
}

/* src/mHM/mo_pet.f90:314-355 */
double or_extraterr_rad_approx(int doy, double latitude)
{
  double dr, delta, omega, arg;
  dr = 1.0 + DUFFIE_DR * cos(TWOPI_D * (double)doy / YEARDAYS);
  delta = DUFFIE_DELTA1 * sin(TWOPI_D * (double)doy / YEARDAYS - DUFFIE_DELTA2);
  arg = -tan(latitude) * tan(delta);
  if (arg < -1.0) arg = -1.0;
  if (arg > 1.0) arg = 1.0;
  omega = acos(arg);
  return DAYSECS / PI_D / SPECHEATET_DP * SOLARCONST_DP * dr *
         (omega * sin(latitude) * sin(delta) + cos(latitude) * cos(delta) * sin(omega));
}

/* src/mHM/mo_pet.f90:73-116 (latitude in degrees, converted at line 113) */
double or_pet_hargreaves(double coeff, double cnst, double tavg, double tmax, double tmin,
                         double latitude, int doy)
{
  double delta_temp = tmax - tmin;
  if (or_LE(delta_temp, 0.0) || or_LE(tavg, -cnst)) return 0.0;
  return coeff * or_extraterr_rad_approx(doy, DEG2RAD_DP * latitude) * (tavg + cnst) *
         sqrt(delta_temp);
}

/* src/mHM/mo_pet.f90:426-441 */
double or_sat_vap_pressure(double tavg)
{
  return TETENS_C1 * exp(TETENS_C2 * tavg / (tavg + TETENS_C3));
}

/* src/mHM/mo_pet.f90:384-399 */
double or_slope_satpressure(double tavg)
{
  return SATPRESSURESLOPE1 * or_sat_vap_pressure(tavg) / exp(2.0 * log(tavg + TETENS_C3));
}

/* src/mHM/mo_pet.f90:149-174 */
double or_pet_priestly(double alpha, double rn, double tavg)
{
  double delta = or_slope_satpressure(tavg);
  return alpha * delta / (PSYCHRO_DP + delta) * (rn * DAYSECS / SPECHEATET_DP);
}

/* src/mHM/mo_pet.f90:235-272 */
double or_pet_penman(double net_rad, double tavg, double act_vap_pressure,
                     double aerodyn_resistance, double bulksurface_resistance, double a_s,
                     double a_sh)
{
  return DAYSECS / SPECHEATET_DP *
         (or_slope_satpressure(tavg) * net_rad +
          RHO0_DP * CP0_DP * (or_sat_vap_pressure(tavg) - act_vap_pressure) * a_sh /
              aerodyn_resistance) /
         (or_slope_satpressure(tavg) +
          PSYCHRO_DP * a_sh / a_s * (1.0 + bulksurface_resistance / aerodyn_resistance));
}

/* ---------------------------------------------------------------------------
 * src/meteo/mo_meteo_temporal_tools.f90:32-161
 */
double or_disagg_meteo_weights(double v_day, double weight, double correction)
{
  return (v_day + correction) * weight - correction;
}

double or_disagg_flux_daynight(int isday, double ntimesteps_day, double v_day, double fday,
                               double fnight)
{
  if (ntimesteps_day > 1.0) {
    if (isday) return 2.0 * v_day * fday / ntimesteps_day;
    return 2.0 * v_day * fnight / ntimesteps_day;
  }
  return v_day;
}

double or_disagg_state_daynight(int isday, double ntimesteps_day, double v_day, double fday,
                                double fnight, int add_correction)
{
  if (ntimesteps_day > 1.0) {
    if (add_correction) {
      if (isday) return v_day + fday;
      return v_day + fnight;
    }
    if (isday) return 2.0 * v_day * fday;
    return 2.0 * v_day * fnight;
  }
  return v_day;
}

/* ---------------------------------------------------------------------------
 * row-layout helpers
 */
int or_n_param_rows(int nH) { return OR_P_NSCALAR + 5 * nH + 60; }
int or_n_state_rows(int nH) { return OR_S_NSCALAR + nH; }
int or_n_flux_rows(int nH) { return OR_X_NSCALAR + 2 * nH; }

#define PROW(r) (params + (int64_t)(r) * ld)
#define SROW(r) (states + (int64_t)(r) * ld)
#define XROW(r) (fluxes + (int64_t)(r) * ld)

/* ---------------------------------------------------------------------------
 * One model time step: the meteo getters (src/meteo/mo_meteo_handler.f90:
 * update_timestep 575-620, get_corrected_pet 915-1123, get_temp 1126-1205,
 * get_prec 1208-1285) followed by mHM() (src/mHM/mo_mhm.f90:229-521), in the
 * order of src/mHM/mo_mhm_interface_run.f90:357-457.
 */
void or_mhm_step(const or_cells_cfg *cfg, const or_step_info *st, const double *params,
                 const double *forcing, double *states, double *fluxes, int nthreads)
{
  const int64_t n = cfg->n_cells, ld = cfg->ld;
  const int nH = cfg->n_h;
  const int m = st->month - 1;
  const int isday = (st->hour > 6) && (st->hour <= 18); /* mo_meteo_handler.f90:1048 */
  const int64_t day = (int64_t)(st->iMeteoTS - 1);
  const int64_t fld = cfg->forcing_ld;
  const double *f_pre = forcing + ((int64_t)OR_F_PRE * cfg->forcing_days + day) * fld;
  const double *f_tavg = forcing + ((int64_t)OR_F_TAVG * cfg->forcing_days + day) * fld;
  const double *f_tmin = forcing + ((int64_t)OR_F_TMIN * cfg->forcing_days + day) * fld;
  const double *f_tmax = forcing + ((int64_t)OR_F_TMAX * cfg->forcing_days + day) * fld;
  const double *f_pet = forcing + ((int64_t)OR_F_PET * cfg->forcing_days + day) * fld;
  const double *f_rn = forcing + ((int64_t)OR_F_NETRAD * cfg->forcing_days + day) * fld;
  const double *f_avp = forcing + ((int64_t)OR_F_ABSVAPPRESS * cfg->forcing_days + day) * fld;
  const double *f_ws = forcing + ((int64_t)OR_F_WINDSPEED * cfg->forcing_days + day) * fld;
  /* L1_x_weights(cell, month, hour + 1), here [var][month * 24 + hour][cell] */
  const double *w_pre = NULL, *w_temp = NULL, *w_pet = NULL;
  /* mo_meteo_handler.f90:404-406 */
  const double fday_prec = 1.0 - cfg->fnight_prec[m];
  const double fday_pet = 1.0 - cfg->fnight_pet[m];
  const double fday_temp = -1.0 * cfg->fnight_temp[m];
  const double fnight_prec = cfg->fnight_prec[m];
  const double fnight_pet = cfg->fnight_pet[m];
  const double fnight_temp = cfg->fnight_temp[m];
  const double nT = cfg->nTstepDay;
  const double c2TSTu = cfg->c2TSTu;
  const double evap = cfg->evap_coeff[m];
  const int hourly = cfg->hourly_forcing;
  const int pet_case = cfg->process_pet;
  const int soil_case = cfg->process_soil;
  const int rSat = OR_P_NSCALAR, rFr = rSat + nH, rFC = rFr + nH, rWP = rFC + nH,
            rExp = rWP + nH, rMaxI = rExp + nH, rLAIc = rMaxI + 12, rPT = rLAIc + 12,
            rAero = rPT + 12, rSurf = rAero + 12;
  double *pet_calc = XROW(OR_X_PET), *temp_calc = XROW(OR_X_TEMP), *prec_calc = XROW(OR_X_PREC);
  int64_t k;
  if (cfg->meteo_weights) {
    const int64_t wr = (int64_t)m * 24 + st->hour;
    w_pre = cfg->meteo_weights + wr * fld;
    w_temp = cfg->meteo_weights + ((int64_t)288 + wr) * fld;
    w_pet = cfg->meteo_weights + ((int64_t)576 + wr) * fld;
  }
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
#endif

  /* get_corrected_pet (mo_meteo_handler.f90:1050-1121) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (k = 0; k < n; ++k) {
    double pet;
    if (pet_case == 1) {
      pet = PROW(OR_P_FASP)[k] *
            or_pet_hargreaves(PROW(OR_P_HARSAM)[k], HARSAMCONST, f_tavg[k], f_tmax[k],
                              f_tmin[k], PROW(OR_P_LAT)[k], st->doy);
    } else if (pet_case == 0) {
      pet = PROW(OR_P_FASP)[k] * f_pet[k];
    } else if (pet_case == 2) { /* Priestley-Taylor (1078-1083), iLAI = month */
      pet = or_pet_priestly(PROW(rPT + m)[k], fmax(f_rn[k], 0.0), f_tavg[k]);
    } else if (pet_case == 3) { /* Penman-Monteith (1085-1095) */
      pet = or_pet_penman(fmax(f_rn[k], 0.0), f_tavg[k], f_avp[k] / 1000.0,
                          PROW(rAero + m)[k] / f_ws[k], PROW(rSurf + m)[k], 1.0, 1.0);
    } else { /* -1: LAI-driven correction, iLAI = month */
      pet = PROW(rLAIc + m)[k] * f_pet[k];
    }
    if (hourly) {
      pet_calc[k] = pet;
    } else if (w_pet) { /* read_meteo_weights (1102-1107) */
      pet_calc[k] = or_disagg_meteo_weights(pet, w_pet[k], 0.0);
    } else {
      pet_calc[k] = or_disagg_flux_daynight(isday, nT, pet, fday_pet, fnight_pet);
    }
  }
  /* get_temp (1175-1203) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (k = 0; k < n; ++k) {
    temp_calc[k] = hourly   ? f_tavg[k]
                   : w_temp ? or_disagg_meteo_weights(f_tavg[k], w_temp[k], T0_DP) /* 1183-1189 */
                            : or_disagg_state_daynight(isday, nT, f_tavg[k], fday_temp,
                                                       fnight_temp, 1);
  }
  /* get_prec (1256-1283) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (k = 0; k < n; ++k) {
    prec_calc[k] = hourly  ? f_pre[k]
                   : w_pre ? or_disagg_meteo_weights(f_pre[k], w_pre[k], 0.0) /* 1265-1270 */
                           : or_disagg_flux_daynight(isday, nT, f_pre[k], fday_prec, fnight_prec);
  }

  /* mHM(): initial soil moisture (mo_mhm.f90:448-450) */
  if (st->tt == 1 && !st->read_states) {
    int h;
    for (h = 0; h < nH; ++h)
      for (k = 0; k < n; ++k) SROW(OR_S_NSCALAR + h)[k] = 0.5 * PROW(rFC + h)[k];
  }

  /* mHM(): per-cell loop (mo_mhm.f90:455-519) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
  for (k = 0; k < n; ++k) {
    double sm_sat[16], froots[16], fc[16], wp[16], bexp[16], sm[16], infil[16], aet[16];
    double interc = SROW(OR_S_INTERC)[k], pack = SROW(OR_S_SNOWPACK)[k];
    double seal = SROW(OR_S_SEALSTW)[k], unsat = SROW(OR_S_UNSATSTW)[k],
           sat = SROW(OR_S_SATSTW)[k];
    double thr, ecan, ddf, melt, peff, rain, snow, rseal, aseal, slow, fast, perc, bflow;
    int h;
    for (h = 0; h < nH; ++h) {
      sm_sat[h] = PROW(rSat + h)[k];
      froots[h] = PROW(rFr + h)[k];
      fc[h] = PROW(rFC + h)[k];
      wp[h] = PROW(rWP + h)[k];
      bexp[h] = PROW(rExp + h)[k];
      sm[h] = SROW(OR_S_NSCALAR + h)[k];
    }
    or_canopy_interc(pet_calc[k], PROW(rMaxI + m)[k], prec_calc[k], &interc, &thr, &ecan);
    or_snow_accum_melt(PROW(OR_P_DEGDAY_INCR)[k], PROW(OR_P_DEGDAY_MAX)[k] * c2TSTu,
                       PROW(OR_P_DEGDAY_NOPRE)[k] * c2TSTu, prec_calc[k], temp_calc[k],
                       PROW(OR_P_TEMP_THRESH)[k], thr, &pack, &ddf, &melt, &peff, &rain, &snow);
    or_soil_moisture(soil_case, nH, PROW(OR_P_FSEALED)[k], PROW(OR_P_SEALED_THRESH)[k],
                     pet_calc[k], evap, sm_sat, froots, fc, wp, bexp, PROW(OR_P_JARVIS_C1)[k],
                     ecan, peff, &rseal, &seal, infil, sm, aet, &aseal);
    or_runoff_unsat_zone(c2TSTu / PROW(OR_P_K1)[k], c2TSTu / PROW(OR_P_KP)[k],
                         c2TSTu / PROW(OR_P_K0)[k], PROW(OR_P_ALPHA)[k], PROW(OR_P_KARST)[k],
                         infil[nH - 1], PROW(OR_P_UNSAT_THRESH)[k], &sat, &unsat, &slow, &fast,
                         &perc);
    or_runoff_sat_zone(c2TSTu / PROW(OR_P_K2)[k], &sat, &bflow);

    SROW(OR_S_INTERC)[k] = interc;
    SROW(OR_S_SNOWPACK)[k] = pack;
    SROW(OR_S_SEALSTW)[k] = seal;
    SROW(OR_S_UNSATSTW)[k] = unsat;
    SROW(OR_S_SATSTW)[k] = sat;
    XROW(OR_X_THROUGHFALL)[k] = thr;
    XROW(OR_X_AET_CANOPY)[k] = ecan;
    XROW(OR_X_DEG_DAY)[k] = ddf;
    XROW(OR_X_MELT)[k] = melt;
    XROW(OR_X_PREC_EFFECT)[k] = peff;
    XROW(OR_X_RAIN)[k] = rain;
    XROW(OR_X_SNOW)[k] = snow;
    XROW(OR_X_RUNOFF_SEALED)[k] = rseal;
    XROW(OR_X_AET_SEALED)[k] = aseal;
    XROW(OR_X_FAST)[k] = fast;
    XROW(OR_X_SLOW)[k] = slow;
    XROW(OR_X_PERC)[k] = perc;
    XROW(OR_X_BASEFLOW)[k] = bflow;
    XROW(OR_X_TOTAL_RUNOFF)[k] =
        or_L1_total_runoff(PROW(OR_P_FSEALED)[k], fast, slow, bflow, rseal);
    for (h = 0; h < nH; ++h) {
      SROW(OR_S_NSCALAR + h)[k] = sm[h];
      XROW(OR_X_NSCALAR + h)[k] = infil[h];
      XROW(OR_X_NSCALAR + nH + h)[k] = aet[h];
    }
  }
}

/* src/mHM/mo_init_states.f90:258-345 (fluxes_states_default_init) with the
 * constants of src/MPR/mo_mpr_constants.f90 (P2=15, P3=10, P4=75, P5=1500,
 * C1_InitStateSM=0.25). */
void or_default_states(int64_t n_cells, int64_t ld, int n_h, const double *horizon_depth,
                       double *states, double *fluxes)
{
  int64_t k;
  int r, h;
  for (k = 0; k < n_cells; ++k) {
    states[OR_S_INTERC * ld + k] = 0.0;
    states[OR_S_SNOWPACK * ld + k] = 15.0;
    states[OR_S_SEALSTW * ld + k] = 0.0;
    states[OR_S_UNSATSTW * ld + k] = 10.0;
    states[OR_S_SATSTW * ld + k] = 75.0;
    for (h = 0; h < n_h - 1; ++h) {
      double d = (h == 0) ? horizon_depth[0] : horizon_depth[h] - horizon_depth[h - 1];
      states[(OR_S_NSCALAR + h) * ld + k] = d * 0.25;
    }
    states[(OR_S_NSCALAR + n_h - 1) * ld + k] =
        (1500.0 - (n_h >= 2 ? horizon_depth[n_h - 2] : 0.0)) * 0.25;
    for (r = 0; r < or_n_flux_rows(n_h); ++r) fluxes[r * ld + k] = 0.0;
  }
}

/* ---------------------------------------------------------------------------
 * src/mRM/mo_mrm_mpr.f90:61-119 (reg_rout)
 */
void or_reg_rout(const double param[5], int64_t n_links, const double *length,
                 const double *slope, const double *fFPimp, double TS, double *C1, double *C2)
{
  int64_t i;
  double ssMax = -DBL_MAX;
  for (i = 0; i < n_links; ++i)
    if (slope[i] > ssMax) ssMax = slope[i];
  for (i = 0; i < n_links; ++i) {
    double K = param[0] + param[1] * (length[i] * 0.001) + param[2] * slope[i] +
               param[3] * fFPimp[i];
    double xi = param[4] * (1.0 + slope[i] / ssMax);
    if (xi > 0.5) xi = 0.5;
    if (xi < 0.005) xi = 0.005;
    if (K > 0.5 * TS / xi) K = 0.5 * TS / xi;
    if (K < 0.5 * TS / (1.0 - xi)) K = 0.5 * TS / (1.0 - xi);
    C1[i] = TS / (K * (1.0 - xi) + 0.5 * TS);
    C2[i] = 1.0 - C1[i] * K / TS;
  }
}

/* src/mRM/mo_mrm_pre_routing.f90:77-143 (L11_runoff_acc) */
void or_L11_runoff_acc(int64_t n_l1, const double *q_all, const double *efec_area,
                       const int32_t *L1_L11_Id, int64_t n_l11, const double *L11_area,
                       const int32_t *L11_L1_Id, int32_t TS, int32_t map_flag, double *q_acc)
{
  int64_t k;
  double TST = HOURSECS * (double)TS;
  if (map_flag) {
    for (k = 0; k < n_l11; ++k) q_acc[k] = 0.0;
    for (k = 0; k < n_l1; ++k) {
      int64_t j = L1_L11_Id[k] - 1;
      q_acc[j] = q_acc[j] + q_all[k] * efec_area[k];
    }
    for (k = 0; k < n_l11; ++k) q_acc[k] = q_acc[k] * 1000.0 / TST;
  } else {
    for (k = 0; k < n_l11; ++k) q_acc[k] = q_all[L11_L1_Id[k] - 1];
    for (k = 0; k < n_l11; ++k) q_acc[k] = q_acc[k] * L11_area[k] * 1000.0 / TST;
  }
}

/* src/mRM/mo_mrm_pre_routing.f90:179-214 (add_inflow) */
void or_add_inflow(int32_t n_inflow, const int32_t *index_list, const int32_t *headwater,
                   const int32_t *node_list, const double *q_inflow, double *q_out)
{
  int32_t ii;
  for (ii = 0; ii < n_inflow; ++ii) {
    int64_t node = node_list[ii] - 1;
    if (headwater[ii]) {
      q_out[node] = q_out[node] + q_inflow[index_list[ii] - 1];
    } else {
      q_out[node] = q_inflow[index_list[ii] - 1];
    }
  }
}

/* src/mRM/mo_mrm_routing.f90:380-481 (L11_routing) */
void or_L11_routing(int64_t n_nodes, int64_t n_links, const int32_t *netPerm,
                    const int32_t *fromN, const int32_t *toN, const double *C1,
                    const double *C2, const double *qOUT, int32_t n_inflow,
                    const int32_t *inflow_headwater, const int32_t *inflow_node_list,
                    double *qTIN, double *qTR, double *Qmod)
{
  double *qTIN1 = qTIN, *qTIN2 = qTIN + n_nodes; /* IT1 = past, IT = current */
  double *qTR1 = qTR, *qTR2 = qTR + n_nodes;
  int64_t k, tNode;
  for (k = 0; k < n_nodes; ++k) qTIN2[k] = 0.0;
  for (k = 0; k < n_links; ++k) {
    int64_t i = netPerm[k] - 1;
    int64_t iNode = fromN[i] - 1;
    int64_t tN = toN[i] - 1;
    int32_t g;
    qTIN2[iNode] = qTIN2[iNode] + qOUT[iNode];
    qTR2[iNode] = qTR1[iNode] + C1[i] * (qTIN1[iNode] - qTR1[iNode]) +
                  C2[i] * (qTIN2[iNode] - qTIN1[iNode]);
    for (g = 0; g < n_inflow; ++g) {
      if ((tN == (int64_t)inflow_node_list[g] - 1) && !inflow_headwater[g]) qTR2[iNode] = 0.0;
    }
    qTIN2[tN] = qTIN2[tN] + qTR2[iNode];
  }
  tNode = toN[netPerm[n_links - 1] - 1] - 1;
  qTIN2[tNode] = qTIN2[tNode] + qOUT[tNode];
  for (k = 0; k < n_nodes; ++k) {
    Qmod[k] = qTIN2[k];
    qTR1[k] = qTR2[k];
    qTIN1[k] = qTIN2[k];
  }
}

/* src/mRM/mo_mrm_routing.f90:104-303 (mRM_routing, processCase 1) */
void or_mrm_routing(const or_route_cfg *rc, const double param[5], int64_t n_l1,
                    const double *L1_total_runoff, const double *L1_area,
                    const int32_t *L1_L11_Id, const double *L11_area,
                    const int32_t *L11_L1_Id, const int32_t *netPerm, const int32_t *fromN,
                    const int32_t *toN, const double *length, const double *slope,
                    const double *fFPimp, const int32_t *inflow_index, const int32_t *inflow_hw,
                    const int32_t *inflow_node, const double *inflow_q,
                    const int32_t *gauge_node, double *C1, double *C2, double *qOUT,
                    double *qTIN, double *qTR, double *Qmod, double *gauge_out)
{
  const int64_t nN = rc->n_nodes;
  const int64_t nL = nN - rc->n_outlets;
  int64_t k;
  int32_t g;
  if (rc->process_case == 1 && !rc->read_states && nN > 1)
    or_reg_rout(param, nL, length, slope, fFPimp, (double)rc->timestep, C1, C2);
  or_L11_runoff_acc(n_l1, L1_total_runoff, L1_area, L1_L11_Id, nN, L11_area, L11_L1_Id,
                    rc->timestep, rc->map_flag, qOUT);
  or_add_inflow(rc->n_inflow, inflow_index, inflow_hw, inflow_node, inflow_q, qOUT);
  if (nN > 1) {
    /* rout_loop routings of the same qOUT, Qmod averaged (lines 224, 243-284):
     * qAcc = 0, qAcc = qAcc + Qmod per loop, Qmod = qAcc / rout_loop */
    const int32_t loops = rc->rout_loop > 1 ? rc->rout_loop : 1;
    double *acc = (double *)calloc((size_t)nN, sizeof(double));
    if (!acc) return;
    for (int32_t it = 0; it < loops; ++it) {
      or_L11_routing(nN, nL, netPerm, fromN, toN, C1, C2, qOUT, rc->n_inflow, inflow_hw,
                     inflow_node, qTIN, qTR, Qmod);
      for (k = 0; k < nN; ++k) acc[k] = acc[k] + Qmod[k];
    }
    for (k = 0; k < nN; ++k) Qmod[k] = acc[k] / (double)loops;
    free(acc);
  } else {
    for (k = 0; k < nN; ++k) Qmod[k] = qOUT[k];
  }
  if (gauge_out)
    for (g = 0; g < rc->n_gauges; ++g) gauge_out[g] = Qmod[gauge_node[g] - 1];
}

/* src/mRM/mo_mrm_net_startup.f90:728-859 (L11_routing_order), O(nLinks^2) sweeps */
void or_routing_order(int64_t n_nodes, int64_t n_links, const int32_t *fromN,
                      const int32_t *toN, int32_t *rOrder, int32_t *netPerm)
{
  int64_t ii, jj, kk = 0;
  for (ii = 0; ii < n_nodes; ++ii) netPerm[ii] = -9999;
  for (ii = 0; ii < n_links; ++ii) rOrder[ii] = 1;
  if (n_nodes > n_links) rOrder[n_nodes - 1] = -9999;
  if (n_nodes <= 1) return;
  for (ii = 0; ii < n_links; ++ii) {
    for (jj = 0; jj < n_links; ++jj) {
      if (jj == ii) continue;
      if (fromN[ii] == toN[jj]) rOrder[ii] = -9;
      if (rOrder[ii] == -9) break;
    }
  }
  for (ii = 0; ii < n_links; ++ii)
    if (rOrder[ii] == 1) rOrder[ii] = (int32_t)(++kk);
  for (;;) {
    int32_t mn = 0x7fffffff;
    for (ii = 0; ii < n_links; ++ii)
      if (rOrder[ii] < mn) mn = rOrder[ii];
    if (!(mn < 0)) break;
    for (ii = 0; ii < n_links; ++ii) {
      int flag = 1;
      if (rOrder[ii] != -9) continue;
      for (jj = 0; jj < n_links; ++jj) {
        if (jj == ii || fromN[ii] != toN[jj]) continue;
        if (!(fromN[ii] == toN[jj] && rOrder[jj] > 0)) {
          flag = 0;
          break;
        }
      }
      if (flag) rOrder[ii] = (int32_t)(++kk);
    }
  }
  for (ii = 0; ii < n_links; ++ii) netPerm[rOrder[ii] - 1] = (int32_t)(ii + 1);
}

/* ------------------------------------------------------------------------
 * Forcing on the meteo grid (L2) to the L1 cells.
 * meteo_forcings_wrapper, mo_meteo_helper.f90:119-143; the spatial tools
 * mo_meteo_spatial_tools.f90:94-200 (aggregation) and 313-377
 * (disaggregation).  Loops in the reference's order: the aggregation adds
 * the valid L2 cells column by column (j outer, i inner) into their L1
 * cell, then divides by the count; nodata_dp = -9999 (mo_common_constants).
 * ------------------------------------------------------------------------ */
int or_meteo_l2_to_l1(int32_t nr2, int32_t nc2, double cellsize2, const uint8_t *mask2,
                      int32_t nr1, int32_t nc1, double cellsize1, const uint8_t *mask1,
                      int32_t ndays, const double *data2, double *out) {
  const double nodata = -9999.0;
  const int64_t n1 = (int64_t)nr1 * nc1, n2 = (int64_t)nr2 * nc2;
  int64_t ncells = 0;
  for (int64_t k = 0; k < n1; ++k) ncells += mask1[k] != 0;
  double *grid1 = (double *)malloc(sizeof(double) * (size_t)n1);
  if (!grid1) return -1;
  const double hbym = cellsize1 / cellsize2;  /* cellFactorHbyM (line 120) */
  int rc = 0;
  if (hbym > 1.0) {
    /* spatial_aggregation_3d */
    const double f = cellsize1 / cellsize2;
    int32_t *ntc = (int32_t *)calloc((size_t)n1, sizeof(int32_t));
    if (!ntc) {
      free(grid1);
      return -1;
    }
    for (int32_t j = 1; j <= nc2 && rc == 0; ++j) {
      const int64_t jc = (int64_t)ceil((double)j / f);
      for (int32_t i = 1; i <= nr2; ++i) {
        const int64_t ic = (int64_t)ceil((double)i / f);
        if (!mask2[(int64_t)(j - 1) * nr2 + (i - 1)]) continue;
        if (ic > nr1 || jc > nc1) {
          rc = -1;
          break;
        }
        ntc[(jc - 1) * nr1 + (ic - 1)] += 1;
      }
    }
    for (int32_t t = 0; t < ndays && rc == 0; ++t) {
      for (int64_t k = 0; k < n1; ++k) grid1[k] = 0.0;
      const double *d2 = data2 + (int64_t)t * n2;
      for (int32_t j = 1; j <= nc2; ++j) {
        const int64_t jc = (int64_t)ceil((double)j / f);
        for (int32_t i = 1; i <= nr2; ++i) {
          const int64_t ic = (int64_t)ceil((double)i / f);
          const int64_t k2 = (int64_t)(j - 1) * nr2 + (i - 1);
          if (!mask2[k2]) continue;
          grid1[(jc - 1) * nr1 + (ic - 1)] = grid1[(jc - 1) * nr1 + (ic - 1)] + d2[k2];
        }
      }
      for (int64_t k = 0; k < n1; ++k)
        grid1[k] = mask1[k] ? grid1[k] / (double)ntc[k] : nodata;
      int64_t c = 0;
      for (int64_t k = 0; k < n1; ++k)
        if (mask1[k]) out[(int64_t)t * ncells + c++] = grid1[k];
    }
    free(ntc);
  } else if (hbym < 1.0) {
    /* spatial_disaggregation_3d */
    const double f = cellsize2 / cellsize1;
    for (int32_t t = 0; t < ndays && rc == 0; ++t) {
      const double *d2 = data2 + (int64_t)t * n2;
      for (int32_t j = 1; j <= nc1 && rc == 0; ++j) {
        const int64_t jc = (int64_t)ceil((double)j / f);
        for (int32_t i = 1; i <= nr1; ++i) {
          const int64_t ic = (int64_t)ceil((double)i / f);
          const int64_t k1 = (int64_t)(j - 1) * nr1 + (i - 1);
          grid1[k1] = nodata;
          if (ic > nr2 || jc > nc2) {
            if (mask1[k1]) {
              rc = -1;
              break;
            }
            continue;
          }
          const int64_t k2 = (jc - 1) * nr2 + (ic - 1);
          if (!mask2[k2]) continue;
          grid1[k1] = d2[k2];
        }
      }
      int64_t c = 0;
      for (int64_t k = 0; k < n1 && rc == 0; ++k)
        if (mask1[k]) out[(int64_t)t * ncells + c++] = grid1[k];
    }
  } else {
    /* equal resolution: L1_data = L2_data (mo_meteo_helper.f90:129-131) */
    if (nr1 != nr2 || nc1 != nc2) rc = -1;
    for (int32_t t = 0; t < ndays && rc == 0; ++t) {
      int64_t c = 0;
      for (int64_t k = 0; k < n1; ++k)
        if (mask1[k]) out[(int64_t)t * ncells + c++] = data2[(int64_t)t * n2 + k];
    }
  }
  free(grid1);
  return rc;
}

/* ------------------------------------------------------------------------
 * Adaptive routing parameters, constant celerity (processCase(8) = 2):
 * mrm_update_param, mo_mrm_mpr.f90:268-329.  given_TS: mo_mrm_constants.F90
 * :42-46.  locate and LE as FORCES mo_utils, restated from the copy in
 * pre-proc/mtCLIM/mo_utils.f90 (lesserequal_dp 924-936, locate_0d_dp 1329-1346).
 * ------------------------------------------------------------------------ */
static const double kGivenTS[19] = {60.0,   120.0,   180.0,   240.0,   300.0,   360.0,  600.0,
                                    720.0,  900.0,   1200.0,  1800.0,  3600.0,  7200.0, 10800.0,
                                    14400.0, 21600.0, 28800.0, 43200.0, 86400.0};

/* lesserequal_dp, pre-proc/mtCLIM/mo_utils.f90:924-936 */
static int or_le(double a, double b) {
  return !(((DBL_EPSILON * fabs(b) - fabs(a - b)) < 0.0) && (a > b));
}

/* locate(x, y): c = minloc(|x - y|) (first minimum), c if LE(x(c), y) else c - 1 (1-based) */
static int or_locate(const double *x, int n, double y) {
  int c = 0;
  for (int i = 1; i < n; ++i)
    if (fabs(x[i] - y) < fabs(x[c] - y)) c = i;
  return or_le(x[c], y) ? c + 1 : c;
}

void or_mrm_update_param_celerity(int64_t n_links, const double *length, double celerity,
                                  int32_t timestep, double *C1, double *C2, double *ts_rout,
                                  double *ts_rout_factor, int32_t *rout_loop) {
  or_mrm_update_param_links(n_links, length, &celerity, 0, timestep, C1, C2, ts_rout,
                            ts_rout_factor, rout_loop);
}

/* mrm_update_param (src/mRM/mo_mrm_mpr.f90:241-329) for processCase(8) = 3:
 * K = L11_length / L11_celerity per link (302-311), the rest as for a
 * constant celerity.  stride 0: one celerity for every link (processCase
 * 2), 1: celerity[i] for link i. */
void or_mrm_update_param_links(int64_t n_links, const double *length, const double *celerity,
                               int32_t stride, int32_t timestep, double *C1, double *C2,
                               double *ts_rout, double *ts_rout_factor, int32_t *rout_loop) {
  const double xi = fabs(0.0); /* rout_space_weight */
  double kmin = 0.0;
  for (int64_t i = 0; i < n_links; ++i) {
    const double K = length[i] / celerity[i * stride];
    if (i == 0 || K < kmin) kmin = K;
  }
  int ind = or_locate(kGivenTS, 19, kmin);
  if (ind < 1) ind = 1;
  const double ts = kGivenTS[ind - 1];
  for (int64_t i = 0; i < n_links; ++i) {
    const double K = length[i] / celerity[i * stride];
    C1[i] = ts / (K * (1.0 - xi) + 0.5 * ts);
    C2[i] = 1.0 - C1[i] * K / ts;
  }
  *ts_rout = ts;
  /* tsRoutFactor = L11_tsRout / (timestep * HourSecs) */
  *ts_rout_factor = ts / ((double)timestep * 3600.0);
  /* rout_loop = max(1, nint(1 / tsRoutFactor)); Fortran nint rounds halves away from zero */
  const double inv = 1.0 / *ts_rout_factor;
  const int32_t nl = (int32_t)floor(inv + 0.5);
  *rout_loop = nl > 1 ? nl : 1;
}
