// chain.h -- the per-cell mHM process chain as device functions (FP64).
//
// Each function restates one reference routine with its operation order,
// branch structure and clamps unchanged, so the result of every +, -, *, /
// and sqrt is the reference's (the library is built with -ffp-contract=off).
// Only exp/log/pow/acos/sin come from the device math library.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace mhmhip {

// src/common/mo_common_constants.f90: eps_dp = epsilon(1.0_dp)
constexpr double kEps = DBL_EPSILON;
// FORCES mo_constants: twothird_dp
constexpr double kTwoThird = 0.6666666666666666666666666666666666667;
// src/mHM/mo_mhm_constants.f90:36-39
constexpr double kHarSamConst = 17.800;
constexpr double kDuffieDr = 0.0330;
constexpr double kDuffieDelta1 = 0.4090;
constexpr double kDuffieDelta2 = 1.3900;
// FORCES mo_constants: DaySecs / PI_D / SpecHeatET_dp * SolarConst_dp, folded
// left to right exactly as the reference expression (mo_pet.f90:353).
constexpr double kPi = 3.141592653589793238462643383279502884197;
constexpr double kRadConst = 86400.0 / kPi / 2.45e06 * 1367.0;

// x**y for y > 0 as exp(y*log(x)).  For the chain's arguments (|y*log(x)| of
// order 10 at most) this is within a few 1e-16 relative of the correctly
// rounded power, at a fraction of pow()'s cost; x <= 0 gives pow()'s result
// for non-integer y (0 for x == 0, NaN below).
// exp(y * log(x)), kept out of line: inlined at every call site inside the
// time loop, the polynomial constants of exp and log are hoisted into
// registers for the whole loop and cut the kernel's occupancy.
#ifndef MHMHIP_OUTLINE_TRANSC
#define MHMHIP_OUTLINE_TRANSC 1
#endif
#if MHMHIP_OUTLINE_TRANSC
#define MHMHIP_OUTLINE __attribute__((noinline))
#else
#define MHMHIP_OUTLINE __forceinline__
#endif
__device__ MHMHIP_OUTLINE double exp_ylogx(double y, double x) {
  return exp(y * log(x));
}

__device__ __forceinline__ double pow_pos(double x, double y) {
  if (x > 0.0) return exp_ylogx(y, x);
  return x == 0.0 ? 0.0 : __builtin_nan("");
}

// FORCES mo_utils LE: a <= b within eps*|b| (pre-proc/mtCLIM/mo_utils.f90:922-934)
__device__ __forceinline__ bool le_eps(double a, double b) {
  return !(((DBL_EPSILON * fabs(b) - fabs(a - b)) < 0.0) && (a > b));
}

// src/mHM/mo_canopy_interc.f90:75-133
__device__ __forceinline__ void canopy_interc(double pet, double interc_max, double precip,
                                              double &interc, double &throughfall,
                                              double &evap_canopy) {
  const double aux = interc + precip;
  if (aux >= interc_max) {
    throughfall = aux - interc_max;
    interc = interc_max;
  } else {
    throughfall = 0.0;
    interc = aux;
  }
  if (interc_max > kEps)
    evap_canopy = pet * pow_pos(interc / interc_max, kTwoThird);
  else
    evap_canopy = 0.0;
  if (evap_canopy < 0.0) evap_canopy = 0.0;
  if (interc > evap_canopy) {
    interc = interc - evap_canopy;
  } else {
    evap_canopy = interc;
    interc = 0.0;
  }
}

// src/mHM/mo_snow_accum_melt.f90:70-158.  ddthr = (deg_day_max - deg_day_noprec)
// / deg_day_incr with the c2TSTu-scaled factors (computed once per cell; the
// reference recomputes the same quotient every step).
__device__ __forceinline__ void snow_accum_melt(double incr, double dmax, double dnp,
                                                double ddthr, double prec, double temp,
                                                double tth, double thr, double &pack,
                                                double &deg_day, double &melt,
                                                double &prec_effect, double &rain,
                                                double &snow) {
  if (temp > tth) {
    snow = 0.0;
    rain = thr;
  } else {
    snow = thr;
    rain = 0.0;
  }
  if (prec <= ddthr)
    deg_day = dnp + incr * prec;
  else
    deg_day = dmax;
  if (temp > tth) {
    if (pack > 0.0) {
      const double aux = deg_day * (temp - tth);
      if (aux > pack) {
        melt = pack;
        pack = 0.0;
      } else {
        melt = aux;
        pack = pack - aux;
      }
    } else {
      melt = 0.0;
      pack = 0.0;
    }
  } else {
    melt = 0.0;
    pack = pack + snow;
  }
  prec_effect = melt + rain;
}

// src/mHM/mo_soil_moisture.f90:333-363
__device__ __forceinline__ double feddes(double sm, double fc, double wp, double fr) {
  if (sm >= fc) return fr;
  if (sm > wp) return fr * (sm - wp) / (fc - wp);
  return 0.0;
}

// src/mHM/mo_soil_moisture.f90:405-446
__device__ __forceinline__ double jarvis(double sm, double sat, double wp, double fr, double c1) {
  double th = (sm - wp) / (sat - wp);
  double r = 0.0;
  if (th < 0.0) th = 0.0;
  if (th > 1.0) th = 1.0;
  if (th >= c1)
    r = fr;
  else if (th < c1)
    r = fr * (th / c1);
  return r;
}

// src/mHM/mo_soil_moisture.f90:94-290
template <int NH>
__device__ __forceinline__ void soil_moisture(bool use_jarvis, double fS, double thrS, double pet,
                                              double evap, const double *sat, const double *fr,
                                              const double *fc, const double *wp,
                                              const double *bexp, double c1, double ecan,
                                              double peff, double &runoff_sealed,
                                              double &storage_sealed, double *infil, double *sm,
                                              double *aet, double &aet_sealed) {
  runoff_sealed = 0.0;
  aet_sealed = 0.0;
  if (fS > 0.0) {
    const double tmp = storage_sealed + peff;
    if (tmp > thrS) {
      runoff_sealed = tmp - thrS;
      storage_sealed = thrS;
    } else {
      runoff_sealed = 0.0;
      storage_sealed = tmp;
    }
    if (thrS > kEps) {
      aet_sealed = (pet / evap - ecan) * (storage_sealed / thrS);
      if (aet_sealed < 0.0) aet_sealed = 0.0;
    } else {
      aet_sealed = DBL_MAX;  // huge(1.0_dp)
    }
    if (storage_sealed > aet_sealed) {
      storage_sealed = storage_sealed - aet_sealed;
    } else {
      aet_sealed = storage_sealed;
      storage_sealed = 0.0;
    }
  }
  double pin = peff;
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    if (h != 0) pin = infil[h - 1];
    if (sm[h] > sat[h]) {
      infil[h] = pin;
    } else {
      double frac;
      if (sm[h] > kEps)
        frac = exp_ylogx(bexp[h], sm[h] / sat[h]);
      else
        frac = 0.0;
      const double tmp = pin * (1.0 - frac);
      if ((sm[h] + tmp) > sat[h]) {
        infil[h] = pin + (sm[h] - sat[h]);
        sm[h] = sat[h];
      } else {
        infil[h] = pin - tmp;
        sm[h] = sm[h] + tmp;
      }
    }
    double a = pet - ecan;
    if (h != 0) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < h; ++j)
        if (aet[j] > 0.0) s = s + aet[j];
      a = a - s;
    }
    const double f = use_jarvis ? jarvis(sm[h], sat[h], wp[h], fr[h], c1)
                                : feddes(sm[h], fc[h], wp[h], fr[h]);
    a = a * f;
    if (a < 0.0) a = 0.0;
    if (sm[h] > a) {
      sm[h] = sm[h] - a;
    } else {
      a = sm[h] - kEps;
      sm[h] = kEps;
    }
    if (sm[h] < kEps) sm[h] = kEps;
    aet[h] = a;
  }
}

// src/mHM/mo_runoff.f90:74-154 (k0, k1, kp already c2TSTu / k)
__device__ __forceinline__ void runoff_unsat_zone(double k1, double kp, double k0, double alpha,
                                                  double karst, double pin, double thr,
                                                  double &sat, double &unsat, double &slow,
                                                  double &fast, double &perc) {
  unsat = unsat + pin;
  fast = 0.0;
  if (unsat > thr) {
    const double a = k0 * (unsat - thr), b = unsat - kEps;
    fast = (b < a) ? b : a;
  }
  unsat = unsat - fast;
  slow = 0.0;
  if (unsat > kEps) {
    const double a = k1 * pow_pos(unsat, 1.0 + alpha), b = unsat - kEps;
    slow = (b < a) ? b : a;
  }
  unsat = unsat - slow;
  perc = kp * unsat;
  if (unsat > perc) {
    unsat = unsat - perc;
    sat = sat + perc * karst;
  } else {
    sat = sat + unsat * karst;
    unsat = 0.0;
  }
}

// src/mHM/mo_runoff.f90:191-212
__device__ __forceinline__ void runoff_sat_zone(double k2, double &sat, double &baseflow) {
  if (sat > 0.0) {
    baseflow = k2 * sat;
    sat = sat - baseflow;
  } else {
    baseflow = 0.0;
    sat = 0.0;
  }
}

// src/mHM/mo_runoff.f90:248-274
__device__ __forceinline__ double total_runoff(double fS, double fast, double slow, double base,
                                               double direct) {
  return ((base + slow + fast) * (1.0 - fS)) + (direct * fS);
}

// src/mHM/mo_pet.f90:73-116 with extraterr_rad_approx (314-355).  The
// latitude terms tan/sin/cos(deg2rad*lat) are per-cell constants and the
// day-of-year terms dr, tan/sin/cos(delta) per-step constants; both are
// evaluated once on the host with the same libm calls the reference makes.
__device__ MHMHIP_OUTLINE double pet_hargreaves(double coeff, double tavg, double tmax,
                                                 double tmin, double tanlat, double sinlat,
                                                 double coslat, double dr, double tand,
                                                 double sind, double cosd) {
  const double dT = tmax - tmin;
  if (le_eps(dT, 0.0) || le_eps(tavg, -kHarSamConst)) return 0.0;
  double arg = -tanlat * tand;
  if (arg < -1.0) arg = -1.0;
  if (arg > 1.0) arg = 1.0;
  const double omega = acos(arg);
  const double ra = kRadConst * dr * (omega * sinlat * sind + coslat * cosd * sin(omega));
  return coeff * ra * (tavg + kHarSamConst) * sqrt(dT);
}

// src/meteo/mo_meteo_temporal_tools.f90:54-79 with the day or night factor f
// already chosen: 2 * v * f / ntimesteps_day
__device__ __forceinline__ double disagg_flux(double nT, double v, double f) {
  if (nT > 1.0) return 2.0 * v * f / nT;
  return v;
}

// src/meteo/mo_meteo_temporal_tools.f90:100-161 with add_correction = .true.
__device__ __forceinline__ double disagg_state_add(double nT, double v, double f) {
  if (nT > 1.0) return v + f;
  return v;
}

}  // namespace mhmhip
