// chain.h -- the per-cell mHM process chain as device functions (FP64).
//
// Each function restates one reference routine with its operation order,
// branch structure and clamps unchanged, so the result of every +, -, *, /
// and sqrt is the reference's (the library is built with -ffp-contract=off).
// pow is evaluated by exp_ylogx below (own log and exp); acos and sin of
// the daily PET come from the device math library.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "mathtab.h"

namespace mhmhip {

// src/common/mo_common_constants.f90: eps_dp = epsilon(1.0_dp)
constexpr double kEps = DBL_EPSILON;
// FORCES mo_constants: twothird_dp
// src/mHM/mo_mhm_constants.f90:36-39
constexpr double kHarSamConst = 17.800;
constexpr double kDuffieDr = 0.0330;
constexpr double kDuffieDelta1 = 0.4090;
constexpr double kDuffieDelta2 = 1.3900;
// FORCES mo_constants: DaySecs / PI_D / SpecHeatET_dp * SolarConst_dp, folded
// left to right exactly as the reference expression (mo_pet.f90:353).
constexpr double kPi = 3.141592653589793238462643383279502884197;
constexpr double kRadConst = 86400.0 / kPi / 2.45e06 * 1367.0;

// Out-of-line device functions: inlined at every call site inside the time
// loop, their polynomial constants are hoisted into registers for the whole
// loop and cut the kernel's occupancy.
#ifndef MHMHIP_OUTLINE_TRANSC
#define MHMHIP_OUTLINE_TRANSC 1
#endif
#if MHMHIP_OUTLINE_TRANSC
#define MHMHIP_OUTLINE __attribute__((noinline))
#else
#define MHMHIP_OUTLINE __forceinline__
#endif

// fma(a, b, c) for a constant addend c.  On the device c is passed in a
// scalar register pair, so each Horner step is one v_fma_f64 (the compiler
// otherwise copies every 64-bit constant into two VGPRs with v_mov_b32 and
// accumulates with v_fmac, three vector instructions per step).  The operands
// a and b never come straight from a transcendental instruction.  Same IEEE
// fused multiply-add on host and device.
__host__ __device__ inline double fma_k(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

// ln 2 split so that k * kLn2Hi is exact for |k| < 2^21 (32 significant bits)
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
constexpr double kInvLn2 = 1.4426950408889634;    // RN(1/ln 2)
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;  // RN(sqrt(1/2))

// N doubles passed and returned by value (registers across an out-of-line call)
template <int N>
struct DV {
  double v[N];
};

// Natural log of finite x > 0 (normal or subnormal), N values in lockstep,
// table-driven (the method of Tang, ACM TOMS 16 (1990) 378-400): x = m 2^e
// with m in [1/2, 1) (frexp); F = j/256 the nearest table point (j = 128..256),
// so f = m - F is exact (|f| <= 1/512); u = f RN(1/F) (|u| <= 1/256) and
//   log x = e ln2 + log F + log1p(u),
// log1p(u) = u + u^2 (-1/2 + u/3 - ... + u^5/7) (truncation below 2e-18
// relative).  log F comes from mathtab.h as hi + lo with hi a multiple of
// 2^-42, so e ln2_hi + hi is exact; at F = 1/2 the table holds exactly
// -(ln2_hi + ln2_lo), so just above x = 1 (e = 1, F = 1/2) the two terms
// cancel with no rounding and log x is log1p(u) of u = 2m - 1 = x - 1.
// Within 1.25 ulp.  Every operation is an IEEE +, -, *, fma, rint or an exact
// frexp, and the table holds the same numbers on host and device (LDS,
// filled by mathtab_fill), so the two builds agree bit for bit.  Each step
// runs over all N values before the next one, which gives the scheduler N
// independent dependency chains to interleave.  Rows are 4 doubles (32 B):
// a row's address is one shift of j.
#if defined(__HIP_DEVICE_COMPILE__)
__shared__ __attribute__((aligned(16))) double s_mathtab[kLogTabN * 4 + kExpTabN * 2];
#endif

// Every workgroup of a kernel that evaluates a log or an exp calls this
// before the first one (all threads, before any of them returns).  The
// tables are copied in 16-byte pieces, a thread's first three loads issued
// before its stores, so that the copy waits for memory once.
__device__ inline void mathtab_fill() {
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr int n = (kLogTabN * 4 + kExpTabN * 2) / 2;
  const double2 *src = reinterpret_cast<const double2 *>(kMathTab);
  double2 *dst = reinterpret_cast<double2 *>(s_mathtab);
  const int i0 = threadIdx.x, bs = blockDim.x;
  // pieces past the table copy its last piece again (no branch between the
  // loads: the stores of one value to one place)
  double2 v[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = src[min(i0 + k * bs, n - 1)];
#pragma unroll
  for (int k = 0; k < 3; ++k) dst[min(i0 + k * bs, n - 1)] = v[k];
  for (int i = i0 + 3 * bs; i < n; i += bs) dst[i] = src[i];
  __syncthreads();
#endif
}

__host__ __device__ inline const double *logtab() {
#if defined(__HIP_DEVICE_COMPILE__)
  return s_mathtab;
#else
  return kMathTab;
#endif
}

__host__ __device__ inline const double *exptab() {
#if defined(__HIP_DEVICE_COMPILE__)
  return s_mathtab + kExpTabOff;
#else
  return kMathTab + kExpTabOff;
#endif
}

template <int N>
__host__ __device__ inline DV<N> log_pos_v(const DV<N> &x) {
  DV<N> u, hi, lo, out;
  int e[N];
  const double *tab = logtab();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double m = frexp(x.v[i], &e[i]);  // [1/2, 1)
    const double jd = rint(m * 256.0);       // 128..256
    const double f = fma(jd, -0.00390625, m);  // m - j/256, exact
    const double *t = tab + ((int)jd - kLogTabJ0) * 4;
    u.v[i] = f * t[0];
    hi.v[i] = t[1];
    lo.v[i] = t[2];
  }
  DV<N> p;
#pragma unroll
  for (int i = 0; i < N; ++i) p.v[i] = 1.0 / 7.0;
  constexpr double kP[5] = {-1.0 / 6.0, 1.0 / 5.0, -1.0 / 4.0, 1.0 / 3.0, -1.0 / 2.0};
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int i = 0; i < N; ++i) p.v[i] = fma_k(p.v[i], u.v[i], kP[k]);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double poly = fma(u.v[i] * u.v[i], p.v[i], u.v[i]);
    const double de = (double)e[i];
    out.v[i] = fma(de, kLn2Hi, hi.v[i]) + (poly + fma(de, kLn2Lo, lo.v[i]));
  }
  return out;
}

// exp(z) for |z| < 5e6 (n fits an int; beyond |z| = 745 the result is 0 or
// +inf), N values in lockstep, table-driven (the method of Tang, ACM TOMS 15 (1989) 144-157):
// n = RN(z 256 / ln2), z = n ln2/256 + r with |r| <= ln2/512 (ln2/256 split
// as ln2_hi/256 + ln2_lo/256; fma keeps r's rounding to its own), and
//   exp(z) = 2^(n >> 8) 2^((n & 255)/256) exp(r),
// 2^(j/256) from mathtab.h as hi + lo, exp(r) - 1 by its Taylor series
// through r^4 (truncation below 4e-17 relative), the sum formed as
// hi + (lo + hi (exp(r) - 1)); ldexp gives 0 and +inf beyond the double
// range.  Within 1 ulp.  The chain's arguments y log x stay within |z| < 1e4
// (x >= 5e-324, y <= 10).
template <int N>
__host__ __device__ inline DV<N> exp_fin_v(const DV<N> &zin) {
  constexpr double kInvLn2x256 = 0x1.71547652b82fep+8;  // RN(256 / ln 2)
  DV<N> r, th, tl, q, out;
  int sc[N];
  const double *tab = exptab();
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double z = zin.v[i];
    const double n = rint(z * kInvLn2x256);
    r.v[i] = fma(-n, kLn2Lo * 0.00390625, fma(-n, kLn2Hi * 0.00390625, z));
    const int ni = (int)n;
    sc[i] = ni >> 8;  // floor(n / 256)
    const double *t = tab + (ni & 255) * 2;
    th.v[i] = t[0];
    tl.v[i] = t[1];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) q.v[i] = fma_k(r.v[i], 1.0 / 24.0, 1.0 / 6.0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    q.v[i] = fma(q.v[i], r.v[i], 0.5);
    const double pm = fma(r.v[i] * r.v[i], q.v[i], r.v[i]);  // exp(r) - 1
    out.v[i] = ldexp(th.v[i] + fma(th.v[i], pm, tl.v[i]), sc[i]);
  }
  return out;
}

// x**y for x > 0 as exp(y log x), N pairs at once: within (2 + 3 |y log x|)
// ulp of the exact power (log's error and the rounding of y log x, amplified
// by exp); for the chain's arguments |y log x| stays of order 10, a few
// 1e-15 relative at most.
template <int N>
__host__ __device__ MHMHIP_OUTLINE DV<N> exp_ylogx_v(DV<N> y, DV<N> x) {
  DV<N> l = log_pos_v<N>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) l.v[i] = y.v[i] * l.v[i];
  return exp_fin_v<N>(l);
}

__host__ __device__ inline double log_pos(double x) { return log_pos_v<1>(DV<1>{{x}}).v[0]; }
__host__ __device__ inline double exp_fin(double z) { return exp_fin_v<1>(DV<1>{{z}}).v[0]; }
__host__ __device__ inline double exp_ylogx(double y, double x) {
  return exp_ylogx_v<1>(DV<1>{{y}}, DV<1>{{x}}).v[0];
}

// x**(2/3) for finite x > 0 (the canopy's (I / Imax)**(2/3),
// mo_canopy_interc.f90:121), without a log and an exp: x = m 2^e with
// e = 3q + r (r in 0..2), a = m 2^r in [1/2, 4), y = a^(-1/3) from a
// single-precision seed 2^(-log2(a) / 3) (relative error below 2^-20),
// refined by one Halley step written as a series in the residual
// r = 1 - a y^3 (the binomial series of (1 - r)^(-1/3)),
//   y' = y + y r (1/3 + 2 r / 9)      (error 14 r^3 / 81, 2^-62, before rounding),
// and x**(2/3) = a y' 2^(2q): no division.  Within 3 ulp of the exact power
// (against mpmath in the tests), exact on exact cubes such as powers of 8,
// 27 and 125.  The seed comes from the hardware's single-precision log2 /
// exp2 on the device and from libm on the host, so the two builds may
// differ in the last bit.
__host__ __device__ inline double pow23(double x) {
  int e;
  const double m = frexp(x, &e);  // [1/2, 1)
  // q = floor(e / 3) and e - 3q for the exponents of finite doubles: with
  // u = e + 1200 in [126, 2224], floor(u / 3) = (u * 43691) >> 17 (exact for
  // u < 2^16), so both are 24-bit products (single-pass on the device, where
  // a 32-bit divide by 3 takes multi-pass multiplies)
  const uint32_t u = (uint32_t)(e + 1200);
#if defined(__HIP_DEVICE_COMPILE__)
  const int q = (int)(__umul24(u, 43691u) >> 17) - 400;
  const int er = e - __mul24(q + 400, 3) + 1200;
#else
  const int q = (int)((u * 43691u) >> 17) - 400;
  const int er = e - 3 * (q + 400) + 1200;
#endif
  const double a = ldexp(m, er);
#if defined(__HIP_DEVICE_COMPILE__)
  const double y =
      (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)a) * (-1.0f / 3.0f));
#else
  const double y = (double)exp2f(log2f((float)a) * (-1.0f / 3.0f));
#endif
  const double r = fma(-a, (y * y) * y, 1.0);
  const double y1 = fma(y * r, fma(r, 2.0 / 9.0, 1.0 / 3.0), y);
  return ldexp(a * y1, 2 * q);
}

// x / d for a divisor whose correctly rounded reciprocal r = RN(1/d) is known:
// q = x r, then one correction with the exact residual e = x - q d (fma):
// RN(q + e r) is the correctly rounded quotient (Markstein, IBM J. Res. Dev.
// 34 (1990) 111-119; finite, normal operands and quotient), i.e. the same
// bits as x / d in three operations instead of a full division.
__host__ __device__ inline double div_rcp(double x, double d, double r) {
  const double q = x * r;
  return fma(fma(-q, d, x), r, q);
}

// pow(x, y) of the reference for y > 0 and the x > 0 power already in hand:
// pow's value for non-integer y when x <= 0 (0 for x == 0, NaN below)
__device__ __forceinline__ double pow_pos_from(double x, double px) {
  if (x > 0.0) return px;
  return x == 0.0 ? 0.0 : __builtin_nan("");
}

__device__ __forceinline__ double pow_pos(double x, double y) {
  return pow_pos_from(x, x > 0.0 ? exp_ylogx(y, x) : 0.0);
}

// FORCES mo_utils LE: a <= b within eps*|b| (pre-proc/mtCLIM/mo_utils.f90:922-934)
__device__ __forceinline__ bool le_eps(double a, double b) {
  return !(((DBL_EPSILON * fabs(b) - fabs(a - b)) < 0.0) && (a > b));
}

// src/mHM/mo_canopy_interc.f90:75-133, in two parts around the power
// (interc / interc_max)**(2/3): canopy_fill updates the storage and the
// throughfall and returns the ratio (1.0 when interc_max <= eps, where the
// reference takes no power); canopy_evap finishes with the ratio's power.
// (rmax = RN(1/interc_max): the ratio as div_rcp, same bits)
// Several of the reference's two-way clamps are a minimum or a maximum:
//   if (a >= b) then d = a - b; s = b else d = 0; s = a   ->  d = max(a - b, 0), s = min(a, b)
//   if (s > e) then s = s - e else e = s; s = 0          ->  e = min(e, s), s = s - e
// They give the same value for every non-NaN operand (a zero may come out
// with the other sign, which no later operation of the chain tells apart).
// On the device they are the min / max instructions themselves: written as
// fmin / fmax, the compiler first quiets a possible signalling NaN in each
// operand it cannot prove canonical (an extra v_max_f64 per loaded value);
// the chain holds no signalling NaN, and for every other operand the result
// is the same.
__device__ __forceinline__ double vmin(double a, double b) {
  double d;
  asm("v_min_f64 %0, %1, %2" : "=v"(d) : "v"(a), "v"(b));
  return d;
}
__device__ __forceinline__ double vmax(double a, double b) {
  double d;
  asm("v_max_f64 %0, %1, %2" : "=v"(d) : "v"(a), "v"(b));
  return d;
}
// max(a, 0) with the zero an inline constant, and max(a, k) for a constant k
// passed in a scalar register pair
__device__ __forceinline__ double vmax0(double a) {
  double d;
  asm("v_max_f64 %0, %1, 0" : "=v"(d) : "v"(a));
  return d;
}
__device__ __forceinline__ double vmax_k(double a, double k) {
  double d;
  asm("v_max_f64 %0, %1, %2" : "=v"(d) : "v"(a), "s"(k));
  return d;
}

__device__ __forceinline__ double canopy_fill(double interc_max, double rmax, double precip,
                                              double &interc, double &throughfall) {
  const double aux = interc + precip;
  throughfall = vmax0(aux - interc_max);
  interc = vmin(aux, interc_max);
  return interc_max > kEps ? div_rcp(interc, interc_max, rmax) : 1.0;
}

// ratio: canopy_fill's result; pw: ratio**(2/3) when ratio > 0
__device__ __forceinline__ void canopy_evap(double pet, double interc_max, double ratio, double pw,
                                            double &interc, double &evap_canopy) {
  if (interc_max > kEps)
    evap_canopy = pet * pow_pos_from(ratio, pw);
  else
    evap_canopy = 0.0;
  evap_canopy = vmin(vmax0(evap_canopy), interc);
  interc = interc - evap_canopy;
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
  const bool warm = temp > tth;
  snow = warm ? 0.0 : thr;
  rain = thr - snow;  // thr (warm) or +0 (cold), as the reference's branches
  if (prec <= ddthr)
    deg_day = dnp + incr * prec;
  else
    deg_day = dmax;
  // warm with a pack: melt = aux unless aux > pack (then the whole pack)
  const double m = vmin(deg_day * (temp - tth), pack);
  if (warm) {
    if (pack > 0.0) {
      melt = m;
      pack = pack - m;
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
// (rfw = RN(1/(fc - wp)): the division as div_rcp, same bits; it is only
// reached with fc > sm > wp)
__device__ __forceinline__ double feddes(double sm, double fc, double wp, double fr, double rfw) {
  if (sm >= fc) return fr;
  if (sm > wp) return div_rcp(fr * (sm - wp), fc - wp, rfw);
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

// src/mHM/mo_soil_moisture.f90:94-290.  frac_pw[h] = (sm[h] / sat[h])**beta[h]
// of the start-of-step moisture (the reference's infiltration fraction,
// line 219), evaluated for all horizons before the horizon loop.
template <int NH, bool JV>
__device__ __forceinline__ void soil_moisture(double fS, double thrS, double rthrS, double pet,
                                              double evap, double revap, const double *sat,
                                              const double *fr,
                                              const double *fc, const double *wp,
                                              const double *rfw, double c1,
                                              double ecan, double peff, const double *frac_pw,
                                              double &runoff_sealed, double &storage_sealed,
                                              double *infil, double *sm, double *aet,
                                              double &aet_sealed) {
  runoff_sealed = 0.0;
  aet_sealed = 0.0;
  if (fS > 0.0) {
    const double tmp = storage_sealed + peff;
    runoff_sealed = vmax0(tmp - thrS);
    storage_sealed = vmin(tmp, thrS);
    if (thrS > kEps) {
      // revap = RN(1/evap): pet / evap as div_rcp, same bits (evap > 0)
      const double pe = div_rcp(pet, evap, revap);
      aet_sealed = vmax0((pe - ecan) * div_rcp(storage_sealed, thrS, rthrS));
    } else {
      aet_sealed = DBL_MAX;  // huge(1.0_dp)
    }
    aet_sealed = vmin(aet_sealed, storage_sealed);
    storage_sealed = storage_sealed - aet_sealed;
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
        frac = frac_pw[h];  // (sm[h] / sat[h])**bexp[h], start-of-step sm[h]
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
      for (int j = 0; j < h; ++j) s = s + vmax0(aet[j]);  // adds only the positive ones
      a = a - s;
    }
    double f;
    if constexpr (JV)
      f = jarvis(sm[h], sat[h], wp[h], fr[h], c1);
    else
      f = feddes(sm[h], fc[h], wp[h], fr[h], rfw[h]);
    a = vmax0(a * f);
    if (sm[h] > a) {
      sm[h] = sm[h] - a;
    } else {
      a = sm[h] - kEps;
      sm[h] = kEps;
    }
    sm[h] = vmax_k(sm[h], kEps);
    aet[h] = a;
  }
}

// src/mHM/mo_runoff.f90:74-154 (k0, k1, kp already c2TSTu / k)
// in two parts around the power unsat**(1 + alpha) of the slow interflow:
// runoff_fast adds the infiltration and takes the fast interflow (the store
// is then the power's base), runoff_slow_perc finishes with the power pw
// (of unsat when unsat > 0).
__device__ __forceinline__ void runoff_fast(double k0, double pin, double thr, double &unsat,
                                            double &fast) {
  unsat = unsat + pin;
  fast = 0.0;
  if (unsat > thr) fast = vmin(unsat - kEps, k0 * (unsat - thr));
  unsat = unsat - fast;
}

__device__ __forceinline__ void runoff_slow_perc(double k1, double kp, double karst, double pw,
                                                 double &sat, double &unsat, double &slow,
                                                 double &perc) {
  slow = 0.0;
  if (unsat > kEps) slow = vmin(unsat - kEps, k1 * pow_pos_from(unsat, pw));
  unsat = unsat - slow;
  perc = kp * unsat;
  const double p = vmin(perc, unsat);  // the whole store when perc exceeds it
  sat = sat + p * karst;
  unsat = unsat - p;
}

__device__ __forceinline__ void runoff_unsat_zone(double k1, double kp, double k0, double alpha,
                                                  double karst, double pin, double thr,
                                                  double &sat, double &unsat, double &slow,
                                                  double &fast, double &perc) {
  runoff_fast(k0, pin, thr, unsat, fast);
  runoff_slow_perc(k1, kp, karst, unsat > 0.0 ? exp_ylogx(1.0 + alpha, unsat) : 0.0, sat, unsat,
                   slow, perc);
}

// src/mHM/mo_runoff.f90:191-212
__device__ __forceinline__ void runoff_sat_zone(double k2, double &sat, double &baseflow) {
  const double s0 = vmax0(sat);  // an empty or negative store drains nothing, ends at 0
  baseflow = k2 * s0;
  sat = s0 - baseflow;
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
// sin(omega) of omega = acos(arg) is sqrt((1 - arg)(1 + arg)), the same
// value within a few ulp (the factored form keeps its relative accuracy as
// arg nears +-1), for a square root instead of a sine's range reduction and
// series.
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
  const double sin_omega = sqrt((1.0 - arg) * (1.0 + arg));
  const double ra = kRadConst * dr * (omega * sinlat * sind + coslat * cosd * sin_omega);
  return coeff * ra * (tavg + kHarSamConst) * sqrt(dT);
}

// FORCES mo_constants: DaySecs, SpecHeatET_dp, Psychro_dp, rho0_dp, cp0_dp, T0_dp;
// src/mHM/mo_mhm_constants.f90:44-48: tetens_c1..c3, satpressureslope1
constexpr double kDaySecs = 86400.0;
constexpr double kSpecHeatET = 2.45e06;
constexpr double kPsychro = 0.0646;
constexpr double kRho0 = 1.225;
constexpr double kCp0 = 1005.0;
constexpr double kT0 = 273.15;
constexpr double kTetensC1 = 0.6108;
constexpr double kTetensC2 = 17.270;
constexpr double kTetensC3 = 237.30;
constexpr double kSatPressureSlope1 = 4098.0;

// src/mHM/mo_pet.f90:426-441
__device__ __forceinline__ double sat_vap_pressure(double tavg) {
  return kTetensC1 * exp_fin(kTetensC2 * tavg / (tavg + kTetensC3));
}

// src/mHM/mo_pet.f90:384-399: exp(2 log(tavg + c3)) as written (NaN below
// tavg = -c3, as log there)
__device__ __forceinline__ double slope_satpressure(double tavg) {
  const double b = tavg + kTetensC3;
  const double sq = b > 0.0 ? exp_fin(2.0 * log_pos(b)) : (b == 0.0 ? 0.0 : __builtin_nan(""));
  return kSatPressureSlope1 * sat_vap_pressure(tavg) / sq;
}

// src/mHM/mo_pet.f90:149-174 (Rn already max(netrad, 0), mo_meteo_handler.f90:1081-1086)
__device__ MHMHIP_OUTLINE double pet_priestly(double alpha, double rn, double tavg) {
  const double delta = slope_satpressure(tavg);
  return alpha * delta / (kPsychro + delta) * (rn * kDaySecs / kSpecHeatET);
}

// src/mHM/mo_pet.f90:235-272 with the getter's arguments
// (mo_meteo_handler.f90:1088-1095): net_rad = max(netrad, 0), act_vap_pressure =
// absvappress / 1000, aerodyn_resistance = aeroResist / windspeed, a_s = a_sh = 1.
// slope_satpressure is evaluated once (the reference's two calls give one value).
__device__ MHMHIP_OUTLINE double pet_penman(double net_rad, double tavg, double act_vap_pressure,
                                             double aerodyn_resistance,
                                             double bulksurface_resistance) {
  const double a_s = 1.0, a_sh = 1.0;
  const double slope = slope_satpressure(tavg);
  return kDaySecs / kSpecHeatET *
         (slope * net_rad +
          kRho0 * kCp0 * (sat_vap_pressure(tavg) - act_vap_pressure) * a_sh / aerodyn_resistance) /
         (slope + kPsychro * a_sh / a_s * (1.0 + bulksurface_resistance / aerodyn_resistance));
}

// src/meteo/mo_meteo_temporal_tools.f90:32-59: (v + c) * w - c
__device__ __forceinline__ double disagg_weights(double v, double w, double c) {
  return (v + c) * w - c;
}

// src/meteo/mo_meteo_temporal_tools.f90:54-79 with the day or night factor f
// already chosen: 2 * v * f / ntimesteps_day
// (rnT = RN(1/nT): the division by nT as div_rcp, same bits)
__device__ __forceinline__ double disagg_flux(double nT, double rnT, double v, double f) {
  if (nT > 1.0) return div_rcp(2.0 * v * f, nT, rnT);
  return v;
}

// src/meteo/mo_meteo_temporal_tools.f90:100-161 with add_correction = .true.
__device__ __forceinline__ double disagg_state_add(double nT, double v, double f) {
  if (nT > 1.0) return v + f;
  return v;
}

}  // namespace mhmhip
