"""Global parameters (gamma) of the active process groups and a light
multiscale parameter regionalisation ("MPR-lite") for synthetic domains.

The bounds and defaults are those of the reference's mhm_parameter.nml
(lines cited per group).  MPR-lite maps gamma plus land-cover fractions,
soil texture and terrain to the L1 effective parameter fields that mHM()
reads.  Where the reference's transfer function needs only land-cover
fractions it is followed as written (snow: src/MPR/mo_multi_param_reg.f90:800-852;
sealed threshold: 883-894); soil hydraulics, recession constants and the
PET aspect factor use simple documented stand-ins that keep the
reference's orderings (PWP < FC < saturation, kPerco >= 2 d
(mo_multi_param_reg.f90:1017), kBaseFlow >= kSlowFlow (618)).  The hot path
consumes the L1 fields only, so MPR-lite does not enter any parity check.
"""
import numpy as np

# name, min, max, default  (mhm_parameter.nml)
GAMMA = [
    # interception1 (mhm_parameter.nml:7-9)
    ("canopyInterceptionFactor", 0.15, 0.40, 0.15),
    # snow1 (11-20)
    ("snowTreshholdTemperature", -2.0, 2.0, 1.0),
    ("degreeDayFactor_forest", 0.0001, 4.0, 1.5),
    ("degreeDayFactor_impervious", 0.0, 1.0, 0.5),
    ("degreeDayFactor_pervious", 0.0, 2.0, 0.5),
    ("increaseDegreeDayFactorByPrecip", 0.1, 0.9, 0.5),
    ("maxDegreeDayFactor_forest", 0.0, 8.0, 3.0),
    ("maxDegreeDayFactor_impervious", 0.0, 8.0, 3.5),
    ("maxDegreeDayFactor_pervious", 0.0, 8.0, 4.0),
    # soilmoisture (root fractions, infiltration shape, Jarvis)
    ("rootFractionCoefficient_forest", 0.90, 0.999, 0.97),
    ("rootFractionCoefficient_impervious", 0.90, 0.95, 0.93),
    ("rootFractionCoefficient_pervious", 0.001, 0.09, 0.02),
    ("infiltrationShapeFactor", 1.0, 4.0, 1.75),
    ("jarvis_sm_threshold_c1", 0.0, 1.0, 0.5),
    # directRunoff1
    ("imperviousStorageCapacity", 0.0, 5.0, 0.5),
    # PET1 (Hargreaves-Samani)
    ("minCorrectionFactorPET", 0.70, 1.30, 0.93),
    ("maxCorrectionFactorPET", 0.0, 0.20, 0.19),
    ("aspectTresholdPET", 160.0, 200.0, 171.0),
    ("HargreavesSamaniCoeff", 0.0016, 0.0030, 0.0023),
    # interflow1
    ("interflowStorageCapacityFactor", 75.0, 200.0, 85.0),
    ("interflowRecession_slope", 0.0, 10.0, 7.0),
    ("fastInterflowRecession_forest", 1.0, 3.0, 1.5),
    ("slowInterflowRecession_Ks", 1.0, 30.0, 15.0),
    ("exponentSlowInterflow", 0.05, 0.30, 0.125),
    # percolation1
    ("rechargeCoefficient", 0.0, 50.0, 35.0),
    # geoparameter (baseflow recession)
    ("GeoParam", 1.0, 1000.0, 100.0),
    # routing1 (mhm_parameter.nml:193-201)
    ("muskingumTravelTime_constant", 0.31, 0.35, 0.325),
    ("muskingumTravelTime_riverLength", 0.07, 0.08, 0.075),
    ("muskingumTravelTime_riverSlope", 1.95, 2.10, 2.0),
    ("muskingumTravelTime_impervious", 0.09, 0.11, 0.1),
    ("muskingumAttenuation_riverSlope", 0.01, 0.50, 0.3),
]
NAMES = [g[0] for g in GAMMA]
INDEX = {n: i for i, n in enumerate(NAMES)}
LOWER = np.array([g[1] for g in GAMMA])
UPPER = np.array([g[2] for g in GAMMA])
DEFAULT = np.array([g[3] for g in GAMMA])
ROUTING = NAMES[-5:]


def draw(rng, n_sets=1):
    """n_sets gamma vectors drawn uniformly within the namelist bounds."""
    u = rng.random((n_sets, len(GAMMA)))
    return LOWER + u * (UPPER - LOWER)


def routing_gamma(g):
    return np.ascontiguousarray(g[[INDEX[n] for n in ROUTING]], dtype=np.float64)


# Monthly LAI per land-cover class (forest, impervious, pervious): smooth
# seasonal curves standing in for the reference's LAI look-up tables.
_MONTHS = np.arange(12)
LAI = np.stack([
    6.0 + 3.0 * np.sin(2 * np.pi * (_MONTHS - 3) / 12.0),
    0.3 + 0.0 * _MONTHS,
    2.0 + 1.5 * np.sin(2 * np.pi * (_MONTHS - 3) / 12.0),
])


def lai_l1(fractions):
    """(n, 12) monthly leaf area index of cells with these cover fractions."""
    fF, fI, fP = fractions
    return fF[:, None] * LAI[0] + fI[:, None] * LAI[1] + fP[:, None] * LAI[2]


def static_fields(fractions, texture, horizon_depth):
    """The fields of mpr_lite that do not depend on gamma (leaf area index,
    PET-LAI factor, soil water contents per horizon), computed once per
    domain and shared read-only by its parameter sets."""
    lai = lai_l1(fractions)
    hd = np.asarray(horizon_depth, dtype=np.float64)
    top = np.concatenate([[0.0], hd[:-1]])
    dz = hd - top
    poro = 0.38 + 0.12 * texture
    fcr = 0.55 + 0.25 * (1.0 - texture)
    wpr = 0.30 + 0.25 * texture
    sat = dz[None, :] * poro[:, None]
    fc = sat * fcr[:, None]
    wp = fc * wpr[:, None]
    out = {"LAI": lai, "PETLAICOR": np.clip(0.6 + 0.08 * lai, 0.3, 1.3), "SOILMOIST_SAT": sat,
           "SOILMOIST_FC": fc, "WILTINGPOINT": wp, "hd": hd, "KARSTLOSS": np.ones(texture.shape[0])}
    for v in out.values():
        v.setflags(write=False)
    return out


def mpr_lite(g, fractions, texture, aspect, horizon_depth, static=None):
    """L1 effective parameter fields for one gamma vector.

    fractions: (3, n) forest / impervious / pervious cover fractions
    texture:   (n,) in [0, 1), a soil texture index
    aspect:    (n,) terrain aspect in degrees
    horizon_depth: lower boundaries of the horizons [mm]
    static: static_fields of the same inputs (computed here when None)
    Returns a dict of arrays: 1-D (n,), horizon (n, nH) or monthly (n, 12);
    the gamma-independent ones are the read-only arrays of `static`.
    """
    st = static if static is not None else static_fields(fractions, texture, horizon_depth)
    fF, fI, fP = fractions
    n = fF.shape[0]
    G = lambda name: g[INDEX[name]]
    out = {}
    out["FSEALED"] = fI.copy()
    # snow_acc_melt_param (mo_multi_param_reg.f90:800-852)
    p = [G("snowTreshholdTemperature"), G("degreeDayFactor_forest"),
         G("degreeDayFactor_impervious"), G("degreeDayFactor_pervious"),
         G("increaseDegreeDayFactorByPrecip"), G("maxDegreeDayFactor_forest"),
         G("maxDegreeDayFactor_impervious"), G("maxDegreeDayFactor_pervious")]
    ddf_f, ddf_i, ddf_p = p[1], p[1] + p[3] + p[2], p[1] + p[3]
    max_f, max_i, max_p = p[1] + p[5], p[1] + p[3] + p[2] + p[6], p[1] + p[3] + p[7]
    out["TEMP_THRESH"] = np.full(n, p[0])
    out["DEGDAY_INC"] = np.full(n, p[4])
    out["DEGDAY_NOPRE"] = ddf_f * fF + ddf_i * fI + ddf_p * fP
    out["DEGDAY_MAX"] = max_f * fF + max_i * fI + max_p * fP
    # iper_thres_runoff (883-894)
    out["SEALED_THRESH"] = np.full(n, G("imperviousStorageCapacity"))
    # interception: LAI * factor
    out["MAXINTER"] = G("canopyInterceptionFactor") * st["LAI"]
    out["PETLAICOR"] = st["PETLAICOR"]
    # soil hydraulics per horizon
    hd = st["hd"]
    nH = hd.shape[0]
    out["SOILMOIST_SAT"] = st["SOILMOIST_SAT"]
    out["SOILMOIST_FC"] = st["SOILMOIST_FC"]
    out["WILTINGPOINT"] = st["WILTINGPOINT"]
    out["SOILMOIST_EXP"] = np.repeat((G("infiltrationShapeFactor") * (1.0 + 0.3 * texture))[:, None],
                                     nH, axis=1)
    # root fractions: cumulative 1 - c**(z/10 mm), class-weighted, normalised
    c = (fF * G("rootFractionCoefficient_forest") + fI * G("rootFractionCoefficient_impervious")
         + fP * G("rootFractionCoefficient_pervious"))
    c = np.clip(c, 1e-3, 0.999)
    cum = 1.0 - c[:, None] ** (hd[None, :] / 10.0)
    cum_top = np.concatenate([np.zeros((n, 1)), cum[:, :-1]], axis=1)
    fr = np.maximum(cum - cum_top, 0.0)
    fr = fr / np.maximum(fr.sum(axis=1, keepdims=True), 1e-12)
    out["FROOTS"] = fr
    out["JARVIS_C1"] = np.full(n, G("jarvis_sm_threshold_c1"))
    # recession constants [d]
    out["KFASTFLOW"] = G("fastInterflowRecession_forest") * (1.0 + 0.5 * (1.0 - fF))
    out["KSLOWFLOW"] = G("slowInterflowRecession_Ks") * (0.5 + texture)
    out["KPERCO"] = np.maximum(G("rechargeCoefficient") * (0.5 + texture), 2.0)
    out["KBASEFLOW"] = np.maximum(G("GeoParam") * (0.5 + 0.5 * texture), out["KSLOWFLOW"])
    out["ALPHA"] = np.full(n, G("exponentSlowInterflow"))
    out["KARSTLOSS"] = st["KARSTLOSS"]
    out["UNSAT_THRESH"] = G("interflowStorageCapacityFactor") * (0.2 + 0.3 * texture)
    # PET aspect correction
    thr = G("aspectTresholdPET")
    mn, mx = G("minCorrectionFactorPET"), G("maxCorrectionFactorPET")
    out["FASP"] = np.where(aspect < thr, mn + mx * (thr - aspect) / thr,
                           mn + mx * (aspect - thr) / (360.0 - thr))
    out["HARSAMCOEFF"] = np.full(n, G("HargreavesSamaniCoeff"))
    return out


def pet_radiation_params(lai, pet_case):
    """Monthly PET parameters of the radiation-based cases, stand-ins shaped
    like the reference's fields: Priestley-Taylor's alpha (L1_PrieTayAlpha,
    process_pet 2) around 1.26; Penman-Monteith's aerodynamic resistance
    per unit wind speed (L1_aeroResist, divided by the day's wind speed in
    get_corrected_pet) around 208 s m-1 and bulk surface resistance
    (L1_surfResist) around 70 s m-1, both falling with the leaf area.
    lai: (n, 12) monthly leaf area index."""
    out = {}
    if pet_case == 2:
        out["PRIETAYALPHA"] = 1.05 + 0.06 * lai
    elif pet_case == 3:
        out["AERORESIST"] = 208.0 / (0.6 + 0.1 * lai)
        out["SURFRESIST"] = 40.0 + 120.0 / (1.0 + lai)
    return out
