/*
 * mhmhip.h -- C-ABI of the MI355X (gfx950) implementation of mHM's per-step
 * hot path: the meteo getters + per-L1-cell process chain of mHM() and the
 * mRM Muskingum routing.
 *
 * The boundary is plain C: pointers, sizes and status codes, no device or
 * framework types.  Host arrays follow the reference's Fortran layout
 * (column-major, cells contiguous, 1-based network indices), so a Fortran
 * caller passes c_loc(X(s1,1)) and size(X,1) through iso_c_binding.
 *
 * Threading: one context per host thread.  Calls on one context are
 * serialised; kernels run asynchronously on the context's own stream and
 * every entry point that returns host data synchronises that stream.
 *
 * Errors: every entry point returns MHMHIP_OK (0) or a negative code;
 * mhmhip_last_error() gives the message.  The reference has no return codes
 * (FORCES error_message aborts); the Fortran shim turns a negative status
 * into error_message, which keeps the reference's abort semantics.
 */
#ifndef MHMHIP_H
#define MHMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHMHIP_OK 0
#define MHMHIP_EINVAL (-1) /* bad argument or shape                        */
#define MHMHIP_EHIP (-2)   /* HIP runtime error (no device, launch failure) */
#define MHMHIP_ESTATE (-3) /* call out of order (e.g. step before bind)     */
#define MHMHIP_ENOMEM (-4) /* device memory exhausted                       */
#define MHMHIP_PENDING 1   /* mhmhip_mrm_step: runoff accumulated, not routed at this step */

#define MHMHIP_MAX_HORIZONS 6

/* Cell variables (mhmhip_set_cell_var / mhmhip_get_cell_var).  ncols is 1,
 * nH (horizon arrays) or 12 (monthly arrays, indexed by iLAI = month). */
enum mhmhip_var {
  /* L1 effective parameters, the MPR output read by mHM() and the meteo
   * getters (src/mHM/mo_mhm_interface_run.f90:366-457) -- set only */
  MHMHIP_P_FSEALED = 0,   /* L1_fSealed(:,1,yId)                  */
  MHMHIP_P_DEGDAY_INC,    /* L1_degDayInc                          */
  MHMHIP_P_DEGDAY_MAX,    /* L1_degDayMax                          */
  MHMHIP_P_DEGDAY_NOPRE,  /* L1_degDayNoPre                        */
  MHMHIP_P_TEMP_THRESH,   /* L1_tempThresh                         */
  MHMHIP_P_SEALED_THRESH, /* L1_sealedThresh                       */
  MHMHIP_P_JARVIS_C1,     /* L1_jarvis_thresh_c1                   */
  MHMHIP_P_KFASTFLOW,     /* L1_kFastFlow  (k0) [d]                */
  MHMHIP_P_KSLOWFLOW,     /* L1_kSlowFlow  (k1) [d]                */
  MHMHIP_P_KBASEFLOW,     /* L1_kBaseFlow  (k2) [d]                */
  MHMHIP_P_KPERCO,        /* L1_kPerco     (kp) [d]                */
  MHMHIP_P_ALPHA,         /* L1_alpha                              */
  MHMHIP_P_KARSTLOSS,     /* L1_karstLoss                          */
  MHMHIP_P_UNSAT_THRESH,  /* L1_unsatThresh                        */
  MHMHIP_P_FASP,          /* L1_fAsp                               */
  MHMHIP_P_HARSAMCOEFF,   /* L1_HarSamCoeff                        */
  MHMHIP_P_LATITUDE,      /* level1%y packed [deg]                 */
  MHMHIP_P_SOILMOIST_SAT, /* L1_soilMoistSat  (nH)                 */
  MHMHIP_P_FROOTS,        /* L1_fRoots        (nH)                 */
  MHMHIP_P_SOILMOIST_FC,  /* L1_soilMoistFC   (nH)                 */
  MHMHIP_P_WILTINGPOINT,  /* L1_wiltingPoint  (nH)                 */
  MHMHIP_P_SOILMOIST_EXP, /* L1_soilMoistExp  (nH)                 */
  MHMHIP_P_MAXINTER,      /* L1_maxInter      (12)                 */
  MHMHIP_P_PETLAICOR,     /* L1_petLAIcorFactor (12)               */
  MHMHIP_P_PRIETAYALPHA,  /* L1_PrieTayAlpha (12, process_pet 2)   */
  MHMHIP_P_AERORESIST,    /* L1_aeroResist   (12, process_pet 3)   */
  MHMHIP_P_SURFRESIST,    /* L1_surfResist   (12, process_pet 3)   */
  MHMHIP_P_COUNT,
  /* states (src/mHM/mo_global_variables.f90) -- set and get */
  MHMHIP_S_INTER = 32,    /* L1_inter                              */
  MHMHIP_S_SNOWPACK,      /* L1_snowPack                           */
  MHMHIP_S_SEALSTW,       /* L1_sealSTW                            */
  MHMHIP_S_UNSATSTW,      /* L1_unsatSTW                           */
  MHMHIP_S_SATSTW,        /* L1_satSTW                             */
  MHMHIP_S_SOILMOIST,     /* L1_soilMoist (nH)                     */
  MHMHIP_S_END,
  /* fluxes of the last step -- get only */
  MHMHIP_X_PET_CALC = 64, /* L1_pet_calc                           */
  MHMHIP_X_TEMP_CALC,     /* L1_temp_calc                          */
  MHMHIP_X_PREC_CALC,     /* L1_prec_calc                          */
  MHMHIP_X_THROUGHFALL,   /* L1_Throughfall                        */
  MHMHIP_X_AETCANOPY,     /* L1_aETCanopy                          */
  MHMHIP_X_DEGDAY,        /* L1_degDay                             */
  MHMHIP_X_MELT,          /* L1_melt                               */
  MHMHIP_X_PREEFFECT,     /* L1_preEffect                          */
  MHMHIP_X_RAIN,          /* L1_rain                               */
  MHMHIP_X_SNOW,          /* L1_snow                               */
  MHMHIP_X_RUNOFFSEAL,    /* L1_runoffSeal                         */
  MHMHIP_X_AETSEALED,     /* L1_aETSealed                          */
  MHMHIP_X_FASTRUNOFF,    /* L1_fastRunoff                         */
  MHMHIP_X_SLOWRUNOFF,    /* L1_slowRunoff                         */
  MHMHIP_X_PERCOL,        /* L1_percol                             */
  MHMHIP_X_BASEFLOW,      /* L1_baseflow                           */
  MHMHIP_X_TOTAL_RUNOFF,  /* L1_total_runoff                       */
  MHMHIP_X_INFILSOIL,     /* L1_infilSoil (nH)                     */
  MHMHIP_X_AETSOIL,       /* L1_aETSoil   (nH)                     */
  MHMHIP_X_END
};

/* Daily (or hourly) forcing variables: L1_pre, L1_temp, L1_tmin, L1_tmax,
 * L1_pet, L1_netrad [W m-2], L1_absvappress [Pa], L1_windspeed [m s-1] of
 * the meteo handler (src/meteo/mo_meteo_handler.f90:120-160).  A context
 * holds the ones its PET case reads (get_corrected_pet, 1058-1097): pre and
 * tavg always; pet for process_pet -1 and 0; tmin and tmax for 1 (Hargreaves-
 * Samani); netrad for 2 (Priestley-Taylor); netrad, absvappress and windspeed
 * for 3 (Penman-Monteith).  The others are accepted by the setters and not
 * stored. */
enum mhmhip_forcing { MHMHIP_F_PRE = 0, MHMHIP_F_TAVG, MHMHIP_F_TMIN, MHMHIP_F_TMAX,
                      MHMHIP_F_PET, MHMHIP_F_NETRAD, MHMHIP_F_ABSVAPPRESS,
                      MHMHIP_F_WINDSPEED, MHMHIP_F_COUNT };

typedef struct mhmhip_ctx mhmhip_ctx;

/* Model configuration of one context: the process switches, time step and
 * the namelist tables the hot path reads. */
typedef struct {
  int64_t n_cells;        /* L1 cells over all packed domains (sum of nCells1)   */
  int32_t n_domains;      /* domains packed in this context (>= 1)              */
  int32_t n_horizons;     /* nSoilHorizons_mHM, 1..MHMHIP_MAX_HORIZONS          */
  int32_t n_sets;         /* parameter sets run side by side on shared forcing  */
  int32_t timestep;       /* model time step [h], a divisor of 24               */
  int32_t process_soil;   /* processMatrix(3,1): 1,4 Feddes; 2,3 Jarvis         */
  int32_t process_pet;    /* processMatrix(5,1): -1, 0, 1, 2 or 3               */
  int32_t process_route;  /* processMatrix(8,1): 0 (off) or 1 (Muskingum)       */
  int32_t hourly_forcing; /* meteo_handler%is_hourly_forcing                    */
  int32_t n_forcing_days; /* forcing columns held on the device                 */
  int32_t device;         /* HIP device ordinal                                 */
  const int64_t *domain_cell_start; /* n_domains+1 offsets into cells (NULL: 1 domain) */
  double evap_coeff[12];  /* panEvapo evap_coeff (mhm.nml:788)                  */
  double fnight_prec[12]; /* nightDayRatio (mhm.nml:794-813)                    */
  double fnight_pet[12];
  double fnight_temp[12];
} mhmhip_config;

/* Per-step scalars the reference derives on the host before mHM():
 * tt (run_cfg%time_step), month/hour/doy from dec2date(newTime - 0.5),
 * iMeteoTS from meteo_handler%update_timestep, read_restart. */
typedef struct {
  int32_t tt;          /* 1-based step counter                          */
  int32_t month;       /* 1..12                                         */
  int32_t hour;        /* 0..23                                         */
  int32_t doy;         /* 1..366                                        */
  int32_t iMeteoTS;    /* 1-based forcing day (iMeteoTS of the run)     */
  int32_t read_states; /* read_restart                                  */
  int32_t n_steps;     /* run_cfg%domainDateTime%nTimeSteps (0: unknown);
                          ends the last routing period early when routing
                          steps are longer than the model step          */
} mhmhip_step;

/* The L11 network of every packed domain, concatenated like the
 * reference's module arrays (src/mRM/mo_mrm_global_variables.f90:190-218):
 * per-domain slices with per-domain 1-based indices.  The routing grid is
 * the L1 grid (L1_L11_Id(k) == k within each domain) or coarser (map_flag
 * = .true.: each L11 node's qOUT aggregates its L1 cells' runoff,
 * mo_mrm_pre_routing.f90:110-126); a routing grid finer than L1
 * (map_flag = .false.) is refused, and so is mhmhip_run_noalloc on a
 * coarser one. */
typedef struct {
  int64_t n_nodes;               /* total L11 nodes (1..n_cells)                 */
  const int64_t *domain_node_start; /* n_domains+1 offsets (NULL: one domain, or
                                       the cell ranges when n_nodes == n_cells)  */
  const int32_t *n_outlets;      /* per domain L11_nOutlets                      */
  const int32_t *netPerm;        /* per domain nLinks entries, 1-based, packed   */
  const int32_t *fromN;          /* like netPerm                                 */
  const int32_t *toN;
  const double *length;          /* L11_length [m], per link                     */
  const double *slope;           /* L11_slope, per link                          */
  const double *fFPimp;          /* L11_nLinkFracFPimp, per link                 */
  const double *L1_area;         /* level1%CellArea * 1e-6 [km2], per cell       */
  const int32_t *L1_L11_Id;      /* per cell, 1-based within its domain          */
  int32_t n_gauges;              /* gauges over all domains                      */
  int32_t n_inflow;              /* inflow gauges over all domains               */
  const int32_t *gauge_domain;   /* 0-based domain of each gauge                 */
  const int32_t *gauge_node;     /* 1-based node within that domain              */
  const int32_t *inflow_domain;  /* 0-based domain of each inflow gauge          */
  const int32_t *inflow_node;    /* 1-based node within that domain              */
  const int32_t *inflow_headwater; /* InflowGaugeHeadwater (0/1)                 */
} mhmhip_network;

/* Options of mhmhip_run. */
typedef struct {
  int32_t mode;          /* MHMHIP_RUN_PLAIN or MHMHIP_RUN_FUSED                 */
  int32_t steps_per_launch; /* fused path: S (0 = 24)                          */
  int32_t block;         /* cell-kernel block size 64/128/256 (0 = 256 fused,
                            128 plain)                                          */
  int32_t use_graph;     /* capture the tile sequence in a hipGraph              */
  int32_t route_mode;    /* 0 auto (= 3), 1 grid diagonals, 2 workgroup-resident
                            trees, 3 tiles (windows of up to 64 steps per launch) */
  int32_t route_window;  /* tiled routing: steps per window, 1..64 (0 = 8); the
                            runoff ring holds (tile depth + 1) windows + S + 1
                            steps (at most the call's steps, rounded up to
                            whole windows) + 16 rows                          */
  int32_t reserved[2];
} mhmhip_run_opts;

#define MHMHIP_RUN_PLAIN 0 /* one launch per step, every flux written each step */
#define MHMHIP_RUN_FUSED 1 /* S steps per launch, states held in registers      */

/* ---- lifecycle ------------------------------------------------------------ */

/* Replaces the allocation part of mhm_initialize / variables_alloc
 * (src/mHM/mo_startup.f90:68-122, src/mHM/mo_init_states.f90:49). */
int mhmhip_create(const mhmhip_config *cfg, mhmhip_ctx **out);
int mhmhip_destroy(mhmhip_ctx *ctx);
/* Message of the last failed call on ctx (ctx NULL: last failed create). */
const char *mhmhip_last_error(const mhmhip_ctx *ctx);
/* Bytes of device memory held by ctx (admission control). */
int64_t mhmhip_device_bytes(const mhmhip_ctx *ctx);
/* Joins the context's stream. */
int mhmhip_sync(mhmhip_ctx *ctx);

/* ---- cell data -------------------------------------------------------------
 * host points at X(1,1) of a Fortran array X(ld, ncols) holding n_cells rows
 * of parameter set `set`.  Replaces the module-array reads/writes of mHM()
 * (src/mHM/mo_mhm.f90:229-242) and get.L1_variable (pybind/src/wrapper.f90:569-661). */
int mhmhip_set_cell_var(mhmhip_ctx *ctx, int32_t var, int32_t set, const double *host,
                        int64_t ld, int32_t ncols);
int mhmhip_get_cell_var(mhmhip_ctx *ctx, int32_t var, int32_t set, double *host, int64_t ld,
                        int32_t ncols);
/* fluxes_states_default_init (src/mHM/mo_init_states.f90:258-345) for every set. */
int mhmhip_default_states(mhmhip_ctx *ctx, const double *horizon_depth);

/* Forcing columns day0..day0+ndays-1 (1-based) of one variable, host X(ld, ndays).
 * Replaces the meteo handler's L1_* arrays and the set_meteo seam
 * (src/meteo/mo_meteo_handler.f90:1462-1605). */
int mhmhip_set_forcing(mhmhip_ctx *ctx, int32_t var, int32_t day0, int32_t ndays,
                       const double *host, int64_t ld);
/* Streamed forcing: days day0..day0+ndays-1 (1-based iMeteoTS, any day of the
 * run) of the variables host[MHMHIP_F_COUNT] points at (NULL: not given; its
 * columns keep their old values), each X(ld, ndays) of all cells, uploaded
 * asynchronously on the context's copy stream into the device's
 * n_forcing_days forcing columns, day d in column (d - 1) mod
 * n_forcing_days -- the chunked read of the meteo handler
 * (src/meteo/mo_meteo_helper.f90:283-327 chunk_config, iMeteoTS relative to
 * the chunk at src/meteo/mo_meteo_handler.f90:609-617) as a ring of device
 * columns.  Returns once the host arrays are copied (to pinned staging, two
 * slots: double buffering); the upload overlaps the compute work queued on
 * the context.  A chunk overwrites its columns only after the work queued
 * before it; a step reading a day waits for the chunk that brought it, and a
 * step whose day is not in its column is refused.  A chunk fits the columns
 * (ndays <= n_forcing_days, n_forcing_days >= 2); stream the next chunk
 * while the current one runs with n_forcing_days >= 2 chunks. */
int mhmhip_stream_forcing(mhmhip_ctx *ctx, int32_t day0, int32_t ndays,
                          const double *const *host, int64_t ld);
/* Forcing columns day0..day0+ndays-1 of one variable for one domain, given
 * on its meteo grid (L2) and mapped to the domain's L1 cells on the device
 * as meteo_forcings_wrapper does after reading them
 * (src/meteo/mo_meteo_helper.f90:119-143): L1 cells coarser than L2 average
 * the valid L2 cells of their block (spatial_aggregation,
 * src/meteo/mo_meteo_spatial_tools.f90:94-200), finer ones take their L2
 * cell or nodata -9999 (spatial_disaggregation, 313-377), equal grids copy;
 * then pack(L1, mask1).  data2: X(nrows2, ncols2, ndays) column-major;
 * mask2, mask1: column-major grids, nonzero = valid; mask1 holds exactly the
 * domain's cells in their order.  A valid cell that maps outside the other
 * grid (where the reference indexes out of bounds) is an error. */
int mhmhip_set_forcing_l2(mhmhip_ctx *ctx, int32_t var, int32_t domain, int32_t day0,
                          int32_t ndays, const double *data2, int32_t nrows2, int32_t ncols2,
                          double cellsize2, const uint8_t *mask2, int32_t nrows1,
                          int32_t ncols1, double cellsize1, const uint8_t *mask1);
/* Disaggregation of the daily meteo values by given weights instead of the
 * day/night factors (read_meteo_weights, src/meteo/mo_meteo_handler.f90:
 * 1102-1107, 1183-1189, 1265-1270 with temporal_disagg_meteo_weights,
 * src/meteo/mo_meteo_temporal_tools.f90:32-59): pre_w, temp_w, pet_w are the
 * handler's L1_pre_weights, L1_temp_weights, L1_pet_weights, each
 * X(ld, 12, 24) column-major (cell, month, hour + 1) over all cells; a step
 * of month m and hour h gets prec = pre * w, temp = (tavg + T0) * w - T0
 * (T0 = 273.15) and pet = pet_day * w.  All three given: on (the weights are
 * copied to the device, 3 x 288 doubles per cell); all three NULL: off. */
int mhmhip_set_meteo_weights(mhmhip_ctx *ctx, const double *pre_w, const double *temp_w,
                             const double *pet_w, int64_t ld);

/* ---- MPR-lite on the device ---------------------------------------------------
 * No counterpart in the reference, whose MPR (src/MPR/mo_multi_param_reg.f90:
 * 67-654) builds the L1 fields from L0 maps and look-up tables and hands them
 * to mHM() through the arrays mhmhip_set_cell_var fills.  For callers whose
 * cells carry the synthetic domains' inputs (land-cover fractions, a soil
 * texture index, aspect, latitude) and whose parameter sets are gamma vectors
 * -- the serving loop -- this builds every parameter row of a set on the
 * device from its gamma vector alone, the rows mhmhip_set_cell_var would
 * store for the fields of the stand-in regionalisation (parameters.mpr_lite:
 * the reference's transfer functions for the snow and sealed-storage fields,
 * src/MPR/mo_multi_param_reg.f90:800-852, 883-894; documented stand-ins for
 * the soil, recession and PET-aspect fields), so a set costs one launch
 * instead of its fields' host computation and upload.  The gamma vector's
 * order (mhm_parameter.nml's groups): */
#define MHMHIP_MPRLITE_CANOPY 0               /* canopyInterceptionFactor        */
#define MHMHIP_MPRLITE_SNOW_TEMP 1            /* snowTreshholdTemperature        */
#define MHMHIP_MPRLITE_DDF_FOREST 2           /* degreeDayFactor_forest          */
#define MHMHIP_MPRLITE_DDF_IMPERVIOUS 3       /* degreeDayFactor_impervious      */
#define MHMHIP_MPRLITE_DDF_PERVIOUS 4         /* degreeDayFactor_pervious        */
#define MHMHIP_MPRLITE_DDF_INCREASE 5         /* increaseDegreeDayFactorByPrecip */
#define MHMHIP_MPRLITE_MAXDDF_FOREST 6        /* maxDegreeDayFactor_forest       */
#define MHMHIP_MPRLITE_MAXDDF_IMPERVIOUS 7    /* maxDegreeDayFactor_impervious   */
#define MHMHIP_MPRLITE_MAXDDF_PERVIOUS 8      /* maxDegreeDayFactor_pervious     */
#define MHMHIP_MPRLITE_ROOT_FOREST 9          /* rootFractionCoefficient_forest  */
#define MHMHIP_MPRLITE_ROOT_IMPERVIOUS 10     /* rootFractionCoefficient_impervious */
#define MHMHIP_MPRLITE_ROOT_PERVIOUS 11       /* rootFractionCoefficient_pervious */
#define MHMHIP_MPRLITE_INFIL_SHAPE 12         /* infiltrationShapeFactor         */
#define MHMHIP_MPRLITE_JARVIS_C1 13           /* jarvis_sm_threshold_c1          */
#define MHMHIP_MPRLITE_IMPERVIOUS_STORAGE 14  /* imperviousStorageCapacity       */
#define MHMHIP_MPRLITE_PET_MIN 15             /* minCorrectionFactorPET          */
#define MHMHIP_MPRLITE_PET_MAX 16             /* maxCorrectionFactorPET          */
#define MHMHIP_MPRLITE_ASPECT_THRESH 17       /* aspectTresholdPET               */
#define MHMHIP_MPRLITE_HARGREAVES 18          /* HargreavesSamaniCoeff           */
#define MHMHIP_MPRLITE_INTERFLOW_CAPACITY 19  /* interflowStorageCapacityFactor  */
#define MHMHIP_MPRLITE_INTERFLOW_SLOPE 20     /* interflowRecession_slope (unused) */
#define MHMHIP_MPRLITE_FAST_FOREST 21         /* fastInterflowRecession_forest   */
#define MHMHIP_MPRLITE_SLOW_KS 22             /* slowInterflowRecession_Ks       */
#define MHMHIP_MPRLITE_EXP_SLOW 23            /* exponentSlowInterflow           */
#define MHMHIP_MPRLITE_RECHARGE 24            /* rechargeCoefficient             */
#define MHMHIP_MPRLITE_GEOPARAM 25            /* GeoParam                        */
#define MHMHIP_MPRLITE_NGAMMA 26              /* (the 5 routing gammas follow:
                                                 mhmhip_set_routing_params)     */
/* The cells' inputs, kept by the context (host copies, laid out on the
 * device as the cells are): cells is X(ld, 6) column-major over all cells:
 * forest, impervious and pervious fractions, soil texture index in [0, 1),
 * aspect [deg], latitude [deg]; horizon_depth[n_horizons] the horizons'
 * lower boundaries [mm]; lai_class[3][12] the monthly leaf area index of the
 * three cover classes. */
int mhmhip_mprlite_inputs(mhmhip_ctx *ctx, const double *cells, int64_t ld,
                          const double *horizon_depth, const double *lai_class);
/* Every parameter row of one set from its gamma vector (the first
 * MHMHIP_MPRLITE_NGAMMA entries, in the order above), on the context's
 * stream; the process_pet 2 and 3 rows too when the context has that case. */
int mhmhip_mprlite_params(mhmhip_ctx *ctx, int32_t set, const double *gamma);

/* ---- one step ----------------------------------------------------------------
 * get_corrected_pet + get_temp + get_prec + mHM() for every cell and set:
 * src/mHM/mo_mhm_interface_run.f90:357-457 calling
 * src/meteo/mo_meteo_handler.f90:915-1285 and src/mHM/mo_mhm.f90:229-521. */
int mhmhip_mhm_step(mhmhip_ctx *ctx, const mhmhip_step *st);

/* ---- routing -------------------------------------------------------------- */

/* L11_routing_order (src/mRM/mo_mrm_net_startup.f90:728-859) in
 * O(nLinks log nLinks): the same rOrder/netPerm as the reference's sweeps.
 * Host-only; fromN/toN/rOrder/netPerm have n_nodes entries, 1-based. */
int mhmhip_routing_order(int64_t n_nodes, int64_t n_links, const int32_t *fromN,
                         const int32_t *toN, int32_t *rOrder, int32_t *netPerm);

/* Device layout of one domain's network as mhmhip_bind_network builds it
 * at the default tile height of 16 levels (no counterpart in the reference;
 * for footprint estimates before a context exists and for inspection).
 * Host-only.  netPerm/fromN/toN as in
 * mhmhip_network for one domain (n_nodes - n_outlets links).  Optional
 * outputs: pos[n_nodes] = device index of each node; slots[4 * n_slots] =
 * per bin slot: device index (-1: proxy or padding), boundary row (-1: none),
 * lag | nchild << 12 | flags (bit 16 node, 17 proxy, 18 tile root),
 * first upstream slot within the bin | tile depth << 16.  MHMHIP_EINVAL when
 * slots_cap < n_slots (plan still filled). */
typedef struct {
  int32_t dmax;     /* deepest node (links to the sink)                      */
  int32_t tau_max;  /* deepest tile                                          */
  int64_t n_tiles;
  int64_t n_bins;   /* workgroups per parameter set (256 slots each)         */
  int64_t n_bnd;    /* tile roots draining into another tile                 */
  int64_t n_slots;  /* n_bins * 256                                          */
} mhmhip_plan;
int mhmhip_plan_network(int64_t n_nodes, int32_t n_outlets, const int32_t *netPerm,
                        const int32_t *fromN, const int32_t *toN, mhmhip_plan *plan,
                        int64_t *pos, int32_t *slots, int64_t slots_cap);

/* Network of every packed domain (input of mRM_routing,
 * src/mRM/mo_mrm_routing.f90:104-110); resets routing states to 0
 * (src/mRM/mo_mrm_init.f90:453-494). */
int mhmhip_bind_network(mhmhip_ctx *ctx, const mhmhip_network *net);

/* Levels per routing tile of the partition the next mhmhip_bind_network
 * builds: 2..64, 0 = 16 (the default).  No counterpart in the reference:
 * the results are the same at every height; the height trades the number
 * of tiled routing launches (tile depth + 1 for one step) against each
 * launch's length.  16 suits the time loop (mhmhip_run), 32 a caller that
 * routes one step per call (mhmhip_mrm_step: 0.94 against 1.16 ms per step
 * on a 1024 x 1024 basin).  Only before the first mhmhip_bind_network. */
int mhmhip_set_tile_height(mhmhip_ctx *ctx, int32_t height);
/* The five routing1 parameters (mhm_parameter.nml:193-201) of one parameter
 * set for one domain (domain < 0: every domain); C1/C2 follow from reg_rout
 * (src/mRM/mo_mrm_mpr.f90:61-119) over that domain's links. */
int mhmhip_set_routing_params(mhmhip_ctx *ctx, int32_t set, int32_t domain,
                              const double gamma[5]);
/* Adaptive-timestep routing with a constant celerity [m s-1]
 * (processCase(8) = 2) for one parameter set and domain (domain < 0: every
 * domain): mrm_update_param (src/mRM/mo_mrm_mpr.f90:268-329) -- K = link
 * length / celerity, the routing step TSrout = given_TS(locate(given_TS,
 * min K)) (src/mRM/mo_mrm_constants.F90:42-46), C1 = TSrout / (K + TSrout/2),
 * C2 = 1 - C1 K / TSrout.  mhmhip_mrm_step then routes each step's qOUT
 * rout_loop = max(1, nint(timestep / TSrout)) times and reports Qmod and the
 * gauges as the average over them (src/mRM/mo_mrm_routing.f90:224-284);
 * mhmhip_run takes rout_loop >= 1 (routing sub-steps inside the time loop);
 * routing steps longer than the model step (runoff accumulated over several
 * steps, src/mHM/mo_mhm_interface_run.f90:494-512) run through
 * mhmhip_mrm_step.  Routing steps that differ between the sets or domains of
 * one context are rejected.  mhmhip_set_routing_params returns a (set,
 * domain) to processCase 1. */
int mhmhip_set_routing_celerity(mhmhip_ctx *ctx, int32_t set, int32_t domain, double celerity);
/* Adaptive-timestep routing with a celerity per link [m s-1] (processCase(8)
 * = 3): mrm_update_param with K = L11_length / L11_celerity
 * (src/mRM/mo_mrm_mpr.f90:306-311), the rest as mhmhip_set_routing_celerity.
 * celerity holds the domain's links (domain >= 0) or every packed link
 * (domain < 0) in link order: L11_celerity as L11_calc_celerity
 * (src/mRM/mo_mrm_net_startup.f90:2212-2423) derives it from the L0 river
 * slopes -- MPR-time work of the caller. */
int mhmhip_set_routing_celerity_links(mhmhip_ctx *ctx, int32_t set, int32_t domain,
                                      const double *celerity);
/* Inflow discharge [m3 s-1] of every inflow gauge for the next mrm step(s):
 * q[step][n_inflow] for nsteps steps (InflowDischarge). */
int mhmhip_set_inflow(mhmhip_ctx *ctx, int32_t nsteps, const double *q);
/* L11 routing states of one set, per node in reference order (1-based node k
 * at index domain_node_start[d] + k - 1): qOUT, qTIN(:,IT1), qTR(:,IT1), Qmod,
 * C1, C2.  Any pointer may be NULL. */
int mhmhip_get_routing_state(mhmhip_ctx *ctx, int32_t set, double *qOUT, double *qTIN,
                             double *qTR, double *Qmod, double *C1, double *C2);
/* Output aggregation of L11_Qmod on the device (replaces the host-side
 * updateVariable / writeVariableTimestep of src/common/mo_nc_output.f90:
 * 140-175, called per step from mo_mhm_interface_run.f90:650-667): while on,
 * every routed step adds each node's Qmod to a per-node sum, in step order.
 * on = 1 starts (sums and count zeroed), on = 0 stops. */
int mhmhip_qmod_accumulate(mhmhip_ctx *ctx, int32_t on);
/* The sums of one set per node in reference order and the number of steps
 * added; sum / count is the average writeVariableTimestep writes. */
int mhmhip_get_qmod_sum(mhmhip_ctx *ctx, int32_t set, double *sum, int64_t *count);
/* Routing states of one set from a restart (mrm_read_restart_states,
 * src/mRM/mo_mrm_restart.f90:459-555): qTIN(:,IT1), qTR(:,IT1) per node in
 * reference order, and C1, C2 per link (as mhmhip_get_routing_state gives
 * them); any pointer may be NULL, C1 and C2 go together.  Loaded C1/C2 stay
 * while steps run with read_states = 1; the first routed step with
 * read_states = 0 replaces them by reg_rout of the set's routing parameters
 * (src/mRM/mo_mrm_routing.f90:211-218), which must have been given. */
int mhmhip_set_routing_state(mhmhip_ctx *ctx, int32_t set, const double *qTIN,
                             const double *qTR, const double *C1, const double *C2);

/* Output aggregation of the L1 states and fluxes on the device
 * (mHM_updateDataset, src/mHM/mo_write_fluxes_states.f90:283-438, with
 * updateVariable, src/common/mo_nc_output.f90:140-148, called every step
 * from src/mHM/mo_mhm_interface_run.f90:702-727): while on, every step adds
 * each selected output's value (states after the step, the step's fluxes, in
 * the reference's expressions) to a per-cell sum, in step order.
 * flags[k - 1] is outputFlxState(k), k = 1..21; k = 18 (neutrons) is not on
 * this path and must be 0.  Starting zeroes the sums and the step count;
 * flags NULL stops (the sums stay readable). */
int mhmhip_output_accumulate(mhmhip_ctx *ctx, const int32_t *flags);
/* The sum of output k (outputFlxState index; horizon 1..nH for k = 3, 4,
 * 17, 19, else ignored) of one set per cell in reference order, and the
 * number of steps added; sum / count is writeVariableTimestep's average. */
int mhmhip_get_output_sum(mhmhip_ctx *ctx, int32_t k, int32_t horizon, int32_t set,
                          double *sum, int64_t *count);

/* One mRM_routing() call for every domain and set after mhmhip_mhm_step:
 * src/mRM/mo_mrm_routing.f90:104-303, called at
 * src/mHM/mo_mhm_interface_run.f90:552-590.  gauge_out (host, may be NULL)
 * receives GaugeDischarge as [set][gauge].  st->read_states is read_restart:
 * 0 evaluates reg_rout (C1/C2 from the routing parameters), 1 keeps C1/C2 as
 * they are (set_routing_params' or a restart's).
 * With an adaptive routing step longer than the model step (tsRoutFactor
 * > 1) the step's L1_total_runoff and inflow add up (RunToRout,
 * InflowDischarge) and routing runs every nint(tsRoutFactor) steps of tt and
 * at tt == st->n_steps (src/mHM/mo_mhm_interface_run.f90:494-512): a step
 * that only accumulates returns MHMHIP_PENDING and leaves gauge_out alone;
 * a routing step returns MHMHIP_OK and its discharge, which the reference
 * then copies to the period's earlier steps (lines 614-618, the caller's). */
int mhmhip_mrm_step(mhmhip_ctx *ctx, const mhmhip_step *st, double *gauge_out);

/* ---- the time loop ------------------------------------------------------------
 * nsteps consecutive steps of do_time_step's hot part (mHM + mRM_routing,
 * src/mHM/mo_mhm_interface_run.f90:341-638) with the routing pipelined
 * across time.  gauge_out (host, may be NULL) receives [step][set][gauge].
 * read_states (read_restart) must be the same for every step of a call. */
int mhmhip_run(mhmhip_ctx *ctx, const mhmhip_step *steps, int32_t nsteps,
               const mhmhip_run_opts *opts, double *gauge_out);

/* mhmhip_run without device memory allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * steps, options (use_graph must be 0), results and states as mhmhip_run;
 * the call's step table, gauge series and runoff ring live in a 64 MiB
 * device array that each GPU the library runs on holds until the process
 * ends, instead of context buffers grown on demand.  A call whose scratch
 * (256-byte aligned step table of 104 B per step, gauge series of
 * nsteps x sets x gauges doubles, ring of (ring rows + 16) x sets x cells
 * rounded to 64 doubles) exceeds 64 MiB fails with MHMHIP_EINVAL, as does a
 * call with profiling on.  Calls on one GPU share the array: two calls that
 * overlap on one GPU (contexts on different host threads) race. */
int mhmhip_run_noalloc(mhmhip_ctx *ctx, const mhmhip_step *steps, int32_t nsteps,
                       const mhmhip_run_opts *opts, double *gauge_out);

/* ---- measurement (no counterpart in the reference, which has only wall
 * timers, mo_mhm_interface.F90:157-162) -------------------------------------
 * With profiling on, kernel launches are bracketed by HIP events on the
 * stream they run on; the device time, launch count and work units
 * (cell-steps for the cell kernels, node-steps for the routing) of the timed
 * launches accumulate per kernel. */
enum mhmhip_kernel { MHMHIP_K_CELL_STEP = 0, MHMHIP_K_CELL_FUSED, MHMHIP_K_ROUTE_DIAG,
                     MHMHIP_K_ROUTE_WG, MHMHIP_K_ROUTE_TILE, MHMHIP_K_COUNT };
/* on > 0 enables timing and resets the totals; each kernel's launches are
 * then timed one in `on` (1: all; an event pair costs the stream a few
 * microseconds).  In a captured time loop the timed launches are chosen at
 * capture and every replay times the same ones. */
int mhmhip_set_profiling(mhmhip_ctx *ctx, int32_t on);
int mhmhip_get_profile(mhmhip_ctx *ctx, int32_t kernel, double *ms, int64_t *launches,
                       double *units);

#ifdef __cplusplus
}
#endif
#endif
