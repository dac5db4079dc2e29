/*
 * mhm_oracle.h -- CPU restatement of the mHM per-cell process chain and the
 * mRM Muskingum routing, in plain C (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the
 * HIP implementation in mhm_amd/ is compared, and the timed CPU baseline in
 * bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links, loads or falls back to it.
 *
 * Every function follows the reference Fortran statement by statement (same
 * operation order, same branch structure, same clamps), cited as
 * path:line into the mHM 5.13.2-dev0 tree.  The reference itself cannot be
 * built here (its FORCES dependency is absent), so the constants it takes
 * from FORCES mo_constants are restated with their published values.
 *
 * Array conventions follow the reference's column-major module arrays:
 *   per-cell 1-D arrays      x[cell]
 *   per-cell 2-D arrays      x[col * ld + cell]    (Fortran X(ld, ncol))
 * Network indices (netPerm, fromN, toN, L1_L11_Id, gauge node lists) are
 * 1-based, exactly as the reference passes them.
 */
#ifndef MHM_ORACLE_H
#define MHM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- leaf process functions (scalar, one cell) ------------------------- */

void or_canopy_interc(double pet, double interc_max, double precip,
                      double *interc, double *throughfall, double *evap_canopy);

void or_snow_accum_melt(double deg_day_incr, double deg_day_max, double deg_day_noprec,
                        double prec, double temperature, double temperature_thresh,
                        double thrfall, double *snow_pack, double *deg_day,
                        double *melt, double *prec_effect, double *rain, double *snow);

double or_feddes_et_reduction(double sm, double sm_fc, double wp, double froots);
double or_jarvis_et_reduction(double sm, double sm_sat, double wp, double froots, double c1);

/* soil_moisture for one cell: per-horizon arrays have nH entries (stride 1). */
void or_soil_moisture(int processCase, int nH, double frac_sealed, double water_thresh_sealed,
                      double pet, double evap_coeff, const double *sm_sat,
                      const double *frac_roots, const double *sm_fc, const double *wilting_point,
                      const double *sm_exponen, double jarvis_thresh_c1, double aet_canopy,
                      double prec_effec, double *runoff_sealed, double *storage_sealed,
                      double *infiltration, double *soil_moist, double *aet, double *aet_sealed);

void or_runoff_unsat_zone(double k1, double kp, double k0, double alpha, double karst_loss,
                          double pefec_soil, double unsat_thresh, double *sat_storage,
                          double *unsat_storage, double *slow_interflow,
                          double *fast_interflow, double *perc);
void or_runoff_sat_zone(double k2, double *sat_storage, double *baseflow);
double or_L1_total_runoff(double fsealed, double fast, double slow, double baseflow,
                          double direct_runoff);

/* PET and helpers */
int    or_LE(double a, double b);
double or_extraterr_rad_approx(int doy, double latitude_rad);
double or_pet_hargreaves(double coeff, double cnst, double tavg, double tmax, double tmin,
                         double latitude_deg, int doy);
double or_pet_priestly(double alpha, double rn, double tavg);
double or_pet_penman(double net_rad, double tavg, double act_vap_pressure,
                     double aerodyn_resistance, double bulksurface_resistance,
                     double a_s, double a_sh);
double or_slope_satpressure(double tavg);
double or_sat_vap_pressure(double tavg);

double or_disagg_meteo_weights(double v_day, double weight, double correction);
double or_disagg_flux_daynight(int isday, double ntimesteps_day, double v_day,
                               double fday, double fnight);
double or_disagg_state_daynight(int isday, double ntimesteps_day, double v_day,
                                double fday, double fnight, int add_correction);

/* ---- one model time step over all cells ---------------------------------
 * Row layout of the packed per-cell arrays (each row holds ld doubles):
 *   params : OR_P_* scalar rows, then 5 horizon blocks of nH rows
 *            (SAT, FROOTS, FC, WP, EXP), then 12 MAXINTER rows (month 1..12),
 *            then 12 PETLAICOR, 12 PRIETAYALPHA, 12 AERORESIST and
 *            12 SURFRESIST rows (month 1..12)
 *   states : OR_S_* rows, then nH SOILMOIST rows
 *   fluxes : OR_X_* rows, then nH INFIL rows, then nH AET_SOIL rows
 *   forcing: [var][day][cell], var in OR_F_* (netrad W m-2, absvappress Pa,
 *            windspeed m s-1)
 *   meteo weights (optional): [var][month * 24 + hour][cell], var = pre,
 *            temp, pet
 */
enum {
  OR_P_FSEALED = 0, OR_P_DEGDAY_INCR, OR_P_DEGDAY_MAX, OR_P_DEGDAY_NOPRE, OR_P_TEMP_THRESH,
  OR_P_SEALED_THRESH, OR_P_JARVIS_C1, OR_P_K0, OR_P_K1, OR_P_K2, OR_P_KP, OR_P_ALPHA,
  OR_P_KARST, OR_P_UNSAT_THRESH, OR_P_FASP, OR_P_HARSAM, OR_P_LAT, OR_P_NSCALAR
};
enum { OR_S_INTERC = 0, OR_S_SNOWPACK, OR_S_SEALSTW, OR_S_UNSATSTW, OR_S_SATSTW, OR_S_NSCALAR };
enum {
  OR_X_PET = 0, OR_X_TEMP, OR_X_PREC, OR_X_THROUGHFALL, OR_X_AET_CANOPY, OR_X_DEG_DAY,
  OR_X_MELT, OR_X_PREC_EFFECT, OR_X_RAIN, OR_X_SNOW, OR_X_RUNOFF_SEALED, OR_X_AET_SEALED,
  OR_X_FAST, OR_X_SLOW, OR_X_PERC, OR_X_BASEFLOW, OR_X_TOTAL_RUNOFF, OR_X_NSCALAR
};
enum { OR_F_PRE = 0, OR_F_TAVG, OR_F_TMIN, OR_F_TMAX, OR_F_PET, OR_F_NETRAD, OR_F_ABSVAPPRESS,
       OR_F_WINDSPEED, OR_F_NVAR };

int or_n_param_rows(int nH);
int or_n_state_rows(int nH);
int or_n_flux_rows(int nH);

typedef struct {
  int64_t n_cells;      /* cells in this call                                */
  int64_t ld;           /* leading dimension of param/state/flux rows        */
  int32_t n_h;          /* soil horizons                                     */
  int32_t process_soil; /* processMatrix(3,1)                                */
  int32_t process_pet;  /* processMatrix(5,1): -1, 0, 1, 2, 3                */
  int32_t hourly_forcing;
  double  c2TSTu;       /* timestep / 24                                     */
  double  nTstepDay;    /* 24 / timestep as real                             */
  int64_t forcing_ld;   /* cells per forcing day row                         */
  int32_t forcing_days; /* days per forcing variable                         */
  int32_t pad;
  double evap_coeff[12];
  double fnight_prec[12];
  double fnight_pet[12];
  double fnight_temp[12];
  const double *meteo_weights; /* NULL: day/night factors                    */
} or_cells_cfg;

typedef struct {
  int32_t tt;        /* 1-based step counter                                   */
  int32_t month;     /* 1..12                                                  */
  int32_t hour;      /* 0..23 at the start of the step                         */
  int32_t doy;       /* 1..366                                                 */
  int32_t iMeteoTS;  /* 1-based forcing column                                 */
  int32_t read_states;
} or_step_info;

/* meteo getters + mHM() for one step, OpenMP static over cells when
 * nthreads > 1 (0 = OpenMP default).  Reference: mo_mhm_interface_run.f90:357-457. */
void or_mhm_step(const or_cells_cfg *cfg, const or_step_info *st, const double *params,
                 const double *forcing, double *states, double *fluxes, int nthreads);

void or_default_states(int64_t n_cells, int64_t ld, int n_h, const double *horizon_depth,
                       double *states, double *fluxes);

/* ---- routing -------------------------------------------------------------- */
void or_reg_rout(const double param[5], int64_t n_links, const double *length,
                 const double *slope, const double *fFPimp, double TS, double *C1, double *C2);

void or_L11_runoff_acc(int64_t n_l1, const double *q_all, const double *efec_area,
                       const int32_t *L1_L11_Id, int64_t n_l11, const double *L11_area,
                       const int32_t *L11_L1_Id, int32_t TS, int32_t map_flag, double *q_acc);

void or_add_inflow(int32_t n_inflow, const int32_t *index_list, const int32_t *headwater,
                   const int32_t *node_list, const double *q_inflow, double *q_out);

/* qTIN, qTR: [2][n_nodes], slot 0 = IT1 (past), slot 1 = IT (current) */
void or_L11_routing(int64_t n_nodes, int64_t n_links, const int32_t *netPerm,
                    const int32_t *fromN, const int32_t *toN, const double *C1,
                    const double *C2, const double *qOUT, int32_t n_inflow,
                    const int32_t *inflow_headwater, const int32_t *inflow_node_list,
                    double *qTIN, double *qTR, double *Qmod);

typedef struct {
  int64_t n_nodes;
  int32_t n_outlets;
  int32_t timestep;       /* hours                                   */
  int32_t process_case;   /* processMatrix(8,1): 1, or 2 with C1/C2 given */
  int32_t read_states;
  int32_t map_flag;
  int32_t n_inflow;
  int32_t n_gauges;
  int32_t rout_loop;      /* routing sub-steps per call (0 or 1: one)  */
} or_route_cfg;

/* Adaptive routing parameters for a constant celerity (processCase(8) = 2):
 * mrm_update_param, src/mRM/mo_mrm_mpr.f90:268-329.  K = length / celerity
 * [s] per link; the routing step ts_rout = given_TS(locate(given_TS, min K))
 * (src/mRM/mo_mrm_constants.F90:42-46; locate and LE as in FORCES mo_utils,
 * restated from pre-proc/mtCLIM/mo_utils.f90:924-936,1329-1346); C1 =
 * ts_rout / (K (1 - xi) + ts_rout / 2), C2 = 1 - C1 K / ts_rout, xi = 0
 * (rout_space_weight).  tsRoutFactor = ts_rout / (timestep * 3600) and
 * rout_loop = max(1, nint(1 / tsRoutFactor)) as mo_mhm_interface_run.f90:
 * 482-493 and mo_mrm_routing.f90:224 derive them. */
void or_mrm_update_param_celerity(int64_t n_links, const double *length, double celerity,
                                  int32_t timestep, double *C1, double *C2, double *ts_rout,
                                  double *ts_rout_factor, int32_t *rout_loop);
/* mrm_update_param (src/mRM/mo_mrm_mpr.f90:241-329), processCase(8) = 3:
 * per-link celerity (stride 1) or one celerity for all links (stride 0). */
void or_mrm_update_param_links(int64_t n_links, const double *length, const double *celerity,
                               int32_t stride, int32_t timestep, double *C1, double *C2,
                               double *ts_rout, double *ts_rout_factor, int32_t *rout_loop);

/* One mRM_routing() call.  Reference: mo_mrm_routing.f90:104-303.
 * n_l1 cells of runoff; gauge_out receives n_gauges values (may be NULL). */
void or_mrm_routing(const or_route_cfg *rc, const double param[5], int64_t n_l1,
                    const double *L1_total_runoff, const double *L1_area,
                    const int32_t *L1_L11_Id, const double *L11_area,
                    const int32_t *L11_L1_Id, const int32_t *netPerm, const int32_t *fromN,
                    const int32_t *toN, const double *length, const double *slope,
                    const double *fFPimp, const int32_t *inflow_index, const int32_t *inflow_hw,
                    const int32_t *inflow_node, const double *inflow_q,
                    const int32_t *gauge_node, double *C1, double *C2, double *qOUT,
                    double *qTIN, double *qTR, double *Qmod, double *gauge_out);

/* L11_routing_order restated (O(n_links^2), for small networks only).
 * Reference: mo_mrm_net_startup.f90:728-859.  rOrder, netPerm: n_nodes entries. */
void or_routing_order(int64_t n_nodes, int64_t n_links, const int32_t *fromN,
                      const int32_t *toN, int32_t *rOrder, int32_t *netPerm);

/* ---- forcing on the meteo grid (L2) to the L1 cells ----------------------
 * meteo_forcings_wrapper's mapping (mo_meteo_helper.f90:119-143): L1 cells
 * coarser than L2 average the valid L2 cells of their block
 * (spatial_aggregation_3d, mo_meteo_spatial_tools.f90:94-200), finer ones take
 * their L2 cell's value or nodata (spatial_disaggregation_3d, 313-377), equal
 * grids copy; then pack(L1, mask1) (column-major over the masked cells).
 * data2: X(nrows2, ncols2, ndays); masks: column-major, nonzero = valid;
 * out: X(n_cells1, ndays) with n_cells1 = count(mask1).  Returns 0, or -1 when
 * a valid cell maps outside the other grid (where the reference indexes out
 * of bounds). */
int or_meteo_l2_to_l1(int32_t nrows2, int32_t ncols2, double cellsize2, const uint8_t *mask2,
                      int32_t nrows1, int32_t ncols1, double cellsize1, const uint8_t *mask1,
                      int32_t ndays, const double *data2, double *out);

#ifdef __cplusplus
}
#endif
#endif
