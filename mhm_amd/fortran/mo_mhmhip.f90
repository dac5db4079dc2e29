!> \file mo_mhmhip.f90
!> \brief iso_c_binding interface to the MI355X hot path (include/mhmhip.h).
!> \details Bindings for every entry point of the C-ABI plus the drop-in
!> routines a Fortran host calls in place of the meteo getters + mHM()
!> (src/mHM/mo_mhm_interface_run.f90:357-457) and mRM_routing()
!> (src/mHM/mo_mhm_interface_run.f90:552-590).  Arrays are passed as
!> c_loc(X(1,1)) with their leading dimension, so no temporaries are made.
module mo_mhmhip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MHMHIP_OK = 0, MHMHIP_PENDING = 1
  integer(c_int32_t), parameter, public :: MHMHIP_RUN_PLAIN = 0, MHMHIP_RUN_FUSED = 1

  ! enum mhmhip_var
  integer(c_int32_t), parameter, public :: &
    P_FSEALED = 0, P_DEGDAY_INC = 1, P_DEGDAY_MAX = 2, P_DEGDAY_NOPRE = 3, P_TEMP_THRESH = 4, &
    P_SEALED_THRESH = 5, P_JARVIS_C1 = 6, P_KFASTFLOW = 7, P_KSLOWFLOW = 8, P_KBASEFLOW = 9, &
    P_KPERCO = 10, P_ALPHA = 11, P_KARSTLOSS = 12, P_UNSAT_THRESH = 13, P_FASP = 14, &
    P_HARSAMCOEFF = 15, P_LATITUDE = 16, P_SOILMOIST_SAT = 17, P_FROOTS = 18, &
    P_SOILMOIST_FC = 19, P_WILTINGPOINT = 20, P_SOILMOIST_EXP = 21, P_MAXINTER = 22, &
    P_PETLAICOR = 23, P_PRIETAYALPHA = 24, P_AERORESIST = 25, P_SURFRESIST = 26, &
    S_INTER = 32, S_SNOWPACK = 33, S_SEALSTW = 34, S_UNSATSTW = 35, S_SATSTW = 36, &
    S_SOILMOIST = 37, &
    X_PET_CALC = 64, X_TEMP_CALC = 65, X_PREC_CALC = 66, X_THROUGHFALL = 67, X_AETCANOPY = 68, &
    X_DEGDAY = 69, X_MELT = 70, X_PREEFFECT = 71, X_RAIN = 72, X_SNOW = 73, X_RUNOFFSEAL = 74, &
    X_AETSEALED = 75, X_FASTRUNOFF = 76, X_SLOWRUNOFF = 77, X_PERCOL = 78, X_BASEFLOW = 79, &
    X_TOTAL_RUNOFF = 80, X_INFILSOIL = 81, X_AETSOIL = 82
  integer(c_int32_t), parameter, public :: F_PRE = 0, F_TAVG = 1, F_TMIN = 2, F_TMAX = 3, &
    F_PET = 4, F_NETRAD = 5, F_ABSVAPPRESS = 6, F_WINDSPEED = 7, F_COUNT = 8

  type, bind(C), public :: mhmhip_config
    integer(c_int64_t) :: n_cells
    integer(c_int32_t) :: n_domains, n_horizons, n_sets, timestep, process_soil, process_pet, &
                          process_route, hourly_forcing, n_forcing_days, device
    type(c_ptr) :: domain_cell_start = c_null_ptr
    real(c_double) :: evap_coeff(12), fnight_prec(12), fnight_pet(12), fnight_temp(12)
  end type

  type, bind(C), public :: mhmhip_step
    integer(c_int32_t) :: tt, month, hour, doy, iMeteoTS, read_states, n_steps = 0
  end type

  type, bind(C), public :: mhmhip_network
    integer(c_int64_t) :: n_nodes
    type(c_ptr) :: domain_node_start = c_null_ptr, n_outlets, netPerm, fromN, toN, length, &
                   slope, fFPimp, L1_area, L1_L11_Id
    integer(c_int32_t) :: n_gauges = 0, n_inflow = 0
    type(c_ptr) :: gauge_domain = c_null_ptr, gauge_node = c_null_ptr, &
                   inflow_domain = c_null_ptr, inflow_node = c_null_ptr, &
                   inflow_headwater = c_null_ptr
  end type

  type, bind(C), public :: mhmhip_run_opts
    integer(c_int32_t) :: mode = 1, steps_per_launch = 0, block = 0, use_graph = 0, &
                          route_mode = 0, route_window = 0, reserved(2) = 0
  end type

  type, bind(C), public :: mhmhip_plan
    integer(c_int32_t) :: dmax, tau_max
    integer(c_int64_t) :: n_tiles, n_bins, n_bnd, n_slots
  end type

  public :: mhmhip_create, mhmhip_destroy, mhmhip_last_error, mhmhip_device_bytes, mhmhip_sync, &
            mhmhip_set_cell_var, mhmhip_get_cell_var, mhmhip_default_states, &
            mhmhip_set_forcing, mhmhip_stream_forcing, mhmhip_set_forcing_l2, &
            mhmhip_set_meteo_weights, &
            mhmhip_set_routing_celerity, mhmhip_set_routing_celerity_links, &
            mhmhip_mhm_step, mhmhip_routing_order, mhmhip_bind_network, mhmhip_set_tile_height, &
            mhmhip_set_routing_params, mhmhip_set_inflow, mhmhip_get_routing_state, &
            mhmhip_set_routing_state, mhmhip_mrm_step, mhmhip_run, mhmhip_run_noalloc, &
            mhmhip_set_profiling, &
            mhmhip_get_profile, mhmhip_plan_network, mhmhip_qmod_accumulate, mhmhip_get_qmod_sum, &
            mhmhip_output_accumulate, mhmhip_get_output_sum
  public :: mhmhip_check, mhm_hip_put, mhm_hip_get, mhm_hip_do_time_step

  interface
    integer(c_int) function mhmhip_create(cfg, ctx) bind(C, name="mhmhip_create")
      import :: c_int, c_ptr, mhmhip_config
      type(mhmhip_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mhmhip_destroy(ctx) bind(C, name="mhmhip_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mhmhip_last_error(ctx) bind(C, name="mhmhip_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int64_t) function mhmhip_device_bytes(ctx) bind(C, name="mhmhip_device_bytes")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhmhip_sync(ctx) bind(C, name="mhmhip_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhmhip_set_cell_var(ctx, var, set, host, ld, ncols) &
        bind(C, name="mhmhip_set_cell_var")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int32_t), value :: var, set, ncols
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function mhmhip_get_cell_var(ctx, var, set, host, ld, ncols) &
        bind(C, name="mhmhip_get_cell_var")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int32_t), value :: var, set, ncols
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function mhmhip_default_states(ctx, horizon_depth) &
        bind(C, name="mhmhip_default_states")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, horizon_depth
    end function
    integer(c_int) function mhmhip_stream_forcing(ctx, day0, ndays, host, ld) &
        bind(C, name="mhmhip_stream_forcing")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: day0, ndays
      type(c_ptr), intent(in) :: host(8)
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function mhmhip_set_meteo_weights(ctx, pre_w, temp_w, pet_w, ld) &
        bind(C, name="mhmhip_set_meteo_weights")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, pre_w, temp_w, pet_w
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function mhmhip_set_forcing(ctx, var, day0, ndays, host, ld) &
        bind(C, name="mhmhip_set_forcing")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, host
      integer(c_int32_t), value :: var, day0, ndays
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function mhmhip_set_forcing_l2(ctx, var, domain, day0, ndays, data2, nrows2, &
        ncols2, cellsize2, mask2, nrows1, ncols1, cellsize1, mask1) &
        bind(C, name="mhmhip_set_forcing_l2")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, data2, mask2, mask1
      integer(c_int32_t), value :: var, domain, day0, ndays, nrows2, ncols2, nrows1, ncols1
      real(c_double), value :: cellsize2, cellsize1
    end function
    integer(c_int) function mhmhip_set_routing_celerity(ctx, set, domain, celerity) &
        bind(C, name="mhmhip_set_routing_celerity")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: set, domain
      real(c_double), value :: celerity
    end function
    integer(c_int) function mhmhip_mhm_step(ctx, st) bind(C, name="mhmhip_mhm_step")
      import :: c_int, c_ptr, mhmhip_step
      type(c_ptr), value :: ctx
      type(mhmhip_step), intent(in) :: st
    end function
    integer(c_int) function mhmhip_routing_order(n_nodes, n_links, fromN, toN, rOrder, netPerm) &
        bind(C, name="mhmhip_routing_order")
      import :: c_int, c_int64_t, c_int32_t
      integer(c_int64_t), value :: n_nodes, n_links
      integer(c_int32_t), intent(in) :: fromN(*), toN(*)
      integer(c_int32_t), intent(out) :: rOrder(*), netPerm(*)
    end function
    integer(c_int) function mhmhip_bind_network(ctx, net) bind(C, name="mhmhip_bind_network")
      import :: c_int, c_ptr, mhmhip_network
      type(c_ptr), value :: ctx
      type(mhmhip_network), intent(in) :: net
    end function
    integer(c_int) function mhmhip_set_tile_height(ctx, height) &
        bind(C, name="mhmhip_set_tile_height")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: height
    end function
    integer(c_int) function mhmhip_set_routing_params(ctx, set, domain, gamma) &
        bind(C, name="mhmhip_set_routing_params")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: set, domain
      real(c_double), intent(in) :: gamma(5)
    end function
    integer(c_int) function mhmhip_set_routing_celerity_links(ctx, set, domain, celerity) &
        bind(C, name="mhmhip_set_routing_celerity_links")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, celerity
      integer(c_int32_t), value :: set, domain
    end function
    integer(c_int) function mhmhip_set_inflow(ctx, nsteps, q) bind(C, name="mhmhip_set_inflow")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, q
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function mhmhip_get_routing_state(ctx, set, qOUT, qTIN, qTR, Qmod, C1, C2) &
        bind(C, name="mhmhip_get_routing_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, qOUT, qTIN, qTR, Qmod, C1, C2
      integer(c_int32_t), value :: set
    end function
    integer(c_int) function mhmhip_set_routing_state(ctx, set, qTIN, qTR, C1, C2) &
        bind(C, name="mhmhip_set_routing_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, qTIN, qTR, C1, C2
      integer(c_int32_t), value :: set
    end function
    integer(c_int) function mhmhip_output_accumulate(ctx, flags) &
        bind(C, name="mhmhip_output_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, flags
    end function
    integer(c_int) function mhmhip_get_output_sum(ctx, k, horizon, set, sum, count) &
        bind(C, name="mhmhip_get_output_sum")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, sum
      integer(c_int32_t), value :: k, horizon, set
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function mhmhip_mrm_step(ctx, st, gauge_out) bind(C, name="mhmhip_mrm_step")
      import :: c_int, c_ptr, mhmhip_step
      type(c_ptr), value :: ctx, gauge_out
      type(mhmhip_step), intent(in) :: st
    end function
    integer(c_int) function mhmhip_run(ctx, steps, nsteps, opts, gauge_out) &
        bind(C, name="mhmhip_run")
      import :: c_int, c_ptr, c_int32_t, mhmhip_step, mhmhip_run_opts
      type(c_ptr), value :: ctx, gauge_out
      type(mhmhip_step), intent(in) :: steps(*)
      integer(c_int32_t), value :: nsteps
      type(mhmhip_run_opts), intent(in) :: opts
    end function
    integer(c_int) function mhmhip_run_noalloc(ctx, steps, nsteps, opts, gauge_out) &
        bind(C, name="mhmhip_run_noalloc")
      import :: c_int, c_ptr, c_int32_t, mhmhip_step, mhmhip_run_opts
      type(c_ptr), value :: ctx, gauge_out
      type(mhmhip_step), intent(in) :: steps(*)
      integer(c_int32_t), value :: nsteps
      type(mhmhip_run_opts), intent(in) :: opts
    end function
    integer(c_int) function mhmhip_set_profiling(ctx, on) bind(C, name="mhmhip_set_profiling")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function mhmhip_get_profile(ctx, kernel, ms, launches, units) &
        bind(C, name="mhmhip_get_profile")
      import :: c_int, c_ptr, c_int32_t, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kernel
      real(c_double), intent(out) :: ms, units
      integer(c_int64_t), intent(out) :: launches
    end function
    integer(c_int) function mhmhip_plan_network(n_nodes, n_outlets, netPerm, fromN, toN, plan, &
                                                pos, slots, slots_cap) &
        bind(C, name="mhmhip_plan_network")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, mhmhip_plan
      integer(c_int64_t), value :: n_nodes, slots_cap
      integer(c_int32_t), value :: n_outlets
      type(c_ptr), value :: netPerm, fromN, toN, pos, slots
      type(mhmhip_plan), intent(out) :: plan
    end function
    integer(c_int) function mhmhip_qmod_accumulate(ctx, on) bind(C, name="mhmhip_qmod_accumulate")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function mhmhip_get_qmod_sum(ctx, set, sum, count) &
        bind(C, name="mhmhip_get_qmod_sum")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, sum
      integer(c_int32_t), value :: set
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  !> Negative status -> message and stop, the abort semantics of FORCES error_message.
  subroutine mhmhip_check(status, ctx, what)
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == MHMHIP_OK) return
    n = int(c_strlen(mhmhip_last_error(ctx)))
    call c_f_pointer(mhmhip_last_error(ctx), msg, [max(n, 1)])
    write(*, '(a, a, i0, a, *(a))') what, ': mhmhip status ', status, ': ', msg(1:n)
    error stop 1
  end subroutine

  !> Upload a module array X(ld, ncols) (parameter or state) of parameter set `set`.
  subroutine mhm_hip_put(ctx, var, set, x)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: var, set
    real(c_double), intent(in), target, contiguous :: x(:, :)
    call mhmhip_check(mhmhip_set_cell_var(ctx, var, set, c_loc(x(1, 1)), &
                      int(size(x, 1), c_int64_t), int(size(x, 2), c_int32_t)), ctx, "mhm_hip_put")
  end subroutine

  !> Download a state or flux of parameter set `set` into X(ld, ncols).
  subroutine mhm_hip_get(ctx, var, set, x)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: var, set
    real(c_double), intent(inout), target, contiguous :: x(:, :)
    call mhmhip_check(mhmhip_get_cell_var(ctx, var, set, c_loc(x(1, 1)), &
                      int(size(x, 1), c_int64_t), int(size(x, 2), c_int32_t)), ctx, "mhm_hip_get")
  end subroutine

  !> The hot part of mhm_interface_run_do_time_step: meteo getters + mHM() +
  !> mRM_routing() for one step, states resident on the device.  gauge(:)
  !> receives GaugeDischarge (mRM_runoff(tt,:)) of parameter set 0.
  !> With routing steps longer than the model step, a step that only
  !> accumulates runoff leaves gauge(:) alone and routed is .false.; the
  !> reference then copies a routing step's discharge to the period's earlier
  !> steps (src/mHM/mo_mhm_interface_run.f90:614-618), the caller's part.
  subroutine mhm_hip_do_time_step(ctx, tt, month, hour, doy, iMeteoTS, read_restart, routing, &
                                  gauge, n_time_steps, routed)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: tt, month, hour, doy, iMeteoTS
    logical, intent(in) :: read_restart, routing
    real(c_double), intent(inout), target, contiguous :: gauge(:)
    integer, intent(in), optional :: n_time_steps
    logical, intent(out), optional :: routed
    type(mhmhip_step) :: st
    integer(c_int) :: status
    st%tt = tt
    st%month = month
    st%hour = hour
    st%doy = doy
    st%iMeteoTS = iMeteoTS
    st%read_states = merge(1, 0, read_restart)
    st%n_steps = 0
    if (present(n_time_steps)) st%n_steps = n_time_steps
    if (present(routed)) routed = .false.
    call mhmhip_check(mhmhip_mhm_step(ctx, st), ctx, "mHM")
    if (.not. routing) return
    status = mhmhip_mrm_step(ctx, st, c_loc(gauge(1)))
    if (status == MHMHIP_PENDING) return
    call mhmhip_check(status, ctx, "mRM_routing")
    if (present(routed)) routed = .true.
  end subroutine

end module mo_mhmhip
