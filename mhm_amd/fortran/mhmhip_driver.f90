!> \file mhmhip_driver.f90
!> \brief A Fortran host running the hot path through mo_mhmhip, the way
!> mhm_interface_run_do_time_step would after the drop-in: module arrays are
!> uploaded once, then one mhm_hip_do_time_step per step.
!> Usage: mhmhip_driver <input bundle> <output file> [<restart file>]
!> With a restart file, the run stops after the bundle's first step list,
!> writes the restart variable set (mHM states, src/mHM/mo_restart.f90; mRM
!> states and C1/C2, src/mRM/mo_mrm_restart.f90:56-428) to that file, ends
!> its context, and a fresh context reads the file back and runs the second
!> step list with read_restart (mhm_interface_run_prepare_domain,
!> src/mHM/mo_mhm_interface_run.f90:221-300): C1/C2 come from the file, not
!> from the routing parameters, which the restarted context never gets.
program mhmhip_driver
  use, intrinsic :: iso_c_binding
  use mo_mhmhip
  implicit none
  integer :: u, k, v, ncol, nargs, nparc, c0
  integer(c_int64_t) :: n_cells
  integer(c_int32_t) :: nh, n_days, n_steps, n_links, n_gauges, n_steps2
  character(len=4096) :: fin, fout, frst
  real(c_double), allocatable, target :: hd(:), pall(:, :), f4(:, :, :), length(:), slope(:), &
                                         fpimp(:), area(:), gauge(:), series(:, :), sm(:, :), &
                                         s1(:, :), qtin(:), qtr(:), qout(:), qmod(:), c1(:), c2(:)
  integer(c_int32_t), allocatable, target :: perm(:), fromn(:), ton(:), l1l11(:), gnode(:), &
                                             gdom(:), steps(:, :), steps2(:, :), nout(:)
  integer :: pcol(0:24)
  real(c_double) :: gam(5)
  type(mhmhip_config) :: cfg
  type(mhmhip_network) :: net
  type(c_ptr) :: ctx

  nargs = command_argument_count()
  if (nargs /= 2 .and. nargs /= 3) &
    error stop "usage: mhmhip_driver <input bundle> <output file> [<restart file>]"
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  if (nargs == 3) call get_command_argument(3, frst)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n_cells, nh, n_days, n_steps, n_links, n_gauges
  allocate(hd(nh))
  read(u) hd
  read(u) cfg%evap_coeff, cfg%fnight_prec, cfg%fnight_pet, cfg%fnight_temp
  allocate(perm(n_links), fromn(n_links), ton(n_links), length(n_links), slope(n_links), &
           fpimp(n_links), area(n_cells), l1l11(n_cells), gnode(n_gauges), gdom(n_gauges), nout(1))
  read(u) perm, fromn, ton, length, slope, fpimp, area, l1l11, gnode, gam
  gdom = 0
  nout = 1
  ! L1 effective parameters, enum order P_FSEALED .. P_PETLAICOR, as columns of pall
  pcol(0) = 1
  do v = 0, 23
    ncol = 1
    if (v >= P_SOILMOIST_SAT .and. v <= P_SOILMOIST_EXP) ncol = nh
    if (v >= P_MAXINTER) ncol = 12
    pcol(v + 1) = pcol(v) + ncol
  end do
  nparc = pcol(24) - 1
  allocate(pall(n_cells, nparc), f4(n_cells, n_days, 4))
  read(u) pall
  read(u) f4
  allocate(steps(6, n_steps))
  read(u) steps
  n_steps2 = 0
  if (nargs == 3) then
    read(u) n_steps2
    allocate(steps2(6, n_steps2))
    read(u) steps2
  end if
  close(u)

  call open_context(.true.)
  call mhmhip_check(mhmhip_default_states(ctx, c_loc(hd(1))), ctx, "default states")
  allocate(gauge(max(n_gauges, 1)), series(max(n_gauges, 1), n_steps + n_steps2))
  do k = 1, n_steps
    call mhm_hip_do_time_step(ctx, steps(1, k), steps(2, k), steps(3, k), steps(4, k), &
                              steps(5, k), steps(6, k) /= 0, .true., gauge)
    series(:, k) = gauge
  end do

  allocate(sm(n_cells, nh), s1(n_cells, 1))
  if (nargs == 3) then
    ! write the restart variable set of parameter set 0
    allocate(qtin(n_cells), qtr(n_cells), qout(n_cells), qmod(n_cells), c1(n_links), c2(n_links))
    call mhmhip_check(mhmhip_get_routing_state(ctx, 0_c_int32_t, c_loc(qout(1)), c_loc(qtin(1)), &
                      c_loc(qtr(1)), c_loc(qmod(1)), c_loc(c1(1)), c_loc(c2(1))), ctx, &
                      "routing state")
    open(newunit=u, file=trim(frst), access='stream', form='unformatted', status='replace')
    call mhm_hip_get(ctx, S_SOILMOIST, 0_c_int32_t, sm)
    write(u) sm
    do v = S_INTER, S_SATSTW
      call mhm_hip_get(ctx, int(v, c_int32_t), 0_c_int32_t, s1)
      write(u) s1
    end do
    write(u) qout, qtin, qtr, qmod, c1, c2
    close(u)
    call mhmhip_check(mhmhip_destroy(ctx), c_null_ptr, "destroy")
    ! a fresh context reads the restart: no routing parameters are given
    qtin = 0.0_c_double
    qtr = 0.0_c_double
    c1 = 0.0_c_double
    c2 = 0.0_c_double
    call open_context(.false.)
    open(newunit=u, file=trim(frst), access='stream', form='unformatted', status='old')
    read(u) sm
    call mhm_hip_put(ctx, S_SOILMOIST, 0_c_int32_t, sm)
    do v = S_INTER, S_SATSTW
      read(u) s1
      call mhm_hip_put(ctx, int(v, c_int32_t), 0_c_int32_t, s1)
    end do
    read(u) qout, qtin, qtr, qmod, c1, c2
    close(u)
    call mhmhip_check(mhmhip_set_routing_state(ctx, 0_c_int32_t, c_loc(qtin(1)), c_loc(qtr(1)), &
                      c_loc(c1(1)), c_loc(c2(1))), ctx, "restart routing state")
    do k = 1, n_steps2
      call mhm_hip_do_time_step(ctx, steps2(1, k), steps2(2, k), steps2(3, k), steps2(4, k), &
                                steps2(5, k), steps2(6, k) /= 0, .true., gauge)
      series(:, n_steps + k) = gauge
    end do
  end if

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) series(1:n_gauges, :)
  call mhm_hip_get(ctx, S_SOILMOIST, 0_c_int32_t, sm)
  write(u) sm
  do v = S_INTER, S_SATSTW
    call mhm_hip_get(ctx, int(v, c_int32_t), 0_c_int32_t, s1)
    write(u) s1
  end do
  call mhm_hip_get(ctx, X_TOTAL_RUNOFF, 0_c_int32_t, s1)
  write(u) s1
  close(u)
  call mhmhip_check(mhmhip_destroy(ctx), c_null_ptr, "destroy")

contains

  !> create + network + parameters + forcing (mhm_initialize and
  !> mhm_interface_run_prepare_domain of the drop-in); routing parameters
  !> only when with_routing_params
  subroutine open_context(with_routing_params)
    logical, intent(in) :: with_routing_params
    real(c_double), allocatable, target :: x(:, :), f(:, :)
    cfg%n_cells = n_cells
    cfg%n_domains = 1
    cfg%n_horizons = nh
    cfg%n_sets = 1
    cfg%timestep = 1
    cfg%process_soil = 1
    cfg%process_pet = 1
    cfg%process_route = 1
    cfg%hourly_forcing = 0
    cfg%n_forcing_days = n_days
    cfg%device = 0
    call mhmhip_check(mhmhip_create(cfg, ctx), c_null_ptr, "create")
    ! network first: cells are laid out in routing order once
    net%n_nodes = n_cells
    net%n_outlets = c_loc(nout(1))
    net%netPerm = c_loc(perm(1))
    net%fromN = c_loc(fromn(1))
    net%toN = c_loc(ton(1))
    net%length = c_loc(length(1))
    net%slope = c_loc(slope(1))
    net%fFPimp = c_loc(fpimp(1))
    net%L1_area = c_loc(area(1))
    net%L1_L11_Id = c_loc(l1l11(1))
    net%n_gauges = n_gauges
    net%gauge_domain = c_loc(gdom(1))
    net%gauge_node = c_loc(gnode(1))
    ! one routing step per call: tiles of 32 levels (fewer tile launches per step)
    call mhmhip_check(mhmhip_set_tile_height(ctx, 32_c_int32_t), ctx, "set_tile_height")
    call mhmhip_check(mhmhip_bind_network(ctx, net), ctx, "bind_network")
    if (with_routing_params) &
      call mhmhip_check(mhmhip_set_routing_params(ctx, 0, -1, gam), ctx, "routing params")
    do v = 0, 23
      c0 = pcol(v)
      allocate(x(n_cells, pcol(v + 1) - c0))
      x = pall(:, c0:pcol(v + 1) - 1)
      call mhm_hip_put(ctx, int(v, c_int32_t), 0_c_int32_t, x)
      deallocate(x)
    end do
    allocate(f(n_cells, n_days))
    do v = 0, 3
      f = f4(:, :, v + 1)
      call mhmhip_check(mhmhip_set_forcing(ctx, int(v, c_int32_t), 1, n_days, c_loc(f(1, 1)), &
                        n_cells), ctx, "forcing")
    end do
    deallocate(f)
  end subroutine

end program mhmhip_driver
