!> \file mhmhip_driver.f90
!> \brief A Fortran host running the hot path through mo_mhmhip, the way
!> mhm_interface_run_do_time_step would after the drop-in: module arrays are
!> uploaded once, then one mhm_hip_do_time_step per step.
!> Usage: mhmhip_driver <input bundle> <output file>
program mhmhip_driver
  use, intrinsic :: iso_c_binding
  use mo_mhmhip
  implicit none
  integer :: u, k, v, ncol, nargs
  integer(c_int64_t) :: n_cells
  integer(c_int32_t) :: nh, n_days, n_steps, n_links, n_gauges
  character(len=4096) :: fin, fout
  real(c_double), allocatable, target :: hd(:), x(:, :), f(:, :), length(:), slope(:), fpimp(:), &
                                         area(:), gauge(:), series(:, :), sm(:, :), s1(:, :)
  integer(c_int32_t), allocatable, target :: perm(:), fromn(:), ton(:), l1l11(:), gnode(:), &
                                             gdom(:), steps(:, :), nout(:)
  real(c_double) :: gam(5)
  type(mhmhip_config) :: cfg
  type(mhmhip_network) :: net
  type(c_ptr) :: ctx

  nargs = command_argument_count()
  if (nargs /= 2) error stop "usage: mhmhip_driver <input bundle> <output file>"
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n_cells, nh, n_days, n_steps, n_links, n_gauges
  allocate(hd(nh))
  read(u) hd

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
  read(u) cfg%evap_coeff, cfg%fnight_prec, cfg%fnight_pet, cfg%fnight_temp
  call mhmhip_check(mhmhip_create(cfg, ctx), c_null_ptr, "create")

  ! network first: cells are laid out in routing order once
  allocate(perm(n_links), fromn(n_links), ton(n_links), length(n_links), slope(n_links), &
           fpimp(n_links), area(n_cells), l1l11(n_cells), gnode(n_gauges), gdom(n_gauges), nout(1))
  read(u) perm, fromn, ton, length, slope, fpimp, area, l1l11, gnode, gam
  gdom = 0
  nout = 1
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
  call mhmhip_check(mhmhip_bind_network(ctx, net), ctx, "bind_network")
  call mhmhip_check(mhmhip_set_routing_params(ctx, 0, -1, gam), ctx, "routing params")

  ! L1 effective parameters, enum order P_FSEALED .. P_PETLAICOR
  do v = 0, 23
    ncol = 1
    if (v >= P_SOILMOIST_SAT .and. v <= P_SOILMOIST_EXP) ncol = nh
    if (v >= P_MAXINTER) ncol = 12
    allocate(x(n_cells, ncol))
    read(u) x
    call mhm_hip_put(ctx, int(v, c_int32_t), 0_c_int32_t, x)
    deallocate(x)
  end do
  call mhmhip_check(mhmhip_default_states(ctx, c_loc(hd(1))), ctx, "default states")
  allocate(f(n_cells, n_days))
  do v = 0, 3
    read(u) f
    call mhmhip_check(mhmhip_set_forcing(ctx, int(v, c_int32_t), 1, n_days, c_loc(f(1, 1)), &
                      n_cells), ctx, "forcing")
  end do
  allocate(steps(6, n_steps))
  read(u) steps
  close(u)

  allocate(gauge(max(n_gauges, 1)), series(max(n_gauges, 1), n_steps))
  do k = 1, n_steps
    call mhm_hip_do_time_step(ctx, steps(1, k), steps(2, k), steps(3, k), steps(4, k), &
                              steps(5, k), steps(6, k) /= 0, .true., gauge)
    series(:, k) = gauge
  end do

  allocate(sm(n_cells, nh), s1(n_cells, 1))
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
end program mhmhip_driver
