! dml_f -- ISO_C_BINDING interface to libdml (include/dml.h), the MI355X
! implementation of the din-mol-Li hot path.  A Fortran caller (dana.F90 or
! dana_dml.F90) uses these interfaces in place of the CPU routines:
!   fuerza (dana.F90:1055-1139)          -> dml_forces
!   cbrownian_hs (dana.F90:798-846)      -> dml_brownian
!   ermak_a / ermak_b (dana.F90:974-1052) -> dml_ermak_a / dml_ermak_b
!   test_update (Neighbor.F90:668-713)   -> dml_test_update
!   overlap_moveback (dana.F90:849-943)  -> dml_overlap_moveback
!   F -> CG (dana.F90:228-236)           -> dml_f_to_cg
!   gcmc_run (dana.F90:590-713)          -> dml_gcmc
!   calc_rho / maxz (dana.F90:521-549, 776-794) -> dml_calc_rho / dml_maxz
!   config_chunk / bloques (dana.F90:552-587, 716-773) -> dml_set_chunk / dml_bloques
!   salida's energia and kion (dana.F90:1141-1183, 1342-1376) -> dml_observables
! A Fortran array real(c_double) :: pos(n,3) is the structure-of-arrays
! layout the ABI takes (element (i,k) at [k*n + i]).
module dml_f
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_ptr, c_null_ptr

  integer(c_int), parameter :: DML_OK = 0, DML_E_ARG = -1, DML_E_HIP = -2, &
    DML_E_NOMEM = -3, DML_E_OUT_OF_TESS = -4, DML_E_LIST_FULL = -5, &
    DML_E_DEPOSIT_Z0 = -6, DML_E_NO_PARTICLE = -7, DML_E_OVERLAP = -8, &
    DML_E_STATE = -9, DML_E_UNSUPPORTED = -10
  ! enums of dml.h (Fortran names are case-blind: DML_BROWNIAN would clash
  ! with dml_brownian, hence the category prefixes)
  integer(c_int32_t), parameter :: DML_INT_BROWNIAN = 0, DML_INT_ERMAK = 1
  integer(c_int32_t), parameter :: DML_RES_PISTON = 0, DML_RES_CHUNKS = 1, DML_RES_GCMC = 2, &
    DML_RES_NONE = 3
  integer(c_int32_t), parameter :: DML_PREC_FP64 = 0, DML_PREC_FP32_FORCE = 1
  integer(c_int32_t), parameter :: DML_OV_EXACT = 0, DML_OV_ROUNDS = 1

  ! struct dml_config
  type, bind(C) :: dml_config
    real(c_double) :: box(3), rcut, skin
    real(c_double) :: eps(9), r0(9), mass(3)
    real(c_double) :: h, gamma, temp, dif_sc, dif_sei, z_switch, prob, z0, zmax
    integer(c_int32_t) :: integrator, reservoir, precision, overlap_mode
    integer(c_int32_t) :: capacity, max_neighbors
    integer(c_int64_t) :: seed
    real(c_double) :: act
    integer(c_int32_t) :: nadj, lanes_per_atom
    integer(c_int32_t) :: gcmc_mode, reserved
  end type dml_config

  ! struct dml_counters
  type, bind(C) :: dml_counters
    integer(c_int64_t) :: try_, depo, choques, choques3, nupd, ins, del, passes
    integer(c_int64_t) :: particle_steps, row_entries, rows, ov_unfinished
  end type dml_counters

  interface
    integer(c_int) function dml_create(cfg, ctx) bind(C, name='dml_create')
      import :: c_int, c_ptr, dml_config
      type(dml_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function dml_create_ensemble(nrep, cfgs, ctx) bind(C, name='dml_create_ensemble')
      import :: c_int, c_int32_t, c_ptr, dml_config
      integer(c_int32_t), value :: nrep
      type(dml_config), intent(in) :: cfgs(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function dml_destroy(ctx) bind(C, name='dml_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_sync(ctx) bind(C, name='dml_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function dml_strerror_c(status) bind(C, name='dml_strerror')
      import :: c_int, c_ptr
      integer(c_int), value :: status
    end function
    integer(c_int) function dml_set_state(ctx, rep, n, pos, vel, acel, pos_old, old_cg, &
                                          species, flags) bind(C, name='dml_set_state')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep, n
      type(c_ptr), value :: pos, vel, acel, pos_old, old_cg, species, flags
    end function
    integer(c_int) function dml_get_state(ctx, rep, n, pos, vel, acel, force, epot, pos_old, &
                                          old_cg, species, flags) bind(C, name='dml_get_state')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep, n
      type(c_ptr), value :: pos, vel, acel, force, epot, pos_old, old_cg, species, flags
    end function
    integer(c_int) function dml_get_slots(ctx, rep, nslots, bkey, pid, seq) bind(C, name='dml_get_slots')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep
      integer(c_int32_t), intent(out) :: nslots
      type(c_ptr), value :: bkey, pid, seq
    end function
    integer(c_int) function dml_get_zmax(ctx, rep, zmax) bind(C, name='dml_get_zmax')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep
      real(c_double), intent(out) :: zmax
    end function
    integer(c_int) function dml_test_update(ctx, rebuilt) bind(C, name='dml_test_update')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rebuilt
    end function
    integer(c_int) function dml_forces(ctx) bind(C, name='dml_forces')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_brownian(ctx, noise, u_dep) bind(C, name='dml_brownian')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, noise, u_dep
    end function
    integer(c_int) function dml_ermak_a(ctx, noise, u_dep) bind(C, name='dml_ermak_a')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, noise, u_dep
    end function
    integer(c_int) function dml_ermak_b(ctx) bind(C, name='dml_ermak_b')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_overlap_moveback(ctx, u_contact, n_u, max_passes) &
        bind(C, name='dml_overlap_moveback')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, u_contact
      integer(c_int64_t), value :: n_u
      integer(c_int32_t), value :: max_passes
    end function
    integer(c_int) function dml_overlap_rounds(ctx, u_round, max_rounds) bind(C, name='dml_overlap_rounds')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, u_round
      integer(c_int32_t), value :: max_rounds
    end function
    integer(c_int) function dml_f_to_cg(ctx) bind(C, name='dml_f_to_cg')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_gcmc(ctx, rep, act, nadj, u, g) bind(C, name='dml_gcmc')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep, nadj
      real(c_double), value :: act
      type(c_ptr), value :: u, g
    end function
    integer(c_int) function dml_calc_rho(ctx, rho) bind(C, name='dml_calc_rho')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho
    end function
    integer(c_int) function dml_observables(ctx, out) bind(C, name='dml_observables')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out
    end function
    integer(c_int) function dml_set_chunk(ctx, rep, n, pos, dist, z1) bind(C, name='dml_set_chunk')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep, n
      type(c_ptr), value :: pos
      real(c_double), value :: dist, z1
    end function
    integer(c_int) function dml_bloques(ctx) bind(C, name='dml_bloques')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_maxz(ctx) bind(C, name='dml_maxz')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_end_step(ctx) bind(C, name='dml_end_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function dml_run(ctx, nsteps, use_graph) bind(C, name='dml_run')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, use_graph
    end function
    integer(c_int) function dml_get_counters(ctx, rep, out) bind(C, name='dml_get_counters')
      import :: c_int, c_int32_t, c_ptr, dml_counters
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep
      type(dml_counters), intent(out) :: out
    end function
    integer(c_int) function dml_replica_status(ctx, rep, status) bind(C, name='dml_replica_status')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rep
      integer(c_int32_t), intent(out) :: status
    end function
  end interface

contains

  ! dana's pair parameters (dana.F90:86-100) and constants (dana.F90:12-17,
  ! 113) in a config; eps(k,m) is symmetric, so the Fortran column-major
  ! flattening equals the ABI's [(k-1)*3 + (m-1)].
  subroutine dml_dana_config(cfg, xi, yi, zmax, z0, skin, h, prob, dif_sc, dif_sei, &
                             ermak, reservoir, capacity, seed)
    type(dml_config), intent(out) :: cfg
    real(c_double), intent(in) :: xi, yi, zmax, z0, skin, h, prob, dif_sc, dif_sei
    logical, intent(in) :: ermak
    integer(c_int32_t), intent(in) :: reservoir, capacity
    integer(c_int64_t), intent(in) :: seed
    real(c_double) :: eps(3,3), r0(3,3)
    eps(:,:) = 0._c_double
    r0(:,:) = 0._c_double
    r0(2,1) = 3.5_c_double
    r0(1,2) = r0(2,1)
    eps(1,1) = 2313.6_c_double
    r0(1,1) = 3.2_c_double
    eps(3,3) = 121._c_double
    r0(3,3) = 3.61_c_double
    eps(1,3) = 529.1_c_double
    eps(3,1) = eps(1,3)
    r0(1,3) = 1.564_c_double
    r0(3,1) = r0(1,3)
    cfg%box = [xi, yi, zmax]
    cfg%rcut = 3.2_c_double
    cfg%skin = skin
    cfg%eps = reshape(eps, [9])
    cfg%r0 = reshape(r0, [9])
    cfg%mass = 6.94_c_double
    cfg%h = h
    cfg%gamma = 1._c_double
    cfg%temp = 300._c_double
    cfg%dif_sc = dif_sc
    cfg%dif_sei = dif_sei
    cfg%z_switch = 80._c_double
    cfg%prob = prob
    cfg%z0 = z0
    cfg%zmax = zmax
    cfg%integrator = merge(DML_INT_ERMAK, DML_INT_BROWNIAN, ermak)
    cfg%reservoir = reservoir
    cfg%precision = DML_PREC_FP64
    cfg%overlap_mode = DML_OV_EXACT
    cfg%capacity = capacity
    cfg%max_neighbors = 0
    cfg%seed = seed
    cfg%act = 0._c_double
    cfg%nadj = 0
    cfg%lanes_per_atom = 0
    cfg%gcmc_mode = 0
    cfg%reserved = 0
  end subroutine dml_dana_config

  ! the status text, as a Fortran string
  function dml_strerror(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=64) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    call c_f_pointer(dml_strerror_c(status), p, [64])
    msg = ' '
    do i = 1, 64
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    enddo
  end function dml_strerror

  ! stop like werr (Errors.f90:59-87) on a failed call
  subroutine dml_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= DML_OK) then
      write(0, '(a,a,a,i0,a,a)') 'dml: ', where, ': status ', status, ' ', trim(dml_strerror(status))
      stop 1
    endif
  end subroutine dml_check

end module dml_f
