! dana_dml -- the main loop of dana (dana.F90:101-263) driving libdml through
! the dml_f interfaces: same input files (entrada.ini, movedor.ini,
! pos_inic.xyz, chunk.xyz for the chunks reservoir), same call order, same
! output (Li.xyz frames and the E, T, rho, try and depo series of salida),
! with every hot routine on the GPU and the state resident there between
! calls.
!
! Usage: dana_dml [nst]   (run in a directory holding the input files;
! the optional argument overrides nst of entrada.ini)
program dana_dml
  use, intrinsic :: iso_c_binding
  use dml_f
  implicit none

  integer(c_int32_t) :: idum, nst, nwr, nadj, n, i, k, nslots, rep, nchunk
  real(c_double) :: prob, h, xi, yi, dist, z0, zmax, dif_sc, dif_sei, nb_dcut, act, z1, t
  real(c_double), target :: obs(4)
  integer(c_int64_t) :: try_last, depo_last
  logical :: integrador, s_piston, s_gcmc, s_chunk
  character(len=100) :: a
  character(len=10) :: sym
  character(len=32) :: arg
  real(c_double), allocatable, target :: pos(:,:), rho(:), chunk(:,:)
  integer(c_int8_t), allocatable, target :: species(:)
  integer(c_int8_t), allocatable, target :: flags(:)
  integer(c_int32_t), allocatable, target :: seq(:)
  type(dml_config) :: cfg
  type(dml_counters) :: cnt
  type(c_ptr) :: ctx
  integer(c_int32_t) :: capacity, reservoir, st

  rep = 0
  ! entrada (dana.F90:309-327)
  open(15, file='entrada.ini', status='old')
  read(15,*) idum
  read(15,*) prob
  read(15,*) h
  read(15,*) nst
  read(15,*) nwr
  read(15,*) xi
  read(15,*) yi
  read(15,*) dist
  read(15,*) z0
  read(15,*) zmax
  read(15,*) dif_sc
  read(15,*) dif_sei
  read(15,*) nb_dcut
  close(15)
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    read(arg,*) nst
  endif

  ! config_run (dana.F90:399-427)
  s_piston = .false.
  s_gcmc = .false.
  s_chunk = .false.
  act = 0._c_double
  nadj = 0
  open(15, file='movedor.ini', status='old')
  read(15,*) integrador
  read(15,*) a
  select case(trim(a))
  case('piston')
    s_piston = .true.
    reservoir = DML_RES_PISTON
  case('gcmc')
    s_gcmc = .true.
    reservoir = DML_RES_GCMC
    read(15,*) act, nadj
  case('chunks')
    s_chunk = .true.
    reservoir = DML_RES_CHUNKS
  case default
    write(0,*) 'Unknown reservoir type'
    stop 1
  end select
  close(15)

  ! config_inic (dana.F90:430-517): the chunks reservoir sizes (437-442)
  if (s_chunk) then
    dist = dist + 3.2_c_double
    z1 = z0 + dist
    zmax = z1 + dist
  endif
  open(11, file='pos_inic.xyz', status='old')
  read(11,*) n
  read(11,*)
  allocate(pos(n,3), species(n))
  do i = 1, n
    read(11,*) sym, pos(i,:)
    select case(trim(sym))
    case('CG')
      species(i) = 2
    case('F')
      species(i) = 3
    case default
      species(i) = 1
    end select
  enddo
  close(11)

  ! config_chunk (dana.F90:552-587) reads chunk.xyz; the template keeps its
  ! file coordinates here, dml_set_chunk places it at zmax + z
  nchunk = 0
  if (s_chunk) then
    open(11, file='chunk.xyz', status='old')
    read(11,*) nchunk
    read(11,*)
    allocate(chunk(nchunk,3))
    do i = 1, nchunk
      read(11,*) sym, chunk(i,:)
    enddo
    close(11)
  endif

  capacity = n
  if (s_gcmc) capacity = n + max(1024, 2 * nadj * nst)
  if (s_chunk) capacity = n + 64 * nchunk
  call dml_dana_config(cfg, xi, yi, zmax, z0, nb_dcut, h, prob, dif_sc, dif_sei, integrador, &
                       reservoir, capacity, int(abs(idum), c_int64_t))
  cfg%act = act
  cfg%nadj = nadj
  call dml_check(dml_create(cfg, ctx), 'dml_create')
  ! pos(n,3) is already the ABI's structure-of-arrays layout
  call dml_check(dml_set_state(ctx, rep, n, c_loc(pos), c_null_ptr, c_null_ptr, c_null_ptr, &
                               c_null_ptr, c_loc(species), c_null_ptr), 'dml_set_state')

  ! test_update, fuerza, calc_rho (dana.F90:140-145)
  call dml_check(dml_test_update(ctx, c_null_ptr), 'test_update')
  if (integrador) call dml_check(dml_forces(ctx), 'fuerza')
  allocate(rho(1))
  call dml_check(dml_calc_rho(ctx, c_loc(rho)), 'calc_rho')
  if (s_chunk) call dml_check(dml_set_chunk(ctx, rep, nchunk, c_loc(chunk), dist, z1), 'config_chunk')

  open(11, file='Li.xyz')
  open(12, file='E.dat')
  open(13, file='T.dat')
  open(14, file='rho.dat')
  open(15, file='try.dat')
  open(16, file='depo.dat')
  t = 0._c_double
  try_last = 0
  depo_last = 0
  call salida()

  ! the main loop (dana.F90:172-262)
  do i = 1, nst
    if (integrador) then
      call dml_check(dml_ermak_a(ctx, c_null_ptr, c_null_ptr), 'ermak_a')
      call dml_check(dml_forces(ctx), 'fuerza')
      call dml_check(dml_ermak_b(ctx), 'ermak_b')
    else
      call dml_check(dml_brownian(ctx, c_null_ptr, c_null_ptr), 'cbrownian_hs')
    endif
    call dml_check(dml_test_update(ctx, c_null_ptr), 'test_update')
    call dml_check(dml_overlap_moveback(ctx, c_null_ptr, 0_c_int64_t, 100000), 'overlap_moveback')
    call dml_check(dml_test_update(ctx, c_null_ptr), 'test_update')
    call dml_check(dml_f_to_cg(ctx), 'F to CG')
    if (s_gcmc) call dml_check(dml_gcmc(ctx, rep, act, nadj, c_null_ptr, c_null_ptr), 'gcmc_run')
    call dml_check(dml_calc_rho(ctx, c_loc(rho)), 'calc_rho')
    if (s_chunk) call dml_check(dml_bloques(ctx), 'bloques')
    if (mod(i, nwr) == 0) call salida()
    if (s_piston) call dml_check(dml_maxz(ctx), 'maxz')
    call dml_check(dml_end_step(ctx), 'end of step')
    t = t + h
  enddo
  close(11)
  close(12)
  close(13)
  close(14)
  close(15)
  close(16)

  call dml_check(dml_get_counters(ctx, rep, cnt), 'counters')
  write(*,'(a,i0,a,i0,a,i0,a,i0,a,i0)') 'steps ', nst, ' nupd ', cnt%nupd, ' depo ', cnt%depo, &
    ' ins ', cnt%ins, ' del ', cnt%del
  st = dml_destroy(ctx)

contains

  ! salida (dana.F90:1141-1183): the frame of every atom in sys order
  subroutine salida()
    real(c_double), allocatable, target :: p(:,:)
    integer(c_int32_t), allocatable :: order(:)
    integer(c_int32_t) :: j, nat, m
    real(c_double) :: zm
    character(len=2) :: s
    call dml_check(dml_get_slots(ctx, rep, nslots, c_null_ptr, c_null_ptr, c_null_ptr), 'slots')
    if (allocated(flags)) deallocate(flags, seq)
    allocate(p(nslots,3), flags(nslots), seq(nslots), order(nslots))
    call dml_check(dml_get_slots(ctx, rep, nslots, c_null_ptr, c_null_ptr, c_loc(seq)), 'slots')
    deallocate(species)
    allocate(species(nslots))
    call dml_check(dml_get_state(ctx, rep, nslots, c_loc(p), c_null_ptr, c_null_ptr, c_null_ptr, &
                                 c_null_ptr, c_null_ptr, c_null_ptr, c_loc(species), c_loc(flags)), &
                   'get_state')
    call dml_check(dml_get_zmax(ctx, rep, zm), 'get_zmax')
    ! alive slots sorted by creation number
    nat = 0
    do j = 1, nslots
      if (iand(int(flags(j)), 1) /= 0) then
        nat = nat + 1
        order(nat) = j
      endif
    enddo
    do j = 2, nat
      m = order(j)
      k = j - 1
      do while (k >= 1)
        if (seq(order(k)) <= seq(m)) exit
        order(k + 1) = order(k)
        k = k - 1
      enddo
      order(k + 1) = m
    enddo
    write(11,*) nat
    write(11,*) 'info:', zm, nat
    do j = 1, nat
      m = order(j)
      select case(species(m))
      case(2)
        s = 'CG'
      case(3)
        s = 'F'
      case default
        s = 'Li'
      end select
      write(11,*) s, p(m,:), int(species(m))
    enddo
    ! energia and kion on the device; try and depo since the last frame
    call dml_check(dml_observables(ctx, c_loc(obs)), 'observables')
    call dml_check(dml_get_counters(ctx, rep, cnt), 'counters')
    write(12,*) t, obs(1)
    write(13,*) t, obs(2)
    write(14,*) t, rho(1)
    write(15,*) t, cnt%try_ - try_last
    write(16,*) t, cnt%depo - depo_last
    try_last = cnt%try_
    depo_last = cnt%depo
    flush(16)
    flush(15)
    flush(14)
    flush(13)
    flush(12)
    flush(11)
  end subroutine salida

end program dana_dml
