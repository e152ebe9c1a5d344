! nbdump -- drive the reference's own cell and neighbour modules on a given
! configuration and print what they build.  TEST INFRASTRUCTURE ONLY.
!
! It is linked against gems_groups / gems_cells / gems_neighbor compiled
! unmodified from the reference sources (see oracle/build_ref.sh).  Atoms
! are attached the way config_inic does (dana.F90:457-489) and
! test_update() (Neighbor.F90:668-713) builds cells and lists.  The harness
! sets the cgroup tessellate procedure pointer itself, because flang does not
! apply its default initialisation (Cells.F90:63).
!
! stdin:   box(1:3) / rcut skin / n / n lines "z x y z"
! stdout:  "cells tessellated ncells(3) cell(3)"
!          n lines "ci(3)" for atom k (file order; -1 when not binned)
!          for every ref atom in ref order: "row id nn" then nn ids
!          (ids are 1-based file positions)
program nbdump
use gems_constants, only: dp
use gems_groups, only: atom, sys, gindex, atom_dclist
use gems_neighbor, only: ngroup, ngindex, test_update, nb_dcut
use gems_cells, only: cgroup_tessellate
use gems_elements, only: elements_init, set_z
use gems_program_types, only: tbox, box_setvars
implicit none
type(ngroup), target        :: hs
type(atom), pointer         :: pa
type(atom_dclist), pointer  :: la
integer                     :: n, i, j, k, zz, c1, c2, c3, nn
integer, allocatable        :: ci(:,:)
real(dp)                    :: bx(3), rc, skin, x(3)

call elements_init()
call set_z(1,sym='Li',mass=6.94_dp)
call set_z(2,sym='CG',mass=6.94_dp)
call set_z(3,sym='F',mass=6.94_dp)
call gindex%init()
call ngindex%init()
call sys%init()

read(*,*) bx(1), bx(2), bx(3)
read(*,*) rc, skin
read(*,*) n
nb_dcut = skin
call hs%init()
hs%b%tessellate => cgroup_tessellate
call hs%setrc(rc)
tbox(:,:) = 0._dp
tbox(1,1) = bx(1)
tbox(2,2) = bx(2)
tbox(3,3) = bx(3)
call box_setvars()

do i = 1, n
  read(*,*) zz, x(1), x(2), x(3)
  allocate(pa)
  call pa%init()
  call pa%setz(zz)
  pa%pos(:) = x(:)
  pa%force(:) = 0._dp
  pa%pos_old(:) = pa%pos(:)
  call sys%attach(pa)
  if (zz == 2) then
    call hs%b%attach(pa)
  else
    call hs%ref%attach(pa)
    call hs%b%attach(pa)
  endif
  call hs%attach(pa)
  pa%pbc(:) = .true.
  pa%pbc(3) = .false.
  pa => null()
enddo

call test_update()

write(*,'(a,1x,l1,3(1x,i0),3(1x,es24.16))') 'cells', hs%b%tessellated, &
     hs%b%ncells(1), hs%b%ncells(2), hs%b%ncells(3), hs%b%cell(1), hs%b%cell(2), hs%b%cell(3)

allocate(ci(3,n))
ci(:,:) = -1
if (hs%b%tessellated) then
  do c3 = 0, hs%b%ncells(3)+1
    do c2 = 0, hs%b%ncells(2)+1
      do c1 = 0, hs%b%ncells(1)+1
        j = hs%b%head(c1,c2,c3)
        do while (j > 0)
          k = hs%b%a(j)%o%gid(sys)
          ci(:,k) = [c1, c2, c3]
          j = hs%b%next(j)
        enddo
      enddo
    enddo
  enddo
endif
do i = 1, n
  write(*,'(3(i0,1x))') ci(:,i)
enddo

la => hs%ref%alist
do i = 1, hs%ref%nat
  la => la%next
  pa => la%o
  k = pa%gid(hs)
  nn = hs%nn(k)
  write(*,'(a,1x,i0,1x,i0)') 'row', pa%gid(sys), nn
  do j = 1, nn
    write(*,'(i0)') hs%a(hs%list(k,j))%o%gid(sys)
  enddo
enddo
end program nbdump
