! Harness for the reference-built RNG table (test infrastructure only).
! Links against the reference's own module `vars`
! (TSLB_newBCS/vars_single_3D_module.f90:78-165) and prints, for a fixed set
! of seeds (i,j,k,l), the IEEE bit patterns of rand_noseeded and
! gauss_noseeded.
program ref_rng_table
  use vars, only: rand_noseeded, gauss_noseeded
  implicit none
  integer :: i, j, k, l, il
  integer, parameter :: nl = 9
  integer, dimension(nl) :: ls = (/ 0, 1, 11, 22, 33, 168, 169, 524, 1732 /)
  real(4) :: r, g
  do il = 1, nl
    l = ls(il)
    do k = 1, 181, 45
      do j = 1, 200, 37
        do i = 1, 200, 13
          r = rand_noseeded(i, j, k, l)
          g = gauss_noseeded(i, j, k, l)
          write(*, '(4I6,1X,Z8.8,1X,Z8.8)') i, j, k, l, transfer(r, 0), transfer(g, 0)
        enddo
      enddo
    enddo
  enddo
  ! inflow seeds: (i,j,1,step)
  do l = 1, 40
    write(*, '(4I6,1X,Z8.8,1X,Z8.8)') 110, 111, 1, 10*l, transfer(rand_noseeded(110,111,1,10*l), 0), &
         transfer(gauss_noseeded(110,111,1,10*l), 0)
  enddo
end program ref_rng_table
