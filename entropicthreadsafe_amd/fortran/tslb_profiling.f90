! tslb_profiling -- named section/routine wall-clock timers for the driver.
!
! Same call interface as the reference's profiling module
! (start_timing2/end_timing2 at TSLB_2C-mpi/profiling_m.f90:937-987,
! print_timing_final at :514, time.dat at :266-308), measured with the wall
! clock (system_clock) after the device has finished the timed work.
module tslb_profiling
  use iso_c_binding
  use tslb_mod, only: tslb_sync
  implicit none
  private
  integer, parameter :: maxt = 64
  character(len=32) :: tsec(maxt), trou(maxt)
  real(8) :: ttot(maxt) = 0.d0, tstart(maxt) = 0.d0
  integer :: tcnt(maxt) = 0, nt = 0
  type(c_ptr), save :: sync_ctx = c_null_ptr
  public :: timing_set_context, start_timing2, end_timing2, print_timing_final, wall_seconds

contains

  real(8) function wall_seconds()
    integer(8) :: c, r
    call system_clock(c, r)
    wall_seconds = dble(c) / dble(r)
  end function wall_seconds

  ! device context whose queue is drained before a timer stops
  subroutine timing_set_context(ctx)
    type(c_ptr), intent(in) :: ctx
    sync_ctx = ctx
  end subroutine timing_set_context

  integer function slot(sec, rou)
    character(len=*), intent(in) :: sec, rou
    integer :: i
    do i = 1, nt
      if (trim(tsec(i)) == sec .and. trim(trou(i)) == rou) then
        slot = i
        return
      endif
    enddo
    if (nt == maxt) error stop 'tslb_profiling: too many timers'
    nt = nt + 1
    tsec(nt) = sec
    trou(nt) = rou
    slot = nt
  end function slot

  subroutine start_timing2(sec, rou)
    character(len=*), intent(in) :: sec, rou
    integer :: i
    i = slot(sec, rou)
    tstart(i) = wall_seconds()
  end subroutine start_timing2

  subroutine end_timing2(sec, rou)
    character(len=*), intent(in) :: sec, rou
    integer :: i, rc
    if (c_associated(sync_ctx)) rc = tslb_sync(sync_ctx)
    i = slot(sec, rou)
    ttot(i) = ttot(i) + (wall_seconds() - tstart(i))
    tcnt(i) = tcnt(i) + 1
  end subroutine end_timing2

  subroutine print_timing_final(iu)
    integer, intent(in) :: iu
    integer :: i, ju
    write(iu, '(a)') '****************************** timing *******************************'
    write(iu, '(a12,1x,a16,1x,a10,1x,a14,1x,a14)') 'section', 'routine', 'calls', 'total [s]', 'mean [ms]'
    do i = 1, nt
      write(iu, '(a12,1x,a16,1x,i10,1x,es14.6,1x,es14.6)') trim(tsec(i)), trim(trou(i)), tcnt(i), ttot(i), &
        1.d3 * ttot(i) / max(1, tcnt(i))
    enddo
    open(newunit=ju, file='time.dat', status='replace', action='write')
    do i = 1, nt
      write(ju, '(a,1x,a,1x,i10,1x,es16.8)') trim(tsec(i)), trim(trou(i)), tcnt(i), ttot(i)
    enddo
    close(ju)
  end subroutine print_timing_final

end module tslb_profiling
