! tslb_mod -- iso_c_binding interface of libtslb.so (include/tslb.h).
!
! Replaces, for the driver, the module-global kernels of the reference
! (moments_TSLB*, fused_TSLB*, bcs_*, pbcs; TSLB_2C-mpi/lb_kernels_module.f90,
! boundary_cds_3D_module.f90) by calls on an opaque device context.
module tslb_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TSLB_RHO = 0, TSLB_U = 1, TSLB_V = 2, TSLB_W = 3
  integer(c_int), parameter :: TSLB_PXX = 4, TSLB_PYY = 5, TSLB_PZZ = 6
  integer(c_int), parameter :: TSLB_PXY = 7, TSLB_PXZ = 8, TSLB_PYZ = 9, TSLB_PHI = 10
  integer(c_int), parameter :: TSLB_FORCE_ADDITIVE = 0, TSLB_FORCE_SHIFT = 1
  integer(c_int), parameter :: TSLB_COLOUR = 3
  integer(c_int), parameter :: TSLB_STORAGE_FP32 = 0, TSLB_STORAGE_BF16 = 1, TSLB_STORAGE_FP16 = 2, &
    TSLB_STORAGE_FX16 = 3

  type, bind(C) :: tslb_params
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: ncomp
    integer(c_int32_t) :: periodic(3)
    integer(c_int32_t) :: force_mode
    integer(c_int32_t) :: storage
    integer(c_int32_t) :: device
    real(c_float) :: tau
    real(c_float) :: sigma
    real(c_float) :: sharp_c
    real(c_float) :: force(3)
    real(c_float) :: lin_force
    real(c_float) :: tau_b
    real(c_float) :: beta
  end type tslb_params

  interface
    subroutine tslb_default_params(p) bind(C, name='tslb_default_params')
      import :: tslb_params
      type(tslb_params), intent(out) :: p
    end subroutine
    integer(c_int) function tslb_create(ctx, p) bind(C, name='tslb_create')
      import :: c_int, c_ptr, tslb_params
      type(c_ptr), intent(out) :: ctx
      type(tslb_params), intent(in) :: p
    end function
    integer(c_int) function tslb_destroy(ctx) bind(C, name='tslb_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tslb_set_isfluid(ctx, isf) bind(C, name='tslb_set_isfluid')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: isf(*)
    end function
    integer(c_int) function tslb_upload_field(ctx, field, a) bind(C, name='tslb_upload_field')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function tslb_download_field(ctx, field, a) bind(C, name='tslb_download_field')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_float), intent(out) :: a(*)
    end function
    integer(c_int) function tslb_init_equilibrium(ctx, prepush) bind(C, name='tslb_init_equilibrium')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prepush
    end function
    integer(c_int) function tslb_seed_noise(ctx, field, amp, l) bind(C, name='tslb_seed_noise')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, l
      real(c_float), value :: amp
    end function
    integer(c_int) function tslb_set_storage_range(ctx, field, half_range) bind(C, name='tslb_set_storage_range')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_float), value :: half_range
    end function
    integer(c_int) function tslb_storage_clipped(ctx, count) bind(C, name='tslb_storage_clipped')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function tslb_step(ctx, nsteps) bind(C, name='tslb_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function tslb_set_zplane_bc(ctx, w_in, phi_in) bind(C, name='tslb_set_zplane_bc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: w_in    ! c_loc of nx*ny floats, or c_null_ptr to disable
      type(c_ptr), value :: phi_in  ! c_loc of nx*ny floats, or c_null_ptr
    end function
    integer(c_int) function tslb_set_jet_bc(ctx, enable) bind(C, name='tslb_set_jet_bc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function
    integer(c_int) function tslb_upload_plane(ctx, field, k, a) bind(C, name='tslb_upload_plane')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, k
      real(c_float), intent(in) :: a(*)
    end function
    integer(c_int) function tslb_checkpoint_write(ctx, path, step) bind(C, name='tslb_checkpoint_write')
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: step
    end function
    integer(c_int) function tslb_checkpoint_read(ctx, path, step) bind(C, name='tslb_checkpoint_read')
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), intent(out) :: step
    end function
    ! learned subgrid closure (no counterpart in the reference drivers)
    integer(c_int) function tslb_closure_create(m, max_samples) bind(C, name='tslb_closure_create')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), intent(out) :: m
      integer(c_int64_t), value :: max_samples
    end function
    integer(c_int) function tslb_closure_destroy(m) bind(C, name='tslb_closure_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function tslb_closure_set_params(m, theta, n) bind(C, name='tslb_closure_set_params')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: theta(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function tslb_closure_add_samples(m, ctx, factor) bind(C, name='tslb_closure_add_samples')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, ctx
      integer(c_int), value :: factor
    end function
    integer(c_int) function tslb_closure_train(m, steps, batch, lr, t_scale, seed, last_loss) &
        bind(C, name='tslb_closure_train')
      import :: c_int, c_ptr, c_float, c_int32_t, c_double
      type(c_ptr), value :: m
      integer(c_int), value :: steps, batch
      real(c_float), value :: lr, t_scale
      integer(c_int32_t), value :: seed
      real(c_double), intent(out) :: last_loss
    end function
    integer(c_int) function tslb_set_closure(ctx, m) bind(C, name='tslb_set_closure')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, m
    end function
    integer(c_int) function tslb_reduce(ctx, out) bind(C, name='tslb_reduce')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(7)
    end function
    integer(c_int) function tslb_sync(ctx) bind(C, name='tslb_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tslb_mem_info(free_b, total_b) bind(C, name='tslb_mem_info')
      import :: c_int, c_size_t
      integer(c_size_t), intent(out) :: free_b, total_b
    end function
    integer(c_int) function tslb_host_mem_mb(rss_mb, ram_mb) bind(C, name='tslb_host_mem_mb')
      import :: c_int, c_double
      real(c_double), intent(out) :: rss_mb, ram_mb
    end function
    real(c_float) function tslb_rand_noseeded(i, j, k, l) bind(C, name='tslb_rand_noseeded')
      import :: c_float, c_int
      integer(c_int), value :: i, j, k, l
    end function
    integer(c_int) function tslb_last_error(buf, n) bind(C, name='tslb_last_error')
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: n
    end function
  end interface

contains

  ! stop with the library's message when a call fails (the reference stops
  ! through dostop/doerror, TSLB_2C-mpi/mpi_module.f90:3226-3270)
  subroutine tslb_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, irc
    if (rc == 0) return
    irc = tslb_last_error(buf, 512)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    enddo
    write(0, '(a,a,a,i0,a,a)') 'tslb: ', trim(what), ' failed (', rc, '): ', trim(msg)
    error stop 1
  end subroutine tslb_check

end module tslb_mod
