! tslb_driver -- Fortran driver of the MI355X TSLB library.
!
! Keeps the reference driver's sequence (recursive_2c_TSLB3D.f90:1-850 and
! TSLB_newBCS/recursive_TSbcs_TSLB3.f90): parameters, allocation, geometry,
! initial conditions, the time loop with 2D slice output on stamp steps and
! timers around each call, then the GLUPS figure and memory report.  The
! populations never exist on the host: the device state is moment fields and
! one tslb_step() is the whole fused step.
!
! Usage: tslb_driver [input.nml]   (namelist /tslb_run/, defaults below)
program tslb_driver
  use iso_c_binding
  use tslb_mod
  use tslb_profiling
  implicit none

  ! ---- run parameters (namelist /tslb_run/) ---------------------------------
  integer :: nx = 64, ny = 64, nz = 64, ncomp = 2
  integer :: nsteps = 100, stamp2D = 50
  integer :: periodic(3) = (/1, 1, 1/)
  integer :: force_mode = 0
  real :: tau = 1.0, sigma = 0.02, sharp_c = 2.1
  real :: fx = 0.0, fy = 0.0, fz = 0.0, lin_force = 0.0
  character(len=16) :: config = 'droplet'
  real :: radius = 10.0, u0 = 0.05, uwall = 0.05, noise = 1.0e-3
  character(len=256) :: dir_out = 'output/'
  ! checkpoint (dump/out_m_.ckpt) every dumpstep steps; restart from a file
  integer :: dumpstep = 0, stamp_vtk = 0
  character(len=256) :: restart = ''
  ! moment storage: 'fp32' (the reference's REAL(4)), 'fp16', 'bf16' or 'fx16'
  character(len=8) :: storage = 'fp32'
  namelist /tslb_run/ nx, ny, nz, ncomp, nsteps, stamp2D, periodic, force_mode, tau, sigma, &
    sharp_c, fx, fy, fz, lin_force, config, radius, u0, uwall, noise, dir_out, dumpstep, &
    stamp_vtk, restart, storage

  type(tslb_params) :: prm
  type(c_ptr) :: ctx
  real(c_float), allocatable, target :: rho(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:), phi(:,:,:)
  real(c_float), allocatable, target :: w_in(:,:), phi_in(:,:)
  integer(c_int32_t), allocatable :: isfluid(:,:,:)
  real(c_double) :: red0(7), red(7), rss, ram
  integer(c_size_t) :: free_b, total_b
  integer :: i, j, k, step, iframe2D, iframe3D, iu, narg, step0
  integer(c_int64_t) :: ck_step
  character(len=256) :: nml
  logical :: has_solid, jet, jet_root
  real(8) :: ts1, ts2, glups, kk, pi

  pi = 4.d0 * atan(1.d0)
  narg = command_argument_count()
  if (narg >= 1) then
    call get_command_argument(1, nml)
    open(newunit=iu, file=trim(nml), status='old', action='read')
    read(iu, nml=tslb_run)
    close(iu)
  endif
  jet = trim(config) == 'jet'
  if (jet) periodic(3) = 0
  ! the root driver's jet: walls on all six faces, regularized inlet/outlet
  jet_root = trim(config) == 'jet_root'
  if (jet_root) periodic = 0

  ! ---- parameters and allocation (recursive_2c_TSLB3D.f90:20-59) ------------
  call tslb_default_params(prm)
  prm%nx = nx
  prm%ny = ny
  prm%nz = nz
  prm%ncomp = ncomp
  prm%periodic = periodic
  prm%force_mode = force_mode
  prm%tau = tau
  prm%sigma = sigma
  prm%sharp_c = sharp_c
  prm%force = (/fx, fy, fz/)
  prm%lin_force = lin_force
  select case (trim(storage))
  case ('fp32')
    prm%storage = TSLB_STORAGE_FP32
  case ('fp16')
    prm%storage = TSLB_STORAGE_FP16
  case ('bf16')
    prm%storage = TSLB_STORAGE_BF16
  case ('fx16')
    prm%storage = TSLB_STORAGE_FX16
  case default
    write(0, '(a,a)') 'tslb_driver: unknown storage ', trim(storage)
    error stop 1
  end select
  allocate(rho(nx,ny,nz), u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), phi(nx,ny,nz))
  allocate(isfluid(nx,ny,nz), w_in(nx,ny), phi_in(nx,ny))

  write(6,*) '*******************LB data*****************'
  write(6,*) 'tau', tau
  write(6,*) 'omega', 1.0 / tau
  write(6,*) 'visc', (tau - 0.5) / 3.0
  write(6,*) 'fx fy fz', fx, fy, fz
  write(6,*) 'sigma sharp_c', sigma, sharp_c
  write(6,*) '*******************INPUT data*****************'
  write(6,*) 'config ', trim(config)
  write(6,*) 'nx ny nz', nx, ny, nz
  write(6,*) 'ncomp', ncomp
  write(6,*) 'storage ', trim(storage)
  write(6,*) 'periodic', periodic
  write(6,*) 'nsteps', nsteps
  write(6,*) 'stamp2D', stamp2D

  call tslb_check(tslb_create(ctx, prm), 'tslb_create')
  call timing_set_context(ctx)

  ! ---- geometry and initial conditions (recursive_2c_TSLB3D.f90:79-104) -----
  isfluid = 1
  rho = 1.0
  u = 0.0
  v = 0.0
  w = 0.0
  phi = 0.0
  has_solid = .false.
  select case (trim(config))
  case ('droplet')
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          if ((real(i) - nx / 2.0)**2 + (real(j) - ny / 2.0)**2 + (real(k) - nz / 2.0)**2 <= radius**2) &
            phi(i, j, k) = 1.0
        enddo
      enddo
    enddo
  case ('tgv')
    kk = 2.d0 * pi / nx
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          u(i, j, k) = real(u0 * sin(kk * (i - 1)) * cos(kk * (j - 1)) * cos(kk * (k - 1)))
          v(i, j, k) = real(-u0 * cos(kk * (i - 1)) * sin(kk * (j - 1)) * cos(kk * (k - 1)))
        enddo
      enddo
    enddo
  case ('channel')
    do j = 1, ny
      u(:, j, :) = real(1.5d0 * u0 * ((j - 0.5d0) / (ny / 2.d0)) * (2.d0 - (j - 0.5d0) / (ny / 2.d0)))
    enddo
  case ('jet')
    ! newBCS jet: inflow disk of radius 10 on k = 1, drawn with amplitude 0.02
    ! from the generator at (i, j, 1, 524) and (i, j, 1, 1732)
    ! (TSLB_newBCS/recursive_TSbcs_TSLB3.f90:108-125)
    call inflow_initial()
  case ('jet_root')
    ! walls on all faces and an inflow disk of radius 20 on k = 1
    ! (recursive_2c_TSLB3D.f90:79-104); the seeded generator replaces
    ! random_number
    isfluid(1, :, :) = 0
    isfluid(nx, :, :) = 0
    isfluid(:, 1, :) = 0
    isfluid(:, ny, :) = 0
    isfluid(:, :, 1) = 0
    isfluid(:, :, nz) = 0
    has_solid = .true.
    w_in = 0.0
    do j = 1, ny
      do i = 1, nx
        if ((real(i) - nx / 2.0)**2 + (real(j) - ny / 2.0)**2 <= 20.0**2) then
          w_in(i, j) = uwall + 0.02 * sqrt(-2.0 * log(max(tslb_rand_noseeded(i, j, 1, 1), 1.0e-30))) * &
            cos(2.0 * real(pi) * tslb_rand_noseeded(i, j, 1, 2))
          phi(i, j, 1) = 1.0
        endif
      enddo
    enddo
    w(:, :, 1) = w_in
  case default
    write(0, *) 'unknown config ', trim(config)
    error stop 2
  end select
  if (has_solid) call tslb_check(tslb_set_isfluid(ctx, isfluid), 'tslb_set_isfluid')
  if (jet_root) call tslb_check(tslb_set_jet_bc(ctx, 1), 'tslb_set_jet_bc')
  call tslb_check(tslb_upload_field(ctx, TSLB_RHO, rho), 'upload rho')
  call tslb_check(tslb_upload_field(ctx, TSLB_U, u), 'upload u')
  call tslb_check(tslb_upload_field(ctx, TSLB_V, v), 'upload v')
  call tslb_check(tslb_upload_field(ctx, TSLB_W, w), 'upload w')
  call tslb_check(tslb_upload_field(ctx, TSLB_PHI, phi), 'upload phi')
  if (noise > 0.0) then
    ! seeded velocity noise, l = 11, 22, 33 (rand_noseeded, vars_single_3D_module.f90:102-165)
    call tslb_check(tslb_seed_noise(ctx, TSLB_U, noise, 11), 'noise u')
    call tslb_check(tslb_seed_noise(ctx, TSLB_V, noise, 22), 'noise v')
    call tslb_check(tslb_seed_noise(ctx, TSLB_W, noise, 33), 'noise w')
  endif
  call tslb_check(tslb_init_equilibrium(ctx, 1), 'tslb_init_equilibrium')
  step0 = 0
  if (len_trim(restart) > 0) then
    ! moment-space restart in place of read_distros_1c_3d (print_module.f90:713-723)
    call tslb_check(tslb_checkpoint_read(ctx, trim(restart) // c_null_char, ck_step), 'tslb_checkpoint_read')
    step0 = int(ck_step)
    write(6, '(a,a,a,i0)') 'restart from ', trim(restart), ' at step ', step0
  endif
  if (jet) call tslb_check(tslb_set_zplane_bc(ctx, c_loc(w_in), c_loc(phi_in)), 'tslb_set_zplane_bc')

  call execute_command_line('mkdir -p ' // trim(dir_out))
  call tslb_check(tslb_reduce(ctx, red0), 'tslb_reduce')
  write(6, '(a,es22.14,a,es22.14)') 'initial mass ', red0(1), ' phase mass ', red0(2)

  ! ---- time loop (recursive_2c_TSLB3D.f90:344-840) ---------------------------
  iframe2D = 0
  iframe3D = 0
  if (dumpstep > 0) call execute_command_line('mkdir -p dump')
  ts1 = wall_seconds()
  do step = step0 + 1, step0 + nsteps
    if (jet .and. mod(step, 10) == 0) then
      ! newBCS order: moments, then the BC with the inflow redraw of this step,
      ! then output and the collision (TSLB_newBCS/recursive_TSbcs_TSLB3.f90:
      ! 563-617; boundary_cds_3D_module.f90:1237-1259).  The state here is the
      ! moments of this step with the last BC applied; set_zplane_bc applies
      ! the BC again with the new plane.
      call start_timing2('LB', 'bcs_inflow')
      call inflow_plane(step)
      call tslb_check(tslb_set_zplane_bc(ctx, c_loc(w_in), c_loc(phi_in)), 'tslb_set_zplane_bc')
      call end_timing2('LB', 'bcs_inflow')
    endif
    if (mod(step, stamp2D) == 0) then
      call start_timing2('IO', 'print2d')
      iframe2D = iframe2D + 1
      call print_raw_slices(iframe2D)
      call end_timing2('IO', 'print2d')
    endif
    if (stamp_vtk > 0) then
      if (mod(step, stamp_vtk) == 0) then
        call start_timing2('IO', 'print_vtk')
        iframe3D = iframe3D + 1
        call print_vti_phi(iframe3D)
        call end_timing2('IO', 'print_vtk')
      endif
    endif
    if (dumpstep > 0) then
      if (mod(step, dumpstep) == 0) then
        ! checkpoint in place of dump_distros_1c_3d (print_module.f90:700-711); it
        ! holds the state before this step's update
        call start_timing2('IO', 'dump')
        call tslb_check(tslb_checkpoint_write(ctx, 'dump/out_m_.ckpt' // c_null_char, &
          int(step - 1, c_int64_t)), 'tslb_checkpoint_write')
        call end_timing2('IO', 'dump')
      endif
    endif
    call start_timing2('LB', 'step')
    call tslb_check(tslb_step(ctx, 1), 'tslb_step')
    call end_timing2('LB', 'step')
    if (jet_root .and. mod(step, 10) == 0) then
      ! inflow refresh inside radius 10 (boundary_cds_3D_module.f90:866-877)
      call start_timing2('LB', 'bcs_inflow')
      call inflow_disk(step)
      call tslb_check(tslb_upload_plane(ctx, TSLB_W, 0, w_in), 'tslb_upload_plane')
      call end_timing2('LB', 'bcs_inflow')
    endif
  enddo
  call tslb_check(tslb_sync(ctx), 'tslb_sync')
  ts2 = wall_seconds()
  call tslb_check(tslb_reduce(ctx, red), 'tslb_reduce')

  glups = dble(nx) * dble(ny) * dble(nz) * dble(nsteps) / 1.d9 / (ts2 - ts1)
  write(6, *) 'time elapsed: ', ts2 - ts1, ' s'
  write(6, '(a,f12.4)') ' glups: ', glups
  write(6, '(a,es22.14,a,es12.4)') 'final mass ', red(1), ' relative drift ', (red(1) - red0(1)) / red0(1)
  if (ncomp == 2) write(6, '(a,es22.14,a,es12.4)') 'final phase mass ', red(2), ' relative drift ', &
    (red(2) - red0(2)) / max(red0(2), 1.d-300)
  write(6, '(a,es22.14)') 'kinetic energy ', red(3)
  call tslb_check(tslb_mem_info(free_b, total_b), 'tslb_mem_info')
  call tslb_check(tslb_host_mem_mb(rss, ram), 'tslb_host_mem_mb')
  write(6, '(a,f12.1,a,f12.1,a)') 'DEVICE memory occupied ', dble(total_b - free_b) / 2.d0**20, &
    ' MB of ', dble(total_b) / 2.d0**20, ' MB'
  write(6, '(a,f12.1,a,f12.1,a)') 'HOST memory peak ', rss, ' MB of ', ram, ' MB'
  call print_timing_final(6)
  call tslb_check(tslb_destroy(ctx), 'tslb_destroy')

contains

  ! phi as VTK ImageData with appended raw Float32 data, in place of
  ! print_vtk_phionly_sync (print_module.f90:919-930): out_phi_NNNNNN.vti
  subroutine print_vti_phi(iframe)
    integer, intent(in) :: iframe
    character(len=512) :: fname
    character(len=1024) :: head
    character(len=128) :: ext
    character(len=1), parameter :: nl = char(10)
    integer :: u
    integer(4) :: nbytes
    call tslb_check(tslb_download_field(ctx, TSLB_PHI, phi), 'download phi')
    write(fname, '(a,a,i6.6,a)') trim(dir_out), 'out_phi_', iframe, '.vti'
    write(ext, '(i0,a,i0,a,i0,a,i0,a,i0,a,i0)') 1, ' ', nx, ' ', 1, ' ', ny, ' ', 1, ' ', nz
    head = '<?xml version="1.0"?>' // nl // &
      '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" header_type="UInt32">' // nl // &
      '  <ImageData WholeExtent="' // trim(ext) // '" Origin="0 0 0" Spacing="1 1 1">' // nl // &
      '    <Piece Extent="' // trim(ext) // '">' // nl // &
      '      <PointData Scalars="phi">' // nl // &
      '        <DataArray type="Float32" Name="phi" NumberOfComponents="1" format="appended" offset="0"/>' // nl // &
      '      </PointData>' // nl // '    </Piece>' // nl // '  </ImageData>' // nl // &
      '  <AppendedData encoding="raw">' // nl // '   _'
    nbytes = int(4, 4) * int(nx, 4) * int(ny, 4) * int(nz, 4)
    open(newunit=u, file=trim(fname), status='replace', action='write', access='stream', form='unformatted')
    write(u) trim(head), nbytes, phi, nl // '  </AppendedData>' // nl // '</VTKFile>' // nl
    close(u)
  end subroutine print_vti_phi

  ! jet_root inflow: redraw w inside radius 10 of the k = 1 plane, the same r
  ! twice as the reference (boundary_cds_3D_module.f90:868-874); w between
  ! radius 10 and 20 keeps its initial value
  subroutine inflow_disk(istep)
    integer, intent(in) :: istep
    integer :: ii, jj
    real :: rr
    do jj = 1, ny
      do ii = 1, nx
        if ((real(ii) - nx / 2.0)**2 + (real(jj) - ny / 2.0)**2 <= 10.0**2) then
          rr = tslb_rand_noseeded(ii, jj, 1, istep)
          w_in(ii, jj) = uwall + 0.004 * sqrt(-2.0 * log(max(rr, 1.0e-30))) * cos(2.0 * real(pi) * rr)
        endif
      enddo
    enddo
  end subroutine inflow_disk

  ! newBCS initial inflow: inside the disk w = uwall + 0.02 sqrt(-2 ln ry)
  ! cos(2 pi rx), rx and ry the generator at (i, j, 1, 524) and (i, j, 1, 1732);
  ! 0 outside; phi = 1 on the disk (two components)
  subroutine inflow_initial()
    integer :: ii, jj
    real :: rx, ry
    w_in = 0.0
    phi_in = 0.0
    do jj = 1, ny
      do ii = 1, nx
        if ((real(ii) - nx / 2.0)**2 + (real(jj) - ny / 2.0)**2 <= 10.0**2) then
          rx = tslb_rand_noseeded(ii, jj, 1, 524)
          ry = tslb_rand_noseeded(ii, jj, 1, 1732)
          w_in(ii, jj) = uwall + 0.02 * sqrt(-2.0 * log(max(ry, 1.0e-30))) * cos(2.0 * real(pi) * rx)
          phi_in(ii, jj) = 1.0
        endif
      enddo
    enddo
    w(:, :, 1) = w_in
    phi(:, :, 1) = phi_in
  end subroutine inflow_initial

  ! inflow redraw: w = uwall + 0.004 sqrt(-2 ln r) cos(2 pi r) inside the disk,
  ! with the same r twice as the reference (boundary_cds_3D_module.f90:1253-1255);
  ! outside the disk w keeps its value
  subroutine inflow_plane(istep)
    integer, intent(in) :: istep
    integer :: ii, jj
    real :: rr
    do jj = 1, ny
      do ii = 1, nx
        if ((real(ii) - nx / 2.0)**2 + (real(jj) - ny / 2.0)**2 <= 10.0**2) then
          rr = tslb_rand_noseeded(ii, jj, 1, istep)
          w_in(ii, jj) = uwall + 0.004 * sqrt(-2.0 * log(max(rr, 1.0e-30))) * cos(2.0 * real(pi) * rr)
        endif
      enddo
    enddo
  end subroutine inflow_plane

  ! mid-plane slices as stream-unformatted REAL(4), output/out_<f><plane>_NNNNNN.raw
  ! (print_raw_slice_2c_sync, print_module.f90:812-898)
  subroutine print_raw_slices(iframe)
    integer, intent(in) :: iframe
    call tslb_check(tslb_download_field(ctx, TSLB_RHO, rho), 'download rho')
    call tslb_check(tslb_download_field(ctx, TSLB_U, u), 'download u')
    call tslb_check(tslb_download_field(ctx, TSLB_V, v), 'download v')
    call tslb_check(tslb_download_field(ctx, TSLB_W, w), 'download w')
    call tslb_check(tslb_download_field(ctx, TSLB_PHI, phi), 'download phi')
    call write_slices('phi', phi, iframe)
    call write_slices('rho', rho, iframe)
    call write_slices('u', u, iframe)
    call write_slices('v', v, iframe)
    call write_slices('w', w, iframe)
  end subroutine print_raw_slices

  subroutine write_slices(name, a, iframe)
    character(len=*), intent(in) :: name
    real(c_float), intent(in) :: a(:,:,:)
    integer, intent(in) :: iframe
    character(len=6) :: num
    integer :: ju
    write(num, '(i6.6)') iframe
    open(newunit=ju, file=trim(dir_out) // 'out_' // name // 'xy_' // num // '.raw', status='replace', &
      action='write', access='stream', form='unformatted')
    write(ju) a(:, :, nz / 2)
    close(ju)
    open(newunit=ju, file=trim(dir_out) // 'out_' // name // 'xz_' // num // '.raw', status='replace', &
      action='write', access='stream', form='unformatted')
    write(ju) a(:, ny / 2, :)
    close(ju)
    open(newunit=ju, file=trim(dir_out) // 'out_' // name // 'yz_' // num // '.raw', status='replace', &
      action='write', access='stream', form='unformatted')
    write(ju) a(nx / 2, :, :)
    close(ju)
  end subroutine write_slices

end program tslb_driver
