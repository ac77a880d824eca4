! monopole_adjoint_hip.f90 -- a self-contained Fortran program that builds the acoustic
! monopole in code (no input files) and runs its forward + discrete-adjoint pass on the GPU
! through magudi_hip's t_Region / t_Solver (the reference's time loops over the stage
! entry points).  Usage: monopole_adjoint_hip [n] [nsteps] [save_interval] [dt]
! Prints the cost functional J and its sensitivity (Grad0) in magudi's output format
! (1X,SP,ES23.15E3).
program monopole_adjoint_hip

  use, intrinsic :: iso_c_binding
  use magudi_hip
  implicit none

  type(t_Region) :: region
  type(t_Solver) :: solver
  type(mgd_region_desc) :: desc
  real(c_double), allocatable :: coords(:,:), target(:,:), ctrl(:), tgt(:)
  real(c_double) :: dx, x, y, r, costJ, sens, loc(3), gamma
  integer :: n, i, j, p, w, ti(2), ci(2), cj(2), nsteps, saveInterval
  character(len = 32) :: arg

  n = 41; nsteps = 16; saveInterval = 4
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg); read(arg, *) n
  end if
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg); read(arg, *) nsteps
  end if
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg); read(arg, *) saveInterval
  end if
  dx = 28.0_c_double / real(n - 1, c_double)
  solver%timeStepSize = 0.5_c_double * dx / sqrt(2.0_c_double)
  if (command_argument_count() >= 4) then
     call get_command_argument(4, arg); read(arg, *) solver%timeStepSize
  end if
  solver%nTimesteps = nsteps
  solver%saveInterval = saveInterval

  gamma = 1.4_c_double
  desc%nd = 2
  desc%n = [n, n, 1]
  desc%dissipation = 1
  desc%dissipation_amount = 0.005_c_double
  desc%gamma = gamma

  allocate(coords(n * n, 2), target(n * n, 4), ctrl(n * n), tgt(n * n))
  do j = 1, n
     do i = 1, n
        p = i + n * (j - 1)
        coords(p, 1) = lin(i)
        coords(p, 2) = lin(j)
     end do
  end do
  call region%setup(desc, coords)

  w = n / 10 + 1
  call region%addPatch(MGD_PATCH_FAR_FIELD, 1, [1, 1, 1, -1, 1, -1], [1.0_c_double, 0.0_c_double])
  call region%addPatch(MGD_PATCH_FAR_FIELD, -1, [-1, -1, 1, -1, 1, -1], [1.0_c_double, 0.0_c_double])
  call region%addPatch(MGD_PATCH_FAR_FIELD, 2, [1, -1, 1, 1, 1, -1], [1.0_c_double, 0.0_c_double])
  call region%addPatch(MGD_PATCH_FAR_FIELD, -2, [1, -1, -1, -1, 1, -1], [1.0_c_double, 0.0_c_double])
  call region%addPatch(MGD_PATCH_SPONGE, 1, [1, w, 1, -1, 1, -1], [0.2_c_double, 2.0_c_double])
  call region%addPatch(MGD_PATCH_SPONGE, -1, [-w, -1, 1, -1, 1, -1], [0.2_c_double, 2.0_c_double])
  call region%addPatch(MGD_PATCH_SPONGE, 2, [1, -1, 1, w, 1, -1], [0.2_c_double, 2.0_c_double])
  call region%addPatch(MGD_PATCH_SPONGE, -2, [1, -1, -w, -1, 1, -1], [0.2_c_double, 2.0_c_double])
  ti = box(-9.6_c_double, 9.6_c_double)
  call region%addPatch(MGD_PATCH_COST_TARGET, 0, [ti(1), ti(2), ti(1), ti(2), 1, -1], [0.0_c_double])
  ci = box(-1.6_c_double, 1.6_c_double)
  cj = box(2.4_c_double, 5.6_c_double)
  call region%addPatch(MGD_PATCH_ACTUATOR, 0, [ci(1), ci(2), cj(1), cj(2), 1, -1], [0.0_c_double])

  target = 0.0_c_double
  target(:, 1) = 1.0_c_double
  target(:, 4) = (1.0_c_double / gamma) / (gamma - 1.0_c_double)
  do p = 1, n * n
     x = coords(p, 1); y = coords(p, 2)
     r = sqrt(x * x + y * y)
     tgt(p) = 0.25_c_double * (1.0_c_double + tanh(4.0_c_double * (r - 7.0_c_double))) *      &
          (1.0_c_double - tanh(4.0_c_double * (r - 9.0_c_double)))
     r = sqrt(x * x + (y - 4.0_c_double) ** 2)
     ctrl(p) = 0.5_c_double * (1.0_c_double - tanh(6.0_c_double * (r - 1.0_c_double)))
  end do
  call check(mgd_region_set_field(region%handle, MGD_FIELD_TARGET_STATE, target), "set target")
  call check(mgd_region_set_field(region%handle, MGD_FIELD_CONTROL_MOLLIFIER, ctrl), "set ctrl")
  call check(mgd_region_set_field(region%handle, MGD_FIELD_TARGET_MOLLIFIER, tgt), "set tgt")
  loc = 0.0_c_double
  call check(mgd_region_add_source(region%handle, loc, 1.0e-2_c_double, 1.0_c_double / 30.0_c_double, &
       1.5_c_double, 0.0_c_double), "add source")
  call region%finalize()
  call check(mgd_state_put(region%handle, MGD_CONSERVED, 0_c_int, target), "initial condition")

  sens = solver%runAdjoint(region, costJ)
  write(*, '(A,1X,SP,ES23.15E3)') "J", costJ
  write(*, '(A,1X,SP,ES23.15E3)') "Grad0", sens
  call region%cleanup()

contains

  function lin(k) result(xk)
    ! k-th of n uniformly spaced points on [-14, 14], rounded as numpy.linspace rounds them
    integer, intent(in) :: k
    real(c_double) :: xk
    if (k == n) then
       xk = 14.0_c_double
    else
       xk = real(k - 1, c_double) * (28.0_c_double / real(n - 1, c_double)) - 14.0_c_double
    end if
  end function lin

  function box(lo, hi) result(ext)
    real(c_double), intent(in) :: lo, hi
    integer :: ext(2), k
    real(c_double) :: xk
    ext = [0, 0]
    do k = 1, n
       xk = lin(k)
       if (xk >= lo .and. xk <= hi) then
          if (ext(1) == 0) ext(1) = k
          ext(2) = k
       end if
    end do
  end function box

end program monopole_adjoint_hip
