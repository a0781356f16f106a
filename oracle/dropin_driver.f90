! Test-infrastructure driver: the reference's own Fortran host with the GPU library dropped in.
!
! The reference sets the case up exactly as its `program main` does (opencfd_ec3d_v1.16a.f90:
! 140-156: read_parameter, init, set_control_para, check_mesh_quality, Init_flow; optionally an
! interior state from U_init.bin followed by Boundary_condition_onemesh + update_buffer_onemesh).
! From that one state it then advances twice:
!   (1) on the device, through the iso_c_binding layer (module ocfd_ec_host): the reference's
!       Global_Var parameters, Block_TYPE geometry and bc_msg tables are handed over once and
!       ocfd_host_step stands in for NS_Time_advance_LU_SGS;
!   (2) on the host, with the reference's own NS_Time_advance(1).
! and writes dropin.txt:
!   res  <block> <m> <max relative difference of Res after the first iteration>
!   u    <block> <m> <max relative difference of U over 0..n after the last iteration>
!   norms_gpu / norms_ref  Res_max(1:NVAR) Res_rms(1:NVAR) after the last iteration
!   force_gpu  CL CD CS CMx CMy CMz after the last iteration (force.log holds the reference's)
! Relative differences use max(|ref|, 1e-8 max|ref|) as the denominator.
! Usage: dropin_driver NSTEPS
program dropin_driver
  use iso_c_binding
  use Global_Var
  use Flow_Var
  use ocfd_ec
  use ocfd_ec_host
  implicit none
  type :: arr4
    real(8), allocatable :: a(:,:,:,:)
  end type
  integer :: ierr, nsteps, n, mb, i, j, k, m, u, nx, ny, nz
  character(len=64) :: arg
  logical :: have_u
  type(Block_TYPE), pointer :: B
  real(8), allocatable :: ubuf(:,:,:,:)
  type(arr4), allocatable :: Ug(:), Rg(:), U0(:)
  real(8), allocatable :: gmax(:), grms(:)
  real(8) :: f12(12)
  type(c_ptr) :: ctx

  call mpi_init(ierr)
  call mpi_comm_rank(MPI_COMM_WORLD, my_id, ierr)
  call mpi_comm_size(MPI_COMM_WORLD, Total_proc, ierr)
  nsteps = 1
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read(arg, *) nsteps
  end if

  call read_parameter
  allocate(Mesh(Num_Mesh))
  call init
  call set_control_para
  call check_mesh_quality
  call Init_flow
  inquire(file='U_init.bin', exist=have_u)
  if (have_u) then
    open(newunit=u, file='U_init.bin', access='stream', form='unformatted', status='old')
    do mb = 1, Mesh(1)%Num_Block
      B => Mesh(1)%Block(mb)
      allocate(ubuf(NVAR, B%nx - 1, B%ny - 1, B%nz - 1))
      read(u) ubuf
      do k = 1, B%nz - 1
        do j = 1, B%ny - 1
          do i = 1, B%nx - 1
            do m = 1, NVAR
              B%U(m, i, j, k) = ubuf(m, i, j, k)
            end do
          end do
        end do
      end do
      deallocate(ubuf)
    end do
    close(u)
    call Boundary_condition_onemesh(1)
    call update_buffer_onemesh(1)
  end if

  ! ---- (1) the device path, from the reference's own host structures ----
  allocate(Ug(Mesh(1)%Num_Block), Rg(Mesh(1)%Num_Block), U0(Mesh(1)%Num_Block))
  do mb = 1, Mesh(1)%Num_Block
    B => Mesh(1)%Block(mb)
    U0(mb)%a = B%U
    allocate(Ug(mb)%a, mold=B%U)
    allocate(Rg(mb)%a, mold=B%Res)
  end do
  call ocfd_host_create(ctx)
  do n = 1, nsteps
    call ocfd_host_step(ctx)
    if (n == 1) then
      do mb = 1, Mesh(1)%Num_Block
        call ocfd_check(ocfd_download_field(ctx, int(mb, c_int32_t), OCFD_F_RES, Rg(mb)%a), 'download Res')
      end do
    end if
  end do
  call ocfd_host_norms(ctx)
  gmax = Mesh(1)%Res_max(1:NVAR)
  grms = Mesh(1)%Res_rms(1:NVAR)
  call ocfd_check(ocfd_forces(ctx, f12), 'ocfd_forces')
  call ocfd_host_forces(ctx)                  ! force.log line from the device (written first)
  do mb = 1, Mesh(1)%Num_Block
    call ocfd_check(ocfd_download_U(ctx, int(mb, c_int32_t), Ug(mb)%a), 'ocfd_download_U')
  end do
  call ocfd_check(ocfd_destroy(ctx), 'ocfd_destroy')

  ! ---- (2) the reference's own iteration from the same state ----
  do mb = 1, Mesh(1)%Num_Block
    Mesh(1)%Block(mb)%U = U0(mb)%a
  end do
  Mesh(1)%Kstep = 0
  Mesh(1)%tt = 0.d0
  open(newunit=u, file='dropin.txt', status='replace')
  do n = 1, nsteps
    call NS_Time_advance(1)
    if (n == 1) then
      do mb = 1, Mesh(1)%Num_Block
        B => Mesh(1)%Block(mb)
        nx = B%nx; ny = B%ny; nz = B%nz
        do m = 1, NVAR
          write(u, '(A,2I6,ES14.5)') 'res ', mb, m, &
            relmax(Rg(mb)%a(m, 1:nx-1, 1:ny-1, 1:nz-1), B%Res(m, 1:nx-1, 1:ny-1, 1:nz-1))
        end do
      end do
    end if
  end do
  call comput_max_Res_onemesh(1)
  call comput_force
  do mb = 1, Mesh(1)%Num_Block
    B => Mesh(1)%Block(mb)
    nx = B%nx; ny = B%ny; nz = B%nz
    do m = 1, NVAR
      write(u, '(A,2I6,ES14.5)') 'u   ', mb, m, relmax(Ug(mb)%a(m, 0:nx, 0:ny, 0:nz), B%U(m, 0:nx, 0:ny, 0:nz))
    end do
  end do
  write(u, '(A,14ES26.17E3)') 'norms_gpu ', gmax, grms
  write(u, '(A,14ES26.17E3)') 'norms_ref ', Mesh(1)%Res_max(1:NVAR), Mesh(1)%Res_rms(1:NVAR)
  write(u, '(A,6ES26.17E3)') 'force_gpu ', f12(10), f12(11), f12(12), f12(7), f12(8), f12(9)
  close(u)
  call mpi_finalize(ierr)

contains

  real(8) function relmax(a, r)
    real(8), intent(in) :: a(:,:,:), r(:,:,:)
    real(8) :: tau
    tau = 1.d-8 * maxval(abs(r))
    if (tau <= 0.d0) tau = 1.d-300
    relmax = maxval(abs(a - r) / max(abs(r), tau))
  end function relmax
end program dropin_driver
