! Test-infrastructure driver for the compiled reference (oracle/_ref).
!
! Replaces the reference's `program main` (opencfd_ec3d_v1.16a.f90:126-204) with a fixed
! sequence that calls the reference's own setup and time-advance routines on the case in the
! current directory and dumps their arrays for parity tests:
!   read_parameter, init, set_control_para, check_mesh_quality, Init_flow   (main :140-156)
!   optional interior state from U_init.bin, then Boundary_condition_onemesh + update_buffer_onemesh
!   nsteps x NS_Time_advance(1)                                            (main :163-172)
! Dumps (stream files of named arrays, Fortran order, with bounds):
!   geom.bin   - every geometric array of every block after setup
!   state0.bin - U after initialisation (BC and halo applied)
!   state1.bin - U, Res, DU, dt, mu, mu_t, Surf1..6 after the first iteration
!   stateN.bin - the same after the last iteration
!   norms.txt  - Kstep, Res_max(1:NVAR), Res_rms(1:NVAR) every kshow steps (comput_max_Res_onemesh)
!   force.log  - written by the reference's comput_force every kshow steps
!   timing.txt - nsteps and the wall-clock seconds of the NS_Time_advance loop (system_clock)
!   flow3d.dat, SA3d.dat/SST3d.dat, Step_mess.dat, wall_dist.dat - the reference's own restart and
!                wall-distance writers (output_flow, write_dw) after the last iteration
! Usage: ref_driver NSTEPS [KSHOW [DUMP]]   (DUMP=0 writes no array dumps, for timing runs)
module dump_util
  implicit none
contains
  subroutine put_hdr(u, name, kind, nd, lb, ub)
    integer, intent(in) :: u, kind, nd, lb(:), ub(:)
    character(*), intent(in) :: name
    integer :: i
    write(u) int(len(name), 4), name, int(kind, 4), int(nd, 4), (int(lb(i), 4), int(ub(i), 4), i = 1, nd)
  end subroutine

  subroutine put_r3(u, name, a)
    integer, intent(in) :: u
    character(*), intent(in) :: name
    real(8), pointer :: a(:,:,:)
    if (.not. associated(a)) return
    call put_hdr(u, name, 1, 3, lbound(a), ubound(a))
    write(u) a
  end subroutine

  subroutine put_r4(u, name, a)
    integer, intent(in) :: u
    character(*), intent(in) :: name
    real(8), pointer :: a(:,:,:,:)
    if (.not. associated(a)) return
    call put_hdr(u, name, 1, 4, lbound(a), ubound(a))
    write(u) a
  end subroutine

  subroutine put_i3(u, name, a)
    integer, intent(in) :: u
    character(*), intent(in) :: name
    integer, pointer :: a(:,:,:)
    if (.not. associated(a)) return
    call put_hdr(u, name, 2, 3, lbound(a), ubound(a))
    write(u) a
  end subroutine

  subroutine put_block(u, B, nvar)
    use Mod_Type_Def
    integer, intent(in) :: u, nvar
    type(Block_TYPE), intent(in) :: B
    integer :: lb(4), ub(4)
    lb = 0; ub = 0
    lb(1) = B%Block_no; ub(1) = B%nx; lb(2) = B%ny; ub(2) = B%nz; lb(3) = nvar; ub(3) = B%IF_OverLimit
    call put_hdr(u, 'block', 0, 3, lb(1:3), ub(1:3))
  end subroutine
end module dump_util

program ref_driver
  use Global_Var
  use Flow_Var
  use dump_util
  implicit none
  integer :: ierr, nsteps, kshow, dump, n, mb, i, j, k, m, u
  integer(8) :: c0, c1, crate
  character(len=64) :: arg
  logical :: have_u
  type(Block_TYPE), pointer :: B
  real(8), allocatable :: ubuf(:,:,:,:)

  call mpi_init(ierr)
  call mpi_comm_rank(MPI_COMM_WORLD, my_id, ierr)
  call mpi_comm_size(MPI_COMM_WORLD, Total_proc, ierr)
  nsteps = 1; kshow = 1; dump = 1
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read(arg, *) nsteps
  end if
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg); read(arg, *) kshow
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg); read(arg, *) dump
  end if

  call read_parameter
!$ call omp_set_num_threads(NUM_THREADS)
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

  if (dump /= 0) then
    call dump_geometry('geom.bin')
    call dump_state('state0.bin')
  end if

  open(newunit=u, file='norms.txt', status='replace')
  call system_clock(c0, crate)
  do n = 1, nsteps
    call NS_Time_advance(1)
    if (n == 1 .and. dump /= 0) call dump_state('state1.bin')
    if (mod(Mesh(1)%Kstep, kshow) == 0) then
      call comput_force
      call comput_max_Res_onemesh(1)
      write(u, '(I8,14ES26.17E3)') Mesh(1)%Kstep, Mesh(1)%Res_max(1:NVAR), Mesh(1)%Res_rms(1:NVAR)
    end if
  end do
  call system_clock(c1)
  close(u)
  open(newunit=u, file='timing.txt', status='replace')
  write(u, '(I8,ES26.17E3)') nsteps, dble(c1 - c0) / dble(crate)
  close(u)
  if (dump /= 0) then
    call dump_state('stateN.bin')
    call output_flow                                   ! flow3d.dat (+ SA3d/SST3d), Step_mess.dat
    if (Mesh(1)%Iflag_turbulence_model /= 0) call write_dw   ! wall_dist.dat
  end if
  call mpi_finalize(ierr)

contains

  subroutine dump_geometry(fname)
    character(*), intent(in) :: fname
    integer :: f
    open(newunit=f, file=fname, access='stream', form='unformatted', status='replace')
    do mb = 1, Mesh(1)%Num_Block
      B => Mesh(1)%Block(mb)
      call put_block(f, B, NVAR)
      call put_r3(f, 'x', B%x); call put_r3(f, 'y', B%y); call put_r3(f, 'z', B%z)
      call put_r3(f, 'xc', B%xc); call put_r3(f, 'yc', B%yc); call put_r3(f, 'zc', B%zc)
      call put_r3(f, 'vol', B%vol)
      call put_r3(f, 'Si', B%Si); call put_r3(f, 'Sj', B%Sj); call put_r3(f, 'Sk', B%Sk)
      call put_r3(f, 'ni1', B%ni1); call put_r3(f, 'ni2', B%ni2); call put_r3(f, 'ni3', B%ni3)
      call put_r3(f, 'nj1', B%nj1); call put_r3(f, 'nj2', B%nj2); call put_r3(f, 'nj3', B%nj3)
      call put_r3(f, 'nk1', B%nk1); call put_r3(f, 'nk2', B%nk2); call put_r3(f, 'nk3', B%nk3)
      call put_r3(f, 'ix1', B%ix1); call put_r3(f, 'iy1', B%iy1); call put_r3(f, 'iz1', B%iz1)
      call put_r3(f, 'jx1', B%jx1); call put_r3(f, 'jy1', B%jy1); call put_r3(f, 'jz1', B%jz1)
      call put_r3(f, 'kx1', B%kx1); call put_r3(f, 'ky1', B%ky1); call put_r3(f, 'kz1', B%kz1)
      call put_r3(f, 'ix2', B%ix2); call put_r3(f, 'iy2', B%iy2); call put_r3(f, 'iz2', B%iz2)
      call put_r3(f, 'jx2', B%jx2); call put_r3(f, 'jy2', B%jy2); call put_r3(f, 'jz2', B%jz2)
      call put_r3(f, 'kx2', B%kx2); call put_r3(f, 'ky2', B%ky2); call put_r3(f, 'kz2', B%kz2)
      call put_r3(f, 'ix3', B%ix3); call put_r3(f, 'iy3', B%iy3); call put_r3(f, 'iz3', B%iz3)
      call put_r3(f, 'jx3', B%jx3); call put_r3(f, 'jy3', B%jy3); call put_r3(f, 'jz3', B%jz3)
      call put_r3(f, 'kx3', B%kx3); call put_r3(f, 'ky3', B%ky3); call put_r3(f, 'kz3', B%kz3)
      if (Mesh(1)%Iflag_turbulence_model /= 0) then
        call put_r3(f, 'ix0', B%ix0); call put_r3(f, 'iy0', B%iy0); call put_r3(f, 'iz0', B%iz0)
        call put_r3(f, 'jx0', B%jx0); call put_r3(f, 'jy0', B%jy0); call put_r3(f, 'jz0', B%jz0)
        call put_r3(f, 'kx0', B%kx0); call put_r3(f, 'ky0', B%ky0); call put_r3(f, 'kz0', B%kz0)
        call put_r3(f, 'dw', B%dw)
      end if
      call put_r3(f, 'dtime_mesh', B%dtime_mesh)
      call put_i3(f, 'BcI', B%BcI); call put_i3(f, 'BcJ', B%BcJ); call put_i3(f, 'BcK', B%BcK)
    end do
    close(f)
  end subroutine

  subroutine dump_state(fname)
    character(*), intent(in) :: fname
    integer :: f
    open(newunit=f, file=fname, access='stream', form='unformatted', status='replace')
    do mb = 1, Mesh(1)%Num_Block
      B => Mesh(1)%Block(mb)
      call put_block(f, B, NVAR)
      call put_r4(f, 'U', B%U)
      call put_r4(f, 'Res', B%Res)
      call put_r4(f, 'DU', B%DU)
      call put_r3(f, 'dt', B%dt)
      if (If_viscous == 1) then
        call put_r3(f, 'mu', B%mu); call put_r3(f, 'mu_t', B%mu_t)
        call put_r3(f, 'Surf1', B%Surf1); call put_r3(f, 'Surf2', B%Surf2)
        call put_r3(f, 'Surf3', B%Surf3); call put_r3(f, 'Surf4', B%Surf4)
        call put_r3(f, 'Surf5', B%Surf5); call put_r3(f, 'Surf6', B%Surf6)
      end if
    end do
    close(f)
  end subroutine
end program ref_driver
