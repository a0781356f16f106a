! Fortran binding of libocfd_ec.so (include/ocfd_ec.h) for the OpenCFD-EC host.
!
! module ocfd_ec       iso_c_binding interfaces of the C ABI, one for one with the header.
! module ocfd_ec_host  the glue a host built on the reference's modules calls:
!   ocfd_host_create(ctx)   Global_Var (after set_const_para) -> ocfd_params, every block of
!                           Mesh(1) -> ocfd_add_block with its Block_TYPE geometry and bc_msg table,
!                           then U of every block is uploaded once
!   ocfd_host_step(ctx)     in place of NS_Time_advance_LU_SGS(1) (sub_time_advance.f90:75-118)
!                           incl. the Kstep/tt bookkeeping and force_vt_kw of NS_Time_advance
!   ocfd_host_fetch(ctx)    U (and Res, dt, mu, mu_t) back into Mesh(1)%Block(:) for output,
!                           comput_force or a restart file
!   ocfd_host_norms(ctx)    in place of comput_max_Res_onemesh(1) (sub_time_advance.f90:330-380)
! Errors stop the program with the library's message, as the reference does on NaN or an
! unknown boundary condition.
module ocfd_ec
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OCFD_OK = 0, OCFD_EHIP = 1, OCFD_EARG = 2, OCFD_ENAN = 3, OCFD_EUNSUPPORTED = 4
  integer(c_int), parameter :: OCFD_F_RES = 1, OCFD_F_DU = 2, OCFD_F_DT = 3, OCFD_F_MU = 4, OCFD_F_MUT = 5, OCFD_F_DW = 6
  integer(c_int), parameter :: OCFD_NSTAGES = 7

  type, bind(C) :: ocfd_params
    real(c_double) :: Ma, Re, gamma, Cv, Cp, PrL, PrT, T_inf
    real(c_double) :: Twall
    real(c_double) :: AoA, A_alfa, A_beta, P_outlet
    real(c_double) :: CFL, dtmax, dtmin, dt_global, w_LU, MUT_MAX, Kt_inf, Wt_inf
    real(c_double) :: Ldmin, Ldmax, Lpmin, Lpmax, Lumax, LSAmax
    real(c_double) :: Ref_S, Ref_L, Centroid(3)
    real(c_double) :: Periodic_d(3)
    integer(c_int32_t) :: NVAR, If_viscous, Iflag_turbulence_model, Iflag_Scheme, Iflag_Flux
    integer(c_int32_t) :: IFlag_Reconstruction, Bound_Scheme, Iflag_local_dt, If_dtime_mesh
    integer(c_int32_t) :: IF_Scheme_Positivity, IFLAG_LIMIT_FLOW, Cood_Y_UP
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: IF_InnerFlow, IF_TurboMachinary
  end type ocfd_params

  ! met(q, s) holds the C member met[s][q]
  type, bind(C) :: ocfd_block_geom
    type(c_ptr) :: x, y, z, xc, yc, zc, vol, dtime_mesh, dw, Si, Sj, Sk
    type(c_ptr) :: n(9)
    type(c_ptr) :: met(9, 0:3)
  end type ocfd_block_geom

  interface
    integer(c_int) function ocfd_create(p, ctx) bind(C, name="ocfd_create")
      import :: c_int, c_ptr, ocfd_params
      type(ocfd_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ocfd_destroy(ctx) bind(C, name="ocfd_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function ocfd_last_error() bind(C, name="ocfd_last_error")
      import :: c_ptr
    end function
    integer(c_int) function ocfd_add_block(ctx, block_no, nx, ny, nz, nsub, bc_msg21, g) bind(C, name="ocfd_add_block")
      import :: c_int, c_int32_t, c_ptr, ocfd_block_geom
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no, nx, ny, nz, nsub
      integer(c_int32_t), intent(in) :: bc_msg21(*)
      type(ocfd_block_geom), intent(in) :: g
    end function
    integer(c_int) function ocfd_add_block_nodes(ctx, block_no, nx, ny, nz, nsub, bc_msg21, x, y, z) &
        bind(C, name="ocfd_add_block_nodes")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no, nx, ny, nz, nsub
      integer(c_int32_t), intent(in) :: bc_msg21(*)
      real(c_double), intent(in) :: x(*), y(*), z(*)
    end function
    integer(c_int) function ocfd_setup_geometry(ctx) bind(C, name="ocfd_setup_geometry")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_upload_U(ctx, block_no, U) bind(C, name="ocfd_upload_U")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no
      real(c_double), intent(in) :: U(*)
    end function
    integer(c_int) function ocfd_download_U(ctx, block_no, U) bind(C, name="ocfd_download_U")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no
      real(c_double), intent(out) :: U(*)
    end function
    integer(c_int) function ocfd_download_field(ctx, block_no, field, out) bind(C, name="ocfd_download_field")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no, field
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function ocfd_init_flow(ctx) bind(C, name="ocfd_init_flow")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_step_lusgs(ctx, niter) bind(C, name="ocfd_step_lusgs")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: niter
    end function
    integer(c_int) function ocfd_residual(ctx) bind(C, name="ocfd_residual")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_update_limit(ctx) bind(C, name="ocfd_update_limit")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_bc(ctx) bind(C, name="ocfd_bc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_halo(ctx) bind(C, name="ocfd_halo")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_force_vt_kw(ctx) bind(C, name="ocfd_force_vt_kw")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_res_norms(ctx, res_max, res_rms) bind(C, name="ocfd_res_norms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: res_max(*), res_rms(*)
    end function
    integer(c_int) function ocfd_forces(ctx, out12) bind(C, name="ocfd_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out12(12)
    end function
    integer(c_int) function ocfd_block_status(ctx, block_no, overlimit, nlimited) bind(C, name="ocfd_block_status")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_no
      integer(c_int32_t), intent(out) :: overlimit, nlimited
    end function
    integer(c_int) function ocfd_get_counters(ctx, kstep, tt) bind(C, name="ocfd_get_counters")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: kstep
      real(c_double), intent(out) :: tt
    end function
    integer(c_int) function ocfd_step_timed(ctx, niter, stage_ms) bind(C, name="ocfd_step_timed")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: niter
      real(c_double), intent(out) :: stage_ms(*)
    end function
    integer(c_int) function ocfd_set_graph(ctx, enable) bind(C, name="ocfd_set_graph")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function ocfd_sync(ctx) bind(C, name="ocfd_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ocfd_set_lu_tile(ctx, tj) bind(C, name="ocfd_set_lu_tile")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tj
    end function
    integer(c_int) function ocfd_autotune_lu(ctx, niter, best_tj, ms) bind(C, name="ocfd_autotune_lu")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: niter
      integer(c_int32_t), intent(out) :: best_tj
      real(c_double), intent(out) :: ms(3)
    end function
    integer(c_int) function ocfd_get_lu_tile(ctx, tj) bind(C, name="ocfd_get_lu_tile")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: tj
    end function
    integer(c_int) function ocfd_set_lu_schedule(ctx, schedule) bind(C, name="ocfd_set_lu_schedule")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: schedule
    end function
    integer(c_int) function ocfd_set_lu_fail_waits(ctx, on) bind(C, name="ocfd_set_lu_fail_waits")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function ocfd_mem_bytes(ctx, bytes) bind(C, name="ocfd_mem_bytes")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int) function ocfd_check_division(n, seed, numerator_exponent, mismatches) &
        bind(C, name="ocfd_check_division")
      import :: c_int, c_int32_t, c_int64_t
      integer(c_int64_t), value :: n, seed
      integer(c_int32_t), value :: numerator_exponent
      integer(c_int64_t), intent(out) :: mismatches
    end function
  end interface

contains

  ! Stop with the library's message when rc is not OCFD_OK.
  subroutine ocfd_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == OCFD_OK) return
    call c_f_pointer(ocfd_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    if (rc == OCFD_ENAN) then
      print *, 'NaN in Residual is found !'
    end if
    print *, 'libocfd_ec: ', where, ': ', trim(text)
    stop
  end subroutine ocfd_check
end module ocfd_ec


module ocfd_ec_host
  use iso_c_binding
  use ocfd_ec
  use Global_Var
  implicit none

contains

  subroutine ocfd_host_create(ctx)
    type(c_ptr), intent(out) :: ctx
    type(ocfd_params) :: p
    type(ocfd_block_geom) :: g
    type(Block_TYPE), pointer :: B
    integer(c_int32_t), allocatable :: tab(:)
    integer :: m, q, nsub

    p%Ma = Ma; p%Re = Re; p%gamma = gamma; p%Cv = Cv; p%Cp = Cp; p%PrL = PrL; p%PrT = PrT; p%T_inf = T_inf
    p%Twall = Twall; p%AoA = AoA; p%A_alfa = A_alfa; p%A_beta = A_beta; p%P_outlet = P_OUTLET
    p%CFL = CFL; p%dtmax = dtmax; p%dtmin = dtmin; p%dt_global = dt_global; p%w_LU = w_LU
    p%MUT_MAX = MUT_MAX; p%Kt_inf = Kt_inf; p%Wt_inf = Wt_inf
    p%Ldmin = Ldmin; p%Ldmax = Ldmax; p%Lpmin = Lpmin; p%Lpmax = Lpmax; p%Lumax = Lumax; p%LSAmax = LSAmax
    p%Ref_S = Ref_S; p%Ref_L = Ref_L; p%Centroid = Centroid
    p%Periodic_d = [Periodic_dX, Periodic_dY, Periodic_dZ]
    p%NVAR = NVAR; p%If_viscous = If_viscous; p%Iflag_turbulence_model = Iflag_turbulence_model
    p%Iflag_Scheme = Iflag_Scheme; p%Iflag_Flux = IFlag_flux; p%IFlag_Reconstruction = IFlag_Reconstruction
    p%Bound_Scheme = Bound_Scheme; p%Iflag_local_dt = Iflag_local_dt; p%If_dtime_mesh = If_dtime_mesh
    p%IF_Scheme_Positivity = IF_Scheme_Positivity; p%IFLAG_LIMIT_FLOW = IFLAG_LIMIT_FLOW
    p%Cood_Y_UP = Cood_Y_UP; p%precision = 0
    p%IF_InnerFlow = IF_InnerFlow; p%IF_TurboMachinary = IF_TurboMachinary
    call ocfd_check(ocfd_create(p, ctx), 'ocfd_create')

    do m = 1, Mesh(1)%Num_Block
      B => Mesh(1)%Block(m)
      nsub = B%subface
      allocate(tab(21 * nsub))
      do q = 1, nsub
        tab(21 * (q - 1) + 1:21 * q) = [B%bc_msg(q)%ib, B%bc_msg(q)%ie, B%bc_msg(q)%jb, B%bc_msg(q)%je, &
          B%bc_msg(q)%kb, B%bc_msg(q)%ke, B%bc_msg(q)%bc, B%bc_msg(q)%face, B%bc_msg(q)%f_no, &
          B%bc_msg(q)%ib1, B%bc_msg(q)%ie1, B%bc_msg(q)%jb1, B%bc_msg(q)%je1, B%bc_msg(q)%kb1, &
          B%bc_msg(q)%ke1, B%bc_msg(q)%nb1, B%bc_msg(q)%face1, B%bc_msg(q)%f_no1, &
          B%bc_msg(q)%L1, B%bc_msg(q)%L2, B%bc_msg(q)%L3]
      end do
      g%x = c_loc(B%x(0, 0, 0)); g%y = c_loc(B%y(0, 0, 0)); g%z = c_loc(B%z(0, 0, 0))
      g%xc = c_loc(B%xc(1 - LAP, 1 - LAP, 1 - LAP)); g%yc = c_loc(B%yc(1 - LAP, 1 - LAP, 1 - LAP))
      g%zc = c_loc(B%zc(1 - LAP, 1 - LAP, 1 - LAP))
      g%vol = c_loc(B%Vol(1, 1, 1)); g%dtime_mesh = c_loc(B%dtime_mesh(1, 1, 1)); g%dw = c_loc(B%dw(1, 1, 1))
      g%Si = c_loc(B%Si(1, 1, 1)); g%Sj = c_loc(B%Sj(1, 1, 1)); g%Sk = c_loc(B%Sk(1, 1, 1))
      g%n = [c_loc(B%ni1(1, 1, 1)), c_loc(B%ni2(1, 1, 1)), c_loc(B%ni3(1, 1, 1)), &
             c_loc(B%nj1(1, 1, 1)), c_loc(B%nj2(1, 1, 1)), c_loc(B%nj3(1, 1, 1)), &
             c_loc(B%nk1(1, 1, 1)), c_loc(B%nk2(1, 1, 1)), c_loc(B%nk3(1, 1, 1))]
      g%met(:, 0) = [c_loc(B%ix0(1, 1, 1)), c_loc(B%iy0(1, 1, 1)), c_loc(B%iz0(1, 1, 1)), &
                     c_loc(B%jx0(1, 1, 1)), c_loc(B%jy0(1, 1, 1)), c_loc(B%jz0(1, 1, 1)), &
                     c_loc(B%kx0(1, 1, 1)), c_loc(B%ky0(1, 1, 1)), c_loc(B%kz0(1, 1, 1))]
      g%met(:, 1) = [c_loc(B%ix1(1, 1, 1)), c_loc(B%iy1(1, 1, 1)), c_loc(B%iz1(1, 1, 1)), &
                     c_loc(B%jx1(1, 1, 1)), c_loc(B%jy1(1, 1, 1)), c_loc(B%jz1(1, 1, 1)), &
                     c_loc(B%kx1(1, 1, 1)), c_loc(B%ky1(1, 1, 1)), c_loc(B%kz1(1, 1, 1))]
      g%met(:, 2) = [c_loc(B%ix2(1, 1, 1)), c_loc(B%iy2(1, 1, 1)), c_loc(B%iz2(1, 1, 1)), &
                     c_loc(B%jx2(1, 1, 1)), c_loc(B%jy2(1, 1, 1)), c_loc(B%jz2(1, 1, 1)), &
                     c_loc(B%kx2(1, 1, 1)), c_loc(B%ky2(1, 1, 1)), c_loc(B%kz2(1, 1, 1))]
      g%met(:, 3) = [c_loc(B%ix3(1, 1, 1)), c_loc(B%iy3(1, 1, 1)), c_loc(B%iz3(1, 1, 1)), &
                     c_loc(B%jx3(1, 1, 1)), c_loc(B%jy3(1, 1, 1)), c_loc(B%jz3(1, 1, 1)), &
                     c_loc(B%kx3(1, 1, 1)), c_loc(B%ky3(1, 1, 1)), c_loc(B%kz3(1, 1, 1))]
      call ocfd_check(ocfd_add_block(ctx, int(B%Block_no, c_int32_t), int(B%nx, c_int32_t), int(B%ny, c_int32_t), &
                                     int(B%nz, c_int32_t), int(nsub, c_int32_t), tab, g), 'ocfd_add_block')
      call ocfd_check(ocfd_upload_U(ctx, int(B%Block_no, c_int32_t), B%U), 'ocfd_upload_U')
      deallocate(tab)
    end do
  end subroutine ocfd_host_create

  ! NS_Time_advance(1) with Time_Method = LU-SGS, on the device
  subroutine ocfd_host_step(ctx)
    type(c_ptr), intent(in) :: ctx
    call ocfd_check(ocfd_step_lusgs(ctx, 1_c_int32_t), 'ocfd_step_lusgs')
    Mesh(1)%tt = Mesh(1)%tt + dt_global
    Mesh(1)%Kstep = Mesh(1)%Kstep + 1
  end subroutine ocfd_host_step

  subroutine ocfd_host_fetch(ctx)
    type(c_ptr), intent(in) :: ctx
    type(Block_TYPE), pointer :: B
    integer :: m
    integer(c_int32_t) :: no
    do m = 1, Mesh(1)%Num_Block
      B => Mesh(1)%Block(m)
      no = int(B%Block_no, c_int32_t)
      call ocfd_check(ocfd_download_U(ctx, no, B%U), 'ocfd_download_U')
      call ocfd_check(ocfd_download_field(ctx, no, OCFD_F_RES, B%Res), 'ocfd_download_field')
      call ocfd_check(ocfd_download_field(ctx, no, OCFD_F_DT, B%dt), 'ocfd_download_field')
      if (If_viscous == 1) then
        call ocfd_check(ocfd_download_field(ctx, no, OCFD_F_MU, B%mu), 'ocfd_download_field')
        call ocfd_check(ocfd_download_field(ctx, no, OCFD_F_MUT, B%mu_t), 'ocfd_download_field')
      end if
    end do
  end subroutine ocfd_host_fetch

  subroutine ocfd_host_norms(ctx)
    type(c_ptr), intent(in) :: ctx
    call ocfd_check(ocfd_res_norms(ctx, Mesh(1)%Res_max, Mesh(1)%Res_rms), 'ocfd_res_norms')
  end subroutine ocfd_host_norms

  ! comput_force on the device (ocfd_forces): the console lines and the force.log and
  ! force-invis-vis.log records comput_force writes (sub_Post.f90:224-246), with no download of
  ! the state; only 12 doubles cross the bus
  subroutine ocfd_host_forces(ctx)
    type(c_ptr), intent(in) :: ctx
    real(c_double) :: f(12)
    integer :: u
    call ocfd_check(ocfd_forces(ctx, f), 'ocfd_forces')
    if (my_id == 0) then
      print*, "--------------Force and Moment -----------------------------"
      print*, "Total Force Coefficient (CL, CD, CS) ="
      print*, f(10), f(11), f(12)
      print*, "Total Moment Coefficient(CMx, CMy, CMz)="
      print*, f(7), f(8), f(9)
      open(newunit=u, file="force.log", position="append")
      write(u, "(I7,6E20.8)") Mesh(1)%Kstep, f(10), f(11), f(12), f(7), f(8), f(9)
      close(u)
      open(newunit=u, file="force-invis-vis.log", position="append")
      write(u, "(I7,6E20.8)") Mesh(1)%Kstep, f(1), f(2), f(3), f(4), f(5), f(6)
      close(u)
    end if
  end subroutine ocfd_host_forces
end module ocfd_ec_host
