! ref_bc_shim.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A C-callable harness around the reference's own boundary-row routines
! APPLY_BC_{1,2,3,TH}_{LOWER,UPPER,LEFT,RIGHT} (src/boundary.F90:124-1760),
! which oracle/Makefile compiles unchanged from the reference tree, together
! with src/modules.F90 and the shipped rundir_unst/grid_def, into
! oracle/_ref/libboundary_ref.so.  The routines read module state (BC codes,
! metrics, fields, the wall model's arrays) and write module arrays (MATL*,
! MATD*, MATU*, VEC*); the harness sets the state from C arrays, fills the
! row arrays with a sentinel, calls one routine and copies the rows out, so
! that tests can compare the oracle's restated rows with the reference's.
! Only tests load it.
module ref_bc_shim
  ! renamed: the reference's run_variable has an array C_int
  use iso_c_binding, only: ci => c_int, cd => c_double
  use ntypes
  use Domain
  use Grid
  use run_variable
  use ADI_var
  implicit none
contains

  ! extents the reference was compiled with: NX, NY, NZ, NXP, NZV
  subroutine ref_bc_dims(d) bind(c, name="ref_bc_dims")
    integer(ci), intent(out) :: d(5)
    d = [int(NX, ci), int(NY, ci), int(NZ, ci), int(NXP, ci), int(NZV, ci)]
  end subroutine ref_bc_dims

  ! allocate the module arrays the rows read and write
  subroutine ref_bc_alloc() bind(c, name="ref_bc_alloc")
    if (allocated(MATLY)) return
    allocate(MATLY(0:NXP, 0:NY+1), MATDY(0:NXP, 0:NY+1), MATUY(0:NXP, 0:NY+1), VECY(0:NXP, 0:NY+1))
    allocate(MATLX(0:NXP, 0:NZ+1), MATDX(0:NXP, 0:NZ+1), MATUX(0:NXP, 0:NZ+1), VECX(0:NXP, 0:NZ+1))
    allocate(INT_JACOB(0:NZ+2, 0:NY+2), CJOB_21(0:NZ+2, 0:NY+2, 2), CJOB_22(0:NZ+2, 0:NY+2, 2))
    allocate(U1X(0:NXP, 0:NZV-1, 0:NY+1), U2X(0:NXP, 0:NZV-1, 0:NY+1), U3X(0:NXP, 0:NZV-1, 0:NY+1))
    allocate(THX(0:NXP, 0:NZV-1, 0:NY+1, 1:N_TH))
    allocate(U_BC_LOWER_WALLMODEL(0:NXP, 0:NZ+1), W_BC_LOWER_WALLMODEL(0:NXP, 0:NZ+1))
    allocate(V_BC_LOWER_WALLMODEL(0:NXP, 0:NZ+1))
    allocate(U_BC_UPPER_NWM(0:NXP, 0:NZ+1), W_BC_UPPER_NWM(0:NXP, 0:NZ+1))
    U1X = 0.d0; U2X = 0.d0; U3X = 0.d0; THX = 0.d0
    U_BC_LOWER_WALLMODEL = 0.d0; W_BC_LOWER_WALLMODEL = 0.d0; V_BC_LOWER_WALLMODEL = 0.d0
    U_BC_UPPER_NWM = 0.d0; W_BC_UPPER_NWM = 0.d0
    ! the mean profiles the Dirichlet z rows of U2, U3 read: zero, as
    ! VEL_PROFILE leaves them with Q = 0 (src/duct.F90:3699-3731)
    allocate(U1_bar(0:NY+2), U2_bar(0:NZV-1, 0:NY+1), U3_bar(0:NZV-1, 0:NY+1))
    U1_bar = 0.d0; U2_bar = 0.d0; U3_bar = 0.d0
    ! grid spacings the upper type-3 rows multiply by the unassigned
    ! U_BC_UPPER_NWM / W_BC_UPPER_NWM (zero here): any finite value
    allocate(DY(0:NY+1), DZ(0:NZ+1), DYF(0:NY), DZF(0:NZ))
    DY = 1.d0; DZ = 1.d0; DYF = 1.d0; DZF = 1.d0
  end subroutine ref_bc_alloc

  ! codes(4,4): rows U1, U2, U3, TH; columns YMIN, YMAX, ZMIN, ZMAX; c1 the C1 values
  subroutine ref_bc_set(codes, c1, ic_type_in, f_type_in, wall_type) bind(c, name="ref_bc_set")
    integer(ci), intent(in) :: codes(4, 4)
    real(cd), intent(in) :: c1(4, 4)
    integer(ci), value :: ic_type_in, f_type_in, wall_type
    U_BC_YMIN = codes(1, 1); U_BC_YMAX = codes(1, 2); U_BC_ZMIN = codes(1, 3); U_BC_ZMAX = codes(1, 4)
    V_BC_YMIN = codes(2, 1); V_BC_YMAX = codes(2, 2); V_BC_ZMIN = codes(2, 3); V_BC_ZMAX = codes(2, 4)
    W_BC_YMIN = codes(3, 1); W_BC_YMAX = codes(3, 2); W_BC_ZMIN = codes(3, 3); W_BC_ZMAX = codes(3, 4)
    TH_BC_YMIN(1) = codes(4, 1); TH_BC_YMAX(1) = codes(4, 2); TH_BC_ZMIN(1) = codes(4, 3); TH_BC_ZMAX(1) = codes(4, 4)
    U_BC_YMIN_C1 = c1(1, 1); U_BC_YMAX_C1 = c1(1, 2); U_BC_ZMIN_C1 = c1(1, 3); U_BC_ZMAX_C1 = c1(1, 4)
    V_BC_YMIN_C1 = c1(2, 1); V_BC_YMAX_C1 = c1(2, 2); V_BC_ZMIN_C1 = c1(2, 3); V_BC_ZMAX_C1 = c1(2, 4)
    W_BC_YMIN_C1 = c1(3, 1); W_BC_YMAX_C1 = c1(3, 2); W_BC_ZMIN_C1 = c1(3, 3); W_BC_ZMAX_C1 = c1(3, 4)
    TH_BC_YMIN_C1(1) = c1(4, 1); TH_BC_YMAX_C1(1) = c1(4, 2)
    TH_BC_ZMIN_C1(1) = c1(4, 3); TH_BC_ZMAX_C1(1) = c1(4, 4)
    IC_TYPE = ic_type_in
    F_TYPE = f_type_in
    WALL_MODEL_TYPE = wall_type
  end subroutine ref_bc_set

  ! metrics in the reference's own layout: INT_JACOB(0:NZ+2,0:NY+2),
  ! CJOB_21 and CJOB_22 (0:NZ+2,0:NY+2,2)
  subroutine ref_bc_set_metrics(jac, c21, c22) bind(c, name="ref_bc_set_metrics")
    real(cd), intent(in) :: jac(0:NZ+2, 0:NY+2), c21(0:NZ+2, 0:NY+2, 2), c22(0:NZ+2, 0:NY+2, 2)
    INT_JACOB = jac
    CJOB_21 = c21
    CJOB_22 = c22
  end subroutine ref_bc_set_metrics

  ! physical field f (1..3: U1X..U3X, 4: THX) from an (0:nx2-1, 0:NZ+1, 0:NY+1)
  ! array, x points 0..NXP
  subroutine ref_bc_set_field(f, nx2, a) bind(c, name="ref_bc_set_field")
    integer(ci), value :: f, nx2
    real(cd), intent(in) :: a(0:nx2-1, 0:NZ+1, 0:NY+1)
    select case (f)
    case (1)
      U1X(0:NXP, 0:NZ+1, :) = a(0:NXP, :, :)
    case (2)
      U2X(0:NXP, 0:NZ+1, :) = a(0:NXP, :, :)
    case (3)
      U3X(0:NXP, 0:NZ+1, :) = a(0:NXP, :, :)
    case (4)
      THX(0:NXP, 0:NZ+1, :, 1) = a(0:NXP, :, :)
    end select
  end subroutine ref_bc_set_field

  ! the wall model's lower-row values of U1 and U3 from (0:nx2-1, 0:NZ+1) arrays
  subroutine ref_bc_set_wall(nx2, u, w) bind(c, name="ref_bc_set_wall")
    integer(ci), value :: nx2
    real(cd), intent(in) :: u(0:nx2-1, 0:NZ+1), w(0:nx2-1, 0:NZ+1)
    U_BC_LOWER_WALLMODEL = u(0:NXP, :)
    W_BC_LOWER_WALLMODEL = w(0:NXP, :)
  end subroutine ref_bc_set_wall

  ! one routine: comp 0..2 velocity (APPLY_BC_1/2/3), 3 the scalar
  ! (APPLY_BC_TH, N = 1); side 0 LOWER, 1 UPPER (line = K), 2 LEFT, 3 RIGHT
  ! (line = J).  The rows of every x point 0..NXP (n = NY+2 or NZ+2 rows),
  ! unwritten entries left at the sentinel.
  subroutine ref_bc_apply(comp, side, line, sentinel, ml, md, mu, vv) bind(c, name="ref_bc_apply")
    integer(ci), value :: comp, side, line
    real(cd), value :: sentinel
    real(cd), intent(out) :: ml(0:NXP, 0:*), md(0:NXP, 0:*), mu(0:NXP, 0:*), vv(0:NXP, 0:*)
    integer :: n
    MATLY = sentinel; MATDY = sentinel; MATUY = sentinel; VECY = sentinel
    MATLX = sentinel; MATDX = sentinel; MATUX = sentinel; VECX = sentinel
    select case (comp * 4 + side)
    case (0); call APPLY_BC_1_LOWER(line)
    case (1); call APPLY_BC_1_UPPER(line)
    case (2); call APPLY_BC_1_LEFT(line)
    case (3); call APPLY_BC_1_RIGHT(line)
    case (4); call APPLY_BC_2_LOWER(line)
    case (5); call APPLY_BC_2_UPPER(line)
    case (6); call APPLY_BC_2_LEFT(line)
    case (7); call APPLY_BC_2_RIGHT(line)
    case (8); call APPLY_BC_3_LOWER(line)
    case (9); call APPLY_BC_3_UPPER(line)
    case (10); call APPLY_BC_3_LEFT(line)
    case (11); call APPLY_BC_3_RIGHT(line)
    case (12); call APPLY_BC_TH_LOWER(1, line)
    case (13); call APPLY_BC_TH_UPPER(1, line)
    case (14); call APPLY_BC_TH_LEFT(1, line)
    case (15); call APPLY_BC_TH_RIGHT(1, line)
    end select
    if (side < 2) then
      n = NY + 1
      ml(:, 0:n) = MATLY; md(:, 0:n) = MATDY; mu(:, 0:n) = MATUY; vv(:, 0:n) = VECY
    else
      n = NZ + 1
      ml(:, 0:n) = MATLX; md(:, 0:n) = MATDX; mu(:, 0:n) = MATUX; vv(:, 0:n) = VECX
    end if
  end subroutine ref_bc_apply
end module ref_bc_shim
