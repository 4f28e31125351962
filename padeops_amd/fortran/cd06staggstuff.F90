! Drop-in for the reference's staggered cd06 module
! (src/derivatives/cd06stagg.F90): same module, type and procedure names,
! argument orders and generic real / complex bindings; the arithmetic runs in
! libpadeops_hip.so.  Both forms: init_periodic (cd06stagg.F90:170-195) and
! init_nonperiodic (:197-232, even / odd / one-sided ends).  Edge (E) arrays
! hold nE = n + 1 planes along z (periodic: the last a copy of the first); a
! complex array is handled as a real one of twice the first extent, the real
! and imaginary parts alike.
module cd06staggstuff
    use iso_c_binding
    use padeops_c
    implicit none
    private
    public :: cd06stagg

    integer, parameter :: rkind = pade_rkind
    integer(c_int), parameter :: D1_E2C = 11, D1_C2E = 12, D2_C2C = 13, D2_E2E = 14, &
                                 INTERP_E2C = 15, INTERP_C2E = 16, D1_C2C = 17, D1_E2E = 18

    interface
        integer(c_int) function apply_host_c(op, oper, axis, f, df, na, nb, bc1, bcn) &
                bind(C, name="pade_op_apply_host")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: op
            integer(c_int), value :: oper, axis, na, nb, bc1, bcn
            complex(c_double_complex), intent(in) :: f(*)
            complex(c_double_complex), intent(out) :: df(*)
        end function
    end interface

    type cd06stagg
        private
        integer :: n = 0, nE = 0
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure, private :: init_periodic
        procedure, private :: init_nonperiodic
        generic :: init => init_periodic, init_nonperiodic
        procedure :: destroy
        procedure, private :: ddz_E2C_REAL, ddz_E2C_CMPLX, ddz_C2E_REAL, ddz_C2E_CMPLX
        procedure, private :: d2dz2_C2C_REAL, d2dz2_C2C_CMPLX, d2dz2_E2E_REAL, d2dz2_E2E_CMPLX
        procedure, private :: InterpZ_E2C_REAL, InterpZ_E2C_CMPLX, InterpZ_C2E_REAL, InterpZ_C2E_CMPLX
        procedure, private :: ddz_C2C_REAL, ddz_C2C_CMPLX, ddz_E2E_REAL, ddz_E2E_CMPLX
        generic :: ddz_C2C => ddz_C2C_REAL, ddz_C2C_CMPLX
        generic :: ddz_E2E => ddz_E2E_REAL, ddz_E2E_CMPLX
        generic :: ddz_E2C => ddz_E2C_REAL, ddz_E2C_CMPLX
        generic :: ddz_C2E => ddz_C2E_REAL, ddz_C2E_CMPLX
        generic :: d2dz2_C2C => d2dz2_C2C_REAL, d2dz2_C2C_CMPLX
        generic :: d2dz2_E2E => d2dz2_E2E_REAL, d2dz2_E2E_CMPLX
        generic :: InterpZ_E2C => InterpZ_E2C_REAL, InterpZ_E2C_CMPLX
        generic :: InterpZ_C2E => InterpZ_C2E_REAL, InterpZ_C2E_CMPLX
    end type

contains

    ! cd06stagg.F90:170-195; GracefulExit 21 for nx <= 4
    subroutine init_periodic(this, nx, dx)
        class(cd06stagg), intent(inout) :: this
        integer, intent(in) :: nx
        real(rkind), intent(in) :: dx
        if (c_associated(this%h)) call this%destroy()
        if (nx <= 4) call pade_fail("CD06_stagg requires at least 4 points", 21)
        call pade_check(pade_op_create(PADE_CD06STAGG, int(nx, c_int), real(dx, c_double), 1_c_int, 0_c_int, &
                                       0_c_int, this%h), "cd06stagg%init")
        this%n = nx
        this%nE = nx + 1
    end subroutine

    ! cd06stagg.F90:197-232: "Bot" is the k = 1 end (bc1), "Top" the k = n end
    ! (bcn); each end code is +-1 (even / odd), +-2 with the one-sided closure
    subroutine init_nonperiodic(this, nx, dx, isTopEven, isBotEven, isTopSided, isBotSided)
        class(cd06stagg), intent(inout) :: this
        integer, intent(in) :: nx
        real(rkind), intent(in) :: dx
        logical, intent(in) :: isTopEven, isBotEven
        logical, intent(in), optional :: isTopSided, isBotSided
        integer(c_int) :: bc1, bcn
        if (c_associated(this%h)) call this%destroy()
        if (nx <= 4) call pade_fail("CD06_stagg requires at least 4 points", 21)
        bc1 = merge(1_c_int, -1_c_int, isBotEven)
        bcn = merge(1_c_int, -1_c_int, isTopEven)
        if (present(isBotSided)) then
            if (isBotSided) bc1 = 2 * bc1
        end if
        if (present(isTopSided)) then
            if (isTopSided) bcn = 2 * bcn
        end if
        call pade_check(pade_op_create(PADE_CD06STAGG, int(nx, c_int), real(dx, c_double), 0_c_int, bc1, &
                                       bcn, this%h), "cd06stagg%init")
        this%n = nx
        this%nE = nx + 1
    end subroutine

    subroutine destroy(this)
        class(cd06stagg), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = pade_op_destroy(this%h)
        this%h = c_null_ptr
    end subroutine

    subroutine run_r(this, oper, f, df, n1, n2, where)
        class(cd06stagg), intent(in) :: this
        integer(c_int), intent(in) :: oper
        real(rkind), intent(in) :: f(*)
        real(rkind), intent(out) :: df(*)
        integer, intent(in) :: n1, n2
        character(len=*), intent(in) :: where
        call pade_apply_host(this%h, oper, 3_c_int, f, df, n1, n2, 0, 0, where)
    end subroutine

    subroutine run_c(this, oper, f, df, n1, n2, where)
        class(cd06stagg), intent(in) :: this
        integer(c_int), intent(in) :: oper
        complex(rkind), intent(in) :: f(*)
        complex(rkind), intent(out) :: df(*)
        integer, intent(in) :: n1, n2
        character(len=*), intent(in) :: where
        if (.not. c_associated(this%h)) call pade_fail(where // ": operator used before init", 400)
        call pade_check(apply_host_c(this%h, oper, 3_c_int, f, df, int(2 * n1, c_int), int(n2, c_int), &
                                     0_c_int, 0_c_int), where)
    end subroutine

    ! cd06stagg.F90:820-848
    subroutine ddz_E2C_REAL(this, fE, dfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        real(rkind), dimension(n1, n2, this%n), intent(out) :: dfC
        call run_r(this, D1_E2C, fE, dfC, n1, n2, "cd06stagg%ddz_E2C")
    end subroutine
    subroutine ddz_E2C_CMPLX(this, fE, dfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        complex(rkind), dimension(n1, n2, this%n), intent(out) :: dfC
        call run_c(this, D1_E2C, fE, dfC, n1, n2, "cd06stagg%ddz_E2C")
    end subroutine

    ! cd06stagg.F90:850-881
    subroutine ddz_C2E_REAL(this, fC, dfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        real(rkind), dimension(n1, n2, this%nE), intent(out) :: dfE
        call run_r(this, D1_C2E, fC, dfE, n1, n2, "cd06stagg%ddz_C2E")
    end subroutine
    subroutine ddz_C2E_CMPLX(this, fC, dfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        complex(rkind), dimension(n1, n2, this%nE), intent(out) :: dfE
        call run_c(this, D1_C2E, fC, dfE, n1, n2, "cd06stagg%ddz_C2E")
    end subroutine

    ! cd06stagg.F90:995-1025
    subroutine d2dz2_C2C_REAL(this, fC, d2fC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        real(rkind), dimension(n1, n2, this%n), intent(out) :: d2fC
        call run_r(this, D2_C2C, fC, d2fC, n1, n2, "cd06stagg%d2dz2_C2C")
    end subroutine
    subroutine d2dz2_C2C_CMPLX(this, fC, d2fC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        complex(rkind), dimension(n1, n2, this%n), intent(out) :: d2fC
        call run_c(this, D2_C2C, fC, d2fC, n1, n2, "cd06stagg%d2dz2_C2C")
    end subroutine

    ! cd06stagg.F90:1027-1060
    subroutine d2dz2_E2E_REAL(this, fE, d2fE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        real(rkind), dimension(n1, n2, this%nE), intent(out) :: d2fE
        call run_r(this, D2_E2E, fE, d2fE, n1, n2, "cd06stagg%d2dz2_E2E")
    end subroutine
    subroutine d2dz2_E2E_CMPLX(this, fE, d2fE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        complex(rkind), dimension(n1, n2, this%nE), intent(out) :: d2fE
        call run_c(this, D2_E2E, fE, d2fE, n1, n2, "cd06stagg%d2dz2_E2E")
    end subroutine

    ! cd06stagg.F90:928-958
    subroutine InterpZ_E2C_REAL(this, fE, IfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        real(rkind), dimension(n1, n2, this%n), intent(out) :: IfC
        call run_r(this, INTERP_E2C, fE, IfC, n1, n2, "cd06stagg%InterpZ_E2C")
    end subroutine
    subroutine InterpZ_E2C_CMPLX(this, fE, IfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        complex(rkind), dimension(n1, n2, this%n), intent(out) :: IfC
        call run_c(this, INTERP_E2C, fE, IfC, n1, n2, "cd06stagg%InterpZ_E2C")
    end subroutine

    ! cd06stagg.F90:883-903 (bounded form)
    subroutine ddz_C2C_REAL(this, fC, dfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        real(rkind), dimension(n1, n2, this%n), intent(out) :: dfC
        call run_r(this, D1_C2C, fC, dfC, n1, n2, "cd06stagg%ddz_C2C")
    end subroutine
    subroutine ddz_C2C_CMPLX(this, fC, dfC, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        complex(rkind), dimension(n1, n2, this%n), intent(out) :: dfC
        call run_c(this, D1_C2C, fC, dfC, n1, n2, "cd06stagg%ddz_C2C")
    end subroutine

    ! cd06stagg.F90:905-925 (bounded form)
    subroutine ddz_E2E_REAL(this, fE, dfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        real(rkind), dimension(n1, n2, this%nE), intent(out) :: dfE
        call run_r(this, D1_E2E, fE, dfE, n1, n2, "cd06stagg%ddz_E2E")
    end subroutine
    subroutine ddz_E2E_CMPLX(this, fE, dfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%nE), intent(in) :: fE
        complex(rkind), dimension(n1, n2, this%nE), intent(out) :: dfE
        call run_c(this, D1_E2E, fE, dfE, n1, n2, "cd06stagg%ddz_E2E")
    end subroutine

    ! cd06stagg.F90:960-993
    subroutine InterpZ_C2E_REAL(this, fC, IfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        real(rkind), dimension(n1, n2, this%nE), intent(out) :: IfE
        call run_r(this, INTERP_C2E, fC, IfE, n1, n2, "cd06stagg%InterpZ_C2E")
    end subroutine
    subroutine InterpZ_C2E_CMPLX(this, fC, IfE, n1, n2)
        class(cd06stagg), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in) :: fC
        complex(rkind), dimension(n1, n2, this%nE), intent(out) :: IfE
        call run_c(this, INTERP_C2E, fC, IfE, n1, n2, "cd06stagg%InterpZ_C2E")
    end subroutine

end module cd06staggstuff
