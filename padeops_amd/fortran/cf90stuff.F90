! Drop-in for the reference's cf90 module (src/filters/cf90.F90): same module,
! type and procedure names, argument orders and optional arguments; the
! arithmetic runs in libpadeops_hip.so.  filter3 is generic as in the
! reference, with its real (cf90.F90:1160-1228) and complex (spectral data,
! cf90.F90:1230-1251) forms.
module cf90stuff
    use iso_c_binding
    use padeops_c
    implicit none
    private
    public :: cf90

    integer, parameter :: rkind = pade_rkind

    type cf90
        private
        integer :: n = 0
        logical :: periodic = .true.
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure :: init
        procedure :: destroy
        procedure :: filter1
        procedure :: filter2
        procedure, private :: filter3_REAL
        procedure, private :: filter3_CMPLX
        generic :: filter3 => filter3_REAL, filter3_CMPLX
        procedure :: apply_dev
    end type

contains

    ! cf90.F90:107-175; ierr 7 for n < 10
    function init(this, n_, periodic_) result(ierr)
        class(cf90), intent(inout) :: this
        integer, intent(in) :: n_
        logical, intent(in) :: periodic_
        integer :: ierr
        integer(c_int) :: rc
        if (c_associated(this%h)) call this%destroy()
        rc = pade_op_create(PADE_CF90, int(n_, c_int), 1.0_c_double, merge(1_c_int, 0_c_int, periodic_), &
                            0_c_int, 0_c_int, this%h)
        if (rc == 2 .or. rc == 3 .or. rc == 7) then
            ierr = int(rc)
            return
        end if
        call pade_check(rc, "cf90%init")
        this%n = n_
        this%periodic = periodic_
        ierr = 0
    end function

    ! cf90.F90:177-196
    subroutine destroy(this)
        class(cf90), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = pade_op_destroy(this%h)
        this%h = c_null_ptr
    end subroutine

    ! cf90.F90:1020-1088
    subroutine filter1(this, f, df, na, nb, bc1_, bcn_)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in) :: f
        real(rkind), dimension(this%n, na, nb), intent(out) :: df
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 1_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cf90%filter1")
    end subroutine

    ! cf90.F90:1090-1158
    subroutine filter2(this, f, df, na, nb, bc1_, bcn_)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in) :: f
        real(rkind), dimension(na, this%n, nb), intent(out) :: df
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 2_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cf90%filter2")
    end subroutine

    ! cf90.F90:1160-1228
    subroutine filter3_REAL(this, f, df, na, nb, bc1_, bcn_)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in) :: f
        real(rkind), dimension(na, nb, this%n), intent(out) :: df
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 3_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cf90%filter3")
    end subroutine

    ! cf90.F90:1230-1251: both parts filtered along z; the right-hand side of
    ! (bc1_, bcn_) when bcn_ is present, else (0, 0), solved with the (0, 0)
    ! left-hand side as the reference does
    subroutine filter3_CMPLX(this, f, df, na, nb, bc1_, bcn_)
        class(cf90), intent(in) :: this
        integer, intent(in) :: na, nb
        complex(rkind), dimension(na, nb, this%n), intent(in) :: f
        complex(rkind), dimension(na, nb, this%n), intent(out) :: df
        integer, optional, intent(in) :: bc1_, bcn_
        integer(c_int) :: b1, bn, has
        if (.not. c_associated(this%h)) call pade_fail("cf90%filter3: operator used before init", 400)
        b1 = 0; bn = 0; has = 0
        if (present(bcn_)) then
            has = 1
            b1 = int(pade_bc(bc1_, "bc1"), c_int)
            bn = int(pade_bc(bcn_, "bcn"), c_int)
        end if
        call pade_check(pade_op_filter3_complex_host(this%h, f, df, int(na, c_int), int(nb, c_int), b1, bn, has), &
                        "cf90%filter3")
    end subroutine

    ! Device-resident arrays (no reference counterpart).
    subroutine apply_dev(this, axis, f, df, na, nb, bc1_, bcn_, stream, adjoint)
        class(cf90), intent(in) :: this
        integer(c_int), intent(in) :: axis
        type(c_ptr), intent(in) :: f, df
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        type(c_ptr), optional, intent(in) :: stream
        logical, optional, intent(in) :: adjoint
        type(c_ptr) :: s
        logical :: adj
        s = c_null_ptr
        if (present(stream)) s = stream
        adj = .false.
        if (present(adjoint)) adj = adjoint
        call pade_apply_device(this%h, PADE_FILTER, axis, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                               pade_bc(bcn_, "bcn"), s, adj, "cf90%apply_dev")
    end subroutine

end module cf90stuff
