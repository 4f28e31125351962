! Drop-in for the reference's Gaussian filter module (src/filters/gaussian.F90):
! same module, type and procedure names and argument orders; the filter,
! periodic or bounded (bc 0 / 1 / -1 at apply time), runs in libpadeops_hip.so
! (kind PADE_GAUSSIAN).
module gaussianstuff
    use iso_c_binding
    use padeops_c
    implicit none
    private
    public :: gaussian

    integer, parameter :: rkind = pade_rkind
    integer(c_int), parameter :: PADE_GAUSSIAN = 5

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

    type gaussian
        private
        integer :: n = 0
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure :: init
        procedure :: destroy
        procedure :: filter1
        procedure :: filter2
        procedure, private :: filter3_real
        procedure, private :: filter3_cmplx
        generic :: filter3 => filter3_real, filter3_cmplx
    end type

contains

    ! gaussian.F90:67-91; the stencil spans 9 points (stops with 21 for
    ! 1 < n < 9, where the reference would index outside its arrays)
    function init(this, n_, periodic_) result(ierr)
        class(gaussian), intent(inout) :: this
        integer, intent(in) :: n_
        logical, intent(in) :: periodic_
        integer :: ierr
        if (c_associated(this%h)) call this%destroy()
        call pade_check(pade_op_create(PADE_GAUSSIAN, int(n_, c_int), 1.0_c_double, merge(1_c_int, 0_c_int, periodic_), &
                                       0_c_int, 0_c_int, this%h), "gaussian%init")
        this%n = n_
        ierr = 0
    end function

    ! gaussian.F90:93-102
    subroutine destroy(this)
        class(gaussian), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = pade_op_destroy(this%h)
        this%h = c_null_ptr
    end subroutine

    ! gaussian.F90:104-334
    subroutine filter1(this, f, fil, n2, n3, bc1_, bcn_)
        class(gaussian), intent(in) :: this
        integer, intent(in) :: n2, n3
        real(rkind), dimension(this%n, n2, n3), intent(in) :: f
        real(rkind), dimension(this%n, n2, n3), intent(out) :: fil
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 1_c_int, f, fil, n2, n3, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "gaussian%filter1")
    end subroutine

    ! gaussian.F90:336-562
    subroutine filter2(this, f, fil, n1, n3, bc1_, bcn_)
        class(gaussian), intent(in) :: this
        integer, intent(in) :: n1, n3
        real(rkind), dimension(n1, this%n, n3), intent(in) :: f
        real(rkind), dimension(n1, this%n, n3), intent(out) :: fil
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 2_c_int, f, fil, n1, n3, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "gaussian%filter2")
    end subroutine

    ! gaussian.F90:564-572
    subroutine filter3_real(this, f, fil, n1, n2, bc1_, bcn_)
        class(gaussian), intent(in) :: this
        integer, intent(in) :: n1, n2
        real(rkind), dimension(n1, n2, this%n), intent(in) :: f
        real(rkind), dimension(n1, n2, this%n), intent(out) :: fil
        integer, optional, intent(in) :: bc1_, bcn_
        call pade_apply_host(this%h, PADE_FILTER, 3_c_int, f, fil, n1, n2, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "gaussian%filter3")
    end subroutine

    ! gaussian.F90:574-582: both parts alike (a complex (n1, n2, n) array is a
    ! real (2 n1, n2, n) one)
    subroutine filter3_cmplx(this, f, fil, n1, n2, bc1_, bcn_)
        class(gaussian), intent(in) :: this
        integer, intent(in) :: n1, n2
        complex(rkind), dimension(n1, n2, this%n), intent(in) :: f
        complex(rkind), dimension(n1, n2, this%n), intent(out) :: fil
        integer, optional, intent(in) :: bc1_, bcn_
        if (.not. c_associated(this%h)) call pade_fail("gaussian%filter3: operator used before init", 400)
        call pade_check(apply_host_c(this%h, PADE_FILTER, 3_c_int, f, fil, int(2 * n1, c_int), int(n2, c_int), &
                                     int(pade_bc(bc1_, "bc1"), c_int), int(pade_bc(bcn_, "bcn"), c_int)), &
                        "gaussian%filter3")
    end subroutine

end module gaussianstuff
