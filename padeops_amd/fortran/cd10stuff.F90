! Drop-in for the reference's cd10 module (src/derivatives/cd10.F90): same
! module, type and procedure names, argument orders and optional arguments;
! the arithmetic runs in libpadeops_hip.so.  Host arrays go through
! pade_op_apply_host; the *_dev procedures take device pointers.
module cd10stuff
    use iso_c_binding
    use padeops_c
    implicit none
    private
    public :: cd10

    integer, parameter :: rkind = pade_rkind

    type cd10
        private
        integer :: n = 0
        real(rkind) :: dx = 1.0_rkind
        logical :: periodic = .true.
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure :: init
        procedure :: destroy
        procedure :: GetSize
        procedure :: dd1
        procedure :: dd2
        procedure :: dd3
        procedure :: d2d1
        procedure :: d2d2
        procedure :: d2d3
        procedure :: apply_dev
    end type

contains

    ! cd10.F90:195-315; ierr 2 for n < 8
    function init(this, n_, dx_, periodic_, bc1_, bcn_) result(ierr)
        class(cd10), intent(inout) :: this
        integer, intent(in) :: n_
        real(rkind), intent(in) :: dx_
        logical, intent(in) :: periodic_
        integer, intent(in) :: bc1_, bcn_
        integer :: ierr
        integer(c_int) :: rc
        if (c_associated(this%h)) call this%destroy()
        rc = pade_op_create(PADE_CD10, int(n_, c_int), real(dx_, c_double), merge(1_c_int, 0_c_int, periodic_), &
                            int(bc1_, c_int), int(bcn_, c_int), this%h)
        if (rc == 2 .or. rc == 3 .or. rc == 7) then
            ierr = int(rc)
            return
        end if
        call pade_check(rc, "cd10%init")
        this%n = n_
        this%dx = dx_
        this%periodic = periodic_
        ierr = 0
    end function

    ! cd10.F90:317-349
    subroutine destroy(this)
        class(cd10), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = pade_op_destroy(this%h)
        this%h = c_null_ptr
    end subroutine

    ! cd10.F90:189-193
    pure function GetSize(this) result(val)
        class(cd10), intent(in) :: this
        integer :: val
        val = this%n
    end function

    ! cd10.F90:2029-2097
    subroutine dd1(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(this%n, na, nb), intent(in) :: f
        real(rkind), dimension(this%n, na, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 1_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%dd1")
    end subroutine

    ! cd10.F90:2099-2167
    subroutine dd2(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(na, this%n, nb), intent(in) :: f
        real(rkind), dimension(na, this%n, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 2_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%dd2")
    end subroutine

    ! cd10.F90:2169-2237
    subroutine dd3(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(na, nb, this%n), intent(in) :: f
        real(rkind), dimension(na, nb, this%n), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 3_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%dd3")
    end subroutine

    ! cd10.F90:2239-2307
    subroutine d2d1(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(this%n, na, nb), intent(in) :: f
        real(rkind), dimension(this%n, na, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 1_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%d2d1")
    end subroutine

    ! cd10.F90:2309-2377
    subroutine d2d2(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(na, this%n, nb), intent(in) :: f
        real(rkind), dimension(na, this%n, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 2_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%d2d2")
    end subroutine

    ! cd10.F90:2379-2447
    subroutine d2d3(this, f, df, na, nb, bc1_, bcn_)
        class(cd10), intent(in) :: this
        integer, intent(in) :: na, nb
        integer, optional, intent(in) :: bc1_, bcn_
        real(rkind), dimension(na, nb, this%n), intent(in) :: f
        real(rkind), dimension(na, nb, this%n), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 3_c_int, f, df, na, nb, pade_bc(bc1_, "bc1"), &
                             pade_bc(bcn_, "bcn"), "cd10%d2d3")
    end subroutine

    ! Device-resident arrays (no reference counterpart): oper PADE_D1 / PADE_D2,
    ! axis 1..3, f and df device pointers, asynchronous on `stream`.
    subroutine apply_dev(this, oper, axis, f, df, na, nb, bc1_, bcn_, stream, adjoint)
        class(cd10), intent(in) :: this
        integer(c_int), intent(in) :: oper, axis
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
        call pade_apply_device(this%h, oper, axis, f, df, na, nb, pade_bc(bc1_, "bc1"), pade_bc(bcn_, "bcn"), &
                               s, adj, "cd10%apply_dev")
    end subroutine

end module cd10stuff
