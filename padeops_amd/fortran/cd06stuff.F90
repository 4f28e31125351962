! Drop-in for the reference's cd06 module (src/derivatives/cd06.F90): same
! module, type and procedure names and argument orders; the arithmetic runs in
! libpadeops_hip.so.  As in the reference, dd1/dd2/dd3 take no boundary codes
! (the closure is fixed by init's bc1/bcn, cd06.F90:264-327).  The second
! derivative is not implemented in the reference (cd06der2 returns zeros,
! cd06.F90:841-853); here d2d1/d2d2/d2d3 and cd06der2 apply this library's
! definition: the periodic sixth-order tridiagonal scheme (alpha = 2/11) and,
! when bounded, the first derivative applied twice.
module cd06stuff
    use iso_c_binding
    use padeops_c
    implicit none
    private
    public :: cd06

    integer, parameter :: rkind = pade_rkind

    type cd06
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
        procedure :: cd06der2
        procedure :: apply_dev
    end type

contains

    ! cd06.F90:129-201; ierr 3 periodic n < 6, 2 bounded n < 8
    function init(this, n_, dx_, periodic_, bc1_, bcn_) result(ierr)
        class(cd06), intent(inout) :: this
        integer, intent(in) :: n_
        real(rkind), intent(in) :: dx_
        logical, intent(in) :: periodic_
        integer, intent(in) :: bc1_, bcn_
        integer :: ierr
        integer(c_int) :: rc
        if (c_associated(this%h)) call this%destroy()
        rc = pade_op_create(PADE_CD06, int(n_, c_int), real(dx_, c_double), merge(1_c_int, 0_c_int, periodic_), &
                            int(bc1_, c_int), int(bcn_, c_int), this%h)
        if (rc == 2 .or. rc == 3 .or. rc == 7) then
            ierr = int(rc)
            return
        end if
        call pade_check(rc, "cd06%init")
        this%n = n_
        this%dx = dx_
        this%periodic = periodic_
        ierr = 0
    end function

    ! cd06.F90:203-219
    subroutine destroy(this)
        class(cd06), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = pade_op_destroy(this%h)
        this%h = c_null_ptr
    end subroutine

    ! cd06.F90:123-127
    pure function GetSize(this) result(val)
        class(cd06), intent(in) :: this
        integer :: val
        val = this%n
    end function

    ! cd06.F90:775-795
    subroutine dd1(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in) :: f
        real(rkind), dimension(this%n, na, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 1_c_int, f, df, na, nb, 0, 0, "cd06%dd1")
    end subroutine

    ! cd06.F90:797-817
    subroutine dd2(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in) :: f
        real(rkind), dimension(na, this%n, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 2_c_int, f, df, na, nb, 0, 0, "cd06%dd2")
    end subroutine

    ! cd06.F90:819-839
    subroutine dd3(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in) :: f
        real(rkind), dimension(na, nb, this%n), intent(out) :: df
        call pade_apply_host(this%h, PADE_D1, 3_c_int, f, df, na, nb, 0, 0, "cd06%dd3")
    end subroutine

    subroutine d2d1(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(this%n, na, nb), intent(in) :: f
        real(rkind), dimension(this%n, na, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 1_c_int, f, df, na, nb, 0, 0, "cd06%d2d1")
    end subroutine

    subroutine d2d2(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, this%n, nb), intent(in) :: f
        real(rkind), dimension(na, this%n, nb), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 2_c_int, f, df, na, nb, 0, 0, "cd06%d2d2")
    end subroutine

    subroutine d2d3(this, f, df, na, nb)
        class(cd06), intent(in) :: this
        integer, intent(in) :: na, nb
        real(rkind), dimension(na, nb, this%n), intent(in) :: f
        real(rkind), dimension(na, nb, this%n), intent(out) :: df
        call pade_apply_host(this%h, PADE_D2, 3_c_int, f, df, na, nb, 0, 0, "cd06%d2d3")
    end subroutine

    ! cd06.F90:841-853 (a 1-D pencil)
    function cd06der2(this, f) result(df)
        class(cd06), intent(in) :: this
        real(rkind), dimension(this%n), intent(in) :: f
        real(rkind), dimension(this%n) :: df
        call pade_apply_host(this%h, PADE_D2, 1_c_int, f, df, 1, 1, 0, 0, "cd06%cd06der2")
    end function

    ! Device-resident arrays (no reference counterpart).
    subroutine apply_dev(this, oper, axis, f, df, na, nb, stream, adjoint)
        class(cd06), intent(in) :: this
        integer(c_int), intent(in) :: oper, axis
        type(c_ptr), intent(in) :: f, df
        integer, intent(in) :: na, nb
        type(c_ptr), optional, intent(in) :: stream
        logical, optional, intent(in) :: adjoint
        type(c_ptr) :: s
        logical :: adj
        s = c_null_ptr
        if (present(stream)) s = stream
        adj = .false.
        if (present(adjoint)) adj = adjoint
        call pade_apply_device(this%h, oper, axis, f, df, na, nb, 0, 0, s, adj, "cd06%apply_dev")
    end subroutine

end module cd06stuff
