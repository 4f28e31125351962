! iso_c_binding interface to libpadeops_hip.so (include/padeops_hip.h) and the
! helpers shared by the operator modules.  Where the reference calls
! GracefulExit (src/utilities/exits.F90:28-35, which aborts the MPI job),
! these modules print the message and `error stop` with the same code.
module padeops_c
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: PADE_CD06 = 1, PADE_CD10 = 2, PADE_CF90 = 3, PADE_CD06STAGG = 4
    integer(c_int), parameter, public :: PADE_D1 = 1, PADE_D2 = 2, PADE_FILTER = 3
    integer, parameter, public :: pade_rkind = c_double
    integer, parameter, public :: pade_clen = 256

    public :: pade_op_create, pade_op_destroy, pade_op_size, pade_op_apply, pade_op_apply_adjoint
    public :: pade_op_apply_host, pade_op_filter3_complex_host, pade_last_error, pade_version
    public :: pade_check, pade_fail, pade_apply_host, pade_apply_device, pade_bc

    interface
        integer(c_int) function pade_op_create(kind, n, dx, periodic, bc1, bcn, op) &
                bind(C, name="pade_op_create")
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: kind, n, periodic, bc1, bcn
            real(c_double), value :: dx
            type(c_ptr) :: op
        end function

        integer(c_int) function pade_op_destroy(op) bind(C, name="pade_op_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: op
        end function

        integer(c_int) function pade_op_size(op) bind(C, name="pade_op_size")
            import :: c_int, c_ptr
            type(c_ptr), value :: op
        end function

        integer(c_int) function pade_op_apply(op, oper, axis, f, df, na, nb, bc1, bcn, stream) &
                bind(C, name="pade_op_apply")
            import :: c_int, c_ptr
            type(c_ptr), value :: op, f, df, stream
            integer(c_int), value :: oper, axis, na, nb, bc1, bcn
        end function

        integer(c_int) function pade_op_apply_adjoint(op, oper, axis, f, df, na, nb, bc1, bcn, stream) &
                bind(C, name="pade_op_apply_adjoint")
            import :: c_int, c_ptr
            type(c_ptr), value :: op, f, df, stream
            integer(c_int), value :: oper, axis, na, nb, bc1, bcn
        end function

        integer(c_int) function pade_op_apply_host(op, oper, axis, f, df, na, nb, bc1, bcn) &
                bind(C, name="pade_op_apply_host")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: op
            integer(c_int), value :: oper, axis, na, nb, bc1, bcn
            real(c_double), intent(in) :: f(*)
            real(c_double), intent(out) :: df(*)
        end function

        integer(c_int) function pade_op_filter3_complex_host(op, f, df, na, nb, bc1, bcn, has_bc) &
                bind(C, name="pade_op_filter3_complex_host")
            import :: c_int, c_ptr, c_double_complex
            type(c_ptr), value :: op
            integer(c_int), value :: na, nb, bc1, bcn, has_bc
            complex(c_double_complex), intent(in) :: f(*)
            complex(c_double_complex), intent(out) :: df(*)
        end function

        type(c_ptr) function pade_last_error() bind(C, name="pade_last_error")
            import :: c_ptr
        end function

        integer(c_int) function pade_version() bind(C, name="pade_version")
            import :: c_int
        end function
    end interface

    interface
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

contains

    ! Print a message and stop with `code` (the reference's GracefulExit).
    subroutine pade_fail(message, code)
        character(len=*), intent(in) :: message
        integer, intent(in) :: code
        write(*, '(a)') "ERROR: " // trim(message)
        error stop code
    end subroutine

    ! Stop with the library's code and message when rc /= 0.
    subroutine pade_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        character(len=512) :: msg
        integer :: i, n
        if (rc == 0) return
        p = pade_last_error()
        msg = ""
        if (c_associated(p)) then
            n = int(min(c_strlen(p), 512_c_size_t))
            call c_f_pointer(p, s, [n])
            do i = 1, n
                msg(i:i) = s(i)
            end do
        end if
        call pade_fail(where // ": " // trim(msg), int(rc))
    end subroutine

    ! Validated boundary code: absent -> 0; outside {-1,0,1} -> stop 324
    ! (src/derivatives/cd10.F90:2042-2058).
    integer function pade_bc(bc, which)
        integer, optional, intent(in) :: bc
        character(len=*), intent(in) :: which
        pade_bc = 0
        if (present(bc)) then
            if (bc /= 0 .and. bc /= 1 .and. bc /= -1) then
                call pade_fail("Incorrect boundary specification for " // which // &
                               " (should be 0, 1 or -1)", 324)
            end if
            pade_bc = bc
        end if
    end function

    ! Host arrays: copy to the device, apply, copy back (synchronous).
    subroutine pade_apply_host(h, oper, axis, f, df, na, nb, bc1, bcn, where)
        type(c_ptr), intent(in) :: h
        integer(c_int), intent(in) :: oper, axis
        integer, intent(in) :: na, nb, bc1, bcn
        real(c_double), intent(in) :: f(*)
        real(c_double), intent(out) :: df(*)
        character(len=*), intent(in) :: where
        if (.not. c_associated(h)) call pade_fail(where // ": operator used before init", 400)
        call pade_check(pade_op_apply_host(h, oper, axis, f, df, int(na, c_int), int(nb, c_int), &
                                           int(bc1, c_int), int(bcn, c_int)), where)
    end subroutine

    ! Device arrays (asynchronous on `stream`, c_null_ptr = default stream).
    subroutine pade_apply_device(h, oper, axis, f, df, na, nb, bc1, bcn, stream, adjoint, where)
        type(c_ptr), intent(in) :: h, f, df, stream
        integer(c_int), intent(in) :: oper, axis
        integer, intent(in) :: na, nb, bc1, bcn
        logical, intent(in) :: adjoint
        character(len=*), intent(in) :: where
        integer(c_int) :: rc
        if (.not. c_associated(h)) call pade_fail(where // ": operator used before init", 400)
        if (adjoint) then
            rc = pade_op_apply_adjoint(h, oper, axis, f, df, int(na, c_int), int(nb, c_int), &
                                       int(bc1, c_int), int(bcn, c_int), stream)
        else
            rc = pade_op_apply(h, oper, axis, f, df, int(na, c_int), int(nb, c_int), &
                               int(bc1, c_int), int(bcn, c_int), stream)
        end if
        call pade_check(rc, where)
    end subroutine

end module padeops_c
